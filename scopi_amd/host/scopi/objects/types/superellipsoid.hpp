// Reference header name; the slot API lives in scopi_hip_host.hpp.
#pragma once
#include "../../scopi_hip_host.hpp"
