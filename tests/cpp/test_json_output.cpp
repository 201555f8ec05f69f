// test_json_output.cpp -- the output document of the C++ host mirror (output_document +
// output::dump) rebuilt from a 2D scene state read from stdin, printed to stdout.
// tests/test_output_files.py feeds it the final state and contacts stored in the reference's
// fixtures (test/references/*.json, converted to the line format below) and compares the
// printed document with the fixture file byte for byte.  No GPU: only the host headers.
//
// Input lines (numbers as C99 hexadecimal floats):
//   problem nofriction|friction <mu>
//   obj sphere|plane <active 0/1> <x> <y> <q0> <q1> <q2> <q3> <radius> <vx> <vy> <vdx> <vdy> <w> <dw> <fx> <fy> <m> <j>
//   contact <i> <j> <dij> <nx> <ny> <pix> <piy> <pjx> <pjy>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include <scopi/scopi_hip_host.hpp>

using namespace scopi;

static double num(std::istringstream& in) {
    std::string t;
    in >> t;
    return std::strtod(t.c_str(), nullptr);
}

template <class problem_t>
static int emit(scopi_container<2>& p, const std::vector<std::vector<double>>& rows, double mu) {
    std::vector<neighbor<2, problem_t>> contacts;
    for (const auto& r : rows) {
        neighbor<2, problem_t> c;
        c.i = (std::size_t)r[0];
        c.j = (std::size_t)r[1];
        c.dij = r[2];
        c.nij[0] = r[3], c.nij[1] = r[4];
        c.pi[0] = r[5], c.pi[1] = r[6];
        c.pj[0] = r[7], c.pj[1] = r[8];
        if constexpr (std::is_same<problem_t, Friction>::value) c.property.mu = mu;
        contacts.push_back(c);
    }
    std::cout << output::dump(output_document(p, contacts), 4);
    return 0;
}

int main() {
    scopi_container<2> particles;
    std::vector<std::vector<double>> contacts;
    std::string line, problem = "nofriction";
    double mu = 0.1;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string tag;
        in >> tag;
        if (tag == "problem") {
            in >> problem;
            mu = num(in);
        } else if (tag == "obj") {
            std::string type;
            int active;
            in >> type >> active;
            fixed<2> x{{num(in), num(in)}};
            type::quaternion_t q{{num(in), num(in), num(in), num(in)}};
            const double radius = num(in);
            fixed<2> v{{num(in), num(in)}}, vd{{num(in), num(in)}};
            const double w = num(in), dw = num(in);
            fixed<2> f{{num(in), num(in)}};
            const double m = num(in), j = num(in);
            auto prop = property<2>().velocity(v).desired_velocity(vd).omega(w).desired_omega(dw).force(f).mass(m)
                            .moment_inertia(j);
            if (!active) prop.deactivate();
            if (type == "sphere")
                particles.push_back(sphere<2>(x, q, radius), prop);
            else
                particles.push_back(plane<2>(x, q), prop);
        } else if (tag == "contact") {
            std::vector<double> r(9);
            for (auto& v : r) v = num(in);
            contacts.push_back(r);
        }
    }
    if (problem == "friction") return emit<Friction>(particles, contacts, mu);
    return emit<NoFriction>(particles, contacts, mu);
}
