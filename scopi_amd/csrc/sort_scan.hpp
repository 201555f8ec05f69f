// sort_scan.hpp -- host launchers of the scan and radix-sort primitives.
#pragma once

#include "common.hpp"

namespace scopi_hip {

// Exclusive prefix sum of int32 counts; keeps its own scratch buffer.
class Scanner {
  public:
    ~Scanner();
    // out[k] = in[0] + ... + in[k-1]; *d_total (device) = sum of all (may be null).
    // in and out must not alias when d_total is requested.
    void exclusive(const int* in, int* out, int64_t n, int* d_total, hipStream_t s);

  private:
    static int64_t tmp_size(int64_t n);
    void reserve(int64_t n);
    void scan_level(const int* in, int* out, int64_t n, int* tmp, hipStream_t s);
    int* tmp_ = nullptr;
    int64_t tmp_cap_ = 0;
};

// Stable LSD radix sort of (key, value) pairs, 8 bits per pass.
class RadixSorter {
  public:
    ~RadixSorter();
    void sort(uint32_t* keys, int32_t* vals, int64_t n, int key_bits, hipStream_t s, Scanner& scan);

  private:
    void reserve(int64_t n);
    int* hist_ = nullptr;
    uint32_t* keys_tmp_ = nullptr;
    int32_t* vals_tmp_ = nullptr;
    int64_t cap_ = 0;
};

}  // namespace scopi_hip
