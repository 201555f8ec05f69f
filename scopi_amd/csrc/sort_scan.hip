// sort_scan.hip -- device primitives of the contact search: exclusive scan of
// int32 counts and a stable LSD radix sort of (uint32 key, int32 value) pairs.
//
// Radix sort: 8-bit digits.  Per pass, every block owns a tile of
// kSortItems*256 keys; a histogram kernel writes per-(digit, block) counts in
// digit-major order, an exclusive scan turns them into global offsets, and the
// scatter kernel ranks equal digits inside each 256-key slice with 64-lane
// ballots (wave-local rank) plus per-wave prefix counts in LDS, which keeps the
// sort stable.
#include "sort_scan.hpp"

namespace scopi_hip {

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;  // 2048
constexpr int kSortItems = 8;
constexpr int kSortTile = kBlock * kSortItems;  // 2048

// ------------------------------------------------------------------ scan
__device__ inline int block_exclusive_scan(int x, int* lds_wave, int& total) {
    // wave-level inclusive scan (64 lanes)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int v = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    if (lane == 63) lds_wave[wave] = v;
    __syncthreads();
    int wave_off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        int s = lds_wave[w];
        if (w < wave) wave_off += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wave_off + v - x;
}

__global__ void __launch_bounds__(kBlock) k_scan_tiles(const int* __restrict__ in, int* __restrict__ out,
                                                      int* __restrict__ sums, int64_t n) {
    __shared__ int lds_wave[kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int vals[kScanItems];
    int local = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t idx = base + k;
        vals[k] = (idx < n) ? in[idx] : 0;
        local += vals[k];
    }
    int total;
    int off = block_exclusive_scan(local, lds_wave, total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t idx = base + k;
        if (idx < n) out[idx] = off;
        off += vals[k];
    }
    if (threadIdx.x == 0 && sums) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kBlock) k_scan_add(int* __restrict__ out, const int* __restrict__ sums, int64_t n) {
    const int add = sums[blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    for (int k = threadIdx.x; k < kScanTile; k += kBlock) {
        int64_t idx = base + k;
        if (idx < n) out[idx] += add;
    }
}

// total = out[n-1] + in[n-1] (0 when n == 0)
__global__ void k_scan_total(const int* __restrict__ in, const int* __restrict__ out, int64_t n, int* __restrict__ total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *total = (n > 0) ? out[n - 1] + in[n - 1] : 0;
}

int64_t Scanner::tmp_size(int64_t n) {
    int64_t s = 0;
    while (n > kScanTile) {
        n = (n + kScanTile - 1) / kScanTile;
        s += n;
    }
    return s + 1;
}

void Scanner::reserve(int64_t n) {
    int64_t need = tmp_size(n);
    if (need > tmp_cap_) {
        if (tmp_) SCOPI_HIP_CHECK(hipFree(tmp_));
        tmp_cap_ = need + need / 2;
        SCOPI_HIP_CHECK(hipMalloc(&tmp_, tmp_cap_ * sizeof(int)));
    }
}

Scanner::~Scanner() {
    if (tmp_) (void)hipFree(tmp_);
}

void Scanner::scan_level(const int* in, int* out, int64_t n, int* tmp, hipStream_t s) {
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles <= 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kBlock), 0, s, in, out, (int*)nullptr, n);
        return;
    }
    int* sums = tmp;
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(kBlock), 0, s, in, out, sums, n);
    scan_level(sums, sums, tiles, tmp + tiles, s);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(kBlock), 0, s, out, sums, n);
}

void Scanner::exclusive(const int* in, int* out, int64_t n, int* d_total, hipStream_t s) {
    if (n > 0) {
        reserve(n);
        scan_level(in, out, n, tmp_, s);
    }
    if (d_total) hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, s, in, out, n, d_total);
}

// ------------------------------------------------------------------ radix sort
__global__ void __launch_bounds__(kBlock) k_radix_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                      int* __restrict__ hist, int nblocks) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int k = threadIdx.x; k < kSortTile; k += kBlock) {
        int64_t idx = base + k;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(kBlock) k_radix_scatter(const uint32_t* __restrict__ keys_in,
                                                         const int32_t* __restrict__ vals_in,
                                                         uint32_t* __restrict__ keys_out, int32_t* __restrict__ vals_out,
                                                         int64_t n, int shift, const int* __restrict__ offsets,
                                                         int nblocks) {
    constexpr int kWaves = kBlock / 64;
    __shared__ int digit_base[256];
    __shared__ int wave_count[kWaves][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    digit_base[tid] = offsets[(int64_t)tid * nblocks + blockIdx.x];
    const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int it = 0; it < kSortItems; ++it) {
        for (int w = 0; w < kWaves; ++w) wave_count[w][tid] = 0;
        __syncthreads();
        const int64_t idx = base + (int64_t)it * kBlock + tid;
        const bool valid = idx < n;
        uint32_t key = valid ? keys_in[idx] : 0u;
        int32_t val = valid ? vals_in[idx] : 0;
        const uint32_t dgt = (key >> shift) & 255u;
        // lanes with the same digit (valid lanes only)
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (dgt >> b) & 1u;
            const uint64_t vote = __ballot(bit);
            peers &= bit ? vote : ~vote;
        }
        const int rank = __popcll(peers & lt_mask);
        const int leader = peers ? (__ffsll((unsigned long long)peers) - 1) : 64;
        if (valid && lane == leader) wave_count[wave][dgt] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int off = digit_base[dgt];
            for (int w = 0; w < wave; ++w) off += wave_count[w][dgt];
            const int pos = off + rank;
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        int tot = 0;
        for (int w = 0; w < kWaves; ++w) tot += wave_count[w][tid];
        digit_base[tid] += tot;
        __syncthreads();
    }
}

RadixSorter::~RadixSorter() {
    if (hist_) (void)hipFree(hist_);
    if (keys_tmp_) (void)hipFree(keys_tmp_);
    if (vals_tmp_) (void)hipFree(vals_tmp_);
}

void RadixSorter::reserve(int64_t n) {
    if (n > cap_) {
        if (hist_) SCOPI_HIP_CHECK(hipFree(hist_));
        if (keys_tmp_) SCOPI_HIP_CHECK(hipFree(keys_tmp_));
        if (vals_tmp_) SCOPI_HIP_CHECK(hipFree(vals_tmp_));
        cap_ = n + n / 4 + 1024;
        const int64_t blocks = (cap_ + kSortTile - 1) / kSortTile;
        SCOPI_HIP_CHECK(hipMalloc(&hist_, 2 * 256 * blocks * sizeof(int)));
        SCOPI_HIP_CHECK(hipMalloc(&keys_tmp_, cap_ * sizeof(uint32_t)));
        SCOPI_HIP_CHECK(hipMalloc(&vals_tmp_, cap_ * sizeof(int32_t)));
    }
}

void RadixSorter::sort(uint32_t* keys, int32_t* vals, int64_t n, int key_bits, hipStream_t s, Scanner& scan) {
    if (n <= 1) return;
    reserve(n);
    const int nblocks = (int)((n + kSortTile - 1) / kSortTile);
    int* hist = hist_;
    int* offs = hist_ + 256 * (int64_t)nblocks;
    uint32_t* kin = keys;
    int32_t* vin = vals;
    uint32_t* kout = keys_tmp_;
    int32_t* vout = vals_tmp_;
    int passes = 0;
    for (int shift = 0; shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nblocks), dim3(kBlock), 0, s, kin, n, shift, hist, nblocks);
        scan.exclusive(hist, offs, 256 * (int64_t)nblocks, nullptr, s);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nblocks), dim3(kBlock), 0, s, kin, vin, kout, vout, n, shift, offs,
                           nblocks);
        std::swap(kin, kout);
        std::swap(vin, vout);
        ++passes;
    }
    if (passes & 1) {
        SCOPI_HIP_CHECK(hipMemcpyAsync(keys, kin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
        SCOPI_HIP_CHECK(hipMemcpyAsync(vals, vin, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    }
}

}  // namespace scopi_hip
