// contacts.hip -- device contact search for the contact slot.
//
// Replaces contact_kdtree::run_impl (include/scopi/contact/contact_kdtree.hpp:200-268)
// and contact_brute_force::run_impl (contact_brute_force.hpp:102-137).  The pair set is
// the reference's exactly:
//   active pairs  : i < j, both active, |c_i - c_j|^2 < kd_tree_radius (kd-tree only,
//                   nanoflann L2 adaptor, strict <) and dij < dmax (strict);
//   obstacle pairs: every (obstacle o, active j) with dij < dmax;
//   order         : sorted by (i, j) (contact/base.hpp:156-175).
// Candidates come from a uniform grid whose cell width is at least the
// candidate cut-off, so every pair that can satisfy the predicates lies in the
// 3^dim neighbouring cells; the predicates themselves are evaluated with the
// reference's arithmetic (common.hpp).  Cells are ordered by Morton code and
// particles are bucketed by a stable radix sort of those codes.  Each active
// particle i counts, then emits, its partners j > i into its own segment of
// the output, which it sorts by j, so the result is deterministic and needs no
// global sort.
#include "contacts.hpp"

#include <algorithm>
#include <cmath>

namespace scopi_hip {

constexpr int kMaxBits3 = 7;   // 128 cells per axis in 3D -> 2^21 cells
constexpr int kMaxBits2 = 10;  // 1024 cells per axis in 2D -> 2^20 cells
constexpr int64_t kMaxCells = int64_t(1) << 21;

__device__ inline uint32_t spread3(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ inline uint32_t spread2(uint32_t v) {
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}
template <int DIM>
__device__ inline uint32_t morton(int cx, int cy, int cz) {
    if (DIM == 3) return spread3((uint32_t)cx) | (spread3((uint32_t)cy) << 1) | (spread3((uint32_t)cz) << 2);
    return spread2((uint32_t)cx) | (spread2((uint32_t)cy) << 1);
}

template <int DIM>
__device__ inline int cell_coord(const GridInfo& g, double x, int d) {
    int c = (int)floor((x - g.lo[d]) / g.w[d]);
    return c < 0 ? 0 : (c >= g.nc[d] ? g.nc[d] - 1 : c);
}

// Bounding radius of an active particle (sphere radius, largest SE semi-axis).
__device__ inline double bounding_radius(const Particles& P, int64_t k) {
    const int s = P.shape[k];
    if (s == SCOPI_SHAPE_SUPERELLIPSOID) {
        double a = P.size[k], b = P.size[P.cap + k], c = P.size[2 * P.cap + k];
        return sqrt(a * a + b * b + c * c);
    }
    return P.size[k];
}

// ---------------------------------------------------------------- grid set-up
// partial[b*7 + {0..2}] = min, [3..5] = max, [6] = max bounding radius
__global__ void __launch_bounds__(kBlock) k_bbox_partial(Particles P, double* __restrict__ partial) {
    __shared__ double red[7][kBlock];
    const int64_t na = P.n - P.n_inactive;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, rm = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < na; t += (int64_t)gridDim.x * kBlock) {
        const int64_t k = P.n_inactive + t;
        for (int d = 0; d < 3; ++d) {
            double x = P.pos[d * P.cap + k];
            lo[d] = fmin(lo[d], x);
            hi[d] = fmax(hi[d], x);
        }
        rm = fmax(rm, bounding_radius(P, k));
    }
    for (int d = 0; d < 3; ++d) {
        red[d][threadIdx.x] = lo[d];
        red[3 + d][threadIdx.x] = hi[d];
    }
    red[6][threadIdx.x] = rm;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            for (int d = 0; d < 3; ++d) red[d][threadIdx.x] = fmin(red[d][threadIdx.x], red[d][threadIdx.x + o]);
            for (int d = 3; d < 7; ++d) red[d][threadIdx.x] = fmax(red[d][threadIdx.x], red[d][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 7) partial[blockIdx.x * 7 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(kBlock) k_grid_setup(const double* __restrict__ partial, int nparts, int dim, int kd,
                                                      double kd_radius, double dmax, int shapes_closed_form,
                                                      GridInfo* __restrict__ g) {
    __shared__ double red[7][kBlock];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, rm = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kBlock) {
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], partial[b * 7 + d]);
            hi[d] = fmax(hi[d], partial[b * 7 + 3 + d]);
        }
        rm = fmax(rm, partial[b * 7 + 6]);
    }
    for (int d = 0; d < 3; ++d) {
        red[d][threadIdx.x] = lo[d];
        red[3 + d][threadIdx.x] = hi[d];
    }
    red[6][threadIdx.x] = rm;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            for (int d = 0; d < 3; ++d) red[d][threadIdx.x] = fmin(red[d][threadIdx.x], red[d][threadIdx.x + o]);
            for (int d = 3; d < 7; ++d) red[d][threadIdx.x] = fmax(red[d][threadIdx.x], red[d][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    for (int d = 0; d < 3; ++d) {
        lo[d] = red[d][0];
        hi[d] = red[3 + d][0];
    }
    rm = red[6][0];
    // Any pair with dij < dmax has |c_i - c_j| < r_i + r_j + dmax for closed-form
    // shapes; the margin covers the roundings of the computed dij.
    double geo = (2.0 * rm + dmax) * (1.0 + 1e-9) + 1e-12;
    double cut = geo;
    if (kd) {
        double kdc = sqrt(kd_radius) * (1.0 + 1e-12) + 1e-300;
        cut = shapes_closed_form ? fmin(geo, kdc) : kdc;
    } else if (!shapes_closed_form) {
        cut = INFINITY;  // brute force over shapes without a distance bound: every pair
    }
    if (!(cut > 0)) cut = 1e-300;
    if (cut == INFINITY) {
        for (int d = 0; d < 3; ++d) {
            g->lo[d] = (d < dim) ? lo[d] : 0.0;
            g->nc[d] = 1;
            g->w[d] = INFINITY;
        }
        g->bits = 0;
        g->cutoff = cut;
        g->rmax = rm;
        return;
    }
    const int bits_max = (dim == 3) ? kMaxBits3 : kMaxBits2;
    int maxc = 1;
    for (int d = 0; d < 3; ++d) {
        g->lo[d] = (d < dim) ? lo[d] : 0.0;
        double ext = (d < dim) ? fmax(hi[d] - lo[d], 0.0) : 0.0;
        double cells = floor(ext / cut) + 1.0;
        int nc = (int)fmin(cells, (double)(1 << bits_max));
        if (d >= dim) nc = 1;
        g->nc[d] = nc;
        double w = cut;
        if (nc == (1 << bits_max)) w = fmax(cut, ext / (double)nc * (1.0 + 1e-12));
        g->w[d] = w;
        maxc = nc > maxc ? nc : maxc;
    }
    int bits = 0;
    while ((1 << bits) < maxc) ++bits;
    g->bits = bits;
    g->cutoff = cut;
    g->rmax = rm;
}

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_cell_keys(Particles P, const GridInfo* __restrict__ gp,
                                                     uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
    const int64_t na = P.n - P.n_inactive;
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= na) return;
    const GridInfo g = *gp;
    const int64_t k = P.n_inactive + t;
    int cx = cell_coord<DIM>(g, P.pos[k], 0);
    int cy = cell_coord<DIM>(g, P.pos[P.cap + k], 1);
    int cz = (DIM == 3) ? cell_coord<DIM>(g, P.pos[2 * P.cap + k], 2) : 0;
    keys[t] = morton<DIM>(cx, cy, cz);
    vals[t] = (int32_t)k;
}

__global__ void __launch_bounds__(kBlock) k_fill_int(int* __restrict__ a, int64_t n, int v) {
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlock) a[t] = v;
}

__global__ void __launch_bounds__(kBlock) k_cell_range(const uint32_t* __restrict__ keys, int64_t na,
                                                      int* __restrict__ start, int* __restrict__ end) {
    const int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (s >= na) return;
    const uint32_t k = keys[s];
    if (s == 0 || keys[s - 1] != k) start[k] = (int)s;
    if (s == na - 1 || keys[s + 1] != k) end[k] = (int)(s + 1);
}

// Visits every j > i of one cell that satisfies both predicates (CAND: the kd-tree predicate only;
// the closest points come later).
template <int DIM, bool KD, bool CAND, class F>
__device__ inline void visit_cell(const Particles& P, const int* __restrict__ start, const int* __restrict__ end,
                                  const int32_t* __restrict__ sorted, int64_t i, const d3& ci, uint32_t key,
                                  double kd_radius, double dmax, F& visit) {
    const int s1 = end[key];
    for (int s = start[key]; s < s1; ++s) {
        const int64_t j = sorted[s];
        if (j <= i) continue;
        if (KD) {
            const d3 cj = load3(P.pos, P.cap, j);
            if (!(kd_dist2<DIM>(ci, cj) < kd_radius)) continue;
        }
        if (CAND) {
            visit(j);
            continue;
        }
        CP cp;
        if (!closest_points<DIM>(P, i, j, cp)) continue;
        if (cp.dij < dmax) visit(j);
    }
}

// Visits every j > i in the neighbouring cells that satisfies both predicates.
template <int DIM, bool KD, bool CAND, class F>
__device__ inline void for_partners(const Particles& P, const GridInfo& g, const int* __restrict__ start,
                                    const int* __restrict__ end, const int32_t* __restrict__ sorted, int64_t i,
                                    double kd_radius, double dmax, F&& visit) {
    const d3 ci = load3(P.pos, P.cap, i);
    const int cx = cell_coord<DIM>(g, ci.x, 0), cy = cell_coord<DIM>(g, ci.y, 1);
    const int cz = (DIM == 3) ? cell_coord<DIM>(g, ci.z, 2) : 0;
    const int zr = (DIM == 3) ? 1 : 0;
    for (int dz = -zr; dz <= zr; ++dz) {
        const int z = cz + dz;
        if (z < 0 || z >= g.nc[2]) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = cy + dy;
            if (y < 0 || y >= g.nc[1]) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = cx + dx;
                if (x < 0 || x >= g.nc[0]) continue;
                visit_cell<DIM, KD, CAND>(P, start, end, sorted, i, ci, morton<DIM>(x, y, z), kd_radius, dmax, visit);
            }
        }
    }
}

// The part `part` (0..8) of for_partners' neighbourhood: in 3D the column of 3 cells along x at
// (dz, dy) = (part / 3 - 1, part % 3 - 1), in 2D the cell (dy, dx) = (part / 3 - 1, part % 3 - 1).
template <int DIM, bool KD, class F>
__device__ inline void for_partners_part(const Particles& P, const GridInfo& g, const int* __restrict__ start,
                                         const int* __restrict__ end, const int32_t* __restrict__ sorted, int64_t i,
                                         int part, double kd_radius, double dmax, F&& visit) {
    const d3 ci = load3(P.pos, P.cap, i);
    const int cx = cell_coord<DIM>(g, ci.x, 0), cy = cell_coord<DIM>(g, ci.y, 1);
    const int cz = (DIM == 3) ? cell_coord<DIM>(g, ci.z, 2) : 0;
    const int a = part / 3 - 1, b = part % 3 - 1;
    const int z = (DIM == 3) ? cz + a : 0;
    const int y = (DIM == 3) ? cy + b : cy + a;
    if (z < 0 || z >= g.nc[2] || y < 0 || y >= g.nc[1]) return;
    const int x0 = (DIM == 3) ? cx - 1 : cx + b, x1 = (DIM == 3) ? cx + 1 : cx + b;
    for (int x = x0; x <= x1; ++x) {
        if (x < 0 || x >= g.nc[0]) continue;
        visit_cell<DIM, KD, false>(P, start, end, sorted, i, ci, morton<DIM>(x, y, z), kd_radius, dmax, visit);
    }
}

template <int DIM, bool KD, bool CAND = false>
__global__ void __launch_bounds__(kBlock) k_pair_count(Particles P, const GridInfo* __restrict__ gp,
                                                      const int* __restrict__ start, const int* __restrict__ end,
                                                      const int32_t* __restrict__ sorted, double kd_radius, double dmax,
                                                      int* __restrict__ counts) {
    const int64_t na = P.n - P.n_inactive;
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= na) return;
    const GridInfo g = *gp;
    int cnt = 0;
    for_partners<DIM, KD, CAND>(P, g, start, end, sorted, P.n_inactive + t, kd_radius, dmax, [&](int64_t) { ++cnt; });
    counts[t] = cnt;
}

template <int DIM, bool KD, bool CAND = false>
__global__ void __launch_bounds__(kBlock) k_pair_emit(Particles P, const GridInfo* __restrict__ gp,
                                                     const int* __restrict__ start, const int* __restrict__ end,
                                                     const int32_t* __restrict__ sorted, double kd_radius, double dmax,
                                                     const int* __restrict__ offsets, int32_t* __restrict__ pair_j) {
    const int64_t na = P.n - P.n_inactive;
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= na) return;
    const GridInfo g = *gp;
    const int64_t base = offsets[t];
    int64_t cnt = 0;
    for_partners<DIM, KD, CAND>(P, g, start, end, sorted, P.n_inactive + t, kd_radius, dmax,
                                [&](int64_t j) { pair_j[base + cnt++] = (int32_t)j; });
    // insertion sort of this particle's partners by j
    for (int64_t a = 1; a < cnt; ++a) {
        const int32_t key = pair_j[base + a];
        int64_t b = a - 1;
        while (b >= 0 && pair_j[base + b] > key) {
            pair_j[base + b + 1] = pair_j[base + b];
            --b;
        }
        pair_j[base + b + 1] = key;
    }
}

// The count pass of the one-scene search also keeps up to kKeepPartners partners of each particle, so
// that the emit needs no second search: it copies them and sorts them by j, which gives the same pairs
// as k_pair_emit (the partners are distinct, so their order before the sort does not matter).  A
// particle with more partners sets `overflow`, and the search then emits with k_pair_emit.
// kCountGroup lanes share a particle, one part of its neighbourhood each (for_partners_part), and
// take the slots of its row from a counter in LDS.
constexpr int kKeepPartners = 16;
constexpr int kCountGroup = 9;
constexpr int kCountPerBlock = kBlock / kCountGroup;  // particles per block; the last lanes idle
template <int DIM, bool KD>
__global__ void __launch_bounds__(kBlock) k_pair_count_keep(Particles P, const GridInfo* __restrict__ gp,
                                                           const int* __restrict__ start, const int* __restrict__ end,
                                                           const int32_t* __restrict__ sorted, double kd_radius,
                                                           double dmax, int* __restrict__ counts,
                                                           int32_t* __restrict__ keep, int* __restrict__ overflow) {
    __shared__ int found[kCountPerBlock];
    const int64_t na = P.n - P.n_inactive;
    const int grp = threadIdx.x / kCountGroup, part = threadIdx.x % kCountGroup;
    const int64_t t = (int64_t)blockIdx.x * kCountPerBlock + grp;
    const bool active = grp < kCountPerBlock && t < na;
    if (threadIdx.x < kCountPerBlock) found[threadIdx.x] = 0;
    __syncthreads();
    if (active) {
        const GridInfo g = *gp;
        int32_t* kp = keep + t * kKeepPartners;
        for_partners_part<DIM, KD>(P, g, start, end, sorted, P.n_inactive + t, part, kd_radius, dmax, [&](int64_t j) {
            const int slot = atomicAdd(&found[grp], 1);
            if (slot < kKeepPartners) kp[slot] = (int32_t)j;
        });
    }
    __syncthreads();
    if (active && part == 0) {
        const int cnt = found[grp];
        counts[t] = cnt;
        if (cnt > kKeepPartners) atomicOr(overflow, 1);
    }
}

__global__ void __launch_bounds__(kBlock) k_pair_emit_kept(int64_t na, const int* __restrict__ counts,
                                                          const int* __restrict__ offsets,
                                                          const int32_t* __restrict__ keep, int32_t* __restrict__ pair_j) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= na) return;
    // counts[t] <= kKeepPartners when no particle overflowed, the only case this kernel runs in
    const int64_t cnt = counts[t] < kKeepPartners ? counts[t] : kKeepPartners;
    const int64_t base = offsets[t];
    const int32_t* kp = keep + t * kKeepPartners;
    for (int64_t m = 0; m < cnt; ++m) pair_j[base + m] = kp[m];
    for (int64_t a = 1; a < cnt; ++a) {  // insertion sort by j, as k_pair_emit
        const int32_t key = pair_j[base + a];
        int64_t b = a - 1;
        while (b >= 0 && pair_j[base + b] > key) {
            pair_j[base + b + 1] = pair_j[base + b];
            --b;
        }
        pair_j[base + b + 1] = key;
    }
}

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_obstacle_flags(Particles P, double dmax, int* __restrict__ flags) {
    const int64_t na = P.n - P.n_inactive;
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= na * P.n_inactive) return;
    const int64_t o = t / na, j = P.n_inactive + (t - o * na);
    CP cp;
    flags[t] = (closest_points<DIM>(P, o, j, cp) && cp.dij < dmax) ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) k_obstacle_emit(int64_t n_inactive, int64_t na, const int* __restrict__ flags,
                                                         const int* __restrict__ offs, Contacts C) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= na * n_inactive || !flags[t]) return;
    const int64_t o = t / na;
    const int64_t pos = offs[t];
    C.i[pos] = (int32_t)o;
    C.j[pos] = (int32_t)(n_inactive + (t - o * na));
}

__global__ void __launch_bounds__(kBlock) k_pair_scatter_i(int64_t n_inactive, int64_t na, int64_t nc_obs,
                                                          const int* __restrict__ counts,
                                                          const int* __restrict__ offsets,
                                                          const int32_t* __restrict__ pair_j, Contacts C) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= na) return;
    const int64_t b = offsets[t], cnt = counts[t];
    for (int64_t k = 0; k < cnt; ++k) {
        C.i[nc_obs + b + k] = (int32_t)(n_inactive + t);
        C.j[nc_obs + b + k] = pair_j[b + k];
    }
}

// compute_exact_distance (contact/base.hpp:103-141): geometry of every emitted pair.
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_contact_geometry(Particles P, Contacts C, int64_t nc, double mu) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    CP cp;
    closest_points<DIM>(P, C.i[c], C.j[c], cp);
    C.dij[c] = cp.dij;
    store3(C.nij, C.cap, c, cp.nij);
    store3(C.pi, C.cap, c, cp.pi);
    store3(C.pj, C.cap, c, cp.pj);
    C.mu[c] = mu;
}

// candidate (i, j) of the superellipsoid path: active pairs after the obstacle pairs
__global__ void __launch_bounds__(kBlock) k_cand_scatter(int64_t n_inactive, int64_t na, const int* __restrict__ counts,
                                                        const int* __restrict__ offsets,
                                                        const int32_t* __restrict__ pair_j, int64_t base,
                                                        SeCandidates S) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= na) return;
    const int64_t b = offsets[t], cnt = counts[t];
    for (int64_t k = 0; k < cnt; ++k) {
        S.i[base + b + k] = (int32_t)(n_inactive + t);
        S.j[base + b + k] = pair_j[b + k];
    }
}

// every (obstacle o, active j) pair is a candidate (contact_kdtree.hpp:249-256 brute-forces them)
__global__ void __launch_bounds__(kBlock) k_cand_obstacles(int64_t n_inactive, int64_t na, SeCandidates S) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n_inactive * na) return;
    const int64_t o = k / na;
    S.i[k] = (int32_t)o;
    S.j[k] = (int32_t)(n_inactive + (k - o * na));
}

__global__ void __launch_bounds__(kBlock) k_cand_flags(SeCandidates S, int64_t n, double dmax, int* __restrict__ flags) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n) return;
    flags[c] = (S.dij[c] < dmax) ? 1 : 0;
}

// compute_exact_distance's dmax filter (contact/base.hpp:118): keeps the candidates in order
__global__ void __launch_bounds__(kBlock) k_cand_emit(SeCandidates S, int64_t n, const int* __restrict__ flags,
                                                     const int* __restrict__ offs, int64_t base, Contacts C,
                                                     double mu) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= n || !flags[c]) return;
    const int64_t o = base + offs[c];
    C.i[o] = S.i[c];
    C.j[o] = S.j[c];
    C.dij[o] = S.dij[c];
    for (int d = 0; d < 3; ++d) {
        C.nij[d * C.cap + o] = S.nij[d * S.cap + c];
        C.pi[d * C.cap + o] = S.pi[d * S.cap + c];
        C.pj[d * C.cap + o] = S.pj[d * S.cap + c];
    }
    C.mu[o] = mu;
}

// ---------------------------------------------------------------- batches of scenes
// One block per scene: bounding box of its active particles, then its grid (cell
// width >= the candidate cut-off, as k_grid_setup; dense cells, at most `cells`).
__global__ void __launch_bounds__(kBlock) k_scene_grid(Particles P, const SceneRanges* __restrict__ scenes, int kd,
                                                      double kd_radius, double dmax, int shapes_closed_form,
                                                      GridInfo* __restrict__ grids) {
    __shared__ double red[7][kBlock];
    const SceneRanges sc = scenes[blockIdx.x];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, rm = 0.0;
    for (int64_t k = sc.a_lo + threadIdx.x; k < sc.a_hi; k += kBlock) {
        for (int d = 0; d < 3; ++d) {
            const double x = P.pos[d * P.cap + k];
            lo[d] = fmin(lo[d], x);
            hi[d] = fmax(hi[d], x);
        }
        rm = fmax(rm, bounding_radius(P, k));
    }
    for (int d = 0; d < 3; ++d) {
        red[d][threadIdx.x] = lo[d];
        red[3 + d][threadIdx.x] = hi[d];
    }
    red[6][threadIdx.x] = rm;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            for (int d = 0; d < 3; ++d) red[d][threadIdx.x] = fmin(red[d][threadIdx.x], red[d][threadIdx.x + o]);
            for (int d = 3; d < 7; ++d) red[d][threadIdx.x] = fmax(red[d][threadIdx.x], red[d][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    GridInfo g;
    rm = red[6][0];
    double cut = (2.0 * rm + dmax) * (1.0 + 1e-9) + 1e-12;
    if (kd) {
        const double kdc = sqrt(kd_radius) * (1.0 + 1e-12) + 1e-300;
        cut = shapes_closed_form ? fmin(cut, kdc) : kdc;
    } else if (!shapes_closed_form) {
        cut = INFINITY;  // brute force over shapes without a distance bound: every pair
    }
    if (!(cut > 0)) cut = 1e-300;
    double ext[3];
    int64_t nc[3];
    for (int d = 0; d < 3; ++d) {
        g.lo[d] = sc.a_hi > sc.a_lo ? red[d][0] : 0.0;
        ext[d] = sc.a_hi > sc.a_lo ? fmax(red[3 + d][0] - red[d][0], 0.0) : 0.0;
        nc[d] = (cut == INFINITY) ? 1 : (int64_t)fmin(floor(ext[d] / cut) + 1.0, 1048576.0);
    }
    // at most `cells` cells: coarser cells along the longest axes (any width >= cut finds every pair)
    const int64_t budget = sc.cells > 1 ? sc.cells : 1;
    while (nc[0] * nc[1] * nc[2] > budget) {
        int a = 0;
        for (int d = 1; d < 3; ++d)
            if (nc[d] > nc[a]) a = d;
        nc[a] = (nc[a] + 1) / 2 < nc[a] ? (nc[a] + 1) / 2 : nc[a] - 1;
        if (nc[a] < 1) nc[a] = 1;
    }
    for (int d = 0; d < 3; ++d) {
        g.nc[d] = (int)nc[d];
        const bool coarse = (double)nc[d] < floor(ext[d] / cut) + 1.0;
        g.w[d] = (cut == INFINITY) ? INFINITY : (coarse ? fmax(cut, ext[d] / (double)nc[d] * (1.0 + 1e-12)) : cut);
    }
    g.bits = 0;
    g.cutoff = cut;
    g.rmax = rm;
    grids[blockIdx.x] = g;
}

__device__ inline int64_t dense_key(const GridInfo& g, int64_t cell0, int cx, int cy, int cz) {
    return cell0 + cx + (int64_t)g.nc[0] * (cy + (int64_t)g.nc[1] * cz);
}

__global__ void __launch_bounds__(kBlock) k_scene_keys(Particles P, const SceneRanges* __restrict__ scenes,
                                                      const GridInfo* __restrict__ grids,
                                                      const int32_t* __restrict__ scene_of, uint32_t* __restrict__ keys,
                                                      int32_t* __restrict__ vals) {
    const int64_t na = P.n - P.n_inactive;
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= na) return;
    const int sc = scene_of[t];
    const GridInfo g = grids[sc];
    const int64_t k = P.n_inactive + t;
    const int cx = cell_coord<3>(g, P.pos[k], 0), cy = cell_coord<3>(g, P.pos[P.cap + k], 1);
    const int cz = cell_coord<3>(g, P.pos[2 * P.cap + k], 2);
    keys[t] = (uint32_t)dense_key(g, scenes[sc].cell0, cx, cy, cz);
    vals[t] = (int32_t)k;
}

// for_partners on a scene's dense grid (same predicates and visiting rule)
template <bool KD, bool CAND, class F>
__device__ inline void for_partners_dense(const Particles& P, const GridInfo& g, int64_t cell0,
                                          const int* __restrict__ start, const int* __restrict__ end,
                                          const int32_t* __restrict__ sorted, int64_t i, double kd_radius, double dmax,
                                          F&& visit) {
    const d3 ci = load3(P.pos, P.cap, i);
    const int cx = cell_coord<3>(g, ci.x, 0), cy = cell_coord<3>(g, ci.y, 1), cz = cell_coord<3>(g, ci.z, 2);
    for (int dz = -1; dz <= 1; ++dz) {
        const int z = cz + dz;
        if (z < 0 || z >= g.nc[2]) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = cy + dy;
            if (y < 0 || y >= g.nc[1]) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = cx + dx;
                if (x < 0 || x >= g.nc[0]) continue;
                const int64_t key = dense_key(g, cell0, x, y, z);
                const int s1 = end[key];
                for (int s = start[key]; s < s1; ++s) {
                    const int64_t j = sorted[s];
                    if (j <= i) continue;
                    if (KD) {
                        const d3 cj = load3(P.pos, P.cap, j);
                        if (!(kd_dist2<3>(ci, cj) < kd_radius)) continue;
                    }
                    if (CAND) {
                        visit(j);
                        continue;
                    }
                    CP cp;
                    if (!closest_points<3>(P, i, j, cp)) continue;
                    if (cp.dij < dmax) visit(j);
                }
            }
        }
    }
}

__device__ inline int unit_scene(const int64_t* __restrict__ ulo, int K, int64_t u) {
    int a = 0, b = K - 1;  // last scene with ulo <= u
    while (a < b) {
        const int m = (a + b + 1) >> 1;
        if (ulo[m] <= u)
            a = m;
        else
            b = m - 1;
    }
    return a;
}

// Work units of every scene: obstacle x active pairs (flag), then one unit per
// active particle (its partners j > i).  EMIT writes the pairs at `offsets` into
// (out_i, out_j).  CAND: candidates of the superellipsoid path (every obstacle pair,
// the kd-tree predicate alone for active pairs; closest points come later).
template <bool KD, bool EMIT, bool CAND>
__global__ void __launch_bounds__(kBlock) k_scene_units(Particles P, const SceneRanges* __restrict__ scenes,
                                                       const int64_t* __restrict__ ulo, int K,
                                                       const GridInfo* __restrict__ grids,
                                                       const int* __restrict__ start, const int* __restrict__ end,
                                                       const int32_t* __restrict__ sorted, double kd_radius,
                                                       double dmax, int* __restrict__ counts,
                                                       const int* __restrict__ offsets, int32_t* __restrict__ out_i,
                                                       int32_t* __restrict__ out_j) {
    const int64_t u = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (u >= ulo[K]) return;
    const int sc = unit_scene(ulo, K, u);
    const SceneRanges r = scenes[sc];
    const int64_t na = r.a_hi - r.a_lo, no = (r.ob_hi - r.ob_lo) * na;
    const int64_t local = u - r.u_lo;
    if (local < no) {
        const int64_t o = r.ob_lo + local / na, j = r.a_lo + local % na;
        if (!EMIT) {
            CP cp;
            counts[u] = CAND ? 1 : ((closest_points<3>(P, o, j, cp) && cp.dij < dmax) ? 1 : 0);
        } else if (counts[u]) {
            const int64_t pos = offsets[u];
            out_i[pos] = (int32_t)o;
            out_j[pos] = (int32_t)j;
        }
        return;
    }
    const int64_t i = r.a_lo + (local - no);
    const GridInfo g = grids[sc];
    if (!EMIT) {
        int cnt = 0;
        if (na > 1)
            for_partners_dense<KD, CAND>(P, g, r.cell0, start, end, sorted, i, kd_radius, dmax,
                                         [&](int64_t) { ++cnt; });
        counts[u] = cnt;
        return;
    }
    const int64_t cnt = counts[u];
    if (cnt == 0) return;
    const int64_t base = offsets[u];
    int64_t m = 0;
    for_partners_dense<KD, CAND>(P, g, r.cell0, start, end, sorted, i, kd_radius, dmax, [&](int64_t j) {
        if (m < cnt) out_j[base + m] = (int32_t)j;  // the count pass found cnt partners
        ++m;
    });
    for (int64_t a = 1; a < cnt; ++a) {  // insertion sort by j
        const int32_t key = out_j[base + a];
        int64_t b = a - 1;
        while (b >= 0 && out_j[base + b] > key) {
            out_j[base + b + 1] = out_j[base + b];
            --b;
        }
        out_j[base + b + 1] = key;
    }
    for (int64_t a = 0; a < cnt; ++a) out_i[base + a] = (int32_t)i;
}

// range[k] (a position in an array of n entries scanned into offs) -> its scanned offset
__global__ void k_remap_ranges(int64_t* __restrict__ range, int n2, const int* __restrict__ offs, int64_t n,
                               const int* __restrict__ total) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n2) return;
    const int64_t x = range[k];
    range[k] = (n > 0) ? (x < n ? (int64_t)offs[x] : (int64_t)*total) : 0;
}

__global__ void k_scene_crange(const int64_t* __restrict__ ulo, int K, const int* __restrict__ offsets,
                               const int* __restrict__ total, int64_t* __restrict__ crange, int64_t* __restrict__ io) {
    const int sc = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t U = ulo[K];
    auto at = [&](int64_t u) -> int64_t { return u < U ? (int64_t)offsets[u] : (int64_t)*total; };
    if (sc == 0 && io) io[0] = (U > 0) ? *total : 0;
    if (sc >= K) return;
    const int64_t a = (U > 0) ? at(ulo[sc]) : 0, b = (U > 0) ? at(ulo[sc + 1]) : 0;
    crange[2 * sc] = a;
    crange[2 * sc + 1] = b;
    if (io) {
        io[1 + 2 * sc] = a;
        io[2 + 2 * sc] = b;
    }
}

// ---------------------------------------------------------------- host
void contacts_free(Contacts& C) {
    for (void* p : {(void*)C.i, (void*)C.j, (void*)C.dij, (void*)C.nij, (void*)C.pi, (void*)C.pj, (void*)C.mu})
        if (p) (void)hipFree(p);
    C = Contacts{};
}

void contacts_reserve(Contacts& C, int64_t n) {
    if (n <= C.cap) return;
    contacts_free(C);
    const int64_t cap = std::max<int64_t>(n + n / 4, 1024);
    C.cap = cap;
    SCOPI_HIP_CHECK(hipMalloc(&C.i, cap * sizeof(int32_t)));
    SCOPI_HIP_CHECK(hipMalloc(&C.j, cap * sizeof(int32_t)));
    SCOPI_HIP_CHECK(hipMalloc(&C.dij, cap * sizeof(double)));
    SCOPI_HIP_CHECK(hipMalloc(&C.nij, 3 * cap * sizeof(double)));
    SCOPI_HIP_CHECK(hipMalloc(&C.pi, 3 * cap * sizeof(double)));
    SCOPI_HIP_CHECK(hipMalloc(&C.pj, 3 * cap * sizeof(double)));
    SCOPI_HIP_CHECK(hipMalloc(&C.mu, cap * sizeof(double)));
}

ContactSearch::~ContactSearch() {
    for (void* p : {(void*)d_grid_, (void*)d_partial_, (void*)keys_, (void*)vals_, (void*)counts_, (void*)offsets_,
                    (void*)keep_, (void*)obst_flags_, (void*)obst_offsets_, (void*)cell_start_, (void*)cell_end_,
                    (void*)pair_j_, (void*)d_totals_, (void*)cand_flags_, (void*)cand_offsets_})
        if (p) (void)hipFree(p);
    if (h_totals_) (void)hipHostFree(h_totals_);
    se_grids_free(se_grids_);
    se_candidates_free(se_cand_);
}

void ContactSearch::reserve_particles(int64_t na, int64_t n_inactive) {
    if (!d_grid_) {
        SCOPI_HIP_CHECK(hipMalloc(&d_grid_, sizeof(GridInfo)));
        SCOPI_HIP_CHECK(hipMalloc(&d_totals_, 3 * sizeof(int)));  // pairs, obstacle pairs, kept-partner overflow
        SCOPI_HIP_CHECK(hipHostMalloc(&h_totals_, 3 * sizeof(int)));
        SCOPI_HIP_CHECK(hipMalloc(&cell_start_, kMaxCells * sizeof(int)));
        SCOPI_HIP_CHECK(hipMalloc(&cell_end_, kMaxCells * sizeof(int)));
        cells_cap_ = kMaxCells;
    }
    if (na + 1 > pcap_) {
        for (void* p : {(void*)keys_, (void*)vals_, (void*)counts_, (void*)offsets_, (void*)keep_})
            if (p) SCOPI_HIP_CHECK(hipFree(p));
        pcap_ = na + na / 4 + 1024;
        SCOPI_HIP_CHECK(hipMalloc(&keys_, pcap_ * sizeof(uint32_t)));
        SCOPI_HIP_CHECK(hipMalloc(&vals_, pcap_ * sizeof(int32_t)));
        SCOPI_HIP_CHECK(hipMalloc(&counts_, pcap_ * sizeof(int)));
        SCOPI_HIP_CHECK(hipMalloc(&offsets_, pcap_ * sizeof(int)));
        SCOPI_HIP_CHECK(hipMalloc(&keep_, pcap_ * kKeepPartners * sizeof(int32_t)));
        const int64_t nparts = std::min<int64_t>(grid_for(pcap_), 1024);
        if (nparts * 7 > partial_cap_) {
            if (d_partial_) SCOPI_HIP_CHECK(hipFree(d_partial_));
            partial_cap_ = nparts * 7;
            SCOPI_HIP_CHECK(hipMalloc(&d_partial_, partial_cap_ * sizeof(double)));
        }
    }
    const int64_t no = na * n_inactive + 1;
    if (no > ocap_) {
        if (obst_flags_) SCOPI_HIP_CHECK(hipFree(obst_flags_));
        if (obst_offsets_) SCOPI_HIP_CHECK(hipFree(obst_offsets_));
        ocap_ = no + no / 4 + 1024;
        SCOPI_HIP_CHECK(hipMalloc(&obst_flags_, ocap_ * sizeof(int)));
        SCOPI_HIP_CHECK(hipMalloc(&obst_offsets_, ocap_ * sizeof(int)));
    }
}

void ContactSearch::reserve_pairs(int64_t n) {
    if (n <= pair_cap_) return;
    if (pair_j_) SCOPI_HIP_CHECK(hipFree(pair_j_));
    pair_cap_ = n + n / 4 + 1024;
    SCOPI_HIP_CHECK(hipMalloc(&pair_j_, pair_cap_ * sizeof(int32_t)));
}

int64_t ContactSearch::run(const Particles& P, Contacts& C, const scopi_hip_params& prm, hipStream_t s,
                           bool closed_form, bool bounded) {
    if (!closed_form) return run_candidates(P, C, prm, s);
    const int64_t na = P.n - P.n_inactive, ni = P.n_inactive;
    const bool kd = prm.contact_method == SCOPI_CONTACT_KDTREE;
    reserve_particles(na, ni);
    int64_t nc_aa = 0;
    if (na > 1) {
        const unsigned nparts = (unsigned)std::min<int64_t>(grid_for(na), 1024);
        hipLaunchKernelGGL(k_bbox_partial, dim3(nparts), dim3(kBlock), 0, s, P, d_partial_);
        hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(kBlock), 0, s, d_partial_, (int)nparts, P.dim, kd ? 1 : 0,
                           prm.kd_tree_radius, prm.dmax, bounded ? 1 : 0, d_grid_);
        if (P.dim == 3)
            hipLaunchKernelGGL(k_cell_keys<3>, dim3(grid_for(na)), dim3(kBlock), 0, s, P, d_grid_, keys_, vals_);
        else
            hipLaunchKernelGGL(k_cell_keys<2>, dim3(grid_for(na)), dim3(kBlock), 0, s, P, d_grid_, keys_, vals_);
        const int key_bits = (P.dim == 3) ? 3 * kMaxBits3 : 2 * kMaxBits2;
        sorter_.sort(keys_, vals_, na, key_bits, s, scanner_);
        const int64_t ncells = (P.dim == 3) ? (int64_t(1) << (3 * kMaxBits3)) : (int64_t(1) << (2 * kMaxBits2));
        hipLaunchKernelGGL(k_fill_int, dim3(1024), dim3(kBlock), 0, s, cell_start_, ncells, 0);
        hipLaunchKernelGGL(k_fill_int, dim3(1024), dim3(kBlock), 0, s, cell_end_, ncells, 0);
        hipLaunchKernelGGL(k_cell_range, dim3(grid_for(na)), dim3(kBlock), 0, s, keys_, na, cell_start_, cell_end_);
#define SCOPI_PAIR_LAUNCH(KERNEL, ...)                                                                         \
    do {                                                                                                       \
        if (P.dim == 3 && kd)                                                                                  \
            hipLaunchKernelGGL((KERNEL<3, true>), dim3(grid_for(na)), dim3(kBlock), 0, s, __VA_ARGS__);       \
        else if (P.dim == 3)                                                                                   \
            hipLaunchKernelGGL((KERNEL<3, false>), dim3(grid_for(na)), dim3(kBlock), 0, s, __VA_ARGS__);      \
        else if (kd)                                                                                           \
            hipLaunchKernelGGL((KERNEL<2, true>), dim3(grid_for(na)), dim3(kBlock), 0, s, __VA_ARGS__);       \
        else                                                                                                   \
            hipLaunchKernelGGL((KERNEL<2, false>), dim3(grid_for(na)), dim3(kBlock), 0, s, __VA_ARGS__);      \
    } while (0)
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, d_totals_ + 2, 1, 0);
        {
            const dim3 gk(grid_for(na, kCountPerBlock));
            if (P.dim == 3 && kd)
                hipLaunchKernelGGL((k_pair_count_keep<3, true>), gk, dim3(kBlock), 0, s, P, d_grid_, cell_start_,
                                   cell_end_, vals_, prm.kd_tree_radius, prm.dmax, counts_, keep_, d_totals_ + 2);
            else if (P.dim == 3)
                hipLaunchKernelGGL((k_pair_count_keep<3, false>), gk, dim3(kBlock), 0, s, P, d_grid_, cell_start_,
                                   cell_end_, vals_, prm.kd_tree_radius, prm.dmax, counts_, keep_, d_totals_ + 2);
            else if (kd)
                hipLaunchKernelGGL((k_pair_count_keep<2, true>), gk, dim3(kBlock), 0, s, P, d_grid_, cell_start_,
                                   cell_end_, vals_, prm.kd_tree_radius, prm.dmax, counts_, keep_, d_totals_ + 2);
            else
                hipLaunchKernelGGL((k_pair_count_keep<2, false>), gk, dim3(kBlock), 0, s, P, d_grid_, cell_start_,
                                   cell_end_, vals_, prm.kd_tree_radius, prm.dmax, counts_, keep_, d_totals_ + 2);
        }
        scanner_.exclusive(counts_, offsets_, na, d_totals_, s);
    } else {
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, d_totals_, 1, 0);
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, d_totals_ + 2, 1, 0);
    }
    const int64_t no = na * ni;
    if (no > 0) {
        if (P.dim == 3)
            hipLaunchKernelGGL(k_obstacle_flags<3>, dim3(grid_for(no)), dim3(kBlock), 0, s, P, prm.dmax, obst_flags_);
        else
            hipLaunchKernelGGL(k_obstacle_flags<2>, dim3(grid_for(no)), dim3(kBlock), 0, s, P, prm.dmax, obst_flags_);
        scanner_.exclusive(obst_flags_, obst_offsets_, no, d_totals_ + 1, s);
    } else {
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, d_totals_ + 1, 1, 0);
    }
    SCOPI_HIP_CHECK(hipMemcpyAsync(h_totals_, d_totals_, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    nc_aa = (na > 1) ? h_totals_[0] : 0;
    const bool kept = h_totals_[2] == 0;  // every particle's partners are in keep_
    nc_obs_ = (no > 0) ? h_totals_[1] : 0;
    const int64_t nc = nc_aa + nc_obs_;
    contacts_reserve(C, nc);
    if (no > 0 && nc_obs_ > 0)
        hipLaunchKernelGGL(k_obstacle_emit, dim3(grid_for(no)), dim3(kBlock), 0, s, ni, na, obst_flags_,
                           obst_offsets_, C);
    if (nc_aa > 0) {
        reserve_pairs(nc_aa);
        if (kept)
            hipLaunchKernelGGL(k_pair_emit_kept, dim3(grid_for(na)), dim3(kBlock), 0, s, na, counts_, offsets_, keep_,
                               pair_j_);
        else
            SCOPI_PAIR_LAUNCH(k_pair_emit, P, d_grid_, cell_start_, cell_end_, vals_, prm.kd_tree_radius, prm.dmax,
                              offsets_, pair_j_);
        hipLaunchKernelGGL(k_pair_scatter_i, dim3(grid_for(na)), dim3(kBlock), 0, s, ni, na, nc_obs_, counts_,
                           offsets_, pair_j_, C);
    }
#undef SCOPI_PAIR_LAUNCH
    if (nc > 0) {
        if (P.dim == 3)
            hipLaunchKernelGGL(k_contact_geometry<3>, dim3(grid_for(nc)), dim3(kBlock), 0, s, P, C, nc, prm.mu);
        else
            hipLaunchKernelGGL(k_contact_geometry<2>, dim3(grid_for(nc)), dim3(kBlock), 0, s, P, C, nc, prm.mu);
    }
    SCOPI_HIP_CHECK(hipGetLastError());
    return nc;
}

// Scenes with superellipsoids (2D or 3D): candidates are every (obstacle, active) pair and the active
// pairs passing the kd-tree predicate (every pair for brute force); closest points of every
// candidate on the device, then the dmax filter in candidate order, which is (i, j) order.
// the candidate-mode pair kernels for the scene's dimension and contact method
#define SCOPI_CAND_LAUNCH(KERNEL, ...)                                                                         \
    do {                                                                                                       \
        if (P.dim == 3 && kd)                                                                                  \
            hipLaunchKernelGGL((KERNEL<3, true, true>), dim3(grid_for(na)), dim3(kBlock), 0, s, __VA_ARGS__); \
        else if (P.dim == 3)                                                                                   \
            hipLaunchKernelGGL((KERNEL<3, false, true>), dim3(grid_for(na)), dim3(kBlock), 0, s, __VA_ARGS__);\
        else if (kd)                                                                                           \
            hipLaunchKernelGGL((KERNEL<2, true, true>), dim3(grid_for(na)), dim3(kBlock), 0, s, __VA_ARGS__); \
        else                                                                                                   \
            hipLaunchKernelGGL((KERNEL<2, false, true>), dim3(grid_for(na)), dim3(kBlock), 0, s, __VA_ARGS__);\
    } while (0)
int64_t ContactSearch::run_candidates(const Particles& P, Contacts& C, const scopi_hip_params& prm, hipStream_t s) {
    const int64_t na = P.n - P.n_inactive, ni = P.n_inactive;
    const bool kd = prm.contact_method == SCOPI_CONTACT_KDTREE;
    nc_obs_ = 0;
    if (na < 1) return 0;
    reserve_particles(na, ni);
    int64_t nact = 0;
    if (na > 1) {
        const unsigned nparts = (unsigned)std::min<int64_t>(grid_for(na), 1024);
        hipLaunchKernelGGL(k_bbox_partial, dim3(nparts), dim3(kBlock), 0, s, P, d_partial_);
        hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(kBlock), 0, s, d_partial_, (int)nparts, P.dim, kd ? 1 : 0,
                           prm.kd_tree_radius, prm.dmax, 0, d_grid_);
        if (P.dim == 3)
            hipLaunchKernelGGL(k_cell_keys<3>, dim3(grid_for(na)), dim3(kBlock), 0, s, P, d_grid_, keys_, vals_);
        else
            hipLaunchKernelGGL(k_cell_keys<2>, dim3(grid_for(na)), dim3(kBlock), 0, s, P, d_grid_, keys_, vals_);
        const int key_bits = (P.dim == 3) ? 3 * kMaxBits3 : 2 * kMaxBits2;
        sorter_.sort(keys_, vals_, na, key_bits, s, scanner_);
        const int64_t ncells = int64_t(1) << key_bits;
        hipLaunchKernelGGL(k_fill_int, dim3(1024), dim3(kBlock), 0, s, cell_start_, ncells, 0);
        hipLaunchKernelGGL(k_fill_int, dim3(1024), dim3(kBlock), 0, s, cell_end_, ncells, 0);
        hipLaunchKernelGGL(k_cell_range, dim3(grid_for(na)), dim3(kBlock), 0, s, keys_, na, cell_start_, cell_end_);
        SCOPI_CAND_LAUNCH(k_pair_count, P, d_grid_, cell_start_, cell_end_, vals_, prm.kd_tree_radius, prm.dmax,
                          counts_);
        scanner_.exclusive(counts_, offsets_, na, d_totals_, s);
        SCOPI_HIP_CHECK(hipMemcpyAsync(h_totals_, d_totals_, sizeof(int), hipMemcpyDeviceToHost, s));
        SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        nact = h_totals_[0];
    }
    const int64_t nobs = ni * na;
    const int64_t ncand = nobs + nact;
    if (ncand == 0) return 0;
    reserve_pairs(std::max<int64_t>(nact, 1));
    se_candidates_reserve(se_cand_, ncand);
    if (ncand + 1 > cand_cap_) {
        if (cand_flags_) SCOPI_HIP_CHECK(hipFree(cand_flags_));
        if (cand_offsets_) SCOPI_HIP_CHECK(hipFree(cand_offsets_));
        cand_cap_ = ncand + ncand / 4 + 1024;
        SCOPI_HIP_CHECK(hipMalloc(&cand_flags_, cand_cap_ * sizeof(int)));
        SCOPI_HIP_CHECK(hipMalloc(&cand_offsets_, cand_cap_ * sizeof(int)));
    }
    if (nobs > 0)
        hipLaunchKernelGGL(k_cand_obstacles, dim3(grid_for(nobs)), dim3(kBlock), 0, s, ni, na, se_cand_);
    if (nact > 0) {
        SCOPI_CAND_LAUNCH(k_pair_emit, P, d_grid_, cell_start_, cell_end_, vals_, prm.kd_tree_radius, prm.dmax,
                          offsets_, pair_j_);
        hipLaunchKernelGGL(k_cand_scatter, dim3(grid_for(na)), dim3(kBlock), 0, s, ni, na, counts_, offsets_,
                           pair_j_, nobs, se_cand_);
    }
    se_closest_points(P, se_grids_, se_cand_, ncand, s);
    hipLaunchKernelGGL(k_cand_flags, dim3(grid_for(ncand)), dim3(kBlock), 0, s, se_cand_, ncand, prm.dmax,
                       cand_flags_);
    scanner_.exclusive(cand_flags_, cand_offsets_, ncand, d_totals_ + 1, s);
    SCOPI_HIP_CHECK(hipMemcpyAsync(h_totals_ + 1, d_totals_ + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    const int64_t nc = h_totals_[1];
    ncand_ = ncand;
    contacts_reserve(C, nc);
    if (nc > 0)
        hipLaunchKernelGGL(k_cand_emit, dim3(grid_for(ncand)), dim3(kBlock), 0, s, se_cand_, ncand, cand_flags_,
                           cand_offsets_, (int64_t)0, C, prm.mu);
    SCOPI_HIP_CHECK(hipGetLastError());
    return nc;
}
#undef SCOPI_CAND_LAUNCH

// ---------------------------------------------------------------- batches of scenes (host)
BatchSearch::~BatchSearch() {
    for (void* p : {(void*)d_scenes_, (void*)d_ulo_, (void*)d_scene_of_, (void*)d_grids_, (void*)d_crange_,
                    (void*)keys_, (void*)vals_, (void*)counts_, (void*)offsets_, (void*)cell_start_,
                    (void*)cell_end_, (void*)d_total_})
        if (p) (void)hipFree(p);
    if (h_io_) (void)hipHostFree(h_io_);
    for (void* p : {(void*)cand_flags_, (void*)cand_offsets_})
        if (p) (void)hipFree(p);
    se_grids_free(se_grids_);
    se_candidates_free(se_cand_);
}

void BatchSearch::set_scenes(const std::vector<SceneRanges>& scenes, int64_t n_inactive, int64_t na, hipStream_t s) {
    const int K = (int)scenes.size();
    K_ = K;
    na_ = na;
    ni_ = n_inactive;
    if (K > kcap_) {
        for (void* p : {(void*)d_scenes_, (void*)d_ulo_, (void*)d_grids_, (void*)d_crange_})
            if (p) SCOPI_HIP_CHECK(hipFree(p));
        if (h_io_) SCOPI_HIP_CHECK(hipHostFree(h_io_));
        kcap_ = K + K / 2 + 16;
        SCOPI_HIP_CHECK(hipMalloc(&d_scenes_, kcap_ * sizeof(SceneRanges)));
        SCOPI_HIP_CHECK(hipMalloc(&d_ulo_, (kcap_ + 1) * sizeof(int64_t)));
        SCOPI_HIP_CHECK(hipMalloc(&d_grids_, kcap_ * sizeof(GridInfo)));
        SCOPI_HIP_CHECK(hipMalloc(&d_crange_, 2 * kcap_ * sizeof(int64_t)));
        SCOPI_HIP_CHECK(hipHostMalloc(&h_io_, (2 * kcap_ + 1) * sizeof(int64_t)));
    }
    if (!d_total_) SCOPI_HIP_CHECK(hipMalloc(&d_total_, sizeof(int)));
    std::vector<int64_t> ulo(K + 1);
    std::vector<int32_t> scene_of(std::max<int64_t>(na, 1));
    int64_t cells = 0;
    for (int k = 0; k < K; ++k) {
        ulo[k] = scenes[k].u_lo;
        cells = std::max(cells, scenes[k].cell0 + scenes[k].cells);
        for (int64_t a = scenes[k].a_lo; a < scenes[k].a_hi; ++a) scene_of[a - n_inactive] = k;
    }
    units_ = K ? scenes[K - 1].u_lo + (scenes[K - 1].ob_hi - scenes[K - 1].ob_lo + 1) *
                                          (scenes[K - 1].a_hi - scenes[K - 1].a_lo)
               : 0;
    ulo[K] = units_;
    cells_ = cells;
    if (units_ >= INT32_MAX || cells_ >= (int64_t(1) << 31)) throw hip_error(hipErrorInvalidValue, "batch search: too many work units");
    if (na + 1 > pcap_) {
        for (void* p : {(void*)keys_, (void*)vals_, (void*)d_scene_of_})
            if (p) SCOPI_HIP_CHECK(hipFree(p));
        pcap_ = na + na / 4 + 1024;
        SCOPI_HIP_CHECK(hipMalloc(&keys_, pcap_ * sizeof(uint32_t)));
        SCOPI_HIP_CHECK(hipMalloc(&vals_, pcap_ * sizeof(int32_t)));
        SCOPI_HIP_CHECK(hipMalloc(&d_scene_of_, pcap_ * sizeof(int32_t)));
    }
    if (units_ + 1 > ucap_) {
        for (void* p : {(void*)counts_, (void*)offsets_})
            if (p) SCOPI_HIP_CHECK(hipFree(p));
        ucap_ = units_ + units_ / 4 + 1024;
        SCOPI_HIP_CHECK(hipMalloc(&counts_, ucap_ * sizeof(int)));
        SCOPI_HIP_CHECK(hipMalloc(&offsets_, ucap_ * sizeof(int)));
    }
    if (cells_ + 1 > ccap_) {
        for (void* p : {(void*)cell_start_, (void*)cell_end_})
            if (p) SCOPI_HIP_CHECK(hipFree(p));
        ccap_ = cells_ + cells_ / 4 + 1024;
        SCOPI_HIP_CHECK(hipMalloc(&cell_start_, ccap_ * sizeof(int)));
        SCOPI_HIP_CHECK(hipMalloc(&cell_end_, ccap_ * sizeof(int)));
    }
    if (K > 0) {
        SCOPI_HIP_CHECK(hipMemcpyAsync(d_scenes_, scenes.data(), K * sizeof(SceneRanges), hipMemcpyHostToDevice, s));
        SCOPI_HIP_CHECK(hipMemcpyAsync(d_ulo_, ulo.data(), (K + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    }
    if (na > 0)
        SCOPI_HIP_CHECK(hipMemcpyAsync(d_scene_of_, scene_of.data(), na * sizeof(int32_t), hipMemcpyHostToDevice, s));
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    h_crange_.assign(2 * K, 0);
}

int64_t BatchSearch::run(const Particles& P, Contacts& C, const scopi_hip_params& prm, hipStream_t s,
                         bool closed_form) {
    const int K = K_;
    if (K == 0) return 0;
    const bool kd = prm.contact_method == SCOPI_CONTACT_KDTREE;
    const int64_t na = na_;
    if (na > 0) {
        hipLaunchKernelGGL(k_scene_grid, dim3(K), dim3(kBlock), 0, s, P, (const SceneRanges*)d_scenes_, kd ? 1 : 0,
                           prm.kd_tree_radius, prm.dmax, closed_form ? 1 : 0, d_grids_);
        hipLaunchKernelGGL(k_scene_keys, dim3(grid_for(na)), dim3(kBlock), 0, s, P, (const SceneRanges*)d_scenes_,
                           (const GridInfo*)d_grids_, (const int32_t*)d_scene_of_, keys_, vals_);
        int bits = 1;
        while ((int64_t(1) << bits) < cells_) ++bits;
        sorter_.sort(keys_, vals_, na, bits, s, scanner_);
        hipLaunchKernelGGL(k_fill_int, dim3(1024), dim3(kBlock), 0, s, cell_start_, cells_, 0);
        hipLaunchKernelGGL(k_fill_int, dim3(1024), dim3(kBlock), 0, s, cell_end_, cells_, 0);
        hipLaunchKernelGGL(k_cell_range, dim3(grid_for(na)), dim3(kBlock), 0, s, keys_, na, cell_start_, cell_end_);
    }
    const unsigned gu = grid_for(std::max<int64_t>(units_, 1));
    auto units = [&](bool emit, int32_t* oi, int32_t* oj) {
        auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(gu), dim3(kBlock), 0, s, P, (const SceneRanges*)d_scenes_,
                               (const int64_t*)d_ulo_, K, (const GridInfo*)d_grids_, (const int*)cell_start_,
                               (const int*)cell_end_, (const int32_t*)vals_, prm.kd_tree_radius, prm.dmax, counts_,
                               (const int*)offsets_, oi, oj);
        };
        const int sel = (kd ? 1 : 0) + (emit ? 2 : 0) + (closed_form ? 0 : 4);
        switch (sel) {
            case 0: go(k_scene_units<false, false, false>); break;
            case 1: go(k_scene_units<true, false, false>); break;
            case 2: go(k_scene_units<false, true, false>); break;
            case 3: go(k_scene_units<true, true, false>); break;
            case 4: go(k_scene_units<false, false, true>); break;
            case 5: go(k_scene_units<true, false, true>); break;
            case 6: go(k_scene_units<false, true, true>); break;
            default: go(k_scene_units<true, true, true>); break;
        }
    };
    if (units_ > 0) {
        units(false, nullptr, nullptr);
        scanner_.exclusive(counts_, offsets_, units_, d_total_, s);
    }
    // pair (or candidate) ranges of the scenes
    hipLaunchKernelGGL(k_scene_crange, dim3(grid_for(K, 64)), dim3(64), 0, s, (const int64_t*)d_ulo_, K,
                       (const int*)offsets_, (const int*)d_total_, d_crange_, (int64_t*)nullptr);
    SCOPI_HIP_CHECK(hipMemcpyAsync(h_io_ + 1, d_crange_, 2 * K * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    const int64_t npairs = (units_ > 0) ? h_io_[2 * K] : 0;
    if (closed_form) {
        for (int k = 0; k < 2 * K; ++k) h_crange_[k] = h_io_[1 + k];
        contacts_reserve(C, std::max<int64_t>(npairs, 1));
        if (npairs > 0) {
            units(true, C.i, C.j);
            hipLaunchKernelGGL(k_contact_geometry<3>, dim3(grid_for(npairs)), dim3(kBlock), 0, s, P, C, npairs,
                               prm.mu);
        }
        SCOPI_HIP_CHECK(hipGetLastError());
        return npairs;
    }
    // superellipsoid path: every candidate's closest points, then the dmax filter in candidate order
    const int64_t ncand = npairs;
    if (ncand > 0) {
        se_candidates_reserve(se_cand_, ncand);
        if (ncand + 1 > cand_cap_) {
            if (cand_flags_) SCOPI_HIP_CHECK(hipFree(cand_flags_));
            if (cand_offsets_) SCOPI_HIP_CHECK(hipFree(cand_offsets_));
            cand_cap_ = ncand + ncand / 4 + 1024;
            SCOPI_HIP_CHECK(hipMalloc(&cand_flags_, cand_cap_ * sizeof(int)));
            SCOPI_HIP_CHECK(hipMalloc(&cand_offsets_, cand_cap_ * sizeof(int)));
        }
        units(true, se_cand_.i, se_cand_.j);
        se_closest_points(P, se_grids_, se_cand_, ncand, s);
        hipLaunchKernelGGL(k_cand_flags, dim3(grid_for(ncand)), dim3(kBlock), 0, s, se_cand_, ncand, prm.dmax,
                           cand_flags_);
        scanner_.exclusive(cand_flags_, cand_offsets_, ncand, d_total_, s);
    }
    hipLaunchKernelGGL(k_remap_ranges, dim3(grid_for(2 * K, 64)), dim3(64), 0, s, d_crange_, 2 * K,
                       (const int*)cand_offsets_, ncand, (const int*)d_total_);
    SCOPI_HIP_CHECK(hipMemcpyAsync(h_io_ + 1, d_crange_, 2 * K * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    for (int k = 0; k < 2 * K; ++k) h_crange_[k] = h_io_[1 + k];
    const int64_t nc = h_crange_[2 * K - 1];
    contacts_reserve(C, std::max<int64_t>(nc, 1));
    if (nc > 0)
        hipLaunchKernelGGL(k_cand_emit, dim3(grid_for(ncand)), dim3(kBlock), 0, s, se_cand_, ncand, cand_flags_,
                           cand_offsets_, (int64_t)0, C, prm.mu);
    SCOPI_HIP_CHECK(hipGetLastError());
    return nc;
}

}  // namespace scopi_hip
