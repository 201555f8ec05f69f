// apgd.hip -- device optim slot: OptimGradient<apgd|pgd>::run
// (include/scopi/solvers/OptimGradient.hpp:252-306) for the NoFriction and
// Friction problems.
//
// The reference applies the matrix-free A / A^T (matrix/velocities.hpp:67-254)
// and recomputes two rotation matrices per contact side on every product.
// Here B is assembled once per step as a block-row sparse matrix (one block
// row per contact with two particle blocks; the column index is the particle,
// so no column array is stored; layouts in apgd.hpp), plus a particle-major
// copy of its blocks.  The solver keeps, for each of its three iterates
// lambda_s, the particle vector w_s = M^-1 B^T lambda_s, and a round of the
// loop is two launches:
//   k_trial : one thread per contact.  Q lambda_k = dt^2 B w_k and Q lambda_{k-1}
//             come from the row product on the fly; y, Q y = (1+beta) Q lambda_k -
//             beta Q lambda_{k-1} and the gradient follow by linearity, then the
//             projection, lambda+, and the scalars of the line search;
//   k_bt    : w+ = M^-1 B^T lambda+, one lane group per particle over its contiguous
//             entries of the particle-major copy (a gather: deterministic, no
//             atomics), and <lambda+, Q lambda+> = dt^2 (B^T lambda+) . w+ summed
//             over particles.
// Each trial costs one Q product instead of the reference's three (gradient + two
// objective evaluations, apgd.hpp:168-181), and no contact-space Q product is stored.
// Block sums are plain stores reduced by the next kernel: block 0 of k_bt sums the
// trial kernel's, and every block of the next round's k_trial sums k_bt's and
// runs the accept / backtrack / restart logic of apgd.hpp:172-209 on its own
// copy of the state (identical in every block), block 0 publishing it.  Kernel
// boundaries order every exchange, and no block ever issues a device-scope fence.
#include "apgd.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

// Multiply-adds are fused only within one source expression (the functions below that follow the
// CPU's rounding turn fusion off altogether).  Under the compiler's default, the backend fuses
// across statements within a basic block, so which products end in an FMA, and with them the
// rounding of the solve, would follow the code's block structure: a change of load order alone
// could move the iterates.
#pragma clang fp contract(on)

namespace scopi_hip {

constexpr int kRedBlocksMax = 2048;
constexpr int kPartialStride = 16;  // doubles of partial_trial_ per block of capacity
__host__ __device__ inline int64_t trial_pstride(const SceneMap& M) { return M.pstride ? M.pstride : kRedBlocksMax; }
// Blocks of the per-contact kernel for nc contacts.  A lane takes `per_iter` contacts per loop
// iteration (two in the rotation-free layout, one in the others) and k iterations, k = kmin (2 by
// default) unless `cap` blocks of kmin iterations cannot hold the contacts; the grid is then the
// fewest blocks that hold nc in k iterations, so that every block runs the same number of iterations
// (a block with one iteration more than the others would hold up the whole launch).  The same rule
// sizes a scene's share of a batch's grid, so a scene solved in a batch runs the one-scene
// arithmetic exactly (with the same kmin).
inline int64_t row_blocks(int64_t nc, int cap, int layout, int kmin = 2) {
    const int64_t per_iter = (layout == kRotFree ? 2 : 1) * (int64_t)kBlock;
    const int64_t k = std::max<int64_t>(kmin, (nc + per_iter * cap - 1) / (per_iter * cap));
    return std::min<int64_t>(std::max<int64_t>((nc + per_iter * k - 1) / (per_iter * k), 1), cap);
}
// Blocks of k_bt for na particles at G lanes per particle, at most kBtBlocksMax (grid-stride
// beyond): each block writes one block sum per round, which every block of the next k_trial reads.
// Its blocks are 1024 lanes wide, so that there are 4x fewer of those sums than with 256-lane
// blocks for the same lanes: every k_trial block loads them all, and with 2048 sums those loads
// on the same few cache lines, from every block, cost about 2 us per round.
constexpr int kBtBlock = 1024;
constexpr int kTransposeGroup = 8;  // lanes per particle of k_transpose
constexpr int kBtBlocksMax = 512;
constexpr int kCtrlR = kBtBlocksMax / kBlock;  // k_bt block sums per thread of a control step
inline unsigned bt_blocks(int64_t na, int G) {
    return std::min<unsigned>(grid_for(na * (int64_t)G, kBtBlock), (unsigned)kBtBlocksMax);
}
constexpr int kStatusStalled = 2;                // FP32 phase reached its precision floor
constexpr int kMaxConsecutiveRejections32 = 40;  // 2^40 growth of L without an accepted step
constexpr int64_t kMaxStagnation32 = 300;        // FP32 accepted iterations without halving the residual

// ------------------------------------------------------------------ reductions
// Sum of NV values over the block; the totals land in thread 0.
// Each level of the wavefront tree shuffles all NV values before adding any, so the NV dependency
// chains overlap (one cross-lane latency per level, not NV); every value is summed in the same
// order as one value alone.
template <int NV, int NTH = kBlock>
__device__ inline void block_sum(double (&v)[NV], double (*lds)[NTH / 64]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double x[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) x[k] = v[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double y[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) y[k] = __shfl_down(x[k], o, 64);
#pragma unroll
        for (int k = 0; k < NV; ++k) x[k] += y[k];
    }
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) lds[k][wave] = x[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int w = 0; w < NTH / 64; ++w) s += lds[k][w];
            v[k] = s;
        }
    }
    __syncthreads();
}

// The thread's share of a fixed-order sum: p[b * stride] for b = threadIdx.x, threadIdx.x + NTH,
// ... < n, accumulated in that order from 0.  The first R loads are issued together, so the sum
// waits for one memory round trip instead of R (the partials were written by blocks on other XCDs
// and come from the Infinity Cache or HBM); any further ones follow one at a time.  Adding 0.0 for
// the missing terms leaves every sum bitwise unchanged (a sum started from +0.0 is never -0.0).
template <int R, int NTH = kBlock>
__device__ inline double thread_sum(const double* __restrict__ p, int n, int stride) {
    double v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int b = (int)threadIdx.x + r * NTH;
        v[r] = b < n ? p[(int64_t)b * stride] : 0.0;
    }
    double t = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) t += v[r];
    for (int b = (int)threadIdx.x + R * NTH; b < n; b += NTH) t += p[(int64_t)b * stride];
    return t;
}

// ------------------------------------------------------------------ geometry helpers
__device__ inline d3 cross3(d3 x, d3 y) {
#pragma clang fp contract(off)
    return d3{x.y * y.z - x.z * y.y, x.z * y.x - x.x * y.z, x.x * y.y - x.y * y.x};
}
__device__ inline d3 mat_mult(const m3& a, d3 x) {
#pragma clang fp contract(off)
    return d3{a.a[0][0] * x.x + a.a[0][1] * x.y + a.a[0][2] * x.z, a.a[1][0] * x.x + a.a[1][1] * x.y + a.a[1][2] * x.z,
              a.a[2][0] * x.x + a.a[2][1] * x.y + a.a[2][2] * x.z};
}
__device__ inline d3 mat_mult_t(const m3& a, d3 x) {
#pragma clang fp contract(off)
    return d3{a.a[0][0] * x.x + a.a[1][0] * x.y + a.a[2][0] * x.z, a.a[0][1] * x.x + a.a[1][1] * x.y + a.a[2][1] * x.z,
              a.a[0][2] * x.x + a.a[1][2] * x.y + a.a[2][2] * x.z};
}

// ------------------------------------------------------------------ assembly
// M_inverse (solvers/minimization_problem.hpp:10-31), NW per particle.
template <class T, int LAY>
__global__ void __launch_bounds__(kBlock) k_minv(Particles P, T* __restrict__ minv) {
    constexpr int NW = lay<LAY>::NW;
    const int64_t na = P.n - P.n_inactive;
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= na) return;
    const int64_t k = P.n_inactive + p;
    const double im = 1. / P.m[k];
    for (int d = 0; d < 3; ++d) minv[p * NW + d] = (T)((d < P.dim) ? im : 0.0);
    if (NW == 6)
        for (int d = 0; d < 3; ++d) {
            double ij = 0.0;
            if (P.dim == 3 || d == 2) ij = 1. / P.J[d * P.cap + k];
            minv[p * NW + 3 + d] = (T)ij;
        }
}

// Coefficients of B and c = g2l(D n + dt A U) (minimization_problem.hpp:33-70,
// velocities.hpp:67-144, lagrange_multiplier.hpp:62-76), with U the a-priori velocities.
template <int DIM, int LAY>
__global__ void __launch_bounds__(kBlock) k_assemble(Particles P, Contacts C, int64_t nc, double dt,
                                                    BMatrix<double> B64, BMatrix<float> B32, int with32) {
#pragma clang fp contract(off)
    constexpr int NC = lay<LAY>::NC, NR = lay<LAY>::NR;
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    const int64_t pc = P.cap;
    const int64_t i = C.i[c], j = C.j[c];
    const bool ai = i >= P.n_inactive, aj = j >= P.n_inactive;
    const d3 n = load3(C.nij, C.cap, c);
    const d3 pi = load3(C.pi, C.cap, c), pj = load3(C.pj, C.cap, c);
    d3 au = mk(0, 0, 0);
    d3 ri = mk(0, 0, 0), rj = mk(0, 0, 0);
    m3 Ri{}, Rj{};
    auto side = [&](int64_t k, d3 p, d3& r, m3& R) -> d3 {
        d3 x = load3(P.pos, pc, k);
        r = mk(p.x - x.x, p.y - x.y, DIM == 3 ? p.z - x.z : 0.0);
        R = rotation3(P.q[k], P.q[pc + k], P.q[2 * pc + k], P.q[3 * pc + k]);
        d3 v = load3(P.v, pc, k);
        if (DIM == 2) v.z = 0.0;
        d3 om = load3(P.omega, pc, k);
        if (DIM == 2) om = mk(0.0, 0.0, om.z);
        d3 cr = cross3(r, mat_mult(R, om));
        return mk(v.x - cr.x, v.y - cr.y, v.z - cr.z);
    };
    if (ai) {
        d3 t = side(i, pi, ri, Ri);
        au = mk(au.x + t.x, au.y + t.y, au.z + t.z);
    }
    if (aj) {
        d3 t = side(j, pj, rj, Rj);
        au = mk(au.x - t.x, au.y - t.y, au.z - t.z);
    }
    const double dij = C.dij[c];
    B64.mu[c] = C.mu[c];
    if (with32) B32.mu[c] = (float)C.mu[c];
    const double Cv[3] = {dij * n.x + dt * au.x, dij * n.y + dt * au.y, (DIM == 3 ? dij * n.z : 0.0) + dt * au.z};
    double co[NC];
    double cr[NR];
    co[0] = n.x, co[1] = n.y, co[2] = n.z;
    if (LAY == kRigid) {
        d3 a_i = ai ? mat_mult_t(Ri, cross3(ri, n)) : mk(0, 0, 0);
        d3 a_j = aj ? mat_mult_t(Rj, cross3(rj, n)) : mk(0, 0, 0);
        co[3] = a_i.x, co[4] = a_i.y, co[5] = a_i.z;
        co[6] = a_j.x, co[7] = a_j.y, co[8] = a_j.z;
    }
    if (LAY == kCone || LAY == kVisc) {
        for (int d = 0; d < 3; ++d) {
            const d3 e = mk(d == 0, d == 1, d == 2);
            d3 ki = (ai && d < DIM) ? mat_mult_t(Ri, cross3(ri, e)) : mk(0, 0, 0);
            d3 kj = (aj && d < DIM) ? mat_mult_t(Rj, cross3(rj, e)) : mk(0, 0, 0);
            co[3 + 3 * d] = ki.x, co[4 + 3 * d] = ki.y, co[5 + 3 * d] = ki.z;
            co[12 + 3 * d] = kj.x, co[13 + 3 * d] = kj.y, co[14 + 3 * d] = kj.z;
        }
    }
    if (LAY == kCone) {
        for (int d = 0; d < 3; ++d) cr[d] = (d < DIM) ? Cv[d] : 0.0;
    } else if (LAY == kVisc) {
        // global2local of C (lagrange_multiplier.hpp:143-163, 403-434) for the contact's kind
        double cl = 0.0;
        cl += Cv[0] * n.x;
        cl += Cv[1] * n.y;
        if (DIM == 3) cl += Cv[2] * n.z;
        const int kd = B64.kind[c] & kKindMask;
        for (int d = 0; d < NR; ++d) cr[d] = 0.0;
        if (kd == 0) {
            cr[0] = cl;
        } else if (kd == 1) {
            cr[0] = cl;
            cr[1] = -cl;
        } else {
            cr[0] = -cl;
            for (int d = 0; d < DIM; ++d) cr[1 + d] = Cv[d];
        }
    } else {
        double cl = 0.0;
        cl += Cv[0] * n.x;
        cl += Cv[1] * n.y;
        if (DIM == 3) cl += Cv[2] * n.z;
        cr[0] = cl;
    }
    for (int k = 0; k < NC; ++k) B64.coef[c * NC + k] = co[k];
    for (int d = 0; d < NR; ++d) B64.c[c * NR + d] = cr[d];
    if (with32) {
        for (int k = 0; k < NC; ++k) B32.coef[c * NC + k] = (float)co[k];
        for (int d = 0; d < NR; ++d) B32.c[c * NR + d] = (float)cr[d];
    }
}

__global__ void __launch_bounds__(kBlock) k_fill_i(int* __restrict__ a, int64_t n, int v) {
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlock) a[t] = v;
}

// contacts whose i is active occupy one contiguous range per particle (sorted by (i, j))
__global__ void __launch_bounds__(kBlock) k_irange(const int32_t* __restrict__ ci, int64_t nc, int64_t ni,
                                                  int* __restrict__ start, int* __restrict__ end) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    const int32_t i = ci[c];
    if (i < ni) return;
    if (c == 0 || ci[c - 1] != i) start[i - ni] = (int)c;
    if (c == nc - 1 || ci[c + 1] != i) end[i - ni] = (int)(c + 1);
}

// j-side entries: key = j's active index; an obstacle j (periodic boxes) gets key na, past every particle
__global__ void __launch_bounds__(kBlock) k_jkeys(const int32_t* __restrict__ cj, int64_t nc, int64_t ni, int64_t na,
                                                 uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    keys[c] = (uint32_t)(cj[c] >= ni ? cj[c] - ni : na);
    vals[c] = (int32_t)c;
}

__global__ void __launch_bounds__(kBlock) k_jrange(const uint32_t* __restrict__ keys, int64_t nc,
                                                  int* __restrict__ start, int* __restrict__ end) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= nc) return;
    const uint32_t k = keys[e];
    if (e == 0 || keys[e - 1] != k) start[k] = (int)e;
    if (e == nc - 1 || keys[e + 1] != k) end[k] = (int)(e + 1);
}

__global__ void __launch_bounds__(kBlock) k_tcount(const int* __restrict__ is, const int* __restrict__ ie,
                                                  const int* __restrict__ js, const int* __restrict__ je, int64_t na,
                                                  int* __restrict__ cnt) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p > na) return;
    cnt[p] = (p < na) ? (ie[p] - is[p]) + (je[p] - js[p]) : 0;
}

// Particle-major copy of B's blocks: for particle p, entries [tstart[p], tstart[p+1]):
// its contacts as i (ascending c), then as j (ascending c), sign folded into the values.
template <class T, int LAY>
__global__ void __launch_bounds__(kBlock) k_transpose(BMatrix<T> B, int64_t na, const int* __restrict__ is,
                                                     const int* __restrict__ ie, const int* __restrict__ js,
                                                     const int* __restrict__ je, const int32_t* __restrict__ jlist,
                                                     const int* __restrict__ tstart, int32_t* __restrict__ tc) {
    // kTransposeGroup lanes per particle, lane k writing the particle's entries k, k + group, ...:
    // a lane's chain of dependent loads (jlist, then the contact's coefficients) is a few entries
    // long instead of the particle's whole list, and a wavefront's stores fall on neighbouring entries
    constexpr int NC = lay<LAY>::NC, NT = lay<LAY>::NT, G = kTransposeGroup;
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t p = g / G;
    const int sub = (int)(g % G);
    if (p >= na) return;
    const int i0 = is[p], ni = ie[p] - i0, j0 = js[p], nt = ni + (je[p] - j0);
    const int64_t e0 = tstart[p];
    for (int k = sub; k < nt; k += G) {
        const int64_t e = e0 + k;
        if (k < ni) {
            const int c = i0 + k;
            tc[e] = c;
            if (LAY == kRotFree)
                for (int q = 0; q < 3; ++q) B.tcoef[e * NT + q] = B.coef[(int64_t)c * NC + q];
            if (LAY == kRigid)
                for (int q = 0; q < 6; ++q) B.tcoef[e * NT + q] = B.coef[(int64_t)c * NC + q];
            if (LAY == kCone || LAY == kVisc)
                for (int q = 0; q < 9; ++q) B.tcoef[e * NT + q] = B.coef[(int64_t)c * NC + 3 + q];
        } else {
            const int c = jlist[j0 + (k - ni)];
            tc[e] = ~c;
            if (LAY == kRotFree)
                for (int q = 0; q < 3; ++q) B.tcoef[e * NT + q] = -B.coef[(int64_t)c * NC + q];
            if (LAY == kRigid) {
                for (int q = 0; q < 3; ++q) B.tcoef[e * NT + q] = -B.coef[(int64_t)c * NC + q];
                for (int q = 0; q < 3; ++q) B.tcoef[e * NT + 3 + q] = -B.coef[(int64_t)c * NC + 6 + q];
            }
            if (LAY == kCone || LAY == kVisc)
                for (int q = 0; q < 9; ++q) B.tcoef[e * NT + q] = -B.coef[(int64_t)c * NC + 12 + q];
        }
    }
}

// ------------------------------------------------------------------ the loop
template <class T>
__host__ __device__ inline T* slot(T* base, int s, int64_t stride) {
    return base + (int64_t)s * stride;
}

// kVisc: the global multiplier (3 rows) of contact c from its 4 local slots (local2global,
// lagrange_multiplier.hpp:165-184, 436-464)
template <class T>
__device__ inline void visc_global(const BMatrix<T>& B, const T* __restrict__ lam, int64_t c, T& g0, T& g1, T& g2) {
    constexpr int NC = lay<kVisc>::NC;
    const int kd = B.kind[c] & kKindMask;
    const T* x = lam + 4 * c;
    const T nx = B.coef[c * NC], ny = B.coef[c * NC + 1], nz = B.coef[c * NC + 2];
    if (kd == 2) {
        const T m = -x[0];
        g0 = m * nx + x[1];
        g1 = m * ny + x[2];
        g2 = m * nz + x[3];
    } else {
        const T a = (kd == 1) ? x[0] - x[1] : x[0];
        g0 = a * nx;
        g1 = a * ny;
        g2 = a * nz;
    }
}

// Coulomb-cone projection of f (lagrange_multiplier.hpp:243-268, 500-519)
template <class T>
__device__ inline void cone_project(T (&f)[3], T nx, T ny, T nz, T mu) {
    const T ln = f[0] * nx + f[1] * ny + f[2] * nz;
    T tx = f[0] - ln * nx, ty = f[1] - ln * ny, tz = f[2] - ln * nz;
    const T norm = sqrt(tx * tx + ty * ty + tz * tz);
    if (norm > mu * ln) {
        if (ln <= -mu * norm) {
            f[0] = f[1] = f[2] = (T)0;
        } else {
            const T nn = (mu * mu) * (norm + ln / mu) / (mu * mu + (T)1);
            const T nl = nn / mu;
            tx /= norm, ty /= norm, tz /= norm;
            f[0] = nn * tx + nl * nx;
            f[1] = nn * ty + nl * ny;
            f[2] = nn * tz + nl * nz;
        }
    }
}

// Control step on the block sums of k_bt (<lambda+, Q lambda+> / dt^2, `nbt` blocks); see below.
__device__ inline void control_step(const ApgdCtrl* __restrict__ in, ApgdCtrl* __restrict__ out,
                                    const double* __restrict__ partial_bt, int nbt, ApgdCtrl& st,
                                    double (*lds)[kBlock / 64], bool leader, int32_t* done_flag);
__device__ inline void control_load(const double* __restrict__ partial_bt, int nbt, double (&pv)[kCtrlR]);
__device__ inline void control_finish(const ApgdCtrl* __restrict__ in, ApgdCtrl* __restrict__ out,
                                      const double* __restrict__ partial_bt, int nbt, const double (&pv)[kCtrlR],
                                      ApgdCtrl& st, double (*lds)[kBlock / 64], bool leader, int32_t* done_flag);

// Logical index of block b among nb: the hardware deals workgroups to the 8 XCDs in turn (XCD
// b mod 8), so XCD x runs blocks x, x + 8, ...; they are given a contiguous run of logical blocks,
// so that in every sweep of a grid-stride loop an XCD works on neighbouring contacts (or
// particles), whose w and lambda gathers then hit its own L2 instead of fetching the same lines
// on every XCD.  Measured on 1.17M contacts: k_trial 19.5 -> 18.3 us.
__device__ inline int xcd_logical(int b, int nb) {
    const int x = b & 7, r = b >> 3, q = nb >> 3, rem = nb & 7;
    return x < rem ? x * (q + 1) + r : rem * (q + 1) + (x - rem) * q + r;
}

// The block's scene and its share of the per-contact grid.
struct SceneBlock {
    int sc, lb, nb;      // scene, block index within the scene, blocks of the scene
    int64_t lo, hi;      // the scene's contacts
    int64_t pbase;       // first block sum of the scene
};
__device__ inline SceneBlock scene_block(const SceneMap& M, int64_t nc) {
    SceneBlock b;
    if (!M.rblk) {
        b.sc = 0, b.lb = blockIdx.x, b.nb = gridDim.x, b.lo = 0, b.hi = nc, b.pbase = 0;
        return b;
    }
    b.sc = M.rblk[blockIdx.x];
    const int f = M.rfirst[b.sc];
    b.lb = blockIdx.x - f;
    b.nb = M.rcount[b.sc];
    b.lo = M.crange[2 * b.sc];
    b.hi = M.crange[2 * b.sc + 1];
    b.pbase = f;
    return b;
}

// B^T lambda over the entries [e0, e1) of one particle of the particle-major copy, lane `sub` of a
// group of G taking entries sub, sub + G, ... (acc holds the lane's share; the group sums the shares).
template <class T, int LAY, int G>
__device__ inline void bt_gather(const BMatrix<T>& B, const T* __restrict__ lam, const int32_t* __restrict__ tc,
                                 int e0, int e1, int sub, T (&acc)[lay<LAY>::NW]) {
    constexpr int NT = lay<LAY>::NT;
    // two entries per iteration, every load issued before the first use (the same
    // accumulation order as one entry at a time)
    for (int e = e0 + sub; e < e1; e += 2 * G) {
        const int eb = e + G;
        const bool hb = eb < e1;
        // the second index is read from a selected address (tc[e] again past the end), not
        // selected after the load, so both loads are in flight together
        const int32_t csa = tc[e];
        const int32_t csb = tc[hb ? eb : e];
        const int64_t ca = csa < 0 ? ~csa : csa, cb = csb < 0 ? ~csb : csb;
        const T* tva = B.tcoef + (int64_t)e * NT;
        const T* tvb = B.tcoef + (int64_t)(hb ? eb : e) * NT;
        T ta[NT], tb[NT];
#pragma unroll
        for (int k = 0; k < NT; ++k) ta[k] = tva[k], tb[k] = tvb[k];
        if (LAY == kRotFree || LAY == kRigid) {
            const T la = lam[ca], lb = lam[cb];
#pragma unroll
            for (int k = 0; k < NT; ++k) acc[k] += la * ta[k];
            // the second entry's term is selected rather than branched over, so that its loads are
            // not sunk behind the first entry's
#pragma unroll
            for (int k = 0; k < NT; ++k) {
                const T nb = acc[k] + lb * tb[k];
                acc[k] = hb ? nb : acc[k];
            }
        } else {
            T a0, a1, a2, b0, b1, b2;
            if (LAY == kVisc) {
                visc_global<T>(B, lam, ca, a0, a1, a2);
                visc_global<T>(B, lam, cb, b0, b1, b2);
            } else {
                a0 = lam[ca * 3], a1 = lam[ca * 3 + 1], a2 = lam[ca * 3 + 2];
                b0 = lam[cb * 3], b1 = lam[cb * 3 + 1], b2 = lam[cb * 3 + 2];
            }
            if (csa < 0) {
                acc[0] -= a0, acc[1] -= a1, acc[2] -= a2;
            } else {
                acc[0] += a0, acc[1] += a1, acc[2] += a2;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[3 + k] += a0 * ta[k] + a1 * ta[3 + k] + a2 * ta[6 + k];
            if (hb) {
                if (csb < 0) {
                    acc[0] -= b0, acc[1] -= b1, acc[2] -= b2;
                } else {
                    acc[0] += b0, acc[1] += b1, acc[2] += b2;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) acc[3 + k] += b0 * tb[k] + b1 * tb[3 + k] + b2 * tb[6 + k];
            }
        }
    }
}

// w_s = M^-1 B^T lambda_s (ATMatrix::mat_mult, velocities.hpp:167-254, then M^-1)
// into the iterate's slot s of w.  G consecutive lanes share one active particle:
// lane k of the group takes its entries k, k + G, ... of the particle-major copy (so
// a wavefront reads 64/G contiguous runs of G entries per load), then the group sums
// its partial results with lane shuffles.  In the loop (ctrl set), every block also
// writes its sum of a . w over its particles, a = B^T lambda: the objective's
// <lambda, Q lambda> = dt^2 a^T M^-1 a without a pass over the contacts; and block 0
// reduces the trial kernel's block sums into ctrl->red.
template <class T, int LAY, int G>
__global__ void __launch_bounds__(kBtBlock) k_bt(ApgdCtrl* __restrict__ ctrl, int use_slot, BMatrix<T> B, int64_t na,
                                              const int* __restrict__ tstart, const int32_t* __restrict__ tc,
                                              const double* __restrict__ partial_trial, int ntrial,
                                              double* __restrict__ partial_bt, SceneMap M) {
    constexpr int NR = lay<LAY>::NR, NW = lay<LAY>::NW;
    __shared__ double lds[1][kBtBlock / 64];
    __shared__ double lds2[2][kBtBlock / 64];
    const bool multi = M.bblk != nullptr;
    const int sc = multi ? M.bblk[blockIdx.x] : 0;
    const int lb = multi ? (int)blockIdx.x - M.bfirst[sc] : (int)blockIdx.x;
    const int64_t p0 = multi ? M.arange[2 * sc] : 0, p1 = multi ? M.arange[2 * sc + 1] : na;
    // grid-stride over the scene's particles in logical block order (XCD-aware for one scene),
    // kBtBlock / G particles per block and pass
    const int nbk = multi ? M.bcount[sc] : (int)gridDim.x;
    const int lblk = multi ? lb : xcd_logical(lb, nbk);
    const int64_t pstep = (int64_t)nbk * (kBtBlock / G);
    const int64_t pend = p1;
    int64_t p = p0 + ((int64_t)lblk * kBtBlock + threadIdx.x) / G;
    const int sub = (int)(threadIdx.x % G);
    // the state and the first particle's entry range are loaded together: neither waits for the other
    if (ctrl) ctrl += sc;
    const int done = ctrl ? ctrl->done : 0;
    const int s = (use_slot >= 0) ? use_slot : (ctrl ? ctrl->trial : 0);
    int e0 = p < pend ? tstart[p] : 0, e1 = p < pend ? tstart[p + 1] : 0;
    if (done) return;
    // Blocks 0-7 also reduce the trial kernel's 8 block sums, one quantity per block (blocks lb, lb + nbk,
    // ... < 8), each in block order; a quantity's sums are contiguous.  The first quantity's loads are
    // issued here, ahead of the gather (a missing term reads the quantity's entry 0 and is replaced by
    // 0.0 where it is summed), and summed at the end with the block's own sum, in the order of
    // thread_sum; so a reducing block waits for no extra memory round trip or barrier.
    constexpr int RT = kRedBlocksMax / kBtBlock;
    const bool red = ctrl && partial_trial && lb < 8;
    const double* pt = multi ? partial_trial + M.rfirst[sc] : partial_trial;
    const int nt = multi ? M.rcount[sc] : ntrial;
    double tv[RT];
    if (red) {
        const double* p = pt + lb * trial_pstride(M);
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const int b = (int)threadIdx.x + r * kBtBlock;
            tv[r] = p[b < nt ? b : 0];
        }
    }
    const T* lam = slot<T>(B.lam, s, NR * B.cap);
    T* w = slot<T>(B.w, s, 6 * B.pcap);
    double q[1] = {0.0};
    // a lane group shares p, so each group passes through the loop as a whole (the shuffles
    // below stay within it)
    for (; p < pend; p += pstep) {
        // the next pass's entry range is loaded now, ahead of this pass's gather (after the stores
        // of w it would wait for them), and the particle's M^-1 with its first entries: after the
        // stores, the compiler could not tell these loads apart from them and would wait for each
        const int f0 = e0, f1 = e1;
        {
            const int64_t pn = p + pstep < pend ? p + pstep : p;
            e0 = tstart[pn], e1 = tstart[pn + 1];
        }
        T mv[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) mv[k] = B.minv[p * NW + k];
        T acc[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) acc[k] = (T)0;
        bt_gather<T, LAY, G>(B, lam, tc, f0, f1, sub, acc);
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
            for (int k = 0; k < NW; ++k) acc[k] += __shfl_xor(acc[k], o, G);
        if (sub == 0)
#pragma unroll
            for (int k = 0; k < NW; ++k) {
                const T wk = mv[k] * acc[k];
                w[p * NW + k] = wk;
                q[0] += (double)acc[k] * (double)wk;
            }
    }
    if (red) {
        double t = 0.0;
#pragma unroll
        for (int r = 0; r < RT; ++r) t += ((int)threadIdx.x + r * kBtBlock < nt) ? tv[r] : 0.0;
        const double* ptk = pt + lb * trial_pstride(M);
        for (int b = (int)threadIdx.x + RT * kBtBlock; b < nt; b += kBtBlock) t += ptk[b];
        double v[2] = {q[0], t};
        block_sum<2, kBtBlock>(v, lds2);
        if (threadIdx.x == 0) {
            partial_bt[(multi ? M.bfirst[sc] : 0) + lblk] = v[0];
            ctrl->red[lb < 7 ? lb : 8] = v[1];
        }
        // further quantities when the scene has fewer than 8 blocks
        for (int k = lb + nbk; k < 8; k += nbk) {
            double tot[1] = {thread_sum<RT, kBtBlock>(pt + k * trial_pstride(M), nt, 1)};
            block_sum<1, kBtBlock>(tot, lds);
            if (threadIdx.x == 0) ctrl->red[k < 7 ? k : 8] = tot[0];
        }
        if (lb == 0 && threadIdx.x == 0) ctrl->pending = 1;
    } else if (ctrl) {
        block_sum<1, kBtBlock>(q, lds);
        if (threadIdx.x == 0) partial_bt[(multi ? M.bfirst[sc] : 0) + lblk] = q[0];
    }
}

template <class T, int LAY>
static void launch_bt(int G, hipStream_t s, ApgdCtrl* ctrl, int use_slot, const BMatrix<T>& B, int64_t na,
                      const int* tstart, const int32_t* tc, const double* partial, int ntrial, double* partial_bt,
                      hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, const SceneMap& M = SceneMap{},
                      int grid_blocks = 0) {
    const unsigned grid = M.bblk ? (unsigned)grid_blocks : bt_blocks(na, G);
    auto go = [&](auto kernel) {
        if (e0)  // start / stop events taken from the kernel's own dispatch
            hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(kBtBlock), 0, s, e0, e1, 0, ctrl, use_slot, B, na, tstart,
                                  tc, partial, ntrial, partial_bt, M);
        else
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBtBlock), 0, s, ctrl, use_slot, B, na, tstart, tc, partial,
                               ntrial, partial_bt, M);
    };
    switch (G) {
        case 1: go(k_bt<T, LAY, 1>); break;
        case 2: go(k_bt<T, LAY, 2>); break;
        case 4: go(k_bt<T, LAY, 4>); break;
        case 16: go(k_bt<T, LAY, 16>); break;
        default: go(k_bt<T, LAY, 8>); break;
    }
}

// Contact c's inputs to its Q rows and its projection: its coefficients (for the cone and
// viscous layouts the first three are the normal), the rows of w_s of its two particles, mu and,
// for kVisc, its kind.  A trial sweep loads them for all its contacts of an iteration before any
// arithmetic or store: a load the compiler finds after a store it cannot tell apart from it is
// issued only then, one more memory latency.  A row of an obstacle (a negative index) is zero; it
// is read at index 0 and replaced where it is used, so neither the load nor its arrival waits on
// a branch.
template <class T, int LAY>
struct RowIn {
    T co[lay<LAY>::NC], wi[lay<LAY>::NW], wj[lay<LAY>::NW];
    T mu;
    int kd;
    bool zi, zj;  // particle i / j is an obstacle: its row of w is zero
};
template <class T, int LAY>
__device__ inline void load_row(const BMatrix<T>& B, const T* __restrict__ w, int64_t c, int64_t i, int64_t j,
                                RowIn<T, LAY>& r) {
    constexpr int NC = lay<LAY>::NC, NW = lay<LAY>::NW;
#pragma unroll
    for (int k = 0; k < NC; ++k) r.co[k] = B.coef[c * NC + k];
    const T* wi = w + (i >= 0 ? i : 0) * NW;
    const T* wj = w + (j >= 0 ? j : 0) * NW;  // j is an obstacle only in periodic boxes
#pragma unroll
    for (int k = 0; k < NW; ++k) r.wi[k] = wi[k], r.wj[k] = wj[k];
    r.zi = i < 0, r.zj = j < 0;
    r.mu = (LAY == kCone || LAY == kVisc) ? B.mu[c] : (T)0;
    r.kd = LAY == kVisc ? (B.kind[c] & kKindMask) : 0;
}

// dt^2 (B w_s) at contact c, its 1 (rotation-free, rigid) or 3 global rows.
template <class T, int LAY>
__device__ inline void row_product(const RowIn<T, LAY>& r, T dt2, T (&q)[3]) {
    constexpr int NW = lay<LAY>::NW;
    const T* co = r.co;
    T wi[NW], wj[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) wi[k] = r.zi ? (T)0 : r.wi[k], wj[k] = r.zj ? (T)0 : r.wj[k];
    if (LAY == kRotFree) {
        q[0] = dt2 * (co[0] * (wi[0] - wj[0]) + co[1] * (wi[1] - wj[1]) + co[2] * (wi[2] - wj[2]));
    } else if (LAY == kRigid) {
        T acc = co[0] * (wi[0] - wj[0]) + co[1] * (wi[1] - wj[1]) + co[2] * (wi[2] - wj[2]);
        acc += co[3] * wi[3] + co[4] * wi[4] + co[5] * wi[5];
        acc -= co[6] * wj[3] + co[7] * wj[4] + co[8] * wj[5];
        q[0] = dt2 * acc;
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            T acc = wi[d] - wj[d];
            acc += co[3 + 3 * d] * wi[3] + co[4 + 3 * d] * wi[4] + co[5 + 3 * d] * wi[5];
            acc -= co[12 + 3 * d] * wj[3] + co[13 + 3 * d] * wj[4] + co[14 + 3 * d] * wj[5];
            q[d] = dt2 * acc;
        }
    }
}

// Q lambda_s at contact c (the 1..4 local rows NR) from w_s = M^-1 B^T lambda_s: dt^2 B w_s,
// then global2local for kVisc on the loaded normal and kind (lagrange_multiplier.hpp:143-163, 403-434).
template <class T, int LAY>
__device__ inline void q_rows(const RowIn<T, LAY>& r, T dt2, T (&q)[lay<LAY>::NR]) {
    constexpr int NR = lay<LAY>::NR;
    T g[3];
    row_product<T, LAY>(r, dt2, g);
    if constexpr (LAY == kVisc) {
        T dn = g[0] * r.co[0];
        dn += g[1] * r.co[1];
        dn += g[2] * r.co[2];
        q[0] = q[1] = q[2] = q[3] = (T)0;
        if (r.kd == 0) {
            q[0] = dn;
        } else if (r.kd == 1) {
            q[0] = dn;
            q[1] = -dn;
        } else {
            q[0] = -dn;
            q[1] = g[0], q[2] = g[1], q[3] = g[2];
        }
    } else {
#pragma unroll
        for (int d = 0; d < NR; ++d) q[d] = g[d];
    }
}

// One contact of a trial point: y = lambda_k + beta (lambda_k - lambda_{k-1}), Q y from the two Q
// products, the gradient g = Q y + c, lambda+ = Proj(y - alpha g) stored in `lt`, and the contact's
// terms of the 8 line-search sums (see k_trial).
template <class T, int LAY>
__device__ inline void trial_contact(T* __restrict__ lt, int64_t cc_, T beta, T alpha, const RowIn<T, LAY>& rin,
                                     const T (&l0)[lay<LAY>::NR], const T (&l1)[lay<LAY>::NR],
                                     const T (&cv)[lay<LAY>::NR], const T (&q0)[lay<LAY>::NR],
                                     const T (&q1)[lay<LAY>::NR], double (&s)[8]) {
    constexpr int NR = lay<LAY>::NR;
    T y[NR], g[NR], z[NR], qy[NR];
#pragma unroll
    for (int d = 0; d < NR; ++d) {
        y[d] = l0[d] + beta * (l0[d] - l1[d]);
        qy[d] = q0[d] + beta * (q0[d] - q1[d]);
        g[d] = qy[d] + cv[d];
        z[d] = y[d] - alpha * g[d];
    }
    if constexpr (LAY == kVisc) {
        const int kd = rin.kd;
        if (kd == 2) {
            // ViscousFriction at gamma_min (lagrange_multiplier.hpp:495-539): the nearer of
            // (0, Proj_cone(f)) and (max(x-, 0), 0)
            T f[3] = {z[1], z[2], z[3]};
            cone_project<T>(f, rin.co[0], rin.co[1], rin.co[2], rin.mu);
            const T zm = z[0] > (T)0 ? z[0] : (T)0;
            T df = z[0] * z[0], dm = (zm - z[0]) * (zm - z[0]);
            for (int d = 0; d < 3; ++d) {
                df += (f[d] - z[1 + d]) * (f[d] - z[1 + d]);
                dm += z[1 + d] * z[1 + d];
            }
            if (sqrt(df) < sqrt(dm)) {
                z[0] = (T)0, z[1] = f[0], z[2] = f[1], z[3] = f[2];
            } else {
                z[0] = zm, z[1] = z[2] = z[3] = (T)0;
            }
        } else {
            // max(lambda, 0) on the contact's 1 or 2 rows (lagrange_multiplier.hpp:197-201, 487-493)
            z[0] = z[0] > (T)0 ? z[0] : (T)0;
            z[1] = (kd == 1 && z[1] > (T)0) ? z[1] : (T)0;
            z[2] = z[3] = (T)0;
        }
    } else if constexpr (LAY != kCone) {
        z[0] = z[0] > (T)0 ? z[0] : (T)0;  // max(lambda, 0): lagrange_multiplier.hpp:104-107
    } else {
        // Coulomb-cone projection: lagrange_multiplier.hpp:243-268
        cone_project<T>(z, rin.co[0], rin.co[1], rin.co[2], rin.mu);
    }
#pragma unroll
    for (int d = 0; d < NR; ++d) {
        lt[cc_ * NR + d] = z[d];
        const double dy = (double)z[d] - (double)y[d], dl = (double)z[d] - (double)l0[d];
        s[0] += (double)g[d] * dy;
        s[1] += dy * dy;
        s[2] += (double)g[d] * dl;
        s[3] += dl * dl;
        s[4] += (double)y[d] * (double)qy[d];
        s[5] += (double)cv[d] * (double)y[d];
        s[6] += (double)cv[d] * (double)z[d];
        s[7] += (double)z[d] * (double)z[d];
    }
}

// One trial point of the line search over the block's contacts:
// y = lambda_k + beta (lambda_k - lambda_{k-1}) and Q y = (1 + beta) Q lambda_k - beta Q lambda_{k-1},
// with Q lambda_k formed on the fly from w_k (dt^2 B w_k, the row product) and stored in lambda_k's
// slot of qlam, where the rounds after lambda_k's successor is accepted read it as Q lambda_{k-1};
// the gradient g = Q y + c, lambda+ = Proj(y - alpha g), and the block sums of
// {<g, l+ - y>, |l+ - y|^2, <g, l+ - l_k>, |l+ - l_k|^2, <y, Qy>, <c, y>, <c, l+>, |l+|^2}.
// The control step of the previous trial runs first, on k_bt's block sums of <l+, Q l+>.
template <class T, int LAY>
__global__ void __launch_bounds__(kBlock) k_trial(const ApgdCtrl* __restrict__ in, ApgdCtrl* __restrict__ out,
                                                 BMatrix<T> B, int64_t nc, const int32_t* __restrict__ ci,
                                                 const int32_t* __restrict__ cj, int64_t ni,
                                                 double* __restrict__ partial, const double* __restrict__ partial_bt,
                                                 int nbt, SceneMap M) {
    constexpr int NR = lay<LAY>::NR;
    __shared__ double lds[8][kBlock / 64];
    __shared__ ApgdCtrl st;
    const SceneBlock sb = scene_block(M, nc);
    // U contacts c + u cstep per iteration (2 for the rotation-free layout, 1 for the others,
    // whose products need more registers); the first contacts' particle indices are loaded
    // together with the state and the block sums, each next group one iteration ahead
    constexpr int U = LAY == kRotFree ? 2 : 1;
    // grid-stride over the scene's contacts in logical block order (XCD-aware for one scene: the
    // blocks of an XCD take neighbouring kBlock-contact groups of every sweep); the grid (row_blocks)
    // gives every block the same iteration count
    const int lblk = M.rblk ? sb.lb : xcd_logical(sb.lb, sb.nb);
    const int64_t cstep = (int64_t)sb.nb * kBlock;
    const int64_t cend = sb.hi;
    int64_t c = sb.lo + (int64_t)lblk * kBlock + threadIdx.x;
    // particle indices are read at every position, those past the end at the last contact (or 0;
    // the arrays hold at least 1024 entries), so that no branch waits for them: a contact past the
    // end is never stored or summed, and reads only valid rows
    auto clampc = [&](int64_t x) { return x < cend ? x : (cend > 0 ? cend - 1 : (int64_t)0); };
    int32_t inx[U], jnx[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t c0 = clampc(c + u * cstep);
        inx[u] = ci[c0], jnx[u] = cj[c0];
    }
    // The first contacts' inputs are loaded before the control step has chosen the iterates' slots,
    // from the slots it will choose when it accepts the pending trial point (lambda_k := the trial
    // point, lambda_{k-1} := lambda_k) or, with no step pending, the current ones; every thread
    // reads the three slot numbers of the state, so the loads wait for one memory round trip and the
    // step's reduction and arithmetic overlap them.  After a rejection (the slots stay) the block
    // loads them again.  The values are the same either way: this only moves the loads earlier.
    const ApgdCtrl* cin = in + sb.sc;
    const bool opend = cin->pending != 0;
    const int pcur = opend ? cin->trial : cin->cur, pprev = opend ? cin->cur : cin->prev;
    const int64_t stride = NR * B.cap, wstride = 6 * B.pcap;
    int64_t cu[U], iu[U], ju[U];
    bool hu[U];
    T l0[U][NR], l1[U][NR], cv[U][NR], q1[U][NR];
    RowIn<T, LAY> rin[U];
    // the group of contacts at c (its particle indices loaded one group ahead)
    auto index = [&]() {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            hu[u] = c + u * cstep < cend;
            cu[u] = hu[u] ? c + u * cstep : c;  // past the end: a duplicate of c, not stored
            iu[u] = (int64_t)inx[u] - ni, ju[u] = (int64_t)jnx[u] - ni;
            const int64_t cn = clampc(c + (U + u) * cstep);
            inx[u] = ci[cn], jnx[u] = cj[cn];
        }
    };
    // its inputs, lambda_k and Q lambda_k from slot scur and lambda_{k-1}, Q lambda_{k-1} from sprev
    auto fetch = [&](int scur, int sprev) {
        const T* lc = slot<T>(B.lam, scur, stride);
        const T* lp = slot<T>(B.lam, sprev, stride);
        const T* qp = slot<T>(B.qlam, sprev, stride);
        const T* wc = slot<T>(B.w, scur, wstride);
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int d = 0; d < NR; ++d) {
                const int64_t r = cu[u] * NR + d;
                l0[u][d] = lc[r], l1[u][d] = lp[r], cv[u][d] = B.c[r], q1[u][d] = qp[r];
            }
            load_row<T, LAY>(B, wc, cu[u], iu[u], ju[u], rin[u]);
        }
    };
    const double* pbt = partial_bt + (M.rblk ? M.bfirst[sb.sc] : 0);
    const int npbt = M.rblk ? M.bcount[sb.sc] : nbt;
    double pv[kCtrlR];
    control_load(pbt, npbt, pv);
    {
        // every lane issues the first group's loads, a lane with no contact at a valid index whose
        // values it leaves unused (contact and particle arrays hold at least 1024 entries)
        const int64_t c0 = c;
        if (c >= cend) c = clampc(c);
        index();
        if (c0 >= cend) hu[0] = false;
        fetch(pcur, pprev);
        c = c0;
    }
    // keep the block sums' use after these loads are issued
    __builtin_amdgcn_sched_barrier(0);
    control_finish(cin, out + sb.sc, pbt, npbt, pv, st, lds, sb.lb == 0, M.done ? M.done + sb.sc : nullptr);
    if (st.done) return;
    const int cur = st.cur, prev = st.prev, trial = st.trial;
    if (c < cend && (cur != pcur || prev != pprev)) fetch(cur, prev);  // the same in the whole block
    const T beta = (T)st.beta, alpha = (T)st.alpha, dt2 = (T)st.dt2;
    T* qc = slot<T>(B.qlam, cur, stride);
    T* lt = slot<T>(B.lam, trial, stride);
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    while (c < cend) {
        T q0[U][NR];
#pragma unroll
        for (int u = 0; u < U; ++u) q_rows<T, LAY>(rin[u], dt2, q0[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (hu[u]) {
#pragma unroll
                for (int d = 0; d < NR; ++d) qc[cu[u] * NR + d] = q0[u][d];
                trial_contact<T, LAY>(lt, cu[u], beta, alpha, rin[u], l0[u], l1[u], cv[u], q0[u], q1[u], s);
            }
        c += U * cstep;
        if (c < cend) {
            index();
            fetch(cur, prev);
        }
    }
    block_sum<8>(s, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 8; ++k) partial[k * trial_pstride(M) + sb.pbase + lblk] = s[k];
}

// Control logic of apgd.hpp:172-209 / pgd apgd.hpp:68-89 on the reduced sums.
// In the FP32 phase (low_precision) the loop also stops when the line search
// can no longer resolve a decrease (repeated rejections, or lambda+ == y):
// that is the precision floor of FP32, and the FP64 phase takes over from the
// last iterate with the last accepted L.
__device__ inline void apgd_control(ApgdCtrl* ctrl) {
    const double* r = ctrl->red;
    const double f_trial = 0.5 * r[7] + r[6];
    const double f_y = 0.5 * r[4] + r[5];
    bool accept = true;
    ctrl->q_applies += 1;
    if (ctrl->method == SCOPI_METHOD_APGD && ctrl->dynamic) {
        if (r[1] == 0.0) {
            if (ctrl->low_precision) {
                ctrl->done = 1;
                ctrl->status = kStatusStalled;
                return;
            }
            // lambda+ == y: the reference's test holds with equality, L doubles
            // until it overflows to +inf, alpha becomes 0 and the NaN right-hand
            // side ends the loop with lambda+ = Proj(y) = y (apgd.hpp:174-181).
            ctrl->L = INFINITY;
            ctrl->alpha = 0.0;
        } else {
            const double rhs = f_y + r[0] + 0.5 * ctrl->L * r[1];
            if (f_trial >= rhs) {
                accept = false;
                ctrl->L *= 2;
                ctrl->alpha = 1. / ctrl->L;
                ctrl->backtracks += 1;
                ctrl->consec_bt += 1;
                if (ctrl->low_precision && ctrl->consec_bt > kMaxConsecutiveRejections32) {
                    ctrl->done = 1;
                    ctrl->status = kStatusStalled;
                }
                // Doubling reaches +inf after ~1100 rejections; beyond that only a
                // non-finite objective can keep rejecting, so stop instead of looping.
                if (ctrl->backtracks > 64 * (ctrl->max_ite + 1100)) {
                    ctrl->done = 1;
                    ctrl->status = SCOPI_HIP_NOT_CONVERGED;
                }
            }
        }
    }
    if (!accept) return;
    ctrl->consec_bt = 0;
    if (isfinite(ctrl->L)) ctrl->L_acc = ctrl->L;
    ctrl->iter += 1;
    ctrl->res = sqrt(r[3]);
    const int old_prev = ctrl->prev;
    // the FP32 phase also ends when the residual has stopped decreasing (its precision floor)
    bool stagnant = false;
    if (ctrl->low_precision) {
        if (!(ctrl->res >= 0.5 * ctrl->best_res)) {
            ctrl->best_res = ctrl->res;
            ctrl->since_best = 0;
        } else {
            ctrl->since_best += 1;
            stagnant = ctrl->since_best > kMaxStagnation32;
        }
    }
    if (ctrl->res < ctrl->tol || ctrl->res < ctrl->rtol * sqrt(r[8]) || stagnant) {
        ctrl->prev = ctrl->cur;
        ctrl->cur = ctrl->trial;
        ctrl->trial = old_prev;
        ctrl->done = 1;
        ctrl->status = stagnant ? kStatusStalled : 0;
        return;
    }
    double theta = ctrl->theta, theta1 = 1.0, beta = 0.0;
    if (ctrl->method == SCOPI_METHOD_APGD) {
        theta1 = 0.5 * (theta * sqrt(4 + theta * theta) - theta * theta);
        beta = theta * (1 - theta) / (theta * theta + theta1);
        if (ctrl->dynamic) {
            if (r[2] > 0) {
                beta = 0.0;
                theta1 = 1.0;
            }
            ctrl->L *= 0.97;
            ctrl->alpha = 1. / ctrl->L;
        }
    }
    ctrl->theta = theta1;
    ctrl->beta = beta;
    ctrl->prev = ctrl->cur;
    ctrl->cur = ctrl->trial;
    ctrl->trial = old_prev;
    if (ctrl->iter >= ctrl->max_ite) {
        ctrl->done = 1;
        ctrl->status = SCOPI_HIP_NOT_CONVERGED;
    }
}

// One block: <lambda+, Q lambda+> = dt^2 (sum of k_bt's block sums), then the control step.
// Publishes the pending control step of the previous product to the block's copy
// `st` (every block computes the same state: same inputs, same reduction order)
// and block 0 writes it to `out`.  In two parts, so that a kernel can issue other loads between
// the block sums' and their use: control_load loads the thread's share of the sums (the first
// kBtBlocksMax of them, all of them for every grid of k_bt), control_finish sums them in the order
// of thread_sum and runs the step.
__device__ inline void control_load(const double* __restrict__ partial_bt, int nbt, double (&pv)[kCtrlR]) {
    // a missing term reads entry 0 (the buffer holds at least 1024) and is replaced by 0.0 where it
    // is summed, so that no branch waits for the load
#pragma unroll
    for (int r = 0; r < kCtrlR; ++r) {
        const int b = (int)threadIdx.x + r * kBlock;
        pv[r] = partial_bt[b < nbt ? b : 0];
    }
}
__device__ inline void control_finish(const ApgdCtrl* __restrict__ in, ApgdCtrl* __restrict__ out,
                                      const double* __restrict__ partial_bt, int nbt, const double (&pv)[kCtrlR],
                                      ApgdCtrl& st, double (*lds)[kBlock / 64], bool leader, int32_t* done_flag) {
    // the state is read by every thread from one address (scalar loads, whose wait does not include
    // the vector loads a kernel issued before this call)
    ApgdCtrl c = *in;
    double t = 0.0;
#pragma unroll
    for (int r = 0; r < kCtrlR; ++r) t += ((int)threadIdx.x + r * kBlock < nbt) ? pv[r] : 0.0;
    for (int b = (int)threadIdx.x + kCtrlR * kBlock; b < nbt; b += kBlock) t += partial_bt[b];
    double tot[1] = {t};
    // summed whether or not the state needs it, so that no lane waits for the state to learn that
    block_sum<1>(tot, lds);
    if (threadIdx.x == 0) {
        // the step runs on a register copy of the state (on the shared copy, every field access would
        // be a round trip to LDS), which is then published to the block and, by the leader, to `out`
        if (c.pending && !c.done) {
            c.red[7] = c.dt2 * tot[0];
            apgd_control(&c);
        }
        c.pending = 0;
        st = c;
        if (leader) {
            *out = c;
            if (done_flag) *done_flag = c.done;
        }
    }
    __syncthreads();
}
__device__ inline void control_step(const ApgdCtrl* __restrict__ in, ApgdCtrl* __restrict__ out,
                                    const double* __restrict__ partial_bt, int nbt, ApgdCtrl& st,
                                    double (*lds)[kBlock / 64], bool leader, int32_t* done_flag) {
    // the state and k_bt's block sums are loaded together, before the state is known to need the
    // sums (a fixed summation order)
    double pv[kCtrlR];
    control_load(partial_bt, nbt, pv);
    control_finish(in, out, partial_bt, nbt, pv, st, lds, leader, done_flag);
}

// The control step alone (end of a batch of rounds, before the host reads the state);
// with scenes, one block per scene of the grid.
__global__ void __launch_bounds__(kBlock) k_ctrl(const ApgdCtrl* __restrict__ in, ApgdCtrl* __restrict__ out,
                                                const double* __restrict__ partial_bt, int nbt, SceneMap M) {
    __shared__ double lds[1][kBlock / 64];
    __shared__ ApgdCtrl st;
    if (!M.rblk) {
        control_step(in, out, partial_bt, nbt, st, lds, true, nullptr);
        return;
    }
    const int sc = M.slist[blockIdx.x];
    control_step(in + sc, out + sc, partial_bt + M.bfirst[sc], M.bcount[sc], st, lds, true, M.done + sc);
}

// ------------------------------------------------------------------ per-scene work queue
// Small scenes of a batch are solved whole by one workgroup each: a persistent grid takes scenes
// from an atomic counter over a list in decreasing contact count, and the workgroup runs every
// round of its scene (the trial sweep over the scene's contacts, the B^T sweep over its particles,
// then the control step of apgd.hpp:172-209 by thread 0) with barriers between the sweeps, so a
// scene's whole solve (both phases of FP32 + FP64 refinement) is one launch with no host polling.
// Workgroup-scope barriers order the sweeps' global stores and loads: a workgroup runs on one CU,
// whose waves share its L1.  Sums are taken in a fixed order within the workgroup, so a scene's
// iterates depend only on the scene and the workgroup size, the same in any mix.

// Fields of the state every thread reads in a round, published by thread 0.
struct QueueRound {
    int32_t cur, prev, trial, done;
    double beta, alpha, dt2;
};

// w_s = M^-1 B^T lambda_s over the particles [plo, phi), G lanes per particle; returns the thread's
// share of sum a . w_s, a = B^T lambda_s.
template <class T, int LAY, int NTH, int G>
__device__ inline double queue_bt(const BMatrix<T>& B, const QueueArgs& A, int s, int64_t plo, int64_t phi) {
    constexpr int NR = lay<LAY>::NR, NW = lay<LAY>::NW;
    const T* lam = slot<T>(B.lam, s, NR * B.cap);
    T* w = slot<T>(B.w, s, 6 * B.pcap);
    const int sub = (int)(threadIdx.x % G);
    double q = 0.0;
    int64_t p = plo + threadIdx.x / G;
    int e0 = p < phi ? A.tstart[p] : 0, e1 = p < phi ? A.tstart[p + 1] : 0;
    // a lane group shares p, so each group passes through the loop as a whole; the next particle's
    // entry range is loaded one pass ahead
    for (; p < phi; p += NTH / G) {
        const int f0 = e0, f1 = e1;
        if (p + NTH / G < phi) e0 = A.tstart[p + NTH / G], e1 = A.tstart[p + NTH / G + 1];
        T mv[NW];  // M^-1 ahead of the stores of w, as in k_bt
#pragma unroll
        for (int k = 0; k < NW; ++k) mv[k] = B.minv[p * NW + k];
        T acc[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) acc[k] = (T)0;
        bt_gather<T, LAY, G>(B, lam, A.tc, f0, f1, sub, acc);
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
            for (int k = 0; k < NW; ++k) acc[k] += __shfl_xor(acc[k], o, G);
        if (sub == 0)
#pragma unroll
            for (int k = 0; k < NW; ++k) {
                const T wk = mv[k] * acc[k];
                w[p * NW + k] = wk;
                q += (double)acc[k] * (double)wk;
            }
    }
    return q;
}

// One phase of the loop (one precision) on the scene's contacts [clo, chi) and particles [plo, phi):
// lambda_0 = 0, or the warm start in slot 0; L and alpha the first step's.  The final state is left
// in `st` (all threads may read it after the phase).
template <class T, int LAY, int NTH, int G>
__device__ void queue_phase(const BMatrix<T>& B, const QueueArgs& A, int64_t clo, int64_t chi, int64_t plo,
                            int64_t phi, bool warm, double tol, double rtol, int64_t max_ite, double L, double alpha,
                            int low, ApgdCtrl& st, QueueRound& sh, double (*lds)[NTH / 64]) {
    constexpr int NR = lay<LAY>::NR, NW = lay<LAY>::NW;
    const int64_t stride = NR * B.cap, wstride = 6 * B.pcap;
    // lambda_0 in slot 0, lambda_{-1} = lambda_0 in slot 1 with a finite Q lambda_{-1} (beta = 0 in
    // the first round), w_0 in slot 0
    for (int64_t r = NR * clo + threadIdx.x; r < NR * chi; r += NTH) {
        const T l0 = warm ? B.lam[r] : (T)0;
        if (!warm) B.lam[r] = (T)0;
        B.lam[stride + r] = l0;
        B.qlam[stride + r] = (T)0;
    }
    if (!warm)
        for (int64_t k = NW * plo + threadIdx.x; k < NW * phi; k += NTH) B.w[k] = (T)0;
    __syncthreads();
    if (warm) {
        (void)queue_bt<T, LAY, NTH, G>(B, A, 0, plo, phi);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ApgdCtrl c{};
        const bool empty = chi <= clo;
        c.L = L;
        c.L_acc = L;
        c.alpha = alpha;
        c.theta = 1.0;
        c.beta = 0.0;
        c.tol = tol;
        c.rtol = rtol;
        c.dt2 = A.dt2;
        c.best_res = INFINITY;
        c.max_ite = max_ite;
        c.q_applies = warm ? 1 : 0;  // the product of the warm start's w_0
        c.cur = 0;
        c.prev = 1;
        c.trial = 2;
        c.done = (empty || max_ite <= 0) ? 1 : 0;
        c.status = (!empty && max_ite <= 0) ? SCOPI_HIP_NOT_CONVERGED : 0;
        c.res = c.done ? 0.0 : INFINITY;
        c.method = A.method;
        c.dynamic = A.dynamic;
        c.low_precision = low;
        st = c;
        sh = QueueRound{c.cur, c.prev, c.trial, c.done, c.beta, c.alpha, c.dt2};
    }
    __syncthreads();
    for (;;) {
        const QueueRound rd = sh;
        if (rd.done) break;
        const T beta = (T)rd.beta, alpha_t = (T)rd.alpha, dt2 = (T)rd.dt2;
        const T* lc = slot<T>(B.lam, rd.cur, stride);
        const T* lp = slot<T>(B.lam, rd.prev, stride);
        const T* wc = slot<T>(B.w, rd.cur, wstride);
        T* qc = slot<T>(B.qlam, rd.cur, stride);
        const T* qp = slot<T>(B.qlam, rd.prev, stride);
        T* lt = slot<T>(B.lam, rd.trial, stride);
        double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        // U contacts c + u NTH per iteration, as in k_trial, every load of an iteration issued before
        // the first use and the next contacts' particle indices one iteration ahead
        constexpr int U = LAY == kRotFree ? 2 : 1;
        int64_t c = clo + threadIdx.x;
        int32_t inx[U], jnx[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            inx[u] = jnx[u] = 0;
            if (c + u * NTH < chi) inx[u] = A.ci[c + u * NTH], jnx[u] = A.cj[c + u * NTH];
        }
        for (; c < chi; c += U * NTH) {
            int64_t cu[U], iu[U], ju[U];
            bool hu[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                hu[u] = c + u * NTH < chi;
                cu[u] = hu[u] ? c + u * NTH : c;  // past the end: a duplicate of c, not stored
                iu[u] = (int64_t)inx[u] - A.ni, ju[u] = (int64_t)jnx[u] - A.ni;
                const int64_t cn = c + (U + u) * NTH;
                if (cn < chi) inx[u] = A.ci[cn], jnx[u] = A.cj[cn];
            }
            T l0[U][NR], l1[U][NR], cv[U][NR], q1[U][NR], q0[U][NR];
            RowIn<T, LAY> rin[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
#pragma unroll
                for (int d = 0; d < NR; ++d) {
                    const int64_t r = cu[u] * NR + d;
                    l0[u][d] = lc[r], l1[u][d] = lp[r], cv[u][d] = B.c[r], q1[u][d] = qp[r];
                }
                load_row<T, LAY>(B, wc, cu[u], iu[u], ju[u], rin[u]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) q_rows<T, LAY>(rin[u], dt2, q0[u]);
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (hu[u]) {
#pragma unroll
                    for (int d = 0; d < NR; ++d) qc[cu[u] * NR + d] = q0[u][d];
                    trial_contact<T, LAY>(lt, cu[u], beta, alpha_t, rin[u], l0[u], l1[u], cv[u], q0[u], q1[u], s);
                }
        }
        // the barriers of the sum also order the block's stores of lambda+ before the B^T sweep
        block_sum<8, NTH>(s, lds);
        double q[1] = {queue_bt<T, LAY, NTH, G>(B, A, rd.trial, plo, phi)};
        block_sum<1, NTH>(q, lds);
        if (threadIdx.x == 0) {
            ApgdCtrl c = st;
#pragma unroll
            for (int k = 0; k < 7; ++k) c.red[k] = s[k];
            c.red[8] = s[7];
            c.red[7] = c.dt2 * q[0];
            apgd_control(&c);
            st = c;
            sh = QueueRound{c.cur, c.prev, c.trial, c.done, c.beta, c.alpha, c.dt2};
        }
        __syncthreads();
    }
}

// The persistent grid: every workgroup takes scenes until the list is exhausted, so each one
// reaches the exit.  A scene's final multipliers (FP64) end in slot 0; its final state in A.result.
template <int LAY, int NTH, int G>
__global__ void __launch_bounds__(NTH) k_apgd_queue(QueueArgs A, BMatrix<double> B64, BMatrix<float> B32) {
    constexpr int NR = lay<LAY>::NR;
    __shared__ double lds[8][NTH / 64];
    __shared__ ApgdCtrl st;
    __shared__ QueueRound sh;
    __shared__ int32_t s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(A.next, 1);
        __syncthreads();
        const int32_t item = s_item;
        if (item >= A.nqueue) return;
        const int sc = A.order[item];
        const int64_t clo = A.crange[2 * sc], chi = A.crange[2 * sc + 1];
        const int64_t plo = A.arange[2 * sc], phi = A.arange[2 * sc + 1];
        int64_t it32 = 0, bt32 = 0, qa32 = 0, max_ite = A.max_ite;
        double L = 1.0 / A.alpha, alpha = A.alpha;
        bool warm = A.warm != 0;
        if (A.fp32) {
            // FP32 phase to its relative stop or precision floor, then FP64 refinement from its
            // last iterate with its last accepted L and the iterations left
            queue_phase<float, LAY, NTH, G>(B32, A, clo, chi, plo, phi, warm, A.tol32, A.rtol32, max_ite, L, alpha, 1,
                                            st, sh, lds);
            const float* lf = B32.lam + (int64_t)st.cur * NR * B32.cap;
            for (int64_t r = NR * clo + threadIdx.x; r < NR * chi; r += NTH) B64.lam[r] = (double)lf[r];
            it32 = st.iter, bt32 = st.backtracks, qa32 = st.q_applies;
            const double La = st.L_acc;
            if (isfinite(La) && La > 0) L = La, alpha = 1. / La;
            max_ite = A.max_ite - it32 < 1 ? 1 : A.max_ite - it32;
            warm = true;
            __syncthreads();
        }
        queue_phase<double, LAY, NTH, G>(B64, A, clo, chi, plo, phi, warm, A.tol, 0.0, max_ite, L, alpha, 0, st, sh,
                                         lds);
        const int fin = st.cur;
        if (fin != 0) {
            const double* lf = B64.lam + (int64_t)fin * NR * B64.cap;
            for (int64_t r = NR * clo + threadIdx.x; r < NR * chi; r += NTH) B64.lam[r] = lf[r];
        }
        if (threadIdx.x == 0) {
            QueueResult q;
            q.iter = st.iter + it32;
            q.iter32 = it32;
            q.backtracks = st.backtracks + bt32;
            q.q_applies = st.q_applies + qa32;
            q.res = st.res;
            q.L_acc = st.L_acc;
            q.status = st.status == kStatusStalled ? 0 : st.status;
            q.pad = 0;
            A.result[sc] = q;
        }
        __syncthreads();  // st, sh and s_item are reused by the next scene
    }
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_zero(T* __restrict__ a, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlock) a[t] = (T)0;
}

template <class TA, class TB>
__global__ void __launch_bounds__(kBlock) k_convert(const TA* __restrict__ a, TB* __restrict__ b, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlock)
        b[t] = (TB)a[t];
}

// Initial state in `ctrl`; with keep_L, L starts from the last accepted L of the
// previous phase's final state `last`.
__global__ void k_ctrl_init(ApgdCtrl* ctrl, const ApgdCtrl* last, double alpha, double L, double tol, double rtol,
                            double dt2, int64_t max_ite, int method, int dynamic, int keep_L, int low_precision) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double L_last = last->L_acc;
    if (keep_L && isfinite(L_last) && L_last > 0) {
        L = L_last;
        alpha = 1. / L;
    }
    ctrl->L_acc = L;
    ctrl->consec_bt = 0;
    ctrl->low_precision = low_precision;
    ctrl->L = L;
    ctrl->alpha = alpha;
    ctrl->theta = 1.0;
    ctrl->beta = 0.0;
    ctrl->tol = tol;
    ctrl->rtol = rtol;
    ctrl->dt2 = dt2;
    ctrl->res = INFINITY;
    ctrl->best_res = INFINITY;
    ctrl->since_best = 0;
    for (int k = 0; k < 9; ++k) ctrl->red[k] = 0.0;
    ctrl->iter = 0;
    ctrl->backtracks = 0;
    ctrl->max_ite = max_ite;
    ctrl->q_applies = 0;
    ctrl->cur = 0;
    ctrl->prev = 1;
    ctrl->trial = 2;
    ctrl->done = (max_ite <= 0) ? 1 : 0;
    ctrl->status = (max_ite <= 0) ? SCOPI_HIP_NOT_CONVERGED : 0;
    ctrl->method = method;
    ctrl->dynamic = dynamic;
    ctrl->pending = 0;
}

// Initial state of every scene in ctrl[0 .. K); with keep_L, from the scene's final
// state of the previous phase, ctrl[last_buf[sc] * K + sc] (L, and the iteration
// budget left).  A scene without contacts is done at once.
__global__ void k_ctrl_init_scenes(ApgdCtrl* ctrl, const int32_t* __restrict__ last_buf, int K, double alpha0,
                                   double L0, double tol, double rtol, double dt2, int64_t max_ite0, int method,
                                   int dynamic, int keep_L, int low_precision, const int64_t* __restrict__ crange,
                                   int32_t* __restrict__ done) {
    const int sc = blockIdx.x * blockDim.x + threadIdx.x;
    if (sc >= K) return;
    double L = L0, alpha = alpha0;
    int64_t max_ite = max_ite0;
    if (keep_L) {
        const ApgdCtrl& last = ctrl[(int64_t)last_buf[sc] * K + sc];
        const double L_last = last.L_acc;
        if (isfinite(L_last) && L_last > 0) {
            L = L_last;
            alpha = 1. / L;
        }
        max_ite = max_ite0 - last.iter;
        if (max_ite < 1) max_ite = 1;
    }
    ApgdCtrl c{};
    c.L_acc = L;
    c.consec_bt = 0;
    c.low_precision = low_precision;
    c.L = L;
    c.alpha = alpha;
    c.theta = 1.0;
    c.beta = 0.0;
    c.tol = tol;
    c.rtol = rtol;
    c.dt2 = dt2;
    c.res = 0.0;
    c.best_res = INFINITY;
    c.since_best = 0;
    for (int k = 0; k < 9; ++k) c.red[k] = 0.0;
    c.iter = 0;
    c.backtracks = 0;
    c.max_ite = max_ite;
    c.q_applies = 0;
    c.cur = 0;
    c.prev = 1;
    c.trial = 2;
    const bool empty = crange[2 * sc + 1] <= crange[2 * sc];
    c.done = (empty || max_ite <= 0) ? 1 : 0;
    c.status = (!empty && max_ite <= 0) ? SCOPI_HIP_NOT_CONVERGED : 0;
    if (!c.done) c.res = INFINITY;
    c.method = method;
    c.dynamic = dynamic;
    c.pending = 0;
    ctrl[sc] = c;
    done[sc] = c.done;
}

// dst slot 0 <- src slot src_slot[scene], row by row over every scene's contacts
// (per-contact grid of the scene map); a copy onto itself is skipped.
template <class TA, class TB, int NR>
__global__ void __launch_bounds__(kBlock) k_scene_slot(const TA* __restrict__ src, int64_t src_stride, TB* dst,
                                                      const int32_t* __restrict__ src_slot, SceneMap M) {
    const SceneBlock sb = scene_block(M, 0);
    const int sl = src_slot[sb.sc];
    if (std::is_same<TA, TB>::value && sl == 0 && (const void*)src == (const void*)dst) return;
    const TA* a = src + (int64_t)sl * src_stride;
    for (int64_t r = NR * sb.lo + (int64_t)sb.lb * kBlock + threadIdx.x; r < NR * sb.hi; r += (int64_t)sb.nb * kBlock)
        dst[r] = (TB)a[r];
}

// Set-up of the two-launch rounds for the scenes of the grid only, so that rows and particles of
// other scenes (a concurrent work-queue launch may own them) are not touched: cold, lambda_0 =
// lambda_{-1} = 0 and w_0 = 0; warm, lambda_{-1} = lambda_0 (the warm start in slot 0); and a finite
// Q lambda_{-1} = 0 (per-contact grid of the scene map).
template <class T, int NR, int NW>
__global__ void __launch_bounds__(kBlock) k_scene_setup(BMatrix<T> B, int warm, SceneMap M) {
    const SceneBlock sb = scene_block(M, 0);
    const int64_t stride = NR * B.cap;
    const int64_t step = (int64_t)sb.nb * kBlock;
    for (int64_t r = NR * sb.lo + (int64_t)sb.lb * kBlock + threadIdx.x; r < NR * sb.hi; r += step) {
        if (warm) {
            B.lam[stride + r] = B.lam[r];
        } else {
            B.lam[r] = (T)0;
            B.lam[stride + r] = (T)0;
        }
        B.qlam[stride + r] = (T)0;
    }
    if (!warm) {
        const int64_t p0 = M.arange[2 * sb.sc], p1 = M.arange[2 * sb.sc + 1];
        for (int64_t k = NW * p0 + (int64_t)sb.lb * kBlock + threadIdx.x; k < NW * p1; k += step) B.w[k] = (T)0;
    }
}

// u = v + dt M^-1 B^T lambda, omega likewise (OptimGradient.hpp:267-284); with no
// contact, or no rotational block (kRotFree), the corresponding part is v / omega.
template <int NW>
__global__ void __launch_bounds__(kBlock) k_velocities(Particles P, const double* __restrict__ w, int64_t pcap,
                                                      double dt, int has_w, double* __restrict__ u,
                                                      double* __restrict__ om) {
#pragma clang fp contract(off)
    const int64_t na = P.n - P.n_inactive;
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= na) return;
    const int64_t k = P.n_inactive + p;
    for (int d = 0; d < 3; ++d) {
        const bool tr = d < P.dim;
        const bool rot = (P.dim == 3) || d == 2;
        u[d * pcap + p] = tr ? P.v[d * P.cap + k] + (has_w ? dt * w[p * NW + d] : 0.0) : 0.0;
        om[d * pcap + p] =
            rot ? P.omega[d * P.cap + k] + ((has_w && NW == 6) ? dt * w[p * NW + 3 + d] : 0.0) : 0.0;
    }
}

// ------------------------------------------------------------------ host
template <class T>
static void bfree(BMatrix<T>& B) {
    for (void* p : {(void*)B.coef, (void*)B.tcoef, (void*)B.c, (void*)B.mu, (void*)B.minv, (void*)B.lam,
                    (void*)B.qlam, (void*)B.w})
        if (p) (void)hipFree(p);
    B = BMatrix<T>{};
}

template <class T>
static void balloc(BMatrix<T>& B, int64_t ccap, int64_t pcap) {
    B.cap = ccap;
    B.pcap = pcap;
    SCOPI_HIP_CHECK(hipMalloc(&B.coef, 21 * ccap * sizeof(T)));
    SCOPI_HIP_CHECK(hipMalloc(&B.tcoef, 9 * 2 * ccap * sizeof(T)));
    SCOPI_HIP_CHECK(hipMalloc(&B.c, 4 * ccap * sizeof(T)));  // NR <= 4 rows per contact
    SCOPI_HIP_CHECK(hipMalloc(&B.mu, ccap * sizeof(T)));
    SCOPI_HIP_CHECK(hipMalloc(&B.minv, 6 * pcap * sizeof(T)));
    SCOPI_HIP_CHECK(hipMalloc(&B.lam, 3 * 4 * ccap * sizeof(T)));
    SCOPI_HIP_CHECK(hipMalloc(&B.qlam, 3 * 4 * ccap * sizeof(T)));
    SCOPI_HIP_CHECK(hipMalloc(&B.w, 3 * 6 * pcap * sizeof(T)));
}

Solver::~Solver() {
    bfree(b64_);
    bfree(b32_);
    for (void* p : {(void*)istart_, (void*)iend_, (void*)jstart_, (void*)jend_, (void*)tcount_, (void*)tstart_,
                    (void*)tc_, (void*)jkeys_, (void*)jlist_, (void*)u_, (void*)w_, (void*)ctrl_,
                    (void*)partial_trial_, (void*)partial_bt_})
        if (p) (void)hipFree(p);
    if (h_ctrl_) (void)hipHostFree(h_ctrl_);
    for (void* p : {(void*)d_arange_, (void*)d_done_, (void*)d_buf_, (void*)d_tab_, (void*)mctrl_, (void*)fp_c0_,
                    (void*)fp_s_, (void*)fp_partial_, (void*)vkind_, (void*)vrows_})
        if (p) (void)hipFree(p);
    if (h_fp_partial_) (void)hipHostFree(h_fp_partial_);
    if (h_vrows_) (void)hipHostFree(h_vrows_);
    for (void* p : {(void*)h_done_, (void*)h_tab_, (void*)h_ctrls_, (void*)h_qorder_, (void*)h_qres_})
        if (p) (void)hipHostFree(p);
    for (void* p : {(void*)d_qorder_, (void*)d_qnext_, (void*)d_qres_})
        if (p) (void)hipFree(p);
    for (hipGraphExec_t e : {graph_exec32_, graph_exec64_})
        if (e) (void)hipGraphExecDestroy(e);
    if (qstream_) {
        (void)hipStreamSynchronize(qstream_);
        (void)hipStreamDestroy(qstream_);
        (void)hipEventDestroy(qfork_);
        (void)hipEventDestroy(qjoin_);
    }
}

void Solver::reserve(int64_t na, int64_t nc, bool fp32) {
    if (!ctrl_) {
        SCOPI_HIP_CHECK(hipMalloc(&ctrl_, 2 * sizeof(ApgdCtrl)));
        SCOPI_HIP_CHECK(hipMemset(ctrl_, 0, 2 * sizeof(ApgdCtrl)));
        SCOPI_HIP_CHECK(hipHostMalloc(&h_ctrl_, sizeof(ApgdCtrl)));
        SCOPI_HIP_CHECK(hipMalloc(&partial_trial_, kRedBlocksMax * kPartialStride * sizeof(double)));
        partial_cap_ = kRedBlocksMax;
    }
    const bool grow_p = na + 2 > pcap_, grow_c = nc + 1 > ccap_;
    if (grow_p) {
        for (void* p : {(void*)istart_, (void*)iend_, (void*)jstart_, (void*)jend_, (void*)tcount_, (void*)tstart_,
                        (void*)u_, (void*)w_})
            if (p) SCOPI_HIP_CHECK(hipFree(p));
        pcap_ = na + na / 4 + 1024;
        for (int** a : {&istart_, &iend_, &jstart_, &jend_, &tcount_, &tstart_})
            SCOPI_HIP_CHECK(hipMalloc(a, pcap_ * sizeof(int)));
        SCOPI_HIP_CHECK(hipMalloc(&u_, 3 * pcap_ * sizeof(double)));
        SCOPI_HIP_CHECK(hipMalloc(&w_, 3 * pcap_ * sizeof(double)));
    }
    if (grow_c) {
        for (void* p : {(void*)jkeys_, (void*)jlist_, (void*)tc_})
            if (p) SCOPI_HIP_CHECK(hipFree(p));
        ccap_ = (nc + nc / 4 + 1024 + 63) / 64 * 64;  // even: slots stay aligned for paired loads
        SCOPI_HIP_CHECK(hipMalloc(&jkeys_, ccap_ * sizeof(uint32_t)));
        SCOPI_HIP_CHECK(hipMalloc(&jlist_, ccap_ * sizeof(int32_t)));
        SCOPI_HIP_CHECK(hipMalloc(&tc_, 2 * ccap_ * sizeof(int32_t)));
    }
    const int64_t nbt = (int64_t)bt_blocks(pcap_, 16) + 1024;
    if (nbt > partial_bt_cap_) {
        if (partial_bt_) SCOPI_HIP_CHECK(hipFree(partial_bt_));
        SCOPI_HIP_CHECK(hipMalloc(&partial_bt_, nbt * sizeof(double)));
        partial_bt_cap_ = nbt;
    }
    if (grow_p || grow_c || !b64_.coef) {
        bfree(b64_);
        balloc(b64_, ccap_, pcap_);
        if (b32_alloc_) {
            bfree(b32_);
            b32_alloc_ = false;
        }
    }
    if (fp32 && !b32_alloc_) {
        balloc(b32_, ccap_, pcap_);
        b32_alloc_ = true;
    }
}

void Solver::timing(double* out) const {
    out[0] = t_bt_ms_;
    out[1] = (double)n_bt_;
    out[2] = t_row_ms_;
    out[3] = (double)n_row_;
    out[4] = t_trial_ms_;
    out[5] = (double)n_trial_;
}

template <class T, int LAY>
void Solver::run_phase(BMatrix<T>& B, const Particles& P, const Contacts& C, int64_t na, int64_t nc,
                       const scopi_hip_params& prm, double dt, double tol, double rtol, int64_t max_ite, bool warm,
                       bool keep_L, hipStream_t s, scopi_hip_solve_info& info) {
    constexpr int NR = lay<LAY>::NR;
    const unsigned gred = (unsigned)row_blocks(nc, reduce_blocks_, LAY, row_iters_);
    const int nbt = (int)bt_blocks(na, bt_group_);
    const int64_t ni = P.n_inactive;
    const int64_t stride = NR * B.cap, wstride = 6 * B.pcap;
    const double dt2 = dt * dt;
    const int32_t* ci = (const int32_t*)C.i;
    const int32_t* cj = (const int32_t*)C.j;
    // The first round reads lambda_0 (slot 0) with w_0, and lambda_{-1}, Q lambda_{-1} (slot 1)
    // with beta = 0, so that y = lambda_0 and Q y = Q lambda_0 whatever slot 1 holds if finite.
    if (!warm) {
        // lambda_0 = lambda_{-1} = 0, w_0 = 0
        hipLaunchKernelGGL(k_zero<T>, dim3(1024), dim3(kBlock), 0, s, B.lam, 3 * stride);
        hipLaunchKernelGGL(k_zero<T>, dim3(1024), dim3(kBlock), 0, s, B.w, wstride);
    } else {
        // lambda_0 is in slot 0: lambda_{-1} = lambda_0, and w_0
        SCOPI_HIP_CHECK(hipMemcpyAsync(B.lam + stride, B.lam, stride * sizeof(T), hipMemcpyDeviceToDevice, s));
        launch_bt<T, LAY>(bt_group_, s, (ApgdCtrl*)nullptr, 0, B, na, tstart_, tc_, nullptr, 0, nullptr);
        info.q_applies += 1;
    }
    hipLaunchKernelGGL(k_zero<T>, dim3(1024), dim3(kBlock), 0, s, B.qlam + stride, stride);
    hipLaunchKernelGGL(k_ctrl_init, dim3(1), dim3(64), 0, s, ctrl_, (const ApgdCtrl*)(ctrl_ + ctrl_cur_), prm.alpha,
                       1. / prm.alpha, tol, rtol, dt2, max_ite, prm.method, prm.dynamic_descent ? 1 : 0,
                       keep_L ? 1 : 0, sizeof(T) == 4 ? 1 : 0);
    SCOPI_HIP_CHECK(hipGetLastError());
    std::vector<hipEvent_t> ev;
    if (timing_) {
        ev.resize(4 * 64);
        for (auto& e : ev) SCOPI_HIP_CHECK(hipEventCreate(&e));
    }
    int64_t before = 0;
    // Rounds enqueued before each poll of the control state.  Without a prediction: 2, 4, 8, ... up
    // to 64.  A solve phase takes about as many rounds as the same phase of the last time step, so
    // with that count the first poll comes after three quarters of it, the next ones after half of
    // what the prediction leaves (each poll idles the device while the host reads the state and
    // enqueues the next batch), and past the prediction the batches double again from 4.  Rounds
    // enqueued after convergence exit at their first instruction: the batches change the number of
    // polls and of such empty rounds, not the solve.  At most 1024 rounds go between two polls.
    const int64_t predicted = timing_ ? 0 : last_rounds_[sizeof(T) == 4 ? 1 : 0];
    const bool predict = predicted > 32;
    bool past = false;
    int batch = predict ? (int)std::min<int64_t>(predicted * 3 / 4, 1024) : 2;
    int cur = 0;  // copy of ApgdCtrl holding the newest state
    if (graph_rounds_ > 0 && !timing_) {
        // hipGraph of graph_rounds_ rounds, the control step and the state's copy to the host,
        // replayed until the loop is done (rounds after `done` exit at their first instruction).
        // graph_rounds_ is odd: the rounds and the control step swap the state copies an even
        // number of times, so every replay starts from copy 0 like the first.  The executable graph
        // of each precision is kept across solves: a new capture (the launch shapes and arguments of
        // this solve) updates it in place when only parameters changed, and is instantiated otherwise.
        hipGraph_t graph = nullptr;
        hipGraphExec_t& exec = sizeof(T) == 4 ? graph_exec32_ : graph_exec64_;
        SCOPI_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int c = 0;
        for (int r = 0; r < graph_rounds_; ++r) {
            hipLaunchKernelGGL((k_trial<T, LAY>), dim3(gred), dim3(kBlock), 0, s, (const ApgdCtrl*)(ctrl_ + c),
                               ctrl_ + (1 - c), B, nc, ci, cj, ni, partial_trial_, (const double*)partial_bt_, nbt,
                               SceneMap{});
            c = 1 - c;
            launch_bt<T, LAY>(bt_group_, s, ctrl_ + c, -1, B, na, tstart_, tc_, partial_trial_, (int)gred, partial_bt_);
        }
        hipLaunchKernelGGL(k_ctrl, dim3(1), dim3(kBlock), 0, s, (const ApgdCtrl*)(ctrl_ + c), ctrl_ + (1 - c),
                           (const double*)partial_bt_, nbt, SceneMap{});
        c = 1 - c;
        SCOPI_HIP_CHECK(hipMemcpyAsync(h_ctrl_, ctrl_ + c, sizeof(ApgdCtrl), hipMemcpyDeviceToHost, s));
        SCOPI_HIP_CHECK(hipStreamEndCapture(s, &graph));
        if (exec) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            if (hipGraphExecUpdate(exec, graph, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
                (void)hipGetLastError();
                SCOPI_HIP_CHECK(hipGraphExecDestroy(exec));
                exec = nullptr;
            }
        }
        if (!exec) SCOPI_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        SCOPI_HIP_CHECK(hipGraphDestroy(graph));
        // With a prediction, the replays that cover three quarters of it go out before the first poll.
        int64_t replays = predict ? std::max<int64_t>(1, predicted * 3 / 4 / graph_rounds_) : 1;
        for (;;) {
            for (int64_t r = 0; r < replays; ++r) SCOPI_HIP_CHECK(hipGraphLaunch(exec, s));
            SCOPI_HIP_CHECK(hipStreamSynchronize(s));
            if (h_ctrl_->done) break;
            replays = 1;
        }
        cur = c;
    }
    for (; !(graph_rounds_ > 0 && !timing_);) {
        for (int r = 0; r < batch; ++r) {
            hipEvent_t* e = timing_ ? &ev[4 * r] : nullptr;
            if (e)
                hipExtLaunchKernelGGL(k_trial<T, LAY>, dim3(gred), dim3(kBlock), 0, s, e[0], e[1], 0,
                                      (const ApgdCtrl*)(ctrl_ + cur), ctrl_ + (1 - cur), B, nc, ci, cj, ni,
                                      partial_trial_, (const double*)partial_bt_, nbt, SceneMap{});
            else
                hipLaunchKernelGGL((k_trial<T, LAY>), dim3(gred), dim3(kBlock), 0, s, (const ApgdCtrl*)(ctrl_ + cur),
                                   ctrl_ + (1 - cur), B, nc, ci, cj, ni, partial_trial_, (const double*)partial_bt_,
                                   nbt, SceneMap{});
            cur = 1 - cur;
            launch_bt<T, LAY>(bt_group_, s, ctrl_ + cur, -1, B, na, tstart_, tc_, partial_trial_, (int)gred,
                              partial_bt_, e ? e[2] : nullptr, e ? e[3] : nullptr);
        }
        hipLaunchKernelGGL(k_ctrl, dim3(1), dim3(kBlock), 0, s, (const ApgdCtrl*)(ctrl_ + cur), ctrl_ + (1 - cur),
                           (const double*)partial_bt_, nbt, SceneMap{});
        cur = 1 - cur;
        SCOPI_HIP_CHECK(hipGetLastError());
        SCOPI_HIP_CHECK(hipMemcpyAsync(h_ctrl_, ctrl_ + cur, sizeof(ApgdCtrl), hipMemcpyDeviceToHost, s));
        SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        const int64_t rounds = h_ctrl_->q_applies;  // one per executed round
        if (timing_) {
            // rounds after `done` exit at their first instruction: only executed ones count
            const int64_t executed = std::min<int64_t>(rounds - before, batch);
            for (int64_t r = 0; r < executed; ++r) {
                float ms;
                SCOPI_HIP_CHECK(hipEventElapsedTime(&ms, ev[4 * r + 0], ev[4 * r + 1]));
                t_trial_ms_ += ms;
                SCOPI_HIP_CHECK(hipEventElapsedTime(&ms, ev[4 * r + 2], ev[4 * r + 3]));
                t_bt_ms_ += ms;
            }
            n_trial_ += executed;
            n_bt_ += executed;
        }
        before = rounds;
        if (h_ctrl_->done) break;
        const int64_t left = predicted - rounds;
        if (predict && !past && left >= 8) {
            batch = (int)std::min<int64_t>(left / 2, 1024);
        } else if (predict && !past) {
            past = true;
            batch = 4;
        } else {
            batch = std::min(batch * 2, 64);
        }
    }
    last_rounds_[sizeof(T) == 4 ? 1 : 0] = h_ctrl_->q_applies;
    ctrl_cur_ = cur;
    for (auto& e : ev) (void)hipEventDestroy(e);
    info.iterations += h_ctrl_->iter;
    info.backtracks += h_ctrl_->backtracks;
    info.q_applies += h_ctrl_->q_applies;
    info.status = h_ctrl_->status == kStatusStalled ? 0 : h_ctrl_->status;
    info.converged = info.status == 0 ? 1 : 0;
    info.residual = h_ctrl_->res;
    info.step = 1.0 / h_ctrl_->L_acc;
    final_slot_ = h_ctrl_->cur;
}

template <int LAY>
void Solver::assemble(const Particles& P, const Contacts& C, int64_t nc, double dt, bool fp32, hipStream_t s) {
    const int64_t na = P.n - P.n_inactive, ni = P.n_inactive;
    // assembly: M^-1, B, c
    hipLaunchKernelGGL((k_minv<double, LAY>), dim3(grid_for(na)), dim3(kBlock), 0, s, P, b64_.minv);
    if (fp32) hipLaunchKernelGGL((k_minv<float, LAY>), dim3(grid_for(na)), dim3(kBlock), 0, s, P, b32_.minv);
    BMatrix<float> b32 = fp32 ? b32_ : BMatrix<float>{};
    const int w32 = fp32 ? 1 : 0;
    if (P.dim == 3)
        hipLaunchKernelGGL((k_assemble<3, LAY>), dim3(grid_for(nc)), dim3(kBlock), 0, s, P, C, nc, dt, b64_, b32, w32);
    else
        hipLaunchKernelGGL((k_assemble<2, LAY>), dim3(grid_for(nc)), dim3(kBlock), 0, s, P, C, nc, dt, b64_, b32, w32);
    // particle-major transpose
    for (int* a : {istart_, iend_, jstart_, jend_})
        hipLaunchKernelGGL(k_fill_i, dim3(256), dim3(kBlock), 0, s, a, na, 0);
    hipLaunchKernelGGL(k_irange, dim3(grid_for(nc)), dim3(kBlock), 0, s, C.i, nc, ni, istart_, iend_);
    hipLaunchKernelGGL(k_jkeys, dim3(grid_for(nc)), dim3(kBlock), 0, s, C.j, nc, ni, na, jkeys_, jlist_);
    int bits = 1;
    while ((int64_t(1) << bits) <= na) ++bits;  // keys 0 .. na
    sorter_.sort(jkeys_, jlist_, nc, bits, s, scanner_);
    hipLaunchKernelGGL(k_jrange, dim3(grid_for(nc)), dim3(kBlock), 0, s, jkeys_, nc, jstart_, jend_);
    hipLaunchKernelGGL(k_tcount, dim3(grid_for(na + 1)), dim3(kBlock), 0, s, istart_, iend_, jstart_, jend_, na,
                       tcount_);
    scanner_.exclusive(tcount_, tstart_, na + 1, nullptr, s);
    hipLaunchKernelGGL((k_transpose<double, LAY>), dim3(grid_for(na * kTransposeGroup)), dim3(kBlock), 0, s, b64_, na,
                       istart_, iend_, jstart_, jend_, jlist_, tstart_, tc_);
    if (fp32)
        hipLaunchKernelGGL((k_transpose<float, LAY>), dim3(grid_for(na * kTransposeGroup)), dim3(kBlock), 0, s, b32_, na,
                           istart_, iend_, jstart_, jend_, jlist_, tstart_, tc_);
    SCOPI_HIP_CHECK(hipGetLastError());
}

template <int LAY>
void Solver::solve_layout(const Particles& P, const Contacts& C, int64_t nc, const scopi_hip_params& prm, double dt,
                          hipStream_t s, scopi_hip_solve_info& info) {
    const bool fp32 = prm.precision == SCOPI_PRECISION_FP32_REFINE;
    assemble<LAY>(P, C, nc, dt, fp32, s);
    solve_phases<LAY>(P, C, nc, prm, dt, s, info);
}

// The projected-gradient solve on the assembled B and c, then u = v + dt M^-1 B^T lambda.
template <int LAY>
void Solver::solve_phases(const Particles& P, const Contacts& C, int64_t nc, const scopi_hip_params& prm, double dt,
                          hipStream_t s, scopi_hip_solve_info& info) {
    constexpr int NR = lay<LAY>::NR, NW = lay<LAY>::NW;
    const int64_t na = P.n - P.n_inactive;
    const bool fp32 = prm.precision == SCOPI_PRECISION_FP32_REFINE;
    if (fp32) {
        // FP32 phase to a relative stop or its precision floor, then FP64 refinement
        // warm-started from it with the last accepted L
        const double rtol32 = prm.fp32_tolerance > 0 ? 0.0 : 1e-5;
        const double tol32 = prm.fp32_tolerance > 0 ? prm.fp32_tolerance : prm.tolerance;
        const int64_t before = info.iterations;
        if (lam0_)
            hipLaunchKernelGGL((k_convert<double, float>), dim3(1024), dim3(kBlock), 0, s, lam0_, b32_.lam, NR * nc);
        run_phase<float, LAY>(b32_, P, C, na, nc, prm, dt, tol32, rtol32, prm.max_ite, lam0_ != nullptr, false, s,
                              info);
        const int64_t it32 = info.iterations - before;
        info.iterations_fp32 += it32;
        const int64_t stride32 = NR * b32_.cap;
        hipLaunchKernelGGL((k_convert<float, double>), dim3(1024), dim3(kBlock), 0, s,
                           b32_.lam + (int64_t)final_slot_ * stride32, b64_.lam, NR * nc);
        const int64_t left = std::max<int64_t>(prm.max_ite - it32, 1);
        run_phase<double, LAY>(b64_, P, C, na, nc, prm, dt, prm.tolerance, 0.0, left, true, true, s, info);
    } else {
        if (lam0_)
            SCOPI_HIP_CHECK(hipMemcpyAsync(b64_.lam, lam0_, NR * nc * sizeof(double), hipMemcpyDeviceToDevice, s));
        run_phase<double, LAY>(b64_, P, C, na, nc, prm, dt, prm.tolerance, 0.0, prm.max_ite, lam0_ != nullptr, false,
                               s, info);
    }
    // velocities from the final lambda (FP64)
    launch_bt<double, LAY>(bt_group_, s, (ApgdCtrl*)nullptr, final_slot_, b64_, na, tstart_, tc_, nullptr, 0, nullptr);
    hipLaunchKernelGGL(k_velocities<NW>, dim3(grid_for(na)), dim3(kBlock), 0, s, P,
                       (const double*)slot(b64_.w, final_slot_, 6 * b64_.pcap), pcap_, dt, 1, u_, w_);
    SCOPI_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ fixed points on s_ij
// FrictionFixedPoint: c = c0 + S with S = s_ij n_ij on the contact's first dim rows
// (LagrangeMultiplier<dim, FrictionFixedPoint> S_Vector, lagrange_multiplier.hpp:290-297; the
// gradient and the objective of minimization_problem.hpp:130-141 add S to Q lambda + c and
// lambda . S to the objective, the same as solving with c + S).
// ViscousFriction (kVisc): S = dt mu_ij s_ij n_ij on the cone rows of the contacts of kind 2
// (lagrange_multiplier.hpp:379-400).
template <int LAY>
__global__ void __launch_bounds__(kBlock) k_apply_s(BMatrix<double> B64, BMatrix<float> B32, int with32,
                                                   const double* __restrict__ c0, const double* __restrict__ sij,
                                                   int64_t nc, int dim, double dt) {
#pragma clang fp contract(off)
    constexpr int NR = lay<LAY>::NR, NC = lay<LAY>::NC;
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    const bool k2 = LAY == kVisc && (B64.kind[c] & kKindMask) == 2;
    for (int d = 0; d < NR; ++d) {
        double v = c0[NR * c + d];
        if (LAY == kCone && d < dim) v = c0[NR * c + d] + sij[c] * B64.coef[c * NC + d];
        if (k2 && d >= 1 && d <= dim) v = c0[NR * c + d] + dt * B64.mu[c] * sij[c] * B64.coef[c * NC + d - 1];
        B64.c[NR * c + d] = v;
        if (with32) B32.c[NR * c + d] = (float)v;
    }
}

// extra_steps_before_solve (OptimGradient.hpp:86-106): s_ij = mu_ij dt (FrictionFixedPoint), or
// 1 on the contacts at gamma == gamma_min (ViscousFriction; the others take no part)
__global__ void __launch_bounds__(kBlock) k_init_s(double* __restrict__ sij, const double* __restrict__ mu,
                                                  const int32_t* __restrict__ kind, int64_t nc, double dt) {
#pragma clang fp contract(off)
    for (int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x; c < nc; c += (int64_t)gridDim.x * kBlock)
        sij[c] = kind ? ((kind[c] & kAtGammaMin) ? 1.0 : 0.0) : mu[c] * dt;
}

// extra_steps_after_solve (OptimGradient.hpp:116-240): A U at every contact from the solve's
// velocities (the rows of B: e_d . (u_i - u_j) + K_i[d] . w_i - K_j[d] . w_j), its tangential
// part T, s_ij = mu_ij dt |T| (FrictionFixedPoint) or |T| on the contacts at gamma == gamma_min
// (ViscousFriction, `kind` given), over the first dim components; block sums of
// |s_new - s_old|^2 and |s_new|^2 (the other ViscousFriction contacts add 0), and mu of
// contact 0 (the FrictionFixedPoint error's scale) after the sums.
__global__ void __launch_bounds__(kBlock) k_fixed_point_s(BMatrix<double> B, const int32_t* __restrict__ ci,
                                                         const int32_t* __restrict__ cj, int64_t ni, int64_t nc,
                                                         const double* __restrict__ u, const double* __restrict__ w,
                                                         int64_t pcap, double dt, int dim,
                                                         const int32_t* __restrict__ kind, double* __restrict__ sij,
                                                         double* __restrict__ partial) {
#pragma clang fp contract(off)
    __shared__ double lds[2][kBlock / 64];
    double acc[2] = {0.0, 0.0};
    for (int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x; c < nc; c += (int64_t)gridDim.x * kBlock) {
        if (kind && !(kind[c] & kAtGammaMin)) continue;
        const int64_t i = (int64_t)ci[c] - ni, j = (int64_t)cj[c] - ni;
        const double* co = B.coef + c * 21;
        double au[3];
        for (int d = 0; d < 3; ++d) {
            double a = 0.0;
            if (i >= 0) {
                a += u[d * pcap + i];
                a += co[3 + 3 * d] * w[i] + co[4 + 3 * d] * w[pcap + i] + co[5 + 3 * d] * w[2 * pcap + i];
            }
            if (j >= 0) {
                a -= u[d * pcap + j];
                a -= co[12 + 3 * d] * w[j] + co[13 + 3 * d] * w[pcap + j] + co[14 + 3 * d] * w[2 * pcap + j];
            }
            au[d] = a;
        }
        double an = 0.0;
        for (int d = 0; d < dim; ++d) an += au[d] * co[d];
        double t2 = 0.0;
        for (int d = 0; d < dim; ++d) {
            const double t = au[d] - co[d] * an;
            t2 += t * t;
        }
        const double s_new = kind ? sqrt(t2) : B.mu[c] * dt * sqrt(t2);
        const double ds = s_new - sij[c];
        sij[c] = s_new;
        acc[0] += ds * ds;
        acc[1] += s_new * s_new;
    }
    block_sum<2>(acc, lds);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = acc[0];
        partial[2 * blockIdx.x + 1] = acc[1];
        if (blockIdx.x == 0) partial[2 * gridDim.x] = B.mu[0];
    }
}

// Viscous problems: the kind word of every contact (kVisc) and the number of multiplier rows
// (LagrangeMultiplier<Viscous> :125-141, <ViscousFriction> :357-378)
__global__ void __launch_bounds__(kBlock) k_visc_kinds(const double* __restrict__ gamma, int64_t nc, double gmin,
                                                      double gtol, int with_friction, int32_t* __restrict__ kind,
                                                      unsigned long long* __restrict__ rows) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    unsigned long long r = 0;
    if (c < nc) {
        const double g = gamma[c];
        int kd = 0;
        r = 1;
        if (g < -gtol) {
            kd = (with_friction && g == gmin) ? 2 : 1;
            r = (kd == 2) ? 4 : 2;
        }
        kind[c] = kd | ((g == gmin) ? kAtGammaMin : 0);
    }
    for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o);
    if ((threadIdx.x & 63) == 0 && r) atomicAdd(rows, r);
}

// update_contact_properties (OptimGradient.hpp:19-54): gamma = min(max(gamma - dt n . lambda_global,
// gamma_min), 0), lambda_global = local2global of the final multipliers
__global__ void __launch_bounds__(kBlock) k_visc_gamma(BMatrix<double> B, const double* __restrict__ lam, int64_t nc,
                                                      int dim, double dt, double gmin, double* __restrict__ gamma) {
#pragma clang fp contract(off)
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    double g[3];
    visc_global<double>(B, lam, c, g[0], g[1], g[2]);
    double dn = 0.0;
    for (int d = 0; d < dim; ++d) dn += B.coef[c * 21 + d] * g[d];
    const double v = gamma[c] - dt * dn;
    const double m = (v < gmin) ? gmin : v;  // std::max(v, gamma_min)
    gamma[c] = (0.0 < m) ? 0.0 : m;          // std::min(., 0.)
}

template <int LAY>
void Solver::solve_fixed_point(const Particles& P, const Contacts& C, int64_t nc, const scopi_hip_params& prm,
                               double dt, hipStream_t s, scopi_hip_solve_info& info) {
    constexpr int NR = lay<LAY>::NR;
    const int64_t ni = P.n_inactive;
    const bool fp32 = prm.precision == SCOPI_PRECISION_FP32_REFINE;
    const bool visc = LAY == kVisc;  // ViscousFriction, else FrictionFixedPoint
    assemble<LAY>(P, C, nc, dt, fp32, s);
    if (nc > fp_cap_) {
        for (void* p : {(void*)fp_c0_, (void*)fp_s_})
            if (p) SCOPI_HIP_CHECK(hipFree(p));
        fp_cap_ = nc + nc / 4 + 1024;
        SCOPI_HIP_CHECK(hipMalloc(&fp_c0_, 4 * fp_cap_ * sizeof(double)));
        SCOPI_HIP_CHECK(hipMalloc(&fp_s_, fp_cap_ * sizeof(double)));
    }
    if (!fp_partial_) {
        SCOPI_HIP_CHECK(hipMalloc(&fp_partial_, (2 * 1024 + 1) * sizeof(double)));
        SCOPI_HIP_CHECK(hipHostMalloc(&h_fp_partial_, (2 * 1024 + 1) * sizeof(double)));
    }
    SCOPI_HIP_CHECK(hipMemcpyAsync(fp_c0_, b64_.c, NR * nc * sizeof(double), hipMemcpyDeviceToDevice, s));
    const int32_t* kind = visc ? (const int32_t*)vkind_ : nullptr;
    hipLaunchKernelGGL(k_init_s, dim3(256), dim3(kBlock), 0, s, fp_s_, (const double*)b64_.mu, kind, nc, dt);
    const unsigned gp = (unsigned)std::min<int64_t>(grid_for(nc), 1024);
    BMatrix<float> b32 = fp32 ? b32_ : BMatrix<float>{};
    info.outer_iterations = 0;
    for (;;) {
        hipLaunchKernelGGL(k_apply_s<LAY>, dim3(grid_for(nc)), dim3(kBlock), 0, s, b64_, b32, fp32 ? 1 : 0,
                           (const double*)fp_c0_, (const double*)fp_s_, nc, P.dim, dt);
        solve_phases<LAY>(P, C, nc, prm, dt, s, info);
        info.outer_iterations += 1;
        hipLaunchKernelGGL(k_fixed_point_s, dim3(gp), dim3(kBlock), 0, s, b64_, (const int32_t*)C.i,
                           (const int32_t*)C.j, ni, nc, (const double*)u_, (const double*)w_, pcap_, dt, P.dim,
                           kind, fp_s_, fp_partial_);
        SCOPI_HIP_CHECK(
            hipMemcpyAsync(h_fp_partial_, fp_partial_, (2 * gp + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
        SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        double ds2 = 0.0, s2 = 0.0;
        for (unsigned b = 0; b < gp; ++b) {
            ds2 += h_fp_partial_[2 * b];
            s2 += h_fp_partial_[2 * b + 1];
        }
        // FrictionFixedPoint: err = (1 / (mu_0 dt)) |s_new - s_old| / (1 + (1 / (mu_0 dt)) |s_new|)
        // (OptimGradient.hpp:157-158); ViscousFriction: |s_new - s_old| / (1 + |s_new|) (:216)
        const double mudt = h_fp_partial_[2 * gp] * dt;
        const double err = visc ? std::sqrt(ds2) / (1 + std::sqrt(s2))
                                : (1 / mudt) * std::sqrt(ds2) / (1 + (1 / mudt) * std::sqrt(s2));
        info.fixed_point_error = err;
        if (err < prm.fixed_point_tol || info.outer_iterations >= prm.fixed_point_max_iter) break;
    }
}

int Solver::set_tuning(int key, int value) {
    if (key == SCOPI_TUNE_REDUCE_BLOCKS) {
        if (value < 1 || value > kRedBlocksMax) return SCOPI_HIP_ERR_ARG;
        reduce_blocks_ = value;
        return SCOPI_HIP_OK;
    }
    if (key == SCOPI_TUNE_GRAPH_ROUNDS) {
        if (value < 0 || value > 255 || (value > 0 && value % 2 == 0)) return SCOPI_HIP_ERR_ARG;
        graph_rounds_ = value;
        return SCOPI_HIP_OK;
    }
    if (key == SCOPI_TUNE_BT_GROUP) {
        if (value != 1 && value != 2 && value != 4 && value != 8 && value != 16) return SCOPI_HIP_ERR_ARG;
        bt_group_ = value;
        return SCOPI_HIP_OK;
    }
    if (key == SCOPI_TUNE_ROW_ITERS) {
        if (value < 1 || value > 64) return SCOPI_HIP_ERR_ARG;
        row_iters_ = value;
        return SCOPI_HIP_OK;
    }
    if (key == SCOPI_TUNE_QUEUE_MAX_CONTACTS) {
        if (value < 0) return SCOPI_HIP_ERR_ARG;
        queue_max_contacts_ = value;
        return SCOPI_HIP_OK;
    }
    if (key == SCOPI_TUNE_QUEUE_BLOCKS) {
        if (value < 1 || value > 65536) return SCOPI_HIP_ERR_ARG;
        queue_blocks_ = value;
        return SCOPI_HIP_OK;
    }
    if (key == SCOPI_TUNE_QUEUE_THREADS) {
        if (value != 256 && value != 512 && value != 1024) return SCOPI_HIP_ERR_ARG;
        queue_threads_ = value;
        return SCOPI_HIP_OK;
    }
    return SCOPI_HIP_ERR_ARG;
}

int64_t Solver::viscous_kinds(const double* gamma, int64_t nc, const scopi_hip_params& prm, hipStream_t s) {
    if (nc > vkind_cap_) {
        if (vkind_) SCOPI_HIP_CHECK(hipFree(vkind_));
        vkind_cap_ = nc + nc / 4 + 1024;
        SCOPI_HIP_CHECK(hipMalloc(&vkind_, vkind_cap_ * sizeof(int32_t)));
    }
    if (!vrows_) {
        SCOPI_HIP_CHECK(hipMalloc(&vrows_, sizeof(unsigned long long)));
        SCOPI_HIP_CHECK(hipHostMalloc(&h_vrows_, sizeof(unsigned long long)));
    }
    hipLaunchKernelGGL(k_zero<unsigned long long>, dim3(1), dim3(kBlock), 0, s, vrows_, (int64_t)1);
    hipLaunchKernelGGL(k_visc_kinds, dim3(grid_for(nc)), dim3(kBlock), 0, s, gamma, nc, prm.gamma_min, prm.gamma_tol,
                       prm.problem == SCOPI_PROBLEM_VISCOUS_FRICTION ? 1 : 0, vkind_, vrows_);
    SCOPI_HIP_CHECK(hipMemcpyAsync(h_vrows_, vrows_, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    b64_.kind = vkind_;
    b32_.kind = vkind_;
    return (int64_t)*h_vrows_;
}

int Solver::solve(const Particles& P, const Contacts& C, int64_t nc, const scopi_hip_params& prm, double dt,
                  bool rot_free, hipStream_t s, scopi_hip_solve_info* info, double* gamma, const double* lam0) {
    const int64_t na = P.n - P.n_inactive;
    const bool fixed_point = prm.problem == SCOPI_PROBLEM_FRICTION_FIXED_POINT;
    const bool friction = prm.problem == SCOPI_PROBLEM_FRICTION || fixed_point;
    const bool viscous = prm.problem == SCOPI_PROBLEM_VISCOUS || prm.problem == SCOPI_PROBLEM_VISCOUS_FRICTION;
    const bool fp32 = prm.precision == SCOPI_PRECISION_FP32_REFINE;
    if (prm.problem != SCOPI_PROBLEM_NOFRICTION && !friction && !viscous) return SCOPI_HIP_ERR_UNSUPPORTED;
    if (prm.method != SCOPI_METHOD_APGD && prm.method != SCOPI_METHOD_PGD) return SCOPI_HIP_ERR_UNSUPPORTED;
    if (!(prm.alpha > 0)) return SCOPI_HIP_ERR_ARG;
    if (fixed_point && !(dt > 0)) return SCOPI_HIP_ERR_ARG;
    if (viscous && nc > 0 && !gamma) return SCOPI_HIP_ERR_ARG;
    if (lam0 && (viscous || fixed_point)) return SCOPI_HIP_ERR_UNSUPPORTED;
    reserve(na, nc, fp32);
    lam0_ = lam0;
    struct Clear {
        const double*& p;
        ~Clear() { p = nullptr; }
    } clear_lam0{lam0_};
    layout_ = viscous ? kVisc : friction ? kCone : ((rot_free && !prm.general_jacobian) ? kRotFree : kRigid);
    nr_ = viscous ? 4 : friction ? 3 : 1;
    last_problem_ = prm.problem;
    nc_ = nc;
    rows_ = nr_ * nc;
    scopi_hip_solve_info loc{};
    loc.n_contacts = nc;
    loc.converged = 1;
    if (viscous && nc > 0) rows_ = viscous_kinds(gamma, nc, prm, s);
    loc.rows = (nc == 0) ? 0 : rows_;
    if (nc == 0 || na == 0) {
        hipLaunchKernelGGL(k_velocities<6>, dim3(grid_for(na)), dim3(kBlock), 0, s, P, (const double*)nullptr, pcap_, dt,
                           0, u_, w_);
        SCOPI_HIP_CHECK(hipGetLastError());
        has_solution_ = true;
        if (info) *info = loc;
        return SCOPI_HIP_OK;
    }
    if (fixed_point)
        solve_fixed_point<kCone>(P, C, nc, prm, dt, s, loc);
    else if (prm.problem == SCOPI_PROBLEM_VISCOUS_FRICTION)
        solve_fixed_point<kVisc>(P, C, nc, prm, dt, s, loc);
    else if (viscous)
        solve_layout<kVisc>(P, C, nc, prm, dt, s, loc);
    else if (layout_ == kCone)
        solve_layout<kCone>(P, C, nc, prm, dt, s, loc);
    else if (layout_ == kRigid)
        solve_layout<kRigid>(P, C, nc, prm, dt, s, loc);
    else
        solve_layout<kRotFree>(P, C, nc, prm, dt, s, loc);
    if (viscous) {
        const double* lam = b64_.lam + (int64_t)final_slot_ * nr_ * b64_.cap;
        hipLaunchKernelGGL(k_visc_gamma, dim3(grid_for(nc)), dim3(kBlock), 0, s, b64_, lam, nc, P.dim, dt,
                           prm.gamma_min, gamma);
        SCOPI_HIP_CHECK(hipGetLastError());
    }
    has_solution_ = true;
    if (info) *info = loc;
    return SCOPI_HIP_OK;
}

// ------------------------------------------------------------------ scenes
void Solver::scene_tables(const std::vector<int>& scenes, hipStream_t s) {
    const int K = scenes_->nscene;
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));  // the staging buffer may still feed a queued copy
    int64_t R = 0, Bn = 0;
    for (int sc : scenes) {
        R += scene_rblocks_[sc];
        Bn += scene_bblocks_[sc];
    }
    const int64_t need = R + Bn + 4 * (int64_t)K + (int64_t)scenes.size();
    if (need > tab_cap_) {
        if (d_tab_) SCOPI_HIP_CHECK(hipFree(d_tab_));
        if (h_tab_) SCOPI_HIP_CHECK(hipHostFree(h_tab_));
        tab_cap_ = need + need / 2 + 1024;
        SCOPI_HIP_CHECK(hipMalloc(&d_tab_, tab_cap_ * sizeof(int32_t)));
        SCOPI_HIP_CHECK(hipHostMalloc(&h_tab_, tab_cap_ * sizeof(int32_t)));
    }
    int32_t* rblk = h_tab_;
    int32_t* bblk = rblk + R;
    int32_t* rfirst = bblk + Bn;
    int32_t* rcount = rfirst + K;
    int32_t* bfirst = rcount + K;
    int32_t* bcount = bfirst + K;
    int32_t* slist = bcount + K;
    for (int k = 0; k < K; ++k) rfirst[k] = rcount[k] = bfirst[k] = bcount[k] = 0;
    int64_t r = 0, b = 0;
    for (size_t k = 0; k < scenes.size(); ++k) {
        const int sc = scenes[k];
        rfirst[sc] = (int32_t)r;
        rcount[sc] = scene_rblocks_[sc];
        for (int m = 0; m < scene_rblocks_[sc]; ++m) rblk[r++] = sc;
        bfirst[sc] = (int32_t)b;
        bcount[sc] = scene_bblocks_[sc];
        for (int m = 0; m < scene_bblocks_[sc]; ++m) bblk[b++] = sc;
        slist[k] = sc;
    }
    SCOPI_HIP_CHECK(hipMemcpyAsync(d_tab_, h_tab_, need * sizeof(int32_t), hipMemcpyHostToDevice, s));
    map_.rblk = d_tab_;
    map_.bblk = d_tab_ + R;
    map_.rfirst = d_tab_ + R + Bn;
    map_.rcount = map_.rfirst + K;
    map_.bfirst = map_.rcount + K;
    map_.bcount = map_.bfirst + K;
    map_.slist = map_.bcount + K;
    rgrid_ = (int)R;
    bgrid_ = (int)Bn;
    nlist_ = (int)scenes.size();
}

template <class T, int LAY>
void Solver::run_phase_scenes(BMatrix<T>& B, const Contacts& C, int64_t na, int64_t nc, const scopi_hip_params& prm,
                              double dt, double tol, double rtol, int64_t max_ite, bool warm, bool keep_L,
                              hipStream_t s, const std::vector<int>& scenes, std::vector<scopi_hip_solve_info>& info) {
    constexpr int NR = lay<LAY>::NR;
    const int K = scenes_->nscene;
    const double dt2 = dt * dt;
    const int32_t* ci = (const int32_t*)C.i;
    const int32_t* cj = (const int32_t*)C.j;
    hipLaunchKernelGGL(k_ctrl_init_scenes, dim3(grid_for(K, 64)), dim3(64), 0, s, mctrl_, (const int32_t*)d_buf_, K,
                       prm.alpha, 1. / prm.alpha, tol, rtol, dt2, max_ite, prm.method, prm.dynamic_descent ? 1 : 0,
                       keep_L ? 1 : 0, sizeof(T) == 4 ? 1 : 0, scenes_->d_crange, d_done_);
    SCOPI_HIP_CHECK(hipGetLastError());
    std::vector<int> live(scenes), buf(K, 0);
    scene_tables(live, s);
    // slots as in the one-scene loop, for these scenes only: lambda_0 and w_0 in slot 0, lambda_{-1}
    // and a finite Q lambda_{-1} in slot 1 (beta = 0 in every scene's first round); warm, every
    // scene's lambda_0 is in slot 0 and w_0 comes from it
    hipLaunchKernelGGL((k_scene_setup<T, NR, lay<LAY>::NW>), dim3(rgrid_), dim3(kBlock), 0, s, B, warm ? 1 : 0, map_);
    if (warm) {
        launch_bt<T, LAY>(bt_group_, s, (ApgdCtrl*)nullptr, 0, B, na, tstart_, tc_, nullptr, 0, nullptr, nullptr,
                          nullptr, map_, bgrid_);
        for (int sc : scenes) info[sc].q_applies += 1;
    }
    int batch = 2;
    int cur = 0;  // ctrl copy (K scenes) holding the newest state of the scenes in the grid
    for (;;) {
        for (int r = 0; r < batch; ++r) {
            hipLaunchKernelGGL((k_trial<T, LAY>), dim3(rgrid_), dim3(kBlock), 0, s,
                               (const ApgdCtrl*)(mctrl_ + (int64_t)cur * K), mctrl_ + (int64_t)(1 - cur) * K, B, nc,
                               ci, cj, p_inactive_, partial_trial_, (const double*)partial_bt_, 0, map_);
            cur = 1 - cur;
            launch_bt<T, LAY>(bt_group_, s, mctrl_ + (int64_t)cur * K, -1, B, na, tstart_, tc_, partial_trial_, 0,
                              partial_bt_, nullptr, nullptr, map_, bgrid_);
        }
        hipLaunchKernelGGL(k_ctrl, dim3(nlist_), dim3(kBlock), 0, s, (const ApgdCtrl*)(mctrl_ + (int64_t)cur * K),
                           mctrl_ + (int64_t)(1 - cur) * K, (const double*)partial_bt_, 0, map_);
        cur = 1 - cur;
        SCOPI_HIP_CHECK(hipGetLastError());
        SCOPI_HIP_CHECK(hipMemcpyAsync(h_done_, d_done_, K * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        std::vector<int> next;
        next.reserve(live.size());
        for (int sc : live) {
            if (h_done_[sc])
                buf[sc] = cur;
            else
                next.push_back(sc);
        }
        if (next.empty()) break;
        if (next.size() != live.size()) {
            live.swap(next);
            scene_tables(live, s);
        }
        batch = std::min(batch * 2, 64);
    }
    // final states: scene sc in copy buf[sc]
    scene_buf_ = buf;
    SCOPI_HIP_CHECK(hipMemcpyAsync(d_buf_, scene_buf_.data(), K * sizeof(int32_t), hipMemcpyHostToDevice, s));
    SCOPI_HIP_CHECK(hipMemcpyAsync(h_ctrls_, mctrl_, 2 * (int64_t)K * sizeof(ApgdCtrl), hipMemcpyDeviceToHost, s));
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    for (int sc : scenes) {
        const ApgdCtrl& c = h_ctrls_[(int64_t)buf[sc] * K + sc];
        scopi_hip_solve_info& in = info[sc];
        in.iterations += c.iter;
        in.backtracks += c.backtracks;
        in.q_applies += c.q_applies;
        in.status = c.status == kStatusStalled ? 0 : c.status;
        in.converged = in.status == 0 ? 1 : 0;
        in.residual = c.res;
        in.step = 1.0 / c.L_acc;
    }
}

template <int LAY>
void Solver::solve_layout_scenes(const Particles& P, const Contacts& C, int64_t nc, const scopi_hip_params& prm,
                                 double dt, hipStream_t s, std::vector<scopi_hip_solve_info>& info) {
    constexpr int NR = lay<LAY>::NR, NW = lay<LAY>::NW;
    const int K = scenes_->nscene;
    const int64_t na = P.n - P.n_inactive;
    const bool fp32 = prm.precision == SCOPI_PRECISION_FP32_REFINE;
    assemble<LAY>(P, C, nc, dt, fp32, s);
    // scenes of at most queue_max_contacts_ contacts go to the work-queue kernel, the others take
    // the two-launch rounds
    std::vector<int> all(K), big, small;
    std::vector<char> queued(K, 0);
    for (int k = 0; k < K; ++k) {
        all[k] = k;
        queued[k] = queue_max_contacts_ > 0 && scene_nc_[k] <= queue_max_contacts_;
        (queued[k] ? small : big).push_back(k);
    }
    // per scene, the slot holding the final lambda of the last phase (a queued scene's is slot 0)
    std::vector<int32_t> sl(K);
    auto slots_to_device = [&](int32_t* dev) {
        for (int k = 0; k < K; ++k) sl[k] = queued[k] ? 0 : h_ctrls_[(int64_t)scene_buf_[k] * K + k].cur;
        SCOPI_HIP_CHECK(hipMemcpyAsync(dev, sl.data(), K * sizeof(int32_t), hipMemcpyHostToDevice, s));
        SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    };
    int32_t* d_slot = d_done_;  // reused as the per-scene slot list between phases
    if (fp32 && lam0_)
        hipLaunchKernelGGL((k_convert<double, float>), dim3(1024), dim3(kBlock), 0, s, lam0_, b32_.lam, NR * nc);
    if (!fp32 && lam0_)
        SCOPI_HIP_CHECK(hipMemcpyAsync(b64_.lam, lam0_, NR * nc * sizeof(double), hipMemcpyDeviceToDevice, s));
    // The queued scenes run on a second stream, concurrently with the rounds of the others: the
    // latency-bound work-queue launch overlaps the bandwidth-bound rounds.  From here on each side
    // writes only its own scenes' rows and particles (the rounds' set-up is per scene), and both
    // only read B, c, M^-1 and the particle-major copy.
    if (!small.empty()) queue_launch<LAY>(C, prm, dt, s, small);
    if (fp32) {
        const double rtol32 = prm.fp32_tolerance > 0 ? 0.0 : 1e-5;
        const double tol32 = prm.fp32_tolerance > 0 ? prm.fp32_tolerance : prm.tolerance;
        if (!big.empty()) {
            run_phase_scenes<float, LAY>(b32_, C, na, nc, prm, dt, tol32, rtol32, prm.max_ite, lam0_ != nullptr, false,
                                         s, big, info);
            for (int k : big) info[k].iterations_fp32 = info[k].iterations;
            slots_to_device(d_slot);
            scene_tables(big, s);
            hipLaunchKernelGGL((k_scene_slot<float, double, NR>), dim3(rgrid_), dim3(kBlock), 0, s,
                               (const float*)b32_.lam, NR * b32_.cap, b64_.lam, (const int32_t*)d_slot, map_);
            run_phase_scenes<double, LAY>(b64_, C, na, nc, prm, dt, prm.tolerance, 0.0, prm.max_ite, true, true, s,
                                          big, info);
        }
    } else if (!big.empty()) {
        run_phase_scenes<double, LAY>(b64_, C, na, nc, prm, dt, prm.tolerance, 0.0, prm.max_ite, lam0_ != nullptr,
                                      false, s, big, info);
    }
    if (!small.empty()) queue_join(s, small, info);
    // every scene's final lambda to slot 0, then the velocities (FP64)
    slots_to_device(d_slot);
    scene_tables(all, s);
    hipLaunchKernelGGL((k_scene_slot<double, double, NR>), dim3(rgrid_), dim3(kBlock), 0, s, (const double*)b64_.lam,
                       NR * b64_.cap, b64_.lam, (const int32_t*)d_slot, map_);
    final_slot_ = 0;
    launch_bt<double, LAY>(bt_group_, s, (ApgdCtrl*)nullptr, 0, b64_, na, tstart_, tc_, nullptr, 0, nullptr);
    hipLaunchKernelGGL(k_velocities<NW>, dim3(grid_for(na)), dim3(kBlock), 0, s, P, (const double*)b64_.w, pcap_, dt, 1,
                       u_, w_);
    SCOPI_HIP_CHECK(hipGetLastError());
}

template <int LAY>
void Solver::queue_launch(const Contacts& C, const scopi_hip_params& prm, double dt, hipStream_t s,
                          const std::vector<int>& scenes) {
    const int K = scenes_->nscene;
    const int nq = (int)scenes.size();
    if (!qstream_) {
        SCOPI_HIP_CHECK(hipStreamCreateWithFlags(&qstream_, hipStreamNonBlocking));  // default priority
        SCOPI_HIP_CHECK(hipEventCreateWithFlags(&qfork_, hipEventDisableTiming));
        SCOPI_HIP_CHECK(hipEventCreateWithFlags(&qjoin_, hipEventDisableTiming));
        SCOPI_HIP_CHECK(hipMalloc(&d_qnext_, sizeof(int32_t)));
    }
    if (std::max(nq, K) > qcap_) {
        for (void* p : {(void*)d_qorder_, (void*)d_qres_})
            if (p) SCOPI_HIP_CHECK(hipFree(p));
        for (void* p : {(void*)h_qorder_, (void*)h_qres_})
            if (p) SCOPI_HIP_CHECK(hipHostFree(p));
        qcap_ = std::max(nq, K) + 64;
        SCOPI_HIP_CHECK(hipMalloc(&d_qorder_, qcap_ * sizeof(int32_t)));
        SCOPI_HIP_CHECK(hipHostMalloc(&h_qorder_, qcap_ * sizeof(int32_t)));
        SCOPI_HIP_CHECK(hipMalloc(&d_qres_, qcap_ * sizeof(QueueResult)));
        SCOPI_HIP_CHECK(hipHostMalloc(&h_qres_, qcap_ * sizeof(QueueResult)));
    }
    // largest scenes first, so that the last workgroups to finish hold small scenes
    std::vector<int> order(scenes);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return scene_nc_[a] > scene_nc_[b]; });
    SCOPI_HIP_CHECK(hipStreamSynchronize(qstream_));  // the staging buffers may still feed the last copies
    for (int k = 0; k < nq; ++k) h_qorder_[k] = order[k];
    // the queue starts once B, the particle-major copy and the warm start are in place on `s`
    SCOPI_HIP_CHECK(hipEventRecord(qfork_, s));
    SCOPI_HIP_CHECK(hipStreamWaitEvent(qstream_, qfork_, 0));
    SCOPI_HIP_CHECK(hipMemcpyAsync(d_qorder_, h_qorder_, nq * sizeof(int32_t), hipMemcpyHostToDevice, qstream_));
    hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(kBlock), 0, qstream_, d_qnext_, (int64_t)1, 0);
    QueueArgs A;
    A.order = d_qorder_;
    A.nqueue = nq;
    A.next = d_qnext_;
    A.crange = scenes_->d_crange;
    A.arange = d_arange_;
    A.result = d_qres_;
    A.tstart = tstart_;
    A.tc = tc_;
    A.ci = (const int32_t*)C.i;
    A.cj = (const int32_t*)C.j;
    A.ni = p_inactive_;
    A.alpha = prm.alpha;
    A.tol = prm.tolerance;
    A.tol32 = prm.fp32_tolerance > 0 ? prm.fp32_tolerance : prm.tolerance;
    A.rtol32 = prm.fp32_tolerance > 0 ? 0.0 : 1e-5;
    A.dt2 = dt * dt;
    A.max_ite = prm.max_ite;
    A.method = prm.method;
    A.dynamic = prm.dynamic_descent ? 1 : 0;
    A.warm = lam0_ != nullptr ? 1 : 0;
    A.fp32 = prm.precision == SCOPI_PRECISION_FP32_REFINE ? 1 : 0;
    const unsigned grid = (unsigned)std::min(nq, queue_blocks_);
    BMatrix<float> b32 = A.fp32 ? b32_ : BMatrix<float>{};
    if (queue_threads_ == 1024)
        hipLaunchKernelGGL((k_apgd_queue<LAY, 1024, 8>), dim3(grid), dim3(1024), 0, qstream_, A, b64_, b32);
    else if (queue_threads_ == 512)
        hipLaunchKernelGGL((k_apgd_queue<LAY, 512, 8>), dim3(grid), dim3(512), 0, qstream_, A, b64_, b32);
    else
        hipLaunchKernelGGL((k_apgd_queue<LAY, 256, 8>), dim3(grid), dim3(256), 0, qstream_, A, b64_, b32);
    SCOPI_HIP_CHECK(hipGetLastError());
    SCOPI_HIP_CHECK(hipMemcpyAsync(h_qres_, d_qres_, K * sizeof(QueueResult), hipMemcpyDeviceToHost, qstream_));
    SCOPI_HIP_CHECK(hipEventRecord(qjoin_, qstream_));
}

void Solver::queue_join(hipStream_t s, const std::vector<int>& scenes, std::vector<scopi_hip_solve_info>& info) {
    // `s` continues (final slot copy, velocities) after the queue's results; the host reads them
    SCOPI_HIP_CHECK(hipStreamWaitEvent(s, qjoin_, 0));
    SCOPI_HIP_CHECK(hipEventSynchronize(qjoin_));
    for (int sc : scenes) {
        const QueueResult& r = h_qres_[sc];
        scopi_hip_solve_info& in = info[sc];
        in.iterations += r.iter;
        in.iterations_fp32 += r.iter32;
        in.backtracks += r.backtracks;
        in.q_applies += r.q_applies;
        in.status = r.status;
        in.converged = in.status == 0 ? 1 : 0;
        in.residual = r.res;
        in.step = 1.0 / r.L_acc;
    }
}

int Solver::solve_scenes(const Particles& P, const Contacts& C, int64_t nc, const scopi_hip_params& prm, double dt,
                         bool rot_free, hipStream_t s, const SceneSet& S, std::vector<scopi_hip_solve_info>& info,
                         const double* lam0) {
    const int64_t na = P.n - P.n_inactive;
    const int K = S.nscene;
    const bool friction = prm.problem == SCOPI_PROBLEM_FRICTION;
    const bool fp32 = prm.precision == SCOPI_PRECISION_FP32_REFINE;
    if (prm.problem != SCOPI_PROBLEM_NOFRICTION && prm.problem != SCOPI_PROBLEM_FRICTION) return SCOPI_HIP_ERR_UNSUPPORTED;
    if (prm.method != SCOPI_METHOD_APGD && prm.method != SCOPI_METHOD_PGD) return SCOPI_HIP_ERR_UNSUPPORTED;
    if (!(prm.alpha > 0) || K < 1 || (int)S.alo.size() != K || (int)S.ahi.size() != K) return SCOPI_HIP_ERR_ARG;
    reserve(na, nc, fp32);
    lam0_ = lam0;
    struct Clear {
        const double*& p;
        ~Clear() { p = nullptr; }
    } clear_lam0{lam0_};
    layout_ = friction ? kCone : ((rot_free && !prm.general_jacobian) ? kRotFree : kRigid);
    nr_ = friction ? 3 : 1;
    nc_ = nc;
    rows_ = nr_ * nc;
    p_inactive_ = P.n_inactive;
    scenes_ = &S;
    if (K > nscene_cap_) {
        for (void* p : {(void*)d_arange_, (void*)d_done_, (void*)d_buf_, (void*)mctrl_})
            if (p) SCOPI_HIP_CHECK(hipFree(p));
        for (void* p : {(void*)h_done_, (void*)h_ctrls_})
            if (p) SCOPI_HIP_CHECK(hipHostFree(p));
        nscene_cap_ = K + K / 2 + 16;
        SCOPI_HIP_CHECK(hipMalloc(&d_arange_, 2 * (int64_t)nscene_cap_ * sizeof(int64_t)));
        SCOPI_HIP_CHECK(hipMalloc(&d_done_, nscene_cap_ * sizeof(int32_t)));
        SCOPI_HIP_CHECK(hipMalloc(&d_buf_, nscene_cap_ * sizeof(int32_t)));
        SCOPI_HIP_CHECK(hipMalloc(&mctrl_, 2 * (int64_t)nscene_cap_ * sizeof(ApgdCtrl)));
        SCOPI_HIP_CHECK(hipMemset(mctrl_, 0, 2 * (int64_t)nscene_cap_ * sizeof(ApgdCtrl)));
        SCOPI_HIP_CHECK(hipHostMalloc(&h_done_, nscene_cap_ * sizeof(int32_t)));
        SCOPI_HIP_CHECK(hipHostMalloc(&h_ctrls_, 2 * (int64_t)nscene_cap_ * sizeof(ApgdCtrl)));
    }
    // per-scene shares of the grids: the per-contact grid from the scene's contacts,
    // the k_bt grid from its active particles (neither depends on the other scenes)
    std::vector<int64_t> ar(2 * K), hc(2 * K);
    SCOPI_HIP_CHECK(hipMemcpyAsync(hc.data(), S.d_crange, 2 * K * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    scene_rblocks_.assign(K, 1);
    scene_bblocks_.assign(K, 1);
    scene_nc_.assign(K, 0);
    int64_t R = 0, Bn = 0;
    for (int k = 0; k < K; ++k) {
        ar[2 * k] = S.alo[k];
        ar[2 * k + 1] = S.ahi[k];
        const int64_t nck = hc[2 * k + 1] - hc[2 * k];
        scene_nc_[k] = nck;
        scene_rblocks_[k] = (int)row_blocks(nck, reduce_blocks_, layout_, row_iters_);
        scene_bblocks_[k] = (int)std::max<int64_t>(bt_blocks(S.ahi[k] - S.alo[k], bt_group_), 1);
        R += scene_rblocks_[k];
        Bn += scene_bblocks_[k];
    }
    if (R > partial_cap_) {
        SCOPI_HIP_CHECK(hipFree(partial_trial_));
        SCOPI_HIP_CHECK(hipMalloc(&partial_trial_, (R + 1024) * kPartialStride * sizeof(double)));
        partial_cap_ = R + 1024;
    }
    if (Bn > partial_bt_cap_) {
        SCOPI_HIP_CHECK(hipFree(partial_bt_));
        SCOPI_HIP_CHECK(hipMalloc(&partial_bt_, (Bn + 1024) * sizeof(double)));
        partial_bt_cap_ = Bn + 1024;
    }
    SCOPI_HIP_CHECK(hipMemcpyAsync(d_arange_, ar.data(), 2 * K * sizeof(int64_t), hipMemcpyHostToDevice, s));
    map_ = SceneMap{};
    map_.pstride = partial_cap_;
    map_.crange = S.d_crange;
    map_.arange = d_arange_;
    map_.done = d_done_;
    hipLaunchKernelGGL(k_fill_i, dim3(grid_for(K)), dim3(kBlock), 0, s, d_buf_, (int64_t)K, 0);
    info.assign(K, scopi_hip_solve_info{});
    for (int k = 0; k < K; ++k) {
        info[k].n_contacts = hc[2 * k + 1] - hc[2 * k];
        info[k].rows = nr_ * info[k].n_contacts;
        info[k].converged = 1;
    }
    if (nc == 0 || na == 0) {
        hipLaunchKernelGGL(k_velocities<6>, dim3(grid_for(na)), dim3(kBlock), 0, s, P, (const double*)nullptr, pcap_, dt,
                           0, u_, w_);
        SCOPI_HIP_CHECK(hipGetLastError());
        SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        has_solution_ = true;
        scenes_ = nullptr;
        return SCOPI_HIP_OK;
    }
    if (layout_ == kCone)
        solve_layout_scenes<kCone>(P, C, nc, prm, dt, s, info);
    else if (layout_ == kRigid)
        solve_layout_scenes<kRigid>(P, C, nc, prm, dt, s, info);
    else
        solve_layout_scenes<kRotFree>(P, C, nc, prm, dt, s, info);
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    has_solution_ = true;
    scenes_ = nullptr;
    return SCOPI_HIP_OK;
}

void Solver::download_lambda_range(double* host, int64_t lo, int64_t hi, hipStream_t s) const {
    if (!has_solution_ || hi <= lo) return;
    const double* lam = b64_.lam + (int64_t)final_slot_ * nr_ * b64_.cap;
    SCOPI_HIP_CHECK(hipMemcpyAsync(host, lam + nr_ * lo, nr_ * (hi - lo) * sizeof(double), hipMemcpyDeviceToHost, s));
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
}

void Solver::download_lambda(double* host, hipStream_t s) const {
    if (!has_solution_ || nc_ == 0) return;
    const double* lam = b64_.lam + (int64_t)final_slot_ * nr_ * b64_.cap;
    if (layout_ != kVisc) {
        // rows are stored per contact (c * NR + d): the reference's order
        SCOPI_HIP_CHECK(hipMemcpyAsync(host, lam, nr_ * nc_ * sizeof(double), hipMemcpyDeviceToHost, s));
        SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        return;
    }
    // kVisc: 4 slots per contact; the reference's local order is Viscous: every contact's first
    // row, then the second rows in contact order; ViscousFriction: each contact's 1, 2 or 4 rows
    std::vector<double> x(4 * nc_);
    std::vector<int32_t> kind(nc_);
    SCOPI_HIP_CHECK(hipMemcpyAsync(x.data(), lam, 4 * nc_ * sizeof(double), hipMemcpyDeviceToHost, s));
    SCOPI_HIP_CHECK(hipMemcpyAsync(kind.data(), vkind_, nc_ * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    int64_t r = 0;
    if (last_problem_ == SCOPI_PROBLEM_VISCOUS) {
        for (int64_t c = 0; c < nc_; ++c) host[r++] = x[4 * c];
        for (int64_t c = 0; c < nc_; ++c)
            if ((kind[c] & kKindMask) == 1) host[r++] = x[4 * c + 1];
    } else {
        for (int64_t c = 0; c < nc_; ++c) {
            const int kd = kind[c] & kKindMask;
            const int nrow = kd == 0 ? 1 : kd == 1 ? 2 : 4;
            for (int d = 0; d < nrow; ++d) host[r++] = x[4 * c + d];
        }
    }
}

}  // namespace scopi_hip
