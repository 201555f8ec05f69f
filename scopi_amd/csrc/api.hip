// api.hip -- the extern "C" layer (include/scopi_hip.h): one context per scene,
// the per-particle kernels of the vap slot and of the integration, and the
// device-resident time loop (ScopiSolver::run, include/scopi/solver.hpp:266-303).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "apgd.hpp"
#include "contacts.hpp"

using namespace scopi_hip;

struct scopi_hip_ctx {
    int dim = 3;
    int device = 0;
    hipStream_t stream = nullptr;
    Particles P;
    Contacts C;
    int64_t nc = 0;
    bool contacts_valid = false;
    bool all_spheres = true;  // every active particle is a sphere (rotation-free B for NoFriction)
    bool has_se = false;      // superellipsoids present: candidate search + device Newton closest points
    ContactSearch search;
    Solver solver;
    std::string err;
    // BoxDomain (include/scopi/box.hpp): periodic axes get periodic copies in the search
    double box_lo[3] = {0, 0, 0}, box_hi[3] = {0, 0, 0};
    int periodic[3] = {0, 0, 0};
    bool any_periodic() const { return periodic[0] || periodic[1] || periodic[2]; }
    double* proj = nullptr;        // vap_projection: u then w, Na x 3 each
    int64_t proj_n = -1;           // active particles the projection velocities were given for
    Contacts C2;                   // the search's pairs over the particles and their copies
    int* pflags = nullptr;         // periodic copies: flags, offsets, originals, sort buffers
    int* poffs = nullptr;
    int32_t* porig = nullptr;
    int* ptotal = nullptr;
    int* h_ptotal = nullptr;
    uint32_t* pkeys = nullptr;
    int32_t* pvals = nullptr;
    int32_t* pi_of = nullptr;
    int64_t cap_flags = 0, cap_offs = 0, cap_orig = 0, cap_keys = 0, cap_vals = 0, cap_iof = 0;
    RadixSorter psort;
    Scanner pscan;
    // Viscous problems: gamma of the current contacts (valid when gamma_valid), and the contact
    // history of the last step for transfer: its (i, j) keys (sorted) and gamma
    double* gamma = nullptr;
    int64_t gamma_cap = 0;
    bool gamma_valid = false;
    uint64_t* hist_keys = nullptr;
    double* hist_gamma = nullptr;
    int64_t hist_n = 0, hist_cap = 0;
    // warm start: the last solve's multipliers (hist_nr rows per history contact) and the current
    // contacts' first iterate
    double* hist_lam = nullptr;
    int64_t hist_lam_cap = 0;
    int hist_nr = 0;
    int hist_kind = 0;  // what hist_keys index: 0 nothing, 1 hist_lam (hist_nr rows each), 2 hist_gamma
    double* lam0 = nullptr;
    int64_t lam0_cap = 0;
};

namespace {

// vap_fixed (vap/vap_fixed.hpp:46-55) / vap_fpd (vap/vap_fpd.hpp:70-82, src/vap/vap_fpd.cpp:6-26) /
// vap_projection (vap/vap_projection.hpp:34-49: vd = u, omega = w(i, 2) broadcast; pu, pw: Na x 3)
__global__ void __launch_bounds__(kBlock) k_vap(Particles P, int kind, double dt, const double* __restrict__ pu,
                                               const double* __restrict__ pw) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t k = P.n_inactive + t;
    if (k >= P.n) return;
    const int64_t c = P.cap;
    if (kind == SCOPI_VAP_PROJECTION) {
        for (int d = 0; d < P.dim; ++d) P.vd[d * c + k] = pu[3 * t + d];
        if (P.dim == 2)
            P.omega[2 * c + k] = pw[3 * t + 2];
        else
            for (int d = 0; d < 3; ++d) P.omega[d * c + k] = pw[3 * t + 2];
        return;
    }
    if (kind == SCOPI_VAP_FIXED) {
        for (int d = 0; d < 3; ++d) {
            P.v[d * c + k] = P.vd[d * c + k];
            P.omega[d * c + k] = P.domega[d * c + k];
        }
        return;
    }
    for (int d = 0; d < P.dim; ++d) P.v[d * c + k] += dt * P.f[d * c + k] / P.m[k];
    if (P.dim == 3) {
        const double o1 = P.omega[k], o2 = P.omega[c + k], o3 = P.omega[2 * c + k];
        const double j1 = P.J[k], j2 = P.J[c + k], j3 = P.J[2 * c + k];
        const double r0 = o2 * o3 * (j3 - j2), r1 = o1 * o3 * (j1 - j3), r2 = o1 * o2 * (j2 - j1);
        P.omega[k] += r0;
        P.omega[c + k] += r1;
        P.omega[2 * c + k] += r2;
    }
}

// q <- normalize(q * exp(w dt / 2)), mult_quaternion (src/quaternion.cpp:92-120),
// normalize (src/quaternion.cpp:6-10), with the ||w|| = 0 -> 1 guard of solver.hpp:463-470.
__device__ inline void rotate_q(double* q, int64_t c, int64_t k, double wx, double wy, double wz, double normw,
                                double dt) {
#pragma clang fp contract(off)
    if (normw == 0) normw = 1;
    const double e0 = cos(0.5 * normw * dt);
    const double sn = sin(0.5 * normw * dt) / normw;
    const double e[4] = {e0, sn * wx, sn * wy, sn * wz};
    const double q1[4] = {q[k], q[c + k], q[2 * c + k], q[3 * c + k]};
    double r[4];
    r[0] = q1[0] * e[0];
    for (int d = 0; d < 3; ++d) {
        r[0] -= q1[d + 1] * e[d + 1];
        r[d + 1] = q1[0] * e[d + 1] + e[0] * q1[d + 1];
    }
    r[1] += (q1[2] * e[3] - q1[3] * e[2]);
    r[2] += (q1[3] * e[1] - q1[1] * e[3]);
    r[3] += (q1[1] * e[2] - q1[2] * e[1]);
    double nn = 0.0;
    for (int d = 0; d < 4; ++d) nn += r[d] * r[d];
    nn = sqrt(nn);
    for (int d = 0; d < 4; ++d) q[d * c + k] = r[d] / nn;
}

// update_velocity (solver.hpp:527-546) + move_active_particles (solver.hpp:440-480)
__global__ void __launch_bounds__(kBlock) k_update_move(Particles P, const double* __restrict__ u,
                                                       const double* __restrict__ w, int64_t pcap, double dt) {
#pragma clang fp contract(off)
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t k = P.n_inactive + p;
    if (k >= P.n) return;
    const int64_t c = P.cap;
    for (int d = 0; d < P.dim; ++d) P.v[d * c + k] = u[d * pcap + p];
    double wx, wy, wz, normw;
    if (P.dim == 2) {
        P.omega[2 * c + k] = w[2 * pcap + p];
        wx = 0.0, wy = 0.0, wz = P.omega[2 * c + k];
        normw = fabs(wz);
    } else {
        for (int d = 0; d < 3; ++d) P.omega[d * c + k] = w[d * pcap + p];
        wx = P.omega[k], wy = P.omega[c + k], wz = P.omega[2 * c + k];
        double s = 0.0;
        s += wx * wx;
        s += wy * wy;
        s += wz * wz;
        normw = sqrt(s);
    }
    for (int d = 0; d < P.dim; ++d) P.pos[d * c + k] += dt * P.v[d * c + k];
    rotate_q(P.q, c, k, wx, wy, wz, normw, dt);
}

// displacement_obstacles (solver.hpp:329-355), get_omega: 2D (0, 0, w), 3D w.
__global__ void k_obstacles(Particles P, double dt) {
#pragma clang fp contract(off)
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= P.n_inactive) return;
    const int64_t c = P.cap;
    double wx = P.domega[k], wy = P.domega[c + k], wz = P.domega[2 * c + k];
    if (P.dim == 2) wx = 0.0, wy = 0.0;
    double s = 0.0;
    s += wx * wx;
    s += wy * wy;
    s += wz * wz;
    const double normw = sqrt(s);
    for (int d = 0; d < P.dim; ++d) P.pos[d * c + k] += dt * P.vd[d * c + k];
    rotate_q(P.q, c, k, wx, wy, wz, normw, dt);
}

// ------------------------------------------------------------------ periodic boxes
// add_objects_from_periodicity (utils.hpp:228-257): along periodic axis d (in order), object k
// gets a copy when pos[d] - dmax < lower(d); flags in [axis slot][k] order, as the copies are made.
struct BoxDev {
    double lo[3], hi[3];
    int axes[3];  // the periodic axes
    int naxes;
};

__global__ void __launch_bounds__(kBlock) k_copy_flags(Particles P, int64_t n, BoxDev b, double dmax,
                                                      int* __restrict__ flags) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= n * b.naxes) return;
    const int d = b.axes[t / n];
    const int64_t k = t % n;
    flags[t] = (P.pos[d * P.cap + k] - dmax < b.lo[d]) ? 1 : 0;
}

// copy g = offs[t] of object k, moved by L along its axis, at particle n + g
__global__ void __launch_bounds__(kBlock) k_make_copies(Particles P, int64_t n, BoxDev b, const int* __restrict__ flags,
                                                       const int* __restrict__ offs, int32_t* __restrict__ orig) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= n * b.naxes || !flags[t]) return;
    const int d = b.axes[t / n];
    const int64_t k = t % n, g = n + offs[t];
    const int64_t c = P.cap;
    P.shape[g] = P.shape[k];
    P.m[g] = P.m[k];
    for (int e = 0; e < 3; ++e) {
        P.pos[e * c + g] = P.pos[e * c + k];
        P.size[e * c + g] = P.size[e * c + k];
        P.v[e * c + g] = P.v[e * c + k];
        P.vd[e * c + g] = P.vd[e * c + k];
        P.omega[e * c + g] = P.omega[e * c + k];
        P.domega[e * c + g] = P.domega[e * c + k];
        P.f[e * c + g] = P.f[e * c + k];
        P.J[e * c + g] = P.J[e * c + k];
    }
    for (int e = 0; e < 4; ++e) P.q[e * c + g] = P.q[e * c + k];
    for (int e = 0; e < 2; ++e) P.sq[e * c + g] = P.sq[e * c + k];
    P.pos[d * c + g] += b.hi[d] - b.lo[d];
    orig[g - n] = (int32_t)k;
}

// compute_exact_distance's periodic rule (contact/base.hpp:116-139): keep a pair when one
// particle is an original (i < j, so i < n); a copy's index maps to its original and its
// contact point moves back by L along every periodic axis
__global__ void __launch_bounds__(kBlock) k_keep_pairs(const int32_t* __restrict__ ci, int64_t nc, int64_t n,
                                                      int* __restrict__ keep) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c < nc) keep[c] = ci[c] < n ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) k_map_pairs(Contacts E, int64_t nc, int64_t n, BoxDev b,
                                                     const int32_t* __restrict__ orig, const int* __restrict__ keep,
                                                     const int* __restrict__ pos, Contacts T, uint32_t* __restrict__ jkey,
                                                     int32_t* __restrict__ ikey, int32_t* __restrict__ vals) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc || !keep[c]) return;
    const int64_t p = pos[c];
    const int32_t i = E.i[c], je = E.j[c];
    const bool copy = je >= n;
    const int32_t j = copy ? orig[je - n] : je;
    T.i[p] = i;
    T.j[p] = j;
    T.dij[p] = E.dij[c];
    T.mu[p] = E.mu[c];
    for (int e = 0; e < 3; ++e) {
        T.nij[e * T.cap + p] = E.nij[e * E.cap + c];
        T.pi[e * T.cap + p] = E.pi[e * E.cap + c];
        double pj = E.pj[e * E.cap + c];
        for (int a = 0; a < b.naxes; ++a)
            if (copy && b.axes[a] == e) pj -= b.hi[e] - b.lo[e];
        T.pj[e * T.cap + p] = pj;
    }
    jkey[p] = (uint32_t)j;
    ikey[p] = i;
    vals[p] = (int32_t)p;
}

__global__ void __launch_bounds__(kBlock) k_permute_keys(const int32_t* __restrict__ ikey, const int32_t* __restrict__ vals,
                                                        int64_t nc, uint32_t* __restrict__ keys) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k < nc) keys[k] = (uint32_t)ikey[vals[k]];
}

__global__ void __launch_bounds__(kBlock) k_gather_pairs(Contacts T, const int32_t* __restrict__ vals, int64_t nc,
                                                        Contacts C) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= nc) return;
    const int64_t p = vals[k];
    C.i[k] = T.i[p];
    C.j[k] = T.j[p];
    C.dij[k] = T.dij[p];
    C.mu[k] = T.mu[p];
    for (int e = 0; e < 3; ++e) {
        C.nij[e * C.cap + k] = T.nij[e * T.cap + p];
        C.pi[e * C.cap + k] = T.pi[e * T.cap + p];
        C.pj[e * C.cap + k] = T.pj[e * T.cap + p];
    }
}

// the end of move_active_particles (solver.hpp:482-522): along a periodic axis an object past
// the upper bound moves back by L, one past the lower bound forward by L (obstacles too)
__global__ void __launch_bounds__(kBlock) k_periodic_wrap(Particles P, BoxDev b) {
#pragma clang fp contract(off)
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= P.n) return;
    for (int a = 0; a < b.naxes; ++a) {
        const int d = b.axes[a];
        double& x = P.pos[d * P.cap + k];
        const bool plus = x > b.hi[d], minus = x < b.lo[d];
        if (plus) x -= b.hi[d] - b.lo[d];
        if (minus) x += b.hi[d] - b.lo[d];
    }
}

void particles_free(Particles& P) {
    for (void* p : {(void*)P.shape, (void*)P.pos, (void*)P.q, (void*)P.size, (void*)P.sq, (void*)P.v, (void*)P.vd,
                    (void*)P.omega, (void*)P.domega, (void*)P.f, (void*)P.m, (void*)P.J})
        if (p) (void)hipFree(p);
    const int dim = P.dim;
    P = Particles{};
    P.dim = dim;
}

void particles_reserve(Particles& P, int64_t n) {
    if (n <= P.cap) return;
    particles_free(P);
    const int64_t cap = std::max<int64_t>(n, 64);
    P.cap = cap;
    SCOPI_HIP_CHECK(hipMalloc(&P.shape, cap * sizeof(int32_t)));
    for (double** a : {&P.pos, &P.size, &P.v, &P.vd, &P.omega, &P.domega, &P.f, &P.J})
        SCOPI_HIP_CHECK(hipMalloc(a, 3 * cap * sizeof(double)));
    SCOPI_HIP_CHECK(hipMalloc(&P.q, 4 * cap * sizeof(double)));
    SCOPI_HIP_CHECK(hipMalloc(&P.sq, 2 * cap * sizeof(double)));
    SCOPI_HIP_CHECK(hipMalloc(&P.m, cap * sizeof(double)));
}

// host n x w row-major -> device component-major (w x cap)
void upload_rows(double* dev, int64_t cap, const double* host, int64_t n, int w, hipStream_t s) {
    std::vector<double> t((size_t)w * n);
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < w; ++d) t[(size_t)d * n + i] = host[i * w + d];
    for (int d = 0; d < w; ++d)
        SCOPI_HIP_CHECK(hipMemcpyAsync(dev + d * cap, t.data() + (size_t)d * n, n * sizeof(double),
                                       hipMemcpyHostToDevice, s));
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
}

void download_rows(const double* dev, int64_t cap, double* host, int64_t n, int w, hipStream_t s,
                   int64_t host_rows_offset = 0) {
    std::vector<double> t((size_t)w * n);
    for (int d = 0; d < w; ++d)
        SCOPI_HIP_CHECK(hipMemcpyAsync(t.data() + (size_t)d * n, dev + d * cap, n * sizeof(double),
                                       hipMemcpyDeviceToHost, s));
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < w; ++d) host[(host_rows_offset + i) * w + d] = t[(size_t)d * n + i];
}

// grows P to `cap` particles keeping its contents
void particles_grow(Particles& P, int64_t cap, hipStream_t s) {
    if (cap <= P.cap) return;
    Particles Q{};
    Q.dim = P.dim;
    particles_reserve(Q, cap);
    Q.n = P.n;
    Q.n_inactive = P.n_inactive;
    const int64_t n = P.n;
    if (n > 0) {
        SCOPI_HIP_CHECK(hipMemcpyAsync(Q.shape, P.shape, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        SCOPI_HIP_CHECK(hipMemcpyAsync(Q.m, P.m, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        struct F {
            double* dst;
            const double* src;
            int w;
        };
        for (const F& f : {F{Q.pos, P.pos, 3}, F{Q.q, P.q, 4}, F{Q.size, P.size, 3}, F{Q.sq, P.sq, 2}, F{Q.v, P.v, 3},
                           F{Q.vd, P.vd, 3}, F{Q.omega, P.omega, 3}, F{Q.domega, P.domega, 3}, F{Q.f, P.f, 3},
                           F{Q.J, P.J, 3}})
            for (int d = 0; d < f.w; ++d)
                SCOPI_HIP_CHECK(hipMemcpyAsync(f.dst + d * Q.cap, f.src + d * P.cap, n * sizeof(double),
                                               hipMemcpyDeviceToDevice, s));
        SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    }
    particles_free(P);
    P = Q;
}

int fail(scopi_hip_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

template <class F>
int guarded(scopi_hip_ctx* ctx, F&& f) {
    if (!ctx) return SCOPI_HIP_ERR_ARG;
    try {
        SCOPI_HIP_CHECK(hipSetDevice(ctx->device));
        return f();
    } catch (const hip_error& e) {
        return fail(ctx, SCOPI_HIP_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.code));
    } catch (const std::exception& e) {
        return fail(ctx, SCOPI_HIP_ERR_HIP, e.what());
    }
}

bool viscous_problem(int problem) {
    return problem == SCOPI_PROBLEM_VISCOUS || problem == SCOPI_PROBLEM_VISCOUS_FRICTION;
}

__global__ void k_fill_d(double* __restrict__ a, int64_t n, double v) {
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlock) a[t] = v;
}

// ScopiSolver::transfer (solver.hpp:241-259): a contact whose (i, j) was a contact of the last step
// takes that contact's gamma (the last of its kind when a pair occurs twice, as the reference's
// map keeps the last index), any other contact the default gamma.  keys: (i << 32 | j), sorted.
__global__ void __launch_bounds__(kBlock) k_transfer_gamma(const int32_t* __restrict__ ci, const int32_t* __restrict__ cj,
                                                          int64_t nc, const uint64_t* __restrict__ keys,
                                                          const double* __restrict__ hg, int64_t hn, double g0,
                                                          double* __restrict__ gamma) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    const uint64_t key = ((uint64_t)(uint32_t)ci[c] << 32) | (uint32_t)cj[c];
    int64_t lo = 0, hi = hn;  // upper bound of key
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] <= key)
            lo = mid + 1;
        else
            hi = mid;
    }
    gamma[c] = (lo > 0 && keys[lo - 1] == key) ? hg[lo - 1] : g0;
}

__global__ void __launch_bounds__(kBlock) k_record_history(const int32_t* __restrict__ ci, const int32_t* __restrict__ cj,
                                                          int64_t nc, const double* __restrict__ gamma,
                                                          uint64_t* __restrict__ keys, double* __restrict__ hg) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    keys[c] = ((uint64_t)(uint32_t)ci[c] << 32) | (uint32_t)cj[c];
    hg[c] = gamma[c];
}

// Warm start: a contact whose (i, j) was a contact of the last solve starts from that contact's
// multipliers (nr rows), any other from 0 -- transfer (solver.hpp:241-259) applied to lambda.
__global__ void __launch_bounds__(kBlock) k_transfer_lambda(const int32_t* __restrict__ ci,
                                                           const int32_t* __restrict__ cj, int64_t nc,
                                                           const uint64_t* __restrict__ keys,
                                                           const double* __restrict__ hl, int64_t hn, int nr,
                                                           double* __restrict__ lam0) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    const uint64_t key = ((uint64_t)(uint32_t)ci[c] << 32) | (uint32_t)cj[c];
    int64_t lo = 0, hi = hn;  // upper bound of key
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] <= key)
            lo = mid + 1;
        else
            hi = mid;
    }
    const bool found = lo > 0 && keys[lo - 1] == key;
    for (int d = 0; d < nr; ++d) lam0[c * nr + d] = found ? hl[(lo - 1) * nr + d] : 0.0;
}

__global__ void __launch_bounds__(kBlock) k_record_lambda(const int32_t* __restrict__ ci, const int32_t* __restrict__ cj,
                                                         int64_t nc, int nr, const double* __restrict__ lam,
                                                         uint64_t* __restrict__ keys, double* __restrict__ hl) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    keys[c] = ((uint64_t)(uint32_t)ci[c] << 32) | (uint32_t)cj[c];
    for (int d = 0; d < nr; ++d) hl[c * nr + d] = lam[c * nr + d];
}

void gamma_reserve(scopi_hip_ctx* ctx, int64_t nc) {
    if (nc <= ctx->gamma_cap) return;
    if (ctx->gamma) SCOPI_HIP_CHECK(hipFree(ctx->gamma));
    ctx->gamma_cap = nc + nc / 4 + 1024;
    SCOPI_HIP_CHECK(hipMalloc(&ctx->gamma, ctx->gamma_cap * sizeof(double)));
}

// the current contacts' gamma before a solve: kept when set for these contacts, else the
// default (from_history: the transfer of the last step's values)
void prepare_gamma(scopi_hip_ctx* ctx, const scopi_hip_params& prm, bool from_history) {
    const int64_t nc = ctx->nc;
    gamma_reserve(ctx, std::max<int64_t>(nc, 1));
    if (nc == 0 || (ctx->gamma_valid && !from_history)) return;
    if (from_history && ctx->hist_n > 0 && ctx->hist_kind == 2)
        hipLaunchKernelGGL(k_transfer_gamma, dim3(grid_for(nc)), dim3(kBlock), 0, ctx->stream,
                           (const int32_t*)ctx->C.i, (const int32_t*)ctx->C.j, nc, (const uint64_t*)ctx->hist_keys,
                           (const double*)ctx->hist_gamma, ctx->hist_n, prm.gamma, ctx->gamma);
    else
        hipLaunchKernelGGL(k_fill_d, dim3(256), dim3(kBlock), 0, ctx->stream, ctx->gamma, nc, prm.gamma);
    SCOPI_HIP_CHECK(hipGetLastError());
    ctx->gamma_valid = true;
}

// m_old_contacts = this step's contacts (solver.hpp:300)
bool warm_problem(const scopi_hip_params& prm) {
    return prm.problem == SCOPI_PROBLEM_NOFRICTION || prm.problem == SCOPI_PROBLEM_FRICTION;
}

// warm start: the current contacts' first iterate from the last solve's multipliers, or null when
// there is nothing to start from (no history, another row count)
const double* prepare_lambda(scopi_hip_ctx* ctx, const scopi_hip_params& prm) {
    const int nr = prm.problem == SCOPI_PROBLEM_FRICTION ? 3 : 1;
    const int64_t nc = ctx->nc;
    if (!prm.warm_start || nc == 0 || ctx->hist_n == 0 || ctx->hist_kind != 1 || ctx->hist_nr != nr) return nullptr;
    if (nr * nc > ctx->lam0_cap) {
        if (ctx->lam0) SCOPI_HIP_CHECK(hipFree(ctx->lam0));
        ctx->lam0_cap = nr * nc + nr * nc / 4 + 1024;
        SCOPI_HIP_CHECK(hipMalloc(&ctx->lam0, ctx->lam0_cap * sizeof(double)));
    }
    hipLaunchKernelGGL(k_transfer_lambda, dim3(grid_for(nc)), dim3(kBlock), 0, ctx->stream, (const int32_t*)ctx->C.i,
                       (const int32_t*)ctx->C.j, nc, (const uint64_t*)ctx->hist_keys, (const double*)ctx->hist_lam,
                       ctx->hist_n, nr, ctx->lam0);
    SCOPI_HIP_CHECK(hipGetLastError());
    return ctx->lam0;
}

// warm start: keep this solve's contacts and multipliers for the next one
void record_lambda(scopi_hip_ctx* ctx) {
    const int64_t nc = ctx->nc;
    const int nr = ctx->solver.rows_per_contact();
    if (nc > ctx->hist_cap) {
        for (void* p : {(void*)ctx->hist_keys, (void*)ctx->hist_gamma})
            if (p) SCOPI_HIP_CHECK(hipFree(p));
        ctx->hist_cap = nc + nc / 4 + 1024;
        SCOPI_HIP_CHECK(hipMalloc(&ctx->hist_keys, ctx->hist_cap * sizeof(uint64_t)));
        SCOPI_HIP_CHECK(hipMalloc(&ctx->hist_gamma, ctx->hist_cap * sizeof(double)));
    }
    if (nr * nc > ctx->hist_lam_cap) {
        if (ctx->hist_lam) SCOPI_HIP_CHECK(hipFree(ctx->hist_lam));
        ctx->hist_lam_cap = nr * nc + nr * nc / 4 + 1024;
        SCOPI_HIP_CHECK(hipMalloc(&ctx->hist_lam, ctx->hist_lam_cap * sizeof(double)));
    }
    if (nc > 0)
        hipLaunchKernelGGL(k_record_lambda, dim3(grid_for(nc)), dim3(kBlock), 0, ctx->stream, (const int32_t*)ctx->C.i,
                           (const int32_t*)ctx->C.j, nc, nr, ctx->solver.final_lambda(), ctx->hist_keys,
                           ctx->hist_lam);
    SCOPI_HIP_CHECK(hipGetLastError());
    ctx->hist_n = nc;
    ctx->hist_nr = nr;
    ctx->hist_kind = 1;
}

void record_history(scopi_hip_ctx* ctx) {
    const int64_t nc = ctx->nc;
    if (nc > ctx->hist_cap) {
        for (void* p : {(void*)ctx->hist_keys, (void*)ctx->hist_gamma})
            if (p) SCOPI_HIP_CHECK(hipFree(p));
        ctx->hist_cap = nc + nc / 4 + 1024;
        SCOPI_HIP_CHECK(hipMalloc(&ctx->hist_keys, ctx->hist_cap * sizeof(uint64_t)));
        SCOPI_HIP_CHECK(hipMalloc(&ctx->hist_gamma, ctx->hist_cap * sizeof(double)));
    }
    if (nc > 0)
        hipLaunchKernelGGL(k_record_history, dim3(grid_for(nc)), dim3(kBlock), 0, ctx->stream, (const int32_t*)ctx->C.i,
                           (const int32_t*)ctx->C.j, nc, (const double*)ctx->gamma, ctx->hist_keys, ctx->hist_gamma);
    SCOPI_HIP_CHECK(hipGetLastError());
    ctx->hist_n = nc;
    ctx->hist_kind = 2;
}

bool params_ok(const scopi_hip_params* p) {
    return p &&
           (p->problem == SCOPI_PROBLEM_NOFRICTION || p->problem == SCOPI_PROBLEM_FRICTION ||
            p->problem == SCOPI_PROBLEM_FRICTION_FIXED_POINT || viscous_problem(p->problem)) &&
           (p->contact_method == SCOPI_CONTACT_KDTREE || p->contact_method == SCOPI_CONTACT_BRUTE_FORCE) &&
           (p->vap == SCOPI_VAP_FIXED || p->vap == SCOPI_VAP_FPD || p->vap == SCOPI_VAP_PROJECTION) &&
           (p->method == SCOPI_METHOD_APGD || p->method == SCOPI_METHOD_PGD) &&
           (p->precision == SCOPI_PRECISION_FP64 || p->precision == SCOPI_PRECISION_FP32_REFINE) &&
           (p->warm_start == 0 || p->warm_start == 1);
}

BoxDev box_dev(const scopi_hip_ctx* ctx) {
    BoxDev b{};
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = ctx->box_lo[d];
        b.hi[d] = ctx->box_hi[d];
        if (d < ctx->dim && ctx->periodic[d]) b.axes[b.naxes++] = d;
    }
    return b;
}

template <class T>
void grow_buf(T*& p, int64_t& cap, int64_t need) {
    if (need <= cap) return;
    if (p) SCOPI_HIP_CHECK(hipFree(p));
    cap = need + need / 4 + 1024;
    SCOPI_HIP_CHECK(hipMalloc(&p, cap * sizeof(T)));
}

// contact_base::run (contact/base.hpp:70-77): with periodic axes, the search runs over the
// particles and their periodic copies (appended after them), and its pairs are kept, mapped
// back and sorted by (i, j) (stable: a pair met directly and through a copy keeps the
// direct one first)
int64_t contacts_step(scopi_hip_ctx* ctx, const scopi_hip_params& prm) {
    hipStream_t s = ctx->stream;
    Particles& P = ctx->P;
    const BoxDev b = box_dev(ctx);
    if (b.naxes == 0) return ctx->search.run(P, ctx->C, prm, s, !ctx->has_se);
    const int64_t n = P.n, nt = n * b.naxes;
    if (!ctx->ptotal) {
        SCOPI_HIP_CHECK(hipMalloc(&ctx->ptotal, sizeof(int)));
        SCOPI_HIP_CHECK(hipHostMalloc(&ctx->h_ptotal, sizeof(int)));
    }
    grow_buf(ctx->pflags, ctx->cap_flags, nt + 1);
    grow_buf(ctx->poffs, ctx->cap_offs, nt + 1);
    hipLaunchKernelGGL(k_copy_flags, dim3(grid_for(std::max<int64_t>(nt, 1))), dim3(kBlock), 0, s, P, n, b, prm.dmax,
                       ctx->pflags);
    ctx->pscan.exclusive(ctx->pflags, ctx->poffs, nt, ctx->ptotal, s);
    SCOPI_HIP_CHECK(hipMemcpyAsync(ctx->h_ptotal, ctx->ptotal, sizeof(int), hipMemcpyDeviceToHost, s));
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    const int64_t G = nt > 0 ? *ctx->h_ptotal : 0;
    if (n + G > P.cap) particles_grow(P, n + G + (n + G) / 4, s);
    if (G > 0) {
        grow_buf(ctx->porig, ctx->cap_orig, G);
        hipLaunchKernelGGL(k_make_copies, dim3(grid_for(nt)), dim3(kBlock), 0, s, P, n, b, (const int*)ctx->pflags,
                           (const int*)ctx->poffs, ctx->porig);
    }
    P.n = n + G;
    int64_t next = 0;
    try {
        // copies of obstacles (planes) are searched as particles: no bounding radius for the cut-off
        next = ctx->search.run(P, ctx->C2, prm, s, !ctx->has_se, P.n_inactive == 0);
    } catch (...) {
        P.n = n;
        throw;
    }
    P.n = n;
    if (next == 0) {
        contacts_reserve(ctx->C, 1);
        return 0;
    }
    grow_buf(ctx->pflags, ctx->cap_flags, next + 1);
    grow_buf(ctx->poffs, ctx->cap_offs, next + 1);
    hipLaunchKernelGGL(k_keep_pairs, dim3(grid_for(next)), dim3(kBlock), 0, s, (const int32_t*)ctx->C2.i, next, n,
                       ctx->pflags);
    ctx->pscan.exclusive(ctx->pflags, ctx->poffs, next, ctx->ptotal, s);
    SCOPI_HIP_CHECK(hipMemcpyAsync(ctx->h_ptotal, ctx->ptotal, sizeof(int), hipMemcpyDeviceToHost, s));
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
    const int64_t nc = *ctx->h_ptotal;
    contacts_reserve(ctx->C, std::max<int64_t>(nc, 1));
    if (nc == 0) return 0;
    grow_buf(ctx->pkeys, ctx->cap_keys, nc);
    grow_buf(ctx->pvals, ctx->cap_vals, nc);
    grow_buf(ctx->pi_of, ctx->cap_iof, nc);
    hipLaunchKernelGGL(k_map_pairs, dim3(grid_for(next)), dim3(kBlock), 0, s, ctx->C2, next, n, b,
                       (const int32_t*)ctx->porig, (const int*)ctx->pflags, (const int*)ctx->poffs, ctx->C, ctx->pkeys,
                       ctx->pi_of, ctx->pvals);
    int bits = 1;
    while ((int64_t(1) << bits) <= n) ++bits;
    ctx->psort.sort(ctx->pkeys, ctx->pvals, nc, bits, s, ctx->pscan);  // by j ...
    hipLaunchKernelGGL(k_permute_keys, dim3(grid_for(nc)), dim3(kBlock), 0, s, (const int32_t*)ctx->pi_of,
                       (const int32_t*)ctx->pvals, nc, ctx->pkeys);
    ctx->psort.sort(ctx->pkeys, ctx->pvals, nc, bits, s, ctx->pscan);  // ... then stably by i
    hipLaunchKernelGGL(k_gather_pairs, dim3(grid_for(nc)), dim3(kBlock), 0, s, ctx->C, (const int32_t*)ctx->pvals, nc,
                       ctx->C2);
    SCOPI_HIP_CHECK(hipGetLastError());
    std::swap(ctx->C, ctx->C2);
    return nc;
}

void periodic_wrap(scopi_hip_ctx* ctx) {
    const BoxDev b = box_dev(ctx);
    if (b.naxes == 0 || ctx->P.n == 0) return;
    hipLaunchKernelGGL(k_periodic_wrap, dim3(grid_for(ctx->P.n)), dim3(kBlock), 0, ctx->stream, ctx->P, b);
    SCOPI_HIP_CHECK(hipGetLastError());
}

}  // namespace

// ------------------------------------------------------------------ allocation-free closest points
// scopi_hip_se_closest_points_noalloc: the same kernels as scopi_hip_se_closest_points on device
// buffers of the caller, with every scratch array (the 2n bodies' particle records, their grids,
// the candidates and the work counter) carved out of one file-scope device array.  Each GPU holds
// its own copy of the array from module load until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns for the current device.
namespace {
constexpr int64_t kNoAllocPairs = SCOPI_HIP_NOALLOC_MAX_PAIRS;
constexpr int64_t kNoAllocBodies = 2 * kNoAllocPairs;
constexpr int64_t kAlign = 32;  // doubles: 256 bytes
constexpr int64_t up(int64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
// sizes in doubles (int arrays rounded up to whole doubles)
constexpr int64_t kNaParticles = up(3 * kNoAllocBodies) + up(4 * kNoAllocBodies) + up(3 * kNoAllocBodies) +
                                 up(2 * kNoAllocBodies) + up(kNoAllocBodies / 2);
constexpr int64_t kNaGrids = up(2 * kSeGridAngles * kNoAllocBodies) + up(kNoAllocBodies) +
                             up((int64_t)6 * kSeGridPoints * kNoAllocBodies);
constexpr int64_t kNaCandidates = 2 * up(kNoAllocPairs / 2) + 2 * up(4 * kNoAllocPairs) + 2 * up(kNoAllocPairs / 2) +
                                  up(kNoAllocPairs) + 3 * up(3 * kNoAllocPairs) + up(1);
constexpr int64_t kNoAllocScratch = kNaParticles + kNaGrids + kNaCandidates;  // about 14.5 MB

static __device__ __align__(256) double scratch[kNoAllocScratch];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once per device).
double* scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_scratch[device] = static_cast<double*>(p);
    }
    return g_scratch[device];
}

// Particle records, grids and candidates for n pairs over the scratch array at `base`; false if the
// layout would not fit the array.
bool carve(double* base, int64_t n, Particles& T, SeGrids& G, SeCandidates& S) {
    double* at = base;
    auto take = [&](int64_t doubles) {
        double* p = at;
        at += up(doubles);
        return p;
    };
    T = Particles{};
    T.dim = 3;
    T.n = 2 * n;
    T.n_inactive = 0;
    T.cap = kNoAllocBodies;
    T.pos = take(3 * kNoAllocBodies);
    T.q = take(4 * kNoAllocBodies);
    T.size = take(3 * kNoAllocBodies);
    T.sq = take(2 * kNoAllocBodies);
    T.shape = reinterpret_cast<int32_t*>(take(kNoAllocBodies / 2));
    G = SeGrids{};
    G.cap = kNoAllocBodies;
    G.ang = take(2 * kSeGridAngles * kNoAllocBodies);
    G.count = reinterpret_cast<int32_t*>(take(kNoAllocBodies));
    G.pts = take((int64_t)6 * kSeGridPoints * kNoAllocBodies);
    S = SeCandidates{};
    S.cap = kNoAllocPairs;
    S.i = reinterpret_cast<int32_t*>(take(kNoAllocPairs / 2));
    S.j = reinterpret_cast<int32_t*>(take(kNoAllocPairs / 2));
    S.u0 = take(4 * kNoAllocPairs);
    S.u = take(4 * kNoAllocPairs);
    S.info = reinterpret_cast<int32_t*>(take(kNoAllocPairs / 2));
    S.nfev = reinterpret_cast<int32_t*>(take(kNoAllocPairs / 2));
    S.dij = take(kNoAllocPairs);
    S.nij = take(3 * kNoAllocPairs);
    S.pi = take(3 * kNoAllocPairs);
    S.pj = take(3 * kNoAllocPairs);
    S.counter = reinterpret_cast<int*>(take(1));
    return at - base <= kNoAllocScratch;
}

// the caller's row-major pair arrays -> body records 0..n-1 (side i) and n..2n-1 (side j)
__global__ void __launch_bounds__(kBlock) k_noalloc_stage(Particles T, SeCandidates S, int64_t n,
                                                          const double* __restrict__ pos_i,
                                                          const double* __restrict__ q_i,
                                                          const double* __restrict__ size_i,
                                                          const double* __restrict__ sq_i,
                                                          const double* __restrict__ pos_j,
                                                          const double* __restrict__ q_j,
                                                          const double* __restrict__ size_j,
                                                          const double* __restrict__ sq_j) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t cap = T.cap;
    for (int side = 0; side < 2; ++side) {
        const int64_t b = side ? n + k : k;
        const double* pos = side ? pos_j : pos_i;
        const double* q = side ? q_j : q_i;
        const double* size = side ? size_j : size_i;
        const double* sq = side ? sq_j : sq_i;
        for (int d = 0; d < 3; ++d) T.pos[d * cap + b] = pos[3 * k + d];
        for (int d = 0; d < 4; ++d) T.q[d * cap + b] = q[4 * k + d];
        for (int d = 0; d < 3; ++d) T.size[d * cap + b] = size[3 * k + d];
        for (int d = 0; d < 2; ++d) T.sq[d * cap + b] = sq[2 * k + d];
        T.shape[b] = SCOPI_SHAPE_SUPERELLIPSOID;
    }
    S.i[k] = (int32_t)k;
    S.j[k] = (int32_t)(n + k);
}

// the candidates' records -> the caller's row-major outputs
__global__ void __launch_bounds__(kBlock) k_noalloc_scatter(SeCandidates S, int64_t n, double* __restrict__ dij,
                                                            double* __restrict__ nij, double* __restrict__ pi,
                                                            double* __restrict__ pj, double* __restrict__ u0,
                                                            double* __restrict__ u, int32_t* __restrict__ info,
                                                            int32_t* __restrict__ nfev) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t cap = S.cap;
    dij[k] = S.dij[k];
    for (int d = 0; d < 3; ++d) {
        nij[3 * k + d] = S.nij[d * cap + k];
        pi[3 * k + d] = S.pi[d * cap + k];
        pj[3 * k + d] = S.pj[d * cap + k];
    }
    if (u0)
        for (int d = 0; d < 4; ++d) u0[4 * k + d] = S.u0[4 * k + d];
    if (u)
        for (int d = 0; d < 4; ++d) u[4 * k + d] = S.u[4 * k + d];
    if (info) info[k] = S.info[k];
    if (nfev) nfev[k] = S.nfev[k];
}

int noalloc_fail(scopi_hip_ctx* ctx, int code, const char* msg) {
    ctx->err.assign(msg);  // within the capacity reserved at create: no allocation
    return code;
}
}  // namespace

extern "C" {

int scopi_hip_default_params(scopi_hip_params* p) {
    if (!p) return SCOPI_HIP_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    p->problem = SCOPI_PROBLEM_NOFRICTION;
    p->contact_method = SCOPI_CONTACT_KDTREE;
    p->vap = SCOPI_VAP_FIXED;
    p->method = SCOPI_METHOD_APGD;
    p->precision = SCOPI_PRECISION_FP64;
    p->dynamic_descent = 1;
    p->dmax = 2.0;
    p->kd_tree_radius = 17.0;
    p->alpha = 0.05;
    p->tolerance = 1e-7;
    p->mu = 0.1;
    p->max_ite = 10000;
    p->fp32_tolerance = 0.0;
    p->warm_start = 0;
    p->general_jacobian = 0;
    p->fixed_point_tol = 1e-6;
    p->fixed_point_max_iter = 1000;
    p->gamma = 0.0;  // contact_property<Viscous> (contact/property.hpp:125-127)
    p->gamma_min = -2.;
    p->gamma_tol = 1e-6;
    return SCOPI_HIP_OK;
}

int scopi_hip_create(int32_t dim, int64_t capacity_particles, int64_t capacity_contacts, scopi_hip_ctx** out) {
    if (!out || (dim != 2 && dim != 3) || capacity_particles < 0 || capacity_contacts < 0) return SCOPI_HIP_ERR_ARG;
    *out = nullptr;
    auto* ctx = new scopi_hip_ctx();
    ctx->dim = dim;
    ctx->P.dim = dim;
    ctx->err.reserve(256);  // error messages up to this length never allocate afterwards
    try {
        SCOPI_HIP_CHECK(hipGetDevice(&ctx->device));
        SCOPI_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        particles_reserve(ctx->P, std::max<int64_t>(capacity_particles, 1));
        contacts_reserve(ctx->C, std::max<int64_t>(capacity_contacts, 1));
    } catch (const hip_error& e) {
        std::fprintf(stderr, "scopi_hip_create: %s: %s\n", e.what, hipGetErrorString(e.code));
        delete ctx;
        return SCOPI_HIP_ERR_HIP;
    }
    *out = ctx;
    return SCOPI_HIP_OK;
}

int scopi_hip_destroy(scopi_hip_ctx* ctx) {
    if (!ctx) return SCOPI_HIP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    particles_free(ctx->P);
    contacts_free(ctx->C);
    contacts_free(ctx->C2);
    if (ctx->proj) (void)hipFree(ctx->proj);
    for (void* p : {(void*)ctx->gamma, (void*)ctx->hist_keys, (void*)ctx->hist_gamma, (void*)ctx->hist_lam,
                    (void*)ctx->lam0})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)ctx->pflags, (void*)ctx->poffs, (void*)ctx->porig, (void*)ctx->ptotal, (void*)ctx->pkeys,
                    (void*)ctx->pvals, (void*)ctx->pi_of})
        if (p) (void)hipFree(p);
    if (ctx->h_ptotal) (void)hipHostFree(ctx->h_ptotal);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SCOPI_HIP_OK;
}

const char* scopi_hip_last_error(const scopi_hip_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int scopi_hip_upload_state(scopi_hip_ctx* ctx, int64_t n, int64_t n_inactive, const int32_t* shape, const double* pos,
                           const double* q, const double* size, const double* squareness, const double* v,
                           const double* vd, const double* omega, const double* domega, const double* f,
                           const double* m, const double* J) {
    return guarded(ctx, [&]() -> int {
        if (n < 0 || n_inactive < 0 || n_inactive > n || !shape || !pos || !q || !size || !v || !vd || !omega ||
            !domega || !f || !m || !J)
            return fail(ctx, SCOPI_HIP_ERR_ARG, "upload_state: bad argument");
        if (n > INT32_MAX / 2) return fail(ctx, SCOPI_HIP_ERR_CAPACITY, "upload_state: too many particles");
        bool has_se = false;
        for (int64_t k = 0; k < n; ++k) {
            const int s = shape[k];
            const bool active = k >= n_inactive;
            if (s < 0 || s > SCOPI_SHAPE_SUPERELLIPSOID)
                return fail(ctx, SCOPI_HIP_ERR_ARG, "upload_state: unknown shape tag");
            if (active && s != SCOPI_SHAPE_SPHERE && s != SCOPI_SHAPE_SUPERELLIPSOID)
                return fail(ctx, SCOPI_HIP_ERR_UNSUPPORTED, "upload_state: active particles must be spheres or "
                                                            "superellipsoids");
            if (s == SCOPI_SHAPE_SEGMENT && ctx->dim != 2)
                return fail(ctx, SCOPI_HIP_ERR_UNSUPPORTED, "upload_state: segments are 2D objects");
            has_se = has_se || s == SCOPI_SHAPE_SUPERELLIPSOID;
        }
        if (has_se && ctx->any_periodic())
            return fail(ctx, SCOPI_HIP_ERR_UNSUPPORTED, "upload_state: periodic axes with superellipsoids are not supported");
        ctx->has_se = has_se;
        if (n != ctx->P.n || n_inactive != ctx->P.n_inactive) ctx->hist_n = 0;  // indices change meaning
        particles_reserve(ctx->P, n);
        Particles& P = ctx->P;
        P.n = n;
        P.n_inactive = n_inactive;
        P.dim = ctx->dim;
        hipStream_t s = ctx->stream;
        if (n > 0) {
            SCOPI_HIP_CHECK(hipMemcpyAsync(P.shape, shape, n * sizeof(int32_t), hipMemcpyHostToDevice, s));
            SCOPI_HIP_CHECK(hipMemcpyAsync(P.m, m, n * sizeof(double), hipMemcpyHostToDevice, s));
            upload_rows(P.pos, P.cap, pos, n, 3, s);
            upload_rows(P.q, P.cap, q, n, 4, s);
            upload_rows(P.size, P.cap, size, n, 3, s);
            std::vector<double> ones(2 * n, 1.0);
            upload_rows(P.sq, P.cap, squareness ? squareness : ones.data(), n, 2, s);
            upload_rows(P.v, P.cap, v, n, 3, s);
            upload_rows(P.vd, P.cap, vd, n, 3, s);
            upload_rows(P.omega, P.cap, omega, n, 3, s);
            upload_rows(P.domega, P.cap, domega, n, 3, s);
            upload_rows(P.f, P.cap, f, n, 3, s);
            upload_rows(P.J, P.cap, J, n, 3, s);
        }
        ctx->contacts_valid = false;
        ctx->nc = 0;
        ctx->all_spheres = true;
        for (int64_t k = n_inactive; k < n; ++k) ctx->all_spheres = ctx->all_spheres && shape[k] == SCOPI_SHAPE_SPHERE;
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_download_state(scopi_hip_ctx* ctx, double* pos, double* q, double* v, double* omega) {
    return guarded(ctx, [&]() -> int {
        const Particles& P = ctx->P;
        if (P.n == 0) return SCOPI_HIP_OK;
        if (pos) download_rows(P.pos, P.cap, pos, P.n, 3, ctx->stream);
        if (q) download_rows(P.q, P.cap, q, P.n, 4, ctx->stream);
        if (v) download_rows(P.v, P.cap, v, P.n, 3, ctx->stream);
        if (omega) download_rows(P.omega, P.cap, omega, P.n, 3, ctx->stream);
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_upload_velocities(scopi_hip_ctx* ctx, const double* v, const double* omega) {
    return guarded(ctx, [&]() -> int {
        Particles& P = ctx->P;
        if (P.n == 0) return SCOPI_HIP_OK;
        if (v) upload_rows(P.v, P.cap, v, P.n, 3, ctx->stream);
        if (omega) upload_rows(P.omega, P.cap, omega, P.n, 3, ctx->stream);
        return SCOPI_HIP_OK;
    });
}

namespace {
// n independent pairs (particles k and n + k of a scratch set) through the candidate kernels
int pairs_batch(scopi_hip_ctx* ctx, int64_t n, const int32_t* shape_i, const double* pos_i, const double* q_i,
                const double* size_i, const double* sq_i, const int32_t* shape_j, const double* pos_j,
                const double* q_j, const double* size_j, const double* sq_j, double* dij, double* nij, double* pi,
                double* pj, double* u0, double* u, int32_t* info, int32_t* nfev) {
    if (n < 0 || n > INT32_MAX / 4 || (n > 0 && (!pos_i || !q_i || !size_i || !pos_j || !q_j || !size_j || !dij ||
                                                 !nij || !pi || !pj)))
        return fail(ctx, SCOPI_HIP_ERR_ARG, "closest_points: bad argument");
    if (n == 0) return SCOPI_HIP_OK;
    std::vector<int32_t> shape(2 * n, SCOPI_SHAPE_SUPERELLIPSOID), ci(n), cj(n);
    for (int64_t k = 0; k < n; ++k) {
        if (shape_i) shape[k] = shape_i[k];
        if (shape_j) shape[n + k] = shape_j[k];
        for (int64_t m : {k, n + k})
            if (shape[m] != SCOPI_SHAPE_SPHERE && shape[m] != SCOPI_SHAPE_PLANE &&
                shape[m] != SCOPI_SHAPE_SUPERELLIPSOID)
                return fail(ctx, SCOPI_HIP_ERR_UNSUPPORTED, "closest_points: spheres, planes, superellipsoids");
        ci[k] = (int32_t)k;
        cj[k] = (int32_t)(n + k);
    }
    hipStream_t s = ctx->stream;
    Particles T{};
    T.dim = ctx->dim;
    SeCandidates S{};
    SeGrids G{};
    struct Cleanup {
        Particles& T;
        SeCandidates& S;
        SeGrids& G;
        ~Cleanup() {
            particles_free(T);
            se_candidates_free(S);
            se_grids_free(G);
        }
    } cleanup{T, S, G};
    particles_reserve(T, 2 * n);
    T.n = 2 * n;
    T.n_inactive = 0;
    auto both = [&](const double* a, const double* b, int w, double fill) {
        std::vector<double> t((size_t)2 * n * w, fill);
        if (a) std::memcpy(t.data(), a, (size_t)n * w * sizeof(double));
        if (b) std::memcpy(t.data() + (size_t)n * w, b, (size_t)n * w * sizeof(double));
        return t;
    };
    upload_rows(T.pos, T.cap, both(pos_i, pos_j, 3, 0.0).data(), 2 * n, 3, s);
    upload_rows(T.q, T.cap, both(q_i, q_j, 4, 0.0).data(), 2 * n, 4, s);
    upload_rows(T.size, T.cap, both(size_i, size_j, 3, 0.0).data(), 2 * n, 3, s);
    upload_rows(T.sq, T.cap, both(sq_i, sq_j, 2, 1.0).data(), 2 * n, 2, s);
    SCOPI_HIP_CHECK(hipMemcpyAsync(T.shape, shape.data(), 2 * n * sizeof(int32_t), hipMemcpyHostToDevice, s));
    se_candidates_reserve(S, n);
    SCOPI_HIP_CHECK(hipMemcpyAsync(S.i, ci.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s));
    SCOPI_HIP_CHECK(hipMemcpyAsync(S.j, cj.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s));
    se_closest_points(T, G, S, n, s);
    SCOPI_HIP_CHECK(hipMemcpyAsync(dij, S.dij, n * sizeof(double), hipMemcpyDeviceToHost, s));
    if (u0) SCOPI_HIP_CHECK(hipMemcpyAsync(u0, S.u0, 4 * n * sizeof(double), hipMemcpyDeviceToHost, s));
    if (u) SCOPI_HIP_CHECK(hipMemcpyAsync(u, S.u, 4 * n * sizeof(double), hipMemcpyDeviceToHost, s));
    if (info) SCOPI_HIP_CHECK(hipMemcpyAsync(info, S.info, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (nfev) SCOPI_HIP_CHECK(hipMemcpyAsync(nfev, S.nfev, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    download_rows(S.nij, S.cap, nij, n, 3, s);
    download_rows(S.pi, S.cap, pi, n, 3, s);
    download_rows(S.pj, S.cap, pj, n, 3, s);
    return SCOPI_HIP_OK;
}
}  // namespace

int scopi_hip_se_closest_points(scopi_hip_ctx* ctx, int64_t n, const double* pos_i, const double* q_i,
                                const double* size_i, const double* sq_i, const double* pos_j, const double* q_j,
                                const double* size_j, const double* sq_j, double* dij, double* nij, double* pi,
                                double* pj, double* u0, double* u, int32_t* info, int32_t* nfev) {
    return guarded(ctx, [&]() -> int {
        if (n > 0 && (!sq_i || !sq_j)) return fail(ctx, SCOPI_HIP_ERR_ARG, "se_closest_points: bad argument");
        return pairs_batch(ctx, n, nullptr, pos_i, q_i, size_i, sq_i, nullptr, pos_j, q_j, size_j, sq_j, dij, nij, pi,
                           pj, u0, u, info, nfev);
    });
}

int scopi_hip_closest_points(scopi_hip_ctx* ctx, int64_t n, const int32_t* shape_i, const double* pos_i,
                             const double* q_i, const double* size_i, const double* sq_i, const int32_t* shape_j,
                             const double* pos_j, const double* q_j, const double* size_j, const double* sq_j,
                             double* dij, double* nij, double* pi, double* pj, double* u0, double* u, int32_t* info,
                             int32_t* nfev) {
    return guarded(ctx, [&]() -> int {
        if (n > 0 && (!shape_i || !shape_j)) return fail(ctx, SCOPI_HIP_ERR_ARG, "closest_points: bad argument");
        return pairs_batch(ctx, n, shape_i, pos_i, q_i, size_i, sq_i, shape_j, pos_j, q_j, size_j, sq_j, dij, nij, pi,
                           pj, u0, u, info, nfev);
    });
}

// allocation-free variant of scopi_hip_se_closest_points (the scratch array and its helpers are
// defined above the extern "C" block)
int scopi_hip_se_closest_points_noalloc(scopi_hip_ctx* ctx, void* stream, int64_t n, const double* pos_i,
                                        const double* q_i, const double* size_i, const double* sq_i,
                                        const double* pos_j, const double* q_j, const double* size_j,
                                        const double* sq_j, double* dij, double* nij, double* pi, double* pj,
                                        double* u0, double* u, int32_t* info, int32_t* nfev) {
    if (!ctx) return SCOPI_HIP_ERR_ARG;
    if (ctx->dim != 3) return noalloc_fail(ctx, SCOPI_HIP_ERR_UNSUPPORTED, "se_closest_points_noalloc: 3D only");
    if (n < 0) return noalloc_fail(ctx, SCOPI_HIP_ERR_ARG, "se_closest_points_noalloc: n < 0");
    if (n > kNoAllocPairs)
        return noalloc_fail(ctx, SCOPI_HIP_ERR_CAPACITY,
                            "se_closest_points_noalloc: more than SCOPI_HIP_NOALLOC_MAX_PAIRS (256) pairs");
    if (n == 0) return SCOPI_HIP_OK;
    if (!pos_i || !q_i || !size_i || !sq_i || !pos_j || !q_j || !size_j || !sq_j || !dij || !nij || !pi || !pj)
        return noalloc_fail(ctx, SCOPI_HIP_ERR_ARG, "se_closest_points_noalloc: NULL input or output");
    if (hipSetDevice(ctx->device) != hipSuccess)
        return noalloc_fail(ctx, SCOPI_HIP_ERR_HIP, "se_closest_points_noalloc: hipSetDevice failed");
    double* base = scratch_address(ctx->device);
    if (!base) return noalloc_fail(ctx, SCOPI_HIP_ERR_HIP, "se_closest_points_noalloc: no scratch address");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    Particles T;
    SeGrids G;
    SeCandidates S;
    if (!carve(base, n, T, G, S))
        return noalloc_fail(ctx, SCOPI_HIP_ERR_CAPACITY, "se_closest_points_noalloc: scratch layout too large");
    const unsigned grid = (unsigned)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_noalloc_stage, dim3(grid), dim3(kBlock), 0, s, T, S, n, pos_i, q_i, size_i, sq_i, pos_j,
                       q_j, size_j, sq_j);
    try {
        se_closest_points(T, G, S, n, s);  // G and S already hold 2n bodies / n pairs: no allocation
    } catch (const hip_error&) {
        return noalloc_fail(ctx, SCOPI_HIP_ERR_HIP, "se_closest_points_noalloc: kernel launch failed");
    }
    hipLaunchKernelGGL(k_noalloc_scatter, dim3(grid), dim3(kBlock), 0, s, S, n, dij, nij, pi, pj, u0, u, info, nfev);
    if (hipGetLastError() != hipSuccess)
        return noalloc_fail(ctx, SCOPI_HIP_ERR_HIP, "se_closest_points_noalloc: kernel launch failed");
    return SCOPI_HIP_OK;
}

int scopi_hip_newton_probe(scopi_hip_ctx* ctx, int32_t kind, const double* in, double* out) {
    return guarded(ctx, [&]() -> int {
        if ((kind != 0 && kind != 1) || !in || !out) return fail(ctx, SCOPI_HIP_ERR_ARG, "newton_probe: bad argument");
        double* d = nullptr;
        SCOPI_HIP_CHECK(hipMalloc(&d, 72 * sizeof(double)));
        struct Free {
            double* p;
            ~Free() { (void)hipFree(p); }
        } f{d};
        hipStream_t s = ctx->stream;
        SCOPI_HIP_CHECK(hipMemcpyAsync(d, in, 44 * sizeof(double), hipMemcpyHostToDevice, s));
        newton_probe(kind, d, d + 44, s);
        SCOPI_HIP_CHECK(hipMemcpyAsync(out, d + 44, 24 * sizeof(double), hipMemcpyDeviceToHost, s));
        SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_portable_math(scopi_hip_ctx* ctx, int32_t fn, int64_t n, const double* x, const double* y,
                            double* out) {
    return guarded(ctx, [&]() -> int {
        if (fn < 0 || fn > 5 || n < 0 || (n > 0 && (!x || !out || (fn == 5 && !y))))
            return fail(ctx, SCOPI_HIP_ERR_ARG, "portable_math: bad argument");
        if (n == 0) return SCOPI_HIP_OK;
        double* d = nullptr;
        SCOPI_HIP_CHECK(hipMalloc(&d, 3 * n * sizeof(double)));
        struct Free {
            double* p;
            ~Free() { (void)hipFree(p); }
        } f{d};
        hipStream_t s = ctx->stream;
        SCOPI_HIP_CHECK(hipMemcpyAsync(d, x, n * sizeof(double), hipMemcpyHostToDevice, s));
        if (y) SCOPI_HIP_CHECK(hipMemcpyAsync(d + n, y, n * sizeof(double), hipMemcpyHostToDevice, s));
        portable_math(fn, n, d, d + n, d + 2 * n, s);
        SCOPI_HIP_CHECK(hipMemcpyAsync(out, d + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost, s));
        SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_compute_contacts(scopi_hip_ctx* ctx, const scopi_hip_params* params, int64_t* n_contacts) {
    return guarded(ctx, [&]() -> int {
        if (!params_ok(params)) return fail(ctx, SCOPI_HIP_ERR_ARG, "compute_contacts: bad params");
        ctx->nc = contacts_step(ctx, *params);
        ctx->contacts_valid = true;
        ctx->gamma_valid = false;
        if (n_contacts) *n_contacts = ctx->nc;
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_download_contacts(scopi_hip_ctx* ctx, int64_t* i, int64_t* j, double* dij, double* nij, double* pi,
                                double* pj) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->contacts_valid) return fail(ctx, SCOPI_HIP_ERR_STATE, "download_contacts: no contact list");
        const int64_t nc = ctx->nc;
        if (nc == 0) return SCOPI_HIP_OK;
        hipStream_t s = ctx->stream;
        std::vector<int32_t> t(nc);
        if (i) {
            SCOPI_HIP_CHECK(hipMemcpyAsync(t.data(), ctx->C.i, nc * sizeof(int32_t), hipMemcpyDeviceToHost, s));
            SCOPI_HIP_CHECK(hipStreamSynchronize(s));
            for (int64_t c = 0; c < nc; ++c) i[c] = t[c];
        }
        if (j) {
            SCOPI_HIP_CHECK(hipMemcpyAsync(t.data(), ctx->C.j, nc * sizeof(int32_t), hipMemcpyDeviceToHost, s));
            SCOPI_HIP_CHECK(hipStreamSynchronize(s));
            for (int64_t c = 0; c < nc; ++c) j[c] = t[c];
        }
        if (dij) {
            SCOPI_HIP_CHECK(hipMemcpyAsync(dij, ctx->C.dij, nc * sizeof(double), hipMemcpyDeviceToHost, s));
            SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        }
        if (nij) download_rows(ctx->C.nij, ctx->C.cap, nij, nc, 3, s);
        if (pi) download_rows(ctx->C.pi, ctx->C.cap, pi, nc, 3, s);
        if (pj) download_rows(ctx->C.pj, ctx->C.cap, pj, nc, 3, s);
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_upload_contacts(scopi_hip_ctx* ctx, int64_t nc, const int64_t* i, const int64_t* j, const double* dij,
                              const double* nij, const double* pi, const double* pj, const double* mu) {
    return guarded(ctx, [&]() -> int {
        if (nc < 0 || (nc > 0 && (!i || !j || !dij || !nij || !pi || !pj)))
            return fail(ctx, SCOPI_HIP_ERR_ARG, "upload_contacts: bad argument");
        const Particles& P = ctx->P;
        std::vector<int32_t> ti(nc), tj(nc);
        for (int64_t c = 0; c < nc; ++c) {
            if (i[c] < 0 || j[c] < 0 || i[c] >= P.n || j[c] >= P.n || i[c] >= j[c] || j[c] < P.n_inactive)
                return fail(ctx, SCOPI_HIP_ERR_ARG, "upload_contacts: index out of range or not i < j, j active");
            if (c > 0 && (i[c] < i[c - 1] || (i[c] == i[c - 1] && j[c] <= j[c - 1])))
                return fail(ctx, SCOPI_HIP_ERR_ARG, "upload_contacts: contacts must be sorted by (i, j)");
            ti[c] = (int32_t)i[c];
            tj[c] = (int32_t)j[c];
        }
        contacts_reserve(ctx->C, nc);
        hipStream_t s = ctx->stream;
        if (nc > 0) {
            SCOPI_HIP_CHECK(hipMemcpyAsync(ctx->C.i, ti.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice, s));
            SCOPI_HIP_CHECK(hipMemcpyAsync(ctx->C.j, tj.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice, s));
            SCOPI_HIP_CHECK(hipMemcpyAsync(ctx->C.dij, dij, nc * sizeof(double), hipMemcpyHostToDevice, s));
            upload_rows(ctx->C.nij, ctx->C.cap, nij, nc, 3, s);
            upload_rows(ctx->C.pi, ctx->C.cap, pi, nc, 3, s);
            upload_rows(ctx->C.pj, ctx->C.cap, pj, nc, 3, s);
            std::vector<double> m(nc, -1.0);
            SCOPI_HIP_CHECK(hipMemcpyAsync(ctx->C.mu, mu ? mu : m.data(), nc * sizeof(double), hipMemcpyHostToDevice, s));
            SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        }
        ctx->nc = nc;
        ctx->contacts_valid = true;
        ctx->gamma_valid = false;
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_set_projection_velocity(scopi_hip_ctx* ctx, const double* u, const double* w) {
    return guarded(ctx, [&]() -> int {
        if (!u || !w) return fail(ctx, SCOPI_HIP_ERR_ARG, "set_projection_velocity: bad argument");
        const int64_t na = ctx->P.n - ctx->P.n_inactive;
        if (ctx->proj) SCOPI_HIP_CHECK(hipFree(ctx->proj));
        ctx->proj = nullptr;
        SCOPI_HIP_CHECK(hipMalloc(&ctx->proj, 6 * std::max<int64_t>(na, 1) * sizeof(double)));
        if (na > 0) {
            SCOPI_HIP_CHECK(hipMemcpyAsync(ctx->proj, u, 3 * na * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            SCOPI_HIP_CHECK(
                hipMemcpyAsync(ctx->proj + 3 * na, w, 3 * na * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        }
        SCOPI_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        ctx->proj_n = na;
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_set_contact_gamma(scopi_hip_ctx* ctx, const double* gamma) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->contacts_valid) return fail(ctx, SCOPI_HIP_ERR_STATE, "set_contact_gamma: no contact list");
        if (!gamma && ctx->nc > 0) return fail(ctx, SCOPI_HIP_ERR_ARG, "set_contact_gamma: bad argument");
        gamma_reserve(ctx, std::max<int64_t>(ctx->nc, 1));
        if (ctx->nc > 0) {
            SCOPI_HIP_CHECK(
                hipMemcpyAsync(ctx->gamma, gamma, ctx->nc * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
            SCOPI_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        }
        ctx->gamma_valid = true;
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_download_contact_gamma(scopi_hip_ctx* ctx, double* gamma) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->contacts_valid || !ctx->gamma_valid)
            return fail(ctx, SCOPI_HIP_ERR_STATE, "download_contact_gamma: no gamma for the current contacts");
        if (!gamma && ctx->nc > 0) return fail(ctx, SCOPI_HIP_ERR_ARG, "download_contact_gamma: bad argument");
        if (ctx->nc > 0) {
            SCOPI_HIP_CHECK(
                hipMemcpyAsync(gamma, ctx->gamma, ctx->nc * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
            SCOPI_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        }
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_set_a_priori_velocity(scopi_hip_ctx* ctx, const scopi_hip_params* params, double dt) {
    return guarded(ctx, [&]() -> int {
        if (!params_ok(params)) return fail(ctx, SCOPI_HIP_ERR_ARG, "set_a_priori_velocity: bad params");
        if (params->vap == SCOPI_VAP_PROJECTION && ctx->proj_n != ctx->P.n - ctx->P.n_inactive)
            return fail(ctx, SCOPI_HIP_ERR_STATE, "set_a_priori_velocity: no projection velocities for these particles");
        const int64_t na = ctx->P.n - ctx->P.n_inactive;
        if (na > 0) hipLaunchKernelGGL(k_vap, dim3(grid_for(na)), dim3(kBlock), 0, ctx->stream, ctx->P, params->vap, dt,
                                      (const double*)ctx->proj, (const double*)(ctx->proj ? ctx->proj + 3 * na : nullptr));
        SCOPI_HIP_CHECK(hipGetLastError());
        return SCOPI_HIP_OK;
    });
}

// Contacts uploaded without mu take params->mu.
__global__ void k_fill_mu(double* mu, int64_t nc, double v) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c < nc && mu[c] < 0) mu[c] = v;
}

int scopi_hip_solve(scopi_hip_ctx* ctx, const scopi_hip_params* params, double dt, scopi_hip_solve_info* info) {
    return guarded(ctx, [&]() -> int {
        if (!params_ok(params)) return fail(ctx, SCOPI_HIP_ERR_ARG, "solve: bad params");
        if (!ctx->contacts_valid) return fail(ctx, SCOPI_HIP_ERR_STATE, "solve: no contact list on the device");
        if (ctx->nc > 0)
            hipLaunchKernelGGL(k_fill_mu, dim3(grid_for(ctx->nc)), dim3(kBlock), 0, ctx->stream, ctx->C.mu, ctx->nc,
                               params->mu);
        const bool visc = viscous_problem(params->problem);
        if (params->warm_start && !warm_problem(*params))
            return fail(ctx, SCOPI_HIP_ERR_UNSUPPORTED, "solve: warm_start with NoFriction or Friction only");
        if (visc) prepare_gamma(ctx, *params, false);
        int rc = ctx->solver.solve(ctx->P, ctx->C, ctx->nc, *params, dt, ctx->all_spheres && !ctx->any_periodic(),
                                   ctx->stream, info, visc ? ctx->gamma : nullptr, prepare_lambda(ctx, *params));
        if (rc != SCOPI_HIP_OK) return fail(ctx, rc, "solve: unsupported option");
        if (params->warm_start)
            record_lambda(ctx);
        else if (ctx->hist_kind == 1)
            ctx->hist_kind = 0;  // no multipliers for these contacts: a later warm start begins at 0
        SCOPI_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_download_solution(scopi_hip_ctx* ctx, double* u, double* omega, double* lambda) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->solver.has_solution()) return fail(ctx, SCOPI_HIP_ERR_STATE, "download_solution: no solve yet");
        const int64_t na = ctx->P.n - ctx->P.n_inactive;
        if (u && na) download_rows(ctx->solver.u(), ctx->solver.pcap(), u, na, 3, ctx->stream);
        if (omega && na) download_rows(ctx->solver.w(), ctx->solver.pcap(), omega, na, 3, ctx->stream);
        if (lambda) ctx->solver.download_lambda(lambda, ctx->stream);
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_update_and_move(scopi_hip_ctx* ctx, double dt) {
    return guarded(ctx, [&]() -> int {
        if (!ctx->solver.has_solution()) return fail(ctx, SCOPI_HIP_ERR_STATE, "update_and_move: no solve yet");
        const int64_t na = ctx->P.n - ctx->P.n_inactive;
        if (na > 0)
            hipLaunchKernelGGL(k_update_move, dim3(grid_for(na)), dim3(kBlock), 0, ctx->stream, ctx->P, ctx->solver.u(),
                               ctx->solver.w(), ctx->solver.pcap(), dt);
        periodic_wrap(ctx);
        SCOPI_HIP_CHECK(hipGetLastError());
        SCOPI_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_displacement_obstacles(scopi_hip_ctx* ctx, double dt) {
    return guarded(ctx, [&]() -> int {
        if (ctx->P.n_inactive > 0)
            hipLaunchKernelGGL(k_obstacles, dim3(grid_for(ctx->P.n_inactive)), dim3(kBlock), 0, ctx->stream, ctx->P, dt);
        SCOPI_HIP_CHECK(hipGetLastError());
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_run(scopi_hip_ctx* ctx, const scopi_hip_params* params, double dt, int64_t total_it, int64_t initial_iter,
                  scopi_hip_solve_info* info) {
    return guarded(ctx, [&]() -> int {
        if (!params_ok(params)) return fail(ctx, SCOPI_HIP_ERR_ARG, "run: bad params");
        if (params->warm_start && !warm_problem(*params))
            return fail(ctx, SCOPI_HIP_ERR_UNSUPPORTED, "run: warm_start with NoFriction or Friction only");
        if (params->vap == SCOPI_VAP_PROJECTION && ctx->proj_n != ctx->P.n - ctx->P.n_inactive)
            return fail(ctx, SCOPI_HIP_ERR_STATE, "run: no projection velocities for these particles");
        scopi_hip_solve_info acc{};
        acc.converged = 1;
        hipStream_t s = ctx->stream;
        Particles& P = ctx->P;
        const int64_t na = P.n - P.n_inactive;
        const bool visc = viscous_problem(params->problem);
        for (int64_t nite = initial_iter; nite < total_it; ++nite) {
            if (P.n_inactive > 0)
                hipLaunchKernelGGL(k_obstacles, dim3(grid_for(P.n_inactive)), dim3(kBlock), 0, s, P, dt);
            ctx->nc = contacts_step(ctx, *params);
            ctx->contacts_valid = true;
            ctx->gamma_valid = false;
            if (visc) prepare_gamma(ctx, *params, true);
            if (na > 0)
                hipLaunchKernelGGL(k_vap, dim3(grid_for(na)), dim3(kBlock), 0, s, P, params->vap, dt,
                                   (const double*)ctx->proj, (const double*)(ctx->proj ? ctx->proj + 3 * na : nullptr));
            scopi_hip_solve_info one{};
            int rc = ctx->solver.solve(P, ctx->C, ctx->nc, *params, dt, ctx->all_spheres && !ctx->any_periodic(), s, &one,
                                       visc ? ctx->gamma : nullptr, prepare_lambda(ctx, *params));
            if (rc != SCOPI_HIP_OK) return fail(ctx, rc, "run: solve failed");
            if (visc) record_history(ctx);
            if (params->warm_start)
                record_lambda(ctx);
            else if (ctx->hist_kind == 1)
                ctx->hist_kind = 0;  // no multipliers for these contacts: a later warm start begins at 0
            if (na > 0)
                hipLaunchKernelGGL(k_update_move, dim3(grid_for(na)), dim3(kBlock), 0, s, P, ctx->solver.u(),
                                   ctx->solver.w(), ctx->solver.pcap(), dt);
            periodic_wrap(ctx);
            SCOPI_HIP_CHECK(hipGetLastError());
            acc.n_contacts = one.n_contacts;
            acc.rows = one.rows;
            acc.iterations += one.iterations;
            acc.iterations_fp32 += one.iterations_fp32;
            acc.backtracks += one.backtracks;
            acc.q_applies += one.q_applies;
            acc.converged = acc.converged && one.converged;
            if (one.status != 0) acc.status = one.status;
            acc.residual = one.residual;
            acc.step = one.step;
            acc.outer_iterations += one.outer_iterations;
            acc.fixed_point_error = one.fixed_point_error;
        }
        SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        if (info) *info = acc;
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_set_box(scopi_hip_ctx* ctx, const double* lo, const double* hi, const int32_t* periodic) {
    return guarded(ctx, [&]() -> int {
        if (!lo || !hi || !periodic) return fail(ctx, SCOPI_HIP_ERR_ARG, "set_box: bad argument");
        bool any = false;
        for (int d = 0; d < ctx->dim; ++d) {
            if (periodic[d] && !(hi[d] > lo[d])) return fail(ctx, SCOPI_HIP_ERR_ARG, "set_box: periodic axis with hi <= lo");
            any = any || periodic[d];
        }
        if (any && ctx->has_se)
            return fail(ctx, SCOPI_HIP_ERR_UNSUPPORTED, "set_box: periodic axes with superellipsoids are not supported");
        for (int d = 0; d < 3; ++d) {
            ctx->box_lo[d] = d < ctx->dim ? lo[d] : 0.0;
            ctx->box_hi[d] = d < ctx->dim ? hi[d] : 0.0;
            ctx->periodic[d] = (d < ctx->dim && periodic[d]) ? 1 : 0;
        }
        ctx->contacts_valid = false;
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_set_kernel_timing(scopi_hip_ctx* ctx, int32_t enable) {
    if (!ctx) return SCOPI_HIP_ERR_ARG;
    ctx->solver.set_timing(enable != 0);
    return SCOPI_HIP_OK;
}

int scopi_hip_set_tuning(scopi_hip_ctx* ctx, int32_t key, int32_t value) {
    if (!ctx) return SCOPI_HIP_ERR_ARG;
    const int rc = ctx->solver.set_tuning(key, value);
    if (rc != SCOPI_HIP_OK) return fail(ctx, rc, "set_tuning: unknown key or value out of range");
    return SCOPI_HIP_OK;
}

int scopi_hip_kernel_timing(scopi_hip_ctx* ctx, double* out) {
    if (!ctx || !out) return SCOPI_HIP_ERR_ARG;
    ctx->solver.timing(out);
    return SCOPI_HIP_OK;
}

}  // extern "C"

// ============================================================== batches of scenes
struct scopi_hip_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    scopi_hip_params prm{};
    double dt = 0.0;
    Particles P;
    Contacts C;
    int64_t nc = 0;
    BatchSearch search;
    Solver solver;
    struct Scene {
        int64_t id = 0, n = 0, n_inactive = 0;
        int64_t ob_lo = 0, a_lo = 0;  // packed positions of its obstacles and active particles
        int64_t steps = 0;
        scopi_hip_solve_info info{};
        bool has_contacts = false;
        bool has_se = false;       // superellipsoids among its particles
        bool all_spheres = true;   // every active particle is a sphere
    };
    std::vector<Scene> scenes;  // packed order
    struct Staged {             // added, not packed yet (host copies)
        Scene sc;
        std::vector<int32_t> shape;
        std::vector<double> pos, q, size, sq, v, vd, omega, domega, f, m, J;
    };
    std::vector<Staged> staged;
    std::vector<int64_t> removed;
    int64_t next_id = 1;
    bool solved = false;
    // warm start: the last step's contacts, keyed (scene id, local i, local j), and their multipliers
    // (hist_nr rows each); the current contacts' keys and first iterate; per scene (id, first
    // obstacle, first active particle, obstacles) in packed order
    uint64_t* hist_keys = nullptr;
    uint64_t* cur_keys = nullptr;
    int64_t key_cap = 0, hist_n = 0;
    double* hist_lam = nullptr;
    double* lam0 = nullptr;
    int64_t lam_cap = 0;
    int hist_nr = 0;
    int64_t* d_tab = nullptr;
    int64_t tab_cap = 0;
    unsigned char* h_stage = nullptr;  // pinned staging of the fields of scenes being added
    int64_t stage_cap = 0;             // bytes
    bool has_se = false;       // any scene with superellipsoids: candidate search for the batch
    bool all_spheres = true;   // every active particle of the batch is a sphere
    std::string err;
};

namespace {

int bfail(scopi_hip_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    return code;
}

template <class F>
int bguarded(scopi_hip_batch* b, F&& f) {
    if (!b) return SCOPI_HIP_ERR_ARG;
    try {
        SCOPI_HIP_CHECK(hipSetDevice(b->device));
        return f();
    } catch (const hip_error& e) {
        return bfail(b, SCOPI_HIP_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.code));
    } catch (const std::exception& e) {
        return bfail(b, SCOPI_HIP_ERR_HIP, e.what());
    }
}

// every field of particle map[k] of S into particle k of D (map[k] < 0: left as is)
__global__ void __launch_bounds__(kBlock) k_gather_particles(Particles D, Particles S, const int32_t* __restrict__ map,
                                                            int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n || map[k] < 0) return;
    const int64_t o = map[k];
    D.shape[k] = S.shape[o];
    D.m[k] = S.m[o];
    for (int d = 0; d < 3; ++d) {
        D.pos[d * D.cap + k] = S.pos[d * S.cap + o];
        D.size[d * D.cap + k] = S.size[d * S.cap + o];
        D.v[d * D.cap + k] = S.v[d * S.cap + o];
        D.vd[d * D.cap + k] = S.vd[d * S.cap + o];
        D.omega[d * D.cap + k] = S.omega[d * S.cap + o];
        D.domega[d * D.cap + k] = S.domega[d * S.cap + o];
        D.f[d * D.cap + k] = S.f[d * S.cap + o];
        D.J[d * D.cap + k] = S.J[d * S.cap + o];
    }
    for (int d = 0; d < 4; ++d) D.q[d * D.cap + k] = S.q[d * S.cap + o];
    for (int d = 0; d < 2; ++d) D.sq[d * D.cap + k] = S.sq[d * S.cap + o];
}

// The fields of the staged scenes (next[kept..]) into Q.  They come after the kept scenes, so their
// obstacles form one contiguous range of Q and their active particles another: every field component
// goes up in two copies (obstacles, active particles) from one pinned staging buffer, in which each
// component is a column of the staged rows (obstacles of every staged scene, then their active
// particles, scenes in order).
void upload_staged(scopi_hip_batch* b, Particles& Q, const std::vector<scopi_hip_batch::Scene>& next, size_t kept) {
    int64_t nob = 0, nac = 0;
    for (size_t k = kept; k < next.size(); ++k) {
        nob += next[k].n_inactive;
        nac += next[k].n - next[k].n_inactive;
    }
    const int64_t rows = nob + nac;
    if (rows == 0) return;
    using Staged = scopi_hip_batch::Staged;
    struct Field {
        double* dev;
        int w;
        const std::vector<double> Staged::*host;
    };
    const Field fields[] = {{Q.pos, 3, &Staged::pos},       {Q.q, 4, &Staged::q},         {Q.size, 3, &Staged::size},
                            {Q.sq, 2, &Staged::sq},         {Q.v, 3, &Staged::v},         {Q.vd, 3, &Staged::vd},
                            {Q.omega, 3, &Staged::omega},   {Q.domega, 3, &Staged::domega}, {Q.f, 3, &Staged::f},
                            {Q.m, 1, &Staged::m},           {Q.J, 3, &Staged::J}};
    int cols = 0;
    for (const auto& f : fields) cols += f.w;
    const int64_t bytes = (int64_t)cols * rows * (int64_t)sizeof(double) + rows * (int64_t)sizeof(int32_t);
    hipStream_t s = b->stream;
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));  // the staging buffer may still feed earlier copies
    if (bytes > b->stage_cap) {
        if (b->h_stage) SCOPI_HIP_CHECK(hipHostFree(b->h_stage));
        b->stage_cap = bytes + bytes / 2;
        SCOPI_HIP_CHECK(hipHostMalloc(&b->h_stage, b->stage_cap));
    }
    double* col = (double*)b->h_stage;
    int32_t* shp = (int32_t*)(col + (int64_t)cols * rows);
    // row of staged scene k's obstacle / active particle t within the staging columns
    int64_t ob = 0, ac = nob;
    for (size_t k = kept; k < next.size(); ++k) {
        const Staged& st = b->staged[k - kept];
        const int64_t ni = next[k].n_inactive, n = next[k].n;
        for (int64_t t = 0; t < n; ++t) {
            const int64_t r = t < ni ? ob + t : ac + (t - ni);
            shp[r] = st.shape[t];
            int64_t c = 0;
            for (const auto& f : fields) {
                const std::vector<double>& h = st.*(f.host);
                for (int d = 0; d < f.w; ++d, ++c) col[c * rows + r] = h[(size_t)t * f.w + d];
            }
        }
        ob += ni;
        ac += n - ni;
    }
    const int64_t ob0 = next[kept].ob_lo, ac0 = next[kept].a_lo;
    auto up = [&](void* dst, const void* src, int64_t nbytes) {
        if (nbytes > 0) SCOPI_HIP_CHECK(hipMemcpyAsync(dst, src, nbytes, hipMemcpyHostToDevice, s));
    };
    up(Q.shape + ob0, shp, nob * (int64_t)sizeof(int32_t));
    up(Q.shape + ac0, shp + nob, nac * (int64_t)sizeof(int32_t));
    int64_t c = 0;
    for (const auto& f : fields)
        for (int d = 0; d < f.w; ++d, ++c) {
            up(f.dev + d * Q.cap + ob0, col + c * rows, nob * (int64_t)sizeof(double));
            up(f.dev + d * Q.cap + ac0, col + c * rows + nob, nac * (int64_t)sizeof(double));
        }
    SCOPI_HIP_CHECK(hipStreamSynchronize(s));
}

// Packs the kept scenes (in order) and the staged ones: obstacles of every scene
// first, then the active particles of every scene.
void batch_pack(scopi_hip_batch* b) {
    if (b->staged.empty() && b->removed.empty()) return;
    std::vector<scopi_hip_batch::Scene> next;
    for (const auto& sc : b->scenes)
        if (std::find(b->removed.begin(), b->removed.end(), sc.id) == b->removed.end()) next.push_back(sc);
    const size_t kept = next.size();
    for (const auto& st : b->staged) next.push_back(st.sc);
    int64_t NI = 0, N = 0;
    for (const auto& sc : next) {
        NI += sc.n_inactive;
        N += sc.n;
    }
    if (N > INT32_MAX / 2) throw hip_error(hipErrorInvalidValue, "batch: too many particles");
    std::vector<int32_t> map((size_t)std::max<int64_t>(N, 1), -1);
    int64_t ob = 0, ac = NI;
    for (size_t k = 0; k < next.size(); ++k) {
        auto& sc = next[k];
        const int64_t na = sc.n - sc.n_inactive;
        if (k < kept) {
            for (int64_t t = 0; t < sc.n_inactive; ++t) map[ob + t] = (int32_t)(sc.ob_lo + t);
            for (int64_t t = 0; t < na; ++t) map[ac + t] = (int32_t)(sc.a_lo + t);
        }
        sc.ob_lo = ob;
        sc.a_lo = ac;
        ob += sc.n_inactive;
        ac += na;
    }
    hipStream_t s = b->stream;
    Particles Q{};
    Q.dim = 3;
    particles_reserve(Q, std::max<int64_t>(N, 1));
    Q.n = N;
    Q.n_inactive = NI;
    if (N > 0 && b->P.n > 0) {
        int32_t* dmap = nullptr;
        SCOPI_HIP_CHECK(hipMalloc(&dmap, N * sizeof(int32_t)));
        SCOPI_HIP_CHECK(hipMemcpyAsync(dmap, map.data(), N * sizeof(int32_t), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_gather_particles, dim3(grid_for(N)), dim3(kBlock), 0, s, Q, b->P, (const int32_t*)dmap, N);
        SCOPI_HIP_CHECK(hipGetLastError());
        SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        SCOPI_HIP_CHECK(hipFree(dmap));
    }
    upload_staged(b, Q, next, kept);
    particles_free(b->P);
    b->P = Q;
    b->scenes.swap(next);
    b->staged.clear();
    b->removed.clear();
    // scene table of the search: work units and cell keys in scene order
    std::vector<SceneRanges> R(b->scenes.size());
    int64_t u = 0, cell = 0;
    for (size_t k = 0; k < b->scenes.size(); ++k) {
        const auto& sc = b->scenes[k];
        const int64_t na = sc.n - sc.n_inactive;
        R[k].ob_lo = sc.ob_lo;
        R[k].ob_hi = sc.ob_lo + sc.n_inactive;
        R[k].a_lo = sc.a_lo;
        R[k].a_hi = sc.a_lo + na;
        R[k].u_lo = u;
        R[k].cell0 = cell;
        R[k].cells = 2 * na + 64;
        u += (sc.n_inactive + 1) * na;
        cell += R[k].cells;
        b->scenes[k].has_contacts = false;
    }
    b->search.set_scenes(R, NI, N - NI, s);
    b->nc = 0;
    b->solved = false;
    b->has_se = false;
    b->all_spheres = true;
    for (const auto& sc : b->scenes) {
        b->has_se = b->has_se || sc.has_se;
        b->all_spheres = b->all_spheres && sc.all_spheres;
    }
}

// Warm start in a batch: contact c's key (scene id, local i, local j), 24 + 20 + 20 bits.  The packed
// contacts are in scene order (ids increase along it) and each scene's in (i, j) order, with its
// obstacles before its active particles as in the scene's own numbering, so the keys are sorted.
constexpr int kKeyIdBits = 24, kKeyLocalBits = 20;
__global__ void __launch_bounds__(kBlock) k_batch_keys(const int32_t* __restrict__ ci, const int32_t* __restrict__ cj,
                                                      int64_t nc, const int64_t* __restrict__ crange, int K,
                                                      const int64_t* __restrict__ tab, int64_t NI,
                                                      uint64_t* __restrict__ keys) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    int lo = 0, hi = K;  // the last scene whose contacts start at or before c
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (crange[2 * mid] <= c)
            lo = mid;
        else
            hi = mid;
    }
    const int64_t* t = tab + 4 * lo;
    auto local = [&](int64_t g) { return (uint64_t)(g < NI ? g - t[1] : t[3] + (g - t[2])); };
    keys[c] = ((uint64_t)t[0] << (2 * kKeyLocalBits)) | (local(ci[c]) << kKeyLocalBits) | local(cj[c]);
}

// lam0 of the current contacts: the rows of the history's last equal key, 0 for a new contact
__global__ void __launch_bounds__(kBlock) k_transfer_keys(const uint64_t* __restrict__ cur, int64_t nc,
                                                         const uint64_t* __restrict__ keys,
                                                         const double* __restrict__ hl, int64_t hn, int nr,
                                                         double* __restrict__ lam0) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nc) return;
    const uint64_t key = cur[c];
    int64_t lo = 0, hi = hn;  // upper bound of key
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] <= key)
            lo = mid + 1;
        else
            hi = mid;
    }
    const bool found = lo > 0 && keys[lo - 1] == key;
    for (int d = 0; d < nr; ++d) lam0[c * nr + d] = found ? hl[(lo - 1) * nr + d] : 0.0;
}

void batch_warm_reserve(scopi_hip_batch* b, int64_t nc, int nr) {
    if (nc > b->key_cap) {
        for (uint64_t** p : {&b->hist_keys, &b->cur_keys}) {
            uint64_t* q = nullptr;
            SCOPI_HIP_CHECK(hipMalloc(&q, (nc + nc / 4 + 1024) * sizeof(uint64_t)));
            if (*p && b->hist_n > 0 && p == &b->hist_keys)
                SCOPI_HIP_CHECK(hipMemcpyAsync(q, *p, b->hist_n * sizeof(uint64_t), hipMemcpyDeviceToDevice, b->stream));
            SCOPI_HIP_CHECK(hipStreamSynchronize(b->stream));
            if (*p) SCOPI_HIP_CHECK(hipFree(*p));
            *p = q;
        }
        b->key_cap = nc + nc / 4 + 1024;
    }
    if (nr * nc > b->lam_cap) {
        for (double** p : {&b->hist_lam, &b->lam0}) {
            double* q = nullptr;
            SCOPI_HIP_CHECK(hipMalloc(&q, (nr * nc + nr * nc / 4 + 1024) * sizeof(double)));
            if (*p && b->hist_n > 0 && p == &b->hist_lam)
                SCOPI_HIP_CHECK(hipMemcpyAsync(q, *p, b->hist_nr * b->hist_n * sizeof(double), hipMemcpyDeviceToDevice,
                                               b->stream));
            SCOPI_HIP_CHECK(hipStreamSynchronize(b->stream));
            if (*p) SCOPI_HIP_CHECK(hipFree(*p));
            *p = q;
        }
        b->lam_cap = nr * nc + nr * nc / 4 + 1024;
    }
}

scopi_hip_batch::Scene* batch_find(scopi_hip_batch* b, int64_t id) {
    for (auto& sc : b->scenes)
        if (sc.id == id) return &sc;
    return nullptr;
}

}  // namespace

extern "C" {

int scopi_hip_batch_create(const scopi_hip_params* params, double dt, scopi_hip_batch** out) {
    if (!out || !params_ok(params) || !(dt > 0) || params->vap == SCOPI_VAP_PROJECTION) return SCOPI_HIP_ERR_ARG;
    if (params->problem == SCOPI_PROBLEM_FRICTION_FIXED_POINT || viscous_problem(params->problem))
        return SCOPI_HIP_ERR_UNSUPPORTED;
    *out = nullptr;
    auto* b = new scopi_hip_batch();
    b->prm = *params;
    b->dt = dt;
    b->P.dim = 3;
    try {
        SCOPI_HIP_CHECK(hipGetDevice(&b->device));
        SCOPI_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        contacts_reserve(b->C, 1024);
    } catch (const hip_error& e) {
        std::fprintf(stderr, "scopi_hip_batch_create: %s: %s\n", e.what, hipGetErrorString(e.code));
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return SCOPI_HIP_ERR_HIP;
    }
    *out = b;
    return SCOPI_HIP_OK;
}

int scopi_hip_batch_destroy(scopi_hip_batch* b) {
    if (!b) return SCOPI_HIP_ERR_ARG;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    particles_free(b->P);
    contacts_free(b->C);
    for (void* p : {(void*)b->hist_keys, (void*)b->cur_keys, (void*)b->hist_lam, (void*)b->lam0, (void*)b->d_tab})
        if (p) (void)hipFree(p);
    if (b->h_stage) (void)hipHostFree(b->h_stage);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return SCOPI_HIP_OK;
}

const char* scopi_hip_batch_last_error(const scopi_hip_batch* b) { return b ? b->err.c_str() : "null batch"; }

int scopi_hip_batch_add_scene(scopi_hip_batch* b, int64_t n, int64_t n_inactive, const int32_t* shape,
                              const double* pos, const double* q, const double* size, const double* squareness,
                              const double* v, const double* vd, const double* omega, const double* domega,
                              const double* f, const double* m, const double* J, int64_t* scene_id) {
    return bguarded(b, [&]() -> int {
        if (n < 0 || n_inactive < 0 || n_inactive > n || !shape || !pos || !q || !size || !v || !vd || !omega ||
            !domega || !f || !m || !J || !scene_id)
            return bfail(b, SCOPI_HIP_ERR_ARG, "batch_add_scene: bad argument");
        scopi_hip_batch::Staged st;
        for (int64_t k = 0; k < n; ++k) {
            const int s = shape[k];
            const bool active = k >= n_inactive;
            if (s != SCOPI_SHAPE_SPHERE && s != SCOPI_SHAPE_PLANE && s != SCOPI_SHAPE_SUPERELLIPSOID)
                return bfail(b, SCOPI_HIP_ERR_UNSUPPORTED, "batch_add_scene: 3D spheres, planes, superellipsoids");
            if (active && s == SCOPI_SHAPE_PLANE)
                return bfail(b, SCOPI_HIP_ERR_UNSUPPORTED, "batch_add_scene: planes must be obstacles");
            st.sc.has_se = st.sc.has_se || s == SCOPI_SHAPE_SUPERELLIPSOID;
            if (active) st.sc.all_spheres = st.sc.all_spheres && s == SCOPI_SHAPE_SPHERE;
        }
        if (st.sc.has_se && !squareness)
            return bfail(b, SCOPI_HIP_ERR_ARG, "batch_add_scene: superellipsoids need squareness");
        st.sc.id = b->next_id++;
        st.sc.n = n;
        st.sc.n_inactive = n_inactive;
        st.shape.assign(shape, shape + n);
        auto cp = [&](std::vector<double>& dst, const double* src, int w, double fill) {
            if (src)
                dst.assign(src, src + n * w);
            else
                dst.assign((size_t)n * w, fill);
        };
        cp(st.pos, pos, 3, 0.0);
        cp(st.q, q, 4, 0.0);
        cp(st.size, size, 3, 0.0);
        cp(st.sq, squareness, 2, 1.0);
        cp(st.v, v, 3, 0.0);
        cp(st.vd, vd, 3, 0.0);
        cp(st.omega, omega, 3, 0.0);
        cp(st.domega, domega, 3, 0.0);
        cp(st.f, f, 3, 0.0);
        cp(st.m, m, 1, 0.0);
        cp(st.J, J, 3, 0.0);
        *scene_id = st.sc.id;
        b->staged.push_back(std::move(st));
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_batch_remove_scene(scopi_hip_batch* b, int64_t scene_id) {
    return bguarded(b, [&]() -> int {
        for (size_t k = 0; k < b->staged.size(); ++k)
            if (b->staged[k].sc.id == scene_id) {
                b->staged.erase(b->staged.begin() + k);
                return SCOPI_HIP_OK;
            }
        if (!batch_find(b, scene_id) ||
            std::find(b->removed.begin(), b->removed.end(), scene_id) != b->removed.end())
            return bfail(b, SCOPI_HIP_ERR_ARG, "batch_remove_scene: unknown scene id");
        b->removed.push_back(scene_id);
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_batch_set_tuning(scopi_hip_batch* b, int32_t key, int32_t value) {
    return bguarded(b, [&]() -> int {
        if (key == SCOPI_TUNE_GRAPH_ROUNDS) return bfail(b, SCOPI_HIP_ERR_ARG, "batch_set_tuning: no graph rounds");
        const int rc = b->solver.set_tuning(key, value);
        if (rc != SCOPI_HIP_OK) return bfail(b, rc, "batch_set_tuning: unknown key or value out of range");
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_batch_step(scopi_hip_batch* b, int64_t n_steps) {
    return bguarded(b, [&]() -> int {
        if (n_steps < 0) return bfail(b, SCOPI_HIP_ERR_ARG, "batch_step: n_steps < 0");
        batch_pack(b);
        const int K = (int)b->scenes.size();
        if (K == 0 || n_steps == 0) return SCOPI_HIP_OK;
        hipStream_t s = b->stream;
        Particles& P = b->P;
        const int64_t na = P.n - P.n_inactive;
        SceneSet S;
        S.nscene = K;
        S.alo.resize(K);
        S.ahi.resize(K);
        for (int k = 0; k < K; ++k) {
            S.alo[k] = b->scenes[k].a_lo - P.n_inactive;
            S.ahi[k] = S.alo[k] + (b->scenes[k].n - b->scenes[k].n_inactive);
        }
        const bool warm = b->prm.warm_start != 0;
        const int nr = b->prm.problem == SCOPI_PROBLEM_FRICTION ? 3 : 1;
        if (warm) {
            // the scenes' table for the contact keys (packed order)
            for (const auto& sc : b->scenes)
                if (sc.n >= ((int64_t)1 << kKeyLocalBits) || sc.id >= ((int64_t)1 << kKeyIdBits))
                    return bfail(b, SCOPI_HIP_ERR_UNSUPPORTED,
                                 "batch_step: warm start needs scenes of fewer than 2^20 particles (ids below 2^24)");
            if (K > b->tab_cap) {
                if (b->d_tab) SCOPI_HIP_CHECK(hipFree(b->d_tab));
                b->tab_cap = K + K / 2 + 16;
                SCOPI_HIP_CHECK(hipMalloc(&b->d_tab, 4 * b->tab_cap * sizeof(int64_t)));
            }
            std::vector<int64_t> tab(4 * (size_t)K);
            for (int k = 0; k < K; ++k) {
                const auto& sc = b->scenes[k];
                tab[4 * k] = sc.id, tab[4 * k + 1] = sc.ob_lo, tab[4 * k + 2] = sc.a_lo, tab[4 * k + 3] = sc.n_inactive;
            }
            SCOPI_HIP_CHECK(hipMemcpyAsync(b->d_tab, tab.data(), tab.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
            SCOPI_HIP_CHECK(hipStreamSynchronize(s));
            if (b->hist_nr != nr) b->hist_n = 0;
        }
        std::vector<scopi_hip_solve_info> infos;
        for (int64_t it = 0; it < n_steps; ++it) {
            if (P.n_inactive > 0)
                hipLaunchKernelGGL(k_obstacles, dim3(grid_for(P.n_inactive)), dim3(kBlock), 0, s, P, b->dt);
            b->nc = b->search.run(P, b->C, b->prm, s, !b->has_se);
            if (na > 0)
                hipLaunchKernelGGL(k_vap, dim3(grid_for(na)), dim3(kBlock), 0, s, P, b->prm.vap, b->dt,
                                   (const double*)nullptr, (const double*)nullptr);
            S.d_crange = b->search.d_crange();
            const double* lam0 = nullptr;
            if (warm && b->nc > 0) {
                batch_warm_reserve(b, b->nc, nr);
                hipLaunchKernelGGL(k_batch_keys, dim3(grid_for(b->nc)), dim3(kBlock), 0, s, (const int32_t*)b->C.i,
                                   (const int32_t*)b->C.j, b->nc, S.d_crange, K, (const int64_t*)b->d_tab,
                                   P.n_inactive, b->cur_keys);
                if (b->hist_n > 0) {
                    hipLaunchKernelGGL(k_transfer_keys, dim3(grid_for(b->nc)), dim3(kBlock), 0, s,
                                       (const uint64_t*)b->cur_keys, b->nc, (const uint64_t*)b->hist_keys,
                                       (const double*)b->hist_lam, b->hist_n, nr, b->lam0);
                    lam0 = b->lam0;
                }
                SCOPI_HIP_CHECK(hipGetLastError());
            }
            int rc = b->solver.solve_scenes(P, b->C, b->nc, b->prm, b->dt, b->all_spheres, s, S, infos, lam0);
            if (rc != SCOPI_HIP_OK) return bfail(b, rc, "batch_step: solve failed");
            if (warm) {
                // this step's contacts and multipliers become the history
                if (b->nc > 0) {
                    std::swap(b->hist_keys, b->cur_keys);
                    SCOPI_HIP_CHECK(hipMemcpyAsync(b->hist_lam, b->solver.final_lambda(), nr * b->nc * sizeof(double),
                                                   hipMemcpyDeviceToDevice, s));
                }
                b->hist_n = b->nc;
                b->hist_nr = nr;
            }
            if (na > 0)
                hipLaunchKernelGGL(k_update_move, dim3(grid_for(na)), dim3(kBlock), 0, s, P, b->solver.u(),
                                   b->solver.w(), b->solver.pcap(), b->dt);
            SCOPI_HIP_CHECK(hipGetLastError());
            for (int k = 0; k < K; ++k) {
                b->scenes[k].info = infos[k];
                b->scenes[k].steps += 1;
                b->scenes[k].has_contacts = true;
            }
            b->solved = true;
        }
        SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_batch_size(scopi_hip_batch* b, int64_t* n_scenes, int64_t* n_particles, int64_t* n_contacts) {
    return bguarded(b, [&]() -> int {
        batch_pack(b);
        if (n_scenes) *n_scenes = (int64_t)b->scenes.size();
        if (n_particles) *n_particles = b->P.n;
        if (n_contacts) *n_contacts = b->nc;
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_batch_download_scene(scopi_hip_batch* b, int64_t scene_id, double* pos, double* q, double* v,
                                   double* omega) {
    return bguarded(b, [&]() -> int {
        batch_pack(b);
        const auto* sc = batch_find(b, scene_id);
        if (!sc) return bfail(b, SCOPI_HIP_ERR_ARG, "batch_download_scene: unknown scene id");
        const Particles& P = b->P;
        const int64_t ni = sc->n_inactive, na = sc->n - ni;
        struct F {
            double* h;
            const double* d;
            int w;
        };
        for (const F& f : {F{pos, P.pos, 3}, F{q, P.q, 4}, F{v, P.v, 3}, F{omega, P.omega, 3}}) {
            if (!f.h) continue;
            if (ni) download_rows(f.d + sc->ob_lo, P.cap, f.h, ni, f.w, b->stream, 0);
            if (na) download_rows(f.d + sc->a_lo, P.cap, f.h, na, f.w, b->stream, ni);
        }
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_batch_scene_info(scopi_hip_batch* b, int64_t scene_id, int64_t* steps, scopi_hip_solve_info* info) {
    return bguarded(b, [&]() -> int {
        batch_pack(b);
        const auto* sc = batch_find(b, scene_id);
        if (!sc) return bfail(b, SCOPI_HIP_ERR_ARG, "batch_scene_info: unknown scene id");
        if (steps) *steps = sc->steps;
        if (info) *info = sc->info;
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_batch_download_contacts(scopi_hip_batch* b, int64_t scene_id, int64_t* n_contacts, int64_t* i,
                                      int64_t* j, double* dij, double* nij, double* pi, double* pj) {
    return bguarded(b, [&]() -> int {
        batch_pack(b);
        const auto* sc = batch_find(b, scene_id);
        if (!sc || !n_contacts) return bfail(b, SCOPI_HIP_ERR_ARG, "batch_download_contacts: bad argument");
        if (!sc->has_contacts) return bfail(b, SCOPI_HIP_ERR_STATE, "batch_download_contacts: no step since packing");
        const int k = (int)(sc - b->scenes.data());
        const int64_t lo = b->search.crange()[2 * k], hi = b->search.crange()[2 * k + 1], nc = hi - lo;
        *n_contacts = nc;
        if (nc == 0) return SCOPI_HIP_OK;
        hipStream_t s = b->stream;
        const int64_t ni = sc->n_inactive;
        auto local = [&](int32_t g) -> int64_t { return g < b->P.n_inactive ? g - sc->ob_lo : ni + (g - sc->a_lo); };
        std::vector<int32_t> t(nc);
        for (auto pr : {std::make_pair(i, b->C.i), std::make_pair(j, b->C.j)}) {
            if (!pr.first) continue;
            SCOPI_HIP_CHECK(hipMemcpyAsync(t.data(), pr.second + lo, nc * sizeof(int32_t), hipMemcpyDeviceToHost, s));
            SCOPI_HIP_CHECK(hipStreamSynchronize(s));
            for (int64_t c = 0; c < nc; ++c) pr.first[c] = local(t[c]);
        }
        if (dij) {
            SCOPI_HIP_CHECK(hipMemcpyAsync(dij, b->C.dij + lo, nc * sizeof(double), hipMemcpyDeviceToHost, s));
            SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        }
        if (nij) download_rows(b->C.nij + lo, b->C.cap, nij, nc, 3, s);
        if (pi) download_rows(b->C.pi + lo, b->C.cap, pi, nc, 3, s);
        if (pj) download_rows(b->C.pj + lo, b->C.cap, pj, nc, 3, s);
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_batch_download_solution(scopi_hip_batch* b, int64_t scene_id, double* u, double* omega,
                                      double* lambda) {
    return bguarded(b, [&]() -> int {
        batch_pack(b);
        const auto* sc = batch_find(b, scene_id);
        if (!sc) return bfail(b, SCOPI_HIP_ERR_ARG, "batch_download_solution: unknown scene id");
        if (!b->solved || !sc->has_contacts) return bfail(b, SCOPI_HIP_ERR_STATE, "batch_download_solution: no solve");
        const int64_t na = sc->n - sc->n_inactive, a0 = sc->a_lo - b->P.n_inactive;
        if (u && na) download_rows(b->solver.u() + a0, b->solver.pcap(), u, na, 3, b->stream);
        if (omega && na) download_rows(b->solver.w() + a0, b->solver.pcap(), omega, na, 3, b->stream);
        if (lambda) {
            const int k = (int)(sc - b->scenes.data());
            b->solver.download_lambda_range(lambda, b->search.crange()[2 * k], b->search.crange()[2 * k + 1],
                                            b->stream);
        }
        return SCOPI_HIP_OK;
    });
}

int64_t scopi_hip_scene_footprint(int64_t n, int64_t n_inactive, const scopi_hip_params* params) {
    if (n < 0 || n_inactive < 0 || n_inactive > n) return -1;
    const int64_t na = n - n_inactive;
    const bool fp32 = params && params->precision == SCOPI_PRECISION_FP32_REFINE;
    const int64_t nc = 12 * na;
    // particles: state (31 doubles + tag), solver per particle (M^-1 and three slots of w, 6 values
    // each, per precision), search (keys, units, cells)
    const int64_t per_particle = 31 * 8 + 4;
    const int64_t per_active =
        (6 + 6 + 6) * 4 + 6 * 8 + (6 + 3 * 6) * 8 * (fp32 ? 2 : 1) + 3 * 4 + 8 * (n_inactive + 1) + 2 * 4 * 3;
    // contacts: record (96 B), B and its vectors (488 B FP64, 244 B more for FP32), transposes (24 B)
    const int64_t per_contact = 96 + 488 + (fp32 ? 244 : 0) + 24;
    const int64_t bytes = n * per_particle + na * per_active + nc * per_contact;
    return bytes + bytes / 4;
}

}  // extern "C"
