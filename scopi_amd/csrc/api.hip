// api.hip -- the extern "C" layer (include/scopi_hip.h): one context per scene,
// the per-particle kernels of the vap slot and of the integration, and the
// device-resident time loop (ScopiSolver::run, include/scopi/solver.hpp:266-303).
#include <cmath>
#include <cstdio>
#include <cstring>
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
};

namespace {

// vap_fixed (vap/vap_fixed.hpp:46-55) / vap_fpd (vap/vap_fpd.hpp:70-82, src/vap/vap_fpd.cpp:6-26)
__global__ void __launch_bounds__(kBlock) k_vap(Particles P, int kind, double dt) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t k = P.n_inactive + t;
    if (k >= P.n) return;
    const int64_t c = P.cap;
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

bool params_ok(const scopi_hip_params* p) {
    return p && (p->problem == SCOPI_PROBLEM_NOFRICTION || p->problem == SCOPI_PROBLEM_FRICTION) &&
           (p->contact_method == SCOPI_CONTACT_KDTREE || p->contact_method == SCOPI_CONTACT_BRUTE_FORCE) &&
           (p->vap == SCOPI_VAP_FIXED || p->vap == SCOPI_VAP_FPD) &&
           (p->method == SCOPI_METHOD_APGD || p->method == SCOPI_METHOD_PGD) &&
           (p->precision == SCOPI_PRECISION_FP64 || p->precision == SCOPI_PRECISION_FP32_REFINE) &&
           p->warm_start == 0;
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
    return SCOPI_HIP_OK;
}

int scopi_hip_create(int32_t dim, int64_t capacity_particles, int64_t capacity_contacts, scopi_hip_ctx** out) {
    if (!out || (dim != 2 && dim != 3) || capacity_particles < 0 || capacity_contacts < 0) return SCOPI_HIP_ERR_ARG;
    *out = nullptr;
    auto* ctx = new scopi_hip_ctx();
    ctx->dim = dim;
    ctx->P.dim = dim;
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
        if (has_se && ctx->dim != 3)
            return fail(ctx, SCOPI_HIP_ERR_UNSUPPORTED, "upload_state: superellipsoids are 3D objects here");
        ctx->has_se = has_se;
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
                return fail(ctx, SCOPI_HIP_ERR_UNSUPPORTED, "closest_points: 3D spheres, planes, superellipsoids");
        ci[k] = (int32_t)k;
        cj[k] = (int32_t)(n + k);
    }
    hipStream_t s = ctx->stream;
    Particles T{};
    T.dim = 3;
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
        ctx->nc = ctx->search.run(ctx->P, ctx->C, *params, ctx->stream, !ctx->has_se);
        ctx->contacts_valid = true;
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
        return SCOPI_HIP_OK;
    });
}

int scopi_hip_set_a_priori_velocity(scopi_hip_ctx* ctx, const scopi_hip_params* params, double dt) {
    return guarded(ctx, [&]() -> int {
        if (!params_ok(params)) return fail(ctx, SCOPI_HIP_ERR_ARG, "set_a_priori_velocity: bad params");
        const int64_t na = ctx->P.n - ctx->P.n_inactive;
        if (na > 0) hipLaunchKernelGGL(k_vap, dim3(grid_for(na)), dim3(kBlock), 0, ctx->stream, ctx->P, params->vap, dt);
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
        int rc = ctx->solver.solve(ctx->P, ctx->C, ctx->nc, *params, dt, ctx->all_spheres, ctx->stream, info);
        if (rc != SCOPI_HIP_OK) return fail(ctx, rc, "solve: unsupported option");
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
        scopi_hip_solve_info acc{};
        acc.converged = 1;
        hipStream_t s = ctx->stream;
        Particles& P = ctx->P;
        const int64_t na = P.n - P.n_inactive;
        for (int64_t nite = initial_iter; nite < total_it; ++nite) {
            if (P.n_inactive > 0)
                hipLaunchKernelGGL(k_obstacles, dim3(grid_for(P.n_inactive)), dim3(kBlock), 0, s, P, dt);
            ctx->nc = ctx->search.run(P, ctx->C, *params, s, !ctx->has_se);
            ctx->contacts_valid = true;
            if (na > 0) hipLaunchKernelGGL(k_vap, dim3(grid_for(na)), dim3(kBlock), 0, s, P, params->vap, dt);
            scopi_hip_solve_info one{};
            int rc = ctx->solver.solve(P, ctx->C, ctx->nc, *params, dt, ctx->all_spheres, s, &one);
            if (rc != SCOPI_HIP_OK) return fail(ctx, rc, "run: solve failed");
            if (na > 0)
                hipLaunchKernelGGL(k_update_move, dim3(grid_for(na)), dim3(kBlock), 0, s, P, ctx->solver.u(),
                                   ctx->solver.w(), ctx->solver.pcap(), dt);
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
        }
        SCOPI_HIP_CHECK(hipStreamSynchronize(s));
        if (info) *info = acc;
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
