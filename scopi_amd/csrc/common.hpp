// common.hpp -- device-side types and geometry shared by the contact-step kernels.
//
// All geometry is FP64.  Closest-point formulas keep the reference's operation
// order and are compiled without FMA contraction (the pragma below), so the two
// contact predicates (squared centre distance < kd_tree_radius, dij < dmax) are
// evaluated with the same roundings as the CPU restatement.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/scopi_hip.h"

namespace scopi_hip {

constexpr int kBlock = 256;

struct d3 {
    double x, y, z;
};

__host__ __device__ inline d3 mk(double x, double y, double z) { return d3{x, y, z}; }

// Particle state, structure of arrays, `cap` entries per component.
struct Particles {
    int64_t n = 0, n_inactive = 0, cap = 0;
    int dim = 3;
    int32_t* shape = nullptr;
    double* pos = nullptr;     // 3*cap
    double* q = nullptr;       // 4*cap (w, x, y, z)
    double* size = nullptr;    // 3*cap
    double* sq = nullptr;      // 2*cap (e, n)
    double* v = nullptr;       // 3*cap
    double* vd = nullptr;      // 3*cap
    double* omega = nullptr;   // 3*cap
    double* domega = nullptr;  // 3*cap
    double* f = nullptr;       // 3*cap
    double* m = nullptr;       // cap
    double* J = nullptr;       // 3*cap
};

// Contact list, structure of arrays (sorted by (i, j)).
struct Contacts {
    int64_t cap = 0;
    int32_t* i = nullptr;
    int32_t* j = nullptr;
    double* dij = nullptr;
    double* nij = nullptr;  // 3*cap
    double* pi = nullptr;   // 3*cap
    double* pj = nullptr;   // 3*cap
    double* mu = nullptr;   // cap
};

__device__ inline d3 load3(const double* a, int64_t cap, int64_t k) { return d3{a[k], a[cap + k], a[2 * cap + k]}; }
__device__ inline void store3(double* a, int64_t cap, int64_t k, d3 v) {
    a[k] = v.x;
    a[cap + k] = v.y;
    a[2 * cap + k] = v.z;
}

// rotation_matrix<3> (include/scopi/quaternion.hpp:86-124), q scalar-first.
struct m3 {
    double a[3][3];
};

__device__ inline m3 rotation3(double qw, double qx, double qy, double qz) {
#pragma clang fp contract(off)
    const double x = qw, y = qx, z = qy, w = qz;
    m3 r;
    r.a[0][0] = 1 - 2 * z * z - 2 * w * w;
    r.a[0][1] = 2 * y * z - 2 * x * w;
    r.a[0][2] = 2 * y * w + 2 * x * z;
    r.a[1][0] = 2 * y * z + 2 * x * w;
    r.a[1][1] = 1 - 2 * y * y - 2 * w * w;
    r.a[1][2] = 2 * z * w - 2 * x * y;
    r.a[2][0] = 2 * y * w - 2 * x * z;
    r.a[2][1] = 2 * z * w + 2 * x * y;
    r.a[2][2] = 1 - 2 * y * y - 2 * z * z;
    return r;
}

// First column of rotation_matrix<dim>(q): plane / segment normal
// (plane.hpp:142-147; 2D form quaternion.hpp:65-84).
__device__ inline d3 first_column(int dim, double qw, double qx, double qy, double qz) {
#pragma clang fp contract(off)
    if (dim == 2) {
        const double x = qw, w = qz;
        return d3{1 - 2 * w * w, 2 * x * w, 0.0};
    }
    m3 r = rotation3(qw, qx, qy, qz);
    return d3{r.a[0][0], r.a[1][0], r.a[2][0]};
}

__device__ inline double sgn(double x) { return (x > 0) ? 1.0 : ((x < 0) ? -1.0 : 0.0); }

template <int DIM>
__device__ inline double dotd(d3 a, d3 b) {
#pragma clang fp contract(off)
    double s = 0.0;
    s += a.x * b.x;
    s += a.y * b.y;
    if (DIM == 3) s += a.z * b.z;
    return s;
}

struct CP {
    double dij;
    d3 nij, pi, pj;
};

// closest_points(sphere, sphere): closest_points.hpp:36-50
template <int DIM>
__device__ inline CP cp_sphere_sphere(d3 ci, double ri, d3 cj, double rj) {
#pragma clang fp contract(off)
    d3 d = mk(cj.x - ci.x, cj.y - ci.y, DIM == 3 ? cj.z - ci.z : 0.0);
    double nd = sqrt(dotd<DIM>(d, d));
    d3 u = mk(d.x / nd, d.y / nd, DIM == 3 ? d.z / nd : 0.0);
    CP c;
    c.pi = mk(ci.x + ri * u.x, ci.y + ri * u.y, DIM == 3 ? ci.z + ri * u.z : 0.0);
    c.pj = mk(cj.x - rj * u.x, cj.y - rj * u.y, DIM == 3 ? cj.z - rj * u.z : 0.0);
    d3 e = mk(c.pj.x - cj.x, c.pj.y - cj.y, DIM == 3 ? c.pj.z - cj.z : 0.0);
    double ne = sqrt(dotd<DIM>(e, e));
    c.nij = mk(e.x / ne, e.y / ne, DIM == 3 ? e.z / ne : 0.0);
    d3 pd = mk(c.pi.x - c.pj.x, c.pi.y - c.pj.y, DIM == 3 ? c.pi.z - c.pj.z : 0.0);
    c.dij = dotd<DIM>(pd, c.nij);
    return c;
}

// closest_points(sphere, plane): closest_points.hpp:768-787
template <int DIM>
__device__ inline CP cp_sphere_plane(d3 cs, double r, d3 pp, d3 nrm) {
#pragma clang fp contract(off)
    d3 d = mk(cs.x - pp.x, cs.y - pp.y, DIM == 3 ? cs.z - pp.z : 0.0);
    double p2s = dotd<DIM>(d, nrm);
    double sg = sgn(p2s);
    CP c;
    c.pi = mk(cs.x - sg * r * nrm.x, cs.y - sg * r * nrm.y, DIM == 3 ? cs.z - sg * r * nrm.z : 0.0);
    c.pj = mk(cs.x - p2s * nrm.x, cs.y - p2s * nrm.y, DIM == 3 ? cs.z - p2s * nrm.z : 0.0);
    c.nij = mk(sg * nrm.x, sg * nrm.y, DIM == 3 ? sg * nrm.z : 0.0);
    d3 pd = mk(c.pi.x - c.pj.x, c.pi.y - c.pj.y, DIM == 3 ? c.pi.z - c.pj.z : 0.0);
    c.dij = dotd<DIM>(pd, c.nij);
    return c;
}

// closest_points(plane, sphere) / (segment, sphere): closest_points.hpp:802-809, 874-881
__device__ inline CP swap_sides(CP c) {
    c.nij = mk(c.nij.x * -1.0, c.nij.y * -1.0, c.nij.z * -1.0);
    d3 t = c.pi;
    c.pi = c.pj;
    c.pj = t;
    return c;
}

// closest_points(sphere, segment), 2D: closest_points.hpp:824-859 with
// segment::normal/tangent/extrema (segment.hpp:154-177).
__device__ inline CP cp_sphere_segment2(d3 cs, double r, d3 sp, double qw, double qz, double length) {
#pragma clang fp contract(off)
    const double x = qw, w = qz;
    const double r00 = 1 - 2 * w * w, r01 = -2 * x * w, r10 = 2 * x * w, r11 = 1 - 2 * w * w;
    d3 nrm = mk(r00, r10, 0.0);
    d3 tangent = mk(-nrm.y, nrm.x, 0.0);
    d3 e0 = mk(r00 * 0 + r01 * (-0.5 * length) + sp.x, r10 * 0 + r11 * (-0.5 * length) + sp.y, 0.0);
    d3 e1 = mk(r00 * 0 + r01 * (0.5 * length) + sp.x, r10 * 0 + r11 * (0.5 * length) + sp.y, 0.0);
    d3 d0 = mk(cs.x - e0.x, cs.y - e0.y, 0.0);
    d3 d1 = mk(cs.x - e1.x, cs.y - e1.y, 0.0);
    CP c;
    auto endpoint = [&](d3 dd, d3 e) {
        double nd = sqrt(dotd<2>(dd, dd));
        c.nij = mk(dd.x / nd, dd.y / nd, 0.0);
        c.pi = mk(cs.x - c.nij.x * r, cs.y - c.nij.y * r, 0.0);
        c.pj = e;
        d3 pd = mk(c.pi.x - c.pj.x, c.pi.y - c.pj.y, 0.0);
        c.dij = dotd<2>(pd, c.nij);
    };
    if (dotd<2>(d0, tangent) < 0) {
        endpoint(d0, e0);
    } else if (dotd<2>(d1, tangent) > 0) {
        endpoint(d1, e1);
    } else {
        d3 d = mk(cs.x - sp.x, cs.y - sp.y, 0.0);
        double s2s = dotd<2>(d, nrm);
        double sg = sgn(s2s);
        c.pi = mk(cs.x - sg * r * nrm.x, cs.y - sg * r * nrm.y, 0.0);
        c.pj = mk(cs.x - s2s * nrm.x, cs.y - s2s * nrm.y, 0.0);
        c.nij = mk(sg * nrm.x, sg * nrm.y, 0.0);
        d3 pd = mk(c.pi.x - c.pj.x, c.pi.y - c.pj.y, 0.0);
        c.dij = dotd<2>(pd, c.nij);
    }
    return c;
}

// closest_points_dispatcher for the closed forms (closest_points.hpp:2037-2043).
// Returns false for pairs without a closed form here.
template <int DIM>
__device__ inline bool closest_points(const Particles& P, int64_t a, int64_t b, CP& out) {
    const int64_t cap = P.cap;
    const int sa = P.shape[a], sb = P.shape[b];
    d3 ca = load3(P.pos, cap, a), cb = load3(P.pos, cap, b);
    if (sa == SCOPI_SHAPE_SPHERE && sb == SCOPI_SHAPE_SPHERE) {
        out = cp_sphere_sphere<DIM>(ca, P.size[a], cb, P.size[b]);
        return true;
    }
    if (sa == SCOPI_SHAPE_SPHERE && sb == SCOPI_SHAPE_PLANE) {
        d3 nrm = first_column(DIM, P.q[b], P.q[cap + b], P.q[2 * cap + b], P.q[3 * cap + b]);
        out = cp_sphere_plane<DIM>(ca, P.size[a], cb, nrm);
        return true;
    }
    if (sa == SCOPI_SHAPE_PLANE && sb == SCOPI_SHAPE_SPHERE) {
        d3 nrm = first_column(DIM, P.q[a], P.q[cap + a], P.q[2 * cap + a], P.q[3 * cap + a]);
        out = swap_sides(cp_sphere_plane<DIM>(cb, P.size[b], ca, nrm));
        return true;
    }
    if (DIM == 2 && sa == SCOPI_SHAPE_SPHERE && sb == SCOPI_SHAPE_SEGMENT) {
        out = cp_sphere_segment2(ca, P.size[a], cb, P.q[b], P.q[3 * cap + b], P.size[b]);
        return true;
    }
    if (DIM == 2 && sa == SCOPI_SHAPE_SEGMENT && sb == SCOPI_SHAPE_SPHERE) {
        out = swap_sides(cp_sphere_segment2(cb, P.size[b], ca, P.q[a], P.q[3 * cap + a], P.size[a]));
        return true;
    }
    // plane / segment pairs: the reference's value-initialised neighbor (closest_points.hpp:750-753,
    // 898-917), dij = 0 and zero vectors; only an obstacle and a periodic copy of one meet this way
    if ((sa == SCOPI_SHAPE_PLANE || sa == SCOPI_SHAPE_SEGMENT) && (sb == SCOPI_SHAPE_PLANE || sb == SCOPI_SHAPE_SEGMENT)) {
        out.dij = 0.0;
        out.nij = out.pi = out.pj = mk(0.0, 0.0, 0.0);
        return true;
    }
    return false;
}

// nanoflann L2_Simple_Adaptor::evalMetric: sum of diff*diff in coordinate order.
template <int DIM>
__device__ inline double kd_dist2(d3 a, d3 b) {
#pragma clang fp contract(off)
    double r = 0.0;
    double dx = a.x - b.x;
    r += dx * dx;
    double dy = a.y - b.y;
    r += dy * dy;
    if (DIM == 3) {
        double dz = a.z - b.z;
        r += dz * dz;
    }
    return r;
}

// --------------------------------------------------------------------- small host helpers
#define SCOPI_HIP_CHECK(expr)                                         \
    do {                                                              \
        hipError_t _e = (expr);                                       \
        if (_e != hipSuccess) throw scopi_hip::hip_error(_e, #expr); \
    } while (0)

struct hip_error {
    hipError_t code;
    const char* what;
    hip_error(hipError_t c, const char* w) : code(c), what(w) {}
};

inline unsigned grid_for(int64_t n, int block = kBlock) {
    int64_t g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace scopi_hip
