// scopi_oracle.cpp -- FP64 CPU restatement of SCoPI's per-time-step contact
// pipeline (reference: hpc-maths/scopi v0.1.0).
//
// TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
// against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it.  It follows the reference's arithmetic order where that
// order can move a result by more than an ulp, and cites the reference file
// and line for every function.  Built with -ffp-contract=off so that no FMA
// contraction changes a predicate.
//
// What pins each function (see tests/test_oracle_*.py):
//   closest points (sphere/plane/segment)  -> test/test_closest_points.cpp known answers,
//                                             contact entries of test/references/*.json
//   contact set (kd-tree / brute force)    -> test/test_contacts_kdtree.cpp known answers,
//                                             brute-force enumeration with the same predicates
//   A / A^T                                -> test/test_matrices.cpp (adjointness, known values),
//                                             finite differences of the closed-form distance
//   APGD, NoFriction and Friction          -> KKT conditions, exhaustive active-set optimum,
//                                             analytic sphere/plane velocity (test_gradient.cpp:23-84)
//   full step                              -> test/references/{two_spheres_*, 2d_case_spheres,
//                                             obstacles_sphere_sphere_moving, 3spheres_nofriction,
//                                             sphere_plane_friction_mu01}.json, inclined-plane
//                                             expected errors (test_obstacles.cpp:256-313)
//   warm start (not in the reference)      -> first step = cold step bitwise, one iteration per
//                                             solve on a resting stack, off by default
#include "scopi_oracle.h"
#include "superellipsoid.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using vec3 = std::array<double, 3>;
using quat = std::array<double, 4>;
using mat3 = std::array<std::array<double, 3>, 3>;

struct Contact {
    int64_t i, j;
    double dij;
    vec3 nij, pi, pj;
    // contact_property<problem> (contact/property.hpp): mu, and for the Viscous problems the
    // adhesion potential gamma with its threshold gamma_min and tolerance gamma_tol
    double mu;
    double gamma, gamma_min, gamma_tol;
};

struct Box {  // BoxDomain (include/scopi/box.hpp)
    vec3 lo{0, 0, 0}, hi{0, 0, 0};
    std::array<bool, 3> periodic{false, false, false};
    bool any() const { return periodic[0] || periodic[1] || periodic[2]; }
};

struct Scene {
    int dim = 2;
    int64_t n = 0, n_inactive = 0;
    Box box;
    std::vector<double> proj_u, proj_w;  // vap_projection::set_u_w, Na x 3
    std::vector<int32_t> shape;
    std::vector<vec3> pos, size, v, vd, omega, domega, f, J;
    std::vector<std::array<double, 2>> sq;  // superellipsoid squareness (e, n)
    std::vector<quat> q;
    std::vector<double> m;
    oracle_params prm{};
    std::vector<Contact> contacts;
    std::vector<Contact> old_contacts;  // ScopiSolver::m_old_contacts: the last step's contacts and properties
    std::vector<double> lambda;
    bool lambda_is_history = false;  // lambda belongs to old_contacts and may start the next solve
    // optim slot outputs (OptimGradient::m_u / m_omega), Na x 3
    std::vector<double> u, w;
    oracle_stats st{};
    int64_t na() const { return n - n_inactive; }
};

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// xt::sign (include/scopi/objects/methods/closest_points.hpp:777), also src/utils.cpp:47-50
inline double sign(double x) { return (x > 0) ? 1.0 : ((x < 0) ? -1.0 : 0.0); }

// BLAS-style dot / nrm2 over the first n entries (closest_points.hpp:42-48 use xt::linalg).
inline double dot(const double* a, const double* b, int n) {
    double s = 0.0;
    for (int d = 0; d < n; ++d) s += a[d] * b[d];
    return s;
}
inline double nrm2(const double* a, int n) { return std::sqrt(dot(a, a, n)); }

// rotation_matrix_impl, 3D: include/scopi/quaternion.hpp:86-124 (q scalar-first)
mat3 rotation3(const quat& q) {
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    mat3 r;
    r[0] = {1 - 2 * z * z - 2 * w * w, 2 * y * z - 2 * x * w, 2 * y * w + 2 * x * z};
    r[1] = {2 * y * z + 2 * x * w, 1 - 2 * y * y - 2 * w * w, 2 * z * w - 2 * x * y};
    r[2] = {2 * y * w - 2 * x * z, 2 * z * w + 2 * x * y, 1 - 2 * y * y - 2 * z * z};
    return r;
}

// Column 0 of rotation_matrix<dim>(q): plane::normal (plane.hpp:142-147),
// segment::normal (segment.hpp:154-158).  2D form: quaternion.hpp:65-84.
vec3 first_column(int dim, const quat& q) {
    if (dim == 2) {
        const double x = q[0], w = q[3];
        return {1 - 2 * w * w, 2 * x * w, 0.0};
    }
    mat3 r = rotation3(q);
    return {r[0][0], r[1][0], r[2][0]};
}

// mult_quaternion: src/quaternion.cpp:92-120
quat mult_quaternion(const quat& q1, const quat& q2) {
    quat r;
    r[0] = q1[0] * q2[0];
    for (int d = 0; d < 3; ++d) {
        r[0] -= q1[d + 1] * q2[d + 1];
        r[d + 1] = q1[0] * q2[d + 1] + q2[0] * q1[d + 1];
    }
    r[1] += (q1[2] * q2[3] - q1[3] * q2[2]);
    r[2] += (q1[3] * q2[1] - q1[1] * q2[3]);
    r[3] += (q1[1] * q2[2] - q1[2] * q2[1]);
    return r;
}

// normalize: src/quaternion.cpp:6-10
void normalize(quat& q) {
    double n = nrm2(q.data(), 4);
    for (auto& c : q) c /= n;
}

// ------------------------------------------------------------------ closest points
// closest_points(sphere, sphere): closest_points.hpp:36-50
Contact cp_sphere_sphere(int dim, const vec3& ci, double ri, const vec3& cj, double rj) {
    vec3 d{}, u{};
    for (int k = 0; k < dim; ++k) d[k] = cj[k] - ci[k];
    double nd = nrm2(d.data(), dim);
    for (int k = 0; k < dim; ++k) u[k] = d[k] / nd;
    Contact c{};
    for (int k = 0; k < dim; ++k) {
        c.pi[k] = ci[k] + ri * u[k];
        c.pj[k] = cj[k] - rj * u[k];
    }
    vec3 e{};
    for (int k = 0; k < dim; ++k) e[k] = c.pj[k] - cj[k];
    double ne = nrm2(e.data(), dim);
    for (int k = 0; k < dim; ++k) c.nij[k] = e[k] / ne;
    vec3 pd{};
    for (int k = 0; k < dim; ++k) pd[k] = c.pi[k] - c.pj[k];
    c.dij = dot(pd.data(), c.nij.data(), dim);
    return c;
}

// closest_points(sphere, plane): closest_points.hpp:768-787
Contact cp_sphere_plane(int dim, const vec3& cs, double r, const vec3& pp, const quat& qp) {
    vec3 nrm = first_column(dim, qp);
    vec3 d{};
    for (int k = 0; k < dim; ++k) d[k] = cs[k] - pp[k];
    double p2s = dot(d.data(), nrm.data(), dim);
    double sg = sign(p2s);
    Contact c{};
    for (int k = 0; k < dim; ++k) {
        c.pi[k] = cs[k] - sg * r * nrm[k];
        c.pj[k] = cs[k] - p2s * nrm[k];
        c.nij[k] = sg * nrm[k];
    }
    vec3 pd{};
    for (int k = 0; k < dim; ++k) pd[k] = c.pi[k] - c.pj[k];
    c.dij = dot(pd.data(), c.nij.data(), dim);
    return c;
}

// closest_points(plane, sphere): closest_points.hpp:802-809
Contact swap_sides(int dim, Contact c) {
    for (int k = 0; k < dim; ++k) c.nij[k] *= -1.0;
    std::swap(c.pi, c.pj);
    return c;
}

// closest_points(sphere, segment): closest_points.hpp:824-859 (2D only; segment.hpp:154-177)
Contact cp_sphere_segment(int dim, const vec3& cs, double r, const vec3& sp, const quat& sq, double length) {
    vec3 nrm = first_column(dim, sq);
    vec3 tangent{-nrm[1], nrm[0], 0.0};
    // extrema: R * (0, -+L/2) + pos, with the 2D rotation of quaternion.hpp:65-84
    const double x = sq[0], w = sq[3];
    const double r00 = 1 - 2 * w * w, r01 = -2 * x * w, r10 = 2 * x * w, r11 = 1 - 2 * w * w;
    vec3 e0{r00 * 0 + r01 * (-0.5 * length) + sp[0], r10 * 0 + r11 * (-0.5 * length) + sp[1], 0.0};
    vec3 e1{r00 * 0 + r01 * (0.5 * length) + sp[0], r10 * 0 + r11 * (0.5 * length) + sp[1], 0.0};
    Contact c{};
    vec3 d0{cs[0] - e0[0], cs[1] - e0[1], 0.0};
    vec3 d1{cs[0] - e1[0], cs[1] - e1[1], 0.0};
    auto endpoint = [&](const vec3& dd, const vec3& e) {
        double nd = nrm2(dd.data(), 2);
        for (int k = 0; k < 2; ++k) c.nij[k] = dd[k] / nd;
        for (int k = 0; k < 2; ++k) c.pi[k] = cs[k] - c.nij[k] * r;
        c.pj = e;
        vec3 pd{c.pi[0] - c.pj[0], c.pi[1] - c.pj[1], 0.0};
        c.dij = dot(pd.data(), c.nij.data(), 2);
    };
    if (dot(d0.data(), tangent.data(), 2) < 0) {
        endpoint(d0, e0);
    } else if (dot(d1.data(), tangent.data(), 2) > 0) {
        endpoint(d1, e1);
    } else {
        vec3 d{cs[0] - sp[0], cs[1] - sp[1], 0.0};
        double s2s = dot(d.data(), nrm.data(), 2);
        double sg = sign(s2s);
        for (int k = 0; k < 2; ++k) {
            c.pi[k] = cs[k] - sg * r * nrm[k];
            c.pj[k] = cs[k] - s2s * nrm[k];
            c.nij[k] = sg * nrm[k];
        }
        vec3 pd{c.pi[0] - c.pj[0], c.pi[1] - c.pj[1], 0.0};
        c.dij = dot(pd.data(), c.nij.data(), 2);
    }
    (void)dim;
    return c;
}

// closest_points_dispatcher (closest_points.hpp:2037-2043, antisymmetric dispatch):
// returns false for pairs the reference has no closed form for.
// Elementary functions of the superellipsoid solve: 0 = C library, 1 = portable (se_math.hpp).
int g_se_math = 0;

se_oracle::SEResult se_solve(const se_oracle::SE& s1, const se_oracle::SE& s2) {
    return g_se_math ? se_oracle::se_se<se_oracle::PortableMath>(s1, s2) : se_oracle::se_se<se_oracle::LibMath>(s1, s2);
}

se_oracle::SE make_se(const vec3& c, const quat& q, const vec3& r, const std::array<double, 2>& sq) {
    se_oracle::SE s;
    s.c = c;
    s.r = r;
    s.e = sq[0];
    s.n = sq[1];
    s.R = rotation3(q);
    return s;
}

se_oracle::Sph make_sph(const vec3& c, const quat& q, const vec3& r) {
    se_oracle::Sph s;
    s.c = c;
    s.r = r[0];
    s.R = rotation3(q);
    return s;
}
se_oracle::Pln make_pln(const vec3& c, const quat& q) {
    se_oracle::Pln p;
    p.c = c;
    p.R = rotation3(q);
    return p;
}

// rotation_matrix<2> (quaternion.hpp:65-84) in the 2 x 2 block of a 3 x 3 matrix
mat3 rotation2(const quat& q) {
    const double x = q[0], w = q[3];
    mat3 r{};
    r[0] = {1 - 2 * w * w, -2 * x * w, 0.0};
    r[1] = {2 * x * w, 1 - 2 * w * w, 0.0};
    r[2] = {0.0, 0.0, 1.0};
    return r;
}

// the 2D objects: superellipsoid<2>(pos, q, radius, e) keeps e = squareness[0]
se_oracle::SE make_se2(const vec3& c, const quat& q, const vec3& r, const std::array<double, 2>& sq) {
    se_oracle::SE s = make_se(c, q, r, sq);
    s.R = rotation2(q);
    return s;
}
se_oracle::Sph make_sph2(const vec3& c, const quat& q, const vec3& r) {
    se_oracle::Sph s = make_sph(c, q, r);
    s.R = rotation2(q);
    return s;
}
se_oracle::Pln make_pln2(const vec3& c, const quat& q) {
    se_oracle::Pln p = make_pln(c, q);
    p.R = rotation2(q);
    return p;
}

// the 2D superellipsoid pairs (closest_points.hpp:64-249, 1316-1464, 1782-1968), in the
// reference's dispatch order: (sphere i, SE j) and (SE i, plane j) directly, (SE i, sphere j) and
// (plane i, SE j) through the other order with nij negated and pi / pj swapped
bool se2_pair(int si, const vec3& pi, const quat& qi, const vec3& szi, const std::array<double, 2>& sqi, int sj,
              const vec3& pj, const quat& qj, const vec3& szj, const std::array<double, 2>& sqj, Contact& out) {
    const bool port = g_se_math != 0;
    bool flip = false;
    se_oracle::NewtonResult r;
    if (si == ORACLE_SUPERELLIPSOID && sj == ORACLE_SUPERELLIPSOID) {
        const se_oracle::SE a = make_se2(pi, qi, szi, sqi), b = make_se2(pj, qj, szj, sqj);
        const se_oracle::SEResult e = port ? se_oracle::se2_se2<se_oracle::PortableMath>(a, b)
                                           : se_oracle::se2_se2<se_oracle::LibMath>(a, b);
        out.pi = e.pi;
        out.pj = e.pj;
        out.nij = e.nij;
        out.dij = e.dij;
        return true;
    }
    if (si == ORACLE_SPHERE && sj == ORACLE_SUPERELLIPSOID) {
        r = port ? se_oracle::sph2_se2<se_oracle::PortableMath>(make_sph2(pi, qi, szi), make_se2(pj, qj, szj, sqj))
                 : se_oracle::sph2_se2<se_oracle::LibMath>(make_sph2(pi, qi, szi), make_se2(pj, qj, szj, sqj));
    } else if (si == ORACLE_SUPERELLIPSOID && sj == ORACLE_SPHERE) {
        r = port ? se_oracle::sph2_se2<se_oracle::PortableMath>(make_sph2(pj, qj, szj), make_se2(pi, qi, szi, sqi))
                 : se_oracle::sph2_se2<se_oracle::LibMath>(make_sph2(pj, qj, szj), make_se2(pi, qi, szi, sqi));
        flip = true;
    } else if (si == ORACLE_SUPERELLIPSOID && sj == ORACLE_PLANE) {
        r = port ? se_oracle::se2_pln2<se_oracle::PortableMath>(make_se2(pi, qi, szi, sqi), make_pln2(pj, qj))
                 : se_oracle::se2_pln2<se_oracle::LibMath>(make_se2(pi, qi, szi, sqi), make_pln2(pj, qj));
    } else if (si == ORACLE_PLANE && sj == ORACLE_SUPERELLIPSOID) {
        r = port ? se_oracle::se2_pln2<se_oracle::PortableMath>(make_se2(pj, qj, szj, sqj), make_pln2(pi, qi))
                 : se_oracle::se2_pln2<se_oracle::LibMath>(make_se2(pj, qj, szj, sqj), make_pln2(pi, qi));
        flip = true;
    } else {
        return false;
    }
    if (flip) {
        for (int d = 0; d < 3; ++d) r.nij[d] *= -1.;
        std::swap(r.pi, r.pj);
    }
    out.pi = r.pi;
    out.pj = r.pj;
    out.nij = r.nij;
    out.dij = r.dij;
    return true;
}

// sphere / plane against a superellipsoid (3D), in the reference's dispatch order:
// (sphere i, SE j) and (SE i, plane j) directly; (SE i, sphere j) and (plane i, SE j) through
// the other order with nij negated and pi / pj swapped (closest_points.hpp:1295-1302, 1761-1768).
bool newton_pair(int si, const vec3& pi, const quat& qi, const vec3& szi, const std::array<double, 2>& sqi, int sj,
                 const vec3& pj, const quat& qj, const vec3& szj, const std::array<double, 2>& sqj,
                 se_oracle::NewtonResult& r) {
    const bool port = g_se_math != 0;
    bool flip = false;
    if (si == ORACLE_SPHERE && sj == ORACLE_SUPERELLIPSOID) {
        r = port ? se_oracle::sph_se<se_oracle::PortableMath>(make_sph(pi, qi, szi), make_se(pj, qj, szj, sqj))
                 : se_oracle::sph_se<se_oracle::LibMath>(make_sph(pi, qi, szi), make_se(pj, qj, szj, sqj));
    } else if (si == ORACLE_SUPERELLIPSOID && sj == ORACLE_SPHERE) {
        r = port ? se_oracle::sph_se<se_oracle::PortableMath>(make_sph(pj, qj, szj), make_se(pi, qi, szi, sqi))
                 : se_oracle::sph_se<se_oracle::LibMath>(make_sph(pj, qj, szj), make_se(pi, qi, szi, sqi));
        flip = true;
    } else if (si == ORACLE_SUPERELLIPSOID && sj == ORACLE_PLANE) {
        r = port ? se_oracle::se_pln<se_oracle::PortableMath>(make_se(pi, qi, szi, sqi), make_pln(pj, qj))
                 : se_oracle::se_pln<se_oracle::LibMath>(make_se(pi, qi, szi, sqi), make_pln(pj, qj));
    } else if (si == ORACLE_PLANE && sj == ORACLE_SUPERELLIPSOID) {
        r = port ? se_oracle::se_pln<se_oracle::PortableMath>(make_se(pj, qj, szj, sqj), make_pln(pi, qi))
                 : se_oracle::se_pln<se_oracle::LibMath>(make_se(pj, qj, szj, sqj), make_pln(pi, qi));
        flip = true;
    } else {
        return false;
    }
    if (flip) {
        for (int d = 0; d < 3; ++d) r.nij[d] *= -1.;
        std::swap(r.pi, r.pj);
    }
    return true;
}

bool closest_points(int dim, int si, const vec3& pi, const quat& qi, const vec3& szi, const std::array<double, 2>& sqi,
                    int sj, const vec3& pj, const quat& qj, const vec3& szj, const std::array<double, 2>& sqj,
                    Contact& out) {
    if (dim == 3 && si == ORACLE_SUPERELLIPSOID && sj == ORACLE_SUPERELLIPSOID) {
        const se_oracle::SEResult r = se_solve(make_se(pi, qi, szi, sqi), make_se(pj, qj, szj, sqj));
        out.pi = r.pi;
        out.pj = r.pj;
        out.nij = r.nij;
        out.dij = r.dij;
        return true;
    }
    if (dim == 3 && (si == ORACLE_SUPERELLIPSOID || sj == ORACLE_SUPERELLIPSOID)) {
        se_oracle::NewtonResult r;
        if (!newton_pair(si, pi, qi, szi, sqi, sj, pj, qj, szj, sqj, r)) return false;
        out.pi = r.pi;
        out.pj = r.pj;
        out.nij = r.nij;
        out.dij = r.dij;
        return true;
    }
    if (dim == 2 && (si == ORACLE_SUPERELLIPSOID || sj == ORACLE_SUPERELLIPSOID))
        return se2_pair(si, pi, qi, szi, sqi, sj, pj, qj, szj, sqj, out);
    if (si == ORACLE_SPHERE && sj == ORACLE_SPHERE) {
        out = cp_sphere_sphere(dim, pi, szi[0], pj, szj[0]);
        return true;
    }
    if (si == ORACLE_SPHERE && sj == ORACLE_PLANE) {
        out = cp_sphere_plane(dim, pi, szi[0], pj, qj);
        return true;
    }
    if (si == ORACLE_PLANE && sj == ORACLE_SPHERE) {
        out = swap_sides(dim, cp_sphere_plane(dim, pj, szj[0], pi, qi));
        return true;
    }
    if (dim == 2 && si == ORACLE_SPHERE && sj == ORACLE_SEGMENT) {
        out = cp_sphere_segment(dim, pi, szi[0], pj, qj, szj[0]);
        return true;
    }
    if (dim == 2 && si == ORACLE_SEGMENT && sj == ORACLE_SPHERE) {
        out = swap_sides(dim, cp_sphere_segment(dim, pj, szj[0], pi, qi, szi[0]));
        return true;
    }
    // plane/plane, segment/segment, plane/segment: the reference returns a value-initialised
    // neighbor (closest_points.hpp:750-753, 898-917): dij = 0, zero vectors.  Such pairs only
    // arise between an obstacle and a periodic copy of an obstacle.
    if ((si == ORACLE_PLANE || si == ORACLE_SEGMENT) && (sj == ORACLE_PLANE || sj == ORACLE_SEGMENT)) {
        out = Contact{};
        return true;
    }
    return false;
}

// compute_exact_distance: contact/base.hpp:103-141 (non-periodic box)
void exact_distance(const Scene& s, int64_t i, int64_t j, std::vector<Contact>& out) {
    Contact c{};
    if (!closest_points(s.dim, s.shape[i], s.pos[i], s.q[i], s.size[i], s.sq[i], s.shape[j], s.pos[j], s.q[j],
                        s.size[j], s.sq[j], c))
        return;
    if (c.dij < s.prm.dmax) {
        c.i = i;
        c.j = j;
        c.mu = s.prm.mu;
        c.gamma = s.prm.gamma;
        c.gamma_min = s.prm.gamma_min;
        c.gamma_tol = s.prm.gamma_tol;
        out.push_back(c);
    }
}

// sort_contacts: contact/base.hpp:156-175
void sort_contacts(std::vector<Contact>& c) {
    std::sort(c.begin(), c.end(), [](const Contact& a, const Contact& b) {
        if (a.i < b.i) return true;
        if (a.i == b.i) return a.j < b.j;
        return false;
    });
}

// nanoflann L2_Simple_Adaptor::evalMetric: sum of diff*diff in coordinate order.
inline double kd_dist2(const Scene& s, int64_t a, int64_t b) {
    double r = 0.0;
    for (int d = 0; d < s.dim; ++d) {
        double diff = s.pos[a][d] - s.pos[b][d];
        r += diff * diff;
    }
    return r;
}

// contact_kdtree::run_impl: contact/contact_kdtree.hpp:200-268.  The set of
// (i, j) candidates is every active pair with i < j and squared centre
// distance < kd_tree_radius (nanoflann radiusSearch, strict <).  The oracle
// enumerates that exact set through a uniform grid whose cell is at least the
// search radius, then applies the same predicate to every candidate.
std::vector<Contact> contacts_kdtree(const Scene& s) {
    std::vector<Contact> out;
    const int64_t a0 = s.n_inactive, na = s.na();
    if (na > 1) {
        const double h = std::sqrt(s.prm.kd_tree_radius) * (1.0 + 1e-12) + 1e-300;
        vec3 lo{1e300, 1e300, 1e300}, hi{-1e300, -1e300, -1e300};
        for (int64_t p = a0; p < s.n; ++p)
            for (int d = 0; d < s.dim; ++d) {
                lo[d] = std::min(lo[d], s.pos[p][d]);
                hi[d] = std::max(hi[d], s.pos[p][d]);
            }
        std::array<int64_t, 3> nc{1, 1, 1};
        for (int d = 0; d < s.dim; ++d) {
            double cells = std::floor((hi[d] - lo[d]) / h) + 1.0;
            nc[d] = (int64_t)std::min(cells, 2048.0);
        }
        std::array<double, 3> cellw{h, h, h};
        for (int d = 0; d < s.dim; ++d)
            cellw[d] = std::max(h, (hi[d] - lo[d]) / (double)nc[d] * (1.0 + 1e-12));
        auto cell_of = [&](int64_t p, int d) {
            int64_t c = (int64_t)std::floor((s.pos[p][d] - lo[d]) / cellw[d]);
            return std::min(std::max<int64_t>(c, 0), nc[d] - 1);
        };
        std::vector<std::pair<int64_t, int64_t>> key(na);
        for (int64_t p = 0; p < na; ++p) {
            int64_t k = 0;
            for (int d = s.dim - 1; d >= 0; --d) k = k * nc[d] + cell_of(a0 + p, d);
            key[p] = {k, a0 + p};
        }
        std::sort(key.begin(), key.end());
        std::vector<int64_t> keys(na);
        for (int64_t p = 0; p < na; ++p) keys[p] = key[p].first;
        // the reference's parallel loop over the queries (contact_kdtree.hpp:220): static blocks of i,
        // each thread's contacts kept apart and joined in thread order, i.e. in i order
        int nthreads = 1;
#ifdef _OPENMP
        nthreads = omp_get_max_threads();
#endif
        std::vector<std::vector<Contact>> part((size_t)nthreads);
#pragma omp parallel
        {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            std::vector<Contact>& mine = part[(size_t)tid];
            std::vector<int64_t> cand;
#pragma omp for schedule(static)
        for (int64_t i = a0; i < s.n - 1; ++i) {
            std::array<int64_t, 3> c{0, 0, 0};
            for (int d = 0; d < s.dim; ++d) c[d] = cell_of(i, d);
            cand.clear();
            const int zr = (s.dim == 3) ? 1 : 0;
            for (int dz = -zr; dz <= zr; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int64_t cx = c[0] + dx, cy = c[1] + dy, cz = c[2] + dz;
                        if (cx < 0 || cy < 0 || cz < 0 || cx >= nc[0] || cy >= nc[1] || cz >= nc[2]) continue;
                        int64_t k = (cz * nc[1] + cy) * nc[0] + cx;
                        if (s.dim == 2) k = cy * nc[0] + cx;
                        auto lo_it = std::lower_bound(keys.begin(), keys.end(), k);
                        for (auto it = lo_it; it != keys.end() && *it == k; ++it) {
                            int64_t j = key[it - keys.begin()].second;
                            if (i < j && kd_dist2(s, i, j) < s.prm.kd_tree_radius) cand.push_back(j);
                        }
                    }
            for (int64_t j : cand) exact_distance(s, i, j, mine);
        }
        }
        for (auto& pt : part) out.insert(out.end(), pt.begin(), pt.end());
    }
    // obstacles: contact_kdtree.hpp:249-256
    for (int64_t i = 0; i < s.n_inactive; ++i)
        for (int64_t j = s.n_inactive; j < s.n; ++j) exact_distance(s, i, j, out);
    sort_contacts(out);
    return out;
}

// contact_brute_force::run_impl: contact/contact_brute_force.hpp:102-137
std::vector<Contact> contacts_brute_force(const Scene& s) {
    std::vector<Contact> out;
    for (int64_t i = s.n_inactive; i < s.n - 1; ++i)
        for (int64_t j = i + 1; j < s.n; ++j) exact_distance(s, i, j, out);
    for (int64_t i = 0; i < s.n_inactive; ++i)
        for (int64_t j = s.n_inactive; j < s.n; ++j) exact_distance(s, i, j, out);
    sort_contacts(out);
    return out;
}

// ------------------------------------------------------------------ matrices
// detail::cross<dim>: matrix/velocities.hpp:26-44
inline vec3 cross(int dim, const vec3& x, const vec3& y) {
    vec3 o;
    if (dim == 2) {
        o[0] = x[1] * y[2];
        o[1] = -x[0] * y[2];
        o[2] = x[0] * y[1] - x[1] * y[0];
    } else {
        o[0] = x[1] * y[2] - x[2] * y[1];
        o[1] = x[2] * y[0] - x[0] * y[2];
        o[2] = x[0] * y[1] - x[1] * y[0];
    }
    return o;
}
// detail::mat_mult: matrix/velocities.hpp:16-24
inline vec3 mat_mult(const mat3& a, const vec3& x) {
    return {a[0][0] * x[0] + a[0][1] * x[1] + a[0][2] * x[2], a[1][0] * x[0] + a[1][1] * x[1] + a[1][2] * x[2],
            a[2][0] * x[0] + a[2][1] * x[1] + a[2][2] * x[2]};
}
inline vec3 mat_mult_t(const mat3& a, const vec3& x) {
    return {a[0][0] * x[0] + a[1][0] * x[1] + a[2][0] * x[2], a[0][1] * x[0] + a[1][1] * x[1] + a[2][1] * x[2],
            a[0][2] * x[0] + a[1][2] * x[1] + a[2][2] * x[2]};
}

// rij = c.pi - pos(c.i); in 2D the third slot is never read by cross<2>.
inline vec3 arm(const Scene& s, const vec3& p, int64_t k) {
    vec3 r{0, 0, 0};
    for (int d = 0; d < s.dim; ++d) r[d] = p[d] - s.pos[k][d];
    return r;
}

// AMatrix::mat_mult: matrix/velocities.hpp:67-144 (u: 6*Na, out: 3*nc)
void amat(const Scene& s, const std::vector<Contact>& cs, const double* u, double* out) {
    const int64_t off = s.n_inactive, rot = 3 * s.na();
    for (size_t c = 0; c < cs.size(); ++c) {
        const Contact& ct = cs[c];
        double* o = out + 3 * c;
        o[0] = o[1] = o[2] = 0.0;
        if (ct.i >= off) {
            int64_t st = (ct.i - off) * 3;
            vec3 v{u[st], u[st + 1], u[st + 2]}, om{u[rot + st], u[rot + st + 1], u[rot + st + 2]};
            vec3 cr = cross(s.dim, arm(s, ct.pi, ct.i), mat_mult(rotation3(s.q[ct.i]), om));
            for (int d = 0; d < 3; ++d) o[d] += v[d] - cr[d];
        }
        if (ct.j >= off) {
            int64_t st = (ct.j - off) * 3;
            vec3 v{u[st], u[st + 1], u[st + 2]}, om{u[rot + st], u[rot + st + 1], u[rot + st + 2]};
            vec3 cr = cross(s.dim, arm(s, ct.pj, ct.j), mat_mult(rotation3(s.q[ct.j]), om));
            for (int d = 0; d < 3; ++d) o[d] -= v[d] - cr[d];
        }
    }
}

// ATMatrix::mat_mult: matrix/velocities.hpp:167-254 (f: 3*nc, out: 6*Na)
void atmat(const Scene& s, const std::vector<Contact>& cs, const double* f, double* out) {
    const int64_t off = s.n_inactive, rot = 3 * s.na();
    std::fill(out, out + 6 * s.na(), 0.0);
    for (size_t c = 0; c < cs.size(); ++c) {
        const Contact& ct = cs[c];
        vec3 fv{f[3 * c], f[3 * c + 1], f[3 * c + 2]};
        if (ct.i >= off) {
            int64_t st = (ct.i - off) * 3;
            vec3 res = mat_mult_t(rotation3(s.q[ct.i]), cross(s.dim, arm(s, ct.pi, ct.i), fv));
            for (int d = 0; d < 3; ++d) {
                out[st + d] += fv[d];
                out[rot + st + d] += res[d];
            }
        }
        if (ct.j >= off) {
            int64_t st = (ct.j - off) * 3;
            vec3 res = mat_mult_t(rotation3(s.q[ct.j]), cross(s.dim, arm(s, ct.pj, ct.j), fv));
            for (int d = 0; d < 3; ++d) {
                out[st + d] -= fv[d];
                out[rot + st + d] -= res[d];
            }
        }
    }
}

// ------------------------------------------------------------------ minimization problem
// minimization_problem<problem, ...>: solvers/minimization_problem.hpp:10-162 with the
// LagrangeMultiplier of each problem (lagrange_multiplier.hpp):
//   NoFriction (:44-114): one row per contact, lambda n;
//   Friction / FrictionFixedPoint (:208-273, :275-345): three rows per contact, the Coulomb cone;
//   Viscous (:116-206): one row per contact, and for a contact with gamma < -gamma_tol a second
//     row appended after all first rows (in contact order): lambda = (x+ - x-) n, both >= 0;
//   ViscousFriction (:347-553): rows in contact order, 1 (gamma >= -gamma_tol), 2 (x+, x-,
//     gamma > gamma_min) or 4 (gamma == gamma_min: x-, then a cone multiplier f of dim rows,
//     lambda = -x- n + f, projected on the nearer of the cone and the half-line x- >= 0).
enum { kModeNormal = 0, kModeCone = 1, kModeViscous = 2, kModeViscousFriction = 3 };

int problem_mode(int problem) {
    switch (problem) {
        case ORACLE_FRICTION:
        case ORACLE_FRICTION_FIXED_POINT: return kModeCone;
        case ORACLE_VISCOUS: return kModeViscous;
        case ORACLE_VISCOUS_FRICTION: return kModeViscousFriction;
        default: return kModeNormal;
    }
}

struct MinProblem {
    const Scene& s;
    const std::vector<Contact>& cs;
    double dt;
    int mode;
    std::vector<double> invM, C;
    mutable std::vector<double> g3, w6, a3;
    mutable int64_t q_applies = 0;
    // Viscous layouts: the kind of every contact (0: one row, 1: two rows, 2: four rows), its first
    // row and, for Viscous, the row of its second multiplier; nloc rows in all
    std::vector<int> kind;
    std::vector<int64_t> row, row2;
    int64_t nloc = 0;

    // S_Vector: s_ij n_ij per contact for FrictionFixedPoint (lagrange_multiplier.hpp:290-297),
    // dt mu s_ij n_ij on the cone rows of a four-row ViscousFriction contact (:379-400)
    std::vector<double> S;

    MinProblem(const Scene& sc, const std::vector<Contact>& c, double dt_, const std::vector<double>* sij = nullptr)
        : s(sc), cs(c), dt(dt_) {
        mode = problem_mode(s.prm.problem);
        const int64_t na = s.na(), off = s.n_inactive, nc = (int64_t)cs.size();
        kind.assign(nc, 0);
        row.assign(nc, 0);
        row2.assign(nc, -1);
        if (mode == kModeNormal) nloc = nc;
        if (mode == kModeCone) nloc = 3 * nc;
        if (mode == kModeViscous) {
            nloc = nc;
            for (int64_t k = 0; k < nc; ++k) {
                row[k] = k;
                if (cs[k].gamma < -cs[k].gamma_tol) {
                    kind[k] = 1;
                    row2[k] = nloc++;
                }
            }
        }
        if (mode == kModeViscousFriction) {
            for (int64_t k = 0; k < nc; ++k) {
                row[k] = nloc;
                if (cs[k].gamma < -cs[k].gamma_tol) {
                    kind[k] = (cs[k].gamma != cs[k].gamma_min) ? 1 : 2;
                    nloc += (kind[k] == 1) ? 2 : 4;
                } else {
                    nloc += 1;
                }
            }
        }
        S.assign(nloc, 0.0);
        if (sij && mode == kModeCone)
            for (int64_t k = 0; k < nc; ++k)
                for (int d = 0; d < s.dim; ++d) S[3 * k + d] = (*sij)[k] * cs[k].nij[d];
        if (sij && mode == kModeViscousFriction)
            for (int64_t k = 0; k < nc; ++k)
                if (kind[k] == 2)
                    for (int d = 0; d < s.dim; ++d) S[row[k] + 1 + d] = dt * cs[k].mu * (*sij)[k] * cs[k].nij[d];
        // M_inverse: minimization_problem.hpp:10-31
        invM.assign(6 * na, 0.0);
        for (int64_t i = 0; i < na; ++i) {
            for (int d = 0; d < s.dim; ++d) invM[3 * i + d] = 1. / s.m[off + i];
            if (s.dim == 2)
                invM[3 * na + 3 * i + 2] = 1. / s.J[off + i][2];
            else
                for (int d = 0; d < 3; ++d) invM[3 * na + 3 * i + d] = 1. / s.J[off + i][d];
        }
        // CVector: minimization_problem.hpp:33-70 (D: velocities.hpp:274-289)
        std::vector<double> U(6 * na, 0.0);
        for (int64_t i = 0; i < na; ++i) {
            for (int d = 0; d < s.dim; ++d) U[3 * i + d] = s.v[off + i][d];
            if (s.dim == 2)
                U[3 * na + 3 * i + 2] = s.omega[off + i][2];
            else
                for (int d = 0; d < 3; ++d) U[3 * na + 3 * i + d] = s.omega[off + i][d];
        }
        std::vector<double> AU(3 * nc);
        amat(s, cs, U.data(), AU.data());
        C.assign(3 * nc, 0.0);
        for (int64_t k = 0; k < nc; ++k)
            for (int d = 0; d < 3; ++d) {
                double dn = (d < s.dim) ? cs[k].dij * cs[k].nij[d] : 0.0;
                C[3 * k + d] = dn + dt * AU[3 * k + d];
            }
        g3.resize(3 * nc);
        w6.resize(6 * na);
        a3.resize(3 * nc);
    }
    int64_t size() const { return nloc; }

    // local2global (NoFriction :78-92, Viscous :165-184, ViscousFriction :436-464); the unset
    // 2D slot is 0.
    void l2g(const double* x, double* out) const {
        const int64_t nc = cs.size();
        if (mode == kModeCone) {
            std::copy(x, x + 3 * nc, out);
            return;
        }
        for (int64_t k = 0; k < nc; ++k) {
            const double* n = cs[k].nij.data();
            out[3 * k] = out[3 * k + 1] = out[3 * k + 2] = 0.0;
            if (mode == kModeNormal) {
                for (int d = 0; d < s.dim; ++d) out[3 * k + d] = x[k] * n[d];
            } else if (mode == kModeViscous) {
                const double a = kind[k] ? x[k] - x[row2[k]] : x[k];
                for (int d = 0; d < s.dim; ++d) out[3 * k + d] = a * n[d];
            } else {
                const int64_t r = row[k];
                if (kind[k] == 0)
                    for (int d = 0; d < s.dim; ++d) out[3 * k + d] = x[r] * n[d];
                else if (kind[k] == 1)
                    for (int d = 0; d < s.dim; ++d) out[3 * k + d] = (x[r] - x[r + 1]) * n[d];
                else
                    for (int d = 0; d < s.dim; ++d) out[3 * k + d] = (-x[r]) * n[d] + x[r + 1 + d];
            }
        }
    }
    // global2local (NoFriction :62-76, Viscous :143-163, ViscousFriction :403-434)
    void g2l(const double* x, double* out) const {
        const int64_t nc = cs.size();
        if (mode == kModeCone) {
            std::copy(x, x + 3 * nc, out);
            return;
        }
        if (mode == kModeViscousFriction) std::fill(out, out + nloc, 0.0);
        for (int64_t k = 0; k < nc; ++k) {
            const double dn = dot(x + 3 * k, cs[k].nij.data(), s.dim);
            if (mode == kModeNormal) {
                out[k] = dn;
            } else if (mode == kModeViscous) {
                out[k] = dn;
                if (kind[k]) out[row2[k]] = -dn;
            } else {
                const int64_t r = row[k];
                if (kind[k] == 0) {
                    out[r] = dn;
                } else if (kind[k] == 1) {
                    out[r] = dn;
                    out[r + 1] = -dn;
                } else {
                    out[r] = -dn;
                    for (int d = 0; d < s.dim; ++d) out[r + 1 + d] = x[3 * k + d];
                }
            }
        }
    }
    // QMatrix::operator(): minimization_problem.hpp:85-92  Q x = dt^2 A(M^-1 A^T x)
    void Q(const double* xg, double* out) const {
        ++q_applies;
        atmat(s, cs, xg, w6.data());
        for (size_t k = 0; k < w6.size(); ++k) w6[k] = invM[k] * w6[k];
        amat(s, cs, w6.data(), out);
        for (size_t k = 0; k < 3 * cs.size(); ++k) out[k] = dt * dt * out[k];
    }
    // gradient: minimization_problem.hpp:122-128, g2l(Q l2g(lambda) + C) + S
    void gradient(const double* lam, double* out) const {
        l2g(lam, g3.data());
        Q(g3.data(), a3.data());
        for (size_t k = 0; k < a3.size(); ++k) a3[k] = a3[k] + C[k];
        g2l(a3.data(), out);
        for (int64_t k = 0; k < size(); ++k) out[k] = out[k] + S[k];
    }
    // operator(): minimization_problem.hpp:130-134, <l2g(lambda), Q l2g(lambda) / 2 + C> + <lambda, S>
    double value(const double* lam) const {
        l2g(lam, g3.data());
        Q(g3.data(), a3.data());
        double r = 0.0;
        for (size_t k = 0; k < a3.size(); ++k) r += g3[k] * (0.5 * a3[k] + C[k]);
        double rs = 0.0;
        for (int64_t k = 0; k < size(); ++k) rs += lam[k] * S[k];
        return r + rs;
    }
    // velocities: minimization_problem.hpp:136-139 -> M^-1 A^T l2g(lambda)
    void velocities(const double* lam, double* out) const {
        l2g(lam, g3.data());
        atmat(s, cs, g3.data(), out);
        for (int64_t k = 0; k < 6 * s.na(); ++k) out[k] = invM[k] * out[k];
    }
    // the Coulomb-cone projection of a dim-vector (lagrange_multiplier.hpp:243-268, :500-519)
    void cone(double* li, const double* n, double mu) const {
        double ln = dot(li, n, s.dim);
        vec3 lt{0, 0, 0};
        for (int d = 0; d < s.dim; ++d) lt[d] = li[d] - ln * n[d];
        double norm = nrm2(lt.data(), s.dim);
        if (norm > mu * ln) {
            if (ln <= -mu * norm) {
                for (int d = 0; d < s.dim; ++d) li[d] = 0;
            } else {
                double new_norm = (mu * mu) * (norm + ln / mu) / (mu * mu + 1);
                double new_ln = new_norm / mu;
                for (int d = 0; d < s.dim; ++d) lt[d] /= norm;
                for (int d = 0; d < s.dim; ++d) li[d] = new_norm * lt[d] + new_ln * n[d];
            }
        }
    }
    // projection: NoFriction / Viscous max(lambda, 0) (:104-107, :197-201); Friction :243-268;
    // ViscousFriction :487-547
    void projection(double* lam) const {
        const int64_t nc = cs.size();
        if (mode == kModeNormal || mode == kModeViscous) {
            for (int64_t k = 0; k < nloc; ++k) lam[k] = std::max(lam[k], 0.);
            return;
        }
        if (mode == kModeCone) {
            for (int64_t k = 0; k < nc; ++k) cone(lam + 3 * k, cs[k].nij.data(), cs[k].mu);
            return;
        }
        for (int64_t k = 0; k < nc; ++k) {
            double* li = lam + row[k];
            if (kind[k] == 0) {
                li[0] = std::max(li[0], 0.);
            } else if (kind[k] == 1) {
                li[0] = std::max(li[0], 0.);
                li[1] = std::max(li[1], 0.);
            } else {
                // the nearer (Euclidean, over the 1 + dim rows) of (0, Proj_cone(f)) and (max(x-, 0), 0)
                double fric[4] = {0.0, li[1], li[2], li[3]};
                cone(fric + 1, cs[k].nij.data(), cs[k].mu);
                double moins[4] = {std::max(li[0], 0.), 0.0, 0.0, 0.0};
                double df[4], dm[4];
                for (int d = 0; d < 1 + s.dim; ++d) {
                    df[d] = fric[d] - li[d];
                    dm[d] = moins[d] - li[d];
                }
                if (nrm2(df, 1 + s.dim) < nrm2(dm, 1 + s.dim))
                    for (int d = 0; d < 1 + s.dim; ++d) li[d] = fric[d];
                else
                    for (int d = 0; d < 1 + s.dim; ++d) li[d] = moins[d];
            }
        }
    }
};

struct SolveResult {
    std::vector<double> lambda;
    int64_t iterations = 0, backtracks = 0, converged = 0;
};

double norm_diff(const std::vector<double>& a, const std::vector<double>& b) {
    double r = 0.0;
    for (size_t k = 0; k < a.size(); ++k) {
        double d = a[k] - b[k];
        r += d * d;
    }
    return std::sqrt(r);
}

// apgd::operator(): solvers/apgd.hpp:147-222 (theta is uniform, kept scalar).  lam0 (not in the
// reference, which starts from 0 at apgd.hpp:152-159) replaces the first iterate lambda_0 = y_0.
SolveResult apgd(const MinProblem& mp, const oracle_params& prm, const std::vector<double>* lam0 = nullptr) {
    const int64_t n = mp.size();
    SolveResult res;
    std::vector<double> lam_n(n, 0.0), lam_np1(n, 0.0), y_n(n, 0.0), y_np1(n, 0.0), dG(n), tmp(n);
    if (lam0) {
        lam_n = *lam0;
        y_n = *lam0;
    }
    double theta_n = 1.0, theta_np1 = 1.0;
    double alpha = prm.alpha;
    double lipsch = 1. / alpha;
    int64_t ite = 0;
    bool conv = false;
    while (ite < prm.max_ite) {
        ++ite;
        mp.gradient(y_n.data(), dG.data());
        for (int64_t k = 0; k < n; ++k) lam_np1[k] = y_n[k] - alpha * dG[k];
        mp.projection(lam_np1.data());
        if (prm.dynamic_descent) {
            while (true) {
                double f_new = mp.value(lam_np1.data());
                if (prm.objective_perturbation != 0.0) f_new *= 1.0 + prm.objective_perturbation;
                double f_y = mp.value(y_n.data());
                double dt_ = 0.0, nr = 0.0;
                for (int64_t k = 0; k < n; ++k) {
                    double d = lam_np1[k] - y_n[k];
                    dt_ += dG[k] * d;
                    nr += d * d;
                }
                nr = std::sqrt(nr);
                if (!(f_new >= f_y + dt_ + 0.5 * lipsch * std::pow(nr, 2))) break;
                lipsch *= 2;
                alpha = 1. / lipsch;
                for (int64_t k = 0; k < n; ++k) lam_np1[k] = y_n[k] - alpha * dG[k];
                mp.projection(lam_np1.data());
                ++res.backtracks;
            }
        }
        if (norm_diff(lam_np1, lam_n) < prm.tolerance) {
            std::swap(lam_n, lam_np1);
            conv = true;
            break;
        }
        theta_np1 = 0.5 * (theta_n * std::sqrt(4 + theta_n * theta_n) - theta_n * theta_n);
        double beta = theta_n * (1 - theta_n) / (theta_n * theta_n + theta_np1);
        for (int64_t k = 0; k < n; ++k) y_np1[k] = lam_np1[k] + beta * (lam_np1[k] - lam_n[k]);
        if (prm.dynamic_descent) {
            double d = 0.0;
            for (int64_t k = 0; k < n; ++k) d += dG[k] * (lam_np1[k] - lam_n[k]);
            if (d > 0) {
                y_np1 = lam_np1;
                theta_np1 = 1.;
            }
            lipsch *= 0.97;
            alpha = 1. / lipsch;
        }
        std::swap(lam_n, lam_np1);
        std::swap(theta_n, theta_np1);
        std::swap(y_n, y_np1);
    }
    res.iterations = ite;
    res.converged = conv ? 1 : 0;
    res.lambda = lam_n;
    return res;
}

// pgd::operator(): solvers/apgd.hpp:60-98
SolveResult pgd(const MinProblem& mp, const oracle_params& prm) {
    const int64_t n = mp.size();
    SolveResult res;
    std::vector<double> lam_n(n, 0.0), lam_np1(n, 0.0), dG(n);
    int64_t ite = 0;
    bool conv = false;
    while (ite < prm.max_ite) {
        ++ite;
        mp.gradient(lam_n.data(), dG.data());
        for (int64_t k = 0; k < n; ++k) lam_np1[k] = lam_n[k] - prm.alpha * dG[k];
        mp.projection(lam_np1.data());
        if (norm_diff(lam_np1, lam_n) < prm.tolerance) {
            std::swap(lam_n, lam_np1);
            conv = true;
            break;
        }
        std::swap(lam_n, lam_np1);
    }
    res.iterations = ite;
    res.converged = conv ? 1 : 0;
    res.lambda = lam_n;
    return res;
}

// OptimGradient::run: solvers/OptimGradient.hpp:252-306
void optim_solve(Scene& s, double dt, const std::vector<double>* sij);

// A U for the optim slot's velocities (the relative velocity at every contact), 3 rows per contact.
std::vector<double> contact_velocities(const Scene& s) {
    const int64_t na = s.na(), nc = (int64_t)s.contacts.size();
    std::vector<double> U(6 * na, 0.0), AU(3 * nc);
    for (int64_t i = 0; i < na; ++i) {
        for (int d = 0; d < s.dim; ++d) U[3 * i + d] = s.u[3 * i + d];
        if (s.dim == 2)
            U[3 * na + 3 * i + 2] = s.w[3 * i + 2];
        else
            for (int d = 0; d < 3; ++d) U[3 * na + 3 * i + d] = s.w[3 * i + d];
    }
    amat(s, s.contacts, U.data(), AU.data());
    return AU;
}

// |T (A U)_k|: the norm of the tangential part of contact k's relative velocity
double tangential_norm(const Scene& s, const std::vector<double>& AU, int64_t k) {
    const double* a = AU.data() + 3 * k;
    const double* n = s.contacts[k].nij.data();
    const double an = dot(a, n, s.dim);
    vec3 t{0, 0, 0};
    for (int d = 0; d < s.dim; ++d) t[d] = a[d] - n[d] * an;
    return nrm2(t.data(), s.dim);
}

// OptimGradient::run inside ScopiSolver::run's should_solve loop (solver.hpp:287-292): one solve,
// or an outer fixed point on s_ij (OptimGradient.hpp:86-240), controlled by the first contact's
// fixed_point_tol / fixed_point_max_iter:
//   FrictionFixedPoint: s_ij = mu dt, solve, s_ij = mu dt |T(A U)_ij| (tangential part of the
//     contact's relative velocity under the solve's velocities), until
//     (1/(mu dt)) |s_new - s_old| / (1 + (1/(mu dt)) |s_new|) < tol or max_iter solves;
//   ViscousFriction: s_ij = 1 on the contacts at gamma == gamma_min, solve, s_ij = |T(A U)_ij| on
//     those contacts, until |s_new - s_old| / (1 + |s_new|) < tol (both vectors 0 on the other
//     contacts) or max_iter solves.
void optim_run(Scene& s, double dt) {
    s.st.fixed_point_iterations = 0;
    const bool ffp = s.prm.problem == ORACLE_FRICTION_FIXED_POINT;
    const bool vf = s.prm.problem == ORACLE_VISCOUS_FRICTION;
    if (!(ffp || vf) || s.contacts.empty()) {
        optim_solve(s, dt, nullptr);
        return;
    }
    const Contact& c0 = s.contacts[0];
    const double mudt = c0.mu * dt;
    const int64_t nc = (int64_t)s.contacts.size();
    std::vector<double> sij(nc, 0.0);
    for (int64_t k = 0; k < nc; ++k) {
        if (ffp) sij[k] = s.contacts[k].mu * dt;
        if (vf && s.contacts[k].gamma == s.contacts[k].gamma_min) sij[k] = 1.;
    }
    int64_t it = 0, total = 0;
    for (;;) {
        optim_solve(s, dt, &sij);
        total += s.st.iterations;
        const std::vector<double> AU = contact_velocities(s);
        double ds2 = 0.0, s2 = 0.0;
        for (int64_t k = 0; k < nc; ++k) {
            if (vf && s.contacts[k].gamma != s.contacts[k].gamma_min) continue;
            const double old = sij[k];
            sij[k] = ffp ? s.contacts[k].mu * dt * tangential_norm(s, AU, k) : tangential_norm(s, AU, k);
            ds2 += (sij[k] - old) * (sij[k] - old);
            s2 += sij[k] * sij[k];
        }
        const double err = ffp ? (1 / mudt) * std::sqrt(ds2) / (1 + (1 / mudt) * std::sqrt(s2))
                               : std::sqrt(ds2) / (1 + std::sqrt(s2));
        ++it;
        if (err < s.prm.fixed_point_tol || (double)it >= s.prm.fixed_point_max_iter) break;
    }
    s.st.fixed_point_iterations = it;
    s.st.iterations = total;
}

// update_contact_properties (OptimGradient.hpp:19-54), Viscous / ViscousFriction only:
// gamma = min(max(gamma - dt n . lambda_global, gamma_min), 0) with the last solve's lambda.
void update_contact_properties(Scene& s, double dt) {
    const int mode = problem_mode(s.prm.problem);
    if ((mode != kModeViscous && mode != kModeViscousFriction) || s.contacts.empty()) return;
    MinProblem mp(s, s.contacts, dt);
    std::vector<double> lg(3 * s.contacts.size());
    mp.l2g(s.lambda.data(), lg.data());
    for (size_t k = 0; k < s.contacts.size(); ++k) {
        Contact& c = s.contacts[k];
        c.gamma = std::min(std::max(c.gamma - dt * dot(c.nij.data(), lg.data() + 3 * k, s.dim), c.gamma_min), 0.);
    }
}

// ScopiSolver::transfer (solver.hpp:241-259): a new contact whose (i, j) was a contact of the last
// step takes that contact's property (the last one of that pair when the pair occurs twice).
void transfer(const std::vector<Contact>& old, std::vector<Contact>& cs) {
    std::map<std::pair<int64_t, int64_t>, size_t> index;
    for (size_t k = 0; k < old.size(); ++k) index[{old[k].i, old[k].j}] = k;
    for (Contact& c : cs) {
        auto it = index.find({c.i, c.j});
        if (it == index.end()) continue;
        const Contact& o = old[it->second];
        c.mu = o.mu;
        c.gamma = o.gamma;
        c.gamma_min = o.gamma_min;
        c.gamma_tol = o.gamma_tol;
    }
}

// Warm start (not in the reference): lambda_0 of the current contacts from the last solve's
// multipliers, matched by (i, j) as ScopiSolver::transfer matches contacts (solver.hpp:241-259, the
// last occurrence of a pair wins), 0 for a new contact.  Only NoFriction and Friction (their rows
// do not change between steps), and only when the last solve was warm-started as well.
bool warm_start_lambda(const Scene& s, std::vector<double>& lam0) {
    if (!s.prm.warm_start || s.prm.method != ORACLE_METHOD_APGD || !s.lambda_is_history) return false;
    if (s.prm.problem != ORACLE_NOFRICTION && s.prm.problem != ORACLE_FRICTION) return false;
    const size_t nr = s.prm.problem == ORACLE_FRICTION ? 3 : 1;
    if (s.old_contacts.empty() || s.lambda.size() != nr * s.old_contacts.size()) return false;
    std::map<std::pair<int64_t, int64_t>, size_t> index;
    for (size_t k = 0; k < s.old_contacts.size(); ++k) index[{s.old_contacts[k].i, s.old_contacts[k].j}] = k;
    lam0.assign(nr * s.contacts.size(), 0.0);
    for (size_t k = 0; k < s.contacts.size(); ++k) {
        auto it = index.find({s.contacts[k].i, s.contacts[k].j});
        if (it == index.end()) continue;
        for (size_t d = 0; d < nr; ++d) lam0[nr * k + d] = s.lambda[nr * it->second + d];
    }
    return true;
}

void optim_solve(Scene& s, double dt, const std::vector<double>* sij) {
    const int64_t na = s.na(), off = s.n_inactive;
    s.u.assign(3 * na, 0.0);
    s.w.assign(3 * na, 0.0);
    if (!s.contacts.empty()) {
        MinProblem mp(s, s.contacts, dt, sij);
        std::vector<double> lam0;
        const bool warm = warm_start_lambda(s, lam0);
        SolveResult r = (s.prm.method == ORACLE_METHOD_PGD) ? pgd(mp, s.prm) : apgd(mp, s.prm, warm ? &lam0 : nullptr);
        s.lambda = r.lambda;
        std::vector<double> vel(6 * na);
        mp.velocities(s.lambda.data(), vel.data());
        for (int64_t i = 0; i < na; ++i) {
            for (int d = 0; d < s.dim; ++d) {
                s.u[3 * i + d] = s.v[i + off][d] + dt * vel[i * 3 + d];
                if (s.dim == 3) s.w[3 * i + d] = s.omega[i + off][d] + dt * vel[(i + na) * 3 + d];
            }
            if (s.dim == 2) s.w[3 * i + 2] = s.omega[i + off][2] + dt * vel[(i + na) * 3 + 2];
        }
        s.st.iterations = r.iterations;
        s.st.backtracks = r.backtracks;
        s.st.q_applies = mp.q_applies;
        s.st.converged = r.converged;
        s.st.total_iterations += r.iterations;
        s.st.total_q_applies += mp.q_applies;
    } else {
        s.lambda.clear();
        for (int64_t i = 0; i < na; ++i) {
            for (int d = 0; d < s.dim; ++d) {
                s.u[3 * i + d] = s.v[i + off][d];
                if (s.dim == 3) s.w[3 * i + d] = s.omega[i + off][d];
            }
            if (s.dim == 2) s.w[3 * i + 2] = s.omega[i + off][2];
        }
        s.st.iterations = 0;
        s.st.backtracks = 0;
        s.st.q_applies = 0;
        s.st.converged = 1;
    }
}

// vap_fixed (vap/vap_fixed.hpp:46-55) / vap_fpd (vap/vap_fpd.hpp:70-82, src/vap/vap_fpd.cpp:6-26)
void set_a_priori_velocity(Scene& s, double dt) {
#pragma omp parallel for schedule(static)
    for (int64_t i = s.n_inactive; i < s.n; ++i) {
        if (s.prm.vap == ORACLE_VAP_PROJECTION) {
            // vap_projection::set_a_priori_velocity_impl (vap/vap_projection.hpp:34-49): the desired
            // velocity takes u, the rotation velocity w(i, 2) (a scalar assigned to a 3-vector in 3D)
            const int64_t a = i - s.n_inactive;
            for (int d = 0; d < s.dim; ++d) s.vd[i][d] = s.proj_u[3 * a + d];
            if (s.dim == 2)
                s.omega[i][2] = s.proj_w[3 * a + 2];
            else
                for (int d = 0; d < 3; ++d) s.omega[i][d] = s.proj_w[3 * a + 2];
        } else if (s.prm.vap == ORACLE_VAP_FIXED) {
            s.v[i] = s.vd[i];
            s.omega[i] = s.domega[i];
        } else {
            for (int d = 0; d < s.dim; ++d) s.v[i][d] += dt * s.f[i][d] / s.m[i];
            if (s.dim == 3) {
                const double o1 = s.omega[i][0], o2 = s.omega[i][1], o3 = s.omega[i][2];
                const double j1 = s.J[i][0], j2 = s.J[i][1], j3 = s.J[i][2];
                vec3 res{o2 * o3 * (j3 - j2), o1 * o3 * (j1 - j3), o1 * o2 * (j2 - j1)};
                for (int d = 0; d < 3; ++d) s.omega[i][d] += res[d];
            }
        }
    }
}

// q <- q * exp(omega dt / 2), normalised: solver.hpp:329-355 and 440-480
void rotate(int dim, quat& q, const vec3& om, double dt) {
    vec3 w{0, 0, 0};
    double normw;
    if (dim == 2) {
        w = {0, 0, om[2]};
        normw = std::abs(om[2]);
    } else {
        w = om;
        normw = nrm2(w.data(), 3);
    }
    if (normw == 0) normw = 1;
    quat expw;
    expw[0] = std::cos(0.5 * normw * dt);
    for (int d = 0; d < 3; ++d) expw[d + 1] = std::sin(0.5 * normw * dt) / normw * w[d];
    q = mult_quaternion(q, expw);
    normalize(q);
}

// displacement_obstacles: solver.hpp:329-355 (get_omega: 2D (0,0,w), 3D w)
void displacement_obstacles(Scene& s, double dt) {
    for (int64_t i = 0; i < s.n_inactive; ++i) {
        vec3 w = s.domega[i];
        if (s.dim == 2) w = {0, 0, s.domega[i][2]};
        double normw = nrm2(w.data(), 3);
        if (normw == 0) normw = 1;
        quat expw;
        expw[0] = std::cos(0.5 * normw * dt);
        for (int d = 0; d < 3; ++d) expw[d + 1] = std::sin(0.5 * normw * dt) / normw * w[d];
        for (int d = 0; d < s.dim; ++d) s.pos[i][d] += dt * s.vd[i][d];
        s.q[i] = mult_quaternion(s.q[i], expw);
        normalize(s.q[i]);
    }
}

// update_velocity (solver.hpp:527-546) + move_active_particles (solver.hpp:440-480)
void update_and_move(Scene& s, double dt) {
    const int64_t off = s.n_inactive;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < s.na(); ++i) {
        for (int d = 0; d < s.dim; ++d) s.v[i + off][d] = s.u[3 * i + d];
        if (s.dim == 2)
            s.omega[i + off][2] = s.w[3 * i + 2];
        else
            for (int d = 0; d < 3; ++d) s.omega[i + off][d] = s.w[3 * i + d];
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = off; i < s.n; ++i) {
        rotate(s.dim, s.q[i], s.omega[i], dt);
        for (int d = 0; d < s.dim; ++d) s.pos[i][d] += dt * s.v[i][d];
    }
}

std::vector<Contact> search(const Scene& s) {
    return (s.prm.contact_method == ORACLE_CONTACT_BRUTE_FORCE) ? contacts_brute_force(s) : contacts_kdtree(s);
}

// contact_base::run with a periodic box (contact/base.hpp:70-77): periodic copies of the
// objects near a lower bound (add_objects_from_periodicity, utils.hpp:228-257: per periodic
// axis d in order, every object io with pos[d] - dmax < lower(d), copied at pos[d] + L(d),
// appended after all objects and searched as active particles), the search on the extended
// set, then compute_exact_distance's periodic rule (base.hpp:116-139): a pair is kept when
// one of its particles is an original, a copy's index maps back to its original, and a
// copy's contact point moves back by L along every periodic axis.  The kept pairs are
// sorted by (i, j) (sort_contacts); pairs that meet twice (directly and through a copy)
// keep the search's order, the direct one first.
std::vector<Contact> compute_contacts(const Scene& s) {
    if (!s.box.any()) return search(s);
    Scene e = s;
    std::vector<int64_t> orig;
    for (int d = 0; d < s.dim; ++d) {
        if (!s.box.periodic[d]) continue;
        const double L = s.box.hi[d] - s.box.lo[d];
        for (int64_t io = 0; io < s.n; ++io) {
            if (!(s.pos[io][d] - s.prm.dmax < s.box.lo[d])) continue;
            vec3 p = s.pos[io];
            p[d] += L;
            e.shape.push_back(s.shape[io]);
            e.pos.push_back(p);
            e.size.push_back(s.size[io]);
            e.sq.push_back(s.sq[io]);
            e.q.push_back(s.q[io]);
            e.v.push_back(s.v[io]);
            e.vd.push_back(s.vd[io]);
            e.omega.push_back(s.omega[io]);
            e.domega.push_back(s.domega[io]);
            e.f.push_back(s.f[io]);
            e.J.push_back(s.J[io]);
            e.m.push_back(s.m[io]);
            orig.push_back(io);
        }
    }
    e.n = (int64_t)e.pos.size();
    std::vector<Contact> ext = search(e), out;
    for (Contact c : ext) {
        if (!(c.i < s.n || c.j < s.n)) continue;
        const bool gi = c.i >= s.n, gj = c.j >= s.n;
        if (gi) c.i = orig[c.i - s.n];
        if (gj) c.j = orig[c.j - s.n];
        for (int d = 0; d < s.dim; ++d) {
            if (!s.box.periodic[d]) continue;
            const double L = s.box.hi[d] - s.box.lo[d];
            if (gi) c.pi[d] -= L;
            if (gj) c.pj[d] -= L;
        }
        out.push_back(c);
    }
    std::stable_sort(out.begin(), out.end(), [](const Contact& a, const Contact& b) {
        return a.i < b.i || (a.i == b.i && a.j < b.j);
    });
    return out;
}

// the end of move_active_particles (solver.hpp:482-522): along a periodic axis, an object
// past the upper bound moves back by L, one past the lower bound forward by L (obstacles too)
void periodic_wrap(Scene& s) {
    if (!s.box.any()) return;
    for (int64_t io = 0; io < s.n; ++io)
        for (int d = 0; d < s.dim; ++d) {
            if (!s.box.periodic[d]) continue;
            const double L = s.box.hi[d] - s.box.lo[d];
            const bool plus = s.pos[io][d] > s.box.hi[d], minus = s.pos[io][d] < s.box.lo[d];
            if (plus) s.pos[io][d] -= L;
            if (minus) s.pos[io][d] += L;
        }
}

}  // namespace

extern "C" {

void* oracle_scene_new(void) {
    Scene* s = new Scene();
    oracle_default_params(&s->prm);
    return s;
}
void oracle_scene_free(void* scene) { delete static_cast<Scene*>(scene); }
void* oracle_scene_clone(const void* scene) { return scene ? new Scene(*static_cast<const Scene*>(scene)) : nullptr; }

// Defaults: ContactsParams (contact_kdtree.hpp:52,60), apgd_params (apgd.hpp:120-123),
// contact_property<Friction>::mu (contact/property.hpp:72).
int oracle_default_params(oracle_params* p) {
    if (!p) return -1;
    std::memset(p, 0, sizeof(*p));
    p->problem = ORACLE_NOFRICTION;
    p->contact_method = ORACLE_CONTACT_KDTREE;
    p->vap = ORACLE_VAP_FIXED;
    p->method = ORACLE_METHOD_APGD;
    p->dmax = 2.0;
    p->kd_tree_radius = 17.0;
    p->alpha = 0.05;
    p->max_ite = 10000;
    p->tolerance = 1e-7;
    p->dynamic_descent = 1;
    p->mu = 0.1;
    p->fixed_point_tol = 1e-6;
    p->fixed_point_max_iter = 1000;
    p->gamma = 0.0;  // contact_property<Viscous> (contact/property.hpp:125-127)
    p->gamma_min = -2.;
    p->gamma_tol = 1e-6;
    return 0;
}

int oracle_set_params(void* scene, const oracle_params* p) {
    if (!scene || !p) return -1;
    static_cast<Scene*>(scene)->prm = *p;
    return 0;
}

int oracle_scene_set(void* scene, int32_t dim, int64_t n, int64_t n_inactive, const int32_t* shape, const double* pos,
                     const double* q, const double* size, const double* squareness, const double* v, const double* vd, const double* omega,
                     const double* domega, const double* f, const double* m, const double* J) {
    if (!scene || (dim != 2 && dim != 3) || n < 0 || n_inactive < 0 || n_inactive > n) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    s.dim = dim;
    s.n = n;
    s.n_inactive = n_inactive;
    s.shape.assign(shape, shape + n);
    auto load3 = [&](std::vector<vec3>& dst, const double* src) {
        dst.resize(n);
        for (int64_t i = 0; i < n; ++i) dst[i] = {src[3 * i], src[3 * i + 1], src[3 * i + 2]};
    };
    load3(s.pos, pos);
    load3(s.size, size);
    load3(s.v, v);
    load3(s.vd, vd);
    load3(s.omega, omega);
    load3(s.domega, domega);
    load3(s.f, f);
    load3(s.J, J);
    s.q.resize(n);
    for (int64_t i = 0; i < n; ++i) s.q[i] = {q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
    s.sq.assign(n, {1.0, 1.0});
    if (squareness)
        for (int64_t i = 0; i < n; ++i) s.sq[i] = {squareness[2 * i], squareness[2 * i + 1]};
    s.m.assign(m, m + n);
    s.contacts.clear();
    s.old_contacts.clear();
    s.lambda.clear();
    s.st = oracle_stats{};
    return 0;
}

int oracle_set_box(void* scene, const double* lo, const double* hi, const int32_t* periodic) {
    if (!scene || !lo || !hi || !periodic) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    for (int d = 0; d < s.dim; ++d) {
        s.box.lo[d] = lo[d];
        s.box.hi[d] = hi[d];
        s.box.periodic[d] = periodic[d] != 0;
    }
    for (int d = s.dim; d < 3; ++d) s.box.periodic[d] = false;
    return 0;
}

int oracle_set_projection_velocity(void* scene, const double* u, const double* w) {
    if (!scene || !u || !w) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    s.proj_u.assign(u, u + 3 * s.na());
    s.proj_w.assign(w, w + 3 * s.na());
    return 0;
}

int oracle_get_state(void* scene, double* pos, double* q, double* v, double* omega) {
    if (!scene) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    for (int64_t i = 0; i < s.n; ++i) {
        for (int d = 0; d < 3; ++d) {
            if (pos) pos[3 * i + d] = s.pos[i][d];
            if (v) v[3 * i + d] = s.v[i][d];
            if (omega) omega[3 * i + d] = s.omega[i][d];
        }
        for (int d = 0; d < 4; ++d)
            if (q) q[4 * i + d] = s.q[i][d];
    }
    return 0;
}

// ScopiSolver::run: solver.hpp:266-303 (no output).  The last step's contacts stay with the
// scene between calls, as m_old_contacts stays with the solver.
int oracle_run(void* scene, double dt, int64_t total_it, int64_t initial_iter) {
    if (!scene) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    for (int64_t nite = initial_iter; nite < total_it; ++nite) {
        auto t0 = std::chrono::steady_clock::now();
        displacement_obstacles(s, dt);
        s.contacts = compute_contacts(s);
        if (!s.old_contacts.empty()) transfer(s.old_contacts, s.contacts);
        s.st.n_contacts = (int64_t)s.contacts.size();
        s.st.t_contacts += seconds_since(t0);
        t0 = std::chrono::steady_clock::now();
        set_a_priori_velocity(s, dt);
        s.st.t_vap += seconds_since(t0);
        t0 = std::chrono::steady_clock::now();
        optim_run(s, dt);
        update_contact_properties(s, dt);
        s.old_contacts = s.contacts;
        s.lambda_is_history =
            s.prm.warm_start && (s.prm.problem == ORACLE_NOFRICTION || s.prm.problem == ORACLE_FRICTION);
        s.st.t_solve += seconds_since(t0);
        t0 = std::chrono::steady_clock::now();
        update_and_move(s, dt);
        periodic_wrap(s);
        s.st.t_move += seconds_since(t0);
    }
    return 0;
}

int64_t oracle_compute_contacts(void* scene) {
    if (!scene) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    s.contacts = compute_contacts(s);
    s.st.n_contacts = (int64_t)s.contacts.size();
    return (int64_t)s.contacts.size();
}

// Replace the contact list with given records (contact properties from the parameters, as a new
// contact's): the optim slot's problem on another code's contacts, for checking its solution.
int oracle_set_contacts(void* scene, int64_t nc, const int64_t* ij, const double* dij, const double* nij,
                        const double* pi, const double* pj) {
    if (!scene || nc < 0 || (nc > 0 && (!ij || !dij || !nij || !pi || !pj))) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    s.contacts.assign((size_t)nc, Contact{});
    for (int64_t c = 0; c < nc; ++c) {
        Contact& ct = s.contacts[c];
        ct.i = ij[2 * c];
        ct.j = ij[2 * c + 1];
        if (ct.i < 0 || ct.i >= s.n || ct.j < 0 || ct.j >= s.n) return -1;
        ct.dij = dij[c];
        for (int d = 0; d < 3; ++d) {
            ct.nij[d] = nij[3 * c + d];
            ct.pi[d] = pi[3 * c + d];
            ct.pj[d] = pj[3 * c + d];
        }
        ct.mu = s.prm.mu;
        ct.gamma = s.prm.gamma;
        ct.gamma_min = s.prm.gamma_min;
        ct.gamma_tol = s.prm.gamma_tol;
    }
    s.st.n_contacts = nc;
    return 0;
}

int64_t oracle_n_contacts(void* scene) {
    if (!scene) return -1;
    return (int64_t) static_cast<Scene*>(scene)->contacts.size();
}

int oracle_get_contacts(void* scene, int64_t* ij, double* dij, double* nij, double* pi, double* pj) {
    if (!scene) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    for (size_t c = 0; c < s.contacts.size(); ++c) {
        const Contact& ct = s.contacts[c];
        if (ij) {
            ij[2 * c] = ct.i;
            ij[2 * c + 1] = ct.j;
        }
        if (dij) dij[c] = ct.dij;
        for (int d = 0; d < 3; ++d) {
            if (nij) nij[3 * c + d] = ct.nij[d];
            if (pi) pi[3 * c + d] = ct.pi[d];
            if (pj) pj[3 * c + d] = ct.pj[d];
        }
    }
    return 0;
}

int oracle_amat(void* scene, const double* u, double* out) {
    if (!scene) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    amat(s, s.contacts, u, out);
    return 0;
}

int oracle_atmat(void* scene, const double* f, double* out) {
    if (!scene) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    atmat(s, s.contacts, f, out);
    return 0;
}

int oracle_solve(void* scene, double dt, double* lambda, double* u, double* w) {
    if (!scene) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    optim_run(s, dt);
    s.lambda_is_history = false;  // these multipliers belong to s.contacts, not to old_contacts
    if (lambda) std::copy(s.lambda.begin(), s.lambda.end(), lambda);
    if (u) std::copy(s.u.begin(), s.u.end(), u);
    if (w) std::copy(s.w.begin(), s.w.end(), w);
    return 0;
}

// MinProblem::gradient (minimization_problem.hpp:136-140) at a given lambda: the test of the KKT
// conditions of another solver's multipliers on the same problem
int oracle_gradient(void* scene, double dt, const double* lambda, double* out) {
    if (!scene || !lambda || !out) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    if (s.prm.problem != ORACLE_NOFRICTION && s.prm.problem != ORACLE_FRICTION) return -2;
    if (s.contacts.empty()) return 0;
    MinProblem mp(s, s.contacts, dt, nullptr);
    mp.gradient(lambda, out);
    return 0;
}

int oracle_update_contact_properties(void* scene, double dt) {
    if (!scene) return -1;
    update_contact_properties(*static_cast<Scene*>(scene), dt);
    return 0;
}

int oracle_set_forces(void* scene, const double* f) {
    if (!scene || !f) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    for (int64_t i = 0; i < s.n; ++i) s.f[i] = {f[3 * i], f[3 * i + 1], f[3 * i + 2]};
    return 0;
}

int64_t oracle_lambda_size(void* scene) {
    if (!scene) return -1;
    return (int64_t) static_cast<Scene*>(scene)->lambda.size();
}

int oracle_get_contact_gamma(void* scene, double* gamma) {
    if (!scene || !gamma) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    for (size_t k = 0; k < s.contacts.size(); ++k) gamma[k] = s.contacts[k].gamma;
    return 0;
}

int oracle_set_contact_gamma(void* scene, const double* gamma) {
    if (!scene || !gamma) return -1;
    Scene& s = *static_cast<Scene*>(scene);
    for (size_t k = 0; k < s.contacts.size(); ++k) s.contacts[k].gamma = gamma[k];
    return 0;
}

int oracle_set_a_priori_velocity(void* scene, double dt) {
    if (!scene) return -1;
    set_a_priori_velocity(*static_cast<Scene*>(scene), dt);
    return 0;
}

int oracle_get_stats(void* scene, oracle_stats* st) {
    if (!scene || !st) return -1;
    *st = static_cast<Scene*>(scene)->st;
    return 0;
}

int oracle_closest_points(int32_t dim, int32_t shape_i, const double* pos_i, const double* q_i, const double* size_i,
                          const double* sq_i, int32_t shape_j, const double* pos_j, const double* q_j,
                          const double* size_j, const double* sq_j, double* dij, double* nij, double* pi, double* pj) {
    vec3 a{pos_i[0], pos_i[1], pos_i[2]}, b{pos_j[0], pos_j[1], pos_j[2]};
    vec3 sa{size_i[0], size_i[1], size_i[2]}, sb{size_j[0], size_j[1], size_j[2]};
    quat qa{q_i[0], q_i[1], q_i[2], q_i[3]}, qb{q_j[0], q_j[1], q_j[2], q_j[3]};
    std::array<double, 2> ea{1.0, 1.0}, eb{1.0, 1.0};
    if (sq_i) ea = {sq_i[0], sq_i[1]};
    if (sq_j) eb = {sq_j[0], sq_j[1]};
    Contact c{};
    if (!closest_points(dim, shape_i, a, qa, sa, ea, shape_j, b, qb, sb, eb, c)) return -2;
    *dij = c.dij;
    for (int d = 0; d < 3; ++d) {
        nij[d] = c.nij[d];
        pi[d] = c.pi[d];
        pj[d] = c.pj[d];
    }
    return 0;
}

namespace {
se_oracle::SE se_from(const double* pos, const double* q, const double* size, const double* sq) {
    return make_se({pos[0], pos[1], pos[2]}, {q[0], q[1], q[2], q[3]}, {size[0], size[1], size[2]},
                   {sq ? sq[0] : 1.0, sq ? sq[1] : 1.0});
}
}  // namespace

int oracle_se_closest_points(const double* pos_i, const double* q_i, const double* size_i, const double* sq_i,
                             const double* pos_j, const double* q_j, const double* size_j, const double* sq_j,
                             double* dij, double* nij, double* pi, double* pj, double* u0, double* u, int32_t* info,
                             int32_t* nfev, int32_t* njev) {
    const se_oracle::SE s1 = se_from(pos_i, q_i, size_i, sq_i), s2 = se_from(pos_j, q_j, size_j, sq_j);
    double start[4];
    if (g_se_math)
        se_oracle::se_grid_start<se_oracle::PortableMath>(s1, s2, start);
    else
        se_oracle::se_grid_start<se_oracle::LibMath>(s1, s2, start);
    const se_oracle::SEResult r = se_solve(s1, s2);
    *dij = r.dij;
    for (int d = 0; d < 3; ++d) {
        nij[d] = r.nij[d];
        pi[d] = r.pi[d];
        pj[d] = r.pj[d];
    }
    for (int k = 0; k < 4; ++k) {
        if (u0) u0[k] = start[k];
        if (u) u[k] = r.u[k];
    }
    if (info) *info = r.info;
    if (nfev) *nfev = r.nfev;
    if (njev) *njev = r.njev;
    return 0;
}

int oracle_newton_closest_points(int32_t shape_i, const double* pos_i, const double* q_i, const double* size_i,
                                 const double* sq_i, int32_t shape_j, const double* pos_j, const double* q_j,
                                 const double* size_j, const double* sq_j, double* dij, double* nij, double* pi,
                                 double* pj, double* u0, double* u, int32_t* iterations) {
    se_oracle::NewtonResult r;
    const vec3 a{pos_i[0], pos_i[1], pos_i[2]}, b{pos_j[0], pos_j[1], pos_j[2]};
    const vec3 sa{size_i[0], size_i[1], size_i[2]}, sb{size_j[0], size_j[1], size_j[2]};
    const quat qa{q_i[0], q_i[1], q_i[2], q_i[3]}, qb{q_j[0], q_j[1], q_j[2], q_j[3]};
    const std::array<double, 2> ea{sq_i ? sq_i[0] : 1.0, sq_i ? sq_i[1] : 1.0}, eb{sq_j ? sq_j[0] : 1.0, sq_j ? sq_j[1] : 1.0};
    if (!newton_pair(shape_i, a, qa, sa, ea, shape_j, b, qb, sb, eb, r)) return -2;
    *dij = r.dij;
    for (int d = 0; d < 3; ++d) {
        nij[d] = r.nij[d];
        pi[d] = r.pi[d];
        pj[d] = r.pj[d];
    }
    for (int k = 0; k < 4; ++k) {
        if (u0) u0[k] = r.u0[k];
        if (u) u[k] = r.u[k];
    }
    if (iterations) *iterations = r.iterations;
    return 0;
}

int oracle_newton_residual(int32_t kind, const double* pos_i, const double* q_i, const double* size_i,
                           const double* sq_i, const double* pos_j, const double* q_j, const double* size_j,
                           const double* u, double* F, double* J) {
    const se_oracle::SE s1 = se_from(pos_i, q_i, size_i, sq_i);
    if (kind == 0) {
        const se_oracle::Sph s2 = make_sph({pos_j[0], pos_j[1], pos_j[2]}, {q_j[0], q_j[1], q_j[2], q_j[3]},
                                           {size_j[0], size_j[1], size_j[2]});
        if (g_se_math)
            se_oracle::sph_se_residual<se_oracle::PortableMath>(s1, s2, u, F, J);
        else
            se_oracle::sph_se_residual<se_oracle::LibMath>(s1, s2, u, F, J);
    } else {
        const se_oracle::Pln p2 = make_pln({pos_j[0], pos_j[1], pos_j[2]}, {q_j[0], q_j[1], q_j[2], q_j[3]});
        const se_oracle::v3 np = se_oracle::pln_normal(p2);
        const double sp = (s1.c[0] - p2.c[0]) * np[0] + (s1.c[1] - p2.c[1]) * np[1] + (s1.c[2] - p2.c[2]) * np[2];
        if (g_se_math)
            se_oracle::se_pln_residual<se_oracle::PortableMath>(s1, p2, sp, u, F, J);
        else
            se_oracle::se_pln_residual<se_oracle::LibMath>(s1, p2, sp, u, F, J);
    }
    return 0;
}

int oracle_se2_closest_points(const double* pos_i, const double* q_i, const double* size_i, double e_i,
                              const double* pos_j, const double* q_j, const double* size_j, double e_j, double* dij,
                              double* nij, double* pi, double* pj, double* u, int32_t* info) {
    const std::array<double, 2> si{e_i, 1.0}, sj{e_j, 1.0};
    const se_oracle::SE a = make_se2({pos_i[0], pos_i[1], 0.0}, {q_i[0], q_i[1], q_i[2], q_i[3]},
                                     {size_i[0], size_i[1], 0.0}, si);
    const se_oracle::SE b = make_se2({pos_j[0], pos_j[1], 0.0}, {q_j[0], q_j[1], q_j[2], q_j[3]},
                                     {size_j[0], size_j[1], 0.0}, sj);
    const se_oracle::SEResult r = g_se_math ? se_oracle::se2_se2<se_oracle::PortableMath>(a, b)
                                            : se_oracle::se2_se2<se_oracle::LibMath>(a, b);
    *dij = r.dij;
    for (int d = 0; d < 2; ++d) {
        nij[d] = r.nij[d];
        pi[d] = r.pi[d];
        pj[d] = r.pj[d];
        u[d] = r.u[d];
    }
    *info = r.info;
    return 0;
}

int oracle_se2_residual(const double* pos_i, const double* q_i, const double* size_i, double e_i,
                        const double* pos_j, const double* q_j, const double* size_j, double e_j, const double* u,
                        double* F, double* J) {
    const std::array<double, 2> si{e_i, 1.0}, sj{e_j, 1.0};
    const se_oracle::SE a = make_se2({pos_i[0], pos_i[1], 0.0}, {q_i[0], q_i[1], q_i[2], q_i[3]},
                                     {size_i[0], size_i[1], 0.0}, si);
    const se_oracle::SE b = make_se2({pos_j[0], pos_j[1], 0.0}, {q_j[0], q_j[1], q_j[2], q_j[3]},
                                     {size_j[0], size_j[1], 0.0}, sj);
    if (g_se_math)
        se_oracle::se2_se2_residual<se_oracle::PortableMath>(a, b, u, F, J);
    else
        se_oracle::se2_se2_residual<se_oracle::LibMath>(a, b, u, F, J);
    return 0;
}

int oracle_lu_solve4(const double* A, const double* b, double* x) {
    double M[16];
    for (int k = 0; k < 16; ++k) M[k] = A[k];
    for (int k = 0; k < 4; ++k) x[k] = b[k];
    se_oracle::lu_solve4(M, x);
    return 0;
}

int oracle_se_residual(const double* pos_i, const double* q_i, const double* size_i, const double* sq_i,
                       const double* pos_j, const double* q_j, const double* size_j, const double* sq_j,
                       const double* u, double* F, double* J) {
    se_oracle::SEPair P{se_from(pos_i, q_i, size_i, sq_i), se_from(pos_j, q_j, size_j, sq_j)};
    if (g_se_math)
        se_oracle::se_se_residual<se_oracle::PortableMath>(P, u, F, J);
    else
        se_oracle::se_se_residual<se_oracle::LibMath>(P, u, F, J);
    return 0;
}

int oracle_set_se_math(int32_t portable) {
    g_se_math = portable ? 1 : 0;
    return 0;
}

#ifdef _OPENMP
// one thread until oracle_set_threads asks for more (the reference's default build has OpenMP off)
[[maybe_unused]] static const int g_one_thread = (omp_set_num_threads(1), 1);
#endif

int oracle_set_threads(int32_t n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int oracle_set_minpack_variant(int32_t mask) {
    se_oracle::minpack_variant() = mask & 15;
    return 0;
}

int oracle_set_jacobian_ad(int32_t ad) {
    se_oracle::jacobian_ad() = ad ? 1 : 0;
    return 0;
}

int64_t oracle_dogleg_fresh_wa4_count(void) { return se_oracle::dogleg_fresh_wa4_count(); }

int oracle_minpack_dogleg(int32_t n, const double* r, const double* diag, const double* qtb, double delta,
                          double wa4_last, double* x, int32_t* gradient_branch) {
    double wa1[4], wa2[4];
    bool g = false;
    if (n == 4)
        g = se_oracle::Minpack<4>::dogleg(r, diag, qtb, delta, x, wa1, wa2, wa4_last);
    else if (n == 2)
        g = se_oracle::Minpack<2>::dogleg(r, diag, qtb, delta, x, wa1, wa2, wa4_last);
    else
        return -1;
    *gradient_branch = g ? 1 : 0;
    return 0;
}

int oracle_math(int32_t fn, double x, double y, double* out) {
    switch (fn) {
        case 0: *out = se_math::cos(x); break;
        case 1: *out = se_math::sin(x); break;
        case 2: *out = se_math::asin(x); break;
        case 3: *out = se_math::log(x); break;
        case 4: *out = se_math::exp(x); break;
        case 5: *out = se_math::pow(x, y); break;
        default: return -1;
    }
    return 0;
}

int oracle_se_point_normal(const double* pos, const double* q, const double* size, const double* sq, double a,
                           double b, double* point, double* normal) {
    const se_oracle::SE s = se_from(pos, q, size, sq);
    const se_oracle::v3 p = g_se_math ? se_oracle::se_point<se_oracle::PortableMath>(s, a, b)
                                      : se_oracle::se_point<se_oracle::LibMath>(s, a, b);
    const se_oracle::v3 n = g_se_math ? se_oracle::se_normal<se_oracle::PortableMath>(s, a, b)
                                      : se_oracle::se_normal<se_oracle::LibMath>(s, a, b);
    for (int d = 0; d < 3; ++d) {
        point[d] = p[d];
        normal[d] = n[d];
    }
    return 0;
}

int oracle_quaternion(double angle, const double* axis, double* q) {
    quat o{std::cos(angle / 2), std::sin(angle / 2) * axis[0], std::sin(angle / 2) * axis[1],
           std::sin(angle / 2) * axis[2]};
    normalize(o);
    for (int d = 0; d < 4; ++d) q[d] = o[d];
    return 0;
}

// test/test_sphere.cpp:430-470: 2*n*n spheres from std::minstd_rand0(123).
int oracle_gen_2d_case_spheres(int32_t n, double* x, double* y, double* r, double* velocity) {
    std::minstd_rand0 generator(123);
    std::uniform_real_distribution<double> distrib_r(0.2, 0.4);
    std::uniform_real_distribution<double> distrib_move_x(-0.1, 0.1);
    std::uniform_real_distribution<double> distrib_move_y(-0.1, 0.1);
    std::uniform_real_distribution<double> distrib_velocity(2., 5.);
    int k = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double rr = distrib_r(generator);
            double xx = (i + 0.5) + distrib_move_x(generator);
            double yy = (j + 0.5) + distrib_move_y(generator);
            double vel = distrib_velocity(generator);
            x[k] = xx, y[k] = yy, r[k] = rr, velocity[k] = vel;
            ++k;
            rr = distrib_r(generator);
            xx = (n + i + 0.5) + distrib_move_x(generator);
            yy = (j + 0.5) + distrib_move_y(generator);
            vel = distrib_velocity(generator);
            x[k] = xx, y[k] = yy, r[k] = rr, velocity[k] = -vel;
            ++k;
        }
    return 0;
}

/* Draws of demos/segment.cpp:30-44: std::default_random_engine (default seed), per particle
 * x ~ U(0.1, 0.9), y ~ U(0, 0.7), radius ~ U(0.01, max_radius), in that order. */
int oracle_gen_segment_demo(int32_t n, double max_radius, double* x, double* y, double* r) {
    std::default_random_engine generator;
    std::uniform_real_distribution<double> distrib_x(0.1, 0.9);
    std::uniform_real_distribution<double> distrib_y(0, 0.7);
    std::uniform_real_distribution<double> distrib_r(0.01, max_radius);
    for (int i = 0; i < n; ++i) {
        x[i] = distrib_x(generator);
        y[i] = distrib_y(generator);
        r[i] = distrib_r(generator);
    }
    return 0;
}

// n draws of uniform_real_distribution<double>(a, b) from a default-constructed
// std::default_random_engine (demos/pile_of_sand_spheres.cpp:30-71 draws its masses so).
int oracle_gen_uniform(int32_t n, double a, double b, double* out) {
    std::default_random_engine generator;
    std::uniform_real_distribution<double> distrib(a, b);
    for (int k = 0; k < n; ++k) out[k] = distrib(generator);
    return 0;
}

// The random draws of demos/box_spheres_3d.cpp:73-95, in the demo's order: for each sphere (i, j, k
// loops) its radius U(0.8 r0, 0.9 r0), mass U(1, 2) and centre offsets dx, dy, dz U(-0.05, 0.05), all
// from one std::default_random_engine.
int oracle_gen_box_spheres_demo(int32_t n_side, double r0, double* r, double* m, double* dx, double* dy, double* dz) {
    std::default_random_engine generator;
    std::uniform_real_distribution<double> distrib_r(0.8 * r0, 0.9 * r0);
    std::uniform_real_distribution<double> distrib_m(1., 2.);
    std::uniform_real_distribution<double> distrib_mouve_center(-0.05, 0.05);
    const int n = n_side * n_side * n_side;
    for (int k = 0; k < n; ++k) {
        r[k] = distrib_r(generator);
        m[k] = distrib_m(generator);
        dx[k] = distrib_mouve_center(generator);
        dy[k] = distrib_mouve_center(generator);
        dz[k] = distrib_mouve_center(generator);
    }
    return 0;
}

}  // extern "C"
