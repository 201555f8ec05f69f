// scopi_hip_host.hpp -- C++ host mirror of SCoPI's per-step slot API over libscopi_hip.so.
//
// The names, template slots and parameter fields follow the reference
// (hpc-maths/scopi v0.1.0) so that scene set-ups written for it compile against
// these headers:
//   ScopiSolver<dim, problem_t, optim_solver_t, contact_method_t, vap_t>  include/scopi/solver.hpp:91-107
//   scopi_container<dim>::push_back(object, property)                     include/scopi/container.hpp:338-384
//   property<dim> fluent setters                                           include/scopi/property.hpp
//   sphere / plane / segment constructors                                  include/scopi/objects/types/*.hpp
//   contact_kdtree / contact_brute_force (ContactsParams: dmax, kd_tree_radius)
//                                                                          include/scopi/contact/contact_kdtree.hpp:27-61
//   OptimGradient<apgd | pgd> (apgd_params: alpha, max_ite, tolerance, dynamic_descent)
//                                                                          include/scopi/solvers/apgd.hpp:20-124
//   vap_fixed / vap_fpd                                                    include/scopi/vap/*.hpp
//   NoFriction / Friction / FrictionFixedPoint / Viscous / ViscousFriction, contact_property<...>
//     (mu; fixed_point_tol, fixed_point_max_iter; gamma, gamma_min, gamma_tol)
//                                                                          include/scopi/contact/property.hpp
//   neighbor<dim, problem_t>                                               include/scopi/objects/neighbor.hpp:23-70
// Every slot runs on the GPU through the C ABI (include/scopi_hip.h).  When the
// contact, optim and vap slots are the ones defined here, ScopiSolver::run keeps
// the whole time loop on the device (one upload, one download); otherwise it
// calls the slots one by one as the reference's run() does (solver.hpp:266-303).
// Vectors are plain fixed arrays (type::position_t etc.) instead of xtensor types.
#pragma once

#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <future>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "json_output.hpp"

#include "../../../include/scopi_hip.h"

namespace scopi {

constexpr double PI = 3.14159265358979323846;

// ------------------------------------------------------------------ types
template <std::size_t N>
struct fixed {
    std::array<double, N> a{};
    double& operator()(std::size_t i) { return a[i]; }
    double operator()(std::size_t i) const { return a[i]; }
    double& operator[](std::size_t i) { return a[i]; }
    double operator[](std::size_t i) const { return a[i]; }
    static constexpr std::size_t size() { return N; }
};

namespace type {
template <std::size_t dim>
using position_t = fixed<dim>;
template <std::size_t dim>
using velocity_t = fixed<dim>;
template <std::size_t dim>
using force_t = fixed<dim>;
template <std::size_t dim>
using moment_t = typename std::conditional<dim == 2, double, fixed<3>>::type;
template <std::size_t dim>
using rotation_t = typename std::conditional<dim == 2, double, fixed<3>>::type;
using quaternion_t = fixed<4>;
}  // namespace type

// quaternion(angle, axis): src/quaternion.cpp:12-36 (scalar first, normalised)
inline type::quaternion_t quaternion(double angle, const fixed<3>& axes) {
    type::quaternion_t q{{std::cos(angle / 2), std::sin(angle / 2) * axes[0], std::sin(angle / 2) * axes[1],
                          std::sin(angle / 2) * axes[2]}};
    double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (auto& c : q.a) c /= n;
    return q;
}
inline type::quaternion_t quaternion(double angle = 0.0) { return quaternion(angle, fixed<3>{{0.0, 0.0, 1.0}}); }

// ------------------------------------------------------------------ problem tags
struct NoFriction {};
struct Friction {};
struct FrictionFixedPoint {};
struct Viscous {};
struct ViscousFriction {};

template <class problem_t>
struct contact_property {};
template <>
struct contact_property<Friction> {
    double mu = 0.1;
};
// contact_property<FrictionFixedPoint> (contact/property.hpp:75-149): Friction's mu and
// FixedPoint's outer-loop controls (fixed_point_max_iter is a double there too)
template <>
struct contact_property<FrictionFixedPoint> {
    double mu = 0.1;
    double fixed_point_tol = 1e-6;
    double fixed_point_max_iter = 1000;
};
// contact_property<Viscous> (contact/property.hpp:100-128): the adhesion potential gamma of the
// contact, its lower bound and the tolerance of gamma < -gamma_tol
template <>
struct contact_property<Viscous> {
    double gamma = 0;
    double gamma_min = -2.;
    double gamma_tol = 1e-6;
};
// contact_property<ViscousFriction> (contact/property.hpp:151-175): Viscous, Friction and FixedPoint
template <>
struct contact_property<ViscousFriction> {
    double gamma = 0;
    double gamma_min = -2.;
    double gamma_tol = 1e-6;
    double mu = 0.1;
    double fixed_point_tol = 1e-6;
    double fixed_point_max_iter = 1000;
};

// problems whose contacts carry a friction coefficient
template <class problem_t>
constexpr bool has_mu() {
    return std::is_same<problem_t, Friction>::value || std::is_same<problem_t, FrictionFixedPoint>::value ||
           std::is_same<problem_t, ViscousFriction>::value;
}
// problems whose contacts carry an adhesion potential gamma
template <class problem_t>
constexpr bool has_gamma() {
    return std::is_same<problem_t, Viscous>::value || std::is_same<problem_t, ViscousFriction>::value;
}
// problems with an outer fixed point (fixed_point_tol, fixed_point_max_iter)
template <class problem_t>
constexpr bool has_fixed_point() {
    return std::is_same<problem_t, FrictionFixedPoint>::value || std::is_same<problem_t, ViscousFriction>::value;
}

template <std::size_t dim, class problem_t_>
struct neighbor {
    using problem_t = problem_t_;
    std::size_t i = 0, j = 0;
    double dij = 0;
    fixed<dim> nij, pi, pj;
    double sij = 0;
    contact_property<problem_t> property;
};

// ------------------------------------------------------------------ property and objects
template <std::size_t dim>
class property {
  public:
    using velocity_type = type::velocity_t<dim>;
    using rotation_type = type::rotation_t<dim>;
    using force_type = type::force_t<dim>;
    using moment_type = type::moment_t<dim>;
    property& velocity(const velocity_type& v) { m_v = v; return *this; }
    property& desired_velocity(const velocity_type& v) { m_dv = v; return *this; }
    property& omega(const rotation_type& w) { m_w = w; return *this; }
    property& desired_omega(const rotation_type& w) { m_dw = w; return *this; }
    property& force(const force_type& f) { m_f = f; return *this; }
    property& mass(double m) { m_m = m; return *this; }
    property& moment_inertia(const moment_type& j) { m_j = j; return *this; }
    property& activate() { m_active = true; return *this; }
    property& deactivate() { m_active = false; return *this; }
    const velocity_type& velocity() const { return m_v; }
    const velocity_type& desired_velocity() const { return m_dv; }
    const rotation_type& omega() const { return m_w; }
    const rotation_type& desired_omega() const { return m_dw; }
    const force_type& force() const { return m_f; }
    double mass() const { return m_m; }
    const moment_type& moment_inertia() const { return m_j; }
    bool is_active() const { return m_active; }

  private:
    velocity_type m_v{}, m_dv{};
    rotation_type m_w{}, m_dw{};
    force_type m_f{};
    double m_m = 1.0;
    moment_type m_j{};
    bool m_active = true;
};

template <std::size_t dim>
struct object {
    int shape;
    type::position_t<dim> pos;
    type::quaternion_t q;
    std::array<double, 3> size;
    std::array<double, 2> sq{1.0, 1.0};  // superellipsoid squareness (e, n)
};

template <std::size_t dim>
struct sphere : object<dim> {
    sphere(type::position_t<dim> pos, double radius) : object<dim>{SCOPI_SHAPE_SPHERE, pos, quaternion(0.0), {radius, 0, 0}} {}
    sphere(type::position_t<dim> pos, type::quaternion_t q, double radius)
        : object<dim>{SCOPI_SHAPE_SPHERE, pos, q, {radius, 0, 0}} {}
    double radius() const { return this->size[0]; }
};

template <std::size_t dim>
struct plane : object<dim> {
    explicit plane(type::position_t<dim> pos, double angle = 0)
        : object<dim>{SCOPI_SHAPE_PLANE, pos, quaternion(angle), {0, 0, 0}} {}
    plane(type::position_t<dim> pos, type::quaternion_t q) : object<dim>{SCOPI_SHAPE_PLANE, pos, q, {0, 0, 0}} {}
};

// superellipsoid<dim>(pos, q, radius (semi-axes), squareness (e in 2D, (e, n) in 3D)):
// include/scopi/objects/types/superellipsoid.hpp:111-139.
template <std::size_t dim>
struct superellipsoid : object<dim> {
    // superellipsoid(pos, q, radius, squareness) (objects/types/superellipsoid.hpp:57-66): dim radii,
    // dim - 1 squareness values (e in 2D; e, n in 3D)
    superellipsoid(type::position_t<dim> pos, type::quaternion_t q, fixed<dim> radius, fixed<dim - 1> squareness)
        : object<dim>{SCOPI_SHAPE_SUPERELLIPSOID, pos, q, {radius[0], radius[1], dim == 3 ? radius[dim - 1] : 0.0},
                      {squareness[0], dim == 3 ? squareness[dim - 2] : 1.0}} {}
    // 2D: the squareness e as a number, as the reference's one-element squareness takes it
    template <std::size_t D = dim, typename = std::enable_if_t<D == 2>>
    superellipsoid(type::position_t<dim> pos, type::quaternion_t q, fixed<dim> radius, double e)
        : superellipsoid(pos, q, radius, fixed<1>{{e}}) {}
    fixed<dim> radius() const {
        fixed<dim> r;
        for (std::size_t d = 0; d < dim; ++d) r[d] = this->size[d];
        return r;
    }
    fixed<dim - 1> squareness() const {
        fixed<dim - 1> e;
        for (std::size_t d = 0; d + 1 < dim; ++d) e[d] = this->sq[d];
        return e;
    }
};

template <std::size_t dim>
struct segment : object<dim> {
    segment(type::position_t<dim> pos, type::quaternion_t q, double length)
        : object<dim>{SCOPI_SHAPE_SEGMENT, pos, q, {length, 0, 0}} {}
    // segment(pos1, pos2) (objects/types/segment.hpp:133-138): centre at the midpoint, angle
    // atan2(x1 - x2, y2 - y1) about z, length |pos2 - pos1|
    segment(const type::position_t<dim>& pos1, const type::position_t<dim>& pos2)
        : segment(midpoint(pos1, pos2), quaternion(std::atan2(pos1[0] - pos2[0], pos2[1] - pos1[1])),
                  distance(pos1, pos2)) {}

  private:
    static type::position_t<dim> midpoint(const type::position_t<dim>& a, const type::position_t<dim>& b) {
        type::position_t<dim> c;
        for (std::size_t d = 0; d < dim; ++d) c[d] = 0.5 * (a[d] + b[d]);
        return c;
    }
    static double distance(const type::position_t<dim>& a, const type::position_t<dim>& b) {
        double s = 0.0;
        for (std::size_t d = 0; d < dim; ++d) s += (b[d] - a[d]) * (b[d] - a[d]);
        return std::sqrt(s);
    }
};

// ------------------------------------------------------------------ container
template <std::size_t dim>
class scopi_container {
  public:
    static constexpr std::size_t dimension = dim;
    void push_back(const object<dim>& s, const property<dim>& p = property<dim>()) {
        if (!p.is_active() && m_nb_inactive != size())
            throw std::runtime_error("All the obstacles must be pushed before the active particles.");
        m_shape.push_back(s.shape);
        m_pos.push_back(s.pos);
        m_q.push_back(s.q);
        m_size.push_back(s.size);
        m_sq.push_back(s.sq);
        m_v.push_back(p.velocity());
        m_vd.push_back(p.desired_velocity());
        m_w.push_back(p.omega());
        m_dw.push_back(p.desired_omega());
        m_f.push_back(p.force());
        m_m.push_back(p.mass());
        m_j.push_back(p.moment_inertia());
        if (!p.is_active()) ++m_nb_inactive;
    }
    std::size_t size() const { return m_pos.size(); }
    std::size_t nb_active() const { return size() - m_nb_inactive; }
    std::size_t nb_inactive() const { return m_nb_inactive; }
    std::vector<type::position_t<dim>>& pos() { return m_pos; }
    std::vector<type::quaternion_t>& q() { return m_q; }
    std::vector<type::velocity_t<dim>>& v() { return m_v; }
    std::vector<type::velocity_t<dim>>& vd() { return m_vd; }
    std::vector<type::rotation_t<dim>>& omega() { return m_w; }
    std::vector<type::rotation_t<dim>>& desired_omega() { return m_dw; }
    std::vector<type::force_t<dim>>& f() { return m_f; }
    std::vector<double>& m() { return m_m; }
    std::vector<type::moment_t<dim>>& j() { return m_j; }
    const std::vector<type::position_t<dim>>& pos() const { return m_pos; }
    const std::vector<type::quaternion_t>& q() const { return m_q; }
    const std::vector<type::velocity_t<dim>>& v() const { return m_v; }
    const std::vector<type::velocity_t<dim>>& vd() const { return m_vd; }
    const std::vector<type::rotation_t<dim>>& omega() const { return m_w; }
    const std::vector<type::rotation_t<dim>>& desired_omega() const { return m_dw; }
    const std::vector<type::force_t<dim>>& f() const { return m_f; }
    const std::vector<double>& m() const { return m_m; }
    const std::vector<type::moment_t<dim>>& j() const { return m_j; }
    // shape tag, size (radius / semi-axes / length) and squareness of object k
    int shape(std::size_t k) const { return m_shape[k]; }
    const std::array<double, 3>& object_size(std::size_t k) const { return m_size[k]; }
    const std::array<double, 2>& object_squareness(std::size_t k) const { return m_sq[k]; }

    // structure of arrays in the C-ABI convention (3 slots, 2D rotation in slot 2)
    struct soa {
        std::vector<int32_t> shape;
        std::vector<double> pos, q, size, sq, v, vd, w, dw, f, m, J;
    };
    soa pack() const {
        soa s;
        const std::size_t n = size();
        s.shape.assign(m_shape.begin(), m_shape.end());
        s.pos.assign(3 * n, 0.0), s.q.assign(4 * n, 0.0), s.size.assign(3 * n, 0.0), s.sq.assign(2 * n, 1.0);
        s.v.assign(3 * n, 0.0), s.vd.assign(3 * n, 0.0), s.w.assign(3 * n, 0.0), s.dw.assign(3 * n, 0.0);
        s.f.assign(3 * n, 0.0), s.m.assign(n, 0.0), s.J.assign(3 * n, 0.0);
        for (std::size_t k = 0; k < n; ++k) {
            for (std::size_t d = 0; d < dim; ++d) {
                s.pos[3 * k + d] = m_pos[k][d];
                s.v[3 * k + d] = m_v[k][d];
                s.vd[3 * k + d] = m_vd[k][d];
                s.f[3 * k + d] = m_f[k][d];
            }
            for (int d = 0; d < 4; ++d) s.q[4 * k + d] = m_q[k][d];
            for (int d = 0; d < 3; ++d) s.size[3 * k + d] = m_size[k][d];
            for (int d = 0; d < 2; ++d) s.sq[2 * k + d] = m_sq[k][d];
            rot_in(s.w, k, m_w[k]);
            rot_in(s.dw, k, m_dw[k]);
            rot_in(s.J, k, m_j[k]);
            s.m[k] = m_m[k];
        }
        return s;
    }
    void unpack(const std::vector<double>& pos, const std::vector<double>& q, const std::vector<double>& v,
                const std::vector<double>& w) {
        for (std::size_t k = 0; k < size(); ++k) {
            for (std::size_t d = 0; d < dim; ++d) {
                m_pos[k][d] = pos[3 * k + d];
                m_v[k][d] = v[3 * k + d];
            }
            for (int d = 0; d < 4; ++d) m_q[k][d] = q[4 * k + d];
            rot_out(w, k, m_w[k]);
        }
    }

  private:
    static void rot_in(std::vector<double>& a, std::size_t k, const double& x) { a[3 * k + 2] = x; }
    static void rot_in(std::vector<double>& a, std::size_t k, const fixed<3>& x) {
        for (int d = 0; d < 3; ++d) a[3 * k + d] = x[d];
    }
    static void rot_out(const std::vector<double>& a, std::size_t k, double& x) { x = a[3 * k + 2]; }
    static void rot_out(const std::vector<double>& a, std::size_t k, fixed<3>& x) {
        for (int d = 0; d < 3; ++d) x[d] = a[3 * k + d];
    }
    std::vector<int> m_shape;
    std::vector<type::position_t<dim>> m_pos;
    std::vector<type::quaternion_t> m_q;
    std::vector<std::array<double, 3>> m_size;
    std::vector<std::array<double, 2>> m_sq;
    std::vector<type::velocity_t<dim>> m_v, m_vd;
    std::vector<type::rotation_t<dim>> m_w, m_dw;
    std::vector<type::force_t<dim>> m_f;
    std::vector<double> m_m;
    std::vector<type::moment_t<dim>> m_j;
    std::size_t m_nb_inactive = 0;
};

// BoxDomain (include/scopi/box.hpp): corners and periodic axes.
template <std::size_t dim>
class BoxDomain {
  public:
    BoxDomain() = default;
    BoxDomain(const std::array<double, dim>& lo, const std::array<double, dim>& hi) : m_lo(lo), m_hi(hi) {}
    BoxDomain& with_periodicity() {
        m_periodic.fill(true);
        return *this;
    }
    BoxDomain& with_periodicity(std::size_t axis) {
        m_periodic[axis] = true;
        return *this;
    }
    BoxDomain& with_periodicity(const std::initializer_list<std::size_t>& axes) {
        for (auto a : axes) m_periodic[a] = true;
        return *this;
    }
    const std::array<bool, dim>& is_periodic() const { return m_periodic; }
    bool is_periodic(std::size_t axis) const { return m_periodic[axis]; }
    const std::array<double, dim>& min_corner() const { return m_lo; }
    const std::array<double, dim>& max_corner() const { return m_hi; }
    double lower_bound(std::size_t axis) const { return m_lo[axis]; }
    double upper_bound(std::size_t axis) const { return m_hi[axis]; }

  private:
    std::array<double, dim> m_lo{}, m_hi{};
    std::array<bool, dim> m_periodic{};
};

// ------------------------------------------------------------------ output files
namespace detail {
// rotation_matrix<dim>(q), row-major (include/scopi/quaternion.hpp:65-124), evaluated as the
// reference's build that wrote test/references/*.json evaluates it: with its products
// contracted into fused multiply-adds (a - b * c -> fma(-b, c, a), a * b + c * d ->
// fma(c, d, a * b)); its 2D diagonal 1 - 2 w w is fma(-2 w, w, 1) in those files.
inline std::array<double, 4> rotation_matrix2(const type::quaternion_t& q) {
    const double x = q(0), w = q(3);
    const double d = std::fma(-(2 * w), w, 1.0);
    return {{d, -2 * x * w, 2 * x * w, d}};
}
inline std::array<double, 9> rotation_matrix3(const type::quaternion_t& q) {
    const double x = q(0), y = q(1), z = q(2), w = q(3);
    auto diag = [](double a, double b) { return std::fma(-(2 * b), b, std::fma(-(2 * a), a, 1.0)); };
    auto sub = [](double a, double b, double c, double d) { return std::fma(2 * a, b, -((2 * c) * d)); };
    auto add = [](double a, double b, double c, double d) { return std::fma(2 * c, d, (2 * a) * b); };
    return {{diag(z, w), sub(y, z, x, w), add(y, w, x, z), add(y, z, x, w), diag(y, w), sub(z, w, x, y),
             sub(y, w, x, z), add(z, w, x, y), diag(y, z)}};
}
template <std::size_t dim>
output::value vec(const fixed<dim>& v) {
    output::value a = output::value::array();
    for (std::size_t d = 0; d < dim; ++d) a.push_back(v[d]);
    return a;
}
inline output::value rot_value(double x) { return x; }
inline output::value rot_value(const fixed<3>& x) { return vec<3>(x); }

template <class problem_t>
output::value property_json(const contact_property<problem_t>&) {  // contact/property.hpp:44-49
    output::value p;
    p["type"] = "no friction";
    return p;
}
inline output::value property_json(const contact_property<Friction>& c) {  // contact/property.hpp:64-70
    output::value p;
    p["type"] = "friction";
    p["mu"] = c.mu;
    return p;
}
// The reference stops the outer loop when its count (1, 2, ...) reaches the double
// fixed_point_max_iter; the device takes the same bound as the smallest such count.
template <class P>
inline void set_fixed_point(scopi_hip_params& prm, const P& c) {
    prm.fixed_point_tol = c.fixed_point_tol;
    const double m = std::ceil(c.fixed_point_max_iter);
    prm.fixed_point_max_iter = !(m > 1.0) ? 1 : m > 1e15 ? (int64_t)1e15 : (int64_t)m;
}
template <class P>
inline void set_gamma(scopi_hip_params& prm, const P& c) {
    prm.gamma = c.gamma;
    prm.gamma_min = c.gamma_min;
    prm.gamma_tol = c.gamma_tol;
}
inline output::value property_json(const contact_property<Viscous>& c) {  // contact/property.hpp:114-123
    output::value p;
    p["type"] = "viscous";
    p["gamma"] = c.gamma;
    p["gamma_min"] = c.gamma_min;
    p["gamma_tol"] = c.gamma_tol;
    return p;
}
inline output::value property_json(const contact_property<ViscousFriction>& c) {  // contact/property.hpp:162-174
    output::value p;
    p["type"] = "viscous friction";
    p["gamma"] = c.gamma;
    p["gamma_min"] = c.gamma_min;
    p["gamma_tol"] = c.gamma_tol;
    p["mu"] = c.mu;
    p["fixed_point_tol"] = c.fixed_point_tol;
    p["fixed_point_max_iter"] = c.fixed_point_max_iter;
    return p;
}
inline output::value property_json(const contact_property<FrictionFixedPoint>& c) {  // contact/property.hpp:140-148
    output::value p;
    p["type"] = "friction fixed point";
    p["mu"] = c.mu;
    p["fixed_point_tol"] = c.fixed_point_tol;
    p["fixed_point_max_iter"] = c.fixed_point_max_iter;
    return p;
}
}  // namespace detail

// The document ScopiSolver::write_output_files writes (include/scopi/solver.hpp:369-421):
// every object as write_objects (objects/methods/write_objects.hpp:45-160) with its
// properties, then every contact as neighbor::to_json (objects/neighbor.hpp:35-46), with
// "lagrange multiplier" = lambda[ic] when `lambda` is given (the reference indexes its flat
// multiplier vector by the contact's position, also for Friction's three rows per contact).
template <std::size_t dim, class problem_t>
output::value output_document(const scopi_container<dim>& p, const std::vector<neighbor<dim, problem_t>>& contacts,
                              const std::vector<double>* lambda = nullptr) {
    output::value doc;
    doc["objects"] = output::value();
    for (std::size_t k = 0; k < p.size(); ++k) {
        output::value o;
        const type::quaternion_t& q = p.q()[k];
        output::value rot = output::value::array();
        fixed<dim> col0;  // first column of the rotation: plane / segment normal
        if constexpr (dim == 2) {
            const auto R = detail::rotation_matrix2(q);
            for (double x : R) rot.push_back(x);
            col0[0] = R[0], col0[1] = R[2];
        } else {
            const auto R = detail::rotation_matrix3(q);
            for (double x : R) rot.push_back(x);
            col0[0] = R[0], col0[1] = R[3], col0[2] = R[6];
        }
        const auto& sz = p.object_size(k);
        switch (p.shape(k)) {
            case SCOPI_SHAPE_SPHERE:
                o["type"] = "sphere";
                o["id"] = k;
                o["position"] = detail::vec<dim>(p.pos()[k]);
                o["radius"] = sz[0];
                o["rotation"] = rot;
                o["quaternion"] = detail::vec<4>(q);
                break;
            case SCOPI_SHAPE_SUPERELLIPSOID: {
                o["type"] = "superellipsoid";
                o["id"] = k;
                o["position"] = detail::vec<dim>(p.pos()[k]);
                fixed<dim> ax;
                for (std::size_t d = 0; d < dim; ++d) ax[d] = sz[d];
                o["radius"] = detail::vec<dim>(ax);
                output::value e = output::value::array();
                for (std::size_t d = 0; d + 1 < dim; ++d) e.push_back(p.object_squareness(k)[d]);
                o["squareness"] = e;
                o["rotation"] = rot;
                o["quaternion"] = detail::vec<4>(q);
                break;
            }
            case SCOPI_SHAPE_PLANE:
                o["type"] = "plane";
                o["id"] = k;
                o["position"] = detail::vec<dim>(p.pos()[k]);
                o["normal"] = detail::vec<dim>(col0);
                o["rotation"] = rot;
                o["quaternion"] = detail::vec<4>(q);
                break;
            default: {  // segment (2D): extrema R (0, -+L/2) + x, tangent (-n1, n0)
                o["type"] = "segment";
                o["id"] = k;
                fixed<dim> p1, p2, tangent;
                const double L = sz[0];
                if constexpr (dim == 2) {
                    const auto R = detail::rotation_matrix2(q);
                    for (int d = 0; d < 2; ++d) {
                        p1[d] = (R[2 * d] * 0 + R[2 * d + 1] * (-0.5 * L)) + p.pos()[k][d];
                        p2[d] = (R[2 * d] * 0 + R[2 * d + 1] * (0.5 * L)) + p.pos()[k][d];
                    }
                    tangent[0] = -col0[1], tangent[1] = col0[0];
                }
                o["p1"] = detail::vec<dim>(p1);
                o["p2"] = detail::vec<dim>(p2);
                o["normal"] = detail::vec<dim>(col0);
                o["tangent"] = detail::vec<dim>(tangent);
                o["quaternion"] = detail::vec<4>(q);
                break;
            }
        }
        output::value& prop = o["properties"];
        prop["velocity"] = detail::vec<dim>(p.v()[k]);
        prop["desired_velocity"] = detail::vec<dim>(p.vd()[k]);
        prop["omega"] = detail::rot_value(p.omega()[k]);
        prop["desired_omega"] = detail::rot_value(p.desired_omega()[k]);
        prop["force"] = detail::vec<dim>(p.f()[k]);
        prop["mass"] = p.m()[k];
        prop["moment_inertia"] = detail::rot_value(p.j()[k]);
        prop["active"] = k >= p.nb_inactive();
        doc["objects"].push_back(o);
    }
    doc["contacts"] = output::value();
    std::size_t ic = 0;
    for (const auto& c : contacts) {
        output::value cj;
        cj["i"] = c.i;
        cj["j"] = c.j;
        cj["pi"] = detail::vec<dim>(c.pi);
        cj["pj"] = detail::vec<dim>(c.pj);
        cj["normal"] = detail::vec<dim>(c.nij);
        cj["distance"] = c.dij;
        cj["property"] = detail::property_json(c.property);
        if (lambda) cj["lagrange multiplier"] = ic < lambda->size() ? (*lambda)[ic] : 0.0;
        ++ic;
        doc["contacts"].push_back(cj);
    }
    return doc;
}

// {path}/{filename}_{nite:04d}.json (nlohmann's dump with an indent of 4) or .bson
inline std::string write_output_document(const output::value& doc, const std::filesystem::path& path,
                                         const std::string& filename, std::size_t nite, bool binary) {
    if (!std::filesystem::exists(path)) std::filesystem::create_directories(path);
    char tail[32];
    std::snprintf(tail, sizeof(tail), "_%04zu.%s", nite, binary ? "bson" : "json");
    const std::string name = (path / filename).string() + tail;
    std::ofstream file(name, std::ios::out | std::ios::binary);
    if (binary) {
        const auto b = output::to_bson(doc);
        file.write(reinterpret_cast<const char*>(b.data()), (std::streamsize)b.size());
    } else {
        file << output::dump(doc, 4);
    }
    return name;
}

// ------------------------------------------------------------------ device context
class hip_error : public std::runtime_error {
  public:
    using std::runtime_error::runtime_error;
};

namespace detail {
inline void check(int rc, scopi_hip_ctx* ctx, const char* what) {
    if (rc < 0) throw hip_error(std::string(what) + ": " + (ctx ? scopi_hip_last_error(ctx) : "error"));
}

// One device context per slot object; a scene is uploaded when its size or the
// slot's view of it is stale.
struct device_scene {
    scopi_hip_ctx* ctx = nullptr;
    int dim = 0;
    ~device_scene() {
        if (ctx) scopi_hip_destroy(ctx);
    }
    template <std::size_t dim_>
    void upload(const scopi_container<dim_>& p) {
        if (!ctx) {
            check(scopi_hip_create((int)dim_, (int64_t)p.size(), 0, &ctx), nullptr, "scopi_hip_create");
            dim = (int)dim_;
        }
        auto s = p.pack();
        check(scopi_hip_upload_state(ctx, (int64_t)p.size(), (int64_t)p.nb_inactive(), s.shape.data(), s.pos.data(),
                                     s.q.data(), s.size.data(), s.sq.data(), s.v.data(), s.vd.data(), s.w.data(),
                                     s.dw.data(), s.f.data(), s.m.data(), s.J.data()),
              ctx, "scopi_hip_upload_state");
    }
};
}  // namespace detail

// ------------------------------------------------------------------ contact slot
template <class D>
struct ContactsParams {
    double dmax{2.};
    double kd_tree_radius{17.};
    void init_options() {}
};

template <class problem_t, int METHOD>
class contact_hip {
  public:
    using params_t = ContactsParams<contact_hip>;
    static constexpr int method = METHOD;
    explicit contact_hip(const params_t& params = params_t()) : m_params(params) {}
    void init_options() {}
    params_t& get_params() { return m_params; }
    contact_property<problem_t>& default_contact_property() { return m_default; }

    // contact_kdtree::run_impl (contact_kdtree.hpp:200-268): sorted (i, j), i < j.
    template <std::size_t dim>
    std::vector<neighbor<dim, problem_t>> run(const BoxDomain<dim>&, scopi_container<dim>& particles) {
        m_dev.upload(particles);
        scopi_hip_params prm;
        scopi_hip_default_params(&prm);
        prm.contact_method = METHOD;
        prm.dmax = m_params.dmax;
        prm.kd_tree_radius = m_params.kd_tree_radius;
        int64_t nc = 0;
        detail::check(scopi_hip_compute_contacts(m_dev.ctx, &prm, &nc), m_dev.ctx, "scopi_hip_compute_contacts");
        std::vector<int64_t> i(nc), j(nc);
        std::vector<double> dij(nc), nij(3 * nc), pi(3 * nc), pj(3 * nc);
        detail::check(scopi_hip_download_contacts(m_dev.ctx, i.data(), j.data(), dij.data(), nij.data(), pi.data(),
                                                  pj.data()),
                      m_dev.ctx, "scopi_hip_download_contacts");
        std::vector<neighbor<dim, problem_t>> out(nc);
        for (int64_t c = 0; c < nc; ++c) {
            out[c].i = (std::size_t)i[c];
            out[c].j = (std::size_t)j[c];
            out[c].dij = dij[c];
            for (std::size_t d = 0; d < dim; ++d) {
                out[c].nij[d] = nij[3 * c + d];
                out[c].pi[d] = pi[3 * c + d];
                out[c].pj[d] = pj[3 * c + d];
            }
            out[c].property = m_default;
        }
        return out;
    }
    template <std::size_t dim>
    std::vector<neighbor<dim, problem_t>> run(scopi_container<dim>& particles) {
        return run(BoxDomain<dim>(), particles);
    }

  private:
    params_t m_params;
    contact_property<problem_t> m_default;
    detail::device_scene m_dev;
};

template <class problem_t>
using contact_kdtree = contact_hip<problem_t, SCOPI_CONTACT_KDTREE>;
template <class problem_t>
using contact_brute_force = contact_hip<problem_t, SCOPI_CONTACT_BRUTE_FORCE>;

// ------------------------------------------------------------------ optim slot
struct apgd_params {
    double alpha = 0.05;
    std::size_t max_ite = 10000;
    double tolerance = 1e-7;
    bool dynamic_descent = true;
    int precision = SCOPI_PRECISION_FP64;  // SCOPI_PRECISION_FP32_REFINE: FP32 phase + FP64 refinement
    void init_options() {}
};
struct pgd_params {
    double alpha = 0.05;
    std::size_t max_ite = 10000;
    double tolerance = 1e-6;
    int precision = SCOPI_PRECISION_FP64;
    void init_options() {}
};
struct apgd {
    using params_t = apgd_params;
    static constexpr int method = SCOPI_METHOD_APGD;
};
struct pgd {
    using params_t = pgd_params;
    static constexpr int method = SCOPI_METHOD_PGD;
};

template <class problem_t>
constexpr int problem_tag() {
    return std::is_same<problem_t, Friction>::value             ? SCOPI_PROBLEM_FRICTION
           : std::is_same<problem_t, FrictionFixedPoint>::value ? SCOPI_PROBLEM_FRICTION_FIXED_POINT
           : std::is_same<problem_t, Viscous>::value            ? SCOPI_PROBLEM_VISCOUS
           : std::is_same<problem_t, ViscousFriction>::value    ? SCOPI_PROBLEM_VISCOUS_FRICTION
                                                                 : SCOPI_PROBLEM_NOFRICTION;
}

template <class Method>
class OptimGradient {
  public:
    using method_t = Method;
    using params_t = typename Method::params_t;
    void set_timestep(double dt) { m_dt = dt; }
    void init_options() {}
    params_t& get_params() { return m_params; }
    template <class C>
    void extra_steps_before_solve(C&) { m_should_solve = true; }
    bool should_solve() const { return m_should_solve; }
    template <class C, std::size_t dim>
    void extra_steps_after_solve(C&, const scopi_container<dim>&) { m_should_solve = false; }
    // update_contact_properties (OptimGradient.hpp:19-54, 309-315): the solve already updated the
    // gammas on the device; they come back into the contacts' properties
    template <class C>
    void update_contact_properties(C& contacts) {
        using problem_t = typename C::value_type::problem_t;
        if constexpr (has_gamma<problem_t>()) {
            if (contacts.empty()) return;
            std::vector<double> g(contacts.size());
            detail::check(scopi_hip_download_contact_gamma(m_dev.ctx, g.data()), m_dev.ctx,
                          "scopi_hip_download_contact_gamma");
            for (std::size_t c = 0; c < contacts.size(); ++c) contacts[c].property.gamma = g[c];
        }
    }

    scopi_hip_params hip_params(int problem, double mu) const {
        scopi_hip_params prm;
        scopi_hip_default_params(&prm);
        prm.problem = problem;
        prm.method = Method::method;
        prm.alpha = m_params.alpha;
        prm.max_ite = (int64_t)m_params.max_ite;
        prm.tolerance = m_params.tolerance;
        prm.precision = m_params.precision;
        prm.mu = mu;
        if constexpr (std::is_same<Method, apgd>::value) prm.dynamic_descent = m_params.dynamic_descent ? 1 : 0;
        return prm;
    }

    // OptimGradient::run (OptimGradient.hpp:252-306) on the GPU.
    template <std::size_t dim, class problem_t>
    void run(const scopi_container<dim>& particles, const std::vector<neighbor<dim, problem_t>>& contacts, std::size_t) {
        m_dev.upload(particles);
        const int64_t nc = (int64_t)contacts.size();
        std::vector<int64_t> i(nc), j(nc);
        std::vector<double> dij(nc), nij(3 * nc, 0.0), pi(3 * nc, 0.0), pj(3 * nc, 0.0), mu(nc, 0.1);
        for (int64_t c = 0; c < nc; ++c) {
            i[c] = (int64_t)contacts[c].i;
            j[c] = (int64_t)contacts[c].j;
            dij[c] = contacts[c].dij;
            for (std::size_t d = 0; d < dim; ++d) {
                nij[3 * c + d] = contacts[c].nij[d];
                pi[3 * c + d] = contacts[c].pi[d];
                pj[3 * c + d] = contacts[c].pj[d];
            }
            if constexpr (has_mu<problem_t>()) mu[c] = contacts[c].property.mu;
        }
        detail::check(scopi_hip_upload_contacts(m_dev.ctx, nc, i.data(), j.data(), dij.data(), nij.data(), pi.data(),
                                                pj.data(), mu.data()),
                      m_dev.ctx, "scopi_hip_upload_contacts");
        auto prm = hip_params(problem_tag<problem_t>(), 0.1);
        // the outer loop reads its controls from the first contact (OptimGradient.hpp:145-146)
        if constexpr (has_fixed_point<problem_t>())
            if (nc > 0) detail::set_fixed_point(prm, contacts[0].property);
        if constexpr (has_gamma<problem_t>()) {
            // every contact's own gamma; gamma_min and gamma_tol of the first one (the defaults)
            if (nc > 0) detail::set_gamma(prm, contacts[0].property);
            std::vector<double> g(nc);
            for (int64_t c = 0; c < nc; ++c) g[c] = contacts[c].property.gamma;
            detail::check(scopi_hip_set_contact_gamma(m_dev.ctx, g.data()), m_dev.ctx, "scopi_hip_set_contact_gamma");
        }
        detail::check(scopi_hip_solve(m_dev.ctx, &prm, m_dt, &m_info), m_dev.ctx, "scopi_hip_solve");
        const std::size_t na = particles.nb_active();
        m_u.assign(3 * na, 0.0);
        m_omega.assign(3 * na, 0.0);
        m_lambda.assign(nc > 0 ? (std::size_t)m_info.rows : 0, 0.0);
        detail::check(scopi_hip_download_solution(m_dev.ctx, m_u.data(), m_omega.data(), m_lambda.data()), m_dev.ctx,
                      "scopi_hip_download_solution");
    }
    // get_uadapt / get_wadapt: Na x 3, row-major (index 3 * i + d)
    const std::vector<double>& get_uadapt() const { return m_u; }
    const std::vector<double>& get_wadapt() const { return m_omega; }
    const std::vector<double>& lagrange_multiplier() const { return m_lambda; }
    const scopi_hip_solve_info& last_info() const { return m_info; }

  private:
    double m_dt = 0;
    bool m_should_solve = false;
    params_t m_params;
    std::vector<double> m_u, m_omega, m_lambda;
    scopi_hip_solve_info m_info{};
    detail::device_scene m_dev;
};

// ------------------------------------------------------------------ vap slot
struct vap_fixed {
    static constexpr int kind = SCOPI_VAP_FIXED;
    struct params_t {};
    params_t& get_params() { return m_params; }
    params_t m_params;
    // vap/vap_fixed.hpp:46-55
    template <std::size_t dim, class C>
    void set_a_priori_velocity(double, scopi_container<dim>& p, const C&) {
        for (std::size_t i = p.nb_inactive(); i < p.size(); ++i) {
            p.v()[i] = p.vd()[i];
            p.omega()[i] = p.desired_omega()[i];
        }
    }
};
struct vap_fpd {
    static constexpr int kind = SCOPI_VAP_FPD;
    struct params_t {};
    params_t& get_params() { return m_params; }
    params_t m_params;
    // vap/vap_fpd.hpp:70-82, src/vap/vap_fpd.cpp:6-26
    template <std::size_t dim, class C>
    void set_a_priori_velocity(double dt, scopi_container<dim>& p, const C&) {
        for (std::size_t i = p.nb_inactive(); i < p.size(); ++i) {
            for (std::size_t d = 0; d < dim; ++d) p.v()[i][d] += dt * p.f()[i][d] / p.m()[i];
            if constexpr (dim == 3) {
                auto& w = p.omega()[i];
                const auto& J = p.j()[i];
                const double r0 = w[1] * w[2] * (J[2] - J[1]), r1 = w[0] * w[2] * (J[0] - J[2]),
                             r2 = w[0] * w[1] * (J[1] - J[0]);
                w[0] += r0, w[1] += r1, w[2] += r2;
            }
        }
    }
};

// vap_projection (vap/vap_projection.hpp:34-49, src/vap/vap_projection.cpp:6-10): the desired
// velocity of every active particle takes u(i, .), its rotation velocity w(i, 2) (in 3D the scalar
// goes to all three components, as the reference assigns a scalar to a 3-vector).  u, w: Na x 3,
// given by set_u_w; ScopiSolver::vap() reaches the slot object.
struct vap_projection {
    static constexpr int kind = SCOPI_VAP_PROJECTION;
    struct params_t {};
    params_t& get_params() { return m_params; }
    params_t m_params;
    template <class U, class W>
    void set_u_w(const U& u, const W& w) {  // u(i, d), w(i, d) indexable, Na rows
        m_u.clear();
        m_w.clear();
        for (std::size_t i = 0; i < u.shape()[0]; ++i)
            for (std::size_t d = 0; d < 3; ++d) {
                m_u.push_back(u(i, d));
                m_w.push_back(w(i, d));
            }
    }
    void set_u_w(const std::vector<double>& u, const std::vector<double>& w) {  // flat Na x 3
        m_u = u;
        m_w = w;
    }
    template <std::size_t dim, class C>
    void set_a_priori_velocity(double, scopi_container<dim>& p, const C&) {
        for (std::size_t i = p.nb_inactive(); i < p.size(); ++i) {
            const std::size_t a = i - p.nb_inactive();
            for (std::size_t d = 0; d < dim; ++d) p.vd()[i][d] = m_u[3 * a + d];
            if constexpr (dim == 2)
                p.omega()[i] = m_w[3 * a + 2];
            else
                for (int d = 0; d < 3; ++d) p.omega()[i][d] = m_w[3 * a + 2];
        }
    }
    std::vector<double> m_u, m_w;
};

// ------------------------------------------------------------------ driver
// ScopiParams (include/scopi/params.hpp:89-130, src/params.cpp:6-14): output every
// `output_frequency` steps into {path}/{filename}_{step:04d}.json (or .bson).
struct ScopiParams {
    std::size_t output_frequency = 1;
    std::filesystem::path path = std::filesystem::current_path() / "Results";
    std::string filename = "scopi_objects";
    bool write_velocity = false;
    bool write_lagrange_multiplier = false;
    bool binary_output = false;
    void init_options() {}
};

template <class Solver>
struct Params {
    ScopiParams& solver_params;
    typename Solver::optim_solver_t::params_t& optim_params;
    contact_property<typename Solver::problem_t>& default_contact_property;
    typename Solver::contact_method_t::params_t& contact_method_params;
};

template <std::size_t dim, class problem_type = NoFriction, class optim_solver_type = OptimGradient<apgd>,
          template <class> class contact_method_type = contact_kdtree, class vap_type = vap_fixed>
class ScopiSolver {
  public:
    using problem_t = problem_type;
    using optim_solver_t = optim_solver_type;
    using contact_method_t = contact_method_type<problem_t>;
    using vap_t = vap_type;
    using params_t = Params<ScopiSolver>;
    using contact_container_t = std::vector<neighbor<dim, problem_t>>;

    explicit ScopiSolver(scopi_container<dim>& particles) : m_particles(particles) {}
    ScopiSolver(const BoxDomain<dim>& box, scopi_container<dim>& particles) : m_box(box), m_particles(particles) {}
    ~ScopiSolver() {
        if (m_pending.valid()) m_pending.wait();
        if (m_ctx) scopi_hip_destroy(m_ctx);
    }

    params_t get_params() {
        return params_t{m_params, m_optim.get_params(), m_contact.default_contact_property(), m_contact.get_params()};
    }

    // ScopiSolver::run (solver.hpp:266-303): the output file of the initial state, then the
    // steps on the device in runs between two output steps; at each output step the state,
    // the last step's contacts (and multipliers) come back to the host and the file is
    // formatted and written by a background task while the device goes on.
    void run(double dt, std::size_t total_it, std::size_t initial_iter = 0) {
        m_optim.set_timestep(dt);
        upload();
        write_output_files(m_old_contacts, initial_iter);
        auto prm = hip_params();
        const std::size_t freq = m_params.output_frequency;
        const bool periodic = freq != std::size_t(-1) && freq > 0;
        std::size_t nite = initial_iter;
        scopi_hip_solve_info info{};
        while (nite < total_it) {
            std::size_t stop = total_it;
            if (periodic) stop = std::min(total_it, (nite / freq + 1) * freq);
            detail::check(scopi_hip_run(m_ctx, &prm, dt, (int64_t)stop, (int64_t)nite, &info), m_ctx,
                          "scopi_hip_run");
            accumulate(info);
            nite = stop;
            if (periodic && nite % freq == 0) {
                download_state();
                download_contacts(info, m_old_contacts, m_params.write_lagrange_multiplier ? &m_lambda : nullptr);
                write_output_files(m_old_contacts, nite);
            }
        }
        download_state();
        if (total_it > initial_iter) download_contacts(info, m_old_contacts, nullptr);
        wait_output();
    }

    const scopi_hip_solve_info& last_info() const { return m_info; }
    const contact_container_t& last_contacts() const { return m_old_contacts; }
    vap_t& vap() { return m_vap; }

  private:
    // Every slot is a GPU slot of this library: the whole loop stays on the device.
    void upload() {
        if (!m_ctx)
            detail::check(scopi_hip_create((int)dim, (int64_t)m_particles.size(), 0, &m_ctx), nullptr,
                          "scopi_hip_create");
        auto s = m_particles.pack();
        detail::check(scopi_hip_upload_state(m_ctx, (int64_t)m_particles.size(), (int64_t)m_particles.nb_inactive(),
                                             s.shape.data(), s.pos.data(), s.q.data(), s.size.data(), s.sq.data(),
                                             s.v.data(), s.vd.data(), s.w.data(), s.dw.data(), s.f.data(), s.m.data(),
                                             s.J.data()),
                      m_ctx, "scopi_hip_upload_state");
        double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        int32_t per[3] = {0, 0, 0};
        for (std::size_t d = 0; d < dim; ++d) {
            lo[d] = m_box.lower_bound(d);
            hi[d] = m_box.upper_bound(d);
            per[d] = m_box.is_periodic(d) ? 1 : 0;
        }
        detail::check(scopi_hip_set_box(m_ctx, lo, hi, per), m_ctx, "scopi_hip_set_box");
        if constexpr (std::is_same<vap_t, vap_projection>::value) {
            const std::size_t na = m_particles.nb_active();
            if (m_vap.m_u.size() != 3 * na || m_vap.m_w.size() != 3 * na)
                throw std::runtime_error("vap_projection: set_u_w needs Na x 3 velocities");
            detail::check(scopi_hip_set_projection_velocity(m_ctx, m_vap.m_u.data(), m_vap.m_w.data()), m_ctx,
                          "scopi_hip_set_projection_velocity");
        }
        m_info = scopi_hip_solve_info{};
        m_info.converged = 1;
    }
    double mu() {
        double mu = 0.1;
        if constexpr (has_mu<problem_t>()) mu = m_contact.default_contact_property().mu;
        return mu;
    }
    scopi_hip_params hip_params() {
        auto prm = m_optim.hip_params(problem_tag<problem_t>(), mu());
        if constexpr (has_fixed_point<problem_t>()) detail::set_fixed_point(prm, m_contact.default_contact_property());
        if constexpr (has_gamma<problem_t>()) detail::set_gamma(prm, m_contact.default_contact_property());
        prm.vap = vap_t::kind;
        prm.contact_method = contact_method_t::method;
        prm.dmax = m_contact.get_params().dmax;
        prm.kd_tree_radius = m_contact.get_params().kd_tree_radius;
        return prm;
    }
    void accumulate(const scopi_hip_solve_info& one) {
        m_info.n_contacts = one.n_contacts;
        m_info.rows = one.rows;
        m_info.iterations += one.iterations;
        m_info.iterations_fp32 += one.iterations_fp32;
        m_info.backtracks += one.backtracks;
        m_info.outer_iterations += one.outer_iterations;
        m_info.fixed_point_error = one.fixed_point_error;
        m_info.q_applies += one.q_applies;
        m_info.converged = m_info.converged && one.converged;
        if (one.status != 0) m_info.status = one.status;
        m_info.residual = one.residual;
    }
    void download_state() {
        const std::size_t n = m_particles.size();
        std::vector<double> pos(3 * n), q(4 * n), v(3 * n), w(3 * n);
        detail::check(scopi_hip_download_state(m_ctx, pos.data(), q.data(), v.data(), w.data()), m_ctx,
                      "scopi_hip_download_state");
        m_particles.unpack(pos, q, v, w);
    }
    // the last step's contacts (as compute_contacts returned them) and multipliers
    void download_contacts(const scopi_hip_solve_info& info, contact_container_t& out, std::vector<double>* lambda) {
        const std::size_t nc = (std::size_t)info.n_contacts;
        out.assign(nc, typename contact_container_t::value_type{});
        if (nc == 0) {
            if (lambda) lambda->clear();
            return;
        }
        std::vector<int64_t> ci(nc), cj(nc);
        std::vector<double> dij(nc), nij(3 * nc), pi(3 * nc), pj(3 * nc);
        detail::check(scopi_hip_download_contacts(m_ctx, ci.data(), cj.data(), dij.data(), nij.data(), pi.data(),
                                                  pj.data()),
                      m_ctx, "scopi_hip_download_contacts");
        for (std::size_t c = 0; c < nc; ++c) {
            auto& nb = out[c];
            nb.i = (std::size_t)ci[c];
            nb.j = (std::size_t)cj[c];
            nb.dij = dij[c];
            for (std::size_t d = 0; d < dim; ++d) {
                nb.nij[d] = nij[3 * c + d];
                nb.pi[d] = pi[3 * c + d];
                nb.pj[d] = pj[3 * c + d];
            }
            if constexpr (has_mu<problem_t>() || has_gamma<problem_t>()) nb.property = m_contact.default_contact_property();
        }
        if constexpr (has_gamma<problem_t>()) {
            std::vector<double> g(nc);
            detail::check(scopi_hip_download_contact_gamma(m_ctx, g.data()), m_ctx, "scopi_hip_download_contact_gamma");
            for (std::size_t c = 0; c < nc; ++c) out[c].property.gamma = g[c];
        }
        if (lambda) {
            lambda->assign((std::size_t)info.rows, 0.0);
            detail::check(scopi_hip_download_solution(m_ctx, nullptr, nullptr, lambda->data()), m_ctx,
                          "scopi_hip_download_solution");
        }
    }
    // write_output_files (solver.hpp:369-438): the document is built here, formatting and
    // writing run in a background task (one at a time)
    void write_output_files(const contact_container_t& contacts, std::size_t nite) {
        auto doc = output_document(m_particles, contacts,
                                   m_params.write_lagrange_multiplier && !contacts.empty() ? &m_lambda : nullptr);
        wait_output();
        m_pending = std::async(std::launch::async, [doc = std::move(doc), path = m_params.path,
                                                    name = m_params.filename, nite, bin = m_params.binary_output]() {
            return write_output_document(doc, path, name, nite, bin);
        });
    }
    void wait_output() {
        if (m_pending.valid()) m_pending.get();
    }

    ScopiParams m_params;
    BoxDomain<dim> m_box;
    scopi_container<dim>& m_particles;
    optim_solver_t m_optim;
    contact_method_t m_contact;
    vap_t m_vap;
    scopi_hip_ctx* m_ctx = nullptr;
    scopi_hip_solve_info m_info{};
    contact_container_t m_old_contacts;
    std::vector<double> m_lambda;
    std::future<std::string> m_pending;
};

inline void initialize(const std::string& = "") {}

}  // namespace scopi
