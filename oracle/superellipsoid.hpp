// superellipsoid.hpp -- FP64 restatement of the superellipsoid closest points of
// SCoPI (include/scopi/objects/methods/closest_points.hpp:263-733) for the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY (included by scopi_oracle.cpp).
//
//   geometry        superellipsoid::point / normal (objects/types/superellipsoid.hpp:281-325),
//                   binit_xy / ainit_yz (superellipsoid.hpp:359-466)
//   SE3-SE3         576 x 576 penalised initial grid (closest_points.hpp:604-678), residual
//                   F(a1, b1, a2, b2) = [(p2 - p1) x n1 ; n1.n2 + |n1||n2|] (:284-382),
//                   hybrd1 (minpack.hpp:851-923: mode 2, unit diagonal, factor 100,
//                   xtol 1e-10, maxfev 200 (n + 1)), result (:724-728)
//   Powell hybrid   MINPACK hybrd / enorm / qrfac / qform / dogleg / r1updt / r1mpyq as the
//                   reference's C++ translation runs them (src/minpack.cpp, minpack.hpp:337-923):
//                   enorm a plain sqrt of the sum of squares, r1updt's giant = r8_huge() = 1e30,
//                   the first termination test that holds sets info, and dogleg's scaled-gradient
//                   loop with its index shift and its left-to-right (delta/qnorm)(sgnorm/delta)^2
//                   (Minpack::dogleg).  A test-only switch
//                   (minpack_variant(), a bit mask) selects forms of the published algorithm (More,
//                   Garbow, Hillstrom, ANL-80-74, 1980) instead, which an independent MINPACK
//                   (scipy's hybrj) reproduces.
//   Jacobian        the reference's hand-written derivative (closest_points.hpp:385-603; 2D
//                   :127-189), term by term in its evaluation order; forward-mode
//                   differentiation of F is kept as a test-side cross-check (jacobian_ad() = 1).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

#include "se_math.hpp"

namespace se_oracle {

using v3 = std::array<double, 3>;
using m33 = std::array<std::array<double, 3>, 3>;

// A superellipsoid: centre, semi-axes, squareness (e = squareness(0), n = squareness(1)), rotation.
struct SE {
    v3 c, r;
    double e, n;
    m33 R;
};

inline double sgn(double x) { return (x > 0) ? 1.0 : ((x < 0) ? -1.0 : 0.0); }

// Elementary-function policies: the C library (the reference's arithmetic) or the
// portable functions of se_math.hpp (bit-reproducible on the device).
struct LibMath {
    static double cos(double x) { return std::cos(x); }
    static double sin(double x) { return std::sin(x); }
    static double asin(double x) { return std::asin(x); }
    static double pow(double x, double p) { return std::pow(x, p); }
};
struct PortableMath {
    static double cos(double x) { return se_math::cos(x); }
    static double sin(double x) { return se_math::sin(x); }
    static double asin(double x) { return se_math::asin(x); }
    static double pow(double x, double p) { return se_math::pow(x, p); }
};

inline v3 rot(const m33& R, const v3& x) {
    return {R[0][0] * x[0] + R[0][1] * x[1] + R[0][2] * x[2], R[1][0] * x[0] + R[1][1] * x[1] + R[1][2] * x[2],
            R[2][0] * x[0] + R[2][1] * x[1] + R[2][2] * x[2]};
}

// superellipsoid::point(a, b): superellipsoid.hpp:292-300
template <class M>
inline v3 se_point(const SE& s, double a, double b) {
    v3 pt;
    pt[0] = s.r[0] * sgn(M::cos(a)) * M::pow(std::abs(M::cos(a)), s.n) * sgn(M::cos(b)) *
            M::pow(std::abs(M::cos(b)), s.e);
    pt[1] = s.r[1] * sgn(M::cos(a)) * M::pow(std::abs(M::cos(a)), s.n) * sgn(M::sin(b)) *
            M::pow(std::abs(M::sin(b)), s.e);
    pt[2] = s.r[2] * sgn(M::sin(a)) * M::pow(std::abs(M::sin(a)), s.n);
    v3 w = rot(s.R, pt);
    return {w[0] + s.c[0], w[1] + s.c[1], w[2] + s.c[2]};
}

// superellipsoid::normal(a, b): superellipsoid.hpp:314-325 (unit outward normal)
template <class M>
inline v3 se_normal(const SE& s, double a, double b) {
    v3 nn;
    nn[0] = s.r[1] * s.r[2] * sgn(M::cos(a)) * M::pow(std::abs(M::cos(a)), 2 - s.n) * sgn(M::cos(b)) *
            M::pow(std::abs(M::cos(b)), 2 - s.e);
    nn[1] = s.r[0] * s.r[2] * sgn(M::cos(a)) * M::pow(std::abs(M::cos(a)), 2 - s.n) * sgn(M::sin(b)) *
            M::pow(std::abs(M::sin(b)), 2 - s.e);
    nn[2] = s.r[0] * s.r[1] * sgn(M::sin(a)) * M::pow(std::abs(M::sin(a)), 2 - s.n);
    v3 w = rot(s.R, nn);
    double l = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    return {w[0] / l, w[1] / l, w[2] / l};
}

// xt::linspace<double>(lo, hi, n, endpoint = true)
inline double linspace(double lo, double hi, int n, int i) { return lo + (hi - lo) / (n - 1) * i; }

// binit_xy(n): superellipsoid.hpp:359-412
template <class M>
inline std::vector<double> binit_xy(const SE& s, int n) {
    const double pi = 4 * std::atan(1);
    std::vector<double> bs;
    for (int i = 0; i < n; ++i) {
        const double b_ell = linspace(0, 0.5 * pi, n, i);
        const double x_ell = s.r[0] * M::cos(b_ell);
        const double y_ell = s.r[1] * M::sin(b_ell);
        const double d = M::pow(M::pow(s.r[1] * x_ell, 2.0 / s.e) + M::pow(s.r[0] * y_ell, 2.0 / s.e), s.e / 2.0);
        const double y_supell = y_ell * s.r[0] * s.r[1] / d;
        const double sinb = std::max(-1.0, std::min(1.0, std::sqrt(M::pow(M::pow(y_supell / s.r[1], 2), 1.0 / s.e))));
        double b = M::asin(sinb);
        if (b > pi) b -= 2 * pi;
        if (b < -pi) b += 2 * pi;
        bs.push_back(b);
        bs.push_back(b + pi / 2);
        bs.push_back(b - pi);
        bs.push_back(b - pi / 2);
    }
    for (auto& b : bs)
        if (std::abs(b) < 1.0e-6) b = 0.01;
    return bs;
}

// ainit_yz(n): superellipsoid.hpp:415-466
template <class M>
inline std::vector<double> ainit_yz(const SE& s, int n) {
    const double pi = 4 * std::atan(1);
    std::vector<double> as;
    for (int i = 0; i < n; ++i) {
        const double a_ell = linspace(0, 0.5 * pi, n, i);
        const double y_ell = s.r[1] * M::cos(a_ell);
        const double z_ell = s.r[2] * M::sin(a_ell);
        const double d = M::pow(M::pow(s.r[2] * y_ell, 2 / s.n) + M::pow(s.r[1] * z_ell, 2 / s.n), s.n / 2);
        const double z_supell = z_ell * s.r[1] * s.r[2] / d;
        const double sina = std::max(-1.0, std::min(1.0, std::sqrt(M::pow(M::pow(z_supell / s.r[2], 2), 1 / s.n))));
        double a = M::asin(sina);
        if (a > pi) a -= 2 * pi;
        if (a < -pi) a += 2 * pi;
        as.push_back(a);
        as.push_back(a + pi / 2);
        as.push_back(a - pi);
        as.push_back(a - pi / 2);
        for (auto& x : as)
            if (std::abs(x) < 1.0e-6) x = 0.01;
    }
    return as;
}

// xt::unique: sorted distinct values
inline std::vector<double> unique_sorted(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    return v;
}

// ------------------------------------------------------------------ forward-mode AD
struct D4 {
    double v;
    double d[4];
};
inline D4 cst(double x) { return D4{x, {0, 0, 0, 0}}; }
inline D4 var(double x, int k) {
    D4 r = cst(x);
    r.d[k] = 1.0;
    return r;
}
inline D4 operator+(const D4& a, const D4& b) {
    D4 r{a.v + b.v, {}};
    for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
inline D4 operator-(const D4& a, const D4& b) {
    D4 r{a.v - b.v, {}};
    for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
inline D4 operator-(const D4& a) {
    D4 r{-a.v, {}};
    for (int k = 0; k < 4; ++k) r.d[k] = -a.d[k];
    return r;
}
inline D4 operator*(const D4& a, const D4& b) {
    D4 r{a.v * b.v, {}};
    for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
inline D4 operator*(double s, const D4& a) {
    D4 r{s * a.v, {}};
    for (int k = 0; k < 4; ++k) r.d[k] = s * a.d[k];
    return r;
}
inline D4 dsqrt(const D4& a) {
    const double s = std::sqrt(a.v);
    D4 r{s, {}};
    for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] / (2 * s);
    return r;
}
template <class M>
inline D4 dcos(const D4& a) {
    D4 r{M::cos(a.v), {}};
    for (int k = 0; k < 4; ++k) r.d[k] = -M::sin(a.v) * a.d[k];
    return r;
}
template <class M>
inline D4 dsin(const D4& a) {
    D4 r{M::sin(a.v), {}};
    for (int k = 0; k < 4; ++k) r.d[k] = M::cos(a.v) * a.d[k];
    return r;
}
// sign(x) |x|^p and its derivative p |x|^(p-1) x'
template <class M>
inline D4 spow(const D4& a, double p) {
    const double ax = std::abs(a.v);
    D4 r{sgn(a.v) * M::pow(ax, p), {}};
    const double dv = p * M::pow(ax, p - 1);
    for (int k = 0; k < 4; ++k) r.d[k] = dv * a.d[k];
    return r;
}
// sign(x)|x|^(2-s) and its derivative (2-s)|x|^(1-s) x', the exponent 1 - s formed as the reference
// writes it (closest_points.hpp:430, 1100)
template <class M>
inline D4 spow2(const D4& a, double s) {
    const double ax = std::abs(a.v);
    D4 r{sgn(a.v) * M::pow(ax, 2 - s), {}};
    const double dv = (2 - s) * M::pow(ax, 1 - s);
    for (int k = 0; k < 4; ++k) r.d[k] = dv * a.d[k];
    return r;
}

// Residual of closest_points.hpp:284-382 with its Jacobian.
struct SEPair {
    SE s1, s2;
};

// Test-only switches of the restatement.  MINPACK forms, a bit mask (0: the reference's
// src/minpack.cpp throughout; each set bit takes one ANL-80-74 form instead): 1 = scaled enorm,
// 2 = giant = DBL_MAX in r1updt, 4 = dogleg's scaled-gradient loop without the index shift and
// (delta/qnorm)(sgnorm/delta)^2 with the square taken first,
// 8 = every termination test runs and the last that holds sets info, delta = factor xnorm
// replaced by factor when it is 0.  Jacobian: 0 = the reference's hand-written one,
// 1 = forward-mode AD.
inline int& minpack_variant() {
    static int v = 0;
    return v;
}
inline int& jacobian_ad() {
    static int v = 0;
    return v;
}
// Number of dogleg calls that took the scaled-gradient branch before hybrd had evaluated a trial
// point: there the reference reads wa4[n-1] of a fresh, uninitialised work array (see dogleg);
// the restatement reads 0.
inline long& dogleg_fresh_wa4_count() {
    static long v = 0;
    return v;
}

template <class M>
inline void se_se_residual_ad(const SEPair& P, const double* u, double* F, double* J /* 4x4 row-major or null */) {
    // Values follow the reference's expression order (left-to-right products and
    // sums, closest_points.hpp:323-381); derivatives ride along by forward-mode AD.
    D4 a1 = var(u[0], 0), b1 = var(u[1], 1), a2 = var(u[2], 2), b2 = var(u[3], 3);
    const SE &s1 = P.s1, &s2 = P.s2;
    D4 ca1 = dcos<M>(a1), sa1 = dsin<M>(a1), cb1 = dcos<M>(b1), sb1 = dsin<M>(b1);
    D4 ca2 = dcos<M>(a2), sa2 = dsin<M>(a2), cb2 = dcos<M>(b2), sb2 = dsin<M>(b2);
    D4 ca1n = spow<M>(ca1, s1.n), ca1n2 = spow2<M>(ca1, s1.n), sa1n = spow<M>(sa1, s1.n), sa1n2 = spow2<M>(sa1, s1.n);
    D4 cb1e = spow<M>(cb1, s1.e), cb1e2 = spow2<M>(cb1, s1.e), sb1e = spow<M>(sb1, s1.e), sb1e2 = spow2<M>(sb1, s1.e);
    D4 ca2n = spow<M>(ca2, s2.n), ca2n2 = spow2<M>(ca2, s2.n), sa2n = spow<M>(sa2, s2.n), sa2n2 = spow2<M>(sa2, s2.n);
    D4 cb2e = spow<M>(cb2, s2.e), cb2e2 = spow2<M>(cb2, s2.e), sb2e = spow<M>(sb2, s2.e), sb2e2 = spow2<M>(sb2, s2.e);
    // body-frame normal (A1..A3) and point (A4..A6) of s1; point (B1..B3) and normal (B4..B6) of s2
    const D4 A1 = ((s1.r[1] * s1.r[2]) * ca1n2) * cb1e2, A2 = ((s1.r[0] * s1.r[2]) * sb1e2) * ca1n2,
             A3 = (s1.r[0] * s1.r[1]) * sa1n2;
    const D4 A4 = (s1.r[0] * ca1n) * cb1e, A5 = (s1.r[1] * sb1e) * ca1n, A6 = s1.r[2] * sa1n;
    const D4 B1 = (s2.r[0] * ca2n) * cb2e, B2 = (s2.r[1] * sb2e) * ca2n, B3 = s2.r[2] * sa2n;
    const D4 B4 = ((s2.r[1] * s2.r[2]) * ca2n2) * cb2e2, B5 = ((s2.r[0] * s2.r[2]) * sb2e2) * ca2n2,
             B6 = (s2.r[0] * s2.r[1]) * sa2n2;
    const m33 &R1 = s1.R, &R2 = s2.R;
    D4 n1[3], n2[3], d[3];
    for (int k = 0; k < 3; ++k) {
        n1[k] = (R1[k][0] * A1 + R1[k][1] * A2) + R1[k][2] * A3;
        n2[k] = (R2[k][0] * B4 + R2[k][1] * B5) + R2[k][2] * B6;
        // -Mk0 A4 - Mk1 A5 - Mk2 A6 + Nk0 B1 + Nk1 B2 + Nk2 B3 - c1k + c2k
        d[k] = (((((((-(R1[k][0] * A4)) - R1[k][1] * A5) - R1[k][2] * A6) + R2[k][0] * B1) + R2[k][1] * B2) + R2[k][2] * B3) -
                cst(s1.c[k])) +
               cst(s2.c[k]);
    }
    D4 f[4];
    f[0] = (-(n1[1] * d[2])) + n1[2] * d[1];
    f[1] = n1[0] * d[2] - n1[2] * d[0];
    f[2] = (-(n1[0] * d[1])) + n1[1] * d[0];
    const D4 dotn = (n1[0] * n2[0] + n1[1] * n2[1]) + n1[2] * n2[2];
    const D4 l1 = dsqrt((n1[0] * n1[0] + n1[1] * n1[1]) + n1[2] * n1[2]);
    const D4 l2 = dsqrt((n2[0] * n2[0] + n2[1] * n2[1]) + n2[2] * n2[2]);
    f[3] = dotn + l1 * l2;
    for (int i = 0; i < 4; ++i) {
        F[i] = f[i].v;
        if (J)
            for (int k = 0; k < 4; ++k) J[4 * i + k] = f[i].d[k];
    }
}

// Terms of one body at the angles (a, b), closest_points.hpp:426-446: cosines and sines, the
// signed powers of the residual (sign(x)|x|^p, p = n, 2 - n for a and e, 2 - e for b) and the
// unsigned factors of its derivative (n|x|^(n-1), (2-n)|x|^(1-n), e|x|^(e-1), (2-e)|x|^(1-e)).
struct BodyTerms {
    double ca, sa, cb, sb;
    double can, can2, san, san2, cbe, cbe2, sbe, sbe2;
    double can1, can3, san1, san3, cbe1, cbe3, sbe1, sbe3;
};
template <class M>
inline BodyTerms body_terms(const SE& s, double a, double b, bool deriv) {
    BodyTerms t{};
    t.ca = M::cos(a);
    t.sa = M::sin(a);
    t.cb = M::cos(b);
    t.sb = M::sin(b);
    const double aca = std::abs(t.ca), asa = std::abs(t.sa), acb = std::abs(t.cb), asb = std::abs(t.sb);
    t.can = sgn(t.ca) * M::pow(aca, s.n);
    t.can2 = sgn(t.ca) * M::pow(aca, 2 - s.n);
    t.san = sgn(t.sa) * M::pow(asa, s.n);
    t.san2 = sgn(t.sa) * M::pow(asa, 2 - s.n);
    t.cbe = sgn(t.cb) * M::pow(acb, s.e);
    t.cbe2 = sgn(t.cb) * M::pow(acb, 2 - s.e);
    t.sbe = sgn(t.sb) * M::pow(asb, s.e);
    t.sbe2 = sgn(t.sb) * M::pow(asb, 2 - s.e);
    if (deriv) {
        t.can1 = s.n * M::pow(aca, s.n - 1);
        t.can3 = (2 - s.n) * M::pow(aca, 1 - s.n);
        t.san1 = s.n * M::pow(asa, s.n - 1);
        t.san3 = (2 - s.n) * M::pow(asa, 1 - s.n);
        t.cbe1 = s.e * M::pow(acb, s.e - 1);
        t.cbe3 = (2 - s.e) * M::pow(acb, 1 - s.e);
        t.sbe1 = s.e * M::pow(asb, s.e - 1);
        t.sbe3 = (2 - s.e) * M::pow(asb, 1 - s.e);
    }
    return t;
}

// Body-frame normal n (A1..A3) and point p (A4..A6) of closest_points.hpp:448-453 and the
// derivative terms: along a, pa = A7..A9 and na = A10..A12; along b, pb = A13, A14 and
// nb = A15, A16 (:454-463).  The second body's B1..B16 (:496-518) are the same forms:
// B1..B3 = p, B4..B6 = n, B7..B9 = pa, B10, B11 = pb, B12..B14 = na, B15, B16 = nb.
struct BodyForms {
    double n[3], p[3];
    double pa[3], na[3], pb[2], nb[2];
};
inline BodyForms body_forms(const SE& s, const BodyTerms& t, bool deriv) {
    const double rx = s.r[0], ry = s.r[1], rz = s.r[2];
    BodyForms f{};
    f.n[0] = ((ry * rz) * t.can2) * t.cbe2;
    f.n[1] = ((rx * rz) * t.sbe2) * t.can2;
    f.n[2] = (rx * ry) * t.san2;
    f.p[0] = (rx * t.can) * t.cbe;
    f.p[1] = (ry * t.sbe) * t.can;
    f.p[2] = rz * t.san;
    if (deriv) {
        f.pa[0] = ((rx * t.sa) * t.can1) * t.cbe;
        f.pa[1] = ((ry * t.sa) * t.sbe) * t.can1;
        f.pa[2] = (rz * t.ca) * t.san1;
        f.na[0] = (((ry * rz) * t.sa) * t.can3) * t.cbe2;
        f.na[1] = (((rx * rz) * t.sa) * t.sbe2) * t.can3;
        f.na[2] = ((rx * ry) * t.ca) * t.san3;
        f.pb[0] = ((rx * t.sb) * t.can) * t.cbe1;
        f.pb[1] = ((ry * t.cb) * t.sbe1) * t.can;
        f.nb[0] = (((ry * rz) * t.sb) * t.can2) * t.cbe3;
        f.nb[1] = (((rx * rz) * t.cb) * t.sbe3) * t.can2;
    }
    return f;
}

// F (closest_points.hpp:284-382) and the reference's Jacobian (:385-603), row-major.  Every
// factor of the reference's entries is a rotated sum of these forms or its negation, and
// negating a sum is exact: -M00 A10 - M01 A11 + M02 A12 = -((M00 A10 + M01 A11) - M02 A12)
// bit for bit.  With q the row of the rotation:
//   pa1[q] = (Mq0 A7 + Mq1 A8) - Mq2 A9      na1[q] = (Mq0 A10 + Mq1 A11) - Mq2 A12
//   pb1[q] = Mq0 A13 - Mq1 A14               nb1[q] = Mq0 A15 - Mq1 A16
//   pa2[q] = (Nq0 B7 + Nq1 B8) - Nq2 B9      na2[q] = (Nq0 B12 + Nq1 B13) - Nq2 B14
//   pb2[q] = Nq0 B10 - Nq1 B11               nb2[q] = Nq0 B15 - Nq1 B16
// and every entry is summed left to right as the reference writes it.
template <class M>
inline void se_se_residual(const SEPair& P, const double* u, double* F, double* J /* 4x4 row-major or null */) {
    if (J && jacobian_ad()) {
        se_se_residual_ad<M>(P, u, F, J);
        return;
    }
    const SE &s1 = P.s1, &s2 = P.s2;
    const bool deriv = J != nullptr;
    const BodyForms A = body_forms(s1, body_terms<M>(s1, u[0], u[1], deriv), deriv);
    const BodyForms B = body_forms(s2, body_terms<M>(s2, u[2], u[3], deriv), deriv);
    const m33 &Mr = s1.R, &Nr = s2.R;
    double n1[3], n2[3], d[3];
    for (int q = 0; q < 3; ++q) {
        n1[q] = (Mr[q][0] * A.n[0] + Mr[q][1] * A.n[1]) + Mr[q][2] * A.n[2];
        n2[q] = (Nr[q][0] * B.n[0] + Nr[q][1] * B.n[1]) + Nr[q][2] * B.n[2];
        d[q] = (((((((-(Mr[q][0] * A.p[0])) - Mr[q][1] * A.p[1]) - Mr[q][2] * A.p[2]) + Nr[q][0] * B.p[0]) +
                  Nr[q][1] * B.p[1]) +
                 Nr[q][2] * B.p[2]) -
                s1.c[q]) +
               s2.c[q];
    }
    F[0] = (-n1[1]) * d[2] + n1[2] * d[1];
    F[1] = n1[0] * d[2] - n1[2] * d[0];
    F[2] = (-n1[0]) * d[1] + n1[1] * d[0];
    const double l1 = std::sqrt((n1[0] * n1[0] + n1[1] * n1[1]) + n1[2] * n1[2]);
    const double l2 = std::sqrt((n2[0] * n2[0] + n2[1] * n2[1]) + n2[2] * n2[2]);
    F[3] = ((n1[0] * n2[0] + n1[1] * n2[1]) + n1[2] * n2[2]) + l1 * l2;
    if (!J) return;
    double pa1[3], na1[3], pb1[3], nb1[3], pa2[3], na2[3], pb2[3], nb2[3];
    for (int q = 0; q < 3; ++q) {
        pa1[q] = (Mr[q][0] * A.pa[0] + Mr[q][1] * A.pa[1]) - Mr[q][2] * A.pa[2];
        na1[q] = (Mr[q][0] * A.na[0] + Mr[q][1] * A.na[1]) - Mr[q][2] * A.na[2];
        pb1[q] = Mr[q][0] * A.pb[0] - Mr[q][1] * A.pb[1];
        nb1[q] = Mr[q][0] * A.nb[0] - Mr[q][1] * A.nb[1];
        pa2[q] = (Nr[q][0] * B.pa[0] + Nr[q][1] * B.pa[1]) - Nr[q][2] * B.pa[2];
        na2[q] = (Nr[q][0] * B.na[0] + Nr[q][1] * B.na[1]) - Nr[q][2] * B.na[2];
        pb2[q] = Nr[q][0] * B.pb[0] - Nr[q][1] * B.pb[1];
        nb2[q] = Nr[q][0] * B.nb[0] - Nr[q][1] * B.nb[1];
    }
    // rows 0..2: the three components of n1 x (p2 - p1) (:528-571)
    J[0] = (((-n1[1]) * pa1[2] + n1[2] * pa1[1]) + na1[1] * d[2]) + (-na1[2]) * d[1];
    J[1] = (((-n1[1]) * pb1[2] + n1[2] * pb1[1]) + nb1[1] * d[2]) + (-nb1[2]) * d[1];
    J[2] = (-n1[1]) * (-pa2[2]) + n1[2] * (-pa2[1]);
    J[3] = (-n1[1]) * (-pb2[2]) + n1[2] * (-pb2[1]);
    J[4] = ((n1[0] * pa1[2] + (-n1[2]) * pa1[0]) + (-na1[0]) * d[2]) + na1[2] * d[0];
    J[5] = ((n1[0] * pb1[2] + (-n1[2]) * pb1[0]) + (-nb1[0]) * d[2]) + nb1[2] * d[0];
    J[6] = n1[0] * (-pa2[2]) + (-n1[2]) * (-pa2[0]);
    J[7] = n1[0] * (-pb2[2]) + (-n1[2]) * (-pb2[0]);
    J[8] = (((-n1[0]) * pa1[1] + n1[1] * pa1[0]) + na1[0] * d[1]) + (-na1[1]) * d[0];
    J[9] = (((-n1[0]) * pb1[1] + n1[1] * pb1[0]) + nb1[0] * d[1]) + (-nb1[1]) * d[0];
    J[10] = (-n1[0]) * (-pa2[1]) + n1[1] * (-pa2[0]);
    J[11] = (-n1[0]) * (-pb2[1]) + n1[1] * (-pb2[0]);
    // row 3: n1 . n2 + |n1| |n2| (:573-601)
    auto row3 = [](const double* a, const double* b, const double* g, double la, double lb) {
        return (((((a[0] * (-g[0]) + a[1] * (-g[1])) + a[2] * (-g[2])) * lb) / la + b[0] * (-g[0])) + b[1] * (-g[1])) +
               b[2] * (-g[2]);
    };
    J[12] = row3(n1, n2, na1, l1, l2);
    J[13] = row3(n1, n2, nb1, l1, l2);
    J[14] = row3(n2, n1, na2, l2, l1);
    J[15] = row3(n2, n1, nb2, l2, l1);
}

// ------------------------------------------------------------------ MINPACK hybrd (n = 4 in 3D, 2 in 2D)

// enorm: src/minpack.cpp:36-45, sqrt of the sum of squares in index order.
inline double enorm_ref(int n, const double* x) {
    double value = 0.0;
    for (int i = 0; i < n; ++i) value = value + x[i] * x[i];
    return std::sqrt(value);
}
// ANL-80-74's enorm (scaled sums for components below rdwarf or above rgiant / n); test-only.
inline double enorm_anl(int n, const double* x) {
    const double rdwarf = 3.834e-20, rgiant = 1.304e19;
    double s1 = 0, s2 = 0, s3 = 0, x1max = 0, x3max = 0;
    const double agiant = rgiant / n;
    for (int i = 0; i < n; ++i) {
        const double xabs = std::abs(x[i]);
        if (xabs > rdwarf && xabs < agiant) {
            s2 += xabs * xabs;
        } else if (xabs <= rdwarf) {
            if (xabs > x3max) {
                s3 = 1.0 + s3 * (x3max / xabs) * (x3max / xabs);
                x3max = xabs;
            } else if (xabs != 0.0) {
                s3 += (xabs / x3max) * (xabs / x3max);
            }
        } else {
            if (xabs > x1max) {
                s1 = 1.0 + s1 * (x1max / xabs) * (x1max / xabs);
                x1max = xabs;
            } else {
                s1 += (xabs / x1max) * (xabs / x1max);
            }
        }
    }
    if (s1 != 0.0) return x1max * std::sqrt(s1 + (s2 / x1max) / x1max);
    if (s2 != 0.0) {
        if (s2 >= x3max) return std::sqrt(s2 * (1.0 + (x3max / s2) * (x3max * s3)));
        return std::sqrt(x3max * ((s2 / x3max) + (x3max * s3)));
    }
    return x3max * std::sqrt(s3);
}
inline double enorm(int n, const double* x) { return (minpack_variant() & 1) ? enorm_anl(n, x) : enorm_ref(n, x); }

// the routines below for one system size N
template <int N>
struct Minpack {
static constexpr int LR = N * (N + 1) / 2;

// qrfac without pivoting (src/minpack.cpp:546-665); a is column-major n x n (a[i + n*j])
static inline void qrfac(double* a, double* rdiag, double* acnorm) {
    for (int j = 0; j < N; ++j) {
        acnorm[j] = enorm(N, a + N * j);
        rdiag[j] = acnorm[j];
    }
    for (int j = 0; j < N; ++j) {
        double ajnorm = enorm(N - j, a + j + N * j);
        if (ajnorm != 0.0) {
            if (a[j + N * j] < 0.0) ajnorm = -ajnorm;
            for (int i = j; i < N; ++i) a[i + N * j] /= ajnorm;
            a[j + N * j] += 1.0;
            for (int k = j + 1; k < N; ++k) {
                double sum = 0.0;
                for (int i = j; i < N; ++i) sum += a[i + N * j] * a[i + N * k];
                const double temp = sum / a[j + N * j];
                for (int i = j; i < N; ++i) a[i + N * k] -= temp * a[i + N * j];
            }
        }
        rdiag[j] = -ajnorm;
    }
}

// qform (src/minpack.cpp:243-314), m = n
static inline void qform(double* q, double* wa) {
    for (int j = 1; j < N; ++j)
        for (int i = 0; i < j; ++i) q[i + N * j] = 0.0;
    for (int l = 0; l < N; ++l) {
        const int k = N - 1 - l;
        for (int i = k; i < N; ++i) {
            wa[i] = q[i + N * k];
            q[i + N * k] = 0.0;
        }
        q[k + N * k] = 1.0;
        if (wa[k] != 0.0) {
            for (int j = k; j < N; ++j) {
                double sum = 0.0;
                for (int i = k; i < N; ++i) sum += q[i + N * j] * wa[i];
                const double temp = sum / wa[k];
                for (int i = k; i < N; ++i) q[i + N * j] -= temp * wa[i];
            }
        }
    }
}

// dogleg (src/minpack.cpp:101-240).  r: upper triangle packed by rows; 1-based index arithmetic
// kept.  The reference's scaled-gradient loop (:173-182) starts with l = 0 and updates
// wa1[i - 1] with r[l - 1]; it is reproduced here as written.  hybrd1 carves every work array
// out of one block (minpack.hpp:880-915: diag, qtf, wa1..wa4, r, fjac in that order) and hybrd
// passes its wa1, wa2, wa3 as dogleg's x, wa1, wa2 (minpack.hpp:628), so the element before
// dogleg's wa1 is x[n - 1] and the element before r is hybrd's wa4[n - 1], the residual at the
// last trial point (wa4_last).  The loop therefore adds wa4[n - 1] qtb[0] to the Gauss-Newton
// step's last component, accumulates each wa1[k] (k < n - 1) fully, divides it by diag[k], and
// then adds r(k, n - 1) qtb[k + 1]; wa1[n - 1] stays 0.  The ANL-80-74 loop (variant 1) is
// l = 1, wa1[i - 1] += r[l - 1] temp (variant bit 4).
// Returns true when the Gauss-Newton step was longer than delta (the scaled-gradient branch).
static inline bool dogleg(const double* r, const double* diag, const double* qtb, double delta, double* x, double* wa1,
                          double* wa2, double wa4_last) {
    const double epsmch = 2.220446049250313e-16;
    int jj = (N * (N + 1)) / 2 + 1;
    for (int k = 1; k <= N; ++k) {
        const int j = N - k + 1;
        const int jp1 = j + 1;
        jj -= k;
        int l = jj + 1;
        double sum = 0.0;
        for (int i = jp1; i <= N; ++i) {
            sum = sum + r[l - 1] * x[i - 1];
            ++l;
        }
        double temp = r[jj - 1];
        if (temp == 0.0) {
            l = j;
            for (int i = 1; i <= j; ++i) {
                temp = std::max(temp, std::abs(r[l - 1]));
                l += N - i;
            }
            temp = epsmch * temp;
            if (temp == 0.0) temp = epsmch;
        }
        x[j - 1] = (qtb[j - 1] - sum) / temp;
    }
    for (int j = 0; j < N; ++j) {
        wa1[j] = 0.0;
        wa2[j] = diag[j] * x[j];
    }
    const double qnorm = enorm(N, wa2);
    if (qnorm <= delta) return false;
    if (minpack_variant() & 4) {
        int l = 1;
        for (int j = 1; j <= N; ++j) {
            const double temp = qtb[j - 1];
            for (int i = j; i <= N; ++i) {
                wa1[i - 1] += r[l - 1] * temp;
                ++l;
            }
            wa1[j - 1] /= diag[j - 1];
        }
    } else {
        // the reference's loop; W(k) and Rk(l) are its wa1[k] and r[l] for k, l >= -1
        auto W = [&](int k) -> double& { return k >= 0 ? wa1[k] : x[N - 1]; };
        auto Rk = [&](int l) -> double { return l >= 0 ? r[l] : wa4_last; };
        int l = 0;
        for (int j = 0; j < N; ++j) {
            const double temp = qtb[j];
            for (int i = j; i < N; ++i) {
                W(i - 1) = W(i - 1) + Rk(l - 1) * temp;
                l = l + 1;
            }
            wa1[j] = wa1[j] / diag[j];
        }
    }
    const double gnorm = enorm(N, wa1);
    double sgnorm = 0.0;
    double alpha = delta / qnorm;
    if (gnorm != 0.0) {
        for (int j = 0; j < N; ++j) wa1[j] = (wa1[j] / gnorm) / diag[j];
        int l = 0;
        for (int j = 0; j < N; ++j) {
            double sum = 0.0;
            for (int i = j; i < N; ++i) {
                sum = sum + r[l] * wa1[i];
                ++l;
            }
            wa2[j] = sum;
        }
        const double temp = enorm(N, wa2);
        sgnorm = (gnorm / temp) / temp;
        alpha = 0.0;
        if (sgnorm < delta) {
            const double bnorm = enorm(N, qtb);
            double t = (bnorm / gnorm) * (bnorm / qnorm) * (sgnorm / delta);
            // src/minpack.cpp:223 multiplies left to right, (delta/qnorm)(sgnorm/delta) first;
            // ANL-80-74's (delta/qnorm)*(sgnorm/delta)**2 squares first (variant bit 4)
            const double dq = delta / qnorm, sd = sgnorm / delta;
            const double corr = (minpack_variant() & 4) ? dq * (sd * sd) : (dq * sd) * sd;
            t = t - corr +
                std::sqrt((t - (delta / qnorm)) * (t - (delta / qnorm)) +
                          (1.0 - (delta / qnorm) * (delta / qnorm)) * (1.0 - (sgnorm / delta) * (sgnorm / delta)));
            alpha = ((delta / qnorm) * (1.0 - (sgnorm / delta) * (sgnorm / delta))) / t;
        }
    }
    const double temp = (1.0 - alpha) * std::min(sgnorm, delta);
    for (int j = 0; j < N; ++j) x[j] = temp * wa1[j] + alpha * x[j];
    return true;
}

// Rank-1 update of the packed triangular s (m = n = N), src/minpack.cpp:373-542; 1-based index
// arithmetic kept.  giant = r8_huge() = 1e30 (:60-65, :397); DBL_MAX with variant bit 2.
static inline void r1updt(double* s, const double* u, double* v, double* w, bool& sing) {
    const double giant = (minpack_variant() & 2) ? 1.79769313486231570e308 : 1.0e30;
    const int m = N, n = N;
    int jj = (n * (2 * m - n + 1)) / 2 - (m - n);
    int l = jj;
    for (int i = n; i <= m; ++i) {
        w[i - 1] = s[l - 1];
        ++l;
    }
    for (int nmj = 1; nmj <= n - 1; ++nmj) {
        const int j = n - nmj;
        jj -= m - j + 1;
        w[j - 1] = 0.0;
        if (v[j - 1] != 0.0) {
            double cs, sn, tau;
            if (std::abs(v[n - 1]) < std::abs(v[j - 1])) {
                const double cotan = v[n - 1] / v[j - 1];
                sn = 0.5 / std::sqrt(0.25 + 0.25 * cotan * cotan);
                cs = sn * cotan;
                tau = 1.0;
                if (1.0 < std::abs(cs) * giant) tau = 1.0 / cs;
            } else {
                const double tn = v[j - 1] / v[n - 1];
                cs = 0.5 / std::sqrt(0.25 + 0.25 * tn * tn);
                sn = cs * tn;
                tau = sn;
            }
            v[n - 1] = sn * v[j - 1] + cs * v[n - 1];
            v[j - 1] = tau;
            l = jj;
            for (int i = j; i <= m; ++i) {
                const double temp = cs * s[l - 1] - sn * w[i - 1];
                w[i - 1] = sn * s[l - 1] + cs * w[i - 1];
                s[l - 1] = temp;
                ++l;
            }
        }
    }
    for (int i = 1; i <= m; ++i) w[i - 1] = w[i - 1] + v[n - 1] * u[i - 1];
    sing = false;
    for (int j = 1; j <= n - 1; ++j) {
        if (w[j - 1] != 0.0) {
            double cs, sn, tau;
            if (std::abs(s[jj - 1]) < std::abs(w[j - 1])) {
                const double cotan = s[jj - 1] / w[j - 1];
                sn = 0.5 / std::sqrt(0.25 + 0.25 * cotan * cotan);
                cs = sn * cotan;
                tau = 1.0;
                if (1.0 < std::abs(cs) * giant) tau = 1.0 / cs;
            } else {
                const double tn = w[j - 1] / s[jj - 1];
                cs = 0.5 / std::sqrt(0.25 + 0.25 * tn * tn);
                sn = cs * tn;
                tau = sn;
            }
            l = jj;
            for (int i = j; i <= m; ++i) {
                const double temp = cs * s[l - 1] + sn * w[i - 1];
                w[i - 1] = -sn * s[l - 1] + cs * w[i - 1];
                s[l - 1] = temp;
                ++l;
            }
            w[j - 1] = tau;
        }
        if (s[jj - 1] == 0.0) sing = true;
        jj += m - j + 1;
    }
    l = jj;
    for (int i = n; i <= m; ++i) {
        s[l - 1] = w[i - 1];
        ++l;
    }
    if (s[jj - 1] == 0.0) sing = true;
}

// r1mpyq (src/minpack.cpp:316-371); a is column-major m x n with leading dimension lda
static inline void r1mpyq(int m, int n, double* a, int lda, const double* v, const double* w) {
    for (int j = n - 2; 0 <= j; --j) {
        double cs, sn;
        if (1.0 < std::abs(v[j])) {
            cs = 1.0 / v[j];
            sn = std::sqrt(1.0 - cs * cs);
        } else {
            sn = v[j];
            cs = std::sqrt(1.0 - sn * sn);
        }
        for (int i = 0; i < m; ++i) {
            const double temp = cs * a[i + j * lda] - sn * a[i + (n - 1) * lda];
            a[i + (n - 1) * lda] = sn * a[i + j * lda] + cs * a[i + (n - 1) * lda];
            a[i + j * lda] = temp;
        }
    }
    for (int j = 0; j < n - 1; ++j) {
        double cs, sn;
        if (1.0 < std::abs(w[j])) {
            cs = 1.0 / w[j];
            sn = std::sqrt(1.0 - cs * cs);
        } else {
            sn = w[j];
            cs = std::sqrt(1.0 - sn * sn);
        }
        for (int i = 0; i < m; ++i) {
            const double temp = cs * a[i + j * lda] + sn * a[i + (n - 1) * lda];
            a[i + (n - 1) * lda] = -sn * a[i + j * lda] + cs * a[i + (n - 1) * lda];
            a[i + j * lda] = temp;
        }
    }
}

// hybrd1 (minpack.hpp:851-923) running hybrd (:337-831) with the analytic Jacobian: returns
// info (1: converged).  The reference's work block is fresh, uninitialised memory; wa4 is read
// by dogleg before its first write only on a scaled-gradient first step, which reads 0 here
// (counted in dogleg_fresh_wa4_count()).
template <class Fn>
static inline int hybrd1(Fn fcn, double* x, int* nfev_out = nullptr, int* njev_out = nullptr) {
    const bool anl = (minpack_variant() & 8) != 0, anl_dogleg = (minpack_variant() & 4) != 0;
    const double xtol = 1.0e-10, factor = 100.0, epsmch = 2.220446049250313e-16;
    const int maxfev = 200 * (N + 1);
    const double p1 = 0.1, p5 = 0.5, p001 = 0.001, p0001 = 0.0001;
    double diag[N];  // mode 2: wa[0..n) = 1 (minpack.hpp:881-884)
    for (int k = 0; k < N; ++k) diag[k] = 1.0;
    double fvec[N], fjac[N * N], r[LR], qtf[N], wa1[N], wa2[N], wa3[N], wa4[N];
    for (int k = 0; k < N; ++k) wa4[k] = 0.0;
    bool wa4_written = false;
    int info = 0, nfev = 0, njev = 0;
    fcn(x, fvec, nullptr);
    nfev = 1;
    double fnorm = enorm(N, fvec);
    const int msum = N;  // min(ml + mu + 1, n) with ml = mu = n - 1
    int iter = 1, ncsuc = 0, ncfail = 0, nslow1 = 0, nslow2 = 0;
    double delta = 0.0, xnorm = 0.0;
    for (;;) {
        bool jeval = true;
        double Jrow[N * N];
        double ftmp[N];
        fcn(x, ftmp, Jrow);
        ++njev;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) fjac[i + N * j] = Jrow[N * i + j];
        nfev += msum;
        qrfac(fjac, wa1, wa2);
        if (iter == 1) {
            for (int j = 0; j < N; ++j) wa3[j] = diag[j] * x[j];
            xnorm = enorm(N, wa3);
            if (anl) {
                delta = factor * xnorm;
                if (delta == 0.0) delta = factor;
            } else {
                delta = (xnorm == 0.0) ? factor : factor * xnorm;
            }
        }
        for (int i = 0; i < N; ++i) qtf[i] = fvec[i];
        for (int j = 0; j < N; ++j) {
            if (fjac[j + N * j] != 0.0) {
                double sum = 0.0;
                for (int i = j; i < N; ++i) sum = sum + fjac[i + N * j] * qtf[i];
                const double temp = -sum / fjac[j + N * j];
                for (int i = j; i < N; ++i) qtf[i] = qtf[i] + fjac[i + N * j] * temp;
            }
        }
        for (int j = 0; j < N; ++j) {
            int l = j;
            for (int i = 0; i < j; ++i) {
                r[l] = fjac[i + N * j];
                l += N - i - 1;
            }
            r[l] = wa1[j];
        }
        qform(fjac, wa1);
        for (;;) {
            if (dogleg(r, diag, qtf, delta, wa1, wa2, wa3, wa4[N - 1]) && !wa4_written && !anl_dogleg)
                ++dogleg_fresh_wa4_count();
            for (int j = 0; j < N; ++j) {
                wa1[j] = -wa1[j];
                wa2[j] = x[j] + wa1[j];
                wa3[j] = diag[j] * wa1[j];
            }
            const double pnorm = enorm(N, wa3);
            if (iter == 1) delta = std::min(delta, pnorm);
            fcn(wa2, wa4, nullptr);
            wa4_written = true;
            ++nfev;
            const double fnorm1 = enorm(N, wa4);
            double actred = -1.0;
            if (fnorm1 < fnorm) actred = 1.0 - (fnorm1 / fnorm) * (fnorm1 / fnorm);
            int l = 0;
            for (int i = 0; i < N; ++i) {
                double sum = 0.0;
                for (int j = i; j < N; ++j) {
                    sum = sum + r[l] * wa1[j];
                    ++l;
                }
                wa3[i] = qtf[i] + sum;
            }
            const double temp = enorm(N, wa3);
            double prered = 0.0;
            if (temp < fnorm) prered = 1.0 - (temp / fnorm) * (temp / fnorm);
            double ratio = 0.0;
            if (0.0 < prered) ratio = actred / prered;
            if (ratio < p1) {
                ncsuc = 0;
                ++ncfail;
                delta = p5 * delta;
            } else {
                ncfail = 0;
                ++ncsuc;
                if (p5 <= ratio || 1 < ncsuc) delta = std::max(delta, pnorm / p5);
                if (std::abs(ratio - 1.0) <= p1) delta = pnorm / p5;
            }
            if (p0001 <= ratio) {
                for (int j = 0; j < N; ++j) {
                    x[j] = wa2[j];
                    wa2[j] = diag[j] * x[j];
                    fvec[j] = wa4[j];
                }
                xnorm = enorm(N, wa2);
                fnorm = fnorm1;
                ++iter;
            }
            ++nslow1;
            if (p001 <= actred) nslow1 = 0;
            if (jeval) ++nslow2;
            if (p1 <= actred) nslow2 = 0;
            if (delta <= xtol * xnorm || fnorm == 0.0) info = 1;
            if (info != 0) break;
            if (anl) {  // ANL-80-74: every test runs, the last that holds sets info
                if (maxfev <= nfev) info = 2;
                if (p1 * std::max(p1 * delta, pnorm) <= epsmch * xnorm) info = 3;
                if (nslow2 == 5) info = 4;
                if (nslow1 == 10) info = 5;
            } else {  // the reference returns at the first test that holds (minpack.hpp:760-791)
                if (maxfev <= nfev)
                    info = 2;
                else if (p1 * std::max(p1 * delta, pnorm) <= epsmch * xnorm)
                    info = 3;
                else if (nslow2 == 5)
                    info = 4;
                else if (nslow1 == 10)
                    info = 5;
            }
            if (info != 0) break;
            if (ncfail == 2) break;
            for (int j = 0; j < N; ++j) {
                double sum = 0.0;
                for (int i = 0; i < N; ++i) sum = sum + fjac[i + N * j] * wa4[i];
                wa2[j] = (sum - wa3[j]) / pnorm;
                wa1[j] = diag[j] * ((diag[j] * wa1[j]) / pnorm);
                if (p0001 <= ratio) qtf[j] = sum;
            }
            bool sg = false;
            r1updt(r, wa1, wa2, wa3, sg);
            r1mpyq(N, N, fjac, N, wa2, wa3);
            r1mpyq(1, N, qtf, 1, wa2, wa3);
            jeval = false;
        }
        if (info != 0) break;
    }
    if (nfev_out) *nfev_out = nfev;
    if (njev_out) *njev_out = njev;
    if (info == 5) info = 4;  // hybrd1 (minpack.hpp:918-921)
    return info;
}

};

// closest_points(superellipsoid<3>, superellipsoid<3>): closest_points.hpp:263-733.
// Returns the reference's (pi, pj, nij, dij); u holds the root and info hybrd's code.
struct SEResult {
    v3 pi, pj, nij;
    double dij;
    double u[4];
    int info;
    int nfev;
    int njev;
};

// First minimum, in row-major order (xt::where), of the penalised distance
// |p1 - p2| + 2 (1 + n1.n2) over the two 576-point grids: closest_points.hpp:604-690.
template <class M>
inline void se_grid_start(const SE& s1, const SE& s2, double* u0) {
    const int num = 6;
    std::vector<double> b1 = unique_sorted(binit_xy<M>(s1, num)), a1 = unique_sorted(ainit_yz<M>(s1, num));
    std::vector<double> b2 = unique_sorted(binit_xy<M>(s2, num)), a2 = unique_sorted(ainit_yz<M>(s2, num));
    // meshgrid(ainit, binit) flattened row-major: k -> (a[k / nb], b[k % nb])
    const int n1 = (int)(a1.size() * b1.size()), n2 = (int)(a2.size() * b2.size());
    std::vector<v3> P1(n1), N1(n1), P2(n2), N2(n2);
    for (int k = 0; k < n1; ++k) {
        const double a = a1[k / b1.size()], b = b1[k % b1.size()];
        P1[k] = se_point<M>(s1, a, b);
        N1[k] = se_normal<M>(s1, a, b);
    }
    for (int k = 0; k < n2; ++k) {
        const double a = a2[k / b2.size()], b = b2[k % b2.size()];
        P2[k] = se_point<M>(s2, a, b);
        N2[k] = se_normal<M>(s2, a, b);
    }
    double dmin = INFINITY;
    int imin = 0, jmin = 0;
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j) {
            const double dx = P1[i][0] - P2[j][0], dy = P1[i][1] - P2[j][1], dz = P1[i][2] - P2[j][2];
            const double dist = std::sqrt(dx * dx + dy * dy + dz * dz) +
                                2 * (1 + (N1[i][0] * N2[j][0] + N1[i][1] * N2[j][1] + N1[i][2] * N2[j][2]));
            if (dist < dmin) {  // first minimum in row-major order
                dmin = dist;
                imin = i;
                jmin = j;
            }
        }
    u0[0] = a1[imin / b1.size()];
    u0[1] = b1[imin % b1.size()];
    u0[2] = a2[jmin / b2.size()];
    u0[3] = b2[jmin % b2.size()];
}

template <class M>
inline SEResult se_se(const SE& s1, const SE& s2) {
    SEResult res;
    se_grid_start<M>(s1, s2, res.u);
    SEPair pair{s1, s2};
    res.info = Minpack<4>::hybrd1([&](const double* u, double* F, double* J) { se_se_residual<M>(pair, u, F, J); }, res.u,
                                   &res.nfev, &res.njev);
    res.pi = se_point<M>(s1, res.u[0], res.u[1]);
    res.pj = se_point<M>(s2, res.u[2], res.u[3]);
    res.nij = se_normal<M>(s2, res.u[2], res.u[3]);
    const v3 d = {res.pi[0] - res.pj[0], res.pi[1] - res.pj[1], res.pi[2] - res.pj[2]};
    const double sd = sgn(d[0] * res.nij[0] + d[1] * res.nij[1] + d[2] * res.nij[2]);
    res.dij = sd * std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    return res;
}

// ================================================================== sphere / plane against a superellipsoid
// closest_points(sphere<3>, superellipsoid<3>) (closest_points.hpp:944-1281) and
// closest_points(superellipsoid<3>, plane<3>) (:1478-1746): damped Newton
// (utils.hpp:81-118: newton_method(u0, F, GradF, args, 200, ftol 1e-10, xtol 1e-7))
// from the first minimum of a 10 x 10 table of point distances over the diagonal
// (a_i, b_i) of linspace(-pi/2, pi/2, 10) x linspace(-pi, pi, 10).
//
// Residuals, with s1 the superellipsoid (the reference names it s1 in both):
//   F = [(p2 - p1) x n1 ; n1 . n2 + |n1| |n2|], p2 / n2 the sphere's point / normal, or the
//   plane's point R (0, a2, b2) + c and ((c_s1 - c_p) . n_p) n_p.
// Both write the superellipsoid normal as n1 = R (|ca1|^(2-n) cb1e2 ry rz,
// |ca1|^(2-n) sb1e2 rx rz, sign(ca1) sa1n2 rx ry) (:1018-1020, :1571-1573), i.e. sign(cos a1)
// times superellipsoid::normal.  GradF (:1053-1252, :1577-1711) is the chain rule of F
// with the reference's partial derivatives of these terms; for the sphere pair the
// derivative of |cos a1|^(2-n) is written (2-n)|cos a1|^(1-n) (-sin a1), without the
// sign(cos a1) factor of the true derivative (:1088, ca1n3), which the plane pair has
// (:1617).  The restatement evaluates the Jacobian by forward-mode differentiation
// with that rule.
struct Sph {
    v3 c;
    double r;
    m33 R;
};
struct Pln {
    v3 c;
    m33 R;
};

// sphere::point / normal (a, b): sphere.hpp:203-234
template <class M>
inline v3 sph_point(const Sph& s, double a, double b) {
    const v3 pt = {s.r * M::cos(a) * M::cos(b), s.r * M::cos(a) * M::sin(b), s.r * M::sin(a)};
    const v3 w = rot(s.R, pt);
    return {w[0] + s.c[0], w[1] + s.c[1], w[2] + s.c[2]};
}
template <class M>
inline v3 sph_normal(const Sph& s, double a, double b) {
    const v3 nn = {M::cos(a) * M::cos(b), M::cos(a) * M::sin(b), M::sin(a)};
    const v3 w = rot(s.R, nn);
    const double l = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    return {w[0] / l, w[1] / l, w[2] / l};
}
// plane::point(a, b) = R (0, a, b) + c (plane.hpp:184-192); plane::normal = column 0 of R
inline v3 pln_point(const Pln& p, double a, double b) {
    const v3 w = rot(p.R, v3{0.0, a, b});
    return {w[0] + p.c[0], w[1] + p.c[1], w[2] + p.c[2]};
}
inline v3 pln_normal(const Pln& p) { return {p.R[0][0], p.R[1][0], p.R[2][0]}; }

// |x|^p (no sign) with derivative p |x|^(p-1) sign(x) x' (exact) or p |x|^(p-1) x' (the
// sphere pair's Jacobian)
template <class M>
inline D4 upow2(const D4& a, double s, bool exact) {
    const double ax = std::abs(a.v);
    D4 r{M::pow(ax, 2 - s), {}};
    const double dv = exact ? ((2 - s) * sgn(a.v)) * M::pow(ax, 1 - s) : (2 - s) * M::pow(ax, 1 - s);
    for (int k = 0; k < 4; ++k) r.d[k] = dv * a.d[k];
    return r;
}
inline D4 dabs(const D4& a) {
    D4 r{std::abs(a.v), {}};
    const double sg = sgn(a.v);
    for (int k = 0; k < 4; ++k) r.d[k] = sg * a.d[k];
    return r;
}

// the superellipsoid terms n1 (A1..A3) and p1 - c1 (A4..A6) of both residuals
template <class M>
inline void se_side_terms(const SE& s1, const D4& a1, const D4& b1, bool exact, D4* A) {
    const D4 ca1 = dcos<M>(a1), sa1 = dsin<M>(a1), cb1 = dcos<M>(b1), sb1 = dsin<M>(b1);
    const D4 ca1n = spow<M>(ca1, s1.n), ca1n2 = upow2<M>(ca1, s1.n, exact);
    const D4 sa1n = spow<M>(sa1, s1.n), sa1n2 = spow2<M>(sa1, s1.n);
    const D4 cb1e = spow<M>(cb1, s1.e), cb1e2 = spow2<M>(cb1, s1.e);
    const D4 sb1e = spow<M>(sb1, s1.e), sb1e2 = spow2<M>(sb1, s1.e);
    A[0] = ((s1.r[1] * s1.r[2]) * ca1n2) * cb1e2;
    A[1] = ((s1.r[0] * s1.r[2]) * sb1e2) * ca1n2;
    A[2] = sgn(ca1.v) * ((s1.r[0] * s1.r[1]) * sa1n2);
    A[3] = (s1.r[0] * ca1n) * cb1e;
    A[4] = (s1.r[1] * sb1e) * ca1n;
    A[5] = s1.r[2] * sa1n;
}

// F and J (row-major or null) of the sphere pair, closest_points.hpp:958-1252
template <class M>
inline void sph_se_residual(const SE& s1, const Sph& s2, const double* u, double* F, double* J) {
    D4 a1 = var(u[0], 0), b1 = var(u[1], 1), a2 = var(u[2], 2), b2 = var(u[3], 3);
    D4 A[6];
    se_side_terms<M>(s1, a1, b1, false, A);
    const D4 ca2 = dcos<M>(a2), sa2 = dsin<M>(a2), cb2 = dcos<M>(b2), sb2 = dsin<M>(b2);
    const D4 ca2n2 = dabs(ca2);
    const double r = s2.r;
    const D4 B1 = (r * ca2) * cb2, B2 = (r * sb2) * ca2, B3 = r * sa2;
    const D4 B4 = ((r * r) * ca2n2) * cb2, B5 = ((r * r) * sb2) * ca2n2, B6 = sgn(ca2.v) * ((r * r) * sa2);
    const m33 &R1 = s1.R, &R2 = s2.R;
    D4 n1[3], n2[3], d[3];
    for (int k = 0; k < 3; ++k) {
        n1[k] = (R1[k][0] * A[0] + R1[k][1] * A[1]) + R1[k][2] * A[2];
        n2[k] = (R2[k][0] * B4 + R2[k][1] * B5) + R2[k][2] * B6;
        d[k] = (((((((-(R1[k][0] * A[3])) - R1[k][1] * A[4]) - R1[k][2] * A[5]) + R2[k][0] * B1) + R2[k][1] * B2) +
                 R2[k][2] * B3) -
                cst(s1.c[k])) +
               cst(s2.c[k]);
    }
    D4 f[4];
    f[0] = (-(n1[1] * d[2])) + n1[2] * d[1];
    f[1] = n1[0] * d[2] - n1[2] * d[0];
    f[2] = (-(n1[0] * d[1])) + n1[1] * d[0];
    const D4 dotn = (n1[0] * n2[0] + n1[1] * n2[1]) + n1[2] * n2[2];
    const D4 l1 = dsqrt((n1[0] * n1[0] + n1[1] * n1[1]) + n1[2] * n1[2]);
    const D4 l2 = dsqrt((n2[0] * n2[0] + n2[1] * n2[1]) + n2[2] * n2[2]);
    f[3] = dotn + l1 * l2;
    for (int i = 0; i < 4; ++i) {
        F[i] = f[i].v;
        if (J)
            for (int k = 0; k < 4; ++k) J[4 * i + k] = f[i].d[k];
    }
}

// F and J of the plane pair, closest_points.hpp:1492-1711.  sp is args(29) of the reference,
// which is the dot product (c_s1 - c_p) . n_p itself (:1484-1493), not its sign as the
// name `sign_p2s` suggests; F3 = sp n_p . n1 + |n_p| |n1| then has a root only when |sp| >= 1.
template <class M>
inline void se_pln_residual(const SE& s1, const Pln& p2, double sp, const double* u, double* F, double* J) {
    D4 a1 = var(u[0], 0), b1 = var(u[1], 1), a2 = var(u[2], 2), b2 = var(u[3], 3);
    D4 A[6];
    se_side_terms<M>(s1, a1, b1, true, A);
    const m33 &R1 = s1.R, &N = p2.R;
    D4 n1[3], d[3];
    for (int k = 0; k < 3; ++k) {
        n1[k] = (R1[k][0] * A[0] + R1[k][1] * A[1]) + R1[k][2] * A[2];
        d[k] = ((((((-(R1[k][0] * A[3])) - R1[k][1] * A[4]) - R1[k][2] * A[5]) + N[k][1] * a2) + N[k][2] * b2) +
                cst(p2.c[k])) -
               cst(s1.c[k]);
    }
    D4 f[4];
    f[0] = (-(n1[1] * d[2])) + n1[2] * d[1];
    f[1] = n1[0] * d[2] - n1[2] * d[0];
    f[2] = (-(n1[0] * d[1])) + n1[1] * d[0];
    const D4 dotn = ((N[0][0] * sp) * n1[0] + (N[1][0] * sp) * n1[1]) + (N[2][0] * sp) * n1[2];
    const double lp = std::sqrt((N[0][0] * N[0][0] + N[1][0] * N[1][0]) + N[2][0] * N[2][0]);
    const D4 l1 = dsqrt((n1[0] * n1[0] + n1[1] * n1[1]) + n1[2] * n1[2]);
    f[3] = dotn + lp * l1;
    for (int i = 0; i < 4; ++i) {
        F[i] = f[i].v;
        if (J)
            for (int k = 0; k < 4; ++k) J[4 * i + k] = f[i].d[k];
    }
}

// xt::linalg::solve (LAPACK gesv): LU with partial pivoting (first largest |pivot|),
// forward and back substitution.  A is row-major N x N, b is overwritten with x.
template <int N>
inline void lu_solve(double* A, double* b) {
    int piv[N];
    for (int k = 0; k < N; ++k) {
        int p = k;
        double best = std::abs(A[N * k + k]);
        for (int i = k + 1; i < N; ++i)
            if (std::abs(A[N * i + k]) > best) {
                best = std::abs(A[N * i + k]);
                p = i;
            }
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < N; ++j) std::swap(A[N * k + j], A[N * p + j]);
        if (A[N * k + k] != 0.0) {
            const double inv = 1.0 / A[N * k + k];
            for (int i = k + 1; i < N; ++i) A[N * i + k] *= inv;
        }
        for (int i = k + 1; i < N; ++i)
            for (int j = k + 1; j < N; ++j) A[N * i + j] -= A[N * i + k] * A[N * k + j];
    }
    for (int k = 0; k < N; ++k)
        if (piv[k] != k) std::swap(b[k], b[piv[k]]);
    for (int i = 1; i < N; ++i)
        for (int j = 0; j < i; ++j) b[i] -= A[N * i + j] * b[j];
    for (int i = N - 1; i >= 0; --i) {
        for (int j = i + 1; j < N; ++j) b[i] -= A[N * i + j] * b[j];
        b[i] /= A[N * i + i];
    }
}
inline void lu_solve4(double* A, double* b) { lu_solve<4>(A, b); }

template <int N>
inline double normv(const double* x) {
    double s = 0.0;
    for (int k = 0; k < N; ++k) s += x[k] * x[k];
    return std::sqrt(s);
}

// newton_method (utils.hpp:81-118); returns the iteration count, -1 when itermax is reached
template <int N, class Fn>
inline int newton(Fn fcn, double* u, int itermax = 200, double ftol = 1.0e-10, double xtol = 1.0e-7) {
    int cc = 0;
    while (cc < itermax) {
        double F[N], J[N * N], d[N];
        fcn(u, F, J);
        for (int k = 0; k < N; ++k) d[k] = -F[k];
        lu_solve<N>(J, d);
        if (normv<N>(d) < xtol) return cc;
        double t = 1;
        const double ferr = normv<N>(F);
        if (ferr < ftol) return cc;
        for (;;) {
            double ut[N], Ft[N];
            for (int k = 0; k < N; ++k) ut[k] = u[k] + t * d[k];
            fcn(ut, Ft, nullptr);
            if (!(normv<N>(Ft) > ferr && t > 0.01)) break;
            t -= 0.01;
        }
        for (int k = 0; k < N; ++k) u[k] += t * d[k];
        cc += 1;
    }
    return -1;
}

// first minimum of the 10 x 10 diagonal-grid distance table (closest_points.hpp:1255-1271, 1720-1736)
template <class P1, class P2>
inline void diag_grid_start(P1 point1, P2 point2, double* u0) {
    const double pi = 4 * std::atan(1);
    const int num = 10;
    double dmin = INFINITY;
    int imin = 0, jmin = 0;
    for (int i = 0; i < num; ++i) {
        const v3 p = point1(linspace(-pi / 2, pi / 2, num, i), linspace(-pi, pi, num, i));
        for (int j = 0; j < num; ++j) {
            const v3 q = point2(linspace(-pi / 2, pi / 2, num, j), linspace(-pi, pi, num, j));
            const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
            const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (dist < dmin) {
                dmin = dist;
                imin = i;
                jmin = j;
            }
        }
    }
    u0[0] = linspace(-pi / 2, pi / 2, num, imin);
    u0[1] = linspace(-pi, pi, num, imin);
    u0[2] = linspace(-pi / 2, pi / 2, num, jmin);
    u0[3] = linspace(-pi, pi, num, jmin);
}

struct NewtonResult {
    v3 pi, pj, nij;
    double dij;
    double u0[4], u[4];
    int iterations;  // newton_method's count, -1 when it did not converge
};

// closest_points(sphere s2, superellipsoid s1): pi on the superellipsoid, pj on the sphere,
// nij the sphere's normal, dij = (pi - pj) . nij (closest_points.hpp:1275-1279)
template <class M>
inline NewtonResult sph_se(const Sph& s2, const SE& s1) {
    NewtonResult res;
    diag_grid_start([&](double a, double b) { return se_point<M>(s1, a, b); },
                    [&](double a, double b) { return sph_point<M>(s2, a, b); }, res.u0);
    for (int k = 0; k < 4; ++k) res.u[k] = res.u0[k];
    res.iterations =
        newton<4>([&](const double* u, double* F, double* J) { sph_se_residual<M>(s1, s2, u, F, J); }, res.u);
    res.pi = se_point<M>(s1, res.u[0], res.u[1]);
    res.pj = sph_point<M>(s2, res.u[2], res.u[3]);
    res.nij = sph_normal<M>(s2, res.u[2], res.u[3]);
    res.dij = (res.pi[0] - res.pj[0]) * res.nij[0] + (res.pi[1] - res.pj[1]) * res.nij[1] +
              (res.pi[2] - res.pj[2]) * res.nij[2];
    return res;
}

// closest_points(superellipsoid s1, plane p2): pi on the superellipsoid, pj on the plane,
// nij the plane's normal (not oriented), dij = (pi - pj) . nij (closest_points.hpp:1740-1744)
template <class M>
inline NewtonResult se_pln(const SE& s1, const Pln& p2) {
    NewtonResult res;
    const v3 np = pln_normal(p2);
    const double sp = (s1.c[0] - p2.c[0]) * np[0] + (s1.c[1] - p2.c[1]) * np[1] + (s1.c[2] - p2.c[2]) * np[2];
    diag_grid_start([&](double a, double b) { return se_point<M>(s1, a, b); },
                    [&](double a, double b) { return pln_point(p2, a, b); }, res.u0);
    for (int k = 0; k < 4; ++k) res.u[k] = res.u0[k];
    res.iterations =
        newton<4>([&](const double* u, double* F, double* J) { se_pln_residual<M>(s1, p2, sp, u, F, J); }, res.u);
    res.pi = se_point<M>(s1, res.u[0], res.u[1]);
    res.pj = pln_point(p2, res.u[2], res.u[3]);
    res.nij = np;
    res.dij = (res.pi[0] - res.pj[0]) * res.nij[0] + (res.pi[1] - res.pj[1]) * res.nij[1] +
              (res.pi[2] - res.pj[2]) * res.nij[2];
    return res;
}

// ------------------------------------------------------------------ 2D: superellipsoid<2>
// A 2D superellipsoid uses c[0..1], r[0..1], the squareness e and the 2 x 2 block of R
// (rotation_matrix<2>); spheres and planes likewise.  The three pairs of closest_points.hpp
// (:64-249 by hybrd1, :1316-1464 and :1782-1958 by newton_method) solve 2 x 2 systems; their
// Jacobians (hand-written in the reference) are the exact derivatives of the residuals, which the
// dual numbers give here.

// superellipsoid<2>::point(b) / normal(b): superellipsoid.hpp:281-289, 303-312
template <class M>
inline v3 se2_point(const SE& s, double b) {
    const double cb = M::cos(b), sb = M::sin(b);
    const double x = s.r[0] * sgn(cb) * M::pow(std::abs(cb), s.e);
    const double y = s.r[1] * sgn(sb) * M::pow(std::abs(sb), s.e);
    return {(s.R[0][0] * x + s.R[0][1] * y) + s.c[0], (s.R[1][0] * x + s.R[1][1] * y) + s.c[1], 0.0};
}
template <class M>
inline v3 se2_normal(const SE& s, double b) {
    const double cb = M::cos(b), sb = M::sin(b);
    const double x = s.r[1] * sgn(cb) * M::pow(std::abs(cb), 2 - s.e);
    const double y = s.r[0] * sgn(sb) * M::pow(std::abs(sb), 2 - s.e);
    const double n0 = s.R[0][0] * x + s.R[0][1] * y, n1 = s.R[1][0] * x + s.R[1][1] * y;
    const double l = std::sqrt(n0 * n0 + n1 * n1);
    return {n0 / l, n1 / l, 0.0};
}
// sphere<2>::point(b) / normal(b): sphere.hpp:194-221; plane<2>::point(b) = R (0, b) + c (plane.hpp:175-182)
template <class M>
inline v3 sph2_point(const Sph& s, double b) {
    const double x = s.r * M::cos(b), y = s.r * M::sin(b);
    return {(s.R[0][0] * x + s.R[0][1] * y) + s.c[0], (s.R[1][0] * x + s.R[1][1] * y) + s.c[1], 0.0};
}
template <class M>
inline v3 sph2_normal(const Sph& s, double b) {
    const double x = s.r * M::cos(b), y = s.r * M::sin(b);
    const double n0 = s.R[0][0] * x + s.R[0][1] * y, n1 = s.R[1][0] * x + s.R[1][1] * y;
    const double l = std::sqrt(n0 * n0 + n1 * n1);
    return {n0 / l, n1 / l, 0.0};
}
inline v3 pln2_point(const Pln& p, double b) {
    return {(p.R[0][0] * 0.0 + p.R[0][1] * b) + p.c[0], (p.R[1][0] * 0.0 + p.R[1][1] * b) + p.c[1], 0.0};
}

// one superellipsoid side in the body frame: the normal direction (r_y c^(2-e), r_x s^(2-e)) and
// the point (r_x c^e, r_y s^e), c = cos b, s = sin b, powers keeping the sign
template <class M>
inline void se2_terms(const SE& s, const D4& b, D4& nx, D4& ny, D4& px, D4& py) {
    const D4 cb = dcos<M>(b), sb = dsin<M>(b);
    nx = s.r[1] * spow2<M>(cb, s.e);
    ny = s.r[0] * spow2<M>(sb, s.e);
    px = s.r[0] * spow<M>(cb, s.e);
    py = s.r[1] * spow<M>(sb, s.e);
}
inline void put2(const D4* f, double* F, double* J) {
    for (int i = 0; i < 2; ++i) {
        F[i] = f[i].v;
        if (J)
            for (int k = 0; k < 2; ++k) J[2 * i + k] = f[i].d[k];
    }
}

// closest_points(superellipsoid<2>, superellipsoid<2>) residual (closest_points.hpp:79-125) with the
// forward-mode Jacobian (test-side cross-check): the normal n1 of s1 at b1 is collinear with p2 - p1
// (cross product 0) and anti-parallel to n2 at b2
template <class M>
inline void se2_se2_residual_ad(const SE& s1, const SE& s2, const double* u, double* F, double* J) {
    D4 n1x, n1y, p1x, p1y, n2x, n2y, p2x, p2y;
    se2_terms<M>(s1, var(u[0], 0), n1x, n1y, p1x, p1y);
    se2_terms<M>(s2, var(u[1], 1), n2x, n2y, p2x, p2y);
    const auto& A = s1.R;
    const auto& B = s2.R;
    const D4 m1x = A[0][1] * n1y + A[0][0] * n1x, m1y = A[1][1] * n1y + A[1][0] * n1x;
    const D4 m2x = B[0][1] * n2y + B[0][0] * n2x, m2y = B[1][1] * n2y + B[1][0] * n2x;
    const D4 dy = (((((-(A[1][0] * p1x)) - A[1][1] * p1y) - cst(s1.c[1])) + B[1][0] * p2x) + B[1][1] * p2y) + cst(s2.c[1]);
    const D4 dx = (((((-(A[0][0] * p1x)) - A[0][1] * p1y) - cst(s1.c[0])) + B[0][0] * p2x) + B[0][1] * p2y) + cst(s2.c[0]);
    D4 f[2];
    f[0] = (-m1x) * dy + m1y * dx;
    f[1] = (m1x * m2x + m1y * m2y) + dsqrt(m1x * m1x + m1y * m1y) * dsqrt(m2x * m2x + m2y * m2y);
    put2(f, F, J);
}

// The same residual with the reference's hand-written Jacobian (closest_points.hpp:127-189).  Body
// terms F1..F4 (s1) and F5..F8 (s2): F1 = rx sb^(2-e), F2 = ry sb^e, F3 = ry cb^(2-e), F4 = rx cb^e
// (signed powers); derivative terms F9..F12 (s1), F13..F16 (s2): F9 = e rx sb |cb|^(e-1),
// F10 = e ry cb |sb|^(e-1), F11 = ry (2-e) sb |cb|^(1-e), F12 = rx (2-e) cb |sb|^(1-e).
// With the rotated sums m = (M01 F1 + M00 F3, M11 F1 + M10 F3), g1 = (M00 F9 - M01 F10,
// M10 F9 - M11 F10), h1 = (M01 F12 - M00 F11, M11 F12 - M10 F11) and the same for s2 with N
// (g2 from F13, F14; h2 from F15, F16), the reference's factors are these sums or their exact
// negations, and each entry is summed left to right as written.
template <class M>
inline void se2_se2_residual(const SE& s1, const SE& s2, const double* u, double* F, double* J) {
    if (J && jacobian_ad()) {
        se2_se2_residual_ad<M>(s1, s2, u, F, J);
        return;
    }
    struct Side {
        double f1, f2, f3, f4, f9, f10, f11, f12;
    };
    auto side = [&](const SE& s, double b) {
        const double cb = M::cos(b), sb = M::sin(b), acb = std::abs(cb), asb = std::abs(sb);
        const double cbe = sgn(cb) * M::pow(acb, s.e), cbe2 = sgn(cb) * M::pow(acb, 2 - s.e);
        const double sbe = sgn(sb) * M::pow(asb, s.e), sbe2 = sgn(sb) * M::pow(asb, 2 - s.e);
        Side t{s.r[0] * sbe2, s.r[1] * sbe, s.r[1] * cbe2, s.r[0] * cbe, 0, 0, 0, 0};
        if (J) {
            t.f9 = ((s.e * s.r[0]) * sb) * M::pow(acb, s.e - 1);
            t.f10 = ((s.e * s.r[1]) * cb) * M::pow(asb, s.e - 1);
            t.f11 = ((s.r[1] * (2 - s.e)) * sb) * M::pow(acb, 1 - s.e);
            t.f12 = ((s.r[0] * (2 - s.e)) * cb) * M::pow(asb, 1 - s.e);
        }
        return t;
    };
    const Side a = side(s1, u[0]), b = side(s2, u[1]);
    const auto& Mr = s1.R;
    const auto& Nr = s2.R;
    const double m1x = Mr[0][1] * a.f1 + Mr[0][0] * a.f3, m1y = Mr[1][1] * a.f1 + Mr[1][0] * a.f3;
    const double m2x = Nr[0][1] * b.f1 + Nr[0][0] * b.f3, m2y = Nr[1][1] * b.f1 + Nr[1][0] * b.f3;
    const double dy = (((((-(Mr[1][0] * a.f4)) - Mr[1][1] * a.f2) - s1.c[1]) + Nr[1][0] * b.f4) + Nr[1][1] * b.f2) + s2.c[1];
    const double dx = (((((-(Mr[0][0] * a.f4)) - Mr[0][1] * a.f2) - s1.c[0]) + Nr[0][0] * b.f4) + Nr[0][1] * b.f2) + s2.c[0];
    F[0] = (-m1x) * dy + m1y * dx;
    const double l1 = std::sqrt(m1x * m1x + m1y * m1y), l2 = std::sqrt(m2x * m2x + m2y * m2y);
    F[1] = (m1x * m2x + m1y * m2y) + l1 * l2;
    if (!J) return;
    const double g1x = Mr[0][0] * a.f9 - Mr[0][1] * a.f10, g1y = Mr[1][0] * a.f9 - Mr[1][1] * a.f10;
    const double h1x = Mr[0][1] * a.f12 - Mr[0][0] * a.f11, h1y = Mr[1][1] * a.f12 - Mr[1][0] * a.f11;
    const double g2x = Nr[0][0] * b.f9 - Nr[0][1] * b.f10, g2y = Nr[1][0] * b.f9 - Nr[1][1] * b.f10;
    const double h2x = Nr[0][1] * b.f12 - Nr[0][0] * b.f11, h2y = Nr[1][1] * b.f12 - Nr[1][0] * b.f11;
    J[0] = (((-m1x) * g1y + m1y * g1x) + (-h1x) * dy) + h1y * dx;
    J[1] = (-m1x) * (-g2y) + m1y * (-g2x);
    J[2] = (((m1x * h1x + m1y * h1y) * l2) / l1 + m2x * h1x) + m2y * h1y;
    J[3] = (((m2x * h2x + m2y * h2y) * l1) / l2 + m1x * h2x) + m1y * h2y;
}

// closest_points(sphere<2> s2, superellipsoid<2> s1) residual (closest_points.hpp:1332-1366)
template <class M>
inline void sph2_se2_residual(const SE& s1, const Sph& s2, const double* u, double* F, double* J) {
    D4 n1x, n1y, p1x, p1y;
    se2_terms<M>(s1, var(u[0], 0), n1x, n1y, p1x, p1y);
    const D4 b2 = var(u[1], 1);
    const D4 q5 = s2.r * dsin<M>(b2), q6 = s2.r * dcos<M>(b2);
    const auto& A = s1.R;
    const auto& B = s2.R;
    const D4 m1x = A[0][0] * n1x + A[0][1] * n1y, m1y = A[1][0] * n1x + A[1][1] * n1y;
    const D4 m2x = B[0][0] * q6 + B[0][1] * q5, m2y = B[1][0] * q6 + B[1][1] * q5;
    const D4 dy = ((((-(A[1][0] * p1x)) - A[1][1] * p1y) + B[1][0] * q6) + B[1][1] * q5 - cst(s1.c[1])) + cst(s2.c[1]);
    const D4 dx = ((((-(A[0][0] * p1x)) - A[0][1] * p1y) + B[0][0] * q6) + B[0][1] * q5 - cst(s1.c[0])) + cst(s2.c[0]);
    D4 f[2];
    f[0] = (-m1x) * dy + m1y * dx;
    f[1] = (m2x * m1x + m2y * m1y) + dsqrt(m2x * m2x + m2y * m2y) * dsqrt(m1x * m1x + m1y * m1y);
    put2(f, F, J);
}

// closest_points(superellipsoid<2> s1, plane<2> p2) residual (closest_points.hpp:1800-1838); sp is
// the value (c_s1 - c_p) . n_p (the reference's `sign_d2s`, not a sign: F[1] has a root only
// when |sp| >= 1, the reason the reference's own tests of this pair are marked as failing)
template <class M>
inline void se2_pln2_residual(const SE& s1, const Pln& p2, double sp, const double* u, double* F, double* J) {
    D4 n1x, n1y, p1x, p1y;
    se2_terms<M>(s1, var(u[0], 0), n1x, n1y, p1x, p1y);
    const D4 a2 = var(u[1], 1);
    const auto& A = s1.R;
    const auto& N = p2.R;
    const D4 m1x = A[0][0] * n1x + A[0][1] * n1y, m1y = A[1][0] * n1x + A[1][1] * n1y;
    const D4 dy = ((((-(A[1][0] * p1x)) - A[1][1] * p1y) + N[1][1] * a2) + cst(p2.c[1])) - cst(s1.c[1]);
    const D4 dx = ((((-(A[0][0] * p1x)) - A[0][1] * p1y) + N[0][1] * a2) + cst(p2.c[0])) - cst(s1.c[0]);
    D4 f[2];
    f[0] = (-m1x) * dy + m1y * dx;
    const double lp = std::sqrt(N[0][0] * N[0][0] + N[1][0] * N[1][0]);
    f[1] = ((N[0][0] * sp) * m1x + (N[1][0] * sp) * m1y) + lp * dsqrt(m1x * m1x + m1y * m1y);
    put2(f, F, J);
}

inline double dist2(const v3& a, const v3& b) {
    const double dx = a[0] - b[0], dy = a[1] - b[1];
    return std::sqrt(dx * dx + dy * dy);
}

// closest_points(superellipsoid<2>, superellipsoid<2>): closest_points.hpp:64-249.  Start: the first
// minimum, row-major, of |p1 - p2| + 2 (1 + n1 . n2) over unique(binit_xy(4)) of each side.
template <class M>
inline SEResult se2_se2(const SE& s1, const SE& s2) {
    SEResult res;
    const std::vector<double> b1 = unique_sorted(binit_xy<M>(s1, 4)), b2 = unique_sorted(binit_xy<M>(s2, 4));
    double dmin = INFINITY;
    size_t imin = 0, jmin = 0;
    for (size_t i = 0; i < b1.size(); ++i) {
        const v3 p = se2_point<M>(s1, b1[i]), n = se2_normal<M>(s1, b1[i]);
        for (size_t j = 0; j < b2.size(); ++j) {
            const v3 q = se2_point<M>(s2, b2[j]), m = se2_normal<M>(s2, b2[j]);
            const double dist = dist2(p, q) + 2 * (1 + (n[0] * m[0] + n[1] * m[1]));
            if (dist < dmin) {
                dmin = dist;
                imin = i;
                jmin = j;
            }
        }
    }
    res.u[0] = b1[imin];
    res.u[1] = b2[jmin];
    res.u[2] = res.u[3] = 0.0;
    res.info = Minpack<2>::hybrd1([&](const double* u, double* F, double* J) { se2_se2_residual<M>(s1, s2, u, F, J); },
                                  res.u, &res.nfev, &res.njev);
    res.pi = se2_point<M>(s1, res.u[0]);
    res.pj = se2_point<M>(s2, res.u[1]);
    res.nij = se2_normal<M>(s2, res.u[1]);
    const double dx = res.pi[0] - res.pj[0], dy = res.pi[1] - res.pj[1];
    res.dij = sgn(dx * res.nij[0] + dy * res.nij[1]) * std::sqrt(dx * dx + dy * dy);
    return res;
}

// closest_points(sphere<2> s2, superellipsoid<2> s1): closest_points.hpp:1316-1448.  Start: the
// first minimum of |s1.point(b_i) - s2.point(b_j)| over b = {-pi, -pi/2, 0.001, pi/2}; pi on the
// superellipsoid, pj on the sphere, nij the sphere's normal, dij = (pi - pj) . nij.
template <class M>
inline NewtonResult sph2_se2(const Sph& s2, const SE& s1) {
    NewtonResult res;
    const double pi = 4 * std::atan(1);
    const double bi[4] = {-pi, -pi / 2, 0.001, pi / 2};
    double dmin = INFINITY;
    int imin = 0, jmin = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const double d = dist2(se2_point<M>(s1, bi[i]), sph2_point<M>(s2, bi[j]));
            if (d < dmin) {
                dmin = d;
                imin = i;
                jmin = j;
            }
        }
    res.u0[0] = res.u[0] = bi[imin];
    res.u0[1] = res.u[1] = bi[jmin];
    res.u0[2] = res.u0[3] = res.u[2] = res.u[3] = 0.0;
    res.iterations =
        newton<2>([&](const double* u, double* F, double* J) { sph2_se2_residual<M>(s1, s2, u, F, J); }, res.u);
    res.pi = se2_point<M>(s1, res.u[0]);
    res.pj = sph2_point<M>(s2, res.u[1]);
    res.nij = sph2_normal<M>(s2, res.u[1]);
    res.dij = (res.pi[0] - res.pj[0]) * res.nij[0] + (res.pi[1] - res.pj[1]) * res.nij[1];
    return res;
}

// closest_points(superellipsoid<2> s1, plane<2> p2): closest_points.hpp:1782-1949.  Start: the first
// minimum of |s1.point(b_i) - p2.point(a_j)| over b = linspace(-pi, pi, 10), a = linspace(-10, 10, 10);
// pi on the superellipsoid, pj on the plane, nij the plane's normal, dij = (pi - pj) . nij.
template <class M>
inline NewtonResult se2_pln2(const SE& s1, const Pln& p2) {
    NewtonResult res;
    const double pi = 4 * std::atan(1);
    const int num = 10;
    const v3 np = {p2.R[0][0], p2.R[1][0], 0.0};
    const double sp = (s1.c[0] - p2.c[0]) * np[0] + (s1.c[1] - p2.c[1]) * np[1];
    double dmin = INFINITY;
    int imin = 0, jmin = 0;
    for (int i = 0; i < num; ++i) {
        const v3 p = se2_point<M>(s1, linspace(-pi, pi, num, i));
        for (int j = 0; j < num; ++j) {
            const double d = dist2(p, pln2_point(p2, linspace(-num, num, num, j)));
            if (d < dmin) {
                dmin = d;
                imin = i;
                jmin = j;
            }
        }
    }
    res.u0[0] = res.u[0] = linspace(-pi, pi, num, imin);
    res.u0[1] = res.u[1] = linspace(-num, num, num, jmin);
    res.u0[2] = res.u0[3] = res.u[2] = res.u[3] = 0.0;
    res.iterations =
        newton<2>([&](const double* u, double* F, double* J) { se2_pln2_residual<M>(s1, p2, sp, u, F, J); }, res.u);
    res.pi = se2_point<M>(s1, res.u[0]);
    res.pj = pln2_point(p2, res.u[1]);
    res.nij = np;
    res.dij = (res.pi[0] - res.pj[0]) * res.nij[0] + (res.pi[1] - res.pj[1]) * res.nij[1];
    return res;
}

}  // namespace se_oracle
