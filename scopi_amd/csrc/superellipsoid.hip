// superellipsoid.hip -- device closest points of superellipsoid pairs (see superellipsoid.hpp).
#include "superellipsoid.hpp"

#include <algorithm>

// Every floating-point operation of this file is a separately rounded IEEE
// operation: no contraction into FMA, so that the device reproduces the CPU
// restatement bit for bit.
#pragma clang fp contract(off)

namespace scopi_hip {
namespace {

// ------------------------------------------------------------------ portable elementary functions
// Textbook methods, identical operation order on every platform:
//   cos / sin: Cody-Waite reduction by pi/2 (33 + 33 + 53 bits), Taylor to degree 17 / 18;
//   asin: Maclaurin series on [0, 1/2], asin x = pi/2 - 2 asin(sqrt((1 - x) / 2)) above;
//   log: 2 atanh((m - 1) / (m + 1)) series on m in [sqrt(2)/2, sqrt(2)), ln2 in 32 + 53 bits;
//   exp: k ln2 + r reduction, Taylor to degree 14, 2^k through the exponent bits;
//   pow(x >= 0, p) = exp(p log x).
constexpr double kPio2_1 = 0x1.921fb54400000p+0;
constexpr double kPio2_2 = 0x1.0b4611a600000p-34;
constexpr double kPio2_3 = 0x1.3198a2e037073p-69;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
constexpr double kPio2 = 0x1.921fb54442d18p+0;
constexpr double kPi = 0x1.921fb54442d18p+1;
constexpr double kLn2Hi = 0x1.62e42ff000000p-1;
constexpr double kLn2Lo = -0x1.718432a1b0e26p-35;
constexpr double kInvLn2 = 0x1.71547652b82fep+0;
constexpr double kSqrt2 = 0x1.6a09e667f3bcdp+0;
constexpr double kF2 = 0.5, kF3 = 1.0 / 6.0, kF4 = 1.0 / 24.0, kF5 = 1.0 / 120.0, kF6 = 1.0 / 720.0,
                 kF7 = 1.0 / 5040.0, kF8 = 1.0 / 40320.0, kF9 = 1.0 / 362880.0, kF10 = 1.0 / 3628800.0,
                 kF11 = 1.0 / 39916800.0, kF12 = 1.0 / 479001600.0, kF13 = 1.0 / 6227020800.0,
                 kF14 = 1.0 / 87178291200.0, kF15 = 1.0 / 1307674368000.0, kF16 = 1.0 / 20922789888000.0,
                 kF17 = 1.0 / 355687428096000.0, kF18 = 1.0 / 6402373705728000.0;

__device__ inline double sin_poly(double r) {
    const double s = r * r;
    double p = kF17;
    p = p * s - kF15;
    p = p * s + kF13;
    p = p * s - kF11;
    p = p * s + kF9;
    p = p * s - kF7;
    p = p * s + kF5;
    p = p * s - kF3;
    return r + (r * s) * p;
}
__device__ inline double cos_poly(double r) {
    const double s = r * r;
    double p = kF18;
    p = p * s - kF16;
    p = p * s + kF14;
    p = p * s - kF12;
    p = p * s + kF10;
    p = p * s - kF8;
    p = p * s + kF6;
    p = p * s - kF4;
    p = p * s + kF2;
    return 1.0 - s * p;
}
__device__ inline double reduce_pio2(double x, int& quadrant) {
    const double k = rint(x * kTwoOverPi);
    const double r = ((x - k * kPio2_1) - k * kPio2_2) - k * kPio2_3;
    quadrant = (int)(((int64_t)k) & 3);
    return r;
}
__device__ inline double pm_cos(double x) {
    if (!(fabs(x) < 1e15)) return x - x;
    int q;
    const double r = reduce_pio2(x, q);
    if (q == 0) return cos_poly(r);
    if (q == 1) return -sin_poly(r);
    if (q == 2) return -cos_poly(r);
    return sin_poly(r);
}
__device__ inline double pm_sin(double x) {
    if (!(fabs(x) < 1e15)) return x - x;
    int q;
    const double r = reduce_pio2(x, q);
    if (q == 0) return sin_poly(r);
    if (q == 1) return cos_poly(r);
    if (q == 2) return -sin_poly(r);
    return -cos_poly(r);
}
__device__ inline double asin_small(double x) {
    double c[30];
    double t = 1.0;
    c[0] = 1.0;
    for (int n = 1; n < 30; ++n) {
        t = t * (double)(2 * n - 1) / (double)(2 * n);
        c[n] = t / (double)(2 * n + 1);
    }
    const double z = x * x;
    double p = c[29];
    for (int n = 28; n >= 1; --n) p = p * z + c[n];
    return x + (x * z) * p;
}
__device__ inline double pm_asin(double x) {
    if (!(x >= -1.0 && x <= 1.0)) return (x - x) / (x - x);
    const double a = fabs(x);
    double r;
    if (a <= 0.5)
        r = asin_small(a);
    else
        r = kPio2 - 2.0 * asin_small(sqrt((1.0 - a) * 0.5));
    return x < 0 ? -r : r;
}
__device__ inline double pm_log(double x) {
    if (x != x || x < 0) return (x - x) / (x - x);
    if (x == 0) return -INFINITY;
    if (x == INFINITY) return x;
    int k = 0;
    if (x < 0x1p-1022) {
        x = x * 0x1p54;
        k = -54;
    }
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    k += (int)((u >> 52) & 0x7ff) - 1023;
    double m = __longlong_as_double((long long)((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL));
    if (m > kSqrt2) {
        m = m * 0.5;
        k += 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 23.0;
    p = p * z + 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    const double lm = 2.0 * s + (2.0 * s) * (z * p);
    const double dk = (double)k;
    return dk * kLn2Hi + (lm + dk * kLn2Lo);
}
__device__ inline double pm_exp(double y) {
    if (y != y) return y;
    if (y > 709.782712893384) return INFINITY;
    if (y < -745.2) return 0.0;
    const double k = rint(y * kInvLn2);
    const double r = (y - k * kLn2Hi) - k * kLn2Lo;
    double p = kF14;
    p = p * r + kF13;
    p = p * r + kF12;
    p = p * r + kF11;
    p = p * r + kF10;
    p = p * r + kF9;
    p = p * r + kF8;
    p = p * r + kF7;
    p = p * r + kF6;
    p = p * r + kF5;
    p = p * r + kF4;
    p = p * r + kF3;
    p = p * r + kF2;
    p = p * r + 1.0;
    p = p * r + 1.0;
    int ki = (int)k;
    if (ki > 1023) {
        p = p * 0x1p1023;
        ki -= 1023;
    } else if (ki < -1022) {
        p = p * 0x1p-1000;
        ki += 1000;
    }
    return p * __longlong_as_double((long long)((uint64_t)(ki + 1023) << 52));
}
__device__ inline double pm_pow(double x, double p) {  // x >= 0
    if (x != x || p != p) return x + p;
    if (x == 0) return p > 0 ? 0.0 : (p == 0 ? 1.0 : INFINITY);
    if (x == INFINITY) return p > 0 ? INFINITY : (p == 0 ? 1.0 : 0.0);
    return pm_exp(p * pm_log(x));
}
// std::min / std::max semantics (first argument on ties and NaN)
__device__ inline double smin(double a, double b) { return (b < a) ? b : a; }
__device__ inline double smax(double a, double b) { return (a < b) ? b : a; }

// ------------------------------------------------------------------ shape geometry
struct SE {
    double c[3], r[3];
    double e, n;
    double R[3][3];
};

__device__ inline SE load_se(const Particles& P, int64_t k) {
    SE s;
    for (int d = 0; d < 3; ++d) {
        s.c[d] = P.pos[d * P.cap + k];
        s.r[d] = P.size[d * P.cap + k];
    }
    s.e = P.sq[k];
    s.n = P.sq[P.cap + k];
    const m3 R = rotation3(P.q[k], P.q[P.cap + k], P.q[2 * P.cap + k], P.q[3 * P.cap + k]);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) s.R[a][b] = R.a[a][b];
    return s;
}

// superellipsoid::point(a, b): superellipsoid.hpp:292-300
__device__ inline void se_point(const SE& s, double a, double b, double* out) {
    const double ca = pm_cos(a), sa = pm_sin(a), cb = pm_cos(b), sb = pm_sin(b);
    double pt[3];
    pt[0] = s.r[0] * sgn(ca) * pm_pow(fabs(ca), s.n) * sgn(cb) * pm_pow(fabs(cb), s.e);
    pt[1] = s.r[1] * sgn(ca) * pm_pow(fabs(ca), s.n) * sgn(sb) * pm_pow(fabs(sb), s.e);
    pt[2] = s.r[2] * sgn(sa) * pm_pow(fabs(sa), s.n);
    for (int d = 0; d < 3; ++d) out[d] = (s.R[d][0] * pt[0] + s.R[d][1] * pt[1] + s.R[d][2] * pt[2]) + s.c[d];
}

// superellipsoid::normal(a, b): superellipsoid.hpp:314-325
__device__ inline void se_normal(const SE& s, double a, double b, double* out) {
    const double ca = pm_cos(a), sa = pm_sin(a), cb = pm_cos(b), sb = pm_sin(b);
    double nn[3];
    nn[0] = s.r[1] * s.r[2] * sgn(ca) * pm_pow(fabs(ca), 2 - s.n) * sgn(cb) * pm_pow(fabs(cb), 2 - s.e);
    nn[1] = s.r[0] * s.r[2] * sgn(ca) * pm_pow(fabs(ca), 2 - s.n) * sgn(sb) * pm_pow(fabs(sb), 2 - s.e);
    nn[2] = s.r[0] * s.r[1] * sgn(sa) * pm_pow(fabs(sa), 2 - s.n);
    double w[3];
    for (int d = 0; d < 3; ++d) w[d] = s.R[d][0] * nn[0] + s.R[d][1] * nn[1] + s.R[d][2] * nn[2];
    const double l = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int d = 0; d < 3; ++d) out[d] = w[d] / l;
}

// binit_xy(6) (superellipsoid.hpp:359-412) when xy, ainit_yz(6) (:415-466) otherwise;
// sorted distinct values (xt::unique).  Returns the count.
template <int num>
__device__ inline int init_angles_n(const SE& s, bool xy, double* out) {
    double v[4 * num];
    int m = 0;
    for (int i = 0; i < num; ++i) {
        const double t_ell = 0.0 + (0.5 * kPi - 0.0) / (num - 1) * i;
        double ang;
        if (xy) {
            const double x_ell = s.r[0] * pm_cos(t_ell);
            const double y_ell = s.r[1] * pm_sin(t_ell);
            const double d = pm_pow(pm_pow(s.r[1] * x_ell, 2.0 / s.e) + pm_pow(s.r[0] * y_ell, 2.0 / s.e), s.e / 2.0);
            const double y_supell = y_ell * s.r[0] * s.r[1] / d;
            const double sinb = smax(-1.0, smin(1.0, sqrt(pm_pow(pm_pow(y_supell / s.r[1], 2), 1.0 / s.e))));
            ang = pm_asin(sinb);
        } else {
            const double y_ell = s.r[1] * pm_cos(t_ell);
            const double z_ell = s.r[2] * pm_sin(t_ell);
            const double d = pm_pow(pm_pow(s.r[2] * y_ell, 2 / s.n) + pm_pow(s.r[1] * z_ell, 2 / s.n), s.n / 2);
            const double z_supell = z_ell * s.r[1] * s.r[2] / d;
            const double sina = smax(-1.0, smin(1.0, sqrt(pm_pow(pm_pow(z_supell / s.r[2], 2), 1 / s.n))));
            ang = pm_asin(sina);
        }
        if (ang > kPi) ang -= 2 * kPi;
        if (ang < -kPi) ang += 2 * kPi;
        v[m++] = ang;
        v[m++] = ang + kPi / 2;
        v[m++] = ang - kPi;
        v[m++] = ang - kPi / 2;
    }
    for (int k = 0; k < m; ++k)
        if (fabs(v[k]) < 1.0e-6) v[k] = 0.01;
    for (int a = 1; a < m; ++a) {  // insertion sort
        const double key = v[a];
        int b = a - 1;
        while (b >= 0 && v[b] > key) {
            v[b + 1] = v[b];
            --b;
        }
        v[b + 1] = key;
    }
    int cnt = 0;
    for (int k = 0; k < m; ++k)
        if (cnt == 0 || !(v[k] == out[cnt - 1])) out[cnt++] = v[k];
    return cnt;
}
__device__ inline int init_angles(const SE& s, bool xy, double* out) { return init_angles_n<6>(s, xy, out); }

// ------------------------------------------------------------------ residual and Jacobian
// F(a1, b1, a2, b2) of closest_points.hpp:282-382 and the reference's hand-written Jacobian
// (:385-603), both in the reference's evaluation order.  The transcendental part (8 cos/sin,
// 16 signed powers and 16 derivative factors) is evaluated once per point; the four powers of
// each |cos| / |sin| share one logarithm, which gives the same bits as separate pow calls.
// D1 (value + one directional derivative) serves the 2D Newton pairs below.
struct D1 {
    double v, d;
};
__device__ inline D1 operator+(const D1& a, const D1& b) { return D1{a.v + b.v, a.d + b.d}; }
__device__ inline D1 operator-(const D1& a, const D1& b) { return D1{a.v - b.v, a.d - b.d}; }
__device__ inline D1 operator-(const D1& a) { return D1{-a.v, -a.d}; }
__device__ inline D1 operator*(const D1& a, const D1& b) { return D1{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline D1 operator*(double s, const D1& a) { return D1{s * a.v, s * a.d}; }
__device__ inline D1 dsqrt(const D1& a) {
    const double r = sqrt(a.v);
    return D1{r, a.d / (2 * r)};
}
__device__ inline double dsqrt(double a) { return sqrt(a); }
template <class V>
__device__ inline V cst(double x);
template <>
__device__ inline double cst<double>(double x) { return x; }
template <>
__device__ inline D1 cst<D1>(double x) { return D1{x, 0.0}; }

// |x|^p, |x|^(2-p) and the derivative factors p |x|^(p-1), (2-p) |x|^(1-p) of one angle function,
// exponents formed as the reference writes them (closest_points.hpp:426-446), sharing log|x|
__device__ inline void pow_pairs(double x, double p, double* v, double* dv) {
    const double ax = fabs(x);
    const double q[2] = {p, 2 - p};
    const double qd[2] = {p - 1, 1 - p};
    if (ax != ax || ax == 0 || ax == INFINITY) {
        for (int k = 0; k < 2; ++k) {
            v[k] = pm_pow(ax, q[k]);
            if (dv) dv[k] = q[k] * pm_pow(ax, qd[k]);
        }
        return;
    }
    const double L = pm_log(ax);
    for (int k = 0; k < 2; ++k) {
        v[k] = pm_exp(q[k] * L);
        if (dv) dv[k] = q[k] * pm_exp(qd[k] * L);
    }
}

// Transcendental terms at u: cs[2m] = cos u_m, cs[2m+1] = sin u_m; powers pw[2f + k] of the
// angle function f (0 cos a1, 1 sin a1, 2 cos b1, 3 sin b1, 4..7 the same for body 2) with
// exponent k (0: the shape exponent n or e, 1: 2 minus it), signed by the function; dp the
// unsigned derivative factors in the same slots.
struct SeTerms {
    double cs[8];
    double pw[16], dp[16];
};
__device__ inline void se_terms(const SE& s1, const SE& s2, const double* u, SeTerms& t, bool deriv) {
    for (int m = 0; m < 4; ++m) {
        t.cs[2 * m] = pm_cos(u[m]);
        t.cs[2 * m + 1] = pm_sin(u[m]);
    }
    // function f: (angle m, cos or sin, exponent): a -> n, b -> e
    const double ex[8] = {s1.n, s1.n, s1.e, s1.e, s2.n, s2.n, s2.e, s2.e};
    for (int f = 0; f < 8; ++f) {
        const double x = t.cs[f];
        pow_pairs(x, ex[f], &t.pw[2 * f], deriv ? &t.dp[2 * f] : nullptr);
        t.pw[2 * f] = sgn(x) * t.pw[2 * f];
        t.pw[2 * f + 1] = sgn(x) * t.pw[2 * f + 1];
    }
}

// Body-frame normal n (A1..A3) and point p (A4..A6) of one body (closest_points.hpp:448-453) and
// the derivative terms pa = A7..A9, na = A10..A12 (along a), pb = A13, A14, nb = A15, A16
// (along b) (:454-463); the second body's B1..B16 (:496-518) are the same forms.
struct SeForms {
    double n[3], p[3], pa[3], na[3], pb[2], nb[2];
};
__device__ inline void se_forms(const SE& s, const SeTerms& t, int body, bool deriv, SeForms& f) {
    const double rx = s.r[0], ry = s.r[1], rz = s.r[2];
    const double* c = &t.cs[4 * body];  // cos a, sin a, cos b, sin b
    const double* w = &t.pw[8 * body];  // can, can2, san, san2, cbe, cbe2, sbe, sbe2
    f.n[0] = ((ry * rz) * w[1]) * w[5];
    f.n[1] = ((rx * rz) * w[7]) * w[1];
    f.n[2] = (rx * ry) * w[3];
    f.p[0] = (rx * w[0]) * w[4];
    f.p[1] = (ry * w[6]) * w[0];
    f.p[2] = rz * w[2];
    if (!deriv) return;
    const double* g = &t.dp[8 * body];  // can1, can3, san1, san3, cbe1, cbe3, sbe1, sbe3
    f.pa[0] = ((rx * c[1]) * g[0]) * w[4];
    f.pa[1] = ((ry * c[1]) * w[6]) * g[0];
    f.pa[2] = (rz * c[0]) * g[2];
    f.na[0] = (((ry * rz) * c[1]) * g[1]) * w[5];
    f.na[1] = (((rx * rz) * c[1]) * w[7]) * g[1];
    f.na[2] = ((rx * ry) * c[0]) * g[3];
    f.pb[0] = ((rx * c[3]) * w[0]) * g[4];
    f.pb[1] = ((ry * c[2]) * g[6]) * w[0];
    f.nb[0] = (((ry * rz) * c[3]) * w[1]) * g[5];
    f.nb[1] = (((rx * rz) * c[2]) * g[7]) * w[1];
}

// F at u, and the Jacobian (row-major) when J is not null.  Each factor of the reference's
// entries is a rotated sum of the forms or its exact negation (negating a sum is exact):
//   pa1 = (Mq0 A7 + Mq1 A8) - Mq2 A9, na1 = (Mq0 A10 + Mq1 A11) - Mq2 A12, pb1 = Mq0 A13 - Mq1 A14,
//   nb1 = Mq0 A15 - Mq1 A16, and pa2, na2, pb2, nb2 the same with N and the B terms;
// every entry is summed left to right as the reference writes it.
__device__ inline void se_residual(const SE& s1, const SE& s2, const double* u, double* F, double* J) {
    const bool deriv = J != nullptr;
    SeTerms t;
    se_terms(s1, s2, u, t, deriv);
    SeForms A, B;
    se_forms(s1, t, 0, deriv, A);
    se_forms(s2, t, 1, deriv, B);
    double n1[3], n2[3], d[3];
    for (int q = 0; q < 3; ++q) {
        n1[q] = (s1.R[q][0] * A.n[0] + s1.R[q][1] * A.n[1]) + s1.R[q][2] * A.n[2];
        n2[q] = (s2.R[q][0] * B.n[0] + s2.R[q][1] * B.n[1]) + s2.R[q][2] * B.n[2];
        d[q] = (((((((-(s1.R[q][0] * A.p[0])) - s1.R[q][1] * A.p[1]) - s1.R[q][2] * A.p[2]) + s2.R[q][0] * B.p[0]) +
                  s2.R[q][1] * B.p[1]) +
                 s2.R[q][2] * B.p[2]) -
                s1.c[q]) +
               s2.c[q];
    }
    F[0] = (-n1[1]) * d[2] + n1[2] * d[1];
    F[1] = n1[0] * d[2] - n1[2] * d[0];
    F[2] = (-n1[0]) * d[1] + n1[1] * d[0];
    const double l1 = sqrt((n1[0] * n1[0] + n1[1] * n1[1]) + n1[2] * n1[2]);
    const double l2 = sqrt((n2[0] * n2[0] + n2[1] * n2[1]) + n2[2] * n2[2]);
    F[3] = ((n1[0] * n2[0] + n1[1] * n2[1]) + n1[2] * n2[2]) + l1 * l2;
    if (!deriv) return;
    // row 3 entry: ((a . -g) lb / la + b . -g), summed left to right
    auto row3 = [](const double* a, const double* b, const double* g, double la, double lb) {
        return (((((a[0] * (-g[0]) + a[1] * (-g[1])) + a[2] * (-g[2])) * lb) / la + b[0] * (-g[0])) + b[1] * (-g[1])) +
               b[2] * (-g[2]);
    };
    double x[3], y[3];
    // column 0 (a1): x = pa1, y = na1
    for (int q = 0; q < 3; ++q) {
        x[q] = (s1.R[q][0] * A.pa[0] + s1.R[q][1] * A.pa[1]) - s1.R[q][2] * A.pa[2];
        y[q] = (s1.R[q][0] * A.na[0] + s1.R[q][1] * A.na[1]) - s1.R[q][2] * A.na[2];
    }
    J[0] = (((-n1[1]) * x[2] + n1[2] * x[1]) + y[1] * d[2]) + (-y[2]) * d[1];
    J[4] = ((n1[0] * x[2] + (-n1[2]) * x[0]) + (-y[0]) * d[2]) + y[2] * d[0];
    J[8] = (((-n1[0]) * x[1] + n1[1] * x[0]) + y[0] * d[1]) + (-y[1]) * d[0];
    J[12] = row3(n1, n2, y, l1, l2);
    // column 1 (b1): x = pb1, y = nb1
    for (int q = 0; q < 3; ++q) {
        x[q] = s1.R[q][0] * A.pb[0] - s1.R[q][1] * A.pb[1];
        y[q] = s1.R[q][0] * A.nb[0] - s1.R[q][1] * A.nb[1];
    }
    J[1] = (((-n1[1]) * x[2] + n1[2] * x[1]) + y[1] * d[2]) + (-y[2]) * d[1];
    J[5] = ((n1[0] * x[2] + (-n1[2]) * x[0]) + (-y[0]) * d[2]) + y[2] * d[0];
    J[9] = (((-n1[0]) * x[1] + n1[1] * x[0]) + y[0] * d[1]) + (-y[1]) * d[0];
    J[13] = row3(n1, n2, y, l1, l2);
    // column 2 (a2): x = pa2, y = na2
    for (int q = 0; q < 3; ++q) {
        x[q] = (s2.R[q][0] * B.pa[0] + s2.R[q][1] * B.pa[1]) - s2.R[q][2] * B.pa[2];
        y[q] = (s2.R[q][0] * B.na[0] + s2.R[q][1] * B.na[1]) - s2.R[q][2] * B.na[2];
    }
    J[2] = (-n1[1]) * (-x[2]) + n1[2] * (-x[1]);
    J[6] = n1[0] * (-x[2]) + (-n1[2]) * (-x[0]);
    J[10] = (-n1[0]) * (-x[1]) + n1[1] * (-x[0]);
    J[14] = row3(n2, n1, y, l2, l1);
    // column 3 (b2): x = pb2, y = nb2
    for (int q = 0; q < 3; ++q) {
        x[q] = s2.R[q][0] * B.pb[0] - s2.R[q][1] * B.pb[1];
        y[q] = s2.R[q][0] * B.nb[0] - s2.R[q][1] * B.nb[1];
    }
    J[3] = (-n1[1]) * (-x[2]) + n1[2] * (-x[1]);
    J[7] = n1[0] * (-x[2]) + (-n1[2]) * (-x[0]);
    J[11] = (-n1[0]) * (-x[1]) + n1[1] * (-x[0]);
    J[15] = row3(n2, n1, y, l2, l1);
}

// ------------------------------------------------------------------ MINPACK hybrd1, n = 4 (3D) or 2 (2D)
// Powell's hybrid method with the analytic Jacobian, as the reference's C++ translation runs it
// (src/minpack.cpp, minpack.hpp:337-923): enorm, qrfac, qform, dogleg, r1updt, r1mpyq.  Where
// that translation departs from the published algorithm (ANL-80-74) it is followed: a plain
// enorm, giant = r8_huge() = 1e30, the first termination test that holds sets info, and dogleg's
// scaled-gradient loop and its (delta/qnorm)(sgnorm/delta)^2 as written (see dogleg).

// enorm: src/minpack.cpp:36-45
__device__ inline double enorm(int n, const double* x) {
    double value = 0.0;
    for (int i = 0; i < n; ++i) value = value + x[i] * x[i];
    return sqrt(value);
}

// the routines for one system size N
template <int N>
struct Minpack {
static constexpr int LR = N * (N + 1) / 2;

static __device__ inline void qrfac(double* a, double* rdiag, double* acnorm) {
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

static __device__ inline void qform(double* q, double* wa) {
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

// dogleg (src/minpack.cpp:101-240).  hybrd1 carves its work arrays out of one block
// (minpack.hpp:880-915: diag, qtf, wa1..wa4, r, fjac) and hybrd passes its wa1, wa2, wa3 as
// dogleg's x, wa1, wa2, so the reference's scaled-gradient loop (:173-182: l = 0, wa1[i - 1] +=
// r[l - 1] temp) reads hybrd's wa4[n - 1] (wa4_last) as r[-1] and updates x[n - 1] as wa1[-1]:
// x[n - 1] gains wa4[n - 1] qtb[0], each wa1[k] (k < n - 1) is accumulated, divided by diag[k] and
// then gains r(k, n - 1) qtb[k + 1], and wa1[n - 1] stays 0.  That loop is kept as written.
static __device__ inline void dogleg(const double* r, const double* diag, const double* qtb, double delta, double* x,
                              double* wa1, double* wa2, double wa4_last) {
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
                temp = smax(temp, fabs(r[l - 1]));
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
    if (qnorm <= delta) return;
    {
        // the reference's loop with its element -1 of wa1 (x[N - 1]) and of r (wa4_last)
        int l = 0;
        for (int j = 0; j < N; ++j) {
            const double temp = qtb[j];
            for (int i = j; i < N; ++i) {
                const double rl = (l >= 1) ? r[l - 1] : wa4_last;
                if (i >= 1)
                    wa1[i - 1] = wa1[i - 1] + rl * temp;
                else
                    x[N - 1] = x[N - 1] + rl * temp;
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
            t = t - ((delta / qnorm) * (sgnorm / delta)) * (sgnorm / delta) +
                sqrt((t - (delta / qnorm)) * (t - (delta / qnorm)) +
                     (1.0 - (delta / qnorm) * (delta / qnorm)) * (1.0 - (sgnorm / delta) * (sgnorm / delta)));
            alpha = ((delta / qnorm) * (1.0 - (sgnorm / delta) * (sgnorm / delta))) / t;
        }
    }
    const double temp = (1.0 - alpha) * smin(sgnorm, delta);
    for (int j = 0; j < N; ++j) x[j] = temp * wa1[j] + alpha * x[j];
}

// r1updt (src/minpack.cpp:373-542); giant = r8_huge() = 1e30 (:60-65, :397)
static __device__ inline void r1updt(double* s, const double* u, double* v, double* w, bool& sing) {
    const double giant = 1.0e30;
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
            if (fabs(v[n - 1]) < fabs(v[j - 1])) {
                const double cotan = v[n - 1] / v[j - 1];
                sn = 0.5 / sqrt(0.25 + 0.25 * cotan * cotan);
                cs = sn * cotan;
                tau = 1.0;
                if (fabs(cs) * giant > 1.0) tau = 1.0 / cs;
            } else {
                const double tn = v[j - 1] / v[n - 1];
                cs = 0.5 / sqrt(0.25 + 0.25 * tn * tn);
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
    for (int i = 1; i <= m; ++i) w[i - 1] += v[n - 1] * u[i - 1];
    sing = false;
    for (int j = 1; j <= n - 1; ++j) {
        if (w[j - 1] != 0.0) {
            double cs, sn, tau;
            if (fabs(s[jj - 1]) < fabs(w[j - 1])) {
                const double cotan = s[jj - 1] / w[j - 1];
                sn = 0.5 / sqrt(0.25 + 0.25 * cotan * cotan);
                cs = sn * cotan;
                tau = 1.0;
                if (fabs(cs) * giant > 1.0) tau = 1.0 / cs;
            } else {
                const double tn = w[j - 1] / s[jj - 1];
                cs = 0.5 / sqrt(0.25 + 0.25 * tn * tn);
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

static __device__ inline void r1mpyq(int m, int n, double* a, int lda, const double* v, const double* w) {
    for (int nmj = 1; nmj <= n - 1; ++nmj) {
        const int j = n - nmj;
        double cs, sn;
        if (fabs(v[j - 1]) > 1.0) {
            cs = 1.0 / v[j - 1];
            sn = sqrt(1.0 - cs * cs);
        } else {
            sn = v[j - 1];
            cs = sqrt(1.0 - sn * sn);
        }
        for (int i = 1; i <= m; ++i) {
            const double temp = cs * a[(i - 1) + lda * (j - 1)] - sn * a[(i - 1) + lda * (n - 1)];
            a[(i - 1) + lda * (n - 1)] = sn * a[(i - 1) + lda * (j - 1)] + cs * a[(i - 1) + lda * (n - 1)];
            a[(i - 1) + lda * (j - 1)] = temp;
        }
    }
    for (int j = 1; j <= n - 1; ++j) {
        double cs, sn;
        if (fabs(w[j - 1]) > 1.0) {
            cs = 1.0 / w[j - 1];
            sn = sqrt(1.0 - cs * cs);
        } else {
            sn = w[j - 1];
            cs = sqrt(1.0 - sn * sn);
        }
        for (int i = 1; i <= m; ++i) {
            const double temp = cs * a[(i - 1) + lda * (j - 1)] + sn * a[(i - 1) + lda * (n - 1)];
            a[(i - 1) + lda * (n - 1)] = -sn * a[(i - 1) + lda * (j - 1)] + cs * a[(i - 1) + lda * (n - 1)];
            a[(i - 1) + lda * (j - 1)] = temp;
        }
    }
}

// hybrd1 with the analytic Jacobian (info 1 converged, 2 maxfev, 3 xtol too small, 4 no
// progress; 5 is folded into 4 as minpack.hpp:918-921 does), as a resumable per-lane state:
// hybrd's outer (Jacobian) and inner (trial step) loops become passes with one residual
// evaluation each (phase 0: F at the start, 1: F and J at x, 2: F at the trial point), so a
// lane can stop after any pass and take another pair; the arithmetic is the loops' own,
// operation for operation.
struct State {
    double x[N], fvec[N], fjac[N * N], r[LR], qtf[N], wa1[N], wa2[N], wa3[N], wa4[N], Jrow[N * N], ftmp[N];
    double delta, xnorm, fnorm, pnorm;
    int info, nfev, iter, ncsuc, ncfail, nslow1, nslow2, phase;
    bool jeval;
};

static __device__ inline void start(State& h, const double* u0) {
    for (int k = 0; k < N; ++k) {
        h.x[k] = u0[k];
        h.wa4[k] = 0.0;  // fresh work block: read by dogleg only on a scaled-gradient first step
    }
    h.delta = h.xnorm = h.fnorm = h.pnorm = 0.0;
    h.info = h.nfev = 0;
    h.iter = 1;
    h.ncsuc = h.ncfail = h.nslow1 = h.nslow2 = 0;
    h.phase = 0;
    h.jeval = true;
}

// one pass; returns true when hybrd1 has ended (h.info, h.nfev, h.x hold its results)
template <class Res>
static __device__ inline bool pass(Res residual, State& h) {
    const double xtol = 1.0e-10, factor = 100.0, epsmch = 2.220446049250313e-16;
    const int maxfev = 200 * (N + 1);
    const double p1 = 0.1, p5 = 0.5, p001 = 0.001, p0001 = 0.0001;
    double diag[N];  // mode 2: wa[0..n) = 1 (minpack.hpp:881-884)
    for (int k = 0; k < N; ++k) diag[k] = 1.0;
    const int msum = N;
    residual(h.phase == 2 ? h.wa2 : h.x, h.phase == 0 ? h.fvec : (h.phase == 1 ? h.ftmp : h.wa4),
             h.phase == 1 ? h.Jrow : nullptr);
    if (h.phase == 0) {
        h.nfev = 1;
        h.fnorm = enorm(N, h.fvec);
        h.phase = 1;
        return false;
    }
    if (h.phase == 1) {
        // top of hybrd's outer loop: factor the Jacobian at x
        h.jeval = true;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) h.fjac[i + N * j] = h.Jrow[N * i + j];
        h.nfev += msum;
        qrfac(h.fjac, h.wa1, h.wa2);
        if (h.iter == 1) {
            for (int j = 0; j < N; ++j) h.wa3[j] = diag[j] * h.x[j];
            h.xnorm = enorm(N, h.wa3);
            h.delta = (h.xnorm == 0.0) ? factor : factor * h.xnorm;
        }
        for (int i = 0; i < N; ++i) h.qtf[i] = h.fvec[i];
        for (int j = 0; j < N; ++j) {
            if (h.fjac[j + N * j] != 0.0) {
                double sum = 0.0;
                for (int i = j; i < N; ++i) sum += h.fjac[i + N * j] * h.qtf[i];
                const double temp = -sum / h.fjac[j + N * j];
                for (int i = j; i < N; ++i) h.qtf[i] += h.fjac[i + N * j] * temp;
            }
        }
        for (int j = 0; j < N; ++j) {
            int l = j;
            for (int i = 0; i < j; ++i) {
                h.r[l] = h.fjac[i + N * j];
                l += N - i - 1;
            }
            h.r[l] = h.wa1[j];
        }
        qform(h.fjac, h.wa1);
    } else {
        // inner loop, after the residual at the trial point
        ++h.nfev;
        const double fnorm1 = enorm(N, h.wa4);
        double actred = -1.0;
        if (fnorm1 < h.fnorm) actred = 1.0 - (fnorm1 / h.fnorm) * (fnorm1 / h.fnorm);
        int l = 0;
        for (int i = 0; i < N; ++i) {
            double sum = 0.0;
            for (int j = i; j < N; ++j) {
                sum += h.r[l] * h.wa1[j];
                ++l;
            }
            h.wa3[i] = h.qtf[i] + sum;
        }
        const double temp = enorm(N, h.wa3);
        double prered = 0.0;
        if (temp < h.fnorm) prered = 1.0 - (temp / h.fnorm) * (temp / h.fnorm);
        double ratio = 0.0;
        if (prered > 0.0) ratio = actred / prered;
        if (ratio < p1) {
            h.ncsuc = 0;
            ++h.ncfail;
            h.delta = p5 * h.delta;
        } else {
            h.ncfail = 0;
            ++h.ncsuc;
            if (ratio >= p5 || h.ncsuc > 1) h.delta = smax(h.delta, h.pnorm / p5);
            if (fabs(ratio - 1.0) <= p1) h.delta = h.pnorm / p5;
        }
        if (ratio >= p0001) {
            for (int j = 0; j < N; ++j) {
                h.x[j] = h.wa2[j];
                h.wa2[j] = diag[j] * h.x[j];
                h.fvec[j] = h.wa4[j];
            }
            h.xnorm = enorm(N, h.wa2);
            h.fnorm = fnorm1;
            ++h.iter;
        }
        ++h.nslow1;
        if (actred >= p001) h.nslow1 = 0;
        if (h.jeval) ++h.nslow2;
        if (actred >= p1) h.nslow2 = 0;
        // the first test that holds ends hybrd (minpack.hpp:760-791)
        if (h.delta <= xtol * h.xnorm || h.fnorm == 0.0)
            h.info = 1;
        else if (maxfev <= h.nfev)
            h.info = 2;
        else if (p1 * smax(p1 * h.delta, h.pnorm) <= epsmch * h.xnorm)
            h.info = 3;
        else if (h.nslow2 == 5)
            h.info = 4;
        else if (h.nslow1 == 10)
            h.info = 5;
        if (h.info != 0) {
            if (h.info == 5) h.info = 4;
            return true;
        }
        if (h.ncfail == 2) {
            h.phase = 1;  // back to the outer loop: a new Jacobian at x
            return false;
        }
        for (int j = 0; j < N; ++j) {
            double sum = 0.0;
            for (int i = 0; i < N; ++i) sum += h.fjac[i + N * j] * h.wa4[i];
            h.wa2[j] = (sum - h.wa3[j]) / h.pnorm;
            h.wa1[j] = diag[j] * ((diag[j] * h.wa1[j]) / h.pnorm);
            if (ratio >= p0001) h.qtf[j] = sum;
        }
        bool sg = false;
        r1updt(h.r, h.wa1, h.wa2, h.wa3, sg);
        r1mpyq(N, N, h.fjac, N, h.wa2, h.wa3);
        r1mpyq(1, N, h.qtf, 1, h.wa2, h.wa3);
        h.jeval = false;
    }
    // propose the next trial point (head of hybrd's inner loop)
    dogleg(h.r, diag, h.qtf, h.delta, h.wa1, h.wa2, h.wa3, h.wa4[N - 1]);
    for (int j = 0; j < N; ++j) {
        h.wa1[j] = -h.wa1[j];
        h.wa2[j] = h.x[j] + h.wa1[j];
        h.wa3[j] = diag[j] * h.wa1[j];
    }
    h.pnorm = enorm(N, h.wa3);
    if (h.iter == 1) h.delta = smin(h.delta, h.pnorm);
    h.phase = 2;
    return false;
}
};

// ------------------------------------------------------------------ sphere / plane against a superellipsoid
// closest_points(sphere<3>, superellipsoid<3>) (closest_points.hpp:944-1281) and
// closest_points(superellipsoid<3>, plane<3>) (:1478-1746): damped Newton (utils.hpp:81-118,
// 200 iterations, ftol 1e-10, xtol 1e-7) from the first minimum of a 10 x 10 distance table over
// the diagonal (a_i, b_i) of linspace(-pi/2, pi/2, 10) x linspace(-pi, pi, 10).  s1 is the
// superellipsoid; its normal is written sign(cos a1) superellipsoid::normal, and GradF is the chain
// rule of F with the reference's partial derivatives: for the sphere pair the derivative of
// |cos a1|^(2-n) lacks the sign(cos a1) factor (ca1n3, :1088), for the plane pair it has it (:1617).
// The plane pair's `sign_p2s` is the dot product (c_s1 - c_p) . n_p itself (:1484-1493).
struct Sph {
    double c[3], r;
    double R[3][3];
};
struct Pln {
    double c[3];
    double R[3][3];
};

__device__ inline void load_rot(const Particles& P, int64_t k, double (&R)[3][3]) {
    const m3 M = rotation3(P.q[k], P.q[P.cap + k], P.q[2 * P.cap + k], P.q[3 * P.cap + k]);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) R[a][b] = M.a[a][b];
}
__device__ inline Sph load_sph(const Particles& P, int64_t k) {
    Sph s;
    for (int d = 0; d < 3; ++d) s.c[d] = P.pos[d * P.cap + k];
    s.r = P.size[k];
    load_rot(P, k, s.R);
    return s;
}
__device__ inline Pln load_pln(const Particles& P, int64_t k) {
    Pln p;
    for (int d = 0; d < 3; ++d) p.c[d] = P.pos[d * P.cap + k];
    load_rot(P, k, p.R);
    return p;
}

// sphere::point / normal (a, b): sphere.hpp:203-234; plane::point(a, b) = R (0, a, b) + c
__device__ inline void sph_point(const Sph& s, double a, double b, double* out) {
    const double pt[3] = {s.r * pm_cos(a) * pm_cos(b), s.r * pm_cos(a) * pm_sin(b), s.r * pm_sin(a)};
    for (int d = 0; d < 3; ++d) out[d] = (s.R[d][0] * pt[0] + s.R[d][1] * pt[1] + s.R[d][2] * pt[2]) + s.c[d];
}
__device__ inline void sph_normal(const Sph& s, double a, double b, double* out) {
    const double nn[3] = {pm_cos(a) * pm_cos(b), pm_cos(a) * pm_sin(b), pm_sin(a)};
    double w[3];
    for (int d = 0; d < 3; ++d) w[d] = s.R[d][0] * nn[0] + s.R[d][1] * nn[1] + s.R[d][2] * nn[2];
    const double l = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int d = 0; d < 3; ++d) out[d] = w[d] / l;
}
__device__ inline void pln_point(const Pln& p, double a, double b, double* out) {
    const double pt[3] = {0.0, a, b};
    for (int d = 0; d < 3; ++d) out[d] = (p.R[d][0] * pt[0] + p.R[d][1] * pt[1] + p.R[d][2] * pt[2]) + p.c[d];
}

// transcendental terms of the superellipsoid side at (a1, b1): cos/sin of the four angles, and
// the powers of cos a1, sin a1, cos b1, sin b1 with their derivative factors; the second power of
// cos a1 is unsigned, its derivative factor with sign(cos a1) when `exact`.
struct NtTerms {
    double cs[8];
    double pw[8], dp[8];
};
__device__ inline void nt_terms(const SE& s1, const double* u, bool exact, NtTerms& t, bool deriv) {
    for (int m = 0; m < 4; ++m) {
        t.cs[2 * m] = pm_cos(u[m]);
        t.cs[2 * m + 1] = pm_sin(u[m]);
    }
    const double ex[4] = {s1.n, s1.n, s1.e, s1.e};
    for (int f = 0; f < 4; ++f) {
        const double x = t.cs[f];
        double v[2], dv[2];
        pow_pairs(x, ex[f], v, deriv ? dv : nullptr);
        t.pw[2 * f] = sgn(x) * v[0];
        t.pw[2 * f + 1] = (f == 0) ? v[1] : sgn(x) * v[1];
        if (deriv) {
            t.dp[2 * f] = dv[0];
            t.dp[2 * f + 1] = dv[1];
            if (f == 0 && exact) {  // ((2-n) sign(x)) |x|^(1-n) instead of (2-n) |x|^(1-n)
                const double p = 2 - s1.n, pd = 1 - s1.n, ax = fabs(x);
                t.dp[1] = (ax != ax || ax == 0 || ax == INFINITY) ? (p * sgn(x)) * pm_pow(ax, pd)
                                                                 : (p * sgn(x)) * pm_exp(pd * pm_log(ax));
            }
        }
    }
}

// F of the sphere pair (kind 0) or the plane pair (kind 1); V = double (values) or D1 (direction k)
template <class V>
__device__ inline void nt_algebra(int kind, const SE& s1, const Sph& s2, const Pln& p2, double sp, const NtTerms& t,
                                  const double* u, int k, V* F) {
    V fn[8];
    for (int m = 0; m < 4; ++m) {
        if constexpr (sizeof(V) == sizeof(D1)) {
            const double seed = (m == k) ? 1.0 : 0.0;
            fn[2 * m] = D1{t.cs[2 * m], -t.cs[2 * m + 1] * seed};
            fn[2 * m + 1] = D1{t.cs[2 * m + 1], t.cs[2 * m] * seed};
        } else {
            fn[2 * m] = cst<V>(t.cs[2 * m]);
            fn[2 * m + 1] = cst<V>(t.cs[2 * m + 1]);
        }
    }
    auto P = [&](int f, int e) -> V {
        if constexpr (sizeof(V) == sizeof(D1))
            return D1{t.pw[2 * f + e], t.dp[2 * f + e] * fn[f].d};
        else
            return t.pw[2 * f + e];
    };
    const V ca1n = P(0, 0), ca1n2 = P(0, 1), sa1n = P(1, 0), sa1n2 = P(1, 1);
    const V cb1e = P(2, 0), cb1e2 = P(2, 1), sb1e = P(3, 0), sb1e2 = P(3, 1);
    const double sgc = sgn(t.cs[0]);
    const V A1 = ((s1.r[1] * s1.r[2]) * ca1n2) * cb1e2, A2 = ((s1.r[0] * s1.r[2]) * sb1e2) * ca1n2,
            A3 = sgc * ((s1.r[0] * s1.r[1]) * sa1n2);
    const V A4 = (s1.r[0] * ca1n) * cb1e, A5 = (s1.r[1] * sb1e) * ca1n, A6 = s1.r[2] * sa1n;
    V n1[3], d[3];
    for (int q = 0; q < 3; ++q) n1[q] = (s1.R[q][0] * A1 + s1.R[q][1] * A2) + s1.R[q][2] * A3;
    V f3;
    if (kind == 0) {
        const V ca2 = fn[4], sa2 = fn[5], cb2 = fn[6], sb2 = fn[7];
        V ca2n2;
        if constexpr (sizeof(V) == sizeof(D1))
            ca2n2 = D1{fabs(ca2.v), sgn(ca2.v) * ca2.d};
        else
            ca2n2 = fabs(ca2);
        const double r = s2.r, sg2 = sgn(t.cs[4]);
        const V B1 = (r * ca2) * cb2, B2 = (r * sb2) * ca2, B3 = r * sa2;
        const V B4 = ((r * r) * ca2n2) * cb2, B5 = ((r * r) * sb2) * ca2n2, B6 = sg2 * ((r * r) * sa2);
        V n2[3];
        for (int q = 0; q < 3; ++q) {
            n2[q] = (s2.R[q][0] * B4 + s2.R[q][1] * B5) + s2.R[q][2] * B6;
            d[q] = (((((((-(s1.R[q][0] * A4)) - s1.R[q][1] * A5) - s1.R[q][2] * A6) + s2.R[q][0] * B1) +
                      s2.R[q][1] * B2) +
                     s2.R[q][2] * B3) -
                    cst<V>(s1.c[q])) +
                   cst<V>(s2.c[q]);
        }
        const V dotn = (n1[0] * n2[0] + n1[1] * n2[1]) + n1[2] * n2[2];
        const V l1 = dsqrt((n1[0] * n1[0] + n1[1] * n1[1]) + n1[2] * n1[2]);
        const V l2 = dsqrt((n2[0] * n2[0] + n2[1] * n2[1]) + n2[2] * n2[2]);
        f3 = dotn + l1 * l2;
    } else {
        V a2, b2;
        if constexpr (sizeof(V) == sizeof(D1)) {
            a2 = D1{u[2], k == 2 ? 1.0 : 0.0};
            b2 = D1{u[3], k == 3 ? 1.0 : 0.0};
        } else {
            a2 = u[2];
            b2 = u[3];
        }
        for (int q = 0; q < 3; ++q)
            d[q] = ((((((-(s1.R[q][0] * A4)) - s1.R[q][1] * A5) - s1.R[q][2] * A6) + p2.R[q][1] * a2) +
                     p2.R[q][2] * b2) +
                    cst<V>(p2.c[q])) -
                   cst<V>(s1.c[q]);
        const V dotn = ((p2.R[0][0] * sp) * n1[0] + (p2.R[1][0] * sp) * n1[1]) + (p2.R[2][0] * sp) * n1[2];
        const double lp = sqrt((p2.R[0][0] * p2.R[0][0] + p2.R[1][0] * p2.R[1][0]) + p2.R[2][0] * p2.R[2][0]);
        const V l1 = dsqrt((n1[0] * n1[0] + n1[1] * n1[1]) + n1[2] * n1[2]);
        f3 = dotn + lp * l1;
    }
    F[0] = (-(n1[1] * d[2])) + n1[2] * d[1];
    F[1] = n1[0] * d[2] - n1[2] * d[0];
    F[2] = (-(n1[0] * d[1])) + n1[1] * d[0];
    F[3] = f3;
}

__device__ inline void nt_residual(int kind, const SE& s1, const Sph& s2, const Pln& p2, double sp, const double* u,
                                   double* F, double* J) {
    NtTerms t;
    nt_terms(s1, u, kind == 1, t, J != nullptr);
    if (!J) {
        nt_algebra<double>(kind, s1, s2, p2, sp, t, u, -1, F);
        return;
    }
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
        D1 f[4];
        nt_algebra<D1>(kind, s1, s2, p2, sp, t, u, k, f);
        for (int i = 0; i < 4; ++i) {
            J[4 * i + k] = f[i].d;
            if (k == 0) F[i] = f[i].v;
        }
    }
}

// xt::linalg::solve: LU with partial pivoting (first largest |pivot|), A row-major, b -> x
__device__ __attribute__((noinline)) void lu_solve4(double* A, double* b) {
    int piv[4];
    for (int k = 0; k < 4; ++k) {
        int p = k;
        double best = fabs(A[4 * k + k]);
        for (int i = k + 1; i < 4; ++i)
            if (fabs(A[4 * i + k]) > best) {
                best = fabs(A[4 * i + k]);
                p = i;
            }
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < 4; ++j) {
                const double tmp = A[4 * k + j];
                A[4 * k + j] = A[4 * p + j];
                A[4 * p + j] = tmp;
            }
        if (A[4 * k + k] != 0.0) {
            const double inv = 1.0 / A[4 * k + k];
            for (int i = k + 1; i < 4; ++i) A[4 * i + k] *= inv;
        }
        for (int i = k + 1; i < 4; ++i)
            for (int j = k + 1; j < 4; ++j) A[4 * i + j] -= A[4 * i + k] * A[4 * k + j];
    }
    for (int k = 0; k < 4; ++k)
        if (piv[k] != k) {
            const double tmp = b[k];
            b[k] = b[piv[k]];
            b[piv[k]] = tmp;
        }
    for (int i = 1; i < 4; ++i)
        for (int j = 0; j < i; ++j) b[i] -= A[4 * i + j] * b[j];
    for (int i = 3; i >= 0; --i) {
        for (int j = i + 1; j < 4; ++j) b[i] -= A[4 * i + j] * b[j];
        b[i] /= A[4 * i + i];
    }
}

__device__ inline double norm4(const double* x) {
    return sqrt(((x[0] * x[0] + x[1] * x[1]) + x[2] * x[2]) + x[3] * x[3]);
}

// newton_method (utils.hpp:81-118): iteration count, -1 at itermax.  One residual call
// site per pass (phase 0: F and J at u, 1: F at the line-search point).
__device__ inline int newton(int kind, const SE& s1, const Sph& s2, const Pln& p2, double sp, double* u) {
    const int itermax = 200;
    const double ftol = 1.0e-10, xtol = 1.0e-7;
    int cc = 0;
    double F[4], J[16], dd[4], ut[4], Ft[4], ferr = 0.0, t = 1.0;
    int phase = 0;
    while (cc < itermax) {
        nt_residual(kind, s1, s2, p2, sp, phase == 0 ? u : ut, phase == 0 ? F : Ft, phase == 0 ? J : nullptr);
        if (phase == 0) {
            for (int q = 0; q < 4; ++q) dd[q] = -F[q];
            lu_solve4(J, dd);
            if (norm4(dd) < xtol) return cc;
            t = 1;
            ferr = norm4(F);
            if (ferr < ftol) return cc;
        } else if (norm4(Ft) > ferr && t > 0.01) {
            t -= 0.01;
        } else {
            for (int q = 0; q < 4; ++q) u[q] += t * dd[q];
            cc += 1;
            phase = 0;
            continue;
        }
        for (int q = 0; q < 4; ++q) ut[q] = u[q] + t * dd[q];
        phase = 1;
    }
    return -1;
}

__device__ inline double lin(double lo, double hi, int n, int i) { return lo + (hi - lo) / (n - 1) * i; }

// first minimum of the 10 x 10 diagonal distance table; kind 0: sphere s2, 1: plane p2
__device__ inline void diag_grid_start(int kind, const SE& s1, const Sph& s2, const Pln& p2, double* u0) {
    const int num = 10;
    double dmin = INFINITY;
    int imin = 0, jmin = 0;
    for (int i = 0; i < num; ++i) {
        double p[3];
        se_point(s1, lin(-kPi / 2, kPi / 2, num, i), lin(-kPi, kPi, num, i), p);
        for (int j = 0; j < num; ++j) {
            double q[3];
            if (kind == 0)
                sph_point(s2, lin(-kPi / 2, kPi / 2, num, j), lin(-kPi, kPi, num, j), q);
            else
                pln_point(p2, lin(-kPi / 2, kPi / 2, num, j), lin(-kPi, kPi, num, j), q);
            const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
            const double dist = sqrt(dx * dx + dy * dy + dz * dz);
            if (dist < dmin) {
                dmin = dist;
                imin = i;
                jmin = j;
            }
        }
    }
    u0[0] = lin(-kPi / 2, kPi / 2, num, imin);
    u0[1] = lin(-kPi, kPi, num, imin);
    u0[2] = lin(-kPi / 2, kPi / 2, num, jmin);
    u0[3] = lin(-kPi, kPi, num, jmin);
}

// One sphere / plane - superellipsoid pair in dispatcher order: writes the record (pi, pj, nij,
// dij) as the reference returns it for (object i, object j), the start and end points and the
// Newton count.
__device__ inline void newton_pair(const Particles& P, int64_t i, int64_t j, double* pi, double* pj, double* nij,
                                   double& dij, double* u0, double* u, int& its) {
    const int si = P.shape[i], sj = P.shape[j];
    const bool se_i = si == SCOPI_SHAPE_SUPERELLIPSOID;
    const int64_t kse = se_i ? j : i;  // the other object
    const int64_t ks = se_i ? i : j;
    const int other = se_i ? sj : si;
    const int kind = (other == SCOPI_SHAPE_SPHERE) ? 0 : 1;
    // reference order: (sphere, SE) and (SE, plane) directly, the others swapped afterwards
    const bool flip = (kind == 0) ? se_i : !se_i;
    const SE s1 = load_se(P, ks);
    Sph s2{};
    Pln p2{};
    double sp = 0.0;
    if (kind == 0) {
        s2 = load_sph(P, kse);
    } else {
        p2 = load_pln(P, kse);
        sp = (s1.c[0] - p2.c[0]) * p2.R[0][0] + (s1.c[1] - p2.c[1]) * p2.R[1][0] + (s1.c[2] - p2.c[2]) * p2.R[2][0];
    }
    diag_grid_start(kind, s1, s2, p2, u0);
    for (int q = 0; q < 4; ++q) u[q] = u0[q];
    its = newton(kind, s1, s2, p2, sp, u);
    double a[3], b[3], n[3];
    se_point(s1, u[0], u[1], a);
    if (kind == 0) {
        sph_point(s2, u[2], u[3], b);
        sph_normal(s2, u[2], u[3], n);
    } else {
        pln_point(p2, u[2], u[3], b);
        for (int q = 0; q < 3; ++q) n[q] = p2.R[q][0];
    }
    dij = (a[0] - b[0]) * n[0] + (a[1] - b[1]) * n[1] + (a[2] - b[2]) * n[2];
    for (int q = 0; q < 3; ++q) {
        pi[q] = flip ? b[q] : a[q];
        pj[q] = flip ? a[q] : b[q];
        nij[q] = flip ? n[q] * -1. : n[q];
    }
}

// ------------------------------------------------------------------ kernels
// ------------------------------------------------------------------ 2D: superellipsoid<2>
// closest_points(superellipsoid<2>, superellipsoid<2>) (closest_points.hpp:64-249): hybrd1 on the
// 2 x 2 system (normal of s1 collinear with p2 - p1, normals anti-parallel) from the first
// minimum, row-major, of |p1 - p2| + 2 (1 + n1 . n2) over unique(binit_xy(4)) of each side;
// closest_points(sphere<2>, superellipsoid<2>) (:1316-1448) and (superellipsoid<2>, plane<2>)
// (:1782-1949): damped Newton from a distance table.  The objects use the 2 x 2 block of the
// rotation (rotation3 of a 2D quaternion holds rotation_matrix<2> there, bit for bit), the
// superellipsoid its squareness e.  Jacobians: the reference's hand-written one for the
// superellipsoid pair (se2_se2_residual); forward mode, one direction at a time, for the Newton
// pairs, whose solve (LAPACK gesv, BLAS dnrm2 in the reference) is restated, not reproduced.
__device__ inline void se2_point(const SE& s, double b, double* out) {  // superellipsoid.hpp:281-289
    const double cb = pm_cos(b), sb = pm_sin(b);
    const double x = s.r[0] * sgn(cb) * pm_pow(fabs(cb), s.e);
    const double y = s.r[1] * sgn(sb) * pm_pow(fabs(sb), s.e);
    out[0] = (s.R[0][0] * x + s.R[0][1] * y) + s.c[0];
    out[1] = (s.R[1][0] * x + s.R[1][1] * y) + s.c[1];
    out[2] = 0.0;
}
__device__ inline void se2_normal(const SE& s, double b, double* out) {  // superellipsoid.hpp:303-312
    const double cb = pm_cos(b), sb = pm_sin(b);
    const double x = s.r[1] * sgn(cb) * pm_pow(fabs(cb), 2 - s.e);
    const double y = s.r[0] * sgn(sb) * pm_pow(fabs(sb), 2 - s.e);
    const double n0 = s.R[0][0] * x + s.R[0][1] * y, n1 = s.R[1][0] * x + s.R[1][1] * y;
    const double l = sqrt(n0 * n0 + n1 * n1);
    out[0] = n0 / l;
    out[1] = n1 / l;
    out[2] = 0.0;
}
__device__ inline void sph2_point(const Sph& s, double b, double* out) {  // sphere.hpp:194-201
    const double x = s.r * pm_cos(b), y = s.r * pm_sin(b);
    out[0] = (s.R[0][0] * x + s.R[0][1] * y) + s.c[0];
    out[1] = (s.R[1][0] * x + s.R[1][1] * y) + s.c[1];
    out[2] = 0.0;
}
__device__ inline void sph2_normal(const Sph& s, double b, double* out) {  // sphere.hpp:213-221
    const double x = s.r * pm_cos(b), y = s.r * pm_sin(b);
    const double n0 = s.R[0][0] * x + s.R[0][1] * y, n1 = s.R[1][0] * x + s.R[1][1] * y;
    const double l = sqrt(n0 * n0 + n1 * n1);
    out[0] = n0 / l;
    out[1] = n1 / l;
    out[2] = 0.0;
}
__device__ inline void pln2_point(const Pln& p, double b, double* out) {  // plane.hpp:175-182
    out[0] = (p.R[0][0] * 0.0 + p.R[0][1] * b) + p.c[0];
    out[1] = (p.R[1][0] * 0.0 + p.R[1][1] * b) + p.c[1];
    out[2] = 0.0;
}

// cos / sin of an angle as a value (V = double) or with its derivative along a seed direction
template <class V>
__device__ inline V v_cos(double b, double seed) {
    if constexpr (sizeof(V) == sizeof(D1))
        return D1{pm_cos(b), -pm_sin(b) * seed};
    else
        return pm_cos(b);
}
template <class V>
__device__ inline V v_sin(double b, double seed) {
    if constexpr (sizeof(V) == sizeof(D1))
        return D1{pm_sin(b), pm_cos(b) * seed};
    else
        return pm_sin(b);
}
template <class V>
__device__ inline V v_var(double x, double seed) {
    if constexpr (sizeof(V) == sizeof(D1))
        return D1{x, seed};
    else
        return x;
}
// sign(x) |x|^p, derivative p |x|^(p-1) x'
__device__ inline double v_spow(double a, double p) { return sgn(a) * pm_pow(fabs(a), p); }
__device__ inline D1 v_spow(const D1& a, double p) {
    const double ax = fabs(a.v);
    return D1{sgn(a.v) * pm_pow(ax, p), (p * pm_pow(ax, p - 1)) * a.d};
}
// sign(x)|x|^(2-s), derivative (2-s)|x|^(1-s) x' with the exponent 1 - s as the reference writes it
__device__ inline double v_spow2(double a, double s) { return sgn(a) * pm_pow(fabs(a), 2 - s); }
__device__ inline D1 v_spow2(const D1& a, double s) {
    const double ax = fabs(a.v);
    return D1{sgn(a.v) * pm_pow(ax, 2 - s), ((2 - s) * pm_pow(ax, 1 - s)) * a.d};
}
__device__ inline double vval(double a) { return a; }
__device__ inline double vval(const D1& a) { return a.v; }
__device__ inline double vder(double) { return 0.0; }
__device__ inline double vder(const D1& a) { return a.d; }

// one superellipsoid side in the body frame: normal direction (r_y c^(2-e), r_x s^(2-e)), point
// (r_x c^e, r_y s^e)
template <class V>
__device__ inline void se2_terms(const SE& s, double b, double seed, V& nx, V& ny, V& px, V& py) {
    const V cb = v_cos<V>(b, seed), sb = v_sin<V>(b, seed);
    nx = s.r[1] * v_spow2(cb, s.e);
    ny = s.r[0] * v_spow2(sb, s.e);
    px = s.r[0] * v_spow(cb, s.e);
    py = s.r[1] * v_spow(sb, s.e);
}

// residual of closest_points.hpp:1332-1366 (kind 0: sphere s2 with superellipsoid s1) or
// :1800-1838 (kind 1: superellipsoid s1 with plane p2, sp = (c_s1 - c_p) . n_p)
template <class V>
__device__ inline void se2_algebra(int kind, const SE& s1, const SE& s2, const Sph& sp2, const Pln& p2, double sp,
                                   const double* u, int k, V* F) {
    V n1x, n1y, p1x, p1y;
    se2_terms<V>(s1, u[0], k == 0 ? 1.0 : 0.0, n1x, n1y, p1x, p1y);
    const double seed2 = (k == 1) ? 1.0 : 0.0;
    const auto& A = s1.R;
    if (kind == 0) {
        const V q5 = sp2.r * v_sin<V>(u[1], seed2), q6 = sp2.r * v_cos<V>(u[1], seed2);
        const auto& B = sp2.R;
        const V m1x = A[0][0] * n1x + A[0][1] * n1y, m1y = A[1][0] * n1x + A[1][1] * n1y;
        const V m2x = B[0][0] * q6 + B[0][1] * q5, m2y = B[1][0] * q6 + B[1][1] * q5;
        const V dy =
            ((((-(A[1][0] * p1x)) - A[1][1] * p1y) + B[1][0] * q6) + B[1][1] * q5 - cst<V>(s1.c[1])) + cst<V>(sp2.c[1]);
        const V dx =
            ((((-(A[0][0] * p1x)) - A[0][1] * p1y) + B[0][0] * q6) + B[0][1] * q5 - cst<V>(s1.c[0])) + cst<V>(sp2.c[0]);
        F[0] = (-m1x) * dy + m1y * dx;
        F[1] = (m2x * m1x + m2y * m1y) + dsqrt(m2x * m2x + m2y * m2y) * dsqrt(m1x * m1x + m1y * m1y);
    } else {
        const V a2 = v_var<V>(u[1], seed2);
        const auto& N = p2.R;
        const V m1x = A[0][0] * n1x + A[0][1] * n1y, m1y = A[1][0] * n1x + A[1][1] * n1y;
        const V dy = ((((-(A[1][0] * p1x)) - A[1][1] * p1y) + N[1][1] * a2) + cst<V>(p2.c[1])) - cst<V>(s1.c[1]);
        const V dx = ((((-(A[0][0] * p1x)) - A[0][1] * p1y) + N[0][1] * a2) + cst<V>(p2.c[0])) - cst<V>(s1.c[0]);
        F[0] = (-m1x) * dy + m1y * dx;
        const double lp = sqrt(N[0][0] * N[0][0] + N[1][0] * N[1][0]);
        F[1] = ((N[0][0] * sp) * m1x + (N[1][0] * sp) * m1y) + lp * dsqrt(m1x * m1x + m1y * m1y);
    }
}

struct Se2Pair {
    int kind;  // 0 sphere s2 / superellipsoid s1, 1 superellipsoid s1 / plane p2, 2 superellipsoid pair
    SE s1, s2;
    Sph sp2;
    Pln p2;
    double sp;
};

// SE2-SE2 residual (closest_points.hpp:79-125) with the reference's hand-written Jacobian
// (:127-189).  Body terms F1 = rx sb^(2-e), F2 = ry sb^e, F3 = ry cb^(2-e), F4 = rx cb^e (signed
// powers) and derivative terms F9 = e rx sb |cb|^(e-1), F10 = e ry cb |sb|^(e-1),
// F11 = ry (2-e) sb |cb|^(1-e), F12 = rx (2-e) cb |sb|^(1-e) (F5..F8, F13..F16 for the second
// body); the factors of the entries are rotated sums of these or their exact negations, summed
// left to right as written.
struct Se2Side {
    double f1, f2, f3, f4, f9, f10, f11, f12;
};
__device__ inline Se2Side se2_side(const SE& s, double b, bool deriv) {
    const double cb = pm_cos(b), sb = pm_sin(b), acb = fabs(cb), asb = fabs(sb);
    const double cbe = sgn(cb) * pm_pow(acb, s.e), cbe2 = sgn(cb) * pm_pow(acb, 2 - s.e);
    const double sbe = sgn(sb) * pm_pow(asb, s.e), sbe2 = sgn(sb) * pm_pow(asb, 2 - s.e);
    Se2Side t{s.r[0] * sbe2, s.r[1] * sbe, s.r[1] * cbe2, s.r[0] * cbe, 0.0, 0.0, 0.0, 0.0};
    if (deriv) {
        t.f9 = ((s.e * s.r[0]) * sb) * pm_pow(acb, s.e - 1);
        t.f10 = ((s.e * s.r[1]) * cb) * pm_pow(asb, s.e - 1);
        t.f11 = ((s.r[1] * (2 - s.e)) * sb) * pm_pow(acb, 1 - s.e);
        t.f12 = ((s.r[0] * (2 - s.e)) * cb) * pm_pow(asb, 1 - s.e);
    }
    return t;
}
__device__ inline void se2_se2_residual(const SE& s1, const SE& s2, const double* u, double* F, double* J) {
    const Se2Side a = se2_side(s1, u[0], J != nullptr), b = se2_side(s2, u[1], J != nullptr);
    const auto& M = s1.R;
    const auto& N = s2.R;
    const double m1x = M[0][1] * a.f1 + M[0][0] * a.f3, m1y = M[1][1] * a.f1 + M[1][0] * a.f3;
    const double m2x = N[0][1] * b.f1 + N[0][0] * b.f3, m2y = N[1][1] * b.f1 + N[1][0] * b.f3;
    const double dy = (((((-(M[1][0] * a.f4)) - M[1][1] * a.f2) - s1.c[1]) + N[1][0] * b.f4) + N[1][1] * b.f2) + s2.c[1];
    const double dx = (((((-(M[0][0] * a.f4)) - M[0][1] * a.f2) - s1.c[0]) + N[0][0] * b.f4) + N[0][1] * b.f2) + s2.c[0];
    F[0] = (-m1x) * dy + m1y * dx;
    const double l1 = sqrt(m1x * m1x + m1y * m1y), l2 = sqrt(m2x * m2x + m2y * m2y);
    F[1] = (m1x * m2x + m1y * m2y) + l1 * l2;
    if (!J) return;
    const double g1x = M[0][0] * a.f9 - M[0][1] * a.f10, g1y = M[1][0] * a.f9 - M[1][1] * a.f10;
    const double h1x = M[0][1] * a.f12 - M[0][0] * a.f11, h1y = M[1][1] * a.f12 - M[1][0] * a.f11;
    const double g2x = N[0][0] * b.f9 - N[0][1] * b.f10, g2y = N[1][0] * b.f9 - N[1][1] * b.f10;
    const double h2x = N[0][1] * b.f12 - N[0][0] * b.f11, h2y = N[1][1] * b.f12 - N[1][0] * b.f11;
    J[0] = (((-m1x) * g1y + m1y * g1x) + (-h1x) * dy) + h1y * dx;
    J[1] = (-m1x) * (-g2y) + m1y * (-g2x);
    J[2] = (((m1x * h1x + m1y * h1y) * l2) / l1 + m2x * h1x) + m2y * h1y;
    J[3] = (((m2x * h2x + m2y * h2y) * l1) / l2 + m1x * h2x) + m1y * h2y;
}

__device__ inline void se2_residual(const Se2Pair& q, const double* u, double* F, double* J) {
    if (q.kind == 2) {
        se2_se2_residual(q.s1, q.s2, u, F, J);
        return;
    }
    if (!J) {
        se2_algebra<double>(q.kind, q.s1, q.s2, q.sp2, q.p2, q.sp, u, -1, F);
        return;
    }
    for (int k = 0; k < 2; ++k) {
        D1 f[2];
        se2_algebra<D1>(q.kind, q.s1, q.s2, q.sp2, q.p2, q.sp, u, k, f);
        for (int i = 0; i < 2; ++i) {
            J[2 * i + k] = f[i].d;
            if (k == 0) F[i] = f[i].v;
        }
    }
}

// newton_method (utils.hpp:81-118) on the 2 x 2 system: LU with partial pivoting (LAPACK gesv)
__device__ inline void lu_solve2(double* A, double* b) {
    constexpr int n = 2;
    int piv[n];
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(A[n * k + k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[n * i + k]) > best) {
                best = fabs(A[n * i + k]);
                p = i;
            }
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) {
                const double t = A[n * k + j];
                A[n * k + j] = A[n * p + j];
                A[n * p + j] = t;
            }
        if (A[n * k + k] != 0.0) {
            const double inv = 1.0 / A[n * k + k];
            for (int i = k + 1; i < n; ++i) A[n * i + k] *= inv;
        }
        for (int i = k + 1; i < n; ++i)
            for (int j = k + 1; j < n; ++j) A[n * i + j] -= A[n * i + k] * A[n * k + j];
    }
    for (int k = 0; k < n; ++k)
        if (piv[k] != k) {
            const double t = b[k];
            b[k] = b[piv[k]];
            b[piv[k]] = t;
        }
    for (int i = 1; i < n; ++i)
        for (int j = 0; j < i; ++j) b[i] -= A[n * i + j] * b[j];
    for (int i = n - 1; i >= 0; --i) {
        for (int j = i + 1; j < n; ++j) b[i] -= A[n * i + j] * b[j];
        b[i] /= A[n * i + i];
    }
}
__device__ inline double norm2v(const double* x) {
    double s = 0.0;
    for (int k = 0; k < 2; ++k) s += x[k] * x[k];
    return sqrt(s);
}
__device__ inline int newton2(const Se2Pair& q, double* u) {
    int cc = 0;
    while (cc < 200) {
        double F[2], J[4], d[2];
        se2_residual(q, u, F, J);
        for (int k = 0; k < 2; ++k) d[k] = -F[k];
        lu_solve2(J, d);
        if (norm2v(d) < 1.0e-7) return cc;
        double t = 1;
        const double ferr = norm2v(F);
        if (ferr < 1.0e-10) return cc;
        for (;;) {
            double ut[2], Ft[2];
            for (int k = 0; k < 2; ++k) ut[k] = u[k] + t * d[k];
            se2_residual(q, ut, Ft, nullptr);
            if (!(norm2v(Ft) > ferr && t > 0.01)) break;
            t -= 0.01;
        }
        for (int k = 0; k < 2; ++k) u[k] += t * d[k];
        cc += 1;
    }
    return -1;
}

__device__ inline double dist2d(const double* a, const double* b) {
    const double dx = a[0] - b[0], dy = a[1] - b[1];
    return sqrt(dx * dx + dy * dy);
}

// The closest points of one 2D pair with a superellipsoid, in the dispatcher's order (i, j);
// returns false for the shape pairs the reference has no form for.
__device__ inline bool se2_pair(const Particles& P, int64_t i, int64_t j, double* pi, double* pj, double* nij,
                                double& dij, double* u0, double* u, int& status) {
    const int si = P.shape[i], sj = P.shape[j];
    const bool se_i = si == SCOPI_SHAPE_SUPERELLIPSOID, se_j = sj == SCOPI_SHAPE_SUPERELLIPSOID;
    Se2Pair q{};
    u0[2] = u0[3] = u[2] = u[3] = 0.0;
    if (se_i && se_j) {
        q.kind = 2;
        q.s1 = load_se(P, i);
        q.s2 = load_se(P, j);
        double b1[16], b2[16];
        const int n1 = init_angles_n<4>(q.s1, true, b1), n2 = init_angles_n<4>(q.s2, true, b2);
        double dmin = INFINITY;
        int im = 0, jm = 0;
        for (int a = 0; a < n1; ++a) {
            double p[3], n[3];
            se2_point(q.s1, b1[a], p);
            se2_normal(q.s1, b1[a], n);
            for (int b = 0; b < n2; ++b) {
                double pq[3], m[3];
                se2_point(q.s2, b2[b], pq);
                se2_normal(q.s2, b2[b], m);
                const double dist = dist2d(p, pq) + 2 * (1 + (n[0] * m[0] + n[1] * m[1]));
                if (dist < dmin) {
                    dmin = dist;
                    im = a;
                    jm = b;
                }
            }
        }
        u0[0] = b1[im];
        u0[1] = b2[jm];
        typename Minpack<2>::State h;
        Minpack<2>::start(h, u0);
        while (!Minpack<2>::pass([&](const double* x, double* F, double* J) { se2_residual(q, x, F, J); }, h)) {
        }
        u[0] = h.x[0];
        u[1] = h.x[1];
        status = h.info;
        se2_point(q.s1, u[0], pi);
        se2_point(q.s2, u[1], pj);
        se2_normal(q.s2, u[1], nij);
        const double dx = pi[0] - pj[0], dy = pi[1] - pj[1];
        dij = sgn(dx * nij[0] + dy * nij[1]) * sqrt(dx * dx + dy * dy);
        return true;
    }
    const int64_t ks = se_i ? i : j, ko = se_i ? j : i;
    const int other = se_i ? sj : si;
    if (other != SCOPI_SHAPE_SPHERE && other != SCOPI_SHAPE_PLANE) return false;
    q.kind = (other == SCOPI_SHAPE_SPHERE) ? 0 : 1;
    const bool flip = (q.kind == 0) ? se_i : !se_i;
    q.s1 = load_se(P, ks);
    double a[3], b[3], n[3];
    if (q.kind == 0) {
        q.sp2 = load_sph(P, ko);
        const double bi[4] = {-kPi, -kPi / 2, 0.001, kPi / 2};
        double dmin = INFINITY;
        int im = 0, jm = 0;
        for (int x = 0; x < 4; ++x)
            for (int y = 0; y < 4; ++y) {
                se2_point(q.s1, bi[x], a);
                sph2_point(q.sp2, bi[y], b);
                const double d = dist2d(a, b);
                if (d < dmin) {
                    dmin = d;
                    im = x;
                    jm = y;
                }
            }
        u0[0] = bi[im];
        u0[1] = bi[jm];
    } else {
        q.p2 = load_pln(P, ko);
        q.sp = (q.s1.c[0] - q.p2.c[0]) * q.p2.R[0][0] + (q.s1.c[1] - q.p2.c[1]) * q.p2.R[1][0];
        constexpr int num = 10;
        double dmin = INFINITY;
        int im = 0, jm = 0;
        for (int x = 0; x < num; ++x) {
            se2_point(q.s1, lin(-kPi, kPi, num, x), a);
            for (int y = 0; y < num; ++y) {
                pln2_point(q.p2, lin(-num, num, num, y), b);
                const double d = dist2d(a, b);
                if (d < dmin) {
                    dmin = d;
                    im = x;
                    jm = y;
                }
            }
        }
        u0[0] = lin(-kPi, kPi, num, im);
        u0[1] = lin(-num, num, num, jm);
    }
    u[0] = u0[0];
    u[1] = u0[1];
    status = newton2(q, u);
    se2_point(q.s1, u[0], a);
    if (q.kind == 0) {
        sph2_point(q.sp2, u[1], b);
        sph2_normal(q.sp2, u[1], n);
    } else {
        pln2_point(q.p2, u[1], b);
        n[0] = q.p2.R[0][0];
        n[1] = q.p2.R[1][0];
        n[2] = 0.0;
    }
    dij = (a[0] - b[0]) * n[0] + (a[1] - b[1]) * n[1];
    for (int d = 0; d < 3; ++d) {
        pi[d] = flip ? b[d] : a[d];
        pj[d] = flip ? a[d] : b[d];
        nij[d] = flip ? n[d] * -1. : n[d];
    }
    return true;
}

__device__ inline bool is_se(const Particles& P, int64_t k) { return P.shape[k] == SCOPI_SHAPE_SUPERELLIPSOID; }

// initial angles: one lane per active particle
__global__ void __launch_bounds__(kBlock) k_se_angles(Particles P, SeGrids G) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t na = P.n - P.n_inactive;
    if (t >= na) return;
    const int64_t k = P.n_inactive + t;
    if (!is_se(P, k)) {
        G.count[2 * t] = G.count[2 * t + 1] = 0;
        return;
    }
    const SE s = load_se(P, k);
    double* ang = G.ang + 2 * kSeGridAngles * t;
    G.count[2 * t] = init_angles(s, false, ang);                 // a: ainit_yz
    G.count[2 * t + 1] = init_angles(s, true, ang + kSeGridAngles);  // b: binit_xy
}

// grid points and normals: one lane per (active particle, grid point), meshgrid(a, b)
// flattened row-major: point k = (a[k / nb], b[k % nb])
__global__ void __launch_bounds__(kBlock) k_se_points(Particles P, SeGrids G) {
    const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t na = P.n - P.n_inactive;
    const int64_t t = gid / kSeGridPoints;
    const int kp = (int)(gid - t * kSeGridPoints);
    if (t >= na) return;
    const int ca = G.count[2 * t], cb = G.count[2 * t + 1];
    if (kp >= ca * cb) return;
    const int64_t k = P.n_inactive + t;
    const SE s = load_se(P, k);
    const double* ang = G.ang + 2 * kSeGridAngles * t;
    const double a = ang[kp / cb], b = ang[kSeGridAngles + kp % cb];
    double p[3], nn[3];
    se_point(s, a, b, p);
    se_normal(s, a, b, nn);
    double* out = G.pts + 6 * kSeGridPoints * t;
    for (int d = 0; d < 3; ++d) {
        out[d * kSeGridPoints + kp] = p[d];
        out[(3 + d) * kSeGridPoints + kp] = nn[d];
    }
}

// Penalised grid minimum: one workgroup of kSeGridBlock lanes per candidate pair, the first
// minimum of fl(sqrt(d2)) + 2 (1 + n1.n2) over the row-major scan of (i, j) (closest_points.hpp:
// 604-690): the minimum value, and among equal values the smallest index i n2 + j.  Lane l owns grid
// points i = l, l + B, l + 2B of body 1 and scans body 2's grid (staged in LDS) in order, keeping its
// own (value, index) minimum; the workgroup's minimum is then the (value, index) minimum over lanes.
// Only that minimum matters, so a term may be skipped when it is provably above an upper bound g of
// it: g is the lowest term value any lane has found, shared through LDS every 32 points of body 2
// (a value read there is some lane's found value, old or new, so any read is a valid bound).  A term
// whose penalty exceeds g is above g (the distance term is >= 0), and so is one whose squared
// distance exceeds (g - penalty + 1e-15 g)^2 (1 + 1e-14), the margins covering the roundings of the
// term; their square roots are skipped.  Terms equal to g are never skipped, so the minimum and its
// index are those of the full scan, bit for bit.  The penalty test itself is screened first with
// the dot product n1.n2 in three FMAs against 0.5 g - 1 + 1e-13: it differs from the term's dot
// product (three products and two sums, rounded each) by less than 1e-15 for unit normals, so a term
// the screen rejects has 2 (1 + n1.n2) > g as computed; a second screen does the same for the
// distance test, and the terms that pass both take the exact test.
constexpr int kSeGridBlock = 192;  // 3 x 192 = 576
__global__ void __launch_bounds__(kSeGridBlock) k_se_grid(Particles P, SeGrids G, SeCandidates S, int64_t n) {
    __shared__ double g2[6][kSeGridPoints];
    __shared__ double red_v[kSeGridBlock];
    __shared__ int red_i[kSeGridBlock];
    __shared__ double wbest[kSeGridBlock / 64];  // read while other waves write: relaxed atomics (LDS reads, writes)
    const int64_t c = blockIdx.x;
    if (c >= n) return;
    const int64_t i = S.i[c], j = S.j[c];
    if (!is_se(P, i) || !is_se(P, j)) return;
    const int64_t t1 = i - P.n_inactive, t2 = j - P.n_inactive;
    const int na1 = G.count[2 * t1], nb1 = G.count[2 * t1 + 1];
    const int na2 = G.count[2 * t2], nb2 = G.count[2 * t2 + 1];
    const int n1 = na1 * nb1, n2 = na2 * nb2;
    const double* src2 = G.pts + 6 * kSeGridPoints * t2;
    for (int k = threadIdx.x; k < n2; k += kSeGridBlock)
        for (int d = 0; d < 6; ++d) g2[d][k] = src2[d * kSeGridPoints + k];
    if (threadIdx.x < kSeGridBlock / 64) wbest[threadIdx.x] = INFINITY;
    const double* src1 = G.pts + 6 * kSeGridPoints * t1;
    double p1[3][3], m1[3][3];
    bool ok[3];
    for (int r = 0; r < 3; ++r) {
        const int ii = threadIdx.x + r * kSeGridBlock;
        ok[r] = ii < n1;
        for (int d = 0; d < 3; ++d) {
            p1[r][d] = ok[r] ? src1[d * kSeGridPoints + ii] : 0.0;
            m1[r][d] = ok[r] ? src1[(3 + d) * kSeGridPoints + ii] : 0.0;
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    double bv = INFINITY;   // the lane's minimum ...
    int bidx = 0x7fffffff;  // ... and its row-major index
    double g = INFINITY;    // an upper bound of the workgroup's minimum
    for (int jj = 0; jj < n2; ++jj) {
        if ((jj & 31) == 0 && jj > 0) {
            double w = bv;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) w = fmin(w, __shfl_xor(w, o, 64));
            if ((threadIdx.x & 63) == 0 &&
                w < __hip_atomic_load(&wbest[wave], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
                __hip_atomic_store(&wbest[wave], w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
            for (int k = 0; k < kSeGridBlock / 64; ++k)
                g = fmin(g, __hip_atomic_load(&wbest[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
        }
        const double nx = g2[3][jj], ny = g2[4][jj], nz = g2[5][jj];
        const double screen = 0.5 * g - 1.0 + 1e-13;  // +inf while g is
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const double dotf = fma(m1[r][2], nz, fma(m1[r][1], ny, m1[r][0] * nx));
            if (dotf > screen) continue;
            const double dx = p1[r][0] - g2[0][jj], dy = p1[r][1] - g2[1][jj], dz = p1[r][2] - g2[2][jj];
            {
                // the distance test in FMAs: penf is within 2e-15 of the exact penalty and d2f within
                // 5e-16 d2 of the exact squared distance, so a term rejected here with these margins
                // fails the exact test below too
                const double penf = fma(2.0, dotf, 2.0);
                const double uf = (g - penf) + 1e-14 * (g + 1.0);
                const double d2f = fma(dz, dz, fma(dy, dy, dx * dx));
                if (uf > 0 && d2f > uf * uf * (1.0 + 1e-13)) continue;
            }
            const double pen = 2 * (1 + (m1[r][0] * nx + m1[r][1] * ny + m1[r][2] * nz));
            if (ok[r] && pen <= g) {
                const double d2 = dx * dx + dy * dy + dz * dz;
                const double u = (g - pen) + 1e-15 * g;
                if (!(u > 0 && d2 > u * u * (1.0 + 1e-14))) {
                    const double dist = sqrt(d2) + pen;
                    const int idx = (threadIdx.x + r * kSeGridBlock) * n2 + jj;
                    if (dist < bv || (dist == bv && idx < bidx)) {
                        bv = dist;
                        bidx = idx;
                        if (dist < g) g = dist;
                    }
                }
            }
        }
    }
    red_v[threadIdx.x] = bv;
    red_i[threadIdx.x] = bidx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = red_v[0];
        int idx = red_i[0];
        for (int l = 1; l < kSeGridBlock; ++l)
            if (red_v[l] < v || (red_v[l] == v && red_i[l] < idx)) {
                v = red_v[l];
                idx = red_i[l];
            }
        int im = 0, jm = 0;
        if (idx != 0x7fffffff) {
            im = idx / n2;
            jm = idx - im * n2;
        }
        const double* ang1 = G.ang + 2 * kSeGridAngles * t1;
        const double* ang2 = G.ang + 2 * kSeGridAngles * t2;
        S.u0[4 * c + 0] = ang1[im / nb1];
        S.u0[4 * c + 1] = ang1[kSeGridAngles + im % nb1];
        S.u0[4 * c + 2] = ang2[jm / nb2];
        S.u0[4 * c + 3] = ang2[kSeGridAngles + jm % nb2];
    }
}

// Closest points of every candidate: a work queue.  Each lane takes the next pair index from a
// device counter; superellipsoid pairs run hybrd1 pass by pass, and a lane whose pair ends takes
// another at once, so lanes of a wavefront are not held by the pair with the most evaluations.
// Sphere / plane pairs with a superellipsoid run their Newton solve, other pairs their closed form.
__device__ inline void write_record(SeCandidates& S, int64_t c, double dij, const double* nij, const double* pi,
                                    const double* pj) {
    S.dij[c] = dij;
    for (int d = 0; d < 3; ++d) {
        S.nij[d * S.cap + c] = nij[d];
        S.pi[d * S.cap + c] = pi[d];
        S.pj[d * S.cap + c] = pj[d];
    }
}

constexpr int kSolveBlock = 64;
__global__ void __launch_bounds__(kSolveBlock) k_se_solve(Particles P, SeCandidates S, int64_t n) {
    SE s1, s2;
    Minpack<4>::State h;
    int64_t c = atomicAdd(S.counter, 1);
    bool busy = false;
    for (;;) {
        if (!busy) {
            if (c >= n) break;
            const int64_t i = S.i[c], j = S.j[c];
            const bool se_i = is_se(P, i), se_j = is_se(P, j);
            if (se_i && se_j) {
                s1 = load_se(P, i);
                s2 = load_se(P, j);
                Minpack<4>::start(h, S.u0 + 4 * c);
                busy = true;
            } else {
                if (se_i != se_j) {  // sphere / plane - superellipsoid: k_newton_pairs
                    c = atomicAdd(S.counter, 1);
                    continue;
                }
                {
                    CP cp;
                    const bool ok = closest_points<3>(P, i, j, cp);
                    const d3 z = mk(0.0, 0.0, 0.0);
                    const d3 a = ok ? cp.nij : z, b = ok ? cp.pi : z, e = ok ? cp.pj : z;
                    const double na[3] = {a.x, a.y, a.z}, pa[3] = {b.x, b.y, b.z}, pb[3] = {e.x, e.y, e.z};
                    write_record(S, c, ok ? cp.dij : INFINITY, na, pa, pb);
                    S.info[c] = -1;
                }
                S.nfev[c] = 0;
                c = atomicAdd(S.counter, 1);
                continue;
            }
        }
        if (!Minpack<4>::pass([&](const double* x, double* F, double* J) { se_residual(s1, s2, x, F, J); }, h))
            continue;
        // hybrd1 ended: the contact record of closest_points.hpp:724-728
        double pi[3], pj[3], nij[3];
        se_point(s1, h.x[0], h.x[1], pi);
        se_point(s2, h.x[2], h.x[3], pj);
        se_normal(s2, h.x[2], h.x[3], nij);
        const double dx = pi[0] - pj[0], dy = pi[1] - pj[1], dz = pi[2] - pj[2];
        const double sd = sgn(dx * nij[0] + dy * nij[1] + dz * nij[2]);
        write_record(S, c, sd * sqrt(dx * dx + dy * dy + dz * dz), nij, pi, pj);
        for (int k = 0; k < 4; ++k) S.u[4 * c + k] = h.x[k];
        S.info[c] = h.info;
        S.nfev[c] = h.nfev;
        busy = false;
        c = atomicAdd(S.counter, 1);
    }
}

__global__ void k_zero_counter(int* counter) { *counter = 0; }

// sphere / plane - superellipsoid candidates: the reference's damped Newton, one lane per pair
__global__ void __launch_bounds__(kSolveBlock) k_newton_pairs(Particles P, SeCandidates S, int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kSolveBlock + threadIdx.x;
    if (c >= n) return;
    const int64_t i = S.i[c], j = S.j[c];
    if (is_se(P, i) == is_se(P, j)) return;
    double pi[3], pj[3], nij[3], u0[4], u[4], dij;
    int its;
    newton_pair(P, i, j, pi, pj, nij, dij, u0, u, its);
    write_record(S, c, dij, nij, pi, pj);
    for (int k = 0; k < 4; ++k) {
        S.u0[4 * c + k] = u0[k];
        S.u[4 * c + k] = u[k];
    }
    S.info[c] = its;
    S.nfev[c] = 0;
}

// the portable elementary functions, elementwise (tests)
// 2D scenes with superellipsoids: one lane per candidate pair, every shape pair
__global__ void __launch_bounds__(kSolveBlock) k_se2_pairs(Particles P, SeCandidates S, int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * kSolveBlock + threadIdx.x;
    if (c >= n) return;
    const int64_t i = S.i[c], j = S.j[c];
    double pi[3] = {0, 0, 0}, pj[3] = {0, 0, 0}, nij[3] = {0, 0, 0}, u0[4] = {0, 0, 0, 0}, u[4] = {0, 0, 0, 0};
    double dij = INFINITY;
    int status = -1;
    if (is_se(P, i) || is_se(P, j)) {
        if (!se2_pair(P, i, j, pi, pj, nij, dij, u0, u, status)) dij = INFINITY;
    } else {
        CP cp;
        if (closest_points<2>(P, i, j, cp)) {
            dij = cp.dij;
            nij[0] = cp.nij.x, nij[1] = cp.nij.y, nij[2] = cp.nij.z;
            pi[0] = cp.pi.x, pi[1] = cp.pi.y, pi[2] = cp.pi.z;
            pj[0] = cp.pj.x, pj[1] = cp.pj.y, pj[2] = cp.pj.z;
        }
    }
    write_record(S, c, dij, nij, pi, pj);
    for (int k = 0; k < 4; ++k) {
        S.u0[4 * c + k] = u0[k];
        S.u[4 * c + k] = u[k];
    }
    S.info[c] = status;
    S.nfev[c] = 0;
}

__global__ void __launch_bounds__(kBlock) k_portable_math(int fn, int64_t n, const double* __restrict__ x,
                                                         const double* __restrict__ y, double* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const double a = x[k];
    double r;
    switch (fn) {
        case 0: r = pm_cos(a); break;
        case 1: r = pm_sin(a); break;
        case 2: r = pm_asin(a); break;
        case 3: r = pm_log(a); break;
        case 4: r = pm_exp(a); break;
        default: r = pm_pow(a, y[k]); break;
    }
    out[k] = r;
}

// residual / Jacobian of a Newton pair and the 4x4 solve, one thread (tests): kind 0 sphere, 1 plane;
// in: SE (c, q, r, sq: 12 doubles), other (c, q, r: 8 doubles), u (4), A (16), b (4);
// out: F (4), J (16), x (4)
__global__ void k_newton_probe(int kind, const double* __restrict__ in, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    SE s1;
    for (int d = 0; d < 3; ++d) {
        s1.c[d] = in[d];
        s1.r[d] = in[7 + d];
    }
    s1.e = in[10];
    s1.n = in[11];
    const m3 R1 = rotation3(in[3], in[4], in[5], in[6]);
    const m3 R2 = rotation3(in[15], in[16], in[17], in[18]);
    Sph s2{};
    Pln p2{};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            s1.R[a][b] = R1.a[a][b];
            s2.R[a][b] = p2.R[a][b] = R2.a[a][b];
        }
    for (int d = 0; d < 3; ++d) s2.c[d] = p2.c[d] = in[12 + d];
    s2.r = in[19];
    double sp = 0.0;
    if (kind == 1)
        sp = (s1.c[0] - p2.c[0]) * p2.R[0][0] + (s1.c[1] - p2.c[1]) * p2.R[1][0] + (s1.c[2] - p2.c[2]) * p2.R[2][0];
    nt_residual(kind, s1, s2, p2, sp, in + 20, out, out + 4);
    double A[16], b[4];
    for (int k = 0; k < 16; ++k) A[k] = in[24 + k];
    for (int k = 0; k < 4; ++k) b[k] = in[40 + k];
    lu_solve4(A, b);
    for (int k = 0; k < 4; ++k) out[20 + k] = b[k];
}

}  // namespace

void newton_probe(int kind, const double* in, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_newton_probe, dim3(1), dim3(64), 0, s, kind, in, out);
    SCOPI_HIP_CHECK(hipGetLastError());
}

void portable_math(int fn, int64_t n, const double* x, const double* y, double* out, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_portable_math, dim3(grid_for(n)), dim3(kBlock), 0, s, fn, n, x, y, out);
    SCOPI_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------------------------ host
void se_grids_free(SeGrids& G) {
    for (void* p : {(void*)G.ang, (void*)G.count, (void*)G.pts})
        if (p) (void)hipFree(p);
    G = SeGrids{};
}

void se_grids_reserve(SeGrids& G, int64_t n) {
    if (n <= G.cap) return;
    se_grids_free(G);
    const int64_t cap = std::max<int64_t>(n, 64);
    G.cap = cap;
    SCOPI_HIP_CHECK(hipMalloc(&G.ang, 2 * kSeGridAngles * cap * sizeof(double)));
    SCOPI_HIP_CHECK(hipMalloc(&G.count, 2 * cap * sizeof(int32_t)));
    SCOPI_HIP_CHECK(hipMalloc(&G.pts, (size_t)6 * kSeGridPoints * cap * sizeof(double)));
}

void se_candidates_free(SeCandidates& S) {
    for (void* p : {(void*)S.i, (void*)S.j, (void*)S.u0, (void*)S.u, (void*)S.info, (void*)S.nfev, (void*)S.dij,
                    (void*)S.nij, (void*)S.pi, (void*)S.pj, (void*)S.counter})
        if (p) (void)hipFree(p);
    S = SeCandidates{};
}

void se_candidates_reserve(SeCandidates& S, int64_t n) {
    if (n <= S.cap) return;
    se_candidates_free(S);
    const int64_t cap = std::max<int64_t>(n + n / 4, 256);
    S.cap = cap;
    SCOPI_HIP_CHECK(hipMalloc(&S.i, cap * sizeof(int32_t)));
    SCOPI_HIP_CHECK(hipMalloc(&S.j, cap * sizeof(int32_t)));
    SCOPI_HIP_CHECK(hipMalloc(&S.u0, 4 * cap * sizeof(double)));
    SCOPI_HIP_CHECK(hipMalloc(&S.u, 4 * cap * sizeof(double)));
    SCOPI_HIP_CHECK(hipMalloc(&S.info, cap * sizeof(int32_t)));
    SCOPI_HIP_CHECK(hipMalloc(&S.nfev, cap * sizeof(int32_t)));
    SCOPI_HIP_CHECK(hipMalloc(&S.dij, cap * sizeof(double)));
    SCOPI_HIP_CHECK(hipMalloc(&S.nij, 3 * cap * sizeof(double)));
    SCOPI_HIP_CHECK(hipMalloc(&S.pi, 3 * cap * sizeof(double)));
    SCOPI_HIP_CHECK(hipMalloc(&S.pj, 3 * cap * sizeof(double)));
    SCOPI_HIP_CHECK(hipMalloc(&S.counter, sizeof(int)));
}

void se_closest_points(const Particles& P, SeGrids& G, SeCandidates& S, int64_t n, hipStream_t s) {
    const int64_t na = P.n - P.n_inactive;
    if (n <= 0 || na <= 0) return;
    if (P.dim == 2) {
        hipLaunchKernelGGL(k_se2_pairs, dim3((unsigned)((n + kSolveBlock - 1) / kSolveBlock)), dim3(kSolveBlock), 0, s,
                           P, S, n);
        SCOPI_HIP_CHECK(hipGetLastError());
        return;
    }
    se_grids_reserve(G, na);
    hipLaunchKernelGGL(k_se_angles, dim3(grid_for(na)), dim3(kBlock), 0, s, P, G);
    hipLaunchKernelGGL(k_se_points, dim3(grid_for(na * kSeGridPoints)), dim3(kBlock), 0, s, P, G);
    hipLaunchKernelGGL(k_se_grid, dim3((unsigned)n), dim3(kSeGridBlock), 0, s, P, G, S, n);
    hipLaunchKernelGGL(k_newton_pairs, dim3((unsigned)((n + kSolveBlock - 1) / kSolveBlock)), dim3(kSolveBlock), 0, s,
                       P, S, n);
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, s, S.counter);
    // enough lanes to fill the device at the solve kernel's occupancy; the queue balances them
    const int64_t lanes = std::min<int64_t>(n, 2048 * kSolveBlock);
    hipLaunchKernelGGL(k_se_solve, dim3((unsigned)((lanes + kSolveBlock - 1) / kSolveBlock)), dim3(kSolveBlock), 0, s,
                       P, S, n);
    SCOPI_HIP_CHECK(hipGetLastError());
}

}  // namespace scopi_hip
