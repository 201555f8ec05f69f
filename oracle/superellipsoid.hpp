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
//   Powell hybrid   MINPACK hybrd / enorm / qrfac / qform / dogleg / r1updt / r1mpyq,
//                   restated from the published algorithm (More, Garbow, Hillstrom,
//                   ANL-80-74, 1980).  The Jacobian is evaluated exactly by forward-mode
//                   automatic differentiation of F (the reference hand-codes the same
//                   derivative, closest_points.hpp:385-603).
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

// Residual of closest_points.hpp:284-382 with its Jacobian.
struct SEPair {
    SE s1, s2;
};

template <class M>
inline void se_se_residual(const SEPair& P, const double* u, double* F, double* J /* 4x4 row-major or null */) {
    // Values follow the reference's expression order (left-to-right products and
    // sums, closest_points.hpp:323-381); derivatives ride along by forward-mode AD.
    D4 a1 = var(u[0], 0), b1 = var(u[1], 1), a2 = var(u[2], 2), b2 = var(u[3], 3);
    const SE &s1 = P.s1, &s2 = P.s2;
    D4 ca1 = dcos<M>(a1), sa1 = dsin<M>(a1), cb1 = dcos<M>(b1), sb1 = dsin<M>(b1);
    D4 ca2 = dcos<M>(a2), sa2 = dsin<M>(a2), cb2 = dcos<M>(b2), sb2 = dsin<M>(b2);
    D4 ca1n = spow<M>(ca1, s1.n), ca1n2 = spow<M>(ca1, 2 - s1.n), sa1n = spow<M>(sa1, s1.n), sa1n2 = spow<M>(sa1, 2 - s1.n);
    D4 cb1e = spow<M>(cb1, s1.e), cb1e2 = spow<M>(cb1, 2 - s1.e), sb1e = spow<M>(sb1, s1.e), sb1e2 = spow<M>(sb1, 2 - s1.e);
    D4 ca2n = spow<M>(ca2, s2.n), ca2n2 = spow<M>(ca2, 2 - s2.n), sa2n = spow<M>(sa2, s2.n), sa2n2 = spow<M>(sa2, 2 - s2.n);
    D4 cb2e = spow<M>(cb2, s2.e), cb2e2 = spow<M>(cb2, 2 - s2.e), sb2e = spow<M>(sb2, s2.e), sb2e2 = spow<M>(sb2, 2 - s2.e);
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

// ------------------------------------------------------------------ MINPACK hybrd (n = 4 in 3D, 2 in 2D)

inline double enorm(int n, const double* x) {
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

// the routines below for one system size N
template <int N>
struct Minpack {
static constexpr int LR = N * (N + 1) / 2;

// a is column-major n x n (a[i + n*j])
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

// r: upper triangle packed by rows (r[l] with row-major order); 1-based index arithmetic kept.
static inline void dogleg(const double* r, const double* diag, const double* qtb, double delta, double* x, double* wa1,
                   double* wa2) {
    const double epsmch = 2.220446049250313e-16;
    int jj = (N * (N + 1)) / 2 + 1;
    for (int k = 1; k <= N; ++k) {
        const int j = N - k + 1;
        const int jp1 = j + 1;
        jj -= k;
        int l = jj + 1;
        double sum = 0.0;
        for (int i = jp1; i <= N; ++i) {
            sum += r[l - 1] * x[i - 1];
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
    if (qnorm <= delta) return;
    int l = 1;
    for (int j = 1; j <= N; ++j) {
        const double temp = qtb[j - 1];
        for (int i = j; i <= N; ++i) {
            wa1[i - 1] += r[l - 1] * temp;
            ++l;
        }
        wa1[j - 1] /= diag[j - 1];
    }
    const double gnorm = enorm(N, wa1);
    double sgnorm = 0.0;
    double alpha = delta / qnorm;
    if (gnorm != 0.0) {
        for (int j = 0; j < N; ++j) wa1[j] = (wa1[j] / gnorm) / diag[j];
        l = 1;
        for (int j = 1; j <= N; ++j) {
            double sum = 0.0;
            for (int i = j; i <= N; ++i) {
                sum += r[l - 1] * wa1[i - 1];
                ++l;
            }
            wa2[j - 1] = sum;
        }
        const double temp = enorm(N, wa2);
        sgnorm = (gnorm / temp) / temp;
        alpha = 0.0;
        if (sgnorm < delta) {
            const double bnorm = enorm(N, qtb);
            double t = (bnorm / gnorm) * (bnorm / qnorm) * (sgnorm / delta);
            t = t - (delta / qnorm) * (sgnorm / delta) * (sgnorm / delta) +
                std::sqrt((t - (delta / qnorm)) * (t - (delta / qnorm)) +
                          (1.0 - (delta / qnorm) * (delta / qnorm)) * (1.0 - (sgnorm / delta) * (sgnorm / delta)));
            alpha = ((delta / qnorm) * (1.0 - (sgnorm / delta) * (sgnorm / delta))) / t;
        }
    }
    const double temp = (1.0 - alpha) * std::min(sgnorm, delta);
    for (int j = 0; j < N; ++j) x[j] = temp * wa1[j] + alpha * x[j];
}

// Rank-1 update of the packed triangular s (m = n = N); 1-based index arithmetic kept.
static inline void r1updt(double* s, const double* u, double* v, double* w, bool& sing) {
    const double giant = 1.79769313486231570e308;
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
                if (std::abs(cs) * giant > 1.0) tau = 1.0 / cs;
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
    for (int i = 1; i <= m; ++i) w[i - 1] += v[n - 1] * u[i - 1];
    sing = false;
    for (int j = 1; j <= n - 1; ++j) {
        if (w[j - 1] != 0.0) {
            double cs, sn, tau;
            if (std::abs(s[jj - 1]) < std::abs(w[j - 1])) {
                const double cotan = s[jj - 1] / w[j - 1];
                sn = 0.5 / std::sqrt(0.25 + 0.25 * cotan * cotan);
                cs = sn * cotan;
                tau = 1.0;
                if (std::abs(cs) * giant > 1.0) tau = 1.0 / cs;
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

// a is column-major m x n with leading dimension lda
static inline void r1mpyq(int m, int n, double* a, int lda, const double* v, const double* w) {
    for (int nmj = 1; nmj <= n - 1; ++nmj) {
        const int j = n - nmj;
        double cs, sn;
        if (std::abs(v[j - 1]) > 1.0) {
            cs = 1.0 / v[j - 1];
            sn = std::sqrt(1.0 - cs * cs);
        } else {
            sn = v[j - 1];
            cs = std::sqrt(1.0 - sn * sn);
        }
        for (int i = 1; i <= m; ++i) {
            const double temp = cs * a[(i - 1) + lda * (j - 1)] - sn * a[(i - 1) + lda * (n - 1)];
            a[(i - 1) + lda * (n - 1)] = sn * a[(i - 1) + lda * (j - 1)] + cs * a[(i - 1) + lda * (n - 1)];
            a[(i - 1) + lda * (j - 1)] = temp;
        }
    }
    for (int j = 1; j <= n - 1; ++j) {
        double cs, sn;
        if (std::abs(w[j - 1]) > 1.0) {
            cs = 1.0 / w[j - 1];
            sn = std::sqrt(1.0 - cs * cs);
        } else {
            sn = w[j - 1];
            cs = std::sqrt(1.0 - sn * sn);
        }
        for (int i = 1; i <= m; ++i) {
            const double temp = cs * a[(i - 1) + lda * (j - 1)] + sn * a[(i - 1) + lda * (n - 1)];
            a[(i - 1) + lda * (n - 1)] = -sn * a[(i - 1) + lda * (j - 1)] + cs * a[(i - 1) + lda * (n - 1)];
            a[(i - 1) + lda * (j - 1)] = temp;
        }
    }
}

// hybrd1 (minpack.hpp:851-923) with the analytic Jacobian: returns info (1: converged).
template <class Fn>
static inline int hybrd1(Fn fcn, double* x, int* nfev_out = nullptr, int* njev_out = nullptr) {
    const double xtol = 1.0e-10, factor = 100.0, epsmch = 2.220446049250313e-16;
    const int maxfev = 200 * (N + 1);
    const double p1 = 0.1, p5 = 0.5, p001 = 0.001, p0001 = 0.0001;
    double diag[N];  // mode 2: wa[0..n) = 1 (minpack.hpp:881-884)
    for (int k = 0; k < N; ++k) diag[k] = 1.0;
    double fvec[N], fjac[N * N], r[LR], qtf[N], wa1[N], wa2[N], wa3[N], wa4[N];
    int info = 0, nfev = 0, njev = 0;
    fcn(x, fvec, nullptr);
    nfev = 1;
    double fnorm = enorm(N, fvec);
    const int msum = N;
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
            delta = factor * xnorm;
            if (delta == 0.0) delta = factor;
        }
        for (int i = 0; i < N; ++i) qtf[i] = fvec[i];
        for (int j = 0; j < N; ++j) {
            if (fjac[j + N * j] != 0.0) {
                double sum = 0.0;
                for (int i = j; i < N; ++i) sum += fjac[i + N * j] * qtf[i];
                const double temp = -sum / fjac[j + N * j];
                for (int i = j; i < N; ++i) qtf[i] += fjac[i + N * j] * temp;
            }
        }
        bool sing = false;
        for (int j = 0; j < N; ++j) {
            int l = j;
            for (int i = 0; i < j; ++i) {
                r[l] = fjac[i + N * j];
                l += N - i - 1;
            }
            r[l] = wa1[j];
            if (wa1[j] == 0.0) sing = true;
        }
        (void)sing;
        qform(fjac, wa1);
        for (;;) {
            dogleg(r, diag, qtf, delta, wa1, wa2, wa3);
            for (int j = 0; j < N; ++j) {
                wa1[j] = -wa1[j];
                wa2[j] = x[j] + wa1[j];
                wa3[j] = diag[j] * wa1[j];
            }
            const double pnorm = enorm(N, wa3);
            if (iter == 1) delta = std::min(delta, pnorm);
            fcn(wa2, wa4, nullptr);
            ++nfev;
            const double fnorm1 = enorm(N, wa4);
            double actred = -1.0;
            if (fnorm1 < fnorm) actred = 1.0 - (fnorm1 / fnorm) * (fnorm1 / fnorm);
            int l = 0;
            for (int i = 0; i < N; ++i) {
                double sum = 0.0;
                for (int j = i; j < N; ++j) {
                    sum += r[l] * wa1[j];
                    ++l;
                }
                wa3[i] = qtf[i] + sum;
            }
            const double temp = enorm(N, wa3);
            double prered = 0.0;
            if (temp < fnorm) prered = 1.0 - (temp / fnorm) * (temp / fnorm);
            double ratio = 0.0;
            if (prered > 0.0) ratio = actred / prered;
            if (ratio < p1) {
                ncsuc = 0;
                ++ncfail;
                delta = p5 * delta;
            } else {
                ncfail = 0;
                ++ncsuc;
                if (ratio >= p5 || ncsuc > 1) delta = std::max(delta, pnorm / p5);
                if (std::abs(ratio - 1.0) <= p1) delta = pnorm / p5;
            }
            if (ratio >= p0001) {
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
            if (actred >= p001) nslow1 = 0;
            if (jeval) ++nslow2;
            if (actred >= p1) nslow2 = 0;
            if (delta <= xtol * xnorm || fnorm == 0.0) info = 1;
            if (info != 0) break;
            if (maxfev <= nfev) info = 2;
            if (p1 * std::max(p1 * delta, pnorm) <= epsmch * xnorm) info = 3;
            if (nslow2 == 5) info = 4;
            if (nslow1 == 10) info = 5;
            if (info != 0) break;
            if (ncfail == 2) break;
            for (int j = 0; j < N; ++j) {
                double sum = 0.0;
                for (int i = 0; i < N; ++i) sum += fjac[i + N * j] * wa4[i];
                wa2[j] = (sum - wa3[j]) / pnorm;
                wa1[j] = diag[j] * ((diag[j] * wa1[j]) / pnorm);
                if (ratio >= p0001) qtf[j] = sum;
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
inline D4 upow(const D4& a, double p, bool exact) {
    const double ax = std::abs(a.v);
    D4 r{M::pow(ax, p), {}};
    const double dv = exact ? (p * sgn(a.v)) * M::pow(ax, p - 1) : p * M::pow(ax, p - 1);
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
    const D4 ca1n = spow<M>(ca1, s1.n), ca1n2 = upow<M>(ca1, 2 - s1.n, exact);
    const D4 sa1n = spow<M>(sa1, s1.n), sa1n2 = spow<M>(sa1, 2 - s1.n);
    const D4 cb1e = spow<M>(cb1, s1.e), cb1e2 = spow<M>(cb1, 2 - s1.e);
    const D4 sb1e = spow<M>(sb1, s1.e), sb1e2 = spow<M>(sb1, 2 - s1.e);
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
    nx = s.r[1] * spow<M>(cb, 2 - s.e);
    ny = s.r[0] * spow<M>(sb, 2 - s.e);
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

// closest_points(superellipsoid<2>, superellipsoid<2>) residual (closest_points.hpp:79-125): the
// normal n1 of s1 at b1 is collinear with p2 - p1 (cross product 0) and anti-parallel to n2 at b2
template <class M>
inline void se2_se2_residual(const SE& s1, const SE& s2, const double* u, double* F, double* J) {
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
