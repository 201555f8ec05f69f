// se_math.hpp -- portable double-precision cos, sin, asin, log, exp and pow for the
// superellipsoid restatement.
//
// TEST INFRASTRUCTURE ONLY (included by superellipsoid.hpp).
//
// The superellipsoid closest-point solve is chaotic at the rounding level: a 1-ulp
// change of one input moves the hybrd1 end point by up to ~1e-6 in the distance
// (hybrd1 mostly stops on "no progress", info 4, so where it stops depends on
// every rounding).  The device kernel evaluates these functions with the same
// operations in the same order (IEEE add/sub/mul/div/sqrt, no contraction), so
// with this math selected the CPU and GPU results agree bit for bit; with the C
// library's functions selected the oracle is the reference's arithmetic.
//
// Methods (textbook, written from the mathematics):
//   cos / sin  Cody-Waite reduction by pi/2 split into 33 + 33 + 53 bits, Taylor
//              polynomials to degree 17 / 18 on |r| <= pi/4 (truncation < 2^-60).
//   asin       Maclaurin series of asin on [0, 1/2] (30 terms), and
//              asin x = pi/2 - 2 asin(sqrt((1 - x) / 2)) above 1/2.
//   log        x = 2^k m, m in [sqrt(2)/2, sqrt(2)), log m = 2 atanh((m-1)/(m+1)) by its
//              series to s^23, k ln2 with ln2 split in 32 + 53 bits.
//   exp        y = k ln2 + r, |r| <= ln2/2, Taylor to degree 14, scaled by 2^k through
//              the exponent bits.
//   pow(x, p)  x >= 0 only (the callers take |.|): exp(p log x), with pow(0, p) = 0, 1,
//              +inf for p > 0, = 0, < 0 as in C.
// Accuracy against the C library is checked in tests/test_oracle_se.py.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace se_math {

inline uint64_t bits_of(double x) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return u;
}
inline double from_bits(uint64_t u) {
    double x;
    std::memcpy(&x, &u, 8);
    return x;
}

constexpr double kPio2_1 = 0x1.921fb54400000p+0;   // pi/2, first 33 bits
constexpr double kPio2_2 = 0x1.0b4611a600000p-34;  // next 33 bits
constexpr double kPio2_3 = 0x1.3198a2e037073p-69;  // next 53 bits
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
constexpr double kPio2 = 0x1.921fb54442d18p+0;
constexpr double kLn2Hi = 0x1.62e42ff000000p-1;
constexpr double kLn2Lo = -0x1.718432a1b0e26p-35;
constexpr double kInvLn2 = 0x1.71547652b82fep+0;
constexpr double kSqrt2 = 0x1.6a09e667f3bcdp+0;

// 1/n! for n = 0..18
constexpr double kInvFact[19] = {1.0,
                                 1.0,
                                 0.5,
                                 1.0 / 6.0,
                                 1.0 / 24.0,
                                 1.0 / 120.0,
                                 1.0 / 720.0,
                                 1.0 / 5040.0,
                                 1.0 / 40320.0,
                                 1.0 / 362880.0,
                                 1.0 / 3628800.0,
                                 1.0 / 39916800.0,
                                 1.0 / 479001600.0,
                                 1.0 / 6227020800.0,
                                 1.0 / 87178291200.0,
                                 1.0 / 1307674368000.0,
                                 1.0 / 20922789888000.0,
                                 1.0 / 355687428096000.0,
                                 1.0 / 6402373705728000.0};

// sin and cos of |r| <= pi/4 + tiny
inline double sin_poly(double r) {
    const double s = r * r;
    double p = kInvFact[17];
    p = p * s - kInvFact[15];
    p = p * s + kInvFact[13];
    p = p * s - kInvFact[11];
    p = p * s + kInvFact[9];
    p = p * s - kInvFact[7];
    p = p * s + kInvFact[5];
    p = p * s - kInvFact[3];
    return r + (r * s) * p;
}
inline double cos_poly(double r) {
    const double s = r * r;
    double p = kInvFact[18];
    p = p * s - kInvFact[16];
    p = p * s + kInvFact[14];
    p = p * s - kInvFact[12];
    p = p * s + kInvFact[10];
    p = p * s - kInvFact[8];
    p = p * s + kInvFact[6];
    p = p * s - kInvFact[4];
    p = p * s + kInvFact[2];
    return 1.0 - s * p;
}

// x = k pi/2 + r; returns r and the quadrant k mod 4
inline double reduce_pio2(double x, int& quadrant) {
    const double k = std::rint(x * kTwoOverPi);
    const double r = ((x - k * kPio2_1) - k * kPio2_2) - k * kPio2_3;
    const int64_t ki = (int64_t)k;
    quadrant = (int)(ki & 3);
    return r;
}

inline double cos(double x) {
    if (!(std::fabs(x) < 1e15)) return x - x;  // inf / nan -> nan; huge arguments are not used
    int q;
    const double r = reduce_pio2(x, q);
    switch (q) {
        case 0: return cos_poly(r);
        case 1: return -sin_poly(r);
        case 2: return -cos_poly(r);
        default: return sin_poly(r);
    }
}

inline double sin(double x) {
    if (!(std::fabs(x) < 1e15)) return x - x;
    int q;
    const double r = reduce_pio2(x, q);
    switch (q) {
        case 0: return sin_poly(r);
        case 1: return cos_poly(r);
        case 2: return -sin_poly(r);
        default: return -cos_poly(r);
    }
}

// asin Maclaurin coefficients c_n = t_n / (2n + 1), t_n = (2n)! / (4^n (n!)^2), n = 0..29,
// built by the recurrence t_n = t_{n-1} (2n - 1) / (2n).
struct AsinCoef {
    double c[30];
    AsinCoef() {
        double t = 1.0;  // (2n)! / (4^n (n!)^2)
        c[0] = 1.0;
        for (int n = 1; n < 30; ++n) {
            t = t * (double)(2 * n - 1) / (double)(2 * n);
            c[n] = t / (double)(2 * n + 1);
        }
    }
};
inline const AsinCoef& asin_coef() {
    static const AsinCoef k;
    return k;
}

inline double asin_small(double x) {  // 0 <= x <= 1/2
    const AsinCoef& K = asin_coef();
    const double z = x * x;
    double p = K.c[29];
    for (int n = 28; n >= 1; --n) p = p * z + K.c[n];
    return x + (x * z) * p;
}

inline double asin(double x) {
    if (!(x >= -1.0 && x <= 1.0)) return (x - x) / (x - x);
    const double a = std::fabs(x);
    double r;
    if (a <= 0.5) {
        r = asin_small(a);
    } else {
        r = kPio2 - 2.0 * asin_small(std::sqrt((1.0 - a) * 0.5));
    }
    return x < 0 ? -r : r;
}

inline double log(double x) {
    if (x != x || x < 0) return (x - x) / (x - x);
    if (x == 0) return -INFINITY;
    if (x == INFINITY) return x;
    int k = 0;
    if (x < 0x1p-1022) {
        x = x * 0x1p54;
        k = -54;
    }
    const uint64_t u = bits_of(x);
    k += (int)((u >> 52) & 0x7ff) - 1023;
    double m = from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  // [1, 2)
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

inline double exp(double y) {
    if (y != y) return y;
    if (y > 709.782712893384) return INFINITY;
    if (y < -745.2) return 0.0;
    const double k = std::rint(y * kInvLn2);
    const double r = (y - k * kLn2Hi) - k * kLn2Lo;
    double p = kInvFact[14];
    for (int n = 13; n >= 0; --n) p = p * r + kInvFact[n];
    int ki = (int)k;
    if (ki > 1023) {
        p = p * 0x1p1023;
        ki -= 1023;
    } else if (ki < -1022) {
        p = p * 0x1p-1000;
        ki += 1000;
    }
    return p * from_bits((uint64_t)(ki + 1023) << 52);
}

inline double pow(double x, double p) {  // x >= 0
    if (x != x || p != p) return x + p;
    if (x == 0) return p > 0 ? 0.0 : (p == 0 ? 1.0 : INFINITY);
    if (x == INFINITY) return p > 0 ? INFINITY : (p == 0 ? 1.0 : 0.0);
    return exp(p * log(x));
}

}  // namespace se_math
