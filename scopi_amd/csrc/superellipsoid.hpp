// superellipsoid.hpp -- device closest points of superellipsoid pairs.
//
// closest_points(superellipsoid<3>, superellipsoid<3>) (include/scopi/objects/methods/
// closest_points.hpp:263-733) in three launches:
//   k_se_prepare  one lane per (active particle, grid point): the unique sorted
//                 initial angles binit_xy(6) / ainit_yz(6) (superellipsoid.hpp:359-466)
//                 and the 576 surface points and normals of that grid, stored once per
//                 particle and reused by every pair the particle is in;
//   k_se_grid     one workgroup per candidate pair: first minimum, in row-major order,
//                 of |p1 - p2| + 2 (1 + n1.n2) over the 576 x 576 point pairs
//                 (closest_points.hpp:604-690), body 2's grid staged in LDS;
//   k_se_solve    a work queue over the candidate pairs, one pair per lane at a time:
//                 MINPACK hybrd1 (minpack.hpp:851-923, Powell hybrid, mode 2, factor 100,
//                 xtol 1e-10, maxfev 1000) on the 4x4 system of closest_points.hpp:282-382,
//                 then the contact record (:724-728); sphere / plane pairs with a
//                 superellipsoid run the reference's damped Newton (utils.hpp:81-118).
// All arithmetic is FP64 without contraction, and cos/sin/asin/pow are the
// portable functions of this file, so that the result is reproducible bit for bit
// (the solve is chaotic at the rounding level: hybrd1 usually stops on "no
// progress" where every rounding decides the end point).
#pragma once

#include "common.hpp"

namespace scopi_hip {

constexpr int kSeGridAngles = 24;                        // 4 x num, num = 6
constexpr int kSeGridPoints = kSeGridAngles * kSeGridAngles;  // 576

// Per-particle grids of the active superellipsoids (index = particle - n_inactive).
struct SeGrids {
    int64_t cap = 0;             // particles
    double* ang = nullptr;       // 2 * 24 per particle: a values (ainit_yz), then b values (binit_xy)
    int32_t* count = nullptr;    // 2 per particle: number of distinct a, b values
    double* pts = nullptr;       // 6 * 576 per particle: px, py, pz, nx, ny, nz planes of 576
};

// Candidate pairs and their closest points (structure of arrays, sorted by (i, j)).
struct SeCandidates {
    int64_t cap = 0;
    int32_t* i = nullptr;
    int32_t* j = nullptr;
    double* u0 = nullptr;     // 4 per pair: grid start (a1, b1, a2, b2)
    double* u = nullptr;      // 4 per pair: hybrd1 end point
    int32_t* info = nullptr;  // hybrd1 info, -1 for pairs that are not SE-SE
    int32_t* nfev = nullptr;
    double* dij = nullptr;
    double* nij = nullptr;    // 3 * cap, component-major
    double* pi = nullptr;
    double* pj = nullptr;
    int* counter = nullptr;   // work queue of the solve kernel
};

void se_grids_reserve(SeGrids& G, int64_t n);
void se_grids_free(SeGrids& G);
void se_candidates_reserve(SeCandidates& S, int64_t n);
void se_candidates_free(SeCandidates& S);

// Device closest points of every candidate pair (i, j) of S[0, n): SE-SE pairs by the
// grid + hybrd1 solve, other shape pairs by their closed forms.
void se_closest_points(const Particles& P, SeGrids& G, SeCandidates& S, int64_t n, hipStream_t s);


// The portable elementary functions on the device, elementwise (device arrays):
// fn 0 cos, 1 sin, 2 asin, 3 log, 4 exp, 5 pow(x, y) (x >= 0).
void portable_math(int fn, int64_t n, const double* x, const double* y, double* out, hipStream_t s);

// Residual / Jacobian of a sphere (kind 0) or plane (kind 1) - superellipsoid pair and the
// 4x4 solve of its Newton step, on the device (tests); layouts in superellipsoid.hip.
void newton_probe(int kind, const double* in, double* out, hipStream_t s);

}  // namespace scopi_hip
