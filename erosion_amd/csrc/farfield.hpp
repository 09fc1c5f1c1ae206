// Pair routines of the FP64 all-pairs double layer, shared by the direct far
// field (operator.hip) and the direct part of the cluster far field
// (tilefmm.hip).
#pragma once

#include "common.hpp"

namespace stokes {
namespace ff {

constexpr int kTile = 256;

// FP64 reciprocal for the far field: hardware estimate (4.6e-8 relative on
// MI355X) plus one Newton step, measured at <= 2.2e-15 relative error.
__device__ __forceinline__ double rcp64_1(double a) {
    double r = __builtin_amdgcn_rcp(a);
    return __builtin_fma(r, __builtin_fma(-a, r, 1.0), r);
}

// Pair term (1/pi)(r.n)(r.eta) r w / rho^4 with w = speed 2 pi / N folded
// into m = (mx, my).  The product (r.m)(r.eta) is expanded over the source's
// three moments A = mx ex, B = mx ey + my ex, C = my ey:
//   (r.m)(r.eta) = rx (A rx + B ry) + C ry^2,
// which reuses ry^2 from rho^2 and saves one multiply per pair.  The source
// tile holds (x, y), (A, B) and C.  TPT targets per lane; CHECK zeroes the
// self pair (source index == target index).
template <int TPT, bool CHECK>
__device__ __forceinline__ void ff_tile(const double2* __restrict__ sxy,
                                        const double2* __restrict__ sab,
                                        const double* __restrict__ sc, const double* xt,
                                        const double* yt, const int* self, double* u,
                                        double* v) {
#pragma unroll 4
    for (int jj = 0; jj < kTile; ++jj) {
        const double2 p = sxy[jj];
        const double2 ab = sab[jj];
        const double cc = sc[jj];
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            const double rx = xt[q] - p.x, ry = yt[q] - p.y;
            const double ry2 = ry * ry;
            const double rho2 = __builtin_fma(rx, rx, ry2);
            const double t = __builtin_fma(ab.x, rx, ab.y * ry);
            const double w = __builtin_fma(t, rx, cc * ry2);
            const double inv = rcp64_1(rho2);
            double c = w * (inv * inv);
            if (CHECK) c = (jj == self[q]) ? 0.0 : c;
            u[q] = __builtin_fma(c, rx, u[q]);
            v[q] = __builtin_fma(c, ry, v[q]);
        }
    }
}

// The same sum for TPT = 2, 4 or 8 targets with one reciprocal per source and
// group of targets: the product P of the group's rho^2, q = 1/P (estimate +
// one Newton step), and 1/rho_k^2 = q times the other factors of P.  Three of
// four hardware reciprocals become three multiplies.  Off-diagonal tiles only
// unless DIAG; padding sources sit at 1e30, so P stays finite.  DIAG: the
// tile holds some of the lane's own targets.  A pair at distance zero (the
// self pair) has rx = ry = 0, so its moment product w is exactly 0; its rho^2
// is raised to 1e-100 (every other pair has rho^2 >= 1e-20 and is unchanged),
// which keeps P, 1/P and 1/rho^4 <= 1e200 finite, and the pair adds 0.
template <int TPT, bool DIAG, int NT = kTile>
__device__ __forceinline__ void ff_tile_shared(const double2* __restrict__ sxy,
                                               const double2* __restrict__ sab,
                                               const double* __restrict__ sc, const double* xt,
                                               const double* yt, double* u, double* v) {
    static_assert(TPT == 2 || TPT == 4 || TPT == 8, "shared reciprocal over 2, 4 or 8 targets");
#pragma unroll 4
    for (int jj = 0; jj < NT; ++jj) {
        const double2 p = sxy[jj];
        const double2 ab = sab[jj];
        const double cc = sc[jj];
        double rx[TPT], ry[TPT], r2[TPT], w[TPT], inv[TPT];
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            rx[q] = xt[q] - p.x;
            ry[q] = yt[q] - p.y;
            const double ry2 = ry[q] * ry[q];
            r2[q] = __builtin_fma(rx[q], rx[q], ry2);
            if (DIAG) r2[q] = __builtin_fmax(r2[q], 1.0e-100);
            const double t = __builtin_fma(ab.x, rx[q], ab.y * ry[q]);
            w[q] = __builtin_fma(t, rx[q], cc * ry2);
        }
        if (TPT >= 4) {
            // groups of four targets (a product of eight could overflow)
#pragma unroll
            for (int g0 = 0; g0 < TPT; g0 += 4) {
                const int i0 = g0 % TPT, i1 = (g0 + 1) % TPT, i2 = (g0 + 2) % TPT,
                          i3 = (g0 + 3) % TPT;
                const double p01 = r2[i0] * r2[i1], p23 = r2[i2] * r2[i3];
                const double qq = rcp64_1(p01 * p23);
                const double q01 = qq * p23, q23 = qq * p01;  // 1/p01, 1/p23
                inv[i0] = q01 * r2[i1];
                inv[i1] = q01 * r2[i0];
                inv[i2] = q23 * r2[i3];
                inv[i3] = q23 * r2[i2];
            }
        } else {
            const double qq = rcp64_1(r2[0] * r2[1]);
            inv[0] = qq * r2[1];
            inv[1] = qq * r2[0];
        }
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            const double c = w[q] * (inv[q] * inv[q]);
            u[q] = __builtin_fma(c, rx[q], u[q]);
            v[q] = __builtin_fma(c, ry[q], v[q]);
        }
    }
}

// The far field of a true-residual evaluation.  Where bodies carry large,
// mostly cancelling densities, the running sum of a target's pair terms is
// many orders above the final value, and rounding in the sum (not in the
// terms) bounds how small a residual b - A x can be shown to be.  Each lane
// therefore keeps, per target, the sum and a correction: every pair term
// t = c rx enters through an error-free product (t + fma(c, rx, -t) = c rx)
// and an error-free sum (TwoSum, Knuth), the lost low parts gathering in the
// correction (Ogita, Rump and Oishi, "Accurate sum and dot product", 2005).
// The pair terms themselves are the plain kernel's (their rounding is a few
// units in the last place of each term).  Contraction is off, so the
// error-free transformations are evaluated as written.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// (hi, lo) += x
__device__ __forceinline__ void dd_add(double& hi, double& lo, double x) {
    double s, e;
    two_sum(hi, x, s, e);
    hi = s;
    lo += e;
}

// (hi, lo) += c * r
__device__ __forceinline__ void dd_add_prod(double& hi, double& lo, double c, double r) {
#pragma clang fp contract(off)
    const double p = c * r;
    const double ep = __builtin_fma(c, r, -p);
    double s, e;
    two_sum(hi, p, s, e);
    hi = s;
    lo += e + ep;
}

// One target per lane, each pair with its own reciprocal (estimate and one
// Newton step); DIAG as in ff_tile_shared.
template <bool DIAG, int NT = kTile>
__device__ __forceinline__ void ff_tile_single(const double2* __restrict__ sxy,
                                               const double2* __restrict__ sab,
                                               const double* __restrict__ sc, double xt,
                                               double yt, double& u, double& v) {
#pragma unroll 4
    for (int jj = 0; jj < NT; ++jj) {
        const double2 p = sxy[jj];
        const double2 ab = sab[jj];
        const double rx = xt - p.x, ry = yt - p.y;
        const double ry2 = ry * ry;
        double r2 = __builtin_fma(rx, rx, ry2);
        if (DIAG) r2 = __builtin_fmax(r2, 1.0e-100);
        const double t = __builtin_fma(ab.x, rx, ab.y * ry);
        const double w = __builtin_fma(t, rx, sc[jj] * ry2);
        const double inv = rcp64_1(r2);
        const double c = w * (inv * inv);
        u = __builtin_fma(c, rx, u);
        v = __builtin_fma(c, ry, v);
    }
}

// The pair sums of a true-residual evaluation, TPT targets per lane: each
// term formed as in the direct far field with its own full-precision
// reciprocal (a shared one adds a few roundings of every term, and the terms
// of a body's own rigid rotation are 10^6 times the sum), entering its
// target's (sum, correction) through dd_add_prod.  DIAG: the tile holds some
// of the lane's own targets; the zero-distance pair's rho^2 is raised to
// 1e-100, so its reciprocal stays finite and its term (moment product 0) is 0.
template <int TPT, bool DIAG, int NT = kTile>
__device__ __forceinline__ void ff_tile_pairs_comp(const double2* __restrict__ sxy,
                                                   const double2* __restrict__ sab,
                                                   const double* __restrict__ sc,
                                                   const double* xt, const double* yt, double* u,
                                                   double* cu, double* v, double* cv) {
#pragma unroll 2
    for (int jj = 0; jj < NT; ++jj) {
        const double2 p = sxy[jj];
        const double2 ab = sab[jj];
        const double cc = sc[jj];
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            const double rx = xt[q] - p.x, ry = yt[q] - p.y;
            const double ry2 = ry * ry;
            double rho2 = __builtin_fma(rx, rx, ry2);
            if (DIAG) rho2 = __builtin_fmax(rho2, 1.0e-100);
            const double t = __builtin_fma(ab.x, rx, ab.y * ry);
            const double w = __builtin_fma(t, rx, cc * ry2);
            const double inv = rcp64(rho2);
            const double c = w * (inv * inv);
            dd_add_prod(u[q], cu[q], c, rx);
            dd_add_prod(v[q], cv[q], c, ry);
        }
    }
}

// Coordinate of a lane's unused target slots in the FP64 far field: far from
// every source and from the padding sources at +1e30.
constexpr double kNoTarget = -1.0e30;

// Source moments of one point for the expanded pair term.
__device__ __forceinline__ void src_moments(double mx, double my, double ex, double ey,
                                            double2* ab, double* c) {
    *ab = make_double2(mx * ex, __builtin_fma(mx, ey, my * ex));
    *c = my * ey;
}


}  // namespace ff
}  // namespace stokes
