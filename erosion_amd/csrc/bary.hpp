// Device pieces shared by the near-pair matvec correction and the shear
// stress: node data of a source curve and the coincident-node branch of
// laplaceExteriorHolomorphic (ErosionPackage/fortran_code/stokesSolver.f:5492-5506).
#pragma once

#include "common.hpp"

namespace stokes {

// Node data of the source curves, read with wave-uniform addresses.
// bd holds the five boundary values as bd[(2f + {0: re, 1: im}) * N + point].
struct NodeView {
    const double *x, *y, *dzr, *dzi, *bd, *iar, *iai, *mx, *my, *ex, *ey;
    size_t N;
};

// Pass-2 numerator term of node k (field f) when |z_k - x| <= 1e-10:
//   (tmp1/tmp2 - (z_k - x) b_k) / (z_k - x)^2 / (x - a) dz_k
//   tmp1 = sum_{l != k} (b_k (z_k - a)/(z_l - a) - b_l) dz_l / (z_l - x)
//   tmp2 = sum_l dz_l / ((z_l - a)(z_l - x))
// The reference clears tmp1/tmp2 once per target, not per node; two nodes
// of one curve within 1e-10 of the same target cannot occur for distinct
// boundary nodes, so one accumulation per call is the same thing.
// Rare (a target on a node of another curve); kept out of line so the hot
// loops around it keep their registers.  The node arrays come by value: a
// reference to the caller's NodeView would make the caller keep the whole
// struct in scratch memory.
__device__ __noinline__ cplx coincident_term_p(const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               const double* __restrict__ dzr,
                                               const double* __restrict__ dzi,
                                               const double* __restrict__ bd, size_t N, int off,
                                               int len, int k, int f, cplx zt, cplx a, cplx d) {
    const cplx bk = cmk(bd[(2 * f) * N + off + k], bd[(2 * f + 1) * N + off + k]);
    const cplx zk = cmk(x[off + k], y[off + k]);
    cplx tmp1 = cmk(0, 0), tmp2 = cmk(0, 0);
    for (int l = 0; l < len; ++l) {
        const int p = off + l;
        const cplx zl = cmk(x[p], y[p]);
        const cplx dzl = cmk(dzr[p], dzi[p]);
        const cplx bl = cmk(bd[(2 * f) * N + p], bd[(2 * f + 1) * N + p]);
        const cplx izl = cinv(csub(zl, zt));
        const cplx izla = cinv(csub(zl, a));
        if (l != k) {
            cplx t = csub(cmul(cmul(bk, csub(zk, a)), izla), bl);
            tmp1 = cadd(tmp1, cmul(cmul(t, dzl), izl));
        }
        tmp2 = cadd(tmp2, cmul(cmul(izla, izl), dzl));
    }
    cplx num = csub(cdiv(tmp1, tmp2), cmul(d, bk));
    cplx d2 = cmul(d, d);
    return cmul(cdiv(cdiv(num, d2), csub(zt, a)), cmk(dzr[off + k], dzi[off + k]));
}

__device__ __forceinline__ cplx coincident_term(const NodeView& nv, int off, int len, int k,
                                                int f, cplx zt, cplx a, cplx d) {
    return coincident_term_p(nv.x, nv.y, nv.dzr, nv.dzi, nv.bd, nv.N, off, len, k, f, zt, a, d);
}

__device__ __forceinline__ cplx bdv(const NodeView& nv, int f, int p) {
    return cmk(nv.bd[(2 * f) * nv.N + p], nv.bd[(2 * f + 1) * nv.N + p]);
}

}  // namespace stokes
