// Wall shear stress on every body point.
//
// Restates computeShearStress with ib=1 and deformation_on_boundary
// (ErosionPackage/fortran_code/stokesSolver.f:3002-3096, 2618-2998):
//   wall -> bodies       barycentric interior rule, nder=2 (:2707-2727)
//   body -> same body    odd-even trapezoid on eta - eta(x) (:2744-2835)
//   body -> other body   barycentric exterior rule, nder=2 (:2893-2925)
//   jump                 (d eta/ds . t) t-tensor / 2 (:2933-2958)
//   stokeslets/rotlets   gradients (:2964-2994)
//   traction             t = -2 E n, tau = t_x p_y - t_y p_x (:3075-3091)
//
// The reference recomputes the boundary values of the source body for every
// target body (M(M-1)*5 O(n^2) sums); they depend only on the source, so
// here they are computed once per curve (k_bdval) and every (target point,
// source curve) pair then costs three passes over the source nodes.
#include <algorithm>
#include <vector>

#include "bary.hpp"
#include "common.hpp"

namespace stokes {

namespace {

constexpr int kBlk = 256;

inline int cdiv_int(int a, int b) { return (a + b - 1) / b; }

// nder=2 barycentric evaluation of source curves at kBT targets per block,
// one per lane.  Block (tb, g) of a grid (target blocks, G) takes the curves
// lst[lst_ptr[tb G + g] .. lst_ptr[tb G + g + 1]) (curve 0 the wall, 1 + b
// body b); with a mask (MODE 0), a lane takes curve c only where
// mask[b (m + 1) + c] is set for its body b.  The node data of a source curve
// (coordinates, dz, the five boundary values, 1/(z - a)) is staged in LDS
// kNodeTile nodes at a time and read by every lane with broadcast loads; a
// curve that fits one tile is staged once for all three passes.  MODE 0
// (stress): targets are the body points (tgt_x == nullptr), a body's own
// curve is skipped, and part[(g*3 + {0,1,2}) * ntar + q] = E11, E12, E22.
// G partial slots per target; the lists of one target block may hold
// different numbers of curves per g.
// MODE 1 (targets in the fluid): every curve counts and part[...] = u, v,
// vorticity.
constexpr int kBT = 256;
constexpr int kNodeTile = 256;

struct NodeTileLds {
    double x[kNodeTile], y[kNodeTile], dzr[kNodeTile], dzi[kNodeTile];
    double iar[kNodeTile], iai[kNodeTile];
    double bd[10][kNodeTile];
};

__device__ __forceinline__ void stage_nodes(NodeTileLds& T, const NodeView& nv, int p0, int cnt) {
    __syncthreads();
    for (int k = threadIdx.x; k < cnt; k += kBT) {
        const int p = p0 + k;
        T.x[k] = nv.x[p];
        T.y[k] = nv.y[p];
        T.dzr[k] = nv.dzr[p];
        T.dzi[k] = nv.dzi[p];
        T.iar[k] = nv.iar[p];
        T.iai[k] = nv.iai[p];
#pragma unroll
        for (int f = 0; f < 10; ++f) T.bd[f][k] = nv.bd[static_cast<size_t>(f) * nv.N + p];
    }
    __syncthreads();
}

template <int MODE>
__global__ __launch_bounds__(kBT) void k_bary_eval(
    const int* __restrict__ lst_ptr, const int* __restrict__ lst,
    const unsigned char* __restrict__ mask, int m, NodeView nv, const double* __restrict__ ax,
    const double* __restrict__ ay, int no, int n, const double* __restrict__ tgt_x,
    const double* __restrict__ tgt_y, int ntar, double* __restrict__ spart) {
    __shared__ NodeTileLds T;
    const int q = blockIdx.x * kBT + threadIdx.x;
    const bool valid = q < ntar;
    const int qq = valid ? q : 0;
    const int bt = MODE == 0 ? qq / n : -2;
    const double xt = MODE == 0 ? nv.x[no + qq] : tgt_x[qq];
    const double yt = MODE == 0 ? nv.y[no + qq] : tgt_y[qq];
    const cplx zt = cmk(xt, yt);
    double E11 = 0.0, E12 = 0.0, E22 = 0.0;
    const int g = blockIdx.y, G = gridDim.y;
    const unsigned char* __restrict__ mrow =
        MODE == 0 && mask ? mask + static_cast<size_t>(bt) * (m + 1) : nullptr;
    for (int li = lst_ptr[blockIdx.x * G + g]; li < lst_ptr[blockIdx.x * G + g + 1]; ++li) {
        const int c = lst[li];
        const bool skip = !valid || c == 1 + bt || (mrow && !mrow[c]);
        if (__syncthreads_count(skip ? 0 : 1) == 0) continue;  // block-uniform
        const int len = c == 0 ? no : n;
        const int off = c == 0 ? 0 : no + (c - 1) * n;
        const bool ext = c != 0;
        const bool one_tile = len <= kNodeTile;
        const cplx a = cmk(ax[c], ay[c]);
        // pass 1: S0 = sum w, N_f = sum b_f w, w = dz/(z - x)
        double v1[12];
#pragma unroll
        for (int i = 0; i < 12; ++i) v1[i] = 0.0;
        for (int t0 = 0; t0 < len; t0 += kNodeTile) {
            const int tn = min(kNodeTile, len - t0);
            stage_nodes(T, nv, off + t0, tn);
            for (int k = 0; k < tn; ++k) {
                const cplx d = cmk(T.x[k] - xt, T.y[k] - yt);
                const cplx w = cmul(cmk(T.dzr[k], T.dzi[k]), cinv(d));
                v1[0] += w.r;
                v1[1] += w.i;
#pragma unroll
                for (int f = 0; f < 5; ++f) {
                    const cplx b = cmk(T.bd[2 * f][k], T.bd[2 * f + 1][k]);
                    v1[2 + 2 * f] = __builtin_fma(b.r, w.r, __builtin_fma(-b.i, w.i, v1[2 + 2 * f]));
                    v1[3 + 2 * f] = __builtin_fma(b.r, w.i, __builtin_fma(b.i, w.r, v1[3 + 2 * f]));
                }
            }
        }
        const cplx S0 = cmk(v1[0], v1[1]);
        const cplx idv = cinv(ext ? cmk(S0.r, S0.i - kTwoPi) : S0);
        cplx u[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) u[f] = cmul(cmk(v1[2 + 2 * f], v1[3 + 2 * f]), idv);
        // pass 2: M_f = sum (b_f - u_f) w/(z-x); exterior D_a = sum w/(z-a)
        double v2[12];
#pragma unroll
        for (int i = 0; i < 12; ++i) v2[i] = 0.0;
        // node of an exterior curve within 1e-10 of the target: its term is
        // the coincident-node form, added after the loop so that the loop
        // holds no call.  Lanes that skip this curve (their own body, when a
        // block holds targets of two bodies) sit exactly on its nodes: they
        // discard the sums, so they never take that form.
        int kc = -1;
        for (int t0 = 0; t0 < len; t0 += kNodeTile) {
            const int tn = min(kNodeTile, len - t0);
            if (!one_tile) stage_nodes(T, nv, off + t0, tn);
            for (int k = 0; k < tn; ++k) {
                const cplx d = cmk(T.x[k] - xt, T.y[k] - yt);
                const double rho2 = __builtin_fma(d.r, d.r, d.i * d.i);
                const double ir = rcp64(rho2);
                const cplx id = cmk(d.r * ir, -d.i * ir);
                const cplx w = cmul(cmk(T.dzr[k], T.dzi[k]), id);
                if (ext && !skip && !(sqrt(rho2) > 1.0e-10)) {
                    kc = t0 + k;
                } else {
                    const cplx w2 = cmul(w, id);
#pragma unroll
                    for (int f = 0; f < 5; ++f) {
                        const cplx t = csub(cmk(T.bd[2 * f][k], T.bd[2 * f + 1][k]), u[f]);
                        v2[2 * f] = __builtin_fma(t.r, w2.r, __builtin_fma(-t.i, w2.i, v2[2 * f]));
                        v2[2 * f + 1] =
                            __builtin_fma(t.r, w2.i, __builtin_fma(t.i, w2.r, v2[2 * f + 1]));
                    }
                }
                if (ext) {
                    const cplx wa = cmul(w, cmk(T.iar[k], T.iai[k]));
                    v2[10] += wa.r;
                    v2[11] += wa.i;
                }
            }
        }
        if (kc >= 0) {
            const cplx d = cmk(nv.x[off + kc] - xt, nv.y[off + kc] - yt);
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                const cplx t = coincident_term(nv, off, len, kc, f, zt, a, d);
                v2[2 * f] += t.r;
                v2[2 * f + 1] += t.i;
            }
        }
        const cplx iden = cinv(ext ? cmul(csub(zt, a), cmk(v2[10], v2[11])) : S0);
        cplx du[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) du[f] = cmul(cmk(v2[2 * f], v2[2 * f + 1]), iden);
        // pass 3 (fields 0..2): P_f = sum (b_f - u_f - du_f (z-x)) w/(z-x)^2
        double v3[6] = {0, 0, 0, 0, 0, 0};
        for (int t0 = 0; t0 < len; t0 += kNodeTile) {
            const int tn = min(kNodeTile, len - t0);
            if (!one_tile) stage_nodes(T, nv, off + t0, tn);
            for (int k = 0; k < tn; ++k) {
                const cplx d = cmk(T.x[k] - xt, T.y[k] - yt);
                const cplx id = cinv(d);
                const cplx w3 = cmul(cmul(cmul(cmk(T.dzr[k], T.dzi[k]), id), id), id);
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    const cplx t =
                        csub(csub(cmk(T.bd[2 * f][k], T.bd[2 * f + 1][k]), u[f]), cmul(du[f], d));
                    v3[2 * f] = __builtin_fma(t.r, w3.r, __builtin_fma(-t.i, w3.i, v3[2 * f]));
                    v3[2 * f + 1] = __builtin_fma(t.r, w3.i, __builtin_fma(t.i, w3.r, v3[2 * f + 1]));
                }
            }
        }
        cplx dd[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) dd[f] = cscale(cmul(cmk(v3[2 * f], v3[2 * f + 1]), iden), 2.0);
        if (skip) continue;
        // StokesExteriorDLP / StokesInteriorDLP nder=2 combination
        // (stokesSolver.f:5297-5376, 5020-5098)
        double u1x = dd[0].r - du[1].r - xt * dd[1].r - yt * dd[2].r + du[3].r;
        double u1y = -dd[0].i + xt * dd[1].i - du[2].r + yt * dd[2].i - du[3].i;
        double u2x = -dd[0].i + du[1].i + xt * dd[1].i + yt * dd[2].i + du[4].r;
        double u2y = -dd[0].r + xt * dd[1].r + du[2].i + yt * dd[2].r - du[4].i;
        if (MODE == 0) {
            E11 += u1x;
            E12 += 0.5 * (u1y + u2x);
            E22 += u2y;
        } else {  // velocity (nder=1 outputs) and vorticity v_x - u_y
            E11 += du[0].r - xt * du[1].r - yt * du[2].r + u[3].r;
            E12 += -du[0].i + xt * du[1].i + yt * du[2].i + u[4].r;
            E22 += u2x - u1y;
        }
    }
    if (valid) {
        spart[(static_cast<size_t>(g) * 3 + 0) * ntar + q] = E11;
        spart[(static_cast<size_t>(g) * 3 + 1) * ntar + q] = E12;
        spart[(static_cast<size_t>(g) * 3 + 2) * ntar + q] = E22;
    }
}

// The wall (curve 0, the barycentric interior rule) at the points of the
// bodies close to it, split over ranges of kWallRange wall nodes so that a
// handful of bodies still fill the chip: grid (ceil(nct / 64), ranges), one
// wave per block, the range's nodes staged in LDS.  tq[i] is the body point
// of target i.  Pass PASS writes its partial sums part[(range * NV + v) nct +
// i] (pass 1: S0, N_f, 12 values; pass 2: M_f, 10; pass 3: P_f, 6), with
// wu = (S0, u_f) [12][nct] from pass 1 and wdu = du_f [10][nct] from pass 2
// (k_bary_eval's passes for an interior curve, stokesSolver.f:5020-5098).
constexpr int kWallRange = 256;
template <int PASS>
__global__ __launch_bounds__(64) void k_wall_bary_pass(NodeView nv, int no,
                                                       const int* __restrict__ tq, int nct,
                                                       const double* __restrict__ wu,
                                                       const double* __restrict__ wdu,
                                                       double* __restrict__ part) {
    constexpr int NV = PASS == 1 ? 12 : PASS == 2 ? 10 : 6;
    constexpr int NB = PASS == 3 ? 6 : 10;  // boundary-value components read
    __shared__ double sx[kWallRange], sy[kWallRange], szr[kWallRange], szi[kWallRange];
    __shared__ double sb[NB][kWallRange];
    const int i = blockIdx.x * 64 + threadIdx.x;
    const bool valid = i < nct;
    const int p0 = blockIdx.y * kWallRange, cnt = min(kWallRange, no - p0);
    for (int k = threadIdx.x; k < cnt; k += 64) {
        const int p = p0 + k;
        sx[k] = nv.x[p];
        sy[k] = nv.y[p];
        szr[k] = nv.dzr[p];
        szi[k] = nv.dzi[p];
#pragma unroll
        for (int f = 0; f < NB; ++f) sb[f][k] = nv.bd[static_cast<size_t>(f) * nv.N + p];
    }
    __syncthreads();
    const int q = valid ? tq[i] : 0;
    // lanes past the targets evaluate a far point and store nothing
    const double xt = valid ? nv.x[no + q] : 1.0e3, yt = valid ? nv.y[no + q] : 1.0e3;
    cplx u[5], du[3];
    if (PASS >= 2) {
#pragma unroll
        for (int f = 0; f < 5; ++f)
            u[f] = valid ? cmk(wu[(2 + 2 * f) * static_cast<size_t>(nct) + i],
                               wu[(3 + 2 * f) * static_cast<size_t>(nct) + i])
                         : cmk(0.0, 0.0);
    }
    if (PASS == 3) {
#pragma unroll
        for (int f = 0; f < 3; ++f)
            du[f] = valid ? cmk(wdu[(2 * f) * static_cast<size_t>(nct) + i],
                                wdu[(2 * f + 1) * static_cast<size_t>(nct) + i])
                          : cmk(0.0, 0.0);
    }
    double v[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = 0.0;
    for (int k = 0; k < cnt; ++k) {
        const cplx d = cmk(sx[k] - xt, sy[k] - yt);
        const cplx dz = cmk(szr[k], szi[k]);
        if (PASS == 1) {
            const cplx w = cmul(dz, cinv(d));
            v[0] += w.r;
            v[1] += w.i;
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                const cplx b = cmk(sb[2 * f][k], sb[2 * f + 1][k]);
                v[2 + 2 * f] = __builtin_fma(b.r, w.r, __builtin_fma(-b.i, w.i, v[2 + 2 * f]));
                v[3 + 2 * f] = __builtin_fma(b.r, w.i, __builtin_fma(b.i, w.r, v[3 + 2 * f]));
            }
        } else if (PASS == 2) {
            const double rho2 = __builtin_fma(d.r, d.r, d.i * d.i);
            const double ir = rcp64(rho2);
            const cplx id = cmk(d.r * ir, -d.i * ir);
            const cplx w2 = cmul(cmul(dz, id), id);
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                const cplx t = csub(cmk(sb[2 * f][k], sb[2 * f + 1][k]), u[f]);
                v[2 * f] = __builtin_fma(t.r, w2.r, __builtin_fma(-t.i, w2.i, v[2 * f]));
                v[2 * f + 1] = __builtin_fma(t.r, w2.i, __builtin_fma(t.i, w2.r, v[2 * f + 1]));
            }
        } else {
            const cplx id = cinv(d);
            const cplx w3 = cmul(cmul(cmul(dz, id), id), id);
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const cplx t = csub(csub(cmk(sb[2 * f][k], sb[2 * f + 1][k]), u[f]), cmul(du[f], d));
                v[2 * f] = __builtin_fma(t.r, w3.r, __builtin_fma(-t.i, w3.i, v[2 * f]));
                v[2 * f + 1] = __builtin_fma(t.r, w3.i, __builtin_fma(t.i, w3.r, v[2 * f + 1]));
            }
        }
    }
    if (valid) {
        const size_t o = static_cast<size_t>(blockIdx.y) * NV * static_cast<size_t>(nct) + i;
#pragma unroll
        for (int j = 0; j < NV; ++j) part[o + static_cast<size_t>(j) * nct] = v[j];
    }
}

// The passes' sums over the node ranges, in range order, one thread per
// target: STAGE 1 gives wu = (S0, u_f = N_f / S0), STAGE 2 wdu = du_f = M_f /
// S0, STAGE 3 dd_f = 2 P_f / S0 and adds the nder=2 combination of the wall
// (E11, E12, E22) to out[{0,1,2} nbp + q].
template <int STAGE>
__global__ __launch_bounds__(kBlk) void k_wall_bary_reduce(NodeView nv, int no,
                                                           const int* __restrict__ tq, int nct,
                                                           int nrange,
                                                           const double* __restrict__ part,
                                                           double* __restrict__ wu,
                                                           double* __restrict__ wdu, int nbp,
                                                           double* __restrict__ out) {
    constexpr int NV = STAGE == 1 ? 12 : STAGE == 2 ? 10 : 6;
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= nct) return;
    double v[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = 0.0;
    for (int r = 0; r < nrange; ++r) {
        const size_t o = static_cast<size_t>(r) * NV * static_cast<size_t>(nct) + i;
#pragma unroll
        for (int j = 0; j < NV; ++j) v[j] += part[o + static_cast<size_t>(j) * nct];
    }
    const size_t n1 = static_cast<size_t>(nct);
    if (STAGE == 1) {
        const cplx idv = cinv(cmk(v[0], v[1]));  // interior: 1 / S0
        wu[i] = v[0];
        wu[n1 + i] = v[1];
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            const cplx uf = cmul(cmk(v[2 + 2 * f], v[3 + 2 * f]), idv);
            wu[(2 + 2 * f) * n1 + i] = uf.r;
            wu[(3 + 2 * f) * n1 + i] = uf.i;
        }
        return;
    }
    const cplx iden = cinv(cmk(wu[i], wu[n1 + i]));
    if (STAGE == 2) {
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            const cplx d = cmul(cmk(v[2 * f], v[2 * f + 1]), iden);
            wdu[(2 * f) * n1 + i] = d.r;
            wdu[(2 * f + 1) * n1 + i] = d.i;
        }
        return;
    }
    cplx du[5], dd[3];
#pragma unroll
    for (int f = 0; f < 5; ++f) du[f] = cmk(wdu[(2 * f) * n1 + i], wdu[(2 * f + 1) * n1 + i]);
#pragma unroll
    for (int f = 0; f < 3; ++f) dd[f] = cscale(cmul(cmk(v[2 * f], v[2 * f + 1]), iden), 2.0);
    const int q = tq[i];
    const double xt = nv.x[no + q], yt = nv.y[no + q];
    const double u1x = dd[0].r - du[1].r - xt * dd[1].r - yt * dd[2].r + du[3].r;
    const double u1y = -dd[0].i + xt * dd[1].i - du[2].r + yt * dd[2].i - du[3].i;
    const double u2x = -dd[0].i + du[1].i + xt * dd[1].i + yt * dd[2].i + du[4].r;
    const double u2y = -dd[0].r + xt * dd[1].r + du[2].i + yt * dd[2].r - du[4].i;
    out[q] += u1x;
    out[static_cast<size_t>(nbp) + q] += 0.5 * (u1y + u2x);
    out[2 * static_cast<size_t>(nbp) + q] += u2y;
}

// The MODE 1 evaluation of k_bary_eval for a few targets: one wave per
// (target, curve) pair, the lanes striding over the curve's nodes and meeting
// in cross-lane butterflies after each of the three passes.  A handful of
// targets would leave k_bary_eval with one wave walking a 4096-point wall
// alone; here every (target, curve) pair is a wave of its own.  Writes
// part[(c*3 + {0,1,2}) * ntar + q] = u, v, vorticity, as k_bary_eval with one
// curve per group.
constexpr int kQoiWaves = 4;

__device__ __forceinline__ double qoi_wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__global__ __launch_bounds__(64 * kQoiWaves) void k_qoi_wave(
    NodeView nv, const double* __restrict__ ax, const double* __restrict__ ay, int no, int n,
    int m, const double* __restrict__ tgt_x, const double* __restrict__ tgt_y, int ntar,
    double* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const long long e = static_cast<long long>(blockIdx.x) * kQoiWaves + (threadIdx.x >> 6);
    if (e >= static_cast<long long>(ntar) * (m + 1)) return;  // uniform per wave
    const int c = static_cast<int>(e / ntar), q = static_cast<int>(e - static_cast<long long>(c) * ntar);
    const int len = c == 0 ? no : n;
    const int off = c == 0 ? 0 : no + (c - 1) * n;
    const bool ext = c != 0;
    const double xt = tgt_x[q], yt = tgt_y[q];
    const cplx zt = cmk(xt, yt);
    const cplx a = cmk(ax[c], ay[c]);
    const size_t N = nv.N;

    // pass 1: S0 = sum w, N_f = sum b_f w, w = dz/(z - x)
    double v1[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) v1[i] = 0.0;
    for (int k = lane; k < len; k += 64) {
        const int p = off + k;
        const cplx d = cmk(nv.x[p] - xt, nv.y[p] - yt);
        const cplx w = cmul(cmk(nv.dzr[p], nv.dzi[p]), cinv(d));
        v1[0] += w.r;
        v1[1] += w.i;
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            const cplx b = cmk(nv.bd[(2 * f) * N + p], nv.bd[(2 * f + 1) * N + p]);
            v1[2 + 2 * f] = __builtin_fma(b.r, w.r, __builtin_fma(-b.i, w.i, v1[2 + 2 * f]));
            v1[3 + 2 * f] = __builtin_fma(b.r, w.i, __builtin_fma(b.i, w.r, v1[3 + 2 * f]));
        }
    }
#pragma unroll
    for (int i = 0; i < 12; ++i) v1[i] = qoi_wave_sum(v1[i]);
    const cplx S0 = cmk(v1[0], v1[1]);
    const cplx idv = cinv(ext ? cmk(S0.r, S0.i - kTwoPi) : S0);
    cplx u[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) u[f] = cmul(cmk(v1[2 + 2 * f], v1[3 + 2 * f]), idv);

    // pass 2: M_f = sum (b_f - u_f) w/(z-x); exterior D_a = sum w/(z-a)
    double v2[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) v2[i] = 0.0;
    for (int k = lane; k < len; k += 64) {
        const int p = off + k;
        const cplx d = cmk(nv.x[p] - xt, nv.y[p] - yt);
        const double rho2 = __builtin_fma(d.r, d.r, d.i * d.i);
        const double ir = rcp64(rho2);
        const cplx id = cmk(d.r * ir, -d.i * ir);
        const cplx w = cmul(cmk(nv.dzr[p], nv.dzi[p]), id);
        if (ext && !(sqrt(rho2) > 1.0e-10)) {
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                const cplx t = coincident_term(nv, off, len, k, f, zt, a, d);
                v2[2 * f] += t.r;
                v2[2 * f + 1] += t.i;
            }
        } else {
            const cplx w2 = cmul(w, id);
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                const cplx t = csub(cmk(nv.bd[(2 * f) * N + p], nv.bd[(2 * f + 1) * N + p]), u[f]);
                v2[2 * f] = __builtin_fma(t.r, w2.r, __builtin_fma(-t.i, w2.i, v2[2 * f]));
                v2[2 * f + 1] = __builtin_fma(t.r, w2.i, __builtin_fma(t.i, w2.r, v2[2 * f + 1]));
            }
        }
        if (ext) {
            const cplx wa = cmul(w, cmk(nv.iar[p], nv.iai[p]));
            v2[10] += wa.r;
            v2[11] += wa.i;
        }
    }
#pragma unroll
    for (int i = 0; i < 12; ++i) v2[i] = qoi_wave_sum(v2[i]);
    const cplx iden = cinv(ext ? cmul(csub(zt, a), cmk(v2[10], v2[11])) : S0);
    cplx du[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) du[f] = cmul(cmk(v2[2 * f], v2[2 * f + 1]), iden);

    // pass 3 (fields 0..2): P_f = sum (b_f - u_f - du_f (z-x)) w/(z-x)^2
    double v3[6] = {0, 0, 0, 0, 0, 0};
    for (int k = lane; k < len; k += 64) {
        const int p = off + k;
        const cplx d = cmk(nv.x[p] - xt, nv.y[p] - yt);
        const cplx id = cinv(d);
        const cplx w3 = cmul(cmul(cmul(cmk(nv.dzr[p], nv.dzi[p]), id), id), id);
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const cplx t = csub(csub(cmk(nv.bd[(2 * f) * N + p], nv.bd[(2 * f + 1) * N + p]), u[f]),
                                cmul(du[f], d));
            v3[2 * f] = __builtin_fma(t.r, w3.r, __builtin_fma(-t.i, w3.i, v3[2 * f]));
            v3[2 * f + 1] = __builtin_fma(t.r, w3.i, __builtin_fma(t.i, w3.r, v3[2 * f + 1]));
        }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) v3[i] = qoi_wave_sum(v3[i]);
    if (lane != 0) return;
    cplx dd[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) dd[f] = cscale(cmul(cmk(v3[2 * f], v3[2 * f + 1]), iden), 2.0);
    // velocity (nder=1 outputs) and vorticity v_x - u_y (stokesSolver.f:5297-5376)
    const double u1y = -dd[0].i + xt * dd[1].i - du[2].r + yt * dd[2].i - du[3].i;
    const double u2x = -dd[0].i + du[1].i + xt * dd[1].i + yt * dd[2].i + du[4].r;
    part[(static_cast<size_t>(c) * 3 + 0) * ntar + q] =
        du[0].r - xt * du[1].r - yt * du[2].r + u[3].r;
    part[(static_cast<size_t>(c) * 3 + 1) * ntar + q] =
        -du[0].i + xt * du[1].i + yt * du[2].i + u[4].r;
    part[(static_cast<size_t>(c) * 3 + 2) * ntar + q] = u2x - u1y;
}

// computeQoiTargets (stokesSolver.f:3099-3323), per target: trapezoid
// pressure from every source node (and, for ibary=0, trapezoid velocity and
// vorticity), closest body node for the inside/outside test of
// classifyPointsOld (stokesSolver.f:4420-4500), stokeslet/rotlet terms.
// One block per target; out[4*t + {0,1,2,3}] = u, v, p, vorticity before the
// barycentric partial sums are added.
__global__ __launch_bounds__(kBlk) void k_qoi_trap(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ mx, const double* __restrict__ my,
    const double* __restrict__ ex, const double* __restrict__ ey, const double* __restrict__ cx,
    const double* __restrict__ cy, const double* __restrict__ den, int no, int n, int m,
    int ibary, const double* __restrict__ tx, const double* __restrict__ ty,
    double* __restrict__ out, int* __restrict__ iside) {
    __shared__ double sh[5][kBlk];
    __shared__ int si[kBlk];
    const int t = blockIdx.x;
    const double xt = tx[t], yt = ty[t];
    const int npts = no + n * m;
    double u = 0, v = 0, p = 0, w = 0;
    for (int j = threadIdx.x; j < npts; j += kBlk) {
        const double rx = xt - x[j], ry = yt - y[j];
        const double rho2 = rx * rx + ry * ry;
        const double ir2 = 1.0 / rho2;
        const double rdotm = rx * mx[j] + ry * my[j];
        const double rdotd = rx * ex[j] + ry * ey[j];
        const double mdotd = mx[j] * ex[j] + my[j] * ey[j];
        p += (2.0 * rdotm * rdotd * ir2 * ir2 - mdotd * ir2);
        if (!ibary) {
            const double routm = -rx * my[j] + ry * mx[j];
            const double routd = -rx * ey[j] + ry * ex[j];
            u += rdotm * ir2 * rdotd * ir2 * rx;
            v += rdotm * ir2 * rdotd * ir2 * ry;
            w += (rdotd * routm + rdotm * routd) * ir2 * ir2;
        }
    }
    // closest body node (first minimum in body-major order wins)
    double dmin = 1.0e300;
    int imin = 0x7fffffff;
    for (int q = threadIdx.x; q < n * m; q += kBlk) {
        const double dx = xt - x[no + q], dy = yt - y[no + q];
        const double d2 = dx * dx + dy * dy;
        if (d2 < dmin) {
            dmin = d2;
            imin = q;
        }
    }
    sh[0][threadIdx.x] = u;
    sh[1][threadIdx.x] = v;
    sh[2][threadIdx.x] = p;
    sh[3][threadIdx.x] = w;
    sh[4][threadIdx.x] = dmin;
    si[threadIdx.x] = imin;
    __syncthreads();
    for (int s = kBlk / 2; s > 0; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) {
            for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
            const double d2 = sh[4][threadIdx.x + s];
            const int i2 = si[threadIdx.x + s];
            if (d2 < sh[4][threadIdx.x] || (d2 == sh[4][threadIdx.x] && i2 < si[threadIdx.x])) {
                sh[4][threadIdx.x] = d2;
                si[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    u = sh[0][0];
    v = sh[1][0];
    p = sh[2][0];
    w = sh[3][0];
    const int base = 2 * no + 2 * n * m;
    for (int b = 0; b < m; ++b) {
        const double sto1 = den[base + 3 * b], sto2 = den[base + 3 * b + 1];
        const double rot = den[base + 3 * b + 2];
        const double rx = xt - cx[b], ry = yt - cy[b];
        const double rho2 = rx * rx + ry * ry;
        const double rdots = rx * sto1 + ry * sto2;
        u += 0.5 / kTwoPi * (-0.5 * log(rho2) * sto1 + rdots / rho2 * rx);
        v += 0.5 / kTwoPi * (-0.5 * log(rho2) * sto2 + rdots / rho2 * ry);
        p += 1.0 / kTwoPi * rdots / rho2;
        w += (ry * sto1 - rx * sto2) / rho2 / kTwoPi;
        u += rot * ry / rho2;
        v -= rot * rx / rho2;
    }
    out[4 * t + 0] = u;
    out[4 * t + 1] = v;
    out[4 * t + 2] = p;
    out[4 * t + 3] = w;
    int side = 1;
    if (m > 0) {
        const int q = no + si[0];
        const double vdn = (xt - x[q]) * px[q] + (yt - y[q]) * py[q];
        side = vdn < 0.0 ? 1 : 0;
    }
    iside[t] = side;
}

// Barycentric partial sums into the QoI, then zero targets inside a body.
__global__ void k_qoi_final(const double* __restrict__ part, int ngroups, int ntar, int m,
                            const int* __restrict__ iside, double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntar) return;
    double u = out[4 * t], v = out[4 * t + 1], p = out[4 * t + 2], w = out[4 * t + 3];
    for (int g = 0; g < ngroups; ++g) {
        u += part[(static_cast<size_t>(g) * 3 + 0) * ntar + t];
        v += part[(static_cast<size_t>(g) * 3 + 1) * ntar + t];
        w += part[(static_cast<size_t>(g) * 3 + 2) * ntar + t];
    }
    if (m > 0 && iside[t] == 0) u = v = p = w = 0.0;
    out[4 * t] = u;
    out[4 * t + 1] = v;
    out[4 * t + 2] = p;
    out[4 * t + 3] = w;
}

// Sum of the curve groups + odd-even self term + jump + stokeslet/rotlet
// gradients, then traction and shear stress.
__global__ __launch_bounds__(kBlk) void k_stress_final(
    const double* __restrict__ spart, int ngroups, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ px, const double* __restrict__ py,
    const double* __restrict__ speed, const double* __restrict__ cx,
    const double* __restrict__ cy, const double* __restrict__ ex, const double* __restrict__ ey,
    const double* __restrict__ den, const double* __restrict__ dtab, int no, int n, int m,
    double* __restrict__ tau) {
    __shared__ double sc[5][kBlk];
    const int nbp = n * m;
    const int q = blockIdx.x * kBlk + threadIdx.x;
    const bool valid = q < nbp;
    const int qq = valid ? q : 0;
    const int p = no + qq;
    const int b = qq / n, k = qq - b * n;
    const int off = no + b * n;
    double E11 = 0.0, E12 = 0.0, E22 = 0.0;
    if (valid) {
        for (int gi = 0; gi < ngroups; ++gi) {
            E11 += spart[(static_cast<size_t>(gi) * 3 + 0) * nbp + q];
            E12 += spart[(static_cast<size_t>(gi) * 3 + 1) * nbp + q];
            E22 += spart[(static_cast<size_t>(gi) * 3 + 2) * nbp + q];
        }
    }
    const double xk = x[p], yk = y[p], dxk = ex[p], dyk = ey[p];
    // odd-even self term: even-indexed targets (1-based odd) use odd sources
    const double wscale = kTwoPi / static_cast<double>(n) / kPi;
    for (int j = (k % 2 == 0) ? 1 : 0; j < n; j += 2) {
        const int pj = off + j;
        const double rx = xk - x[pj], ry = yk - y[pj];
        const double rho2 = rx * rx + ry * ry;
        const double ir2 = rcp64(rho2);
        const double ir4 = ir2 * ir2, ir6 = ir4 * ir2;
        const double nxj = px[pj], nyj = py[pj];
        const double rdotn = rx * nxj + ry * nyj;
        const double routn = rx * nyj + ry * nxj;
        const double sx = ex[pj] - dxk, sy = ey[pj] - dyk;
        const double rdots = rx * sx + ry * sy;
        const double routs = rx * sy + ry * sx;
        const double w = speed[pj] * wscale;
        E11 += (2.0 * rdotn * rdots * ir4 + 2.0 * rdots * ir4 * rx * nxj +
                2.0 * rdotn * ir4 * rx * sx - 8.0 * rdotn * rdots * ir6 * rx * rx) * w;
        E12 += (rdots * routn * ir4 + rdotn * routs * ir4 - 8.0 * rdotn * rdots * rx * ry * ir6) * w;
        E22 += (2.0 * rdotn * rdots * ir4 + 2.0 * rdots * ir4 * ry * nyj +
                2.0 * rdotn * ir4 * ry * sy - 8.0 * rdotn * rdots * ir6 * ry * ry) * w;
    }
    // jump: derivative of eta along the body (fourierDiff of eta_x + i eta_y)
    {
        // sum_j d(k-j)(eta_j - eta_k): rows of d sum to zero
        double dr = 0.0, di = 0.0;
        for (int j = 0; j < n; ++j) {
            int idx = k - j;
            idx += idx < 0 ? n : 0;
            const double wd = dtab[idx];
            dr = __builtin_fma(wd, ex[off + j] - dxk, dr);
            di = __builtin_fma(wd, ey[off + j] - dyk, di);
        }
        const double tx = -py[p], ty = px[p];
        const double dsdtx = dr / speed[p], dsdty = di / speed[p];
        const double dd = dsdtx * tx + dsdty * ty;
        E11 += 0.5 * dd * (tx * tx - ty * ty);
        E12 += 0.5 * dd * 2.0 * tx * ty;
        E22 += 0.5 * dd * (-tx * tx + ty * ty);
    }
    // stokeslets and rotlets
    const int base = 2 * no + 2 * n * m;
    for (int b0 = 0; b0 < m; b0 += kBlk) {
        __syncthreads();
        const int bb = b0 + threadIdx.x;
        if (bb < m) {
            sc[0][threadIdx.x] = cx[bb];
            sc[1][threadIdx.x] = cy[bb];
            sc[2][threadIdx.x] = den[base + 3 * bb];
            sc[3][threadIdx.x] = den[base + 3 * bb + 1];
            sc[4][threadIdx.x] = den[base + 3 * bb + 2];
        }
        __syncthreads();
        const int cnt = min(kBlk, m - b0);
        for (int t = 0; t < cnt; ++t) {
            const double rx = xk - sc[0][t], ry = yk - sc[1][t];
            const double sto1 = sc[2][t], sto2 = sc[3][t], rot = sc[4][t];
            const double rho2 = rx * rx + ry * ry;
            const double ir2 = 1.0 / rho2, ir4 = ir2 * ir2;
            const double rdots = rx * sto1 + ry * sto2;
            E11 += 0.5 / kTwoPi * (rdots * ir2 - 2.0 * rx * rx * rdots * ir4);
            E12 -= 1.0 / kTwoPi * (rdots * ir4 * rx * ry);
            E22 += 0.5 / kTwoPi * (rdots * ir2 - 2.0 * ry * ry * rdots * ir4);
            E11 -= 2.0 * rot * rx * ry * ir4;
            E12 += rot * (rx * rx - ry * ry) * ir4;
            E22 += 2.0 * rot * rx * ry * ir4;
        }
    }
    if (!valid) return;
    const double nx = px[p], ny = py[p];
    const double tx = -2.0 * E11 * nx - 2.0 * E12 * ny;
    const double ty = -2.0 * E12 * nx - 2.0 * E22 * ny;
    tau[q] = tx * ny - ty * nx;
}

// computePressure (stokesSolver.f:4644-4882), source part: pressure at body
// point q from the sources of one split, kernel (2 (r.n)(r.s)/rho^4 -
// (n.s)/rho^2) speed 2/N; the wall and other bodies by the trapezoid rule,
// the target's own body by the odd-even rule on s = eta_j - eta_q (nodes of
// the other parity, weight doubled; :4731-4787).  grid (ceil(M n / 256),
// splits); part[split][q].
__global__ __launch_bounds__(kBlk) void k_pressure_pairs(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ speed,
    const double* __restrict__ ex, const double* __restrict__ ey, int no, int n, int m,
    int split_len, double* __restrict__ part) {
    const int nbp = n * m, npts = no + nbp;
    const int q = blockIdx.x * kBlk + threadIdx.x;
    if (q >= nbp) return;
    const int gi = no + q, tb = q / n, tk = q - tb * n;
    const double xt = x[gi], yt = y[gi], exq = ex[gi], eyq = ey[gi];
    const int s0 = blockIdx.y * split_len, s1 = min(s0 + split_len, npts);
    const double wno = 2.0 / static_cast<double>(no), wn = 2.0 / static_cast<double>(n);
    double p = 0.0;
    for (int j = s0; j < s1; ++j) {
        double sx = ex[j], sy = ey[j], w;
        if (j < no) {
            w = speed[j] * wno;
        } else {
            const int sb = (j - no) / n, sk = (j - no) - sb * n;
            w = speed[j] * wn;
            if (sb == tb) {
                if (((sk ^ tk) & 1) == 0) continue;  // same parity: not an odd-even partner
                sx -= exq;
                sy -= eyq;
                w *= 2.0;
            }
        }
        const double rx = xt - x[j], ry = yt - y[j];
        const double rho2 = __builtin_fma(rx, rx, ry * ry);
        const double ir = rcp64(rho2);
        const double rdotn = __builtin_fma(rx, px[j], ry * py[j]);
        const double rdots = __builtin_fma(rx, sx, ry * sy);
        const double sdotn = __builtin_fma(px[j], sx, py[j] * sy);
        p = __builtin_fma((2.0 * rdotn * rdots * ir - sdotn) * ir, w, p);
    }
    part[static_cast<size_t>(blockIdx.y) * nbp + q] = p;
}

// Sum of the splits, jump -(d eta/ds . t) with t = (-p_y, p_x) and the
// spectral derivative of the body's density (:4826-4854), stokeslet pressure
// (r.lambda)/(2 pi rho^2) of every body (:4857-4879).
__global__ __launch_bounds__(kBlk) void k_pressure_final(
    const double* __restrict__ part, int nsplit, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ px, const double* __restrict__ py,
    const double* __restrict__ speed, const double* __restrict__ ex,
    const double* __restrict__ ey, const double* __restrict__ dtab,
    const double* __restrict__ cx, const double* __restrict__ cy,
    const double* __restrict__ den, int no, int n, int m, double* __restrict__ out) {
    const int nbp = n * m;
    const int q = blockIdx.x * kBlk + threadIdx.x;
    if (q >= nbp) return;
    double p = 0.0;
    for (int s = 0; s < nsplit; ++s) p += part[static_cast<size_t>(s) * nbp + q];
    const int gi = no + q, tb = q / n, tk = q - tb * n, off = no + tb * n;
    const double exq = ex[gi], eyq = ey[gi];
    double dx = 0.0, dy = 0.0;
    for (int k = 0; k < n; ++k) {
        int idx = tk - k;
        idx += idx < 0 ? n : 0;
        const double w = dtab[idx];
        dx = __builtin_fma(w, ex[off + k] - exq, dx);
        dy = __builtin_fma(w, ey[off + k] - eyq, dy);
    }
    const double tx = -py[gi], ty = px[gi];
    p = p - 1.0 * ((dx / speed[gi]) * tx + (dy / speed[gi]) * ty);
    const double* ext = den + 2 * no + 2 * nbp;
    const double xt = x[gi], yt = y[gi];
    for (int b = 0; b < m; ++b) {
        const double rx = xt - cx[b], ry = yt - cy[b];
        const double rho2 = rx * rx + ry * ry;
        p = p + 1.0 / kTwoPi * (rx * ext[3 * b] + ry * ext[3 * b + 1]) / rho2;
    }
    out[q] = p;
}

}  // namespace

// Source-curve groups balanced by node count (every curve, the same groups
// for every target block).  Returns the number of groups.
static int plan_curve_groups(Context& c, int tiles, std::vector<int>& gb, std::vector<int>& ge) {
    const Geometry& g = c.g;
    const int n = g.n, m = g.m, no = g.no;
    int G = std::max(1, std::min(m + 1, cdiv_int(2048, tiles)));
    // cum[c] = nodes of curves 0..c-1; group gi ends at the first curve
    // boundary past a fraction (gi+1)/G of all nodes
    std::vector<long long> cum(static_cast<size_t>(m + 2), 0);
    for (int cc = 0; cc <= m; ++cc)
        cum[static_cast<size_t>(cc + 1)] = cum[static_cast<size_t>(cc)] + (cc == 0 ? no : n);
    const long long total = cum[static_cast<size_t>(m + 1)];
    gb.clear();
    ge.clear();
    for (int gi = 0, cur = 0; gi < G && cur <= m; ++gi) {
        const long long target = total * (gi + 1) / G;
        int end = cur + 1;
        while (end <= m && cum[static_cast<size_t>(end)] < target) ++end;
        if (gi == G - 1) end = m + 1;
        gb.push_back(cur);
        ge.push_back(end);
        cur = end;
    }
    return static_cast<int>(gb.size());
}

// CSR lists (target block, group) -> curves, uploaded to c.lst_ptr / c.lst
// (which then no longer hold the cached near-curve lists).
static void upload_lists(Context& c, const std::vector<int>& ptr, const std::vector<int>& idx) {
    c.stress_lists_version = -1;
    c.lst_ptr.ensure(ptr.size());
    c.lst.ensure(std::max<size_t>(idx.size(), 1));
    STOKES_HIP(hipMemcpyAsync(c.lst_ptr.p, ptr.data(), ptr.size() * sizeof(int),
                              hipMemcpyHostToDevice, c.stream));
    if (!idx.empty())
        STOKES_HIP(hipMemcpyAsync(c.lst.p, idx.data(), idx.size() * sizeof(int),
                                  hipMemcpyHostToDevice, c.stream));
    STOKES_HIP(hipStreamSynchronize(c.stream));
}

// Every curve for every target block, in G balanced groups.
static int all_curve_lists(Context& c, int tiles) {
    std::vector<int> gb, ge;
    const int G = plan_curve_groups(c, tiles, gb, ge);
    std::vector<int> ptr(1, 0), idx;
    for (int tb = 0; tb < tiles; ++tb)
        for (int gi = 0; gi < G; ++gi) {
            for (int cc = gb[static_cast<size_t>(gi)]; cc < ge[static_cast<size_t>(gi)]; ++cc)
                idx.push_back(cc);
            ptr.push_back(static_cast<int>(idx.size()));
        }
    upload_lists(c, ptr, idx);
    return G;
}

// The shear stress's near curves of every target block: the near bodies (by
// the cluster rule) of each body in the block, in groups balanced by node
// count (largest first onto the lightest group), with the per-(body, curve)
// mask, and the points of the bodies close to the wall, whose wall goes by
// node ranges.  Cached per geometry.  Returns the number of groups.
static int near_curve_lists(Context& c, int tiles) {
    const Geometry& g = c.g;
    if (c.stress_lists_version == g.version) return c.stress_groups;
    const TileFmm& F = c.fmm;
    const int n = g.n, m = g.m, no = g.no, nbp = n * m;
    std::vector<unsigned char> mask(static_cast<size_t>(m) * (m + 1), 0);
    std::vector<std::vector<int>> blk(static_cast<size_t>(tiles));
    long long nodes = 0;
    for (int tb = 0; tb < tiles; ++tb) {
        const int b0 = tb * kBT / n, b1 = (std::min(nbp, (tb + 1) * kBT) - 1) / n;
        std::vector<int>& L = blk[static_cast<size_t>(tb)];
        const size_t first = L.size();
        for (int b = b0; b <= b1; ++b)
            for (int i = F.h_snear_ptr[static_cast<size_t>(b)]; i < F.h_snear_ptr[static_cast<size_t>(b) + 1]; ++i)
                if (F.h_snear_idx[static_cast<size_t>(i)] != b) L.push_back(1 + F.h_snear_idx[static_cast<size_t>(i)]);
        std::sort(L.begin() + static_cast<long>(first), L.end());
        L.erase(std::unique(L.begin() + static_cast<long>(first), L.end()), L.end());
        nodes += static_cast<long long>(n) * static_cast<long long>(L.size() - first);
    }
    // the wall of the bodies close to it goes by node ranges (k_wall_bary_pass)
    std::vector<int> tq;
    for (int b = 0; b < m; ++b)
        if (F.h_swall[static_cast<size_t>(b)])
            for (int k = 0; k < n; ++k) tq.push_back(b * n + k);
    c.wb_nct = static_cast<int>(tq.size());
    c.wb_tq.ensure(std::max<size_t>(tq.size(), 1));
    if (!tq.empty())
        STOKES_HIP(hipMemcpyAsync(c.wb_tq.p, tq.data(), tq.size() * sizeof(int),
                                  hipMemcpyHostToDevice, c.stream));
    for (int b = 0; b < m; ++b) {
        for (int i = F.h_snear_ptr[static_cast<size_t>(b)]; i < F.h_snear_ptr[static_cast<size_t>(b) + 1]; ++i)
            mask[static_cast<size_t>(b) * (m + 1) + 1 + F.h_snear_idx[static_cast<size_t>(i)]] = 1;
    }
    // groups: about (nodes of a block) / (heaviest curve), at most 8
    const double per_block = static_cast<double>(nodes) / tiles;
    const int G = std::max(1, std::min(8, static_cast<int>(std::lround(per_block / n))));
    std::vector<int> ptr(1, 0), idx;
    std::vector<long long> load(static_cast<size_t>(G));
    std::vector<std::vector<int>> grp(static_cast<size_t>(G));
    for (int tb = 0; tb < tiles; ++tb) {
        std::fill(load.begin(), load.end(), 0);
        for (auto& v : grp) v.clear();
        // the wall first (the heaviest), then bodies, each onto the lightest group
        for (int cc : blk[static_cast<size_t>(tb)]) {
            const size_t gi = static_cast<size_t>(std::min_element(load.begin(), load.end()) - load.begin());
            grp[gi].push_back(cc);
            load[gi] += cc == 0 ? no : n;
        }
        for (int gi = 0; gi < G; ++gi) {
            idx.insert(idx.end(), grp[static_cast<size_t>(gi)].begin(), grp[static_cast<size_t>(gi)].end());
            ptr.push_back(static_cast<int>(idx.size()));
        }
    }
    upload_lists(c, ptr, idx);
    c.lst_mask.ensure(mask.size());
    STOKES_HIP(hipMemcpyAsync(c.lst_mask.p, mask.data(), mask.size(), hipMemcpyHostToDevice, c.stream));
    STOKES_HIP(hipStreamSynchronize(c.stream));
    c.stress_lists_version = g.version;
    c.stress_groups = G;
    return G;
}

// The far bodies of the shear stress go through the cluster expansions where
// the cluster far field is allowed, available and the bodies many.
constexpr int kStressFmmBodies = 32;
static bool stress_by_clusters(const Context& c) {
    return c.ff_method == 1 && c.g.m >= kStressFmmBodies && tilefmm_available(c);
}

// Density gathered, boundary values of every curve (wall first).
static void all_boundary_values(Context& c, const double* den_dev) {
    const int m = c.g.m;
    gather_density(c, den_dev);
    std::vector<int> curves(static_cast<size_t>(m + 1));
    for (int i = 0; i <= m; ++i) curves[static_cast<size_t>(i)] = i;
    c.curves_all.ensure(curves.size());
    STOKES_HIP(hipMemcpyAsync(c.curves_all.p, curves.data(), curves.size() * sizeof(int),
                              hipMemcpyHostToDevice, c.stream));
    boundary_values(c, c.curves_all.p, m + 1, 1);
}

void shear_stress(Context& c, const double* den_dev, double* tau_dev) {
    Geometry& g = c.g;
    const int n = g.n, m = g.m, no = g.no;
    const int nbp = n * m;
    if (nbp == 0) return;
    hipStream_t s = c.stream;
    all_boundary_values(c, den_dev);
    const int tiles = cdiv_int(nbp, kBT);
    const bool clusters = stress_by_clusters(c);
    if (clusters) tilefmm_prepare(c);
    const int G = clusters ? near_curve_lists(c, tiles) : all_curve_lists(c, tiles);
    const int slots = G + (clusters ? 1 : 0);
    c.spart.ensure(static_cast<size_t>(slots) * 3 * static_cast<size_t>(nbp));
    NodeView nv{g.x.p, g.y.p, g.dzr.p, g.dzi.p, c.bd.p, g.iar.p, g.iai.p, g.mx.p, g.my.p,
                c.ex.p, c.ey.p, static_cast<size_t>(g.npts)};
    if (c.time_kernels) c.t_stress_eval.start(s);
    k_bary_eval<0><<<dim3(tiles, G), kBT, 0, s>>>(c.lst_ptr.p, c.lst.p,
                                                       clusters ? c.lst_mask.p : nullptr, m, nv,
                                                       g.ax.p, g.ay.p, no, n, nullptr, nullptr,
                                                       nbp, c.spart.p);
    STOKES_LAUNCH_CHECK();
    if (clusters && c.wb_nct > 0) {
        // the wall at the points of the bodies close to it, added to slot 0
        const int nct = c.wb_nct, nr = cdiv_int(no, kWallRange);
        c.wb_part.ensure(static_cast<size_t>(nr) * 12 * nct);
        c.wb_u.ensure(12 * static_cast<size_t>(nct));
        c.wb_du.ensure(10 * static_cast<size_t>(nct));
        const dim3 pg(cdiv_int(nct, 64), nr);
        const int rb = cdiv_int(nct, kBlk);
        k_wall_bary_pass<1><<<pg, 64, 0, s>>>(nv, no, c.wb_tq.p, nct, nullptr, nullptr, c.wb_part.p);
        k_wall_bary_reduce<1><<<rb, kBlk, 0, s>>>(nv, no, c.wb_tq.p, nct, nr, c.wb_part.p,
                                                   c.wb_u.p, c.wb_du.p, nbp, c.spart.p);
        k_wall_bary_pass<2><<<pg, 64, 0, s>>>(nv, no, c.wb_tq.p, nct, c.wb_u.p, nullptr, c.wb_part.p);
        k_wall_bary_reduce<2><<<rb, kBlk, 0, s>>>(nv, no, c.wb_tq.p, nct, nr, c.wb_part.p,
                                                   c.wb_u.p, c.wb_du.p, nbp, c.spart.p);
        k_wall_bary_pass<3><<<pg, 64, 0, s>>>(nv, no, c.wb_tq.p, nct, c.wb_u.p, c.wb_du.p,
                                               c.wb_part.p);
        k_wall_bary_reduce<3><<<rb, kBlk, 0, s>>>(nv, no, c.wb_tq.p, nct, nr, c.wb_part.p,
                                                   c.wb_u.p, c.wb_du.p, nbp, c.spart.p);
        STOKES_LAUNCH_CHECK();
    }
    if (c.time_kernels) {
        // (target point, source node) pairs of the barycentric rule
        c.t_stress_eval.stop(s, clusters ? static_cast<double>(c.fmm.stress_direct_pairs)
                                         : static_cast<double>(nbp) * static_cast<double>(g.npts - n));
    }
    if (clusters) tilefmm_stress(c, c.spart.p + static_cast<size_t>(G) * 3 * nbp);
    k_stress_final<<<cdiv_int(nbp, kBlk), kBlk, 0, s>>>(c.spart.p, slots, g.x.p, g.y.p, g.px.p,
                                                        g.py.p, g.speed.p, g.cx.p, g.cy.p, c.ex.p,
                                                        c.ey.p, den_dev, g.dtab_n.p, no, n, m,
                                                        tau_dev);
    STOKES_LAUNCH_CHECK();
    STOKES_HIP(hipStreamSynchronize(s));
}

void qoi_targets(Context& c, int ibary, const double* den_dev, int ntar, const double* tx_dev,
                 const double* ty_dev, double* out_dev) {
    Geometry& g = c.g;
    const int n = g.n, m = g.m, no = g.no;
    if (ntar <= 0) return;
    hipStream_t s = c.stream;
    c.iside.ensure(static_cast<size_t>(ntar));
    int G = 0;
    if (ibary == 1) {
        all_boundary_values(c, den_dev);
    } else {
        gather_density(c, den_dev);
    }
    k_qoi_trap<<<ntar, kBlk, 0, s>>>(g.x.p, g.y.p, g.px.p, g.py.p, g.mx.p, g.my.p, c.ex.p, c.ey.p,
                                     g.cx.p, g.cy.p, den_dev, no, n, m, ibary, tx_dev, ty_dev,
                                     out_dev, c.iside.p);
    STOKES_LAUNCH_CHECK();
    if (ibary == 1) {
        const int tiles = cdiv_int(ntar, kBT);
        // one group per curve keeps the grid wide for a handful of targets
        G = m + 1;
        if (ntar > 64) {
            std::vector<int> ptr(1, 0), idx;
            for (int tb = 0; tb < tiles; ++tb)
                for (int cc = 0; cc <= m; ++cc) {
                    idx.push_back(cc);
                    ptr.push_back(static_cast<int>(idx.size()));
                }
            upload_lists(c, ptr, idx);
        }
        c.spart.ensure(static_cast<size_t>(G) * 3 * static_cast<size_t>(ntar));
        NodeView nv{g.x.p, g.y.p, g.dzr.p, g.dzi.p, c.bd.p, g.iar.p, g.iai.p, g.mx.p, g.my.p,
                    c.ex.p, c.ey.p, static_cast<size_t>(g.npts)};
        if (ntar <= 64) {
            // few targets: one wave per (target, curve)
            const long long pairs = static_cast<long long>(ntar) * G;
            k_qoi_wave<<<static_cast<int>((pairs + kQoiWaves - 1) / kQoiWaves), 64 * kQoiWaves, 0,
                         s>>>(nv, g.ax.p, g.ay.p, no, n, m, tx_dev, ty_dev, ntar, c.spart.p);
        } else {
            k_bary_eval<1><<<dim3(tiles, G), kBT, 0, s>>>(c.lst_ptr.p, c.lst.p, nullptr, m, nv,
                                                           g.ax.p, g.ay.p, no, n, tx_dev, ty_dev,
                                                           ntar, c.spart.p);
        }
        STOKES_LAUNCH_CHECK();
    }
    k_qoi_final<<<cdiv_int(ntar, kBlk), kBlk, 0, s>>>(c.spart.p, G, ntar, m, c.iside.p, out_dev);
    STOKES_LAUNCH_CHECK();
    STOKES_HIP(hipStreamSynchronize(s));
}

void pressure(Context& c, const double* den_dev, double* out_dev) {
    Geometry& g = c.g;
    const int n = g.n, m = g.m, no = g.no, nbp = n * m;
    if (nbp == 0) return;
    hipStream_t s = c.stream;
    gather_density(c, den_dev);
    const int split_len = 1024;
    const int nsplit = cdiv_int(g.npts, split_len);
    c.spart.ensure(static_cast<size_t>(nsplit) * nbp);
    k_pressure_pairs<<<dim3(cdiv_int(nbp, kBlk), nsplit), kBlk, 0, s>>>(
        g.x.p, g.y.p, g.px.p, g.py.p, g.speed.p, c.ex.p, c.ey.p, no, n, m, split_len, c.spart.p);
    STOKES_LAUNCH_CHECK();
    k_pressure_final<<<cdiv_int(nbp, kBlk), kBlk, 0, s>>>(
        c.spart.p, nsplit, g.x.p, g.y.p, g.px.p, g.py.p, g.speed.p, c.ex.p, c.ey.p, g.dtab_n.p,
        g.cx.p, g.cy.p, den_dev, no, n, m, out_dev);
    STOKES_LAUNCH_CHECK();
    STOKES_HIP(hipStreamSynchronize(s));
}

}  // namespace stokes
