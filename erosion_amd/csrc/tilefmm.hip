// Far field by clusters: the trapezoid double layer of well-separated
// clusters through multipole and local expansions, the rest summed directly.
//
// This replaces the reference's FMM (stokesDLPnew -> bhfmm2dparttarg,
// ErosionPackage/fortran_code/fmm/bhfmm2dpart.f:159-203, tolerance 0.5e-12)
// in the role the reference gives it: the far field of matvec_DLP_fmmbary
// when ifmm = 1 (stokesSolver.f:636-681).  It is a one-level scheme over the
// clusters of a porous geometry, not a tree: every body is a cluster, and so
// is every segment of kWallSeg consecutive wall points.
//
// Complex form of the pair term.  With r = z - w_j (target minus source),
// m_j the weighted normal and eta_j the density as complex numbers,
//   (r.m)(r.eta) r / |r|^4 = (1/4) [ a_j / r + s_j / conj(r) + conj(a_j) r / conj(r)^2 ],
//   a_j = m_j eta_j,  s_j = 2 Re(conj(m_j) eta_j).
// So a cluster s contributes, through the Cauchy sums F(z) = sum a_j/(z-w_j),
// S(z) = sum s_j/(z-w_j) and E(z) = sum a_j conj(w_j - c_s)/(z-w_j),
//   u(z) = (1/4) [ F + conj(S + E') - (z - c_s) conj(F') ],
// and S and E enter only as Q = S + E', one series.  Summed over far
// clusters into local expansions about the target cluster's centre c_t,
// with G = sum_s conj(c_t - c_s) F_s,
//   u(z) = (1/4) [ F_L + conj(Q_L - G_L') - (z - c_t) conj(F_L') ].
// For a rigid motion of a closed curve F and Q vanish term by term of the
// expansion, so the large rigid rotation densities of small bodies cancel
// inside the multipole sums instead of between translated series.
//
// Expansions: multipole about c_s scaled by the cluster radius R_s,
// f(z) = sum_n M_n R_s^n / (z - c_s)^(n+1), M_n = sum_j coef_j ((w_j-c_s)/R_s)^n;
// local about c_t scaled by R_t, f(z) = sum_k L_k ((z - c_t)/R_t)^k, with
// L_k = -(1/d) (R_t/d)^k sum_n binom(n+k, k) (-R_s/d)^n M_n, d = c_s - c_t.
// kP = 32 terms; a pair of clusters is far when R_s + R_t <= |d| / 3, so the
// truncation is below (1/3)^32 = 5e-16 of the cluster's term sizes -- the
// size of the rounding of the direct sum.  Other pairs (and the cluster
// itself) are summed directly.
//
// Kernels per operator application: on a side stream k_fmm_p2m (one block
// per cluster) and k_fmm_m2l (matrix cores; one block per target cluster and
// series); beside them k_fmm_direct (one block per target cluster and source
// tile); after both, k_fmm_l2p (local expansions, and the sum of each
// point's tiles).  Lists are built on the host once per geometry.
//
// The shear stress uses the same clusters: the far bodies' five
// boundary-value series through k_fmm_p2m_bd, k_fmm_m2l and
// k_fmm_l2p_strain (tilefmm_stress).
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "farfield.hpp"

namespace stokes {

namespace {

using namespace ff;

constexpr int kBlk = 256;
constexpr int kP = 32;                 // expansion terms
constexpr double kSep = 1.0 / 3.0;     // far: R_s + R_t <= kSep |c_s - c_t|
constexpr int kMaxBodyPts = 512;
constexpr int kWallSeg = 64;           // wall points per cluster
constexpr int kDTile = 128;            // sources per direct-sum block
constexpr int kStressSeries = 5;       // boundary-value series of the shear stress
constexpr int kGroupMax = 8;           // most children of a parent cluster
constexpr double kWallTrapH = 6.0;     // wall node spacings: trapezoid rule for the stress's wall

inline int cdiv_int(int a, int b) { return (a + b - 1) / b; }

// z^k by repeated squaring (k < 64)
__device__ __forceinline__ cplx cpow_int(cplx z, int k) {
    cplx r = cmk(1.0, 0.0), b = z;
    while (k) {
        if (k & 1) r = cmul(r, b);
        b = cmul(b, b);
        k >>= 1;
    }
    return r;
}

// Multipole coefficients of every cluster: grid nc, block 256 = 4 waves.
// Wave w forms the orders k = 8w .. 8w + 7 of two series: F, and Q = S + E'
// (the S and E terms of a far cluster enter the velocity only as
// conj(S + E'), so they are translated as one series),
//   F_k = sum_j a_j zeta_j^k,   Q_k = sum_j [ s_j zeta_j^k - (k/R) e_j zeta_j^(k-1) ],
// zeta_j = (w_j - c)/R, e_j = a_j conj(w_j - c).  For a rigid motion of a
// closed curve F and Q vanish: the large rigid rotation densities of small
// bodies cancel inside these sums.  COMP (a true-residual evaluation): each
// term enters a (sum, correction) pair through error-free products and sums,
// and the coefficients come out accurate to the terms' own rounding.  Lane l
// sums the cluster's points l, l + 64, ... (its first power by squaring, the
// next seven by products), then the wave sums its lanes.  cr holds (cx, cy,
// R) and cl (first point, count) per cluster; mp[c][q][k], q < 2: F, Q.
//
// A body also carries its stokeslet lambda and rotlet xi (stk[3b..3b+2],
// cluster nwall + b, at its centre): with r = z - c,
//   (1/4pi) [ -log|r| lambda + (r.lambda) r / |r|^2 ] + xi (r_y, -r_x) / |r|^2
//   = (1/4) [ -(1/2pi) (lambda log r + conj(conj(lambda) log r))
//             + (1/2pi) lambda + (1/2pi) conj(lambda) r / conj(r) ] + conj(i xi / r),
// which is the velocity formula above for F + c_F log(z - c), Q + c_Q
// log(z - c) + 4 i xi / (z - c), c_F = -lambda/(2pi), c_Q = -conj(lambda)/(2pi),
// plus the constant lambda / (8pi): mlog[c] = (c_F, c_Q) and Q_0 += 4 i xi.
// stk == nullptr: no stokeslets (zero).
template <bool COMP>
__global__ __launch_bounds__(256) void k_fmm_p2m(const double* __restrict__ x,
                                                 const double* __restrict__ y,
                                                 const double* __restrict__ mx,
                                                 const double* __restrict__ my,
                                                 const double* __restrict__ ex,
                                                 const double* __restrict__ ey,
                                                 const double* __restrict__ cr,
                                                 const int* __restrict__ cl,
                                                 const double* __restrict__ stk, int nwall,
                                                 double* __restrict__ mp,
                                                 double* __restrict__ mlog) {
    constexpr int kPer = kP / 4;
    const int c = blockIdx.x, w = threadIdx.x / 64, l = threadIdx.x % 64;
    // issue priority over the direct sums' waves beside it on the same SIMDs:
    // the expansion chain (P2M, M2M, M2L) sets the application's pace
    __builtin_amdgcn_s_setprio(3);
    const double cx = cr[3 * c], cy = cr[3 * c + 1], iR = 1.0 / cr[3 * c + 2];
    const int off = cl[2 * c], len = cl[2 * c + 1];
    // (sum, correction) of the real and imaginary parts: [series][order][re, im][hi, lo]
    double acc[2][kPer][2][2];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int i = 0; i < kPer; ++i)
#pragma unroll
            for (int h = 0; h < 2; ++h) acc[q][i][h][0] = acc[q][i][h][1] = 0.0;
    for (int j = l; j < len; j += 64) {
        const int p = off + j;
        const double dx = x[p] - cx, dy = y[p] - cy;
        const cplx z = cmk(dx * iR, dy * iR);
        const double mxx = mx[p], myy = my[p], exx = ex[p], eyy = ey[p];
        const cplx a = cmk(mxx * exx - myy * eyy, mxx * eyy + myy * exx);
        const double sj = 2.0 * (mxx * exx + myy * eyy);
        const cplx e = cmk(a.r * dx + a.i * dy, a.i * dx - a.r * dy);  // a conj(w - c)
        // zeta^(k-1) and zeta^k for k = 8w (the same count on the whole wave)
        cplx zkm = w == 0 ? cmk(0.0, 0.0) : cpow_int(z, kPer * w - 1);
        cplx zk = w == 0 ? cmk(1.0, 0.0) : cmul(zkm, z);
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const cplx ek = cscale(e, -static_cast<double>(kPer * w + i) * iR);  // -(k/R) e
            if constexpr (!COMP) {
                acc[0][i][0][0] = __builtin_fma(a.r, zk.r, __builtin_fma(-a.i, zk.i, acc[0][i][0][0]));
                acc[0][i][1][0] = __builtin_fma(a.r, zk.i, __builtin_fma(a.i, zk.r, acc[0][i][1][0]));
                acc[1][i][0][0] = __builtin_fma(sj, zk.r, acc[1][i][0][0]);
                acc[1][i][0][0] = __builtin_fma(ek.r, zkm.r, __builtin_fma(-ek.i, zkm.i, acc[1][i][0][0]));
                acc[1][i][1][0] = __builtin_fma(sj, zk.i, acc[1][i][1][0]);
                acc[1][i][1][0] = __builtin_fma(ek.r, zkm.i, __builtin_fma(ek.i, zkm.r, acc[1][i][1][0]));
                zkm = zk;
                zk = cmul(zk, z);
                continue;
            }
            // F += a zeta^k
            dd_add_prod(acc[0][i][0][0], acc[0][i][0][1], a.r, zk.r);
            dd_add_prod(acc[0][i][0][0], acc[0][i][0][1], -a.i, zk.i);
            dd_add_prod(acc[0][i][1][0], acc[0][i][1][1], a.r, zk.i);
            dd_add_prod(acc[0][i][1][0], acc[0][i][1][1], a.i, zk.r);
            // Q += s zeta^k - (k/R) e zeta^(k-1)
            dd_add_prod(acc[1][i][0][0], acc[1][i][0][1], sj, zk.r);
            dd_add_prod(acc[1][i][0][0], acc[1][i][0][1], ek.r, zkm.r);
            dd_add_prod(acc[1][i][0][0], acc[1][i][0][1], -ek.i, zkm.i);
            dd_add_prod(acc[1][i][1][0], acc[1][i][1][1], sj, zk.i);
            dd_add_prod(acc[1][i][1][0], acc[1][i][1][1], ek.r, zkm.i);
            dd_add_prod(acc[1][i][1][0], acc[1][i][1][1], ek.i, zkm.r);
            zkm = zk;
            zk = cmul(zk, z);
        }
    }
    // wave sums of the (sum, correction) pairs: TwoSum of the sums, the
    // corrections added
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int i = 0; i < kPer; ++i)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                double hi = acc[q][i][h][0], lo = acc[q][i][h][1];
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) {
                    if constexpr (COMP) {
                        double ss, ee;
                        two_sum(hi, __shfl_xor(hi, o), ss, ee);
                        lo = lo + __shfl_xor(lo, o) + ee;
                        hi = ss;
                    } else {
                        hi += __shfl_xor(hi, o);
                    }
                }
                acc[q][i][h][0] = COMP ? hi + lo : hi;
            }
    double lr = 0.0, li = 0.0, xi = 0.0;  // lambda, xi of a body
    if (stk != nullptr && c >= nwall) {
        const int b = c - nwall;
        lr = stk[3 * b];
        li = stk[3 * b + 1];
        xi = stk[3 * b + 2];
    }
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int i = 0; i < kPer; ++i)
            if (l == q * kPer + i) {
                const size_t o = (static_cast<size_t>(c) * 2 + q) * kP + kPer * w + i;
                const bool rot = q == 1 && w == 0 && i == 0;  // Q_0 += 4 i xi
                mp[2 * o] = acc[q][i][0][0];
                mp[2 * o + 1] = acc[q][i][1][0] + (rot ? 4.0 * xi : 0.0);
            }
    if (threadIdx.x == 0) {
        mlog[4 * c] = -lr / kTwoPi;      // c_F = -lambda / (2 pi)
        mlog[4 * c + 1] = -li / kTwoPi;
        mlog[4 * c + 2] = -lr / kTwoPi;  // c_Q = -conj(lambda) / (2 pi)
        mlog[4 * c + 3] = li / kTwoPi;
    }
}

// Multipoles of the parents (groups of nearby clusters, cluster indices nc +
// p) from their children's: grid npar, block 256.  First lane (child, n)
// scales the child's order-n coefficients of both series and forms tau^n;
// then lane (g, series, order k) sums children g and g + 4, and the four
// partials meet in LDS.  A child s (centre c_s, radius
// R_s) inside its parent (|c_s - c_P| + R_s <= R_P) moves to c_P by
//   M'_k = sum_{n <= k} binom(k, n) (rho^n M_n) tau^(k-n),
//   rho = R_s / R_P,  tau = (c_s - c_P) / R_P,
// each term at most the child's largest coefficient (rho + |tau| <= 1).  Q
// depends on the centre through E's factor conj(w - c): about c_P it is
// Q_s + conj(c_s - c_P) F_s', and F_s' has the coefficients c_F (order 0,
// the derivative of the log term) and -(n / R_s) F_(n-1); with the factor
// and the rho^n scaling these are conj(c_s - c_P) c_F (n = 0) and
// -n conj(tau) rho^(n-1) F_(n-1).  A log term c log(z - c_s) is
// c log(z - c_P) - sum_n [c R_P tau^(n+1) / (n + 1)] R_P^n / (z - c_P)^(n+1):
// the parent's log coefficient is the sum of its children's and the rest
// enters M'_n.  binm[k][n] = binom(k, n).
__global__ __launch_bounds__(256) void k_fmm_m2m(const double* __restrict__ cr,
                                                const int* __restrict__ par_ptr,
                                                const int* __restrict__ par_idx,
                                                const double* __restrict__ binm, int nc,
                                                double* __restrict__ mp,
                                                double* __restrict__ mlog) {
    static_assert(kGroupMax * kP == 256, "a lane per (child, order)");
    __builtin_amdgcn_s_setprio(3);  // as in k_fmm_p2m
    __shared__ double sbin[kP][kP];
    __shared__ cplx A[kGroupMax][2][kP];
    __shared__ cplx tp[kGroupMax][kP + 1];
    __shared__ cplx cl2[kGroupMax][2];   // log coefficients (c_F, c_Q) of each child
    __shared__ cplx part[4][2][kP];
    const int p = blockIdx.x, P = nc + p, t = threadIdx.x;
    const int e0 = par_ptr[p], nch = min(par_ptr[p + 1] - e0, kGroupMax);
    for (int e = t; e < kP * kP; e += 256) sbin[e / kP][e % kP] = binm[e];
    const double cpx = cr[3 * P], cpy = cr[3 * P + 1], RP = cr[3 * P + 2];
    const double iRP = 1.0 / RP;
    // phase 1: lane (child ci, order n) scales the child's coefficients and
    // forms tau^n
    {
        const int ci = t / kP, n = t % kP;
        if (ci < nch) {
            const int s = par_idx[e0 + ci];
            const cplx delta = cmk(cr[3 * s] - cpx, cr[3 * s + 1] - cpy);
            const cplx tau = cscale(delta, iRP);
            const double rho = cr[3 * s + 2] * iRP;
            const double* M = mp + static_cast<size_t>(s) * 2 * kP * 2;
            // rho^(n-1) and rho^n by binary powering
            double rn1 = 1.0;
            {
                double bb = rho;
                for (int k = n - 1; k > 0; k >>= 1) {
                    if (k & 1) rn1 *= bb;
                    bb *= bb;
                }
            }
            const double rn = n == 0 ? 1.0 : rn1 * rho;
            A[ci][0][n] = cscale(cmk(M[2 * n], M[2 * n + 1]), rn);
            cplx qn = cscale(cmk(M[2 * (kP + n)], M[2 * (kP + n) + 1]), rn);
            if (n == 0) {
                const cplx cF = cmk(mlog[4 * s], mlog[4 * s + 1]);
                qn = cadd(qn, cmul(cmk(delta.r, -delta.i), cF));
                cl2[ci][0] = cF;
                cl2[ci][1] = cmk(mlog[4 * s + 2], mlog[4 * s + 3]);
            } else {
                const cplx f1 = cmk(M[2 * (n - 1)], M[2 * (n - 1) + 1]);
                qn = csub(qn, cscale(cmul(cmk(tau.r, -tau.i), f1), static_cast<double>(n) * rn1));
            }
            A[ci][1][n] = qn;
            tp[ci][n] = cpow_int(tau, n);
            if (n == kP - 1) tp[ci][kP] = cpow_int(tau, kP);
        }
    }
    __syncthreads();
    // phase 2: lane (part g, series q, order k) sums children g, g + 4
    {
        const int g = t / (2 * kP), q = (t / kP) % 2, k = t % kP;
        cplx acc = cmk(0.0, 0.0);
        for (int ci = g; ci < nch; ci += 4) {
            for (int n = 0; n <= k; ++n) cfma(acc, cscale(A[ci][q][n], sbin[k][n]), tp[ci][k - n]);
            // the log term of the child about the parent's centre
            acc = csub(acc, cscale(cmul(cl2[ci][q], tp[ci][k + 1]), RP / static_cast<double>(k + 1)));
        }
        part[g][q][k] = acc;
    }
    __syncthreads();
    if (t < 2 * kP) {
        const int q = t / kP, k = t % kP;
        cplx acc = part[0][q][k];
        for (int g = 1; g < 4; ++g) acc = cadd(acc, part[g][q][k]);
        const size_t o = (static_cast<size_t>(P) * 2 + q) * kP + k;
        mp[2 * o] = acc.r;
        mp[2 * o + 1] = acc.i;
    } else if (t < 2 * kP + 4) {
        const int j = t - 2 * kP;
        double sum = 0.0;
        for (int ci = 0; ci < nch; ++ci) sum += (j & 1) ? cl2[ci][j / 2].i : cl2[ci][j / 2].r;
        mlog[4 * P + j] = sum;
    }
}

// Local expansions of every target cluster from its far clusters, on the
// matrix cores.  For a batch of 16 far sources j and one series, the
// translation is a real matrix product followed by a complex scaling:
//   Y[k][j] = sum_n B[k][n] W[n][j],   B[k][n] = binom(n + k, k),
//   W[n][j] = (-R_j/d_j)^n M_j[n]      (real and imaginary parts apart),
//   L[k]   += f_j(k) Y[k][j],          f_j(k) = -(1/d_j) (R_t/d_j)^k,
// so that v_mfma_f64_16x16x4f64 forms Y (k in two tiles of 16, n in eight
// steps of 4) and the vector ALU the powers and the scaling.  Lane l of a
// wave holds source j = l % 16 of the batch: W rows n = 4 s + l / 16 (step s)
// and Y rows k = 16 kt + l / 16 + 4 r (register r).  Grid (targets, NQ):
// block (t, q) translates series q of mp[c][NQ][k] into lc[t][NQO][k]; 4
// waves take batches of 16 sources in turn and meet in LDS.  With G (the
// operator's series F, Q), block q = 0 also forms the series
// -G = -sum_s conj(c_t - c_s) F_s into lc[t][NQ][k] (NQO = NQ + 1).  bin
// holds B in the A-operand layout: bin[(kt * 8 + s) * 64 + l] =
// binom(4 s + l / 16 + 16 kt + l % 16, 16 kt + l % 16).
typedef double v4d __attribute__((ext_vector_type(4)));

//
// mlog (or nullptr): per source cluster a log term c_q log(z - c_s) of
// series q < 2, whose local expansion about c_t is
//   c log(-d) - c sum_{k>=1} (1/k) ((z - c_t)/d)^k,
// and for series 0 the constant -c_0 (the stokeslet's lambda / (8 pi) in
// the velocity, times 4).
// logd[i] holds log(c_t - c_s) of far pair i (in far_idx's order).
template <int NQ, bool G, bool LOGS>
__global__ __launch_bounds__(256) void k_fmm_m2l(const double* __restrict__ cr,
                                                 const double* __restrict__ mp,
                                                 const int* __restrict__ far_ptr,
                                                 const int* __restrict__ far_idx,
                                                 const double* __restrict__ bin,
                                                 const double* __restrict__ mlog,
                                                 const double* __restrict__ logd,
                                                 double* __restrict__ lc) {
    constexpr int NQO = NQ + (G ? 1 : 0);
    __builtin_amdgcn_s_setprio(3);  // as in k_fmm_p2m
    __shared__ double RED[4][2][kP][2];
    const int t = blockIdx.x, q = blockIdx.y, w = threadIdx.x / 64, l = threadIdx.x % 64;
    const int j = l % 16, h = l / 16;
    const bool with_g = G && q == 0;
    double A[2][8];
#pragma unroll
    for (int kt = 0; kt < 2; ++kt)
#pragma unroll
        for (int st = 0; st < 8; ++st) A[kt][st] = bin[(kt * 8 + st) * 64 + l];
    const double ctx = cr[3 * t], cty = cr[3 * t + 1], Rt = cr[3 * t + 2];
    cplx LQ[8], LG[8];
    double ik[8];  // 1/k of this lane's rows k = 16 kt + h + 4 r (0 for k = 0)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        LQ[e] = LG[e] = cmk(0.0, 0.0);
        const int k = 16 * (e / 4) + h + 4 * (e % 4);
        ik[e] = LOGS && k > 0 ? 1.0 / k : 0.0;
    }
    const int f0 = far_ptr[t], nf = far_ptr[t + 1] - f0;
    for (int b0 = 16 * w; b0 < nf; b0 += 64) {  // uniform over the wave
        const int i = b0 + j;
        const int s = i < nf ? far_idx[f0 + i] : -1;
        cplx id = cmk(0.0, 0.0), u = cmk(0.0, 0.0), v = cmk(0.0, 0.0), cmd = cmk(0.0, 0.0);
        cplx d = cmk(0.0, 0.0), lg = cmk(0.0, 0.0);
        if (s >= 0) {
            d = cmk(cr[3 * s] - ctx, cr[3 * s + 1] - cty);
            id = cinv(d);
            u = cscale(id, -cr[3 * s + 2]);  // -R_s / d
            v = cscale(id, Rt);              // R_t / d
            cmd = cmk(-d.r, d.i);            // conj(c_t - c_s)
            if (LOGS && q < 2) lg = cmk(mlog[4 * s + 2 * q], mlog[4 * s + 2 * q + 1]);
        }
        // log term: c d f_k / k for k >= 1 (= -c (R_t/d)^k / k), and at k = 0
        // c log(-d), less c for series 0
        cplx lgd = cmk(0.0, 0.0), l0 = cmk(0.0, 0.0);
        if (LOGS && s >= 0) {
            lgd = cmul(lg, d);
            l0 = cmul(lg, cmk(logd[2 * (f0 + i)], logd[2 * (f0 + i) + 1]));
            if (q == 0) l0 = csub(l0, lg);
        }
        // W rows n = 4 st + h: (-R_s/d)^n M[n]
        const cplx u2 = cmul(u, u), u4 = cmul(u2, u2);
        cplx up = h == 0 ? cmk(1.0, 0.0) : h == 1 ? u : h == 2 ? u2 : cmul(u2, u);
        const double* M = mp + (static_cast<size_t>(s >= 0 ? s : 0) * NQ + q) * kP * 2;
        double wr[8], wi[8];
#pragma unroll
        for (int st = 0; st < 8; ++st) {
            const int nn = 4 * st + h;
            const cplx mm = s >= 0 ? cmk(M[2 * nn], M[2 * nn + 1]) : cmk(0.0, 0.0);
            const cplx wv = cmul(up, mm);
            wr[st] = wv.r;
            wi[st] = wv.i;
            up = cmul(up, u4);
        }
        v4d yr[2], yi[2];
#pragma unroll
        for (int kt = 0; kt < 2; ++kt) {
            yr[kt] = v4d{0.0, 0.0, 0.0, 0.0};
            yi[kt] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int st = 0; st < 8; ++st) {
                yr[kt] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[kt][st], wr[st], yr[kt], 0, 0, 0);
                yi[kt] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[kt][st], wi[st], yi[kt], 0, 0, 0);
            }
        }
        // f(k) for this lane's rows k = 16 kt + h + 4 r
        const cplx v2 = cmul(v, v), v4 = cmul(v2, v2), v8 = cmul(v4, v4), v16 = cmul(v8, v8);
        cplx f = cmul(cmk(-id.r, -id.i), h == 0 ? cmk(1.0, 0.0) : h == 1 ? v : h == 2 ? v2 : cmul(v2, v));
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int kt = 0; kt < 2; ++kt) {
                const int e = 4 * kt + r;
                const cplx fk = kt == 0 ? f : cmul(f, v16);
                cplx fy = cmul(fk, cmk(yr[kt][r], yi[kt][r]));
                if (LOGS) {
                    const int k = 16 * kt + h + 4 * r;
                    fy = k > 0 ? cadd(fy, cscale(cmul(lgd, fk), ik[e])) : cadd(fy, l0);
                }
                LQ[e] = cadd(LQ[e], fy);
                if (with_g) LG[e] = csub(LG[e], cmul(cmd, fy));
            }
            f = cmul(f, v4);
        }
    }
    // sum over the 16 sources of a batch (lanes with the same l / 16)
#pragma unroll
    for (int e = 0; e < 8; ++e)
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            LQ[e].r += __shfl_xor(LQ[e].r, o);
            LQ[e].i += __shfl_xor(LQ[e].i, o);
            if (with_g) {
                LG[e].r += __shfl_xor(LG[e].r, o);
                LG[e].i += __shfl_xor(LG[e].i, o);
            }
        }
    if (j == 0) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = 16 * (e / 4) + h + 4 * (e % 4);
            RED[w][0][k][0] = LQ[e].r;
            RED[w][0][k][1] = LQ[e].i;
            RED[w][1][k][0] = LG[e].r;
            RED[w][1][k][1] = LG[e].i;
        }
    }
    __syncthreads();
    const int tid = threadIdx.x;
    if (tid < (with_g ? 2 : 1) * kP) {
        const int o2 = tid / kP, k = tid % kP;
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int ww = 0; ww < 4; ++ww) {
            sr += RED[ww][o2][k][0];
            si += RED[ww][o2][k][1];
        }
        const size_t o = (static_cast<size_t>(t) * NQO + (o2 == 0 ? q : NQ)) * kP + k;
        lc[2 * o] = sr;
        lc[2 * o + 1] = si;
    }
}

// Multipoles of the five boundary-value series of the shear stress: grid
// (m + nwall, 5), block 256 as k_fmm_p2m (wave w: orders 8w .. 8w + 7).
// Source i < m is body i (cluster nwall + i): the series f of point p has the
// coefficient a = b_f(p) dz(p) / (2 pi i), so that sum_p a / (z - w_p) is
// the exterior Cauchy integral of b_f by the trapezoid rule -- the value the
// barycentric exterior rule converges to away from the curve.  Source m + k
// is wall segment k (cluster k) with the coefficient -a: the wall's interior
// Cauchy integral, (1/2 pi i) sum b_f dz / (w - z).  bd holds b_f as
// bd[(2f + {0, 1}) N + p]; mp[i][f][k].
__global__ __launch_bounds__(256) void k_fmm_p2m_bd(const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ dzr,
                                                    const double* __restrict__ dzi,
                                                    const double* __restrict__ bd, size_t N,
                                                    const double* __restrict__ cr,
                                                    const int* __restrict__ cl, int nwall, int m,
                                                    double* __restrict__ mp) {
    constexpr int kPer = kP / 4;
    const int b = blockIdx.x, f = blockIdx.y, c = b < m ? nwall + b : b - m;
    const double sgn = b < m ? 1.0 : -1.0;
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    const double cx = cr[3 * c], cy = cr[3 * c + 1], iR = 1.0 / cr[3 * c + 2];
    const int off = cl[2 * c], len = cl[2 * c + 1];
    cplx acc[kPer];
#pragma unroll
    for (int i = 0; i < kPer; ++i) acc[i] = cmk(0.0, 0.0);
    for (int j = l; j < len; j += 64) {
        const int p = off + j;
        const cplx z = cmk((x[p] - cx) * iR, (y[p] - cy) * iR);
        const cplx bdz = cmul(cmk(bd[(2 * f) * N + p], bd[(2 * f + 1) * N + p]),
                              cmk(dzr[p], dzi[p]));
        const cplx a = cmk(sgn * bdz.i / kTwoPi, -sgn * bdz.r / kTwoPi);  // +-bdz / (2 pi i)
        cplx zk = cpow_int(z, kPer * w);
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            cfma(acc[i], a, zk);
            zk = cmul(zk, z);
        }
    }
#pragma unroll
    for (int i = 0; i < kPer; ++i)
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            acc[i].r += __shfl_xor(acc[i].r, o);
            acc[i].i += __shfl_xor(acc[i].i, o);
        }
#pragma unroll
    for (int i = 0; i < kPer; ++i)
        if (l == i) {
            const size_t o = (static_cast<size_t>(b) * kStressSeries + f) * kP + kPer * w + i;
            mp[2 * o] = acc[i].r;
            mp[2 * o + 1] = acc[i].i;
        }
}

// Strain rate at every body point from the local expansions of its far
// bodies' boundary-value series (one thread per body point q, blocks of
// 256): with C_f the local series of field f about the body's centre, du_f =
// C_f' and dd_f = C_f'', combined as the exterior rule's nder=2 output
// (stokesSolver.f:5297-5376) into E11, E12, E22 at out[{0,1,2} nbp + q].
__global__ __launch_bounds__(kBlk) void k_fmm_l2p_strain(const double* __restrict__ x,
                                                         const double* __restrict__ y,
                                                         const double* __restrict__ cr,
                                                         const double* __restrict__ lc, int no,
                                                         int n, int nwall, int nbp,
                                                         double* __restrict__ out) {
    const int q = blockIdx.x * kBlk + threadIdx.x;
    if (q >= nbp) return;
    const int b = q / n, c = nwall + b, p = no + q;
    const double xt = x[p], yt = y[p];
    const double iR = 1.0 / cr[3 * c + 2];
    const cplx z = cmk((xt - cr[3 * c]) * iR, (yt - cr[3 * c + 1]) * iR);
    cplx du[kStressSeries], dd[kStressSeries];
#pragma unroll
    for (int f = 0; f < kStressSeries; ++f) {
        const double* L = lc + (static_cast<size_t>(b) * kStressSeries + f) * kP * 2;
        // Horner: value, first derivative, half the second derivative
        cplx pv = cmk(L[2 * (kP - 1)], L[2 * (kP - 1) + 1]), d1 = cmk(0.0, 0.0), d2 = cmk(0.0, 0.0);
#pragma unroll 4
        for (int k = kP - 2; k >= 0; --k) {
            d2 = cadd(cmul(d2, z), d1);
            d1 = cadd(cmul(d1, z), pv);
            pv = cadd(cmul(pv, z), cmk(L[2 * k], L[2 * k + 1]));
        }
        du[f] = cscale(d1, iR);
        dd[f] = cscale(d2, 2.0 * iR * iR);
    }

    const double u1x = dd[0].r - du[1].r - xt * dd[1].r - yt * dd[2].r + du[3].r;
    const double u1y = -dd[0].i + xt * dd[1].i - du[2].r + yt * dd[2].i - du[3].i;
    const double u2x = -dd[0].i + du[1].i + xt * dd[1].i + yt * dd[2].i + du[4].r;
    const double u2y = -dd[0].r + xt * dd[1].r + du[2].i + yt * dd[2].r - du[4].i;
    out[q] = u1x;
    out[static_cast<size_t>(nbp) + q] = 0.5 * (u1y + u2x);
    out[2 * static_cast<size_t>(nbp) + q] = u2y;
}

// The wall segments near a body away from the wall (by the cluster rule),
// summed directly by the trapezoid rule: grid m, block 256, body b's points
// on lanes tid, tid + 256 (n <= 512).  Per segment (wnear_ptr, wnear_idx:
// CSR of segment indices per body) the 64 nodes' positions and coefficients
// a_f = b_f dz / (2 pi i) are staged in LDS; the wall's interior Cauchy
// integral C_f = sum a_f / (w - z) has C_f' = sum a_f / (z - w)^2 and
// C_f'' = -2 sum a_f / (z - w)^3, combined as in k_fmm_l2p_strain and added
// to out.
__global__ __launch_bounds__(kBlk) void k_wall_strain_direct(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ dzr,
    const double* __restrict__ dzi, const double* __restrict__ bd, size_t N, int no, int n,
    int nbp, const int* __restrict__ wnear_ptr, const int* __restrict__ wnear_idx,
    double* __restrict__ out) {
    __shared__ double sx[kWallSeg], sy[kWallSeg];
    __shared__ cplx sa[kStressSeries][kWallSeg];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int e0 = wnear_ptr[b], e1 = wnear_ptr[b + 1];
    if (e0 == e1) return;  // block-uniform
    constexpr int kPts = kMaxBodyPts / kBlk;
    cplx du[kPts][kStressSeries], dd[kPts][kStressSeries];
    double xt[kPts], yt[kPts];
#pragma unroll
    for (int r = 0; r < kPts; ++r) {
        const int i = tid + r * kBlk;
        xt[r] = i < n ? x[no + b * n + i] : 1.0e30;
        yt[r] = i < n ? y[no + b * n + i] : 1.0e30;
#pragma unroll
        for (int f = 0; f < kStressSeries; ++f) du[r][f] = dd[r][f] = cmk(0.0, 0.0);
    }
    for (int e = e0; e < e1; ++e) {
        const int w0 = wnear_idx[e] * kWallSeg, cnt = min(kWallSeg, no - w0);
        __syncthreads();
        if (tid < cnt) {
            const int wp = w0 + tid;
            sx[tid] = x[wp];
            sy[tid] = y[wp];
            const cplx dz = cmk(dzr[wp], dzi[wp]);
#pragma unroll
            for (int f = 0; f < kStressSeries; ++f) {
                const cplx bdz = cmul(cmk(bd[(2 * f) * N + wp], bd[(2 * f + 1) * N + wp]), dz);
                sa[f][tid] = cmk(bdz.i / kTwoPi, -bdz.r / kTwoPi);  // bdz / (2 pi i)
            }
        }
        __syncthreads();
        for (int k = 0; k < cnt; ++k) {
#pragma unroll
            for (int r = 0; r < kPts; ++r) {
                const cplx id = cinv(cmk(xt[r] - sx[k], yt[r] - sy[k]));
                const cplx id2 = cmul(id, id);
                const cplx id3 = cmul(id2, id);
#pragma unroll
                for (int f = 0; f < kStressSeries; ++f) {
                    du[r][f] = cadd(du[r][f], cmul(sa[f][k], id2));
                    dd[r][f] = csub(dd[r][f], cscale(cmul(sa[f][k], id3), 2.0));
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kPts; ++r) {
        const int i = tid + r * kBlk;
        if (i >= n) continue;
        const int q = b * n + i;
        const double xq = xt[r], yq = yt[r];
        const cplx* u = du[r];
        const cplx* d = dd[r];
        const double u1x = d[0].r - u[1].r - xq * d[1].r - yq * d[2].r + u[3].r;
        const double u1y = -d[0].i + xq * d[1].i - u[2].r + yq * d[2].i - u[3].i;
        const double u2x = -d[0].i + u[1].i + xq * d[1].i + yq * d[2].i + u[4].r;
        const double u2y = -d[0].r + xq * d[1].r + u[2].i + yq * d[2].r - u[4].i;
        out[q] += u1x;
        out[static_cast<size_t>(nbp) + q] += 0.5 * (u1y + u2x);
        out[2 * static_cast<size_t>(nbp) + q] += u2y;
    }
}

// Direct sums: one block per work item (cluster c, tile j of c's source
// list: the tiles of its near clusters, its own included), block BT <= 128
// threads with TPT targets per lane (lane l owns the cluster's points l, l +
// BT, ..., sharing one reciprocal per source).  Writes part_u/part_v[j][point]
// for the cluster's points.  work holds (cluster, j) per item and tiles
// (first point, count) per tile.
// COMP (a true-residual evaluation): each pair term enters through
// dd_add_prod, and the corrections go to part_cu/part_cv.
template <int TPT, bool COMP>
__global__ __launch_bounds__(128) void k_fmm_direct(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ mx,
    const double* __restrict__ my, const double* __restrict__ ex, const double* __restrict__ ey,
    const int* __restrict__ cl, const int* __restrict__ work, const int* __restrict__ tile_ptr,
    const int* __restrict__ tiles, int npts, double* __restrict__ part_u,
    double* __restrict__ part_v, double* __restrict__ part_cu, double* __restrict__ part_cv) {
    __shared__ double2 sxy[kDTile], sab[kDTile];
    __shared__ double sc[kDTile];
    const int c = work[2 * blockIdx.x], j = work[2 * blockIdx.x + 1];
    const int tid = threadIdx.x, bt = blockDim.x;
    const int off = cl[2 * c], len = cl[2 * c + 1];
    const int ti = tile_ptr[c] + j;
    const int ts = tiles[2 * ti], tl = tiles[2 * ti + 1];
    for (int e = tid; e < kDTile; e += bt) {
        if (e < tl) {
            const int k = ts + e;
            sxy[e] = make_double2(x[k], y[k]);
            src_moments(mx[k], my[k], ex[k], ey[k], &sab[e], &sc[e]);
        } else {
            sxy[e] = make_double2(1.0e30, 1.0e30);
            sab[e] = make_double2(0.0, 0.0);
            sc[e] = 0.0;
        }
    }
    double xt[TPT], yt[TPT], u[TPT], v[TPT], cu[TPT], cv[TPT];
    int tp[TPT];
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        // unused slots sit where no source is (they share a product of rho^2)
        const int k = tid + q * bt;
        tp[q] = k < len ? off + k : -1;
        xt[q] = tp[q] >= 0 ? x[tp[q]] : kNoTarget;
        yt[q] = tp[q] >= 0 ? y[tp[q]] : kNoTarget;
        u[q] = v[q] = cu[q] = cv[q] = 0.0;
    }
    __syncthreads();
    // a tile that holds points of the cluster itself holds self pairs (zero
    // distance, zero moment product): the DIAG form keeps their shared
    // product finite
    const bool diag = ts < off + len && ts + tl > off;
    if constexpr (COMP) {
        if (diag)
            ff_tile_pairs_comp<TPT, true, kDTile>(sxy, sab, sc, xt, yt, u, cu, v, cv);
        else
            ff_tile_pairs_comp<TPT, false, kDTile>(sxy, sab, sc, xt, yt, u, cu, v, cv);
    } else if constexpr (TPT == 1) {
        if (diag)
            ff_tile_single<true, kDTile>(sxy, sab, sc, xt[0], yt[0], u[0], v[0]);
        else
            ff_tile_single<false, kDTile>(sxy, sab, sc, xt[0], yt[0], u[0], v[0]);
    } else {
        if (diag)
            ff_tile_shared<TPT, true, kDTile>(sxy, sab, sc, xt, yt, u, v);
        else
            ff_tile_shared<TPT, false, kDTile>(sxy, sab, sc, xt, yt, u, v);
    }
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        if (tp[q] >= 0) {
            const size_t o = static_cast<size_t>(j) * npts + tp[q];
            part_u[o] = u[q];
            part_v[o] = v[q];
            if (COMP) {
                part_cu[o] = cu[q];
                part_cv[o] = cv[q];
            }
        }
    }
}

// Local expansions at every target point, plus the point's direct partial
// sums (one thread per point, blocks of 256): wall point p belongs to cluster
// p / kWallSeg, body point p >= no to cluster nwall + (p - no) / n.
// lc[c][q][k], q < 3: F, Q and -G, and
//   u = (1/4) [ F + conj(Q - G') - (z - c_t) conj(F') ].
// Sums part_u/part_v[j][p] over the cluster's tiles j and writes the total
// to part_u/part_v[0][p].  COMP: the tiles' (sum, correction) pairs and the
// local value are added in double-double, the correction left in
// part_cu/part_cv[0][p].
// STAGED: the locals of the block's clusters (at most kL2pStage, which holds
// for bodies of 64 points or more) are first copied to LDS, so that the
// Horner loops read them there.
constexpr int kL2pStage = 16;
template <bool COMP, bool STAGED>
__global__ __launch_bounds__(kBlk) void k_fmm_l2p(const double* __restrict__ x,
                                                  const double* __restrict__ y,
                                                  const double* __restrict__ cr,
                                                  const double* __restrict__ lc,
                                                  const int* __restrict__ tile_ptr, int no, int n,
                                                  int nwall, int npts, double* __restrict__ part_u,
                                                  double* __restrict__ part_v,
                                                  double* __restrict__ part_cu,
                                                  double* __restrict__ part_cv) {
    constexpr int kLc = 3 * kP * 2;  // doubles of one cluster's locals
    __shared__ double sL[STAGED ? kL2pStage * kLc : 1];
    const int p0 = blockIdx.x * kBlk;
    const int p = p0 + threadIdx.x;
    const int c = p < no ? p / kWallSeg : nwall + (p - no) / n;
    const double* L;
    if constexpr (STAGED) {
        const int c0 = p0 < no ? p0 / kWallSeg : nwall + (p0 - no) / n;
        const int pl = min(p0 + kBlk, npts) - 1;
        const int c1 = pl < no ? pl / kWallSeg : nwall + (pl - no) / n;
        const int cnt = min(c1 - c0 + 1, kL2pStage) * kLc;
        const double* src = lc + static_cast<size_t>(c0) * kLc;
        for (int e = threadIdx.x; e < cnt; e += kBlk) sL[e] = src[e];
        __syncthreads();
        L = sL + (c - c0) * kLc;
    } else {
        L = lc + static_cast<size_t>(c) * kLc;
    }
    if (p >= npts) return;
    const double dx = x[p] - cr[3 * c], dy = y[p] - cr[3 * c + 1];
    const double iR = 1.0 / cr[3 * c + 2];
    const cplx z = cmk(dx * iR, dy * iR);
    // Horner: F and its derivative, Q, the derivative of -G
    cplx fv = cmk(L[2 * (kP - 1)], L[2 * (kP - 1) + 1]), fd = cmk(0.0, 0.0);
    cplx qv = cmk(L[2 * (2 * kP - 1)], L[2 * (2 * kP - 1) + 1]);
    cplx gv = cmk(L[2 * (3 * kP - 1)], L[2 * (3 * kP - 1) + 1]), gd = cmk(0.0, 0.0);
#pragma unroll 4
    for (int k = kP - 2; k >= 0; --k) {
        fd = cadd(cmul(fd, z), fv);
        fv = cadd(cmul(fv, z), cmk(L[2 * k], L[2 * k + 1]));
        qv = cadd(cmul(qv, z), cmk(L[2 * (kP + k)], L[2 * (kP + k) + 1]));
        gd = cadd(cmul(gd, z), gv);
        gv = cadd(cmul(gv, z), cmk(L[2 * (2 * kP + k)], L[2 * (2 * kP + k) + 1]));
    }
    const cplx A = cadd(qv, cscale(gd, iR));                    // Q - G'
    const cplx fdz = cscale(fd, iR);                            // F'
    const cplx B = cmul(cmk(dx, dy), cmk(fdz.r, -fdz.i));       // (z - c_t) conj(F')
    double u = 0.25 * (fv.r + A.r - B.r), v = 0.25 * (fv.i - A.i - B.i);
    const int nt = tile_ptr[c + 1] - tile_ptr[c];
    if constexpr (COMP) {
        double lu = 0.0, lv = 0.0;
        for (int j = 0; j < nt; ++j) {
            const size_t o = static_cast<size_t>(j) * npts + p;
            dd_add(u, lu, part_u[o]);
            dd_add(v, lv, part_v[o]);
            lu += part_cu[o];
            lv += part_cv[o];
        }
        part_cu[p] = lu;
        part_cv[p] = lv;
    } else {
        // the tiles' partials in order, loaded 8 at a time (one wait per
        // batch instead of one per tile)
        for (int j0 = 0; j0 < nt; j0 += 8) {
            double tu[8], tv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const size_t o = static_cast<size_t>(min(j0 + q, nt - 1)) * npts + p;
                tu[q] = part_u[o];
                tv[q] = part_v[o];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (j0 + q < nt) {
                    u += tu[q];
                    v += tv[q];
                }
            }
        }
    }
    part_u[p] = u;
    part_v[p] = v;
}

// log(c_t - c_s) of every far pair (t, s) of a CSR list, the branch the same
// for both series of a pair: grid = target clusters, block 256.
__global__ __launch_bounds__(256) void k_fmm_logd(const double* __restrict__ cr,
                                                 const int* __restrict__ far_ptr,
                                                 const int* __restrict__ far_idx,
                                                 double* __restrict__ logd) {
    const int t = blockIdx.x;
    const double ctx = cr[3 * t], cty = cr[3 * t + 1];
    for (int i = far_ptr[t] + static_cast<int>(threadIdx.x); i < far_ptr[t + 1]; i += 256) {
        const int s = far_idx[i];
        const double dx = cr[3 * s] - ctx, dy = cr[3 * s + 1] - cty;
        logd[2 * static_cast<size_t>(i)] = 0.5 * log(dx * dx + dy * dy);
        logd[2 * static_cast<size_t>(i) + 1] = atan2(-dy, -dx);
    }
}

}  // namespace

bool tilefmm_available(const Context& c) {
    const Geometry& g = c.g;
    return g.m >= 1 && g.n >= 4 && g.n <= kMaxBodyPts;
}

namespace {

struct Cluster {
    double cx, cy, r;
    int start, len;
};

// Centre (given, or the bounding box's) and radius of points [start, start + len).
Cluster make_cluster(const Geometry& g, int start, int len, bool bbox, double cx, double cy) {
    if (bbox) {
        double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
        for (int k = start; k < start + len; ++k) {
            x0 = std::min(x0, g.h_x[static_cast<size_t>(k)]);
            x1 = std::max(x1, g.h_x[static_cast<size_t>(k)]);
            y0 = std::min(y0, g.h_y[static_cast<size_t>(k)]);
            y1 = std::max(y1, g.h_y[static_cast<size_t>(k)]);
        }
        cx = 0.5 * (x0 + x1);
        cy = 0.5 * (y0 + y1);
    }
    double r2 = 0.0;
    for (int k = start; k < start + len; ++k) {
        const double dx = g.h_x[static_cast<size_t>(k)] - cx, dy = g.h_y[static_cast<size_t>(k)] - cy;
        r2 = std::max(r2, dx * dx + dy * dy);
    }
    return Cluster{cx, cy, std::max(std::sqrt(r2), 1e-300), start, len};
}

// Groups of at most kGroup clusters: the clusters idx[lo, hi) are halved by
// their centres along the longer side of the centres' bounding box until a
// part holds kGroup or fewer.
constexpr int kGroup = kGroupMax;
void bisect_groups(const std::vector<Cluster>& C, std::vector<int>& idx, int lo, int hi,
                   std::vector<std::pair<int, int>>& out) {
    if (hi - lo <= kGroup) {
        out.emplace_back(lo, hi);
        return;
    }
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
    for (int i = lo; i < hi; ++i) {
        const Cluster& c = C[static_cast<size_t>(idx[static_cast<size_t>(i)])];
        x0 = std::min(x0, c.cx);
        x1 = std::max(x1, c.cx);
        y0 = std::min(y0, c.cy);
        y1 = std::max(y1, c.cy);
    }
    const bool byx = x1 - x0 >= y1 - y0;
    const int mid = lo + (hi - lo) / 2;
    std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi, [&](int a, int b) {
        const Cluster &ca = C[static_cast<size_t>(a)], &cb = C[static_cast<size_t>(b)];
        const double ka = byx ? ca.cx : ca.cy, kb = byx ? cb.cx : cb.cy;
        return ka < kb || (ka == kb && a < b);
    });
    bisect_groups(C, idx, lo, mid, out);
    bisect_groups(C, idx, mid, hi, out);
}

// The far rule R_s + R_t <= kSep |c_s - c_t|, squared (no square root); every
// list uses this one predicate, so a pair is near or far the same way in all.
bool far_pair(const Cluster& S, const Cluster& T) {
    const double dx = S.cx - T.cx, dy = S.cy - T.cy, rr = S.r + T.r;
    return rr * rr <= kSep * kSep * (dx * dx + dy * dy);
}

}  // namespace

// Clusters, their lists and the expansion buffers of the current geometry
// (once per geometry).
void tilefmm_prepare(Context& c) {
    Geometry& g = c.g;
    TileFmm& F = c.fmm;
    if (!tilefmm_available(c)) throw Error("the cluster far field needs bodies of 4 to 512 points");
    if (F.version == g.version) return;
    const int n = g.n, m = g.m, no = g.no;
    std::vector<Cluster> C;
    for (int w0 = 0; w0 < no; w0 += kWallSeg)
        C.push_back(make_cluster(g, w0, std::min(kWallSeg, no - w0), true, 0.0, 0.0));
    const int nwall = static_cast<int>(C.size());
    for (int b = 0; b < m; ++b)
        C.push_back(make_cluster(g, no + b * n, n, false, g.h_cx[static_cast<size_t>(b)],
                                 g.h_cy[static_cast<size_t>(b)]));
    const int nc = static_cast<int>(C.size());
    // parents: groups of two or more nearby clusters, each enclosing its
    // children (|c_s - c_P| + R_s <= R_P), centred in the box of the
    // children's discs
    std::vector<int> order(static_cast<size_t>(nc));
    for (int i = 0; i < nc; ++i) order[static_cast<size_t>(i)] = i;
    std::vector<std::pair<int, int>> groups;
    bisect_groups(C, order, 0, nc, groups);
    std::vector<Cluster> PC;                 // parent clusters
    std::vector<int> gpar(groups.size(), -1);  // parent of each group (-1: none)
    std::vector<int> par_ptr(1, 0), par_idx;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        const int lo = groups[gi].first, hi = groups[gi].second;
        if (hi - lo < 2) continue;
        double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
        for (int i = lo; i < hi; ++i) {
            const Cluster& c = C[static_cast<size_t>(order[static_cast<size_t>(i)])];
            x0 = std::min(x0, c.cx - c.r);
            x1 = std::max(x1, c.cx + c.r);
            y0 = std::min(y0, c.cy - c.r);
            y1 = std::max(y1, c.cy + c.r);
        }
        Cluster P{0.5 * (x0 + x1), 0.5 * (y0 + y1), 0.0, 0, 0};
        for (int i = lo; i < hi; ++i) {
            const int ch = order[static_cast<size_t>(i)];
            const Cluster& c = C[static_cast<size_t>(ch)];
            P.r = std::max(P.r, std::hypot(c.cx - P.cx, c.cy - P.cy) + c.r);
            par_idx.push_back(ch);
        }
        // enclosing with a margin of one part in 1e12: the shift's series
        // needs rho + |tau| <= 1
        P.r *= 1.0 + 1e-12;
        par_ptr.push_back(static_cast<int>(par_idx.size()));
        gpar[gi] = static_cast<int>(PC.size());
        PC.push_back(P);
    }
    const int npar = static_cast<int>(PC.size());
    std::vector<double> cr(3 * static_cast<size_t>(nc + npar));
    std::vector<int> cl(2 * static_cast<size_t>(nc));
    for (int t = 0; t < nc + npar; ++t) {
        const Cluster& c = t < nc ? C[static_cast<size_t>(t)] : PC[static_cast<size_t>(t - nc)];
        cr[3 * static_cast<size_t>(t)] = c.cx;
        cr[3 * static_cast<size_t>(t) + 1] = c.cy;
        cr[3 * static_cast<size_t>(t) + 2] = c.r;
        if (t < nc) {
            cl[2 * static_cast<size_t>(t)] = c.start;
            cl[2 * static_cast<size_t>(t) + 1] = c.len;
        }
    }
    // two-level far lists: a far parent stands for all its children (a
    // parent far from t implies every child far from t); the others are
    // taken child by child
    std::vector<int> far2_ptr(static_cast<size_t>(nc) + 1, 0), far2_idx;
    auto push_far2 = [&](int, int src, const Cluster&) { far2_idx.push_back(src); };
    for (int t = 0; t < nc; ++t) {
        const Cluster& T = C[static_cast<size_t>(t)];
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            const int pi = gpar[gi];
            if (pi >= 0 && far_pair(PC[static_cast<size_t>(pi)], T)) {
                push_far2(t, nc + pi, PC[static_cast<size_t>(pi)]);
                continue;
            }
            for (int i = groups[gi].first; i < groups[gi].second; ++i) {
                const int sidx = order[static_cast<size_t>(i)];
                if (sidx != t && far_pair(C[static_cast<size_t>(sidx)], T))
                    push_far2(t, sidx, C[static_cast<size_t>(sidx)]);
            }
        }
        far2_ptr[static_cast<size_t>(t) + 1] = static_cast<int>(far2_idx.size());
    }
    std::vector<int> far_ptr(static_cast<size_t>(nc) + 1, 0), far_idx,
        tile_ptr(static_cast<size_t>(nc) + 1, 0), tiles, nb_ptr(static_cast<size_t>(nc) + 1, 0),
        nb_idx;
    std::vector<std::pair<int, int>> near;  // (first point, count)
    long long direct = 0;
    for (int t = 0; t < nc; ++t) {
        const Cluster& T = C[static_cast<size_t>(t)];
        near.clear();
        long long nsrc = 0;
        for (int s = 0; s < nc; ++s) {
            const Cluster& S = C[static_cast<size_t>(s)];
            if (s != t && far_pair(S, T)) {
                far_idx.push_back(s);
            } else {
                near.emplace_back(S.start, S.len);
                nsrc += S.len;
                if (s >= nwall) nb_idx.push_back(s - nwall);
            }
        }
        // near clusters adjacent in memory (runs of wall segments) merge
        // into one range before it is cut into tiles
        std::sort(near.begin(), near.end());
        for (size_t i = 0; i < near.size();) {
            const int st = near[i].first;
            int en = st + near[i].second;
            size_t j = i + 1;
            while (j < near.size() && near[j].first == en) en += near[j++].second;
            for (int k0 = st; k0 < en; k0 += kDTile) {
                tiles.push_back(k0);
                tiles.push_back(std::min(kDTile, en - k0));
            }
            i = j;
        }
        far_ptr[static_cast<size_t>(t) + 1] = static_cast<int>(far_idx.size());
        tile_ptr[static_cast<size_t>(t) + 1] = static_cast<int>(tiles.size() / 2);
        nb_ptr[static_cast<size_t>(t) + 1] = static_cast<int>(nb_idx.size());
        direct += static_cast<long long>(T.len) * nsrc;
    }
    // shear stress: far bodies of each body by the same rule (expansions),
    // near bodies (itself included) for the barycentric rule.  The wall:
    // a body whose points all lie kWallTrapH node spacings h or more from
    // every wall node (h the largest spacing of the node's segment) takes the
    // wall by the trapezoid rule -- its error decays like exp(-2 pi d / h),
    // 4e-17 of the terms at 6 h -- far segments through the expansions, near
    // ones summed directly; a body closer than that takes the whole wall by
    // the barycentric interior rule.
    // per segment: its largest node spacing times kWallTrapH, and its box
    std::vector<double> dseg(static_cast<size_t>(nwall), 0.0);
    std::vector<double> wbox(4 * static_cast<size_t>(nwall));
    for (int k = 0; k < nwall; ++k) {
        double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
        for (int p = k * kWallSeg; p < std::min(no, (k + 1) * kWallSeg); ++p)
            dseg[static_cast<size_t>(k)] = std::max(dseg[static_cast<size_t>(k)],
                                                    g.h_speed[static_cast<size_t>(p)]);
        dseg[static_cast<size_t>(k)] *= kWallTrapH * kTwoPi / no;
        for (int p = k * kWallSeg; p < std::min(no, (k + 1) * kWallSeg); ++p) {
            x0 = std::min(x0, g.h_x[static_cast<size_t>(p)]);
            x1 = std::max(x1, g.h_x[static_cast<size_t>(p)]);
            y0 = std::min(y0, g.h_y[static_cast<size_t>(p)]);
            y1 = std::max(y1, g.h_y[static_cast<size_t>(p)]);
        }
        wbox[4 * static_cast<size_t>(k)] = x0;
        wbox[4 * static_cast<size_t>(k) + 1] = x1;
        wbox[4 * static_cast<size_t>(k) + 2] = y0;
        wbox[4 * static_cast<size_t>(k) + 3] = y1;
    }
    auto wall_close = [&](int b) {
        const size_t o = static_cast<size_t>(no) + static_cast<size_t>(b) * n;
        const Cluster& B = C[static_cast<size_t>(nwall + b)];
        for (int k = 0; k < nwall; ++k) {
            const Cluster& W = C[static_cast<size_t>(k)];
            const double dwall = dseg[static_cast<size_t>(k)];
            if (std::hypot(B.cx - W.cx, B.cy - W.cy) - B.r - W.r >= dwall) continue;
            const double x0 = wbox[4 * static_cast<size_t>(k)] - dwall,
                         x1 = wbox[4 * static_cast<size_t>(k) + 1] + dwall,
                         y0 = wbox[4 * static_cast<size_t>(k) + 2] - dwall,
                         y1 = wbox[4 * static_cast<size_t>(k) + 3] + dwall;
            for (int i = 0; i < n; ++i) {
                const double xb = g.h_x[o + i], yb = g.h_y[o + i];
                if (xb < x0 || xb > x1 || yb < y0 || yb > y1) continue;
                for (int p = k * kWallSeg; p < std::min(no, (k + 1) * kWallSeg); ++p) {
                    const double dx = xb - g.h_x[static_cast<size_t>(p)],
                                 dy = yb - g.h_y[static_cast<size_t>(p)];
                    if (dx * dx + dy * dy < dwall * dwall) return true;
                }
            }
        }
        return false;
    };
    std::vector<int> sfar_ptr(static_cast<size_t>(m) + 1, 0), sfar_idx;
    std::vector<int> swn_ptr(static_cast<size_t>(m) + 1, 0), swn_idx;
    F.h_snear_ptr.assign(static_cast<size_t>(m) + 1, 0);
    F.h_snear_idx.clear();
    F.h_swall.assign(static_cast<size_t>(m), 1);
    long long sdirect = 0;
    int away = 0;
    for (int b = 0; b < m; ++b) {
        const Cluster& T = C[static_cast<size_t>(nwall + b)];
        for (int e = 0; e < m; ++e) {
            const Cluster& S = C[static_cast<size_t>(nwall + e)];
            if (e != b && far_pair(S, T))
                sfar_idx.push_back(e);
            else
                F.h_snear_idx.push_back(e);
        }
        const bool close = wall_close(b);
        F.h_swall[static_cast<size_t>(b)] = close ? 1 : 0;
        if (!close) {
            ++away;
            for (int k = 0; k < nwall; ++k) {
                if (far_pair(C[static_cast<size_t>(k)], T))
                    sfar_idx.push_back(m + k);
                else
                    swn_idx.push_back(k);
            }
        }
        sfar_ptr[static_cast<size_t>(b) + 1] = static_cast<int>(sfar_idx.size());
        swn_ptr[static_cast<size_t>(b) + 1] = static_cast<int>(swn_idx.size());
        F.h_snear_ptr[static_cast<size_t>(b) + 1] = static_cast<int>(F.h_snear_idx.size());
        const long long nnear = F.h_snear_ptr[static_cast<size_t>(b) + 1] - F.h_snear_ptr[static_cast<size_t>(b)];
        sdirect += static_cast<long long>(n) * ((close ? no : 0) + static_cast<long long>(n) * (nnear - 1));
    }
    F.stress_direct_pairs = sdirect;
    F.stress_wall = away > 0;
    // sources of the stress expansions: bodies, then wall segments
    std::vector<double> scr(3 * static_cast<size_t>(m + nwall));
    for (int i = 0; i < m + nwall; ++i) {
        const Cluster& S = C[static_cast<size_t>(i < m ? nwall + i : i - m)];
        scr[3 * static_cast<size_t>(i)] = S.cx;
        scr[3 * static_cast<size_t>(i) + 1] = S.cy;
        scr[3 * static_cast<size_t>(i) + 2] = S.r;
    }
    hipStream_t s = c.stream;
    auto up = [&](auto& buf, const auto& vec) {
        buf.ensure(std::max<size_t>(vec.size(), 1));
        if (!vec.empty())
            STOKES_HIP(hipMemcpyAsync(buf.p, vec.data(), vec.size() * sizeof(vec[0]),
                                      hipMemcpyHostToDevice, s));
    };
    up(F.cr, cr);
    up(F.cl, cl);
    up(F.far_ptr, far_ptr);
    up(F.far_idx, far_idx);
    up(F.tile_ptr, tile_ptr);
    up(F.tiles, tiles);
    up(F.sfar_ptr, sfar_ptr);
    up(F.sfar_idx, sfar_idx);
    up(F.swn_ptr, swn_ptr);
    up(F.swn_idx, swn_idx);
    up(F.scr, scr);
    up(F.nb_ptr, nb_ptr);
    up(F.nb_idx, nb_idx);
    up(F.par_ptr, par_ptr);
    up(F.par_idx, par_idx);
    up(F.far2_ptr, far2_ptr);
    up(F.far2_idx, far2_idx);
    // log(c_t - c_s) of every far pair of both lists, on the device
    F.logd.ensure(2 * std::max<size_t>(far_idx.size(), 1));
    F.logd2.ensure(2 * std::max<size_t>(far2_idx.size(), 1));
    k_fmm_logd<<<nc, 256, 0, s>>>(F.cr.p, F.far_ptr.p, F.far_idx.p, F.logd.p);
    STOKES_LAUNCH_CHECK();
    k_fmm_logd<<<nc, 256, 0, s>>>(F.cr.p, F.far2_ptr.p, F.far2_idx.p, F.logd2.p);
    STOKES_LAUNCH_CHECK();
    STOKES_HIP(hipStreamSynchronize(s));
    if (!F.binm.p) {
        // binom(k, n), k, n < 32 (exact in double: < 2^53)
        std::vector<double> bm(static_cast<size_t>(kP) * kP, 0.0);
        for (int k = 0; k < kP; ++k) {
            bm[static_cast<size_t>(k) * kP] = 1.0;
            for (int n = 1; n <= k; ++n)
                bm[static_cast<size_t>(k) * kP + n] =
                    bm[static_cast<size_t>(k - 1) * kP + n - 1] +
                    (n < k ? bm[static_cast<size_t>(k - 1) * kP + n] : 0.0);
        }
        up(F.binm, bm);
        STOKES_HIP(hipStreamSynchronize(s));
    }
    if (!F.bin.p) {
        // binom(n + k, k) for n, k < 32 (exact in long double: < 2^63), in
        // the matrix cores' A-operand layout
        long double pas[2 * kP - 1][2 * kP - 1] = {};
        for (int a = 0; a < 2 * kP - 1; ++a) {
            pas[a][0] = 1.0L;
            for (int b = 1; b <= a; ++b) pas[a][b] = pas[a - 1][b - 1] + (b < a ? pas[a - 1][b] : 0.0L);
        }
        std::vector<double> bin(2 * 8 * 64);
        for (int kt = 0; kt < 2; ++kt)
            for (int st = 0; st < 8; ++st)
                for (int l = 0; l < 64; ++l) {
                    const int k = 16 * kt + l % 16, nn = 4 * st + l / 16;
                    bin[static_cast<size_t>((kt * 8 + st) * 64 + l)] = static_cast<double>(pas[nn + k][k]);
                }
        up(F.bin, bin);
        STOKES_HIP(hipStreamSynchronize(s));
    }
    F.mp.ensure(static_cast<size_t>(nc + npar) * 2 * kP * 2);
    F.mlog.ensure(static_cast<size_t>(nc + npar) * 4);
    F.lc.ensure(static_cast<size_t>(nc) * 3 * kP * 2);
    F.smp.ensure(std::max(1, m + nwall) * static_cast<size_t>(kStressSeries) * kP * 2);
    F.slc.ensure(std::max(1, m) * static_cast<size_t>(kStressSeries) * kP * 2);
    F.nc = nc;
    F.npar = npar;
    F.nwall = nwall;
    // one work item per (target cluster, source tile): bodies, then the wall
    std::vector<int> work;
    int max_tiles = 1;
    for (int part = 0; part < 2; ++part) {
        const int c0 = part == 0 ? nwall : 0, c1 = part == 0 ? nc : nwall;
        for (int t = c0; t < c1; ++t) {
            const int nt = tile_ptr[static_cast<size_t>(t) + 1] - tile_ptr[static_cast<size_t>(t)];
            max_tiles = std::max(max_tiles, nt);
            for (int j = 0; j < nt; ++j) {
                work.push_back(t);
                work.push_back(j);
            }
        }
        (part == 0 ? F.work_body : F.work_wall) = static_cast<int>(work.size() / 2) -
                                                  (part == 0 ? 0 : F.work_body);
    }
    up(F.work, work);
    STOKES_HIP(hipStreamSynchronize(s));
    F.max_tiles = max_tiles;
    F.far_pairs = static_cast<long long>(far_idx.size());
    F.far2_pairs = static_cast<long long>(far2_idx.size());
    F.direct_pairs = direct - static_cast<long long>(g.npts);  // no self pairs
    F.version = g.version;
}

namespace {

inline int round32(int v) { return (v + 31) / 32 * 32; }

// Direct sums of work items [w0, w0 + count) (clusters of len <= TPT bt).
template <bool COMP>
void direct_launch(Context& c, int tpt, int bt, int w0, int count) {
    Geometry& g = c.g;
    TileFmm& F = c.fmm;
    if (count <= 0) return;
    const int* work = F.work.p + 2 * static_cast<size_t>(w0);
    double* pcu = COMP ? c.part_cu.p : nullptr;
    double* pcv = COMP ? c.part_cv.p : nullptr;
    if (tpt == 4)
        k_fmm_direct<4, COMP><<<count, bt, 0, c.stream>>>(
            g.x.p, g.y.p, g.mx.p, g.my.p, c.ex.p, c.ey.p, F.cl.p, work, F.tile_ptr.p, F.tiles.p,
            g.npts, c.part_u.p, c.part_v.p, pcu, pcv);
    else if (tpt == 1)
        k_fmm_direct<1, COMP><<<count, bt, 0, c.stream>>>(
            g.x.p, g.y.p, g.mx.p, g.my.p, c.ex.p, c.ey.p, F.cl.p, work, F.tile_ptr.p, F.tiles.p,
            g.npts, c.part_u.p, c.part_v.p, pcu, pcv);
    else
        k_fmm_direct<2, COMP><<<count, bt, 0, c.stream>>>(
            g.x.p, g.y.p, g.mx.p, g.my.p, c.ex.p, c.ey.p, F.cl.p, work, F.tile_ptr.p, F.tiles.p,
            g.npts, c.part_u.p, c.part_v.p, pcu, pcv);
    STOKES_LAUNCH_CHECK();
}

}  // namespace

// The far field of the current density (c.ex, c.ey) into part_u/part_v[0]
// (*wall_nsplit = *body_nsplit = 1): the direct sums one source tile per
// block beside the expansions, then the local expansions, which also sum
// each point's tiles.  The stokeslets and rotlets of each target cluster's
// far bodies (den's last 3 m entries; den == nullptr: none) enter through
// the same expansions; k_assemble adds the near bodies' (tilefmm_near_bodies).
// compensated: the direct sums and the final sum carry corrections
// (part_cu/part_cv[0]) for a true-residual evaluation.
void tilefmm_ready(Context& c, bool compensated) {
    TileFmm& F = c.fmm;
    tilefmm_prepare(c);
    const size_t np = static_cast<size_t>(c.g.npts);
    c.part_u.ensure(np * F.max_tiles);
    c.part_v.ensure(np * F.max_tiles);
    if (compensated) {
        c.part_cu.ensure(np * F.max_tiles);
        c.part_cv.ensure(np * F.max_tiles);
    }
    if (!F.side) {
        STOKES_HIP(hipStreamCreateWithFlags(&F.side, hipStreamNonBlocking));
        STOKES_HIP(hipEventCreateWithFlags(&F.fork, hipEventDisableTiming));
        STOKES_HIP(hipEventCreateWithFlags(&F.join, hipEventDisableTiming));
    }
}

void tilefmm_apply(Context& c, const double* den, int* wall_nsplit, int* body_nsplit,
                   bool compensated) {
    Geometry& g = c.g;
    TileFmm& F = c.fmm;
    tilefmm_ready(c, compensated);
    hipStream_t s = c.stream;
    const int np = g.npts, nc = F.nc, n = g.n;
    // multipoles and translations (matrix cores) on the side stream while
    // the direct sums (vector ALU) run on the context's stream
    // timers count the operator's applications, not true-residual ones
    const bool timed = c.time_kernels && !compensated;
    hipStream_t side = F.one_stream ? s : F.side;
    if (!F.one_stream) {
        STOKES_HIP(hipEventRecord(F.fork, s));
        STOKES_HIP(hipStreamWaitEvent(side, F.fork, 0));
    }
    if (timed) c.t_fmm_expand.start(side);
    const double* stk = den != nullptr ? den + 2 * static_cast<size_t>(g.no) +
                                             2 * static_cast<size_t>(g.n) * g.m
                                       : nullptr;
    if (compensated)
        k_fmm_p2m<true><<<nc, 256, 0, side>>>(g.x.p, g.y.p, g.mx.p, g.my.p, c.ex.p, c.ey.p,
                                                F.cr.p, F.cl.p, stk, F.nwall, F.mp.p, F.mlog.p);
    else
        k_fmm_p2m<false><<<nc, 256, 0, side>>>(g.x.p, g.y.p, g.mx.p, g.my.p, c.ex.p, c.ey.p,
                                                 F.cr.p, F.cl.p, stk, F.nwall, F.mp.p, F.mlog.p);
    STOKES_LAUNCH_CHECK();
    // an operator application translates the parents of far groups (their
    // multipoles formed from the children's); a true-residual evaluation
    // keeps the one-level lists, whose multipoles carry error-free sums
    const bool two_level = !compensated && F.npar > 0;
    if (two_level) {
        k_fmm_m2m<<<F.npar, 256, 0, side>>>(F.cr.p, F.par_ptr.p, F.par_idx.p, F.binm.p, nc,
                                             F.mp.p, F.mlog.p);
        STOKES_LAUNCH_CHECK();
    }
    k_fmm_m2l<2, true, true><<<dim3(nc, 2), 256, 0, side>>>(
        F.cr.p, F.mp.p, two_level ? F.far2_ptr.p : F.far_ptr.p,
        two_level ? F.far2_idx.p : F.far_idx.p, F.bin.p, F.mlog.p,
        two_level ? F.logd2.p : F.logd.p, F.lc.p);
    STOKES_LAUNCH_CHECK();
    if (timed)
        c.t_fmm_expand.stop(side, static_cast<double>(two_level ? F.far2_pairs : F.far_pairs));
    if (!F.one_stream) STOKES_HIP(hipEventRecord(F.join, side));
    if (timed) c.t_fmm_direct.start(s);
    // bodies: 2 targets per lane up to 256 points (blocks of at most 128
    // threads: twice the waves of 4 per lane, for the same shared-reciprocal
    // cost within 1%), 4 above; wall segments: one per lane, a full wave per
    // segment
    static_assert(kMaxBodyPts <= 4 * 128, "4 targets per lane in blocks of at most 128");
    auto direct = compensated ? direct_launch<true> : direct_launch<false>;
    if (n > 256)
        direct(c, 4, round32(cdiv_int(n, 4)), 0, F.work_body);
    else
        direct(c, 2, round32(cdiv_int(n, 2)), 0, F.work_body);
    direct(c, 1, kWallSeg, F.work_body, F.work_wall);
    if (timed) c.t_fmm_direct.stop(s, static_cast<double>(F.direct_pairs));
    if (!F.one_stream) STOKES_HIP(hipStreamWaitEvent(s, F.join, 0));
    // a block of 256 points spans at most 5 wall segments and, for bodies of
    // 64 points or more, 5 bodies: their locals fit the LDS stage
    static_assert(kBlk / kWallSeg + 1 + kBlk / 64 + 1 <= kL2pStage, "L2P stage too small");
    const bool staged = n >= 64;
#define STOKES_L2P(COMP, ST, PCU, PCV)                                                        \
    k_fmm_l2p<COMP, ST><<<cdiv_int(np, kBlk), kBlk, 0, s>>>(g.x.p, g.y.p, F.cr.p, F.lc.p,   \
                                                            F.tile_ptr.p, g.no, n, F.nwall, \
                                                            np, c.part_u.p, c.part_v.p, PCU, PCV)
    if (compensated) {
        if (staged)
            STOKES_L2P(true, true, c.part_cu.p, c.part_cv.p);
        else
            STOKES_L2P(true, false, c.part_cu.p, c.part_cv.p);
    } else {
        if (staged)
            STOKES_L2P(false, true, nullptr, nullptr);
        else
            STOKES_L2P(false, false, nullptr, nullptr);
    }
#undef STOKES_L2P
    STOKES_LAUNCH_CHECK();
    *wall_nsplit = 1;
    *body_nsplit = 1;
}

TileNear tilefmm_near_bodies(const Context& c) {
    const TileFmm& F = c.fmm;
    return TileNear{F.nb_ptr.p, F.nb_idx.p, F.nwall, kWallSeg};
}

// Strain rate at every body point from the far bodies of the shear stress
// (the bodies that are far by the clusters' rule), from the boundary values
// in c.bd: out[{0,1,2} nbp + q] = E11, E12, E22.
void tilefmm_stress(Context& c, double* out) {
    Geometry& g = c.g;
    TileFmm& F = c.fmm;
    tilefmm_prepare(c);
    hipStream_t s = c.stream;
    const int m = g.m, n = g.n, nbp = n * m;
    if (m == 0) return;
    // multipoles of the bodies and of the wall segments (the wall's only
    // where some body is away from the wall)
    const int nsrc = m + (F.stress_wall ? F.nwall : 0);
    k_fmm_p2m_bd<<<dim3(nsrc, kStressSeries), 256, 0, s>>>(g.x.p, g.y.p, g.dzr.p, g.dzi.p, c.bd.p,
                                                          static_cast<size_t>(g.npts), F.cr.p,
                                                          F.cl.p, F.nwall, m, F.smp.p);
    STOKES_LAUNCH_CHECK();
    k_fmm_m2l<kStressSeries, false, false><<<dim3(m, kStressSeries), 256, 0, s>>>(
        F.scr.p, F.smp.p, F.sfar_ptr.p, F.sfar_idx.p, F.bin.p, nullptr, nullptr, F.slc.p);
    STOKES_LAUNCH_CHECK();
    k_fmm_l2p_strain<<<cdiv_int(nbp, kBlk), kBlk, 0, s>>>(g.x.p, g.y.p, F.cr.p, F.slc.p, g.no, n,
                                                         F.nwall, nbp, out);
    STOKES_LAUNCH_CHECK();
    if (F.stress_wall) {
        k_wall_strain_direct<<<m, kBlk, 0, s>>>(g.x.p, g.y.p, g.dzr.p, g.dzi.p, c.bd.p,
                                                static_cast<size_t>(g.npts), g.no, n, nbp,
                                                F.swn_ptr.p, F.swn_idx.p, out);
        STOKES_LAUNCH_CHECK();
    }
}

}  // namespace stokes
