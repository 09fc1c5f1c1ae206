// Far field by body clusters: the trapezoid double layer of well-separated
// bodies through multipole and local expansions, the rest summed directly.
//
// This replaces the reference's FMM (stokesDLPnew -> bhfmm2dparttarg,
// ErosionPackage/fortran_code/fmm/bhfmm2dpart.f:159-203, tolerance 0.5e-12)
// in the role the reference gives it: the far field of matvec_DLP_fmmbary
// when ifmm = 1 (stokesSolver.f:636-681).  It is a one-level scheme over the
// bodies of a porous geometry, not a tree: every body is a cluster, the wall
// is summed directly.
//
// Complex form of the pair term.  With r = z - w_j (target minus source),
// m_j the weighted normal and eta_j the density as complex numbers,
//   (r.m)(r.eta) r / |r|^4 = (1/4) [ a_j / r + s_j / conj(r) + conj(a_j) r / conj(r)^2 ],
//   a_j = m_j eta_j,  s_j = 2 Re(conj(m_j) eta_j).
// So a cluster s contributes, through the Cauchy sums F(z) = sum a_j/(z-w_j),
// S(z) = sum s_j/(z-w_j) and E(z) = sum a_j conj(w_j - c_s)/(z-w_j),
//   u(z) = (1/4) [ F + conj(S) + conj(E' - conj(z - c_s) F') ].
// Summed over far clusters into local expansions about the target body's
// centre c_t, with G = sum_s conj(c_t - c_s) F_s,
//   u(z) = (1/4) [ F_L + conj(S_L) + conj(E_L' - conj(z - c_t) F_L' - G_L') ].
//
// Expansions: multipole about c_s scaled by the cluster radius R_s,
// f(z) = sum_n M_n R_s^n / (z - c_s)^(n+1), M_n = sum_j coef_j ((w_j-c_s)/R_s)^n;
// local about c_t scaled by R_t, f(z) = sum_k L_k ((z - c_t)/R_t)^k, with
// L_k = -(1/d) (R_t/d)^k sum_n binom(n+k, k) (-R_s/d)^n M_n, d = c_s - c_t.
// kP = 32 terms; a pair of bodies is far when R_s + R_t <= |d| / 3, so the
// truncation is below (1/3)^32 = 5e-16 of the cluster's term sizes -- the
// size of the rounding of the direct sum.  Other pairs (and the body itself)
// are summed directly, as are the wall as a source of every target and every
// source at the wall's targets.
//
// Kernels per operator application: k_farfield restricted to the wall's
// targets (source splits), k_fmm_p2m (one block per body), k_fmm_m2l (one
// block per target body), k_fmm_eval (target body x source-tile split: the
// direct part, and the local expansion on split 0).  Lists are built on the
// host once per geometry.
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
constexpr int kSlots = kBlk / kP;      // sources in flight per M2L block
constexpr int kMaxBodyPts = 512;

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

// Multipole coefficients of every body: grid m, block 128; thread q*32 + k
// (q < 3: F, S, E) forms M^q_k.  cr holds (cx, cy, R) per body.
__global__ __launch_bounds__(128) void k_fmm_p2m(const double* __restrict__ x,
                                                 const double* __restrict__ y,
                                                 const double* __restrict__ mx,
                                                 const double* __restrict__ my,
                                                 const double* __restrict__ ex,
                                                 const double* __restrict__ ey,
                                                 const double* __restrict__ cr, int no, int n,
                                                 double* __restrict__ mp) {
    __shared__ double zr[kMaxBodyPts], zi[kMaxBodyPts], ar[kMaxBodyPts], ai[kMaxBodyPts],
        sr[kMaxBodyPts], er[kMaxBodyPts], ei[kMaxBodyPts];
    const int b = blockIdx.x;
    const double cx = cr[3 * b], cy = cr[3 * b + 1], iR = 1.0 / cr[3 * b + 2];
    const int off = no + b * n;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int p = off + j;
        const double dx = x[p] - cx, dy = y[p] - cy;
        zr[j] = dx * iR;
        zi[j] = dy * iR;
        const double mxx = mx[p], myy = my[p], exx = ex[p], eyy = ey[p];
        const double are = mxx * exx - myy * eyy, aim = mxx * eyy + myy * exx;
        ar[j] = are;
        ai[j] = aim;
        sr[j] = 2.0 * (mxx * exx + myy * eyy);
        er[j] = are * dx + aim * dy;  // a conj(w - c)
        ei[j] = aim * dx - are * dy;
    }
    __syncthreads();
    const int q = threadIdx.x / kP, k = threadIdx.x % kP;
    if (q >= 3) return;
    double accr = 0.0, acci = 0.0;
    for (int j = 0; j < n; ++j) {
        const cplx zk = cpow_int(cmk(zr[j], zi[j]), k);
        const double c_r = q == 0 ? ar[j] : (q == 1 ? sr[j] : er[j]);
        const double c_i = q == 0 ? ai[j] : (q == 1 ? 0.0 : ei[j]);
        accr = __builtin_fma(c_r, zk.r, __builtin_fma(-c_i, zk.i, accr));
        acci = __builtin_fma(c_r, zk.i, __builtin_fma(c_i, zk.r, acci));
    }
    const size_t o = (static_cast<size_t>(b) * 3 + q) * kP + k;
    mp[2 * o] = accr;
    mp[2 * o + 1] = acci;
}

// Local expansions of every target body from its far bodies: grid m, block
// 256 = kSlots groups of kP lanes; group g translates far sources g, g +
// kSlots, ...; lane k forms coefficient k.  lc[t][q][k], q < 4: F, S, E, G.
__global__ __launch_bounds__(kBlk) void k_fmm_m2l(const double* __restrict__ cr,
                                                  const double* __restrict__ mp,
                                                  const int* __restrict__ far_ptr,
                                                  const int* __restrict__ far_idx,
                                                  double* __restrict__ lc) {
    __shared__ double binom[2 * kP - 1][kP];
    __shared__ double U[kSlots][kP][2];
    __shared__ double MM[kSlots][3][kP][2];
    __shared__ double RED[kSlots][4][kP][2];
    const int t = blockIdx.x, tid = threadIdx.x;
    for (int e = tid; e < (2 * kP - 1) * kP; e += kBlk) {
        const int i = e / kP, k = e % kP;
        double v = 0.0;
        if (k <= i) {
            v = 1.0;
            for (int r = 1; r <= k; ++r) v = v * static_cast<double>(i - k + r) / r;
        }
        binom[i][k] = v;
    }
    const int slot = tid / kP, k = tid % kP;
    const double ctx = cr[3 * t], cty = cr[3 * t + 1], Rt = cr[3 * t + 2];
    double L[4][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    const int f0 = far_ptr[t], nf = far_ptr[t + 1] - f0;
    __syncthreads();
    for (int i0 = 0; i0 < nf; i0 += kSlots) {  // uniform over the block
        const int i = i0 + slot;
        const bool act = i < nf;
        cplx v = cmk(0.0, 0.0), id = cmk(0.0, 0.0), md = cmk(0.0, 0.0);
        if (act) {
            const int s = far_idx[f0 + i];
            const cplx d = cmk(cr[3 * s] - ctx, cr[3 * s + 1] - cty);
            id = cinv(d);
            const cplx u = cscale(id, -cr[3 * s + 2]);  // -R_s / d
            v = cscale(id, Rt);                        // R_t / d
            md = cmk(-d.r, -d.i);                      // c_t - c_s
            const cplx un = cpow_int(u, k);
            U[slot][k][0] = un.r;
            U[slot][k][1] = un.i;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const size_t o = (static_cast<size_t>(s) * 3 + q) * kP + k;
                MM[slot][q][k][0] = mp[2 * o];
                MM[slot][q][k][1] = mp[2 * o + 1];
            }
        }
        __syncthreads();
        if (act) {
            cplx acc[3] = {cmk(0.0, 0.0), cmk(0.0, 0.0), cmk(0.0, 0.0)};
            // a few terms at a time: the whole unrolled translation spills
#pragma unroll 4
            for (int nn = 0; nn < kP; ++nn) {
                const double bb = binom[nn + k][k];
                const cplx tn = cmk(bb * U[slot][nn][0], bb * U[slot][nn][1]);
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    cfma(acc[q], tn, cmk(MM[slot][q][nn][0], MM[slot][q][nn][1]));
            }
            const cplx f = cmul(cpow_int(v, k), cmk(-id.r, -id.i));  // -(R_t/d)^k / d
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const cplx w = cmul(f, acc[q]);
                L[q][0] += w.r;
                L[q][1] += w.i;
            }
            const cplx g = cmul(cmk(md.r, -md.i), cmul(f, acc[0]));
            L[3][0] += g.r;
            L[3][1] += g.i;
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        RED[slot][q][k][0] = L[q][0];
        RED[slot][q][k][1] = L[q][1];
    }
    __syncthreads();
    if (tid < 4 * kP) {
        const int q = tid / kP, kk = tid % kP;
        double sr = 0.0, si = 0.0;
        for (int sl = 0; sl < kSlots; ++sl) {
            sr += RED[sl][q][kk][0];
            si += RED[sl][q][kk][1];
        }
        const size_t o = (static_cast<size_t>(t) * 4 + q) * kP + kk;
        lc[2 * o] = sr;
        lc[2 * o + 1] = si;
    }
}

// Body targets: grid (m, ysplit), block BT threads; lane l owns the body's
// points l and l + BT (two targets per lane, so that they share one
// reciprocal per source; n <= 2 BT).  Block (b, y) sums the source tiles y,
// y + ysplit, ... of body b's list (wall tiles and the tiles of its near
// bodies, itself included) directly; blocks y = 0 also add the local
// expansion.  Writes part_u/part_v[y][point].  tiles holds (first point,
// count) per tile.
template <int BT>
__global__ __launch_bounds__(BT) void k_fmm_eval(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ mx,
    const double* __restrict__ my, const double* __restrict__ ex, const double* __restrict__ ey,
    const double* __restrict__ cr, const double* __restrict__ lc,
    const int* __restrict__ tile_ptr, const int* __restrict__ tiles, int no, int n, int npts,
    int ysplit, double* __restrict__ part_u, double* __restrict__ part_v) {
    constexpr int TPT = 2;
    __shared__ double2 sxy[kTile], sab[kTile];
    __shared__ double sc[kTile];
    __shared__ double loc[4][kP][2];
    const int b = blockIdx.x, ys = blockIdx.y, tid = threadIdx.x;
    const int off = no + b * n;
    double xt[TPT], yt[TPT], u[TPT], v[TPT];
    int tp[TPT];
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        // unused slots sit where no source is (they share a product of rho^2)
        const int k = tid + q * BT;
        tp[q] = k < n ? off + k : -1;
        xt[q] = tp[q] >= 0 ? x[tp[q]] : kNoTarget;
        yt[q] = tp[q] >= 0 ? y[tp[q]] : kNoTarget;
        u[q] = v[q] = 0.0;
    }
    const int t0 = tile_ptr[b], nt = tile_ptr[b + 1] - t0;
    for (int i = ys; i < nt; i += ysplit) {  // uniform over the block
        const int ts = tiles[2 * (t0 + i)], tl = tiles[2 * (t0 + i) + 1];
        __syncthreads();
        for (int e = tid; e < kTile; e += BT) {
            if (e < tl) {
                const int j = ts + e;
                sxy[e] = make_double2(x[j], y[j]);
                src_moments(mx[j], my[j], ex[j], ey[j], &sab[e], &sc[e]);
            } else {
                sxy[e] = make_double2(1.0e30, 1.0e30);
                sab[e] = make_double2(0.0, 0.0);
                sc[e] = 0.0;
            }
        }
        __syncthreads();
        // a tile of the body itself holds the self pairs (zero distance, zero
        // moment product): the DIAG form keeps their shared product finite
        if (ts >= off && ts < off + n)
            ff_tile_shared<TPT, true>(sxy, sab, sc, xt, yt, u, v);
        else
            ff_tile_shared<TPT, false>(sxy, sab, sc, xt, yt, u, v);
    }
    if (ys == 0) {  // uniform over the block
        for (int e = tid; e < 4 * kP; e += BT) {
            const size_t o = static_cast<size_t>(b) * 4 * kP + e;
            loc[e / kP][e % kP][0] = lc[2 * o];
            loc[e / kP][e % kP][1] = lc[2 * o + 1];
        }
        __syncthreads();
        const double ctx = cr[3 * b], cty = cr[3 * b + 1], Rt = cr[3 * b + 2];
        const double iR = 1.0 / Rt;
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            if (tp[q] < 0) continue;
            const double dx = xt[q] - ctx, dy = yt[q] - cty;
            const cplx z = cmk(dx * iR, dy * iR);
            // Horner for the value and the derivative of each local series
            cplx val[4], der[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                cplx pv = cmk(loc[e][kP - 1][0], loc[e][kP - 1][1]), pd = cmk(0.0, 0.0);
                for (int kk = kP - 2; kk >= 0; --kk) {
                    pd = cadd(cmul(pd, z), pv);
                    pv = cadd(cmul(pv, z), cmk(loc[e][kk][0], loc[e][kk][1]));
                }
                val[e] = pv;
                der[e] = cscale(pd, iR);
            }
            // w = E' - conj(z - c_t) F' - G'
            const cplx w = csub(csub(der[2], cmul(cmk(dx, -dy), der[0])), der[3]);
            u[q] += 0.25 * (val[0].r + val[1].r + w.r);
            v[q] += 0.25 * (val[0].i - val[1].i - w.i);
        }
    }
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        if (tp[q] >= 0) {
            part_u[static_cast<size_t>(ys) * npts + tp[q]] = u[q];
            part_v[static_cast<size_t>(ys) * npts + tp[q]] = v[q];
        }
    }
}

}  // namespace

bool tilefmm_available(const Context& c) {
    const Geometry& g = c.g;
    return g.m >= 2 && g.n >= 4 && g.n <= kMaxBodyPts;
}

// Cluster data and lists of the current geometry (once per geometry).
void tilefmm_prepare(Context& c) {
    Geometry& g = c.g;
    TileFmm& F = c.fmm;
    if (!tilefmm_available(c)) throw Error("the body-cluster far field needs 2+ bodies of <= 512 points");
    if (F.version == g.version) return;
    const int n = g.n, m = g.m, no = g.no;
    std::vector<double> cr(3 * static_cast<size_t>(m));
    for (int b = 0; b < m; ++b) {
        const double cx = g.h_cx[static_cast<size_t>(b)], cy = g.h_cy[static_cast<size_t>(b)];
        double r2 = 0.0;
        for (int k = 0; k < n; ++k) {
            const size_t p = static_cast<size_t>(no + b * n + k);
            const double dx = g.h_x[p] - cx, dy = g.h_y[p] - cy;
            r2 = std::max(r2, dx * dx + dy * dy);
        }
        cr[3 * static_cast<size_t>(b)] = cx;
        cr[3 * static_cast<size_t>(b) + 1] = cy;
        cr[3 * static_cast<size_t>(b) + 2] = std::max(std::sqrt(r2), 1e-300);
    }
    std::vector<int> far_ptr(static_cast<size_t>(m) + 1, 0), far_idx, tile_ptr(static_cast<size_t>(m) + 1, 0),
        tiles;
    long long direct = 0;
    for (int t = 0; t < m; ++t) {
        for (int w0 = 0; w0 < no; w0 += kTile) {  // the wall, directly
            tiles.push_back(w0);
            tiles.push_back(std::min(kTile, no - w0));
        }
        int nsrc = no;
        for (int s = 0; s < m; ++s) {
            const double dx = cr[3 * static_cast<size_t>(s)] - cr[3 * static_cast<size_t>(t)];
            const double dy = cr[3 * static_cast<size_t>(s) + 1] - cr[3 * static_cast<size_t>(t) + 1];
            const double rs = cr[3 * static_cast<size_t>(s) + 2] + cr[3 * static_cast<size_t>(t) + 2];
            if (s != t && rs <= kSep * std::sqrt(dx * dx + dy * dy)) {
                far_idx.push_back(s);
            } else {
                for (int k0 = 0; k0 < n; k0 += kTile) {
                    tiles.push_back(no + s * n + k0);
                    tiles.push_back(std::min(kTile, n - k0));
                }
                nsrc += n;
            }
        }
        far_ptr[static_cast<size_t>(t) + 1] = static_cast<int>(far_idx.size());
        tile_ptr[static_cast<size_t>(t) + 1] = static_cast<int>(tiles.size() / 2);
        direct += static_cast<long long>(n) * nsrc;
    }
    hipStream_t s = c.stream;
    auto up = [&](auto& buf, const auto& vec) {
        buf.ensure(std::max<size_t>(vec.size(), 1));
        if (!vec.empty())
            STOKES_HIP(hipMemcpyAsync(buf.p, vec.data(), vec.size() * sizeof(vec[0]),
                                      hipMemcpyHostToDevice, s));
    };
    up(F.cr, cr);
    up(F.far_ptr, far_ptr);
    up(F.far_idx, far_idx);
    up(F.tile_ptr, tile_ptr);
    up(F.tiles, tiles);
    STOKES_HIP(hipStreamSynchronize(s));
    F.mp.ensure(static_cast<size_t>(m) * 3 * kP * 2);
    F.lc.ensure(static_cast<size_t>(m) * 4 * kP * 2);
    int maxtiles = 0;
    for (int t = 0; t < m; ++t)
        maxtiles = std::max(maxtiles, tile_ptr[static_cast<size_t>(t) + 1] - tile_ptr[static_cast<size_t>(t)]);
    // about 2048 blocks over the bodies and their tile lists
    F.ysplit = std::max(1, std::min(maxtiles, cdiv_int(2048, m)));
    F.far_pairs = static_cast<long long>(far_idx.size());
    // direct pairs: body targets (their lists) and wall targets (every point)
    F.direct_pairs = direct + static_cast<long long>(no) * (g.npts - 1);
    F.wall_nsplit = 0;
    F.version = g.version;
}

// The far field of the current density (c.ex, c.ey) into part_u/part_v:
// wall targets in *wall_nsplit source splits, body targets in *body_nsplit.
void tilefmm_apply(Context& c, int* wall_nsplit, int* body_nsplit) {
    Geometry& g = c.g;
    TileFmm& F = c.fmm;
    tilefmm_prepare(c);
    hipStream_t s = c.stream;
    const int np = g.npts, m = g.m, n = g.n, no = g.no;
    // wall targets against every source: k_farfield, 2 targets per lane with
    // the shared reciprocal, about 512 blocks
    const int tb = cdiv_int(no, 2 * kBlk);
    const int want = std::max(1, std::min(64, cdiv_int(512, tb)));
    const int split_len = cdiv_int(cdiv_int(np, want), kTile) * kTile;
    const int wsplit = cdiv_int(np, split_len);
    const int need = std::max(wsplit, F.ysplit);
    c.part_u.ensure(static_cast<size_t>(np) * need);
    c.part_v.ensure(static_cast<size_t>(np) * need);
    farfield_wall_targets(c, split_len, wsplit);
    k_fmm_p2m<<<m, 128, 0, s>>>(g.x.p, g.y.p, g.mx.p, g.my.p, c.ex.p, c.ey.p, F.cr.p, no, n,
                                F.mp.p);
    STOKES_LAUNCH_CHECK();
    k_fmm_m2l<<<m, kBlk, 0, s>>>(F.cr.p, F.mp.p, F.far_ptr.p, F.far_idx.p, F.lc.p);
    STOKES_LAUNCH_CHECK();
    const dim3 grid(m, F.ysplit);
    if (c.time_kernels) c.t_fmm_direct.start(s);
    // two targets per lane: 128-thread blocks up to 256 points per body
    if (n <= 256)
        k_fmm_eval<128><<<grid, 128, 0, s>>>(g.x.p, g.y.p, g.mx.p, g.my.p, c.ex.p, c.ey.p, F.cr.p,
                                             F.lc.p, F.tile_ptr.p, F.tiles.p, no, n, np, F.ysplit,
                                             c.part_u.p, c.part_v.p);
    else
        k_fmm_eval<kBlk><<<grid, kBlk, 0, s>>>(g.x.p, g.y.p, g.mx.p, g.my.p, c.ex.p, c.ey.p,
                                               F.cr.p, F.lc.p, F.tile_ptr.p, F.tiles.p, no, n, np,
                                               F.ysplit, c.part_u.p, c.part_v.p);
    STOKES_LAUNCH_CHECK();
    if (c.time_kernels)
        c.t_fmm_direct.stop(s, static_cast<double>(F.direct_pairs) -
                                   static_cast<double>(no) * static_cast<double>(np - 1));
    F.wall_nsplit = wsplit;
    *wall_nsplit = wsplit;
    *body_nsplit = F.ysplit;
}

}  // namespace stokes
