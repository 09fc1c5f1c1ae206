// The completed double-layer operator y = A x on the device.
//
// Restates matvec_DLP_fmmbary (ibary=1) and matvec_DLP (ibary=0),
// ErosionPackage/fortran_code/stokesSolver.f:570-1124 and 1519-1912:
//   far field     all-pairs trapezoid double layer, j != i, replacing the FMM
//                 (stokesSolver.f:636-681, pair form of :1563-1584)
//   near pairs    subtract the trapezoid, add barycentric StokesExteriorDLP /
//                 StokesInteriorDLP with nder=1 (:687-926, :5233-5528)
//   local terms   curvature diagonal and jump -eta/2 (:932-987)
//   completion    stokeslets + rotlets (:992-1049), constraint rows
//                 (:1055-1096), rank-one wall term (:1102-1118)
//
// Kernel map, one matvec:
//   k_gather       den -> (etax, etay) in point order (skipped when the
//                  preconditioner's apply has written them)
//   k_farfield     LDS-tiled all-pairs sum, target per lane, partial sums per
//                  source split                                (VALU bound)
//   k_fields       the five Cauchy densities per point
//   k_bdval        boundary values of the Cauchy integrals of the active
//                  source curves, node-sliced across waves      (VALU bound)
//   k_near_wave    barycentric value/derivative at near targets minus the
//                  trapezoid part, one wave per (target, source curve) entry
//   k_reduce_small rank-one integral and constraint rows (skipped when the
//                  preconditioner's apply has formed them)
//   k_assemble     partial sums + near corrections + local terms + stokeslets
#include <algorithm>
#include <cmath>
#include <functional>

#include "bary.hpp"
#include "common.hpp"
#include "farfield.hpp"

namespace stokes {

namespace {

constexpr int kBlk = 256;

inline int cdiv_int(int a, int b) { return (a + b - 1) / b; }

__device__ __forceinline__ int row_x(int i, int no, int n) {
    if (i < no) return i;
    int b = (i - no) / n, k = (i - no) - b * n;
    return 2 * no + 2 * b * n + k;
}
__device__ __forceinline__ int row_y(int i, int no, int n) {
    if (i < no) return no + i;
    int b = (i - no) / n, k = (i - no) - b * n;
    return 2 * no + 2 * b * n + n + k;
}

__global__ void k_gather(const double* __restrict__ den, int no, int n, int npts,
                         double* __restrict__ ex, double* __restrict__ ey) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    ex[i] = den[row_x(i, no, n)];
    ey[i] = den[row_y(i, no, n)];
}

// ---------------------------------------------------------------- far field
using namespace ff;

// grid: (ceil(ntgt / (256 TPT)), source splits); block 256 threads.  Lane l
// of block b owns targets (b TPT + q) 256 + l, q < TPT; the block stages
// 256-source tiles of its split in LDS (double2 SoA, wave-uniform reads).
// Targets are the points [0, ntgt) (every point, or the wall alone), sources
// all npts points.  SHR: off-diagonal tiles share one reciprocal over the
// lane's targets.
template <int TPT, bool SHR>
__global__ __launch_bounds__(kBlk) void k_farfield(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ mx,
    const double* __restrict__ my, const double* __restrict__ ex, const double* __restrict__ ey,
    int npts, int ntgt, int split_len, double* __restrict__ part_u, double* __restrict__ part_v) {
    __shared__ double2 sxy[kTile], sab[kTile];
    __shared__ double sc[kTile];
    const int tid = threadIdx.x;
    const int tb0 = blockIdx.x * TPT * kBlk;
    const int tb1 = min(tb0 + TPT * kBlk, ntgt);
    double xt[TPT], yt[TPT], u[TPT], v[TPT];
    int t[TPT];
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        // targets past ntgt sit where no source is (a source at the same
        // point would zero the product of rho^2 the lane's targets share)
        t[q] = tb0 + q * kBlk + tid;
        const bool valid = t[q] < ntgt;
        xt[q] = valid ? x[t[q]] : kNoTarget;
        yt[q] = valid ? y[t[q]] : kNoTarget;
        u[q] = v[q] = 0.0;
    }
    const int s0 = blockIdx.y * split_len;
    const int s1 = min(s0 + split_len, npts);
    for (int base = s0; base < s1; base += kTile) {
        const int j = base + tid;
        __syncthreads();
        if (j < s1) {
            sxy[tid] = make_double2(x[j], y[j]);
            src_moments(mx[j], my[j], ex[j], ey[j], &sab[tid], &sc[tid]);
        } else {  // padding source: zero weight far away
            sxy[tid] = make_double2(1.0e30, 1.0e30);
            sab[tid] = make_double2(0.0, 0.0);
            sc[tid] = 0.0;
        }
        __syncthreads();
        const bool diag = (base < tb1) && (base + kTile > tb0);
        if (diag && SHR) {
            if constexpr (TPT > 1) ff_tile_shared<TPT, true>(sxy, sab, sc, xt, yt, u, v);
        } else if (diag) {
            int self[TPT];
#pragma unroll
            for (int q = 0; q < TPT; ++q) self[q] = t[q] - base;
            ff_tile<TPT, true>(sxy, sab, sc, xt, yt, self, u, v);
        } else if (SHR) {
            if constexpr (TPT > 1) ff_tile_shared<TPT, false>(sxy, sab, sc, xt, yt, u, v);
        } else {
            ff_tile<TPT, false>(sxy, sab, sc, xt, yt, nullptr, u, v);
        }
    }
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        if (t[q] < ntgt) {
            part_u[static_cast<size_t>(blockIdx.y) * npts + t[q]] = u[q];
            part_v[static_cast<size_t>(blockIdx.y) * npts + t[q]] = v[q];
        }
    }
}

// ---------------------------------------------- compensated FP64 far field
template <int TPT, bool CHECK>
__device__ __forceinline__ void ff_tile_comp(const double2* __restrict__ sxy,
                                             const double2* __restrict__ sab,
                                             const double* __restrict__ sc, const double* xt,
                                             const double* yt, const int* self, double* u,
                                             double* cu, double* v, double* cv) {
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
            const double inv = rcp64(rho2);
            double c = w * (inv * inv);
            if (CHECK) c = (jj == self[q]) ? 0.0 : c;
            dd_add_prod(u[q], cu[q], c, rx);
            dd_add_prod(v[q], cv[q], c, ry);
        }
    }
}

// grid and target layout as k_farfield (no shared reciprocal); writes each
// split's sums to part_u/part_v and their corrections to part_cu/part_cv.
template <int TPT>
__global__ __launch_bounds__(kBlk) void k_farfield_comp(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ mx,
    const double* __restrict__ my, const double* __restrict__ ex, const double* __restrict__ ey,
    int npts, int split_len, double* __restrict__ part_u, double* __restrict__ part_v,
    double* __restrict__ part_cu, double* __restrict__ part_cv) {
    __shared__ double2 sxy[kTile], sab[kTile];
    __shared__ double sc[kTile];
    const int tid = threadIdx.x;
    const int tb0 = blockIdx.x * TPT * kBlk;
    const int tb1 = min(tb0 + TPT * kBlk, npts);
    double xt[TPT], yt[TPT], u[TPT], v[TPT], cu[TPT], cv[TPT];
    int t[TPT];
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        t[q] = tb0 + q * kBlk + tid;
        const bool valid = t[q] < npts;
        xt[q] = valid ? x[t[q]] : kNoTarget;
        yt[q] = valid ? y[t[q]] : kNoTarget;
        u[q] = v[q] = cu[q] = cv[q] = 0.0;
    }
    const int s0 = blockIdx.y * split_len;
    const int s1 = min(s0 + split_len, npts);
    for (int base = s0; base < s1; base += kTile) {
        const int j = base + tid;
        __syncthreads();
        if (j < s1) {
            sxy[tid] = make_double2(x[j], y[j]);
            src_moments(mx[j], my[j], ex[j], ey[j], &sab[tid], &sc[tid]);
        } else {
            sxy[tid] = make_double2(1.0e30, 1.0e30);
            sab[tid] = make_double2(0.0, 0.0);
            sc[tid] = 0.0;
        }
        __syncthreads();
        if ((base < tb1) && (base + kTile > tb0)) {
            int self[TPT];
#pragma unroll
            for (int q = 0; q < TPT; ++q) self[q] = t[q] - base;
            ff_tile_comp<TPT, true>(sxy, sab, sc, xt, yt, self, u, cu, v, cv);
        } else {
            ff_tile_comp<TPT, false>(sxy, sab, sc, xt, yt, nullptr, u, cu, v, cv);
        }
    }
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        if (t[q] < npts) {
            const size_t o = static_cast<size_t>(blockIdx.y) * npts + t[q];
            part_u[o] = u[q];
            part_v[o] = v[q];
            part_cu[o] = cu[q];
            part_cv[o] = cv[q];
        }
    }
}

// ------------------------------------------------------ FP32 far field
// The same pair term in FP32.  Sources are stored relative to the centre of
// their 256-point tile (FP32 offsets of a few body radii at most), each
// target's offset from that centre is formed in FP64 once per tile and
// rounded, so rx = x_t - x_s keeps its relative accuracy for pairs from
// different tiles; the tile that holds a block's own targets (a curve's
// neighbouring points) is summed in FP64.  Pair terms are summed in FP32
// over 64 sources and the partial sums are accumulated in FP64.
constexpr int kSub32 = 64;

template <int TPT, bool CHECK>
__device__ __forceinline__ void ff32_tile(const float2* __restrict__ sxy,
                                          const float2* __restrict__ sab,
                                          const float* __restrict__ sc, const float* dx,
                                          const float* dy, const int* self, double* u, double* v) {
    for (int sub = 0; sub < kTile; sub += kSub32) {
        float uf[TPT], vf[TPT];
#pragma unroll
        for (int q = 0; q < TPT; ++q) uf[q] = vf[q] = 0.0f;
#pragma unroll 4
        for (int jj = sub; jj < sub + kSub32; ++jj) {
            const float2 p = sxy[jj];
            const float2 ab = sab[jj];
            const float cc = sc[jj];
#pragma unroll
            for (int q = 0; q < TPT; ++q) {
                const float rx = dx[q] - p.x, ry = dy[q] - p.y;
                const float ry2 = ry * ry;
                const float rho2 = __builtin_fmaf(rx, rx, ry2);
                const float t = __builtin_fmaf(ab.x, rx, ab.y * ry);
                const float w = __builtin_fmaf(t, rx, cc * ry2);
                const float inv = __builtin_amdgcn_rcpf(rho2);
                float cf = w * (inv * inv);
                if (CHECK) cf = (jj == self[q]) ? 0.0f : cf;
                uf[q] = __builtin_fmaf(cf, rx, uf[q]);
                vf[q] = __builtin_fmaf(cf, ry, vf[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            u[q] += static_cast<double>(uf[q]);
            v[q] += static_cast<double>(vf[q]);
        }
    }
}

// Tile centres (point means) and FP32 source offsets; padding sources past
// npts sit far away (their moments are zero).  One block per tile.
__global__ __launch_bounds__(kBlk) void k_ff32_prep(const double* __restrict__ x,
                                                    const double* __restrict__ y, int npts,
                                                    double2* __restrict__ tctr,
                                                    float2* __restrict__ sxy) {
    __shared__ double shx[kBlk], shy[kBlk];
    const int j = blockIdx.x * kTile + threadIdx.x;
    const bool valid = j < npts;
    shx[threadIdx.x] = valid ? x[j] : 0.0;
    shy[threadIdx.x] = valid ? y[j] : 0.0;
    __syncthreads();
    for (int w = kBlk / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            shx[threadIdx.x] += shx[threadIdx.x + w];
            shy[threadIdx.x] += shy[threadIdx.x + w];
        }
        __syncthreads();
    }
    const int cnt = min(kTile, npts - static_cast<int>(blockIdx.x) * kTile);
    const double cx = shx[0] / cnt, cy = shy[0] / cnt;
    if (threadIdx.x == 0) tctr[blockIdx.x] = make_double2(cx, cy);
    sxy[j] = valid ? make_float2(static_cast<float>(x[j] - cx), static_cast<float>(y[j] - cy))
                   : make_float2(1.0e3f, 1.0e3f);
}

template <int TPT>
__global__ __launch_bounds__(kBlk) void k_farfield32(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ mx,
    const double* __restrict__ my, const float2* __restrict__ s32xy,
    const double2* __restrict__ tctr, const double* __restrict__ ex,
    const double* __restrict__ ey, int npts, int split_len, double* __restrict__ part_u,
    double* __restrict__ part_v) {
    __shared__ float2 sxy[kTile], sab[kTile];
    __shared__ float sc[kTile];
    __shared__ double2 dxy[kTile], dab[kTile];
    __shared__ double dc[kTile];
    const int tid = threadIdx.x;
    const int tb0 = blockIdx.x * TPT * kBlk;
    const int tb1 = min(tb0 + TPT * kBlk, npts);
    double xt[TPT], yt[TPT], u[TPT], v[TPT];
    int t[TPT];
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        t[q] = tb0 + q * kBlk + tid;
        const bool valid = t[q] < npts;
        xt[q] = valid ? x[t[q]] : 0.0;
        yt[q] = valid ? y[t[q]] : 0.0;
        u[q] = v[q] = 0.0;
    }
    const int s0 = blockIdx.y * split_len;
    const int s1 = min(s0 + split_len, npts);
    for (int base = s0; base < s1; base += kTile) {
        const int j = base + tid;
        // a tile holding the block's own targets (neighbouring points of the
        // same curve, where FP32 offsets would cancel) is summed in FP64
        const bool diag = (base < tb1) && (base + kTile > tb0);
        __syncthreads();
        if (diag) {
            if (j < s1) {
                dxy[tid] = make_double2(x[j], y[j]);
                src_moments(mx[j], my[j], ex[j], ey[j], &dab[tid], &dc[tid]);
            } else {
                dxy[tid] = make_double2(1.0e30, 1.0e30);
                dab[tid] = make_double2(0.0, 0.0);
                dc[tid] = 0.0;
            }
            __syncthreads();
            int self[TPT];
#pragma unroll
            for (int q = 0; q < TPT; ++q) self[q] = t[q] - base;
            ff_tile<TPT, true>(dxy, dab, dc, xt, yt, self, u, v);
            continue;
        }
        sxy[tid] = s32xy[j];
        if (j < s1) {  // moments formed in FP64, then rounded
            double2 m2;
            double m1;
            src_moments(mx[j], my[j], ex[j], ey[j], &m2, &m1);
            sab[tid] = make_float2(static_cast<float>(m2.x), static_cast<float>(m2.y));
            sc[tid] = static_cast<float>(m1);
        } else {
            sab[tid] = make_float2(0.0f, 0.0f);
            sc[tid] = 0.0f;
        }
        __syncthreads();
        const double2 ctr = tctr[base / kTile];
        float dx[TPT], dy[TPT];
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            dx[q] = static_cast<float>(xt[q] - ctr.x);
            dy[q] = static_cast<float>(yt[q] - ctr.y);
        }
        ff32_tile<TPT, false>(sxy, sab, sc, dx, dy, nullptr, u, v);
    }
#pragma unroll
    for (int q = 0; q < TPT; ++q) {
        if (t[q] < npts) {
            part_u[static_cast<size_t>(blockIdx.y) * npts + t[q]] = u[q];
            part_v[static_cast<size_t>(blockIdx.y) * npts + t[q]] = v[q];
        }
    }
}

// Far-field launch shape: targets per lane, the grid size it aims for, and
// (FP64 kernel) whether a lane's targets share one reciprocal per source;
// the FP32 kernel ignores `shared`.
struct FFShape {
    int tpt;
    int blocks;
    bool shared;
};
constexpr FFShape kFFShapes[] = {{1, 2048, false}, {1, 4096, false}, {2, 1024, false},
                                 {2, 2048, false}, {2, 4096, false}, {4, 1024, false},
                                 {4, 2048, false}, {2, 2048, true},  {4, 1024, true},
                                 {4, 2048, true},  {8, 1024, true},  {8, 2048, true}};

struct FFPlan {
    int tpt, tblocks, split_len, nsplit;
    bool shared;
};

// The FP32 kernel takes at most 4 targets per lane and no shared reciprocal.
FFPlan ff_plan(int np, FFShape sh, bool fp32) {
    FFPlan p{};
    p.tpt = fp32 ? std::min(sh.tpt, 4) : sh.tpt;
    p.shared = !fp32 && sh.shared && sh.tpt > 1;
    p.tblocks = cdiv_int(np, kBlk * p.tpt);
    const int want = std::max(1, cdiv_int(sh.blocks, p.tblocks));
    p.split_len = cdiv_int(cdiv_int(np, want), kTile) * kTile;
    p.nsplit = cdiv_int(np, p.split_len);
    return p;
}

void ff32_launch(const Context& c, const FFPlan& p, double* pu, double* pv) {
    const Geometry& g = c.g;
    const dim3 grid(p.tblocks, p.nsplit);
    const float2* sxy = reinterpret_cast<const float2*>(c.ff32_xy.p);
    const double2* tc = reinterpret_cast<const double2*>(c.ff32_ctr.p);
    switch (p.tpt) {
    case 1:
        k_farfield32<1><<<grid, kBlk, 0, c.stream>>>(g.x.p, g.y.p, g.mx.p, g.my.p, sxy, tc,
                                                     c.ex.p, c.ey.p,
                                                     g.npts, p.split_len, pu, pv);
        break;
    case 2:
        k_farfield32<2><<<grid, kBlk, 0, c.stream>>>(g.x.p, g.y.p, g.mx.p, g.my.p, sxy, tc,
                                                     c.ex.p, c.ey.p,
                                                     g.npts, p.split_len, pu, pv);
        break;
    default:
        k_farfield32<4><<<grid, kBlk, 0, c.stream>>>(g.x.p, g.y.p, g.mx.p, g.my.p, sxy, tc,
                                                     c.ex.p, c.ey.p,
                                                     g.npts, p.split_len, pu, pv);
        break;
    }
    STOKES_LAUNCH_CHECK();
}

// ntgt: targets [0, ntgt) (default: every point); the plan's target blocks
// are then ceil(ntgt / (256 tpt))
void ff_launch(const Context& c, const FFPlan& p, double* pu, double* pv, int ntgt = -1) {
    const Geometry& g = c.g;
    if (ntgt < 0) ntgt = g.npts;
    const dim3 grid(cdiv_int(ntgt, kBlk * p.tpt), p.nsplit);
#define STOKES_FF(T, S)                                                                      \
    k_farfield<T, S><<<grid, kBlk, 0, c.stream>>>(g.x.p, g.y.p, g.mx.p, g.my.p, c.ex.p, c.ey.p, \
                                                  g.npts, ntgt, p.split_len, pu, pv)
    if (p.tpt == 1)
        STOKES_FF(1, false);
    else if (p.tpt == 2 && p.shared)
        STOKES_FF(2, true);
    else if (p.tpt == 2)
        STOKES_FF(2, false);
    else if (p.tpt == 8)
        STOKES_FF(8, true);
    else if (p.shared)
        STOKES_FF(4, true);
    else
        STOKES_FF(4, false);
#undef STOKES_FF
    STOKES_LAUNCH_CHECK();
}

// ------------------------------------------------------- Cauchy densities
// s.eta, eta_x, eta_y, tau1 = eta n_x/(n_x+i n_y), tau2 = eta n_y/(n_x+i n_y)
// (stokesSolver.f:5268-5276, 4992-4997); stored as fld[(2f+{0,1})*npts + p].
__global__ void k_fields(const double* __restrict__ x, const double* __restrict__ y,
                         const double* __restrict__ px, const double* __restrict__ py,
                         const double* __restrict__ ex, const double* __restrict__ ey, int npts,
                         double* __restrict__ fld) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const size_t N = static_cast<size_t>(npts);
    const double dx = ex[i], dy = ey[i], nx = px[i], ny = py[i];
    const cplx eta = cmk(dx, dy);
    const cplx inor = cinv(cmk(nx, ny));
    const cplx t1 = cmul(cscale(eta, nx), inor);
    const cplx t2 = cmul(cscale(eta, ny), inor);
    fld[0 * N + i] = x[i] * dx + y[i] * dy;
    fld[1 * N + i] = 0.0;
    fld[2 * N + i] = dx;
    fld[3 * N + i] = 0.0;
    fld[4 * N + i] = dy;
    fld[5 * N + i] = 0.0;
    fld[6 * N + i] = t1.r;
    fld[7 * N + i] = t1.i;
    fld[8 * N + i] = t2.r;
    fld[9 * N + i] = t2.i;
}

// ------------------------------------------------------ boundary values
// compute_bdval_ex / compute_bdval_in (stokesSolver.f:5382-5427, 5103-5146)
// for the five fields of one curve at once:
//   B_j = sum_{k!=j} -(f_k - f_j) dz_k/(z_k - z_j) - dtheta f'_j
//   exterior: bd = i B / 2pi       interior: bd = -i B / 2pi - f_j
// grid: (ceil(len/64), curves, ksplit); block 64*NSL threads: lane -> node j,
// wave -> slice of the block's part of the k range; slices combined through
// LDS.  With part == nullptr (ksplit 1) the block finishes the boundary
// values; otherwise it writes its 20 partial sums per node to
// part[(z*20 + q)*len + j] and k_bdval_finish adds the ksplit parts.
template <int NSL>
__global__ __launch_bounds__(64 * NSL) void k_bdval(
    const int* __restrict__ curves, int no, int n, int npts, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ dzr, const double* __restrict__ dzi,
    const double* __restrict__ dno, const double* __restrict__ dn,
    const double* __restrict__ fld, double* __restrict__ bd, int ksplit,
    double* __restrict__ part) {
    __shared__ double red[NSL][64];
    const int c = curves[blockIdx.y];
    const int len = c == 0 ? no : n;
    const int off = c == 0 ? 0 : no + (c - 1) * n;
    const double* __restrict__ dt = c == 0 ? dno : dn;
    const int lane = threadIdx.x & 63, slice = threadIdx.x >> 6;
    const int jl = blockIdx.x * 64 + lane;
    if (static_cast<int>(blockIdx.x) * 64 >= len) return;  // uniform per block
    const bool valid = jl < len;
    const int j = valid ? jl : 0;
    const size_t N = static_cast<size_t>(npts);
    const cplx zj = cmk(x[off + j], y[off + j]);
    cplx fj[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) fj[f] = cmk(fld[(2 * f) * N + off + j], fld[(2 * f + 1) * N + off + j]);
    cplx B[5], D[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) B[f] = D[f] = cmk(0.0, 0.0);
    const int c0 = (blockIdx.z * len) / ksplit, c1 = ((blockIdx.z + 1) * len) / ksplit;
    const int k0 = c0 + (slice * (c1 - c0)) / NSL, k1 = c0 + ((slice + 1) * (c1 - c0)) / NSL;
    for (int k = k0; k < k1; ++k) {
        const int p = off + k;
        cplx fk[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) fk[f] = cmk(fld[(2 * f) * N + p], fld[(2 * f + 1) * N + p]);
        int idx = j - k;
        idx += idx < 0 ? len : 0;
        const double w = dt[idx];
        // derivative as sum_k d(j-k)(f_k - f_j): rows of d sum to zero
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            D[f].r = __builtin_fma(w, fk[f].r - fj[f].r, D[f].r);
            D[f].i = __builtin_fma(w, fk[f].i - fj[f].i, D[f].i);
        }
        if (k != j) {
            const cplx g = cmul(cmk(dzr[p], dzi[p]), cinv(csub(cmk(x[p], y[p]), zj)));
#pragma unroll
            for (int f = 0; f < 5; ++f) cfma(B[f], csub(fj[f], fk[f]), g);  // -(fk - fj) g
        }
    }
    // combine slices: 20 doubles per lane
    double vals[20];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        vals[4 * f + 0] = B[f].r;
        vals[4 * f + 1] = B[f].i;
        vals[4 * f + 2] = D[f].r;
        vals[4 * f + 3] = D[f].i;
    }
#pragma unroll
    for (int q = 0; q < 20; ++q) {
        red[slice][lane] = vals[q];
        __syncthreads();
        if (slice == 0) {
            double s = 0.0;
            for (int t = 0; t < NSL; ++t) s += red[t][lane];
            vals[q] = s;
        }
        __syncthreads();
    }
    if (slice != 0 || !valid) return;
    if (part) {
#pragma unroll
        for (int q = 0; q < 20; ++q)
            part[(static_cast<size_t>(blockIdx.z) * 20 + q) * len + j] = vals[q];
        return;
    }
    const double dth = kTwoPi / static_cast<double>(len);
    const bool ext = c != 0;
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        cplx Bf = cmk(vals[4 * f + 0] - dth * vals[4 * f + 2], vals[4 * f + 1] - dth * vals[4 * f + 3]);
        // i B / 2pi
        cplx iB = cmk(-Bf.i / kTwoPi, Bf.r / kTwoPi);
        cplx out = ext ? iB : cmk(-iB.r - fj[f].r, -iB.i - fj[f].i);
        bd[(2 * f) * N + off + j] = out.r;
        bd[(2 * f + 1) * N + off + j] = out.i;
    }
}

// Boundary values of the wall (curve 0, interior rule) from the ksplit
// partial sums of k_bdval: bd = -i B / 2pi - f_j, B = sum_parts(B) - dtheta
// sum_parts(D).  grid (ceil(no / 256), 5): one thread per (node, field).
__global__ void k_bdval_finish(int no, int npts, const double* __restrict__ fld,
                               const double* __restrict__ part, int ksplit,
                               double* __restrict__ bd) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (j >= no) return;
    const size_t N = static_cast<size_t>(npts);
    double vals[4] = {0.0, 0.0, 0.0, 0.0};
    for (int z = 0; z < ksplit; ++z) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            vals[q] += part[(static_cast<size_t>(z) * 20 + 4 * f + q) * no + j];
    }
    const double dth = kTwoPi / static_cast<double>(no);
    {
        const cplx fj = cmk(fld[(2 * f) * N + j], fld[(2 * f + 1) * N + j]);
        const cplx Bf = cmk(vals[0] - dth * vals[2], vals[1] - dth * vals[3]);
        const cplx iB = cmk(-Bf.i / kTwoPi, Bf.r / kTwoPi);
        bd[(2 * f) * N + j] = -iB.r - fj.r;
        bd[(2 * f + 1) * N + j] = -iB.i - fj.i;
    }
}

// ------------------------------------------------------- near evaluation
// One wave per near entry (target point, source curve): the lanes stride over
// the source's nodes and combine their sums with cross-lane butterflies, so
// the launch has one wave per entry however few entries share a source (a
// handful of targets near a 4096-point wall no longer sit on a few blocks).
// Pass 1: S0 = sum dz/(z-x), N_f = sum b_f dz/(z-x) (5 fields) and the
// trapezoid pair sum the correction replaces; pass 2: the derivative sums
// M_f = sum (b_f - u_f) dz/(z-x)^2 of fields 0..2 and, for an exterior
// source, D_a = sum dz/((z-a)(z-x)).
constexpr int kNearWaves = 4;

__device__ __forceinline__ double wave_sum64(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__global__ __launch_bounds__(64 * kNearWaves) void k_near_wave(
    const int* __restrict__ ent_src, const int* __restrict__ ent_tgt, int nentries, NodeView nv,
    const double* __restrict__ ax, const double* __restrict__ ay, int no, int n,
    double* __restrict__ corr_u, double* __restrict__ corr_v) {
    const int lane = threadIdx.x & 63;
    const int e = blockIdx.x * kNearWaves + (threadIdx.x >> 6);
    if (e >= nentries) return;  // uniform per wave
    const int src = ent_src[e], tp = ent_tgt[e];
    const int len = src == 0 ? no : n;
    const int off = src == 0 ? 0 : no + (src - 1) * n;
    const bool ext = src != 0;
    const double xt = nv.x[tp], yt = nv.y[tp];
    const cplx zt = cmk(xt, yt);
    const cplx a = cmk(ax[src], ay[src]);
    const size_t N = nv.N;

    double v1[14];
#pragma unroll
    for (int q = 0; q < 14; ++q) v1[q] = 0.0;
    for (int k = lane; k < len; k += 64) {
        const int p = off + k;
        const cplx d = cmk(nv.x[p] - xt, nv.y[p] - yt);
        const double rho2 = __builtin_fma(d.r, d.r, d.i * d.i);
        const double irho2 = rcp64(rho2);
        const cplx id = cmk(d.r * irho2, -d.i * irho2);
        const cplx w = cmul(cmk(nv.dzr[p], nv.dzi[p]), id);
        v1[0] += w.r;
        v1[1] += w.i;
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            const cplx b = cmk(nv.bd[(2 * f) * N + p], nv.bd[(2 * f + 1) * N + p]);
            v1[2 + 2 * f] = __builtin_fma(b.r, w.r, __builtin_fma(-b.i, w.i, v1[2 + 2 * f]));
            v1[3 + 2 * f] = __builtin_fma(b.r, w.i, __builtin_fma(b.i, w.r, v1[3 + 2 * f]));
        }
        // trapezoid pair with r = x_t - z_k = -d
        const double rx = -d.r, ry = -d.i;
        const double aa = __builtin_fma(rx, nv.mx[p], ry * nv.my[p]);
        const double bb = __builtin_fma(rx, nv.ex[p], ry * nv.ey[p]);
        const double cc = (aa * bb) * (irho2 * irho2);
        v1[12] = __builtin_fma(cc, rx, v1[12]);
        v1[13] = __builtin_fma(cc, ry, v1[13]);
    }
#pragma unroll
    for (int q = 0; q < 14; ++q) v1[q] = wave_sum64(v1[q]);
    const cplx S0 = cmk(v1[0], v1[1]);
    // exterior: num/(den - 2 pi i) (stokesSolver.f:5481); interior: num/den
    const cplx idv = cinv(ext ? cmk(S0.r, S0.i - kTwoPi) : S0);
    cplx u[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) u[f] = cmul(cmk(v1[2 + 2 * f], v1[3 + 2 * f]), idv);

    double v2[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v2[q] = 0.0;
    for (int k = lane; k < len; k += 64) {
        const int p = off + k;
        const cplx d = cmk(nv.x[p] - xt, nv.y[p] - yt);
        const double rho2 = __builtin_fma(d.r, d.r, d.i * d.i);
        const double irho2 = rcp64(rho2);
        const cplx id = cmk(d.r * irho2, -d.i * irho2);
        const cplx w = cmul(cmk(nv.dzr[p], nv.dzi[p]), id);
        if (ext && !(sqrt(rho2) > 1.0e-10)) {
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const cplx t = coincident_term(nv, off, len, k, f, zt, a, d);
                v2[2 * f] += t.r;
                v2[2 * f + 1] += t.i;
            }
        } else {
            const cplx w2 = cmul(w, id);
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const cplx t = csub(cmk(nv.bd[(2 * f) * N + p], nv.bd[(2 * f + 1) * N + p]), u[f]);
                v2[2 * f] = __builtin_fma(t.r, w2.r, __builtin_fma(-t.i, w2.i, v2[2 * f]));
                v2[2 * f + 1] = __builtin_fma(t.r, w2.i, __builtin_fma(t.i, w2.r, v2[2 * f + 1]));
            }
        }
        if (ext) {
            const cplx wa = cmul(w, cmk(nv.iar[p], nv.iai[p]));
            v2[6] += wa.r;
            v2[7] += wa.i;
        }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) v2[q] = wave_sum64(v2[q]);
    if (lane != 0) return;
    const cplx iden = cinv(ext ? cmul(csub(zt, a), cmk(v2[6], v2[7])) : S0);  // (x-a) D_a or S0
    cplx du[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) du[f] = cmul(cmk(v2[2 * f], v2[2 * f + 1]), iden);
    // u = Re v1' - x Re v2' - y Re v3' + Re v4 ; v = -Im v1' + x Im v2' + y Im v3' + Re v5
    const double ux = du[0].r - xt * du[1].r - yt * du[2].r + u[3].r;
    const double uy = -du[0].i + xt * du[1].i + yt * du[2].i + u[4].r;
    corr_u[e] = ux - v1[12];
    corr_v[e] = uy - v1[13];
}

// ------------------------------------------------- small reductions
// block 0: rank-one integral sigdotn (stokesSolver.f:1107-1111), unscaled,
// in scal[0] (k_assemble applies 2 pi / no);
// block 1+b: constraint rows of body b (stokesSolver.f:1055-1096).
__global__ __launch_bounds__(kBlk) void k_reduce_small(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ speed,
    const double* __restrict__ ex, const double* __restrict__ ey, const double* __restrict__ den,
    int no, int n, int m, double* __restrict__ scal, double* __restrict__ out) {
    __shared__ double s[3][kBlk / 64];
    const int b = blockIdx.x;
    double a0 = 0, a1 = 0, a2 = 0;
    if (b == 0) {
        for (int k = threadIdx.x; k < no; k += kBlk)
            a0 += (px[k] * ex[k] + py[k] * ey[k]) * speed[k];
    } else {
        const int off = no + (b - 1) * n;
        for (int k = threadIdx.x; k < n; k += kBlk) {
            const int p = off + k;
            a0 -= ex[p] * speed[p];
            a1 -= ey[p] * speed[p];
            a2 -= (ex[p] * y[p] - ey[p] * x[p]) * speed[p];
        }
    }
    // butterflies within each wave, then the waves' sums in order (one barrier)
    for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_xor(a0, off);
        a1 += __shfl_xor(a1, off);
        a2 += __shfl_xor(a2, off);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s[0][wave] = a0;
        s[1][wave] = a1;
        s[2][wave] = a2;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < kBlk / 64; ++w) {
        s[0][0] += s[0][w];
        s[1][0] += s[1][w];
        s[2][0] += s[2][w];
    }
    if (b == 0) {
        scal[0] = s[0][0];
    } else {
        const int base = 2 * no + 2 * n * m;
        const int r = base + 3 * (b - 1);
        const double sc = kTwoPi / static_cast<double>(n);
        out[r + 0] = s[0][0] / kTwoPi * sc + den[r + 0];
        out[r + 1] = s[1][0] / kTwoPi * sc + den[r + 1];
        out[r + 2] = s[2][0] / kTwoPi * sc + den[r + 2];
    }
}

// ----------------------------------------------------------- assembly
// lg[i m + b] = log |z_i - c_b|^2 for every boundary point i and body b: the
// geometry-only part of the stokeslet, computed once per geometry.
__global__ __launch_bounds__(kBlk) void k_stokes_logs(const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const double* __restrict__ cx,
                                                      const double* __restrict__ cy, int npts,
                                                      int m, double* __restrict__ lg) {
    const size_t e = static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x;
    if (e >= static_cast<size_t>(npts) * m) return;
    const int i = static_cast<int>(e / m), b = static_cast<int>(e - static_cast<size_t>(i) * m);
    const double rx = x[i] - cx[b], ry = y[i] - cy[b];
    lg[e] = log(rx * rx + ry * ry);
}

// SUB lanes per target: the stokeslet/rotlet sum over the bodies and the
// sum of the far field's per-split partials are split across neighbouring
// lanes and reduced with cross-lane moves, so a launch has SUB npts / 64
// waves to hide the latency of the logarithms and of the partial-sum loads.
// kAsmSub for the direct far field (every body's stokeslet, up to a hundred
// split partials per point); kAsmSubNear with the cluster far field, whose
// points sum a few near bodies and one partial: there a short grid of fatter
// lanes runs in one round of blocks instead of several rounds of
// mostly idle lanes waiting on dependent loads.
constexpr int kAsmSub = 16;
constexpr int kAsmSubNear = 4;
// COMP (a true-residual evaluation): the far field's split sums come with
// their corrections (part_cu, part_cv), and every sum of the assembly is kept
// as a value and a correction (TwoSum), rounded once at the end.
template <bool COMP, int SUB>
__global__ __launch_bounds__(kBlk) void k_assemble(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ cur,
    const double* __restrict__ speed, const double* __restrict__ ex,
    const double* __restrict__ ey, const double* __restrict__ cx, const double* __restrict__ cy,
    const double* __restrict__ den, const double* __restrict__ part_u,
    const double* __restrict__ part_v, const double* __restrict__ part_cu,
    const double* __restrict__ part_cv, int nsplit, int nsplit_body,
    const int* __restrict__ tgt_ptr,
    const int* __restrict__ tgt_ent, const double* __restrict__ corr_u,
    const double* __restrict__ corr_v, int has_near, const double* __restrict__ scal, int nsig,
    const double* __restrict__ lgtab, int no, int n, int m, int npts,
    const int* __restrict__ nb_ptr, const int* __restrict__ nb_idx, int nwall, int wall_seg,
    double* __restrict__ out) {
    static_assert(SUB >= 1 && SUB <= 64 && (SUB & (SUB - 1)) == 0, "SUB: a power of two up to 64");
    __shared__ double sc[5][kBlk];
    const int sub = threadIdx.x & (SUB - 1);
    const int i = blockIdx.x * (kBlk / SUB) + threadIdx.x / SUB;
    const bool valid = i < npts;
    const int ii = valid ? i : 0;
    const double xi = x[ii], yi = y[ii];
    // stokeslets and rotlets of every body, bodies q = sub mod SUB on this
    // lane; with the cluster far field (nb_ptr), only the bodies near the
    // point's cluster (the far ones came with the expansions), their
    // logarithms formed here
    double us = 0.0, vs = 0.0, lus = 0.0, lvs = 0.0;
    const int base = 2 * no + 2 * n * m;
    if (nb_ptr != nullptr) {
        const int cc = ii < no ? ii / wall_seg : nwall + (ii - no) / n;
        for (int t = nb_ptr[cc] + sub; t < nb_ptr[cc + 1]; t += SUB) {
            const int b = nb_idx[t];
            const double rx = xi - cx[b], ry = yi - cy[b];
            const double sto1 = den[base + 3 * b], sto2 = den[base + 3 * b + 1],
                         rot = den[base + 3 * b + 2];
            const double rho2 = rx * rx + ry * ry;
            const double irho2 = rcp64(rho2);
            const double lg = log(rho2);
            const double rdots = rx * sto1 + ry * sto2;
            const double tu = 0.5 / kTwoPi * (-0.5 * lg * sto1 + rdots * irho2 * rx) + rot * ry * irho2;
            const double tv = 0.5 / kTwoPi * (-0.5 * lg * sto2 + rdots * irho2 * ry) - rot * rx * irho2;
            if (COMP) {
                dd_add(us, lus, tu);
                dd_add(vs, lvs, tv);
            } else {
                us += tu;
                vs += tv;
            }
        }
    }
    for (int b0 = 0; nb_ptr == nullptr && b0 < m; b0 += kBlk) {  // block-uniform
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
        for (int q = sub; q < cnt; q += SUB) {
            const double rx = xi - sc[0][q], ry = yi - sc[1][q];
            const double sto1 = sc[2][q], sto2 = sc[3][q], rot = sc[4][q];
            const double rho2 = rx * rx + ry * ry;
            const double irho2 = rcp64(rho2);
            const double lg = lgtab[static_cast<size_t>(ii) * m + b0 + q];
            const double rdots = rx * sto1 + ry * sto2;
            const double tu = 0.5 / kTwoPi * (-0.5 * lg * sto1 + rdots * irho2 * rx) + rot * ry * irho2;
            const double tv = 0.5 / kTwoPi * (-0.5 * lg * sto2 + rdots * irho2 * ry) - rot * rx * irho2;
            if (COMP) {
                dd_add(us, lus, tu);
                dd_add(vs, lvs, tv);
            } else {
                us += tu;
                vs += tv;
            }
        }
    }
    // far-field partial sums of the source splits s = sub mod SUB (wall
    // and body points may have different split counts)
    if (valid) {
        const int ns = i < no ? nsplit : nsplit_body;
        for (int s = sub; s < ns; s += SUB) {
            const size_t o = static_cast<size_t>(s) * npts + i;
            if (COMP) {
                dd_add(us, lus, part_u[o]);
                dd_add(vs, lvs, part_v[o]);
                lus += part_cu[o];
                lvs += part_cv[o];
            } else {
                us += part_u[o];
                vs += part_v[o];
            }
        }
    }
#pragma unroll
    for (int off = 1; off < SUB; off <<= 1) {
        if (COMP) {
            // (hi, lo) pairs: TwoSum of the values, corrections added
            double s, e;
            two_sum(us, __shfl_xor(us, off), s, e);
            lus = lus + __shfl_xor(lus, off) + e;
            us = s;
            two_sum(vs, __shfl_xor(vs, off), s, e);
            lvs = lvs + __shfl_xor(lvs, off) + e;
            vs = s;
        } else {
            us += __shfl_xor(us, off);
            vs += __shfl_xor(vs, off);
        }
    }
    if (!valid || sub != 0) return;
    double u = 0.0, v = 0.0, lu = 0.0, lv = 0.0;
    // near-pair replacement
    if (has_near) {
        for (int e = tgt_ptr[i]; e < tgt_ptr[i + 1]; ++e) {
            if (COMP) {
                dd_add(u, lu, corr_u[tgt_ent[e]]);
                dd_add(v, lv, corr_v[tgt_ent[e]]);
            } else {
                u += corr_u[tgt_ent[e]];
                v += corr_v[tgt_ent[e]];
            }
        }
    }
    // curvature diagonal and jump
    {
        const double nsrc = i < no ? static_cast<double>(no) : static_cast<double>(n);
        const double dx = ex[i], dy = ey[i];
        const double tdotden = py[i] * dx - px[i] * dy;
        if (COMP) {
            dd_add(u, lu, -(cur[i] * tdotden * py[i] * speed[i] / nsrc));
            dd_add(v, lv, cur[i] * tdotden * px[i] * speed[i] / nsrc);
            dd_add(u, lu, -0.5 * dx);
            dd_add(v, lv, -0.5 * dy);
        } else {
            u = u - cur[i] * tdotden * py[i] * speed[i] / nsrc;
            v = v + cur[i] * tdotden * px[i] * speed[i] / nsrc;
            u = u - 0.5 * dx;
            v = v - 0.5 * dy;
        }
    }
    if (COMP) {
        dd_add(u, lu, us);
        dd_add(v, lv, vs);
        lu += lus;
        lv += lvs;
    } else {
        u += us;
        v += vs;
    }
    if (i < no) {
        // the rank-one integral from its nsig unscaled parts, in order,
        // loaded 16 at a time (indices clamped, values past nsig read as 0)
        double sig = 0.0;
        for (int j0 = 0; j0 < nsig; j0 += 16) {
            double tv[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const double v = scal[min(j0 + e, nsig - 1)];
                tv[e] = j0 + e < nsig ? v : 0.0;
            }
#pragma unroll
            for (int e = 0; e < 16; ++e) sig += tv[e];
        }
        sig = sig * kTwoPi / static_cast<double>(no);
        if (COMP) {
            dd_add(u, lu, sig * px[i]);
            dd_add(v, lv, sig * py[i]);
        } else {
            u += sig * px[i];
            v += sig * py[i];
        }
    }
    if (COMP) {
        u += lu;
        v += lv;
    }
    out[row_x(i, no, n)] = u;
    out[row_y(i, no, n)] = v;
}

__global__ void k_rhs(const double* __restrict__ y, int no, int ibc, int nunk,
                      double* __restrict__ rhs) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nunk) return;
    double v = 0.0;
    if (i < no) v = ibc == 0 ? 1.0 * (1.0 + y[i]) * (1.0 - y[i]) : (ibc == 1 ? 1.0 : 0.0);
    rhs[i] = v;
}

}  // namespace

// bd_condition (stokesSolver.f:208-314): background flow on the wall rows.
void bd_condition(Context& c, int ibc, double* rhs_dev) {
    const Geometry& g = c.g;
    k_rhs<<<cdiv_int(g.nunk, kBlk), kBlk, 0, c.stream>>>(g.y.p, g.no, ibc, g.nunk, rhs_dev);
    STOKES_LAUNCH_CHECK();
}

void gather_density(Context& c, const double* den_dev) {
    const Geometry& g = c.g;
    c.pc_token_x = c.pc_token_y = nullptr;
    c.ex.ensure(static_cast<size_t>(g.npts));
    c.ey.ensure(static_cast<size_t>(g.npts));
    k_gather<<<cdiv_int(g.npts, kBlk), kBlk, 0, c.stream>>>(den_dev, g.no, g.n, g.npts, c.ex.p,
                                                            c.ey.p);
    STOKES_LAUNCH_CHECK();
}

// Five Cauchy densities of the gathered density and their boundary values on
// the listed curves; the wall, when listed, must come first and gets 16
// k-slices per block, bodies get 4.
void boundary_values(Context& c, const int* curves_dev, int ncurves, int wall_first) {
    const Geometry& g = c.g;
    const int np = g.npts;
    const size_t N = static_cast<size_t>(np);
    hipStream_t s = c.stream;
    c.fld.ensure(10 * N);
    c.bd.ensure(10 * N);
    k_fields<<<cdiv_int(np, kBlk), kBlk, 0, s>>>(g.x.p, g.y.p, g.px.p, g.py.p, c.ex.p, c.ey.p, np,
                                                 c.fld.p);
    STOKES_LAUNCH_CHECK();
    if (wall_first) {
        // the wall alone is one curve of up to a few thousand nodes: split its
        // k range over blocks so the launch has ~256 blocks, then add the parts
        const int jb = cdiv_int(g.no, 64);
        const int ks = std::max(1, std::min(32, cdiv_int(256, jb)));
        c.bdpart.ensure(static_cast<size_t>(ks) * 20 * static_cast<size_t>(g.no));
        k_bdval<4><<<dim3(jb, 1, ks), 256, 0, s>>>(curves_dev, g.no, g.n, np, g.x.p, g.y.p,
                                                   g.dzr.p, g.dzi.p, g.dtab_no.p, g.dtab_n.p,
                                                   c.fld.p, c.bd.p, ks, c.bdpart.p);
        STOKES_LAUNCH_CHECK();
        k_bdval_finish<<<dim3(cdiv_int(g.no, kBlk), 5), kBlk, 0, s>>>(g.no, np, c.fld.p,
                                                                      c.bdpart.p, ks, c.bd.p);
        STOKES_LAUNCH_CHECK();
    }
    if (ncurves > wall_first) {
        k_bdval<4><<<dim3(cdiv_int(g.n, 64), ncurves - wall_first, 1), 256, 0, s>>>(
            curves_dev + wall_first, g.no, g.n, np, g.x.p, g.y.p, g.dzr.p, g.dzi.p, g.dtab_no.p,
            g.dtab_n.p, c.fld.p, c.bd.p, 1, nullptr);
        STOKES_LAUNCH_CHECK();
    }
}

// FP32 far-field source tiles of the current geometry (once per geometry).
// Every FP32 far-field launch reads them, the autotuner's included: they must
// be formed for the current geometry before the first such launch.
static void ff32_sources(Context& c) {
    Geometry& g = c.g;
    if (c.ff32_version == g.version) return;
    const int np = g.npts;
    const int ntiles = cdiv_int(np, kTile);
    c.ff32_ctr.ensure(2 * static_cast<size_t>(ntiles));
    c.ff32_xy.ensure(2 * static_cast<size_t>(ntiles) * kTile);
    k_ff32_prep<<<ntiles, kBlk, 0, c.stream>>>(g.x.p, g.y.p, np,
                                               reinterpret_cast<double2*>(c.ff32_ctr.p),
                                               reinterpret_cast<float2*>(c.ff32_xy.p));
    STOKES_LAUNCH_CHECK();
    c.ff32_version = g.version;
}

// Autotuner: the first operator application at a given number of boundary
// points times every far-field shape (one warm-up launch, then three timed
// launches each, HIP events on the library stream) and keeps the fastest.
// Shapes include the shared-reciprocal FP64 tiles, so they are used only
// where they measure faster.
constexpr int kFFDirectShapes = static_cast<int>(sizeof(kFFShapes) / sizeof(kFFShapes[0]));
constexpr int kFmmIndex = kFFDirectShapes;  // the cluster far field

// Tuning results are kept per size class -- points per body, wall points,
// boundary points in buckets of 2^(1/16) (about 4%) -- and, for FP64,
// whether the cluster far field may be chosen: an erosion run loses a body
// every few steps, and every exact size would otherwise time every shape
// again.  Key: ((n 2^20 + no) 2^12 + bucket) 2 + cluster flag, so that the
// neighbouring size classes of one body and wall size are key +- 2 d.
constexpr int kTuneReach = 3;

static long long tune_key(const Context& c, bool fp32) {
    const bool fmm = !fp32 && c.ff_fmm && tilefmm_available(c);
    const long long bucket = static_cast<long long>(std::floor(16.0 * std::log2(static_cast<double>(c.g.npts))));
    const long long geo = (static_cast<long long>(c.g.n) << 20) | static_cast<long long>(c.g.no);
    return (((geo << 12) | bucket) << 1) | (fmm ? 1 : 0);
}

static int tune_farfield(Context& c, bool fp32) {
    if (c.ff_force >= 0) {
        if (c.ff_force == kFmmIndex && !tilefmm_available(c))
            throw Error("the cluster far field needs bodies of 4 to 512 points");
        return c.ff_force;
    }
    const int np = c.g.npts;
    const long long key = tune_key(c, fp32);
    std::map<long long, int>& choice = fp32 ? c.ff32_choice : c.ff_choice;
    auto it = choice.find(key);
    if (it != choice.end()) return it->second;
    // a size class within kTuneReach classes (about 13%) of a tuned one
    // keeps that one's choice: the ranking of the shapes does not turn over
    // at that distance, and timing every shape again costs a third of a
    // second at 10^5 points
    std::map<long long, float>& tuned = fp32 ? c.ff32_tuned_ms : c.ff_tuned_ms;
    for (int d = 1; d <= kTuneReach; ++d)
        for (int sgn : {-1, 1}) {
            auto nb = choice.find(key + 2 * sgn * d);
            if (nb != choice.end()) {
                choice[key] = nb->second;
                tuned[key] = tuned[key + 2 * sgn * d];
                return nb->second;
            }
        }
    const bool fmm = key & 1;
    if (fmm) tilefmm_prepare(c);
    if (fp32) ff32_sources(c);
    // the tuning launches are not operator applications: the kernel timers
    // of the cluster far field's parts stay off while they run
    struct TimingOff {
        Context& c;
        const bool was;
        explicit TimingOff(Context& cc) : c(cc), was(cc.time_kernels) { c.time_kernels = false; }
        ~TimingOff() { c.time_kernels = was; }
    } timing_off(c);
    int best = 0;
    float best_ms = 1e30f;
    hipEvent_t e0, e1;
    STOKES_HIP(hipEventCreate(&e0));
    STOKES_HIP(hipEventCreate(&e1));
    for (int i = 0; i < kFFDirectShapes + (fmm ? 1 : 0); ++i) {
        std::function<void()> launch;
        if (i == kFmmIndex) {
            launch = [&]() {
                int a = 0, b = 0;
                tilefmm_apply(c, nullptr, &a, &b);
            };
        } else {
            // the FP32 kernel has no shared form (and no 8-target form)
            if (fp32 && kFFShapes[i].shared) continue;
            const FFPlan p = ff_plan(np, kFFShapes[i], fp32);
            c.part_u.ensure(static_cast<size_t>(np) * static_cast<size_t>(p.nsplit));
            c.part_v.ensure(static_cast<size_t>(np) * static_cast<size_t>(p.nsplit));
            launch = [&c, p, fp32]() {
                if (fp32)
                    ff32_launch(c, p, c.part_u.p, c.part_v.p);
                else
                    ff_launch(c, p, c.part_u.p, c.part_v.p);
            };
        }
        launch();
        STOKES_HIP(hipEventRecord(e0, c.stream));
        for (int r = 0; r < 3; ++r) launch();
        STOKES_HIP(hipEventRecord(e1, c.stream));
        STOKES_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        STOKES_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
            best_ms = ms;
            best = i;
        }
    }
    STOKES_HIP(hipEventDestroy(e0));
    STOKES_HIP(hipEventDestroy(e1));
    choice[key] = best;
    (fp32 ? c.ff32_tuned_ms : c.ff_tuned_ms)[key] = best_ms / 3.0f;
    return best;
}

bool farfield_fp32_pays(const Context& c) {
    // a direct shape fixed by the caller: mixed precision uses the FP32 far
    // field wherever its residual switch allows, whatever the timings
    if (c.ff_force >= 0 && c.ff_force != kFmmIndex) return true;
    auto a = c.ff32_tuned_ms.find(tune_key(c, true));
    auto b = c.ff_tuned_ms.find(tune_key(c, false));
    if (a == c.ff32_tuned_ms.end() || b == c.ff_tuned_ms.end()) return true;
    return a->second < b->second;
}

// Stokeslet logarithms of the current geometry (once per geometry).
static void stokes_logs(Context& c) {
    Geometry& g = c.g;
    if (g.m == 0 || c.stk_log_version == g.version) return;
    const size_t tot = static_cast<size_t>(g.npts) * static_cast<size_t>(g.m);
    c.stk_log.ensure(tot);
    k_stokes_logs<<<static_cast<int>((tot + kBlk - 1) / kBlk), kBlk, 0, c.stream>>>(
        g.x.p, g.y.p, g.cx.p, g.cy.p, g.npts, g.m, c.stk_log.p);
    STOKES_LAUNCH_CHECK();
    c.stk_log_version = g.version;
}

void prepare_operator(Context& c, const double* x_dev, bool fp32) {
    c.scal.ensure(std::max<size_t>(8, static_cast<size_t>(precond_wall_segments(c.g))));
    // the tuning launches read the density in point order: gather (and on
    // the first call allocate) it before timing any far-field shape
    gather_density(c, x_dev);
    if (fp32) ff32_sources(c);
    const int choice = tune_farfield(c, fp32);
    if (choice == kFmmIndex) {
        // lists and part buffers of the cluster far field
        tilefmm_ready(c);
    } else {
        const FFPlan plan = ff_plan(c.g.npts, kFFShapes[choice], fp32);
        c.part_u.ensure(static_cast<size_t>(c.g.npts) * static_cast<size_t>(plan.nsplit));
        c.part_v.ensure(static_cast<size_t>(c.g.npts) * static_cast<size_t>(plan.nsplit));
        // the direct far field's assembly reads every body's stokeslet
        // logarithms from a table (the cluster far field needs none)
        stokes_logs(c);
    }
}

int farfield_shape_count() { return kFFDirectShapes + 1; }

int farfield_choice(const Context& c, bool fp32) {
    if (c.ff_force >= 0) return c.ff_force;
    const std::map<long long, int>& choice = fp32 ? c.ff32_choice : c.ff_choice;
    auto it = choice.find(tune_key(c, fp32));
    return it == choice.end() ? -1 : it->second;
}

void apply_operator(Context& c, const double* x_dev, double* y_dev, int ibary, bool from_precond,
                    bool compensated) {
    Geometry& g = c.g;
    NearList& L = c.near;
    hipStream_t s = c.stream;
    const int np = g.npts;
    const size_t N = static_cast<size_t>(np);
    c.scal.ensure(std::max<size_t>(8, static_cast<size_t>(precond_wall_segments(g))));

    // the preconditioner's work is reused only when it was the last to write
    // c.ex/c.ey and its output is this x with these y rows (the token)
    from_precond = from_precond && c.pc_token_x == x_dev && c.pc_token_y == y_dev;
    c.pc_token_x = nullptr;
    c.pc_token_y = nullptr;
    if (!from_precond) gather_density(c, x_dev);
    bool fp32 = c.ff_fp32_now && !compensated;
    if (fp32) ff32_sources(c);
    // a compensated evaluation runs the cluster far field with compensated
    // direct sums where the FP64 far field in use is the cluster one, else
    // its own direct far field: 4 targets per lane, per-pair reciprocals,
    // about 2048 blocks
    int choice = tune_farfield(c, fp32);
    if (compensated && choice != kFmmIndex) choice = -1;
    if (choice == kFmmIndex) fp32 = false;  // the cluster far field is FP64
    int nsplit = 0, nsplit_body = 0;
    if (choice == kFmmIndex) {
        const bool timed = c.time_kernels && !compensated;
        if (timed) c.t_fmm.start(s);
        tilefmm_apply(c, x_dev, &nsplit, &nsplit_body, compensated);
        if (timed) c.t_fmm.stop(s, static_cast<double>(np) * static_cast<double>(np - 1));
        c.ff_splits = nsplit;
        c.ff_tpt = -1;
    }
    const FFPlan plan = choice == kFmmIndex ? FFPlan{}
                        : compensated ? ff_plan(np, FFShape{4, 2048, false}, false)
                                      : ff_plan(np, kFFShapes[choice], fp32);
    if (choice != kFmmIndex) {
        nsplit = nsplit_body = plan.nsplit;
        c.part_u.ensure(N * static_cast<size_t>(nsplit));
        c.part_v.ensure(N * static_cast<size_t>(nsplit));
    }
    if (choice == kFmmIndex) {
        // done above
    } else if (compensated) {
        c.part_cu.ensure(N * static_cast<size_t>(nsplit));
        c.part_cv.ensure(N * static_cast<size_t>(nsplit));
        k_farfield_comp<4><<<dim3(plan.tblocks, plan.nsplit), kBlk, 0, s>>>(
            g.x.p, g.y.p, g.mx.p, g.my.p, c.ex.p, c.ey.p, np, plan.split_len, c.part_u.p,
            c.part_v.p, c.part_cu.p, c.part_cv.p);
        STOKES_LAUNCH_CHECK();
    } else {
        c.ff_splits = nsplit;
        c.ff_tpt = plan.tpt;
        KernelTimer& tff = fp32 ? c.t_farfield32 : c.t_farfield;
        if (c.time_kernels) tff.start(s);
        if (fp32)
            ff32_launch(c, plan, c.part_u.p, c.part_v.p);
        else
            ff_launch(c, plan, c.part_u.p, c.part_v.p);
        if (c.time_kernels) tff.stop(s, static_cast<double>(np) * static_cast<double>(np - 1));
    }
    const bool has_near = ibary == 1 && L.nentries > 0;
    if (has_near) {
        boundary_values(c, L.active.p, L.nactive, L.wall_active);
        NodeView nv{g.x.p, g.y.p, g.dzr.p, g.dzi.p, c.bd.p, g.iar.p, g.iai.p, g.mx.p, g.my.p,
                    c.ex.p, c.ey.p, N};
        // the timer covers the near-evaluation (gather) pass only
        if (c.time_kernels) c.t_near.start(s);
        k_near_wave<<<cdiv_int(L.nentries, kNearWaves), 64 * kNearWaves, 0, s>>>(
            L.ent_src.p, L.ent_tgt.p, L.nentries, nv, g.ax.p, g.ay.p, g.no, g.n, L.corr_u.p,
            L.corr_v.p);
        STOKES_LAUNCH_CHECK();
    }
    if (has_near && c.time_kernels) c.t_near.stop(s, L.gather_bytes);
    // after the preconditioner's apply, the body constraint rows of y and the
    // wall segments' parts of the rank-one integral are already formed
    if (!from_precond) {
        k_reduce_small<<<g.m + 1, kBlk, 0, s>>>(g.x.p, g.y.p, g.px.p, g.py.p, g.speed.p, c.ex.p,
                                                c.ey.p, x_dev, g.no, g.n, g.m, c.scal.p, y_dev);
        STOKES_LAUNCH_CHECK();
    }
    const int nsig = from_precond ? precond_wall_segments(g) : 1;
    // with the cluster far field the far bodies' stokeslets came with it, and
    // the assembly sums the near bodies' with logarithms of its own
    const TileNear tn = choice == kFmmIndex ? tilefmm_near_bodies(c) : TileNear{nullptr, nullptr, 0, 1};
    if (choice != kFmmIndex) stokes_logs(c);
#define STOKES_ASM(COMP, SUB)                                                                  \
    k_assemble<COMP, SUB><<<cdiv_int(np, kBlk / SUB), kBlk, 0, s>>>(                            \
        g.x.p, g.y.p, g.px.p, g.py.p, g.cur.p, g.speed.p, c.ex.p, c.ey.p, g.cx.p, g.cy.p, x_dev, \
        c.part_u.p, c.part_v.p, c.part_cu.p, c.part_cv.p, nsplit, nsplit_body, L.tgt_ptr.p,      \
        L.tgt_ent.p,                                                                             \
        L.corr_u.p, L.corr_v.p, has_near ? 1 : 0, c.scal.p, nsig,                                 \
        g.m > 0 && choice != kFmmIndex ? c.stk_log.p : nullptr, g.no, g.n, g.m, np, tn.ptr,      \
        tn.idx, tn.nwall, tn.wall_seg, y_dev)
    if (choice == kFmmIndex) {
        if (compensated)
            STOKES_ASM(true, kAsmSubNear);
        else
            STOKES_ASM(false, kAsmSubNear);
    } else {
        if (compensated)
            STOKES_ASM(true, kAsmSub);
        else
            STOKES_ASM(false, kAsmSub);
    }
#undef STOKES_ASM
    STOKES_LAUNCH_CHECK();
}

}  // namespace stokes
