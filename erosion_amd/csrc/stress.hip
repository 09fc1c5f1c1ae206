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

constexpr int kNS = 4;  // waves per block: node slices of each source curve
constexpr int kBlk = 256;

inline int cdiv_int(int a, int b) { return (a + b - 1) / b; }

// Sum NV values per lane across the kNS slices; every thread gets totals.
template <int NV>
__device__ __forceinline__ void slice_sum(double (&v)[NV], double (*red)[kNS][64], int slice,
                                          int lane) {
#pragma unroll
    for (int q = 0; q < NV; ++q) red[q][slice][lane] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < kNS; ++t) s += red[q][t][lane];
        v[q] = s;
    }
    __syncthreads();
}

// Deformation tensor contributions of source curves [grp_beg[g], grp_end[g])
// at 64 body points per block; spart[(g*3 + {0,1,2}) * nbp + q] = E11, E12, E22.
__global__ __launch_bounds__(64 * kNS) void k_stress_eval(
    const int* __restrict__ grp_beg, const int* __restrict__ grp_end, NodeView nv,
    const double* __restrict__ ax, const double* __restrict__ ay, int no, int n, int m,
    double* __restrict__ spart) {
    __shared__ double red[12][kNS][64];
    const int lane = threadIdx.x & 63, slice = threadIdx.x >> 6;
    const int nbp = n * m;
    const int q = blockIdx.x * 64 + lane;
    const bool valid = q < nbp;
    const int qq = valid ? q : 0;
    const int tp = no + qq;
    const int bt = qq / n;
    const double xt = nv.x[tp], yt = nv.y[tp];
    const cplx zt = cmk(xt, yt);
    double E11 = 0.0, E12 = 0.0, E22 = 0.0;
    const int g = blockIdx.y;
    for (int c = grp_beg[g]; c < grp_end[g]; ++c) {
        const bool skip = !valid || c == 1 + bt;
        if (__all(skip)) continue;  // same target set in every wave: uniform
        const int len = c == 0 ? no : n;
        const int off = c == 0 ? 0 : no + (c - 1) * n;
        const bool ext = c != 0;
        const cplx a = cmk(ax[c], ay[c]);
        const int k0 = (slice * len) / kNS, k1 = ((slice + 1) * len) / kNS;
        // pass 1: S0 = sum w, N_f = sum b_f w, w = dz/(z - x)
        double v1[12];
#pragma unroll
        for (int i = 0; i < 12; ++i) v1[i] = 0.0;
        for (int k = k0; k < k1; ++k) {
            const int p = off + k;
            const cplx d = cmk(nv.x[p] - xt, nv.y[p] - yt);
            const cplx w = cmul(cmk(nv.dzr[p], nv.dzi[p]), cinv(d));
            v1[0] += w.r;
            v1[1] += w.i;
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                const cplx b = bdv(nv, f, p);
                v1[2 + 2 * f] = __builtin_fma(b.r, w.r, __builtin_fma(-b.i, w.i, v1[2 + 2 * f]));
                v1[3 + 2 * f] = __builtin_fma(b.r, w.i, __builtin_fma(b.i, w.r, v1[3 + 2 * f]));
            }
        }
        slice_sum<12>(v1, red, slice, lane);
        const cplx S0 = cmk(v1[0], v1[1]);
        const cplx idv = cinv(ext ? cmk(S0.r, S0.i - kTwoPi) : S0);
        cplx u[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) u[f] = cmul(cmk(v1[2 + 2 * f], v1[3 + 2 * f]), idv);
        // pass 2: M_f = sum (b_f - u_f) w/(z-x); exterior D_a = sum w/(z-a)
        double v2[12];
#pragma unroll
        for (int i = 0; i < 12; ++i) v2[i] = 0.0;
        for (int k = k0; k < k1; ++k) {
            const int p = off + k;
            const cplx d = cmk(nv.x[p] - xt, nv.y[p] - yt);
            const double rho2 = __builtin_fma(d.r, d.r, d.i * d.i);
            const double ir = rcp64(rho2);
            const cplx id = cmk(d.r * ir, -d.i * ir);
            const cplx w = cmul(cmk(nv.dzr[p], nv.dzi[p]), id);
            if (ext && !(sqrt(rho2) > 1.0e-10)) {
#pragma unroll
                for (int f = 0; f < 5; ++f) {
                    cplx t = coincident_term(nv, off, len, k, f, zt, a, d);
                    v2[2 * f] += t.r;
                    v2[2 * f + 1] += t.i;
                }
            } else {
                const cplx w2 = cmul(w, id);
#pragma unroll
                for (int f = 0; f < 5; ++f) {
                    const cplx t = csub(bdv(nv, f, p), u[f]);
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
        slice_sum<12>(v2, red, slice, lane);
        const cplx iden = cinv(ext ? cmul(csub(zt, a), cmk(v2[10], v2[11])) : S0);
        cplx du[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) du[f] = cmul(cmk(v2[2 * f], v2[2 * f + 1]), iden);
        // pass 3 (fields 0..2): P_f = sum (b_f - u_f - du_f (z-x)) w/(z-x)^2
        double v3[6] = {0, 0, 0, 0, 0, 0};
        for (int k = k0; k < k1; ++k) {
            const int p = off + k;
            const cplx d = cmk(nv.x[p] - xt, nv.y[p] - yt);
            const cplx id = cinv(d);
            const cplx w3 = cmul(cmul(cmul(cmk(nv.dzr[p], nv.dzi[p]), id), id), id);
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                const cplx t = csub(csub(bdv(nv, f, p), u[f]), cmul(du[f], d));
                v3[2 * f] = __builtin_fma(t.r, w3.r, __builtin_fma(-t.i, w3.i, v3[2 * f]));
                v3[2 * f + 1] = __builtin_fma(t.r, w3.i, __builtin_fma(t.i, w3.r, v3[2 * f + 1]));
            }
        }
        slice_sum<6>(v3, red, slice, lane);
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
        E11 += u1x;
        E12 += 0.5 * (u1y + u2x);
        E22 += u2y;
    }
    if (slice == 0 && valid) {
        spart[(static_cast<size_t>(g) * 3 + 0) * nbp + q] = E11;
        spart[(static_cast<size_t>(g) * 3 + 1) * nbp + q] = E12;
        spart[(static_cast<size_t>(g) * 3 + 2) * nbp + q] = E22;
    }
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
        double dr = 0.0, di = 0.0;
        for (int j = 0; j < n; ++j) {
            int idx = k - j;
            idx += idx < 0 ? n : 0;
            const double wd = dtab[idx];
            dr = __builtin_fma(wd, ex[off + j], dr);
            di = __builtin_fma(wd, ey[off + j], di);
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

}  // namespace

void shear_stress(Context& c, const double* den_dev, double* tau_dev) {
    Geometry& g = c.g;
    const int n = g.n, m = g.m, no = g.no;
    const int nbp = n * m;
    if (nbp == 0) return;
    hipStream_t s = c.stream;
    gather_density(c, den_dev);
    // every curve is a source: wall first, then the bodies
    std::vector<int> curves(static_cast<size_t>(m + 1));
    for (int i = 0; i <= m; ++i) curves[static_cast<size_t>(i)] = i;
    DevBuf<int> dcurves;
    dcurves.ensure(curves.size());
    STOKES_HIP(hipMemcpyAsync(dcurves.p, curves.data(), curves.size() * sizeof(int),
                              hipMemcpyHostToDevice, s));
    boundary_values(c, dcurves.p, m + 1, 1);
    // curve groups balanced by node count so that the grid has ~2048 blocks
    const int tiles = cdiv_int(nbp, 64);
    int G = std::max(1, std::min(m + 1, cdiv_int(2048, tiles)));
    // cum[c] = nodes of curves 0..c-1; group gi ends at the first curve
    // boundary past a fraction (gi+1)/G of all nodes
    std::vector<long long> cum(static_cast<size_t>(m + 2), 0);
    for (int cc = 0; cc <= m; ++cc)
        cum[static_cast<size_t>(cc + 1)] = cum[static_cast<size_t>(cc)] + (cc == 0 ? no : n);
    const long long total = cum[static_cast<size_t>(m + 1)];
    std::vector<int> gb, ge;
    for (int gi = 0, cur = 0; gi < G && cur <= m; ++gi) {
        const long long target = total * (gi + 1) / G;
        int end = cur + 1;
        while (end <= m && cum[static_cast<size_t>(end)] < target) ++end;
        if (gi == G - 1) end = m + 1;
        gb.push_back(cur);
        ge.push_back(end);
        cur = end;
    }
    G = static_cast<int>(gb.size());
    DevBuf<int> dgb, dge;
    dgb.ensure(gb.size());
    dge.ensure(ge.size());
    STOKES_HIP(hipMemcpyAsync(dgb.p, gb.data(), gb.size() * sizeof(int), hipMemcpyHostToDevice, s));
    STOKES_HIP(hipMemcpyAsync(dge.p, ge.data(), ge.size() * sizeof(int), hipMemcpyHostToDevice, s));
    c.spart.ensure(static_cast<size_t>(G) * 3 * static_cast<size_t>(nbp));
    NodeView nv{g.x.p, g.y.p, g.dzr.p, g.dzi.p, c.bd.p, g.iar.p, g.iai.p, g.mx.p, g.my.p,
                c.ex.p, c.ey.p, static_cast<size_t>(g.npts)};
    if (c.time_kernels) c.t_stress_eval.start(s);
    k_stress_eval<<<dim3(tiles, G), 64 * kNS, 0, s>>>(dgb.p, dge.p, nv, g.ax.p, g.ay.p, no, n, m,
                                                      c.spart.p);
    STOKES_LAUNCH_CHECK();
    if (c.time_kernels) {
        // (target point, source node) pairs: every other curve's nodes
        c.t_stress_eval.stop(s, static_cast<double>(nbp) * static_cast<double>(total - n));
    }
    k_stress_final<<<cdiv_int(nbp, kBlk), kBlk, 0, s>>>(c.spart.p, G, g.x.p, g.y.p, g.px.p,
                                                        g.py.p, g.speed.p, g.cx.p, g.cy.p, c.ex.p,
                                                        c.ey.p, den_dev, g.dtab_n.p, no, n, m,
                                                        tau_dev);
    STOKES_LAUNCH_CHECK();
    STOKES_HIP(hipStreamSynchronize(s));
}

}  // namespace stokes
