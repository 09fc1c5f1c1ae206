// Boundary geometry on the device and the near-pair lists.
//
// Restates inner_geometry / outer_geometry / compute_derivs / fourierDiff
// (ErosionPackage/fortran_code/stokesSolver.f:136-204, 348-388, 417-457).
// The spectral derivative with the Nyquist mode removed is applied as the
// circulant matrix d(j-k), d(k) = (-1)^k cot(pi k / N) / 2, d(0) = 0 -- the
// same linear map fourierDiff computes with FFTPACK -- so one kernel family
// serves every curve length with coalesced reads and no FFT plans.
#include <algorithm>
#include <cstring>
#include <cmath>

#include "common.hpp"

namespace stokes {

namespace {

constexpr int kBlk = 256;

// Wall shape (stokesSolver.f:184-191) and copy of the body points.
__global__ void k_points(const double* __restrict__ xx, const double* __restrict__ yy, int no,
                         int nbody_pts, double* __restrict__ x, double* __restrict__ y) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < no) {
        double th = static_cast<double>(i) * (kTwoPi / static_cast<double>(no));
        double c = cos(th), s = sin(th);
        double r = pow(pow(c, 8.0) + pow(s, 8.0), -1.0 / 8.0);
        x[i] = 3.0 * r * c;
        y[i] = r * s;
    } else if (i < no + nbody_pts) {
        x[i] = xx[i - no];
        y[i] = yy[i - no];
    }
}

// One spectral derivative of (fx, fy) on every curve:
//   out_j = sum_k d(j-k) (f_k - f_j)
// (the rows of d sum to zero, so subtracting f_j changes nothing in exact
// arithmetic and removes the O(N) cancellation of the large entries near the
// diagonal).  grid: (ceil(maxlen/256), ncurves); curve 0 = wall (length no,
// table dno), curve c>0 = body c-1 (length n, table dn).
__global__ void k_deriv(const double* __restrict__ fx, const double* __restrict__ fy, int no,
                        int n, const double* __restrict__ dno, const double* __restrict__ dn,
                        int first_curve, double* __restrict__ ox, double* __restrict__ oy) {
    const int c = first_curve + static_cast<int>(blockIdx.y);
    const int len = c == 0 ? no : n;
    const int off = c == 0 ? 0 : no + (c - 1) * n;
    const double* __restrict__ d = c == 0 ? dno : dn;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    const double xj = fx[off + j], yj = fy[off + j];
    double sx = 0.0, sy = 0.0;
    for (int k = 0; k < len; ++k) {
        int idx = j - k;
        idx += idx < 0 ? len : 0;
        double w = d[idx];
        sx = __builtin_fma(w, fx[off + k] - xj, sx);
        sy = __builtin_fma(w, fy[off + k] - yj, sy);
    }
    ox[off + j] = sx;
    oy[off + j] = sy;
}

// compute_derivs (stokesSolver.f:447-453) plus the weights every kernel
// uses; wall normal and curvature negated (stokesSolver.f:195-199).
__global__ void k_frame(const double* __restrict__ dx, const double* __restrict__ dy,
                        const double* __restrict__ ddx, const double* __restrict__ ddy, int no,
                        int npts, int n, double* __restrict__ px, double* __restrict__ py,
                        double* __restrict__ cur, double* __restrict__ speed,
                        double* __restrict__ dzr, double* __restrict__ dzi,
                        double* __restrict__ mx, double* __restrict__ my) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    const bool wall = i < no;
    const double len = wall ? static_cast<double>(no) : static_cast<double>(n);
    double Dx = dx[i], Dy = dy[i];
    double sp = sqrt(Dx * Dx + Dy * Dy);
    double nx = -Dy / sp, ny = Dx / sp;
    double k = (Dy * ddx[i] - Dx * ddy[i]) / (sp * sp * sp);
    if (wall) {
        nx = -nx;
        ny = -ny;
        k = -k;
    }
    px[i] = nx;
    py[i] = ny;
    cur[i] = k;
    speed[i] = sp;
    const double dth = kTwoPi / len;
    dzr[i] = Dx * dth;
    dzi[i] = Dy * dth;
    // trapezoid weight speed*2pi/N over pi, folded into the normal
    const double w = sp * kTwoPi / len / kPi;
    mx[i] = nx * w;
    my[i] = ny * w;
}

// Per-curve point means: body centres as inner_geometry (sum / n,
// stokesSolver.f:151-160) and the point a of laplaceExteriorHolomorphic
// (sum of x/n, stokesSolver.f:5452-5459).  One block per curve.
__global__ void k_curve_means(const double* __restrict__ x, const double* __restrict__ y, int no,
                              int n, double* __restrict__ ax, double* __restrict__ ay,
                              double* __restrict__ cx, double* __restrict__ cy) {
    __shared__ double s[4][kBlk];
    const int c = blockIdx.x;
    const int len = c == 0 ? no : n;
    const int off = c == 0 ? 0 : no + (c - 1) * n;
    const double inv = 1.0 / static_cast<double>(len);
    double sx = 0, sy = 0, qx = 0, qy = 0;
    for (int k = threadIdx.x; k < len; k += blockDim.x) {
        sx += x[off + k];
        sy += y[off + k];
        qx += x[off + k] * inv;
        qy += y[off + k] * inv;
    }
    s[0][threadIdx.x] = sx;
    s[1][threadIdx.x] = sy;
    s[2][threadIdx.x] = qx;
    s[3][threadIdx.x] = qy;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w)
            for (int q = 0; q < 4; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ax[c] = s[2][0];
        ay[c] = s[3][0];
        if (c > 0) {
            cx[c - 1] = s[0][0] / static_cast<double>(len);
            cy[c - 1] = s[1][0] / static_cast<double>(len);
        }
    }
}

// 1/(z - a) at every point of every body curve (wall entries unused).
__global__ void k_inv_za(const double* __restrict__ x, const double* __restrict__ y,
                         const double* __restrict__ ax, const double* __restrict__ ay, int no,
                         int n, int npts, double* __restrict__ iar, double* __restrict__ iai) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    int c = i < no ? 0 : 1 + (i - no) / n;
    cplx inv = cinv(cmk(x[i] - ax[c], y[i] - ay[c]));
    iar[i] = inv.r;
    iai[i] = inv.i;
}

std::vector<double> deriv_table(int len) {
    std::vector<double> d(static_cast<size_t>(len), 0.0);
    for (int k = 1; k < len; ++k) {
        if (2 * k == len) continue;  // cot(pi/2) = 0
        double t = std::tan(kPi * static_cast<double>(k) / static_cast<double>(len));
        d[static_cast<size_t>(k)] = 0.5 * ((k & 1) ? -1.0 : 1.0) / t;
    }
    return d;
}

inline int cdiv_int(int a, int b) { return (a + b - 1) / b; }

}  // namespace

void build_geometry(Context& c, int n, int m, int no, const double* xx, const double* yy,
                    bool on_device) {
    Geometry& g = c.g;
    const size_t nb = static_cast<size_t>(n) * static_cast<size_t>(m);
    // the time stepper hands the same boundary to the solve, the stress and
    // the target evaluations of one state: keep the built geometry for it
    // (device-resident input is always built)
    if (!on_device && g.version > 0 && n == g.key_n && m == g.key_m && no == g.key_no &&
        (nb == 0 || (std::memcmp(xx, g.key_x.data(), nb * sizeof(double)) == 0 &&
                     std::memcmp(yy, g.key_y.data(), nb * sizeof(double)) == 0)))
        return;
    g.key_n = -1;  // invalid until this build completes
    g.n = n;
    g.m = m;
    g.no = no;
    g.npts = no + n * m;
    g.nunk = 2 * no + 2 * n * m + 3 * m;
    const size_t np = static_cast<size_t>(g.npts);
    for (DevBuf<double>* b : {&g.x, &g.y, &g.px, &g.py, &g.cur, &g.speed, &g.dzr, &g.dzi, &g.mx,
                              &g.my, &g.iar, &g.iai})
        b->ensure(np);
    g.ax.ensure(static_cast<size_t>(m + 1));
    g.ay.ensure(static_cast<size_t>(m + 1));
    g.cx.ensure(static_cast<size_t>(std::max(m, 1)));
    g.cy.ensure(static_cast<size_t>(std::max(m, 1)));
    g.dtab_n.ensure(static_cast<size_t>(std::max(n, 1)));
    g.dtab_no.ensure(static_cast<size_t>(no));
    hipStream_t s = c.stream;

    // inputs
    DevBuf<double>&dxx = g.in_x, &dyy = g.in_y, &t1x = g.t1x, &t1y = g.t1y, &t2x = g.t2x,
                   &t2y = g.t2y;
    dxx.ensure(nb);
    dyy.ensure(nb);
    if (nb) {
        const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        STOKES_HIP(hipMemcpyAsync(dxx.p, xx, nb * sizeof(double), kind, s));
        STOKES_HIP(hipMemcpyAsync(dyy.p, yy, nb * sizeof(double), kind, s));
    }
    std::vector<double> tn = deriv_table(std::max(n, 1)), tno = deriv_table(no);
    STOKES_HIP(hipMemcpyAsync(g.dtab_n.p, tn.data(), tn.size() * sizeof(double),
                              hipMemcpyHostToDevice, s));
    STOKES_HIP(hipMemcpyAsync(g.dtab_no.p, tno.data(), tno.size() * sizeof(double),
                              hipMemcpyHostToDevice, s));
    t1x.ensure(np);
    t1y.ensure(np);
    t2x.ensure(np);
    t2y.ensure(np);

    k_points<<<cdiv_int(g.npts, kBlk), kBlk, 0, s>>>(dxx.p, dyy.p, no, n * m, g.x.p, g.y.p);
    STOKES_LAUNCH_CHECK();
    // the wall's derivatives are kept from the last build with the same no;
    // only the bodies' are recomputed (the wall's O(no^2) sums were the
    // longest blocks of the launch)
    const bool wall_kept = g.wall_no == no;
    const int c0 = wall_kept ? 1 : 0;
    if (m + 1 - c0 > 0) {
        dim3 dgrid(cdiv_int(wall_kept ? n : std::max(n, no), kBlk), m + 1 - c0);
        k_deriv<<<dgrid, kBlk, 0, s>>>(g.x.p, g.y.p, no, n, g.dtab_no.p, g.dtab_n.p, c0, t1x.p,
                                       t1y.p);
        STOKES_LAUNCH_CHECK();
        k_deriv<<<dgrid, kBlk, 0, s>>>(t1x.p, t1y.p, no, n, g.dtab_no.p, g.dtab_n.p, c0, t2x.p,
                                       t2y.p);
        STOKES_LAUNCH_CHECK();
    }
    double* const wd[4] = {t1x.p, t1y.p, t2x.p, t2y.p};
    const size_t wb = static_cast<size_t>(no) * sizeof(double);
    if (wall_kept) {
        for (int q = 0; q < 4; ++q)
            STOKES_HIP(hipMemcpyAsync(wd[q], g.wall_d.p + static_cast<size_t>(q) * no, wb,
                                      hipMemcpyDeviceToDevice, s));
    } else {
        g.wall_d.ensure(4 * static_cast<size_t>(no));
        for (int q = 0; q < 4; ++q)
            STOKES_HIP(hipMemcpyAsync(g.wall_d.p + static_cast<size_t>(q) * no, wd[q], wb,
                                      hipMemcpyDeviceToDevice, s));
        g.wall_no = no;
    }
    k_frame<<<cdiv_int(g.npts, kBlk), kBlk, 0, s>>>(t1x.p, t1y.p, t2x.p, t2y.p, no, g.npts, n,
                                                    g.px.p, g.py.p, g.cur.p, g.speed.p, g.dzr.p,
                                                    g.dzi.p, g.mx.p, g.my.p);
    STOKES_LAUNCH_CHECK();
    k_curve_means<<<m + 1, kBlk, 0, s>>>(g.x.p, g.y.p, no, n, g.ax.p, g.ay.p, g.cx.p, g.cy.p);
    STOKES_LAUNCH_CHECK();
    k_inv_za<<<cdiv_int(g.npts, kBlk), kBlk, 0, s>>>(g.x.p, g.y.p, g.ax.p, g.ay.p, no, n, g.npts,
                                                     g.iar.p, g.iai.p);
    STOKES_LAUNCH_CHECK();

    // host copies for the near-pair selection
    g.h_x.resize(np);
    g.h_y.resize(np);
    g.h_speed.resize(np);
    g.h_cx.resize(static_cast<size_t>(m));
    g.h_cy.resize(static_cast<size_t>(m));
    STOKES_HIP(hipMemcpyAsync(g.h_x.data(), g.x.p, np * sizeof(double), hipMemcpyDeviceToHost, s));
    STOKES_HIP(hipMemcpyAsync(g.h_y.data(), g.y.p, np * sizeof(double), hipMemcpyDeviceToHost, s));
    STOKES_HIP(hipMemcpyAsync(g.h_speed.data(), g.speed.p, np * sizeof(double),
                              hipMemcpyDeviceToHost, s));
    if (m) {
        STOKES_HIP(hipMemcpyAsync(g.h_cx.data(), g.cx.p, static_cast<size_t>(m) * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
        STOKES_HIP(hipMemcpyAsync(g.h_cy.data(), g.cy.p, static_cast<size_t>(m) * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
    }
    STOKES_HIP(hipStreamSynchronize(s));
    // the body points are copied unchanged into x, y after the wall points
    g.key_x.assign(g.h_x.begin() + no, g.h_x.end());
    g.key_y.assign(g.h_y.begin() + no, g.h_y.end());
    g.key_n = n;
    g.key_m = m;
    g.key_no = no;
    ++g.version;
}

// Near-pair selection of matvec_DLP_fmmbary (stokesSolver.f:694-926), same
// thresholds and the same arithmetic order:
//   wall -> body point q of body b:  |c_y(b)| >= 1-dsum and |y_q| >= 1-dsum,
//        dsum = rbar_b + 4*lw*2pi/no, lw the INTEGER mean wall speed
//   body b -> wall point j:          same body test, |c_b - z_j| <= 1-|c_y|+2 rbar_b
//   body s -> point q of body t!=s:  |c_s-c_t| <= R and |c_s - z_q| <= R,
//        R = (rbar_s + rbar_t)(1 + 4*2pi/n)
// ibary=0 gives an empty list (matvec_DLP: trapezoid everywhere).
void build_near_list(Context& c, int ibary) {
    Geometry& g = c.g;
    NearList& L = c.near;
    if (L.built_version == g.version && L.built_ibary == ibary) return;
    L.built_version = -1;
    const int n = g.n, m = g.m, no = g.no;
    std::vector<std::vector<int>> by_src(static_cast<size_t>(m + 1));
    if (ibary == 1 && m > 0) {
        const double alpha1 = 4.0, alpha2 = 4.0;
        std::vector<double> rtar(static_cast<size_t>(m));
        for (int b = 0; b < m; ++b) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s = s + g.h_speed[static_cast<size_t>(no + b * n + k)];
            rtar[static_cast<size_t>(b)] = s / static_cast<double>(n);
        }
        // 'lw' is implicitly INTEGER in the reference (stokesSolver.f:705-709)
        int lw = 0;
        for (int k = 0; k < no; ++k)
            lw = static_cast<int>(static_cast<double>(lw) + g.h_speed[static_cast<size_t>(k)]);
        lw = lw / no;
        for (int it = 0; it < m; ++it) {
            const double rt = rtar[static_cast<size_t>(it)];
            const double cyt = g.h_cy[static_cast<size_t>(it)], cxt = g.h_cx[static_cast<size_t>(it)];
            double dsum = rt + alpha2 * static_cast<double>(lw) * kTwoPi / static_cast<double>(no);
            double dsum2 = 1.0 - dsum;
            if (!(std::fabs(cyt) >= dsum2)) continue;
            for (int k = 0; k < n; ++k) {
                int q = no + it * n + k;
                if (std::fabs(g.h_y[static_cast<size_t>(q)]) >= dsum2) by_src[0].push_back(q);
            }
            double r = 1.0 - std::fabs(cyt) + 2.0 * rt;
            for (int j = 0; j < no; ++j) {
                double dx = cxt - g.h_x[static_cast<size_t>(j)];
                double dy = cyt - g.h_y[static_cast<size_t>(j)];
                double d2 = dx * dx + dy * dy;
                if (d2 <= r * r) by_src[static_cast<size_t>(1 + it)].push_back(j);
            }
        }
        for (int is = 0; is < m; ++is) {
            const double rs = rtar[static_cast<size_t>(is)];
            const double cxs = g.h_cx[static_cast<size_t>(is)], cys = g.h_cy[static_cast<size_t>(is)];
            for (int it = 0; it < m; ++it) {
                if (it == is) continue;
                const double rt = rtar[static_cast<size_t>(it)];
                double dx = cxs - g.h_cx[static_cast<size_t>(it)];
                double dy = cys - g.h_cy[static_cast<size_t>(it)];
                double d2 = dx * dx + dy * dy;
                double rsum = rs + rt + alpha1 * (rs + rt) * kTwoPi / static_cast<double>(n);
                if (!(d2 <= rsum * rsum)) continue;
                for (int k = 0; k < n; ++k) {
                    int q = no + it * n + k;
                    dx = cxs - g.h_x[static_cast<size_t>(q)];
                    dy = cys - g.h_y[static_cast<size_t>(q)];
                    d2 = dx * dx + dy * dy;
                    if (d2 <= rsum * rsum) by_src[static_cast<size_t>(1 + is)].push_back(q);
                }
            }
        }
    }
    // flatten: entries sorted by source curve
    std::vector<int> ent_tgt, ent_src, active;
    for (int src = 0; src <= m; ++src) {
        auto& v = by_src[static_cast<size_t>(src)];
        if (v.empty()) continue;
        active.push_back(src);
        ent_src.insert(ent_src.end(), v.size(), src);
        ent_tgt.insert(ent_tgt.end(), v.begin(), v.end());
    }
    const int ne = static_cast<int>(ent_tgt.size());
    std::vector<int> tgt_ptr(static_cast<size_t>(g.npts + 1), 0), tgt_ent(static_cast<size_t>(ne));
    for (int e = 0; e < ne; ++e) tgt_ptr[static_cast<size_t>(ent_tgt[static_cast<size_t>(e)] + 1)]++;
    for (int i = 0; i < g.npts; ++i) tgt_ptr[static_cast<size_t>(i + 1)] += tgt_ptr[static_cast<size_t>(i)];
    {
        std::vector<int> fill(tgt_ptr.begin(), tgt_ptr.end() - 1);
        for (int e = 0; e < ne; ++e)
            tgt_ent[static_cast<size_t>(fill[static_cast<size_t>(ent_tgt[static_cast<size_t>(e)])]++)] = e;
    }
    L.nentries = ne;
    L.nactive = static_cast<int>(active.size());
    L.wall_active = by_src[0].empty() ? 0 : 1;
    {
        const double wall_ent = static_cast<double>(by_src[0].size());
        L.gather_bytes = 14.0 * 8.0 * (wall_ent * no + (ne - wall_ent) * n) + 16.0 * ne;
    }
    auto up = [&](DevBuf<int>& d, const std::vector<int>& h) {
        d.ensure(std::max<size_t>(h.size(), 1));
        if (!h.empty())
            STOKES_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice,
                                      c.stream));
    };
    up(L.ent_tgt, ent_tgt);
    up(L.ent_src, ent_src);
    up(L.tgt_ptr, tgt_ptr);
    up(L.tgt_ent, tgt_ent);
    up(L.active, active);
    L.corr_u.ensure(static_cast<size_t>(std::max(ne, 1)));
    L.corr_v.ensure(static_cast<size_t>(std::max(ne, 1)));
    STOKES_HIP(hipStreamSynchronize(c.stream));
    L.built_version = g.version;
    L.built_ibary = ibary;
}

}  // namespace stokes
