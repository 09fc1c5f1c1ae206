// One erosion step on the device: the theta-L midpoint RK2 of
// ErosionPackage/src/TimeStepping.jl -- rungekutta2 (:189-202), timestep!
// (:130-185), getderivs (:63-95), tangvel (:42-52), delete_indices (:103-126)
// -- with the spectral operations of SpectralMethods.jl (specdiff :35-41,
// specint :45-54, expsmooth :57-64, gaussfilter :67), getxy of
// ThetaLen.jl:92-101 and getstress! / getumax / getpdrop of
// DensityStress.jl:23-36, 129-167.
//
// Each per-body spectral step is one block per body: the body's samples and a
// table of e^{2 pi i j/n} sit in LDS and every transform is a direct DFT
// (n^2 complex products, n <= 512).  The solve, the shear stress and the
// pressure-drop targets run through the library's own device paths, so the
// boundary, the densities and the stresses of both stages stay in HBM.  Per
// body only the scalars (length, surface means, mean stress, dL/dt, centre
// velocity) and the 26 pressure-drop target values reach the host, which runs
// the deletion test and the scalar updates.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace stokes {

namespace {

constexpr int kSB = 256;        // threads per body block
constexpr int kMaxSpecN = 512;  // points per body the LDS tiles hold

inline int cdiv_int(int a, int b) { return (a + b - 1) / b; }

struct SpecLds {
    double twc[kMaxSpecN], tws[kMaxSpecN];  // cos, sin of 2 pi j / n
    double ar[kMaxSpecN], ai[kMaxSpecN], br[kMaxSpecN], bi[kMaxSpecN];
    double cr[kMaxSpecN], ci[kMaxSpecN], dr[kMaxSpecN];
    double red[kSB / 64];
};

__device__ void spec_twiddles(SpecLds& L, int n) {
    for (int j = threadIdx.x; j < n; j += kSB) {
        double s, c;
        sincospi(2.0 * static_cast<double>(j) / static_cast<double>(n), &s, &c);
        L.twc[j] = c;
        L.tws[j] = s;
    }
}

// out_k = sum_j in_j e^{sgn 2 pi i j k / n}, k = 0..n-1; sgn -1 forward, +1
// inverse (unnormalised).  Synchronises before and after.
__device__ void spec_dft(const SpecLds& L, int n, const double* ir, const double* ii,
                         double* orr, double* oi, double sgn) {
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += kSB) {
        double sr = 0.0, si = 0.0;
        int idx = 0;  // j k mod n
        for (int j = 0; j < n; ++j) {
            const double c = L.twc[idx], s = sgn * L.tws[idx];
            sr = __builtin_fma(ir[j], c, __builtin_fma(-ii[j], s, sr));
            si = __builtin_fma(ir[j], s, __builtin_fma(ii[j], c, si));
            idx += k;
            if (idx >= n) idx -= n;
        }
        orr[k] = sr;
        oi[k] = si;
    }
    __syncthreads();
}

// Wavenumber of DFT index k (kvec of SpectralMethods.jl:13-22 in FFT order):
// k below n/2, k - n above; the Nyquist index of an even n gets hmode n/2.
__device__ __forceinline__ double wavenum(int k, int n, int hmode) {
    if (2 * k < n) return static_cast<double>(k);
    if (2 * k == n) return static_cast<double>(hmode * (n / 2));
    return static_cast<double>(k - n);
}

// Sum over the block; every thread receives it.
__device__ double block_sum(double v, SpecLds& L) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) L.red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kSB / 64; ++w) s += L.red[w];
    return s;
}

// Spectrum (br, bi) times the operator's symbol, in place:
//   op 0 specdiff  i k (Nyquist 0)      op 1 specint  1/(i k) (k = 0 and Nyquist 0)
//   op 2 expsmooth exp(-fac k^2) (Nyquist n/2)
__device__ void spec_symbol(SpecLds& L, int n, int op, double fac) {
    for (int k = threadIdx.x; k < n; k += kSB) {
        const double r = L.br[k], i = L.bi[k];
        if (op == 0) {
            const double kv = wavenum(k, n, 0);
            L.br[k] = -i * kv;
            L.bi[k] = r * kv;
        } else if (op == 1) {
            const double kv = wavenum(k, n, 0);
            if (kv == 0.0) {
                L.br[k] = L.bi[k] = 0.0;
            } else {
                L.br[k] = i / kv;
                L.bi[k] = -r / kv;
            }
        } else {
            const double kv = wavenum(k, n, 1);
            const double e = exp(-fac * kv * kv);
            L.br[k] = r * e;
            L.bi[k] = i * e;
        }
    }
}

// (ar, ai) <- op applied to (ar, ai) as a complex signal; both real channels
// are transformed independently because every symbol is Hermitian.
__device__ void spec_apply(SpecLds& L, int n, int op, double fac) {
    spec_dft(L, n, L.ar, L.ai, L.br, L.bi, -1.0);
    spec_symbol(L, n, op, fac);
    spec_dft(L, n, L.br, L.bi, L.ar, L.ai, 1.0);
    const double inv = 1.0 / static_cast<double>(n);
    for (int j = threadIdx.x; j < n; j += kSB) {
        L.ar[j] *= inv;
        L.ai[j] *= inv;
    }
    __syncthreads();
}

// getxy (ThetaLen.jl:92-101): x = specint(L/2pi (cos th - mean)) + xsm, y likewise.
__global__ __launch_bounds__(kSB) void k_getxy(const double* __restrict__ theta,
                                               const double* __restrict__ len,
                                               const double* __restrict__ xsm,
                                               const double* __restrict__ ysm, int n,
                                               double* __restrict__ xx, double* __restrict__ yy) {
    __shared__ SpecLds L;
    const int b = blockIdx.x;
    const double* th = theta + static_cast<size_t>(b) * n;
    spec_twiddles(L, n);
    double pc = 0.0, ps = 0.0;
    for (int j = threadIdx.x; j < n; j += kSB) {
        double s, c;
        sincos(th[j], &s, &c);
        L.ar[j] = c;
        L.ai[j] = s;
        pc += c;
        ps += s;
    }
    const double mc = block_sum(pc, L) / n;
    const double ms = block_sum(ps, L) / n;
    const double f = len[b] / kTwoPi;
    for (int j = threadIdx.x; j < n; j += kSB) {
        L.ar[j] = f * (L.ar[j] - mc);
        L.ai[j] = f * (L.ai[j] - ms);
    }
    spec_apply(L, n, 1, 0.0);
    for (int j = threadIdx.x; j < n; j += kSB) {
        xx[static_cast<size_t>(b) * n + j] = L.ar[j] + xsm[b];
        yy[static_cast<size_t>(b) * n + j] = L.ai[j] + ysm[b];
    }
}

// getstress! (DensityStress.jl:31-34): atau = gaussfilter(|scale tau|, sigma)
// per body, matau = its mean; fac = sigma^2 / 2.
__global__ __launch_bounds__(kSB) void k_stress_filter(const double* __restrict__ tau,
                                                       double scale, double fac, int n,
                                                       double* __restrict__ atau,
                                                       double* __restrict__ matau) {
    __shared__ SpecLds L;
    const int b = blockIdx.x;
    spec_twiddles(L, n);
    for (int j = threadIdx.x; j < n; j += kSB) {
        L.ar[j] = fabs(scale * tau[static_cast<size_t>(b) * n + j]);
        L.ai[j] = 0.0;
    }
    spec_apply(L, n, 2, fac);
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += kSB) {
        atau[static_cast<size_t>(b) * n + j] = L.ar[j];
        s += L.ar[j];
    }
    s = block_sum(s, L);
    if (threadIdx.x == 0) matau[b] = s / n;
}

// getderivs (TimeStepping.jl:63-84) with tangvel (:42-52) for one body:
//   dtheta = specdiff(theta - alpha) + 1
//   vnorm  = atau - fixarea matau + eps matau (dtheta - 1)
//   mterm  = -2 pi mean(dtheta vnorm), vtang = specint(dtheta vnorm - mean)
//   nterm  = 2 pi / L (specdiff(atau) + dtheta vtang)
//   xsmdot = mean(-vnorm sin th + vtang cos th), ysmdot = mean(vnorm cos th + vtang sin th)
// dscal[3b + {0,1,2}] = mterm, xsmdot, ysmdot.
__global__ __launch_bounds__(kSB) void k_derivs(const double* __restrict__ theta,
                                                const double* __restrict__ len,
                                                const double* __restrict__ matau,
                                                const double* __restrict__ atau, int n, double eps,
                                                double fixarea, double* __restrict__ nterm,
                                                double* __restrict__ dscal) {
    __shared__ SpecLds L;
    const int b = blockIdx.x;
    const double* th = theta + static_cast<size_t>(b) * n;
    const double* at = atau + static_cast<size_t>(b) * n;
    const double dal = kTwoPi / static_cast<double>(n);
    const double ma = matau[b];
    spec_twiddles(L, n);
    for (int j = threadIdx.x; j < n; j += kSB) {
        const double alpha = 0.5 * dal + dal * static_cast<double>(j);
        L.ar[j] = th[j] - alpha;
        L.ai[j] = at[j];
    }
    spec_apply(L, n, 0, 0.0);  // ar = specdiff(theta - alpha), ai = specdiff(atau)
    double ps = 0.0;
    for (int j = threadIdx.x; j < n; j += kSB) {
        const double dth = L.ar[j] + 1.0;
        const double vn = (at[j] - fixarea * ma) + eps * ma * (dth - 1.0);
        L.cr[j] = dth;
        L.ci[j] = L.ai[j];
        L.dr[j] = vn;
        ps += dth * vn;
    }
    const double mdthvn = block_sum(ps, L) / n;
    for (int j = threadIdx.x; j < n; j += kSB) {
        L.ar[j] = L.cr[j] * L.dr[j] - mdthvn;
        L.ai[j] = 0.0;
    }
    spec_apply(L, n, 1, 0.0);  // ar = vtang
    const double sc = kTwoPi / len[b];
    double px = 0.0, py = 0.0;
    for (int j = threadIdx.x; j < n; j += kSB) {
        const double vt = L.ar[j], vn = L.dr[j];
        nterm[static_cast<size_t>(b) * n + j] = sc * (L.ci[j] + L.cr[j] * vt);
        double s, c;
        sincos(th[j], &s, &c);
        px += -vn * s + vt * c;
        py += vn * c + vt * s;
    }
    px = block_sum(px, L);
    py = block_sum(py, L);
    if (threadIdx.x == 0) {
        dscal[3 * b + 0] = -kTwoPi * mdthvn;
        dscal[3 * b + 1] = px / n;
        dscal[3 * b + 2] = py / n;
    }
}

// The theta update of timestep! (TimeStepping.jl:174-175):
//   theta1 = expsmooth(theta0 - alpha, fac1) + alpha + dt1 expsmooth(nterm, fac2)
__global__ __launch_bounds__(kSB) void k_theta_update(const double* __restrict__ theta0,
                                                      const double* __restrict__ nterm,
                                                      const double* __restrict__ fac1,
                                                      const double* __restrict__ fac2, double dt1,
                                                      int n, double* __restrict__ theta1) {
    __shared__ SpecLds L;
    const int b = blockIdx.x;
    const double dal = kTwoPi / static_cast<double>(n);
    spec_twiddles(L, n);
    for (int j = threadIdx.x; j < n; j += kSB) {
        const double alpha = 0.5 * dal + dal * static_cast<double>(j);
        L.ar[j] = theta0[static_cast<size_t>(b) * n + j] - alpha;
        L.ai[j] = 0.0;
        L.cr[j] = nterm[static_cast<size_t>(b) * n + j];
    }
    spec_apply(L, n, 2, fac1[b]);
    for (int j = threadIdx.x; j < n; j += kSB) {
        L.dr[j] = L.ar[j];  // expsmooth(theta0 - alpha, fac1)
        L.ar[j] = L.cr[j];
        L.ai[j] = 0.0;
    }
    spec_apply(L, n, 2, fac2[b]);
    for (int j = threadIdx.x; j < n; j += kSB) {
        const double alpha = 0.5 * dal + dal * static_cast<double>(j);
        theta1[static_cast<size_t>(b) * n + j] = (L.dr[j] + alpha) + dt1 * L.ar[j];
    }
}

// One spectral operation on every row (measurement and test hook): op 0
// specdiff, 1 specint, 2 expsmooth with factor fac[b] per row.
__global__ __launch_bounds__(kSB) void k_spec_rows(const double* __restrict__ in, int op,
                                                   const double* __restrict__ fac, int n,
                                                   double* __restrict__ out) {
    __shared__ SpecLds L;
    const int b = blockIdx.x;
    spec_twiddles(L, n);
    for (int j = threadIdx.x; j < n; j += kSB) {
        L.ar[j] = in[static_cast<size_t>(b) * n + j];
        L.ai[j] = 0.0;
    }
    spec_apply(L, n, op, op == 2 ? fac[b] : 0.0);
    for (int j = threadIdx.x; j < n; j += kSB) out[static_cast<size_t>(b) * n + j] = L.ar[j];
}

// dst row r = src row idx[r] (bodies kept after a deletion).
__global__ void k_keep_rows(const double* __restrict__ src, const int* __restrict__ idx, int n,
                            int rows, double* __restrict__ dst) {
    const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<size_t>(rows) * n) return;
    const int r = static_cast<int>(e / n), j = static_cast<int>(e - static_cast<size_t>(r) * n);
    dst[e] = src[static_cast<size_t>(idx[r]) * n + j];
}

// u0 = 2 h1 - h0 (order 2) or h1 (order 1): the guess from the last two solutions.
__global__ void k_guess(const double* __restrict__ h0, const double* __restrict__ h1, int order,
                        int N, double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    x[i] = order == 2 ? 2.0 * h1[i] - h0[i] : h1[i];
}

// dst[i] = src[idx[i]] (the unknowns of the bodies kept after a deletion).
__global__ void k_gather_idx(const double* __restrict__ src, const int* __restrict__ idx, int N,
                             double* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) dst[i] = src[idx[i]];
}

__global__ void k_fill0(double* __restrict__ x, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) x[i] = 0.0;
}

// Per-body scalars of one theta-L state (host side).
struct Scalars {
    std::vector<double> len, xsm, ysm, matau;
    void keep(const std::vector<int>& idx) {
        for (std::vector<double>* v : {&len, &xsm, &ysm, &matau}) {
            std::vector<double> w(idx.size());
            for (size_t r = 0; r < idx.size(); ++r) w[r] = (*v)[static_cast<size_t>(idx[r])];
            *v = std::move(w);
        }
    }
};

// Host vector to device; waits for the copy, so the vector may go away.
void upload(Context& c, DevBuf<double>& d, const std::vector<double>& h) {
    d.ensure(std::max<size_t>(h.size(), 1));
    if (!h.empty()) {
        STOKES_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(double),
                                  hipMemcpyHostToDevice, c.stream));
        STOKES_HIP(hipStreamSynchronize(c.stream));
    }
}

// Rows of a device array kept after a deletion (rows x n), through scratch.
void keep_rows(Context& c, DevBuf<double>& a, const std::vector<int>& idx, int n) {
    StepperBufs& S = c.sp;
    const int rows = static_cast<int>(idx.size());
    if (rows == 0) return;
    S.idx.ensure(idx.size());
    STOKES_HIP(hipMemcpyAsync(S.idx.p, idx.data(), idx.size() * sizeof(int),
                              hipMemcpyHostToDevice, c.stream));
    STOKES_HIP(hipStreamSynchronize(c.stream));
    S.tmp.ensure(static_cast<size_t>(rows) * n);
    const size_t tot = static_cast<size_t>(rows) * n;
    k_keep_rows<<<static_cast<int>((tot + 255) / 256), 256, 0, c.stream>>>(a.p, S.idx.p, n, rows,
                                                                            S.tmp.p);
    STOKES_LAUNCH_CHECK();
    STOKES_HIP(hipMemcpyAsync(a.p, S.tmp.p, tot * sizeof(double), hipMemcpyDeviceToDevice,
                              c.stream));
}

// Derivatives of the state (theta_d on device, scalars sd): solve, stress,
// fixed-pressure-drop rescaling, filter, getderivs.  Leaves nterm in c.sp.nterm
// and the density in c.sp.den; fills sd.matau and mterm / xdot / ydot.
void stage_derivs(Context& c, const StepParams& P, int n, int no, int m, const double* theta_d,
                  Scalars& sd, std::vector<double>& mterm, std::vector<double>& xdot,
                  std::vector<double>& ydot, StepReport& rep, int stage,
                  const double* guess_host) {
    StepperBufs& S = c.sp;
    hipStream_t s = c.stream;
    const size_t nb = static_cast<size_t>(n) * static_cast<size_t>(m);
    const int N = 2 * no + 2 * n * m + 3 * m;
    std::vector<double> sc(3 * static_cast<size_t>(m));
    for (int b = 0; b < m; ++b) {
        sc[static_cast<size_t>(b)] = sd.len[static_cast<size_t>(b)];
        sc[static_cast<size_t>(m + b)] = sd.xsm[static_cast<size_t>(b)];
        sc[static_cast<size_t>(2 * m + b)] = sd.ysm[static_cast<size_t>(b)];
    }
    upload(c, S.scal, sc);
    S.xx.ensure(nb);
    S.yy.ensure(nb);
    k_getxy<<<m, kSB, 0, s>>>(theta_d, S.scal.p, S.scal.p + m, S.scal.p + 2 * m, n, S.xx.p,
                              S.yy.p);
    STOKES_LAUNCH_CHECK();
    build_geometry(c, n, m, no, S.xx.p, S.yy.p, true);
    build_near_list(c, P.ibary);

    // GMRES initial guess in the preconditioned variable u = P x (the density
    // itself without a preconditioner), from the solves of earlier states
    // with these bodies
    S.den.ensure(static_cast<size_t>(N));
    if (P.warm > 0 && S.hist_count > 0 && S.hist_m == m && S.hist_n == N) {
        k_guess<<<cdiv_int(N, 256), 256, 0, s>>>(S.hist0.p, S.hist1.p,
                                                 P.warm == 2 && S.hist_count == 2 ? 2 : 1, N,
                                                 S.den.p);
        ++rep.warm_starts;
    } else if (guess_host) {
        STOKES_HIP(hipMemcpyAsync(S.den.p, guess_host, N * sizeof(double), hipMemcpyHostToDevice,
                                  s));
        ++rep.given_guesses;
    } else {
        k_fill0<<<cdiv_int(N, 256), 256, 0, s>>>(S.den.p, N);
    }
    STOKES_LAUNCH_CHECK();
    const int space = c.guess_space;
    c.guess_space = 1;
    long long it = 0;
    try {
        it = gmres_solve(c, P.ibary, P.ibc, P.maxl, c.tol, S.den.p, true);
    } catch (...) {
        c.guess_space = space;
        throw;
    }
    c.guess_space = space;
    if (c.status == 1)
        std::fprintf(stderr,
                     "libstokes_hip: GMRES stopped after %lld iterations (maxl=%d) with relative "
                     "residual %.3e (true %.3e)\n",
                     it, P.maxl, c.last_implicit_relres, c.last_true_relres);
    else if (c.status == 2)
        throw Error("GMRES breakdown after " + std::to_string(it) + " iterations");
    rep.iters[stage] = it;
    rep.true_relres_max = std::max(rep.true_relres_max, c.last_true_relres);

    // shear stress of the solver's density (computeshearstress_)
    S.tau.ensure(nb);
    shear_stress(c, S.den.p, S.tau.p);

    // fixed pressure drop: umax = 4 x0 / (pdrop + 1e-4) from the mean pressure
    // at 13 points across x = -x0 and x = +x0 (x0 = 2), raw density
    double umax = 1.0;
    if (P.fixpdrop) {
        constexpr int kNy = 13;
        const double x0 = 2.0, dy = 2.0 / kNy;
        std::vector<double> t(4 * kNy);
        for (int i = 0; i < kNy; ++i) {
            const double yl = -1.0 + 0.5 * dy + dy * i;
            t[static_cast<size_t>(i)] = x0;            // targets x: +x0 then -x0
            t[static_cast<size_t>(kNy + i)] = -x0;
            t[static_cast<size_t>(2 * kNy + i)] = yl;  // targets y
            t[static_cast<size_t>(3 * kNy + i)] = yl;
        }
        upload(c, S.tgt, t);
        S.qout.ensure(4 * 2 * kNy);
        qoi_targets(c, P.ibary, S.den.p, 2 * kNy, S.tgt.p, S.tgt.p + 2 * kNy, S.qout.p);
        std::vector<double> q(4 * 2 * kNy);
        STOKES_HIP(hipMemcpyAsync(q.data(), S.qout.p, q.size() * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
        STOKES_HIP(hipStreamSynchronize(s));
        double pp = 0.0, pm = 0.0, up = 0.0, um = 0.0;
        for (int i = 0; i < kNy; ++i) {
            up += q[static_cast<size_t>(4 * i)];
            pp += q[static_cast<size_t>(4 * i + 2)];
            um += q[static_cast<size_t>(4 * (kNy + i))];
            pm += q[static_cast<size_t>(4 * (kNy + i) + 2)];
        }
        const double pdrop = pm / kNy - pp / kNy;
        const double qplus = up / kNy, qminus = um / kNy;
        const double qavg = 0.5 * (qplus + qminus);
        if ((qplus - qminus) / qavg > 1e-3) ++rep.flux_warnings;
        umax = 4.0 * x0 / (pdrop + 1e-4);
    }
    rep.umax[stage] = umax;

    S.atau.ensure(nb);
    S.matau.ensure(static_cast<size_t>(m));
    k_stress_filter<<<m, kSB, 0, s>>>(S.tau.p, umax, 0.5 * P.sigma * P.sigma, n, S.atau.p,
                                      S.matau.p);
    STOKES_LAUNCH_CHECK();
    S.nterm.ensure(nb);
    S.dscal.ensure(3 * static_cast<size_t>(m));
    k_derivs<<<m, kSB, 0, s>>>(theta_d, S.scal.p, S.matau.p, S.atau.p, n, P.eps, P.fixarea,
                               S.nterm.p, S.dscal.p);
    STOKES_LAUNCH_CHECK();
    sd.matau.resize(static_cast<size_t>(m));
    std::vector<double> ds(3 * static_cast<size_t>(m));
    STOKES_HIP(hipMemcpyAsync(sd.matau.data(), S.matau.p, m * sizeof(double),
                              hipMemcpyDeviceToHost, s));
    STOKES_HIP(hipMemcpyAsync(ds.data(), S.dscal.p, ds.size() * sizeof(double),
                              hipMemcpyDeviceToHost, s));
    STOKES_HIP(hipStreamSynchronize(s));
    mterm.resize(static_cast<size_t>(m));
    xdot.resize(static_cast<size_t>(m));
    ydot.resize(static_cast<size_t>(m));
    for (int b = 0; b < m; ++b) {
        mterm[static_cast<size_t>(b)] = ds[static_cast<size_t>(3 * b)];
        xdot[static_cast<size_t>(b)] = ds[static_cast<size_t>(3 * b + 1)];
        ydot[static_cast<size_t>(b)] = ds[static_cast<size_t>(3 * b + 2)];
    }
}

// delete_indices (TimeStepping.jl:103-126): keep the bodies with
// len > -sqrt(2) mterm dt and mterm <= 10.
std::vector<int> kept_bodies(const std::vector<double>& len0, const std::vector<double>& mterm,
                             double dt, StepReport& rep) {
    std::vector<int> keep;
    for (size_t b = 0; b < len0.size(); ++b) {
        const double minlen = -std::sqrt(2.0) * mterm[b] * 1.0 * dt;
        if (mterm[b] > 0.0) ++rep.positive_mterm;
        if (!(len0[b] <= minlen || mterm[b] > 10.0)) keep.push_back(static_cast<int>(b));
    }
    return keep;
}

// After a solve at N unknowns and m bodies: keep its preconditioned variable
// (c.uvec) for the guesses of later solves.
void remember_solution(Context& c, int N, int m) {
    StepperBufs& S = c.sp;
    if (!c.u_valid) {
        S.hist_count = 0;
        S.hist_m = S.hist_n = -1;
        return;
    }
    if (S.hist_m != m || S.hist_n != N) S.hist_count = 0;
    S.hist0.ensure(static_cast<size_t>(N));
    S.hist1.ensure(static_cast<size_t>(N));
    if (S.hist_count > 0)
        STOKES_HIP(hipMemcpyAsync(S.hist0.p, S.hist1.p, N * sizeof(double),
                                  hipMemcpyDeviceToDevice, c.stream));
    STOKES_HIP(hipMemcpyAsync(S.hist1.p, c.uvec.p, N * sizeof(double), hipMemcpyDeviceToDevice,
                              c.stream));
    S.hist_count = std::min(S.hist_count + 1, 2);
    S.hist_m = m;
    S.hist_n = N;
}

// Bodies `keep` (of m) remain: the kept solutions drop the unknowns of the
// others (their density rows and constraint triples), so the guesses of later
// solves still cover every remaining body.
void keep_solutions(Context& c, int n, int no, int m, const std::vector<int>& keep) {
    StepperBufs& S = c.sp;
    if (S.hist_count == 0) return;
    const int mk = static_cast<int>(keep.size());
    const int Nk = 2 * no + 2 * n * mk + 3 * mk;
    std::vector<int> idx(static_cast<size_t>(Nk));
    for (int i = 0; i < 2 * no; ++i) idx[static_cast<size_t>(i)] = i;
    for (int r = 0; r < mk; ++r) {
        const int b = keep[static_cast<size_t>(r)];
        for (int j = 0; j < 2 * n; ++j)
            idx[static_cast<size_t>(2 * no + 2 * n * r + j)] = 2 * no + 2 * n * b + j;
        for (int t = 0; t < 3; ++t)
            idx[static_cast<size_t>(2 * no + 2 * n * mk + 3 * r + t)] = 2 * no + 2 * n * m + 3 * b + t;
    }
    S.idx.ensure(idx.size());
    STOKES_HIP(hipMemcpyAsync(S.idx.p, idx.data(), idx.size() * sizeof(int),
                              hipMemcpyHostToDevice, c.stream));
    S.tmp.ensure(static_cast<size_t>(Nk));
    for (DevBuf<double>* h : {&S.hist1, &S.hist0}) {
        if (h == &S.hist0 && S.hist_count < 2) break;
        k_gather_idx<<<cdiv_int(Nk, 256), 256, 0, c.stream>>>(h->p, S.idx.p, Nk, S.tmp.p);
        STOKES_LAUNCH_CHECK();
        STOKES_HIP(hipMemcpyAsync(h->p, S.tmp.p, Nk * sizeof(double), hipMemcpyDeviceToDevice,
                                  c.stream));
    }
    // the host vector and index buffer must outlive the copies above
    STOKES_HIP(hipStreamSynchronize(c.stream));
    S.hist_m = mk;
    S.hist_n = Nk;
}

// theta1, len1, xsm1, ysm1 of timestep! (TimeStepping.jl:150-182) from the
// state at t0 (theta in th0, scalars s0) and the derivatives of the state sd.
void advance(Context& c, const StepParams& P, int n, int m, const DevBuf<double>& th0,
             const Scalars& s0, const Scalars& sd, const std::vector<double>& mterm,
             const std::vector<double>& xdot, const std::vector<double>& ydot, double dt1,
             double dt2, DevBuf<double>& th1, Scalars& s1) {
    StepperBufs& S = c.sp;
    std::vector<double> fac(2 * static_cast<size_t>(m));
    s1.len.resize(static_cast<size_t>(m));
    s1.xsm.resize(static_cast<size_t>(m));
    s1.ysm.resize(static_cast<size_t>(m));
    s1.matau.assign(static_cast<size_t>(m), std::nan(""));
    for (int b = 0; b < m; ++b) {
        const size_t i = static_cast<size_t>(b);
        s1.len[i] = s0.len[i] + dt1 * mterm[i];
        if (!(s1.len[i] > 0.0)) throw Error("body length became non-positive");
        const double zeta0 = kTwoPi / s0.len[i] * s0.matau[i];
        const double zetad = kTwoPi / sd.len[i] * sd.matau[i];
        fac[i] = P.eps * dt1 * zetad;
        fac[static_cast<size_t>(m) + i] = P.eps * dt2 * (2.0 * zetad - zeta0);
        s1.xsm[i] = s0.xsm[i] + dt1 * xdot[i];
        s1.ysm[i] = s0.ysm[i] + dt1 * ydot[i];
    }
    upload(c, S.fac, fac);
    th1.ensure(static_cast<size_t>(n) * m);
    k_theta_update<<<m, kSB, 0, c.stream>>>(th0.p, S.nterm.p, S.fac.p, S.fac.p + m, dt1, n,
                                            th1.p);
    STOKES_LAUNCH_CHECK();
}

}  // namespace

void rk2_step(Context& c, const StepParams& P, int n, int no, int m, double* theta, double* len,
              double* xsm, double* ysm, const double* guess0, int* m_out, StepReport& rep) {
    if (n > kMaxSpecN) throw Error("the device stepper takes at most 512 points per body");
    if (P.ibary != 1) throw Error("the device stepper needs ibary=1 (barycentric shear stress)");
    StepperBufs& S = c.sp;
    hipStream_t s = c.stream;
    *m_out = m;
    if (m == 0) return;
    S.th0.ensure(static_cast<size_t>(n) * m);
    STOKES_HIP(hipMemcpyAsync(S.th0.p, theta, static_cast<size_t>(n) * m * sizeof(double),
                              hipMemcpyHostToDevice, s));
    Scalars s0;
    s0.len.assign(len, len + m);
    s0.xsm.assign(xsm, xsm + m);
    s0.ysm.assign(ysm, ysm + m);
    std::vector<double> mterm, xdot, ydot;

    // stage 1: timestep!(thld0, thld0, dt/2, dt/2)
    const double h = 0.5 * P.dt;
    stage_derivs(c, P, n, no, m, S.th0.p, s0, mterm, xdot, ydot, rep, 0, guess0);
    std::vector<int> keep = kept_bodies(s0.len, mterm, h, rep);
    const int N0 = 2 * no + 2 * n * m + 3 * m;
    remember_solution(c, N0, m);
    if (static_cast<int>(keep.size()) != m) {
        keep_solutions(c, n, no, m, keep);
        rep.deleted += m - static_cast<int>(keep.size());
        keep_rows(c, S.th0, keep, n);
        keep_rows(c, S.nterm, keep, n);
        s0.keep(keep);
        for (std::vector<double>* v : {&mterm, &xdot, &ydot}) {
            std::vector<double> w(keep.size());
            for (size_t r = 0; r < keep.size(); ++r) w[r] = (*v)[static_cast<size_t>(keep[r])];
            *v = std::move(w);
        }
        m = static_cast<int>(keep.size());
    }
    *m_out = m;
    if (m == 0) return;
    Scalars s05;
    advance(c, P, n, m, S.th0, s0, s0, mterm, xdot, ydot, h, h, S.th05, s05);

    // stage 2: timestep!(thld0, thld05, dt, dt/2)
    stage_derivs(c, P, n, no, m, S.th05.p, s05, mterm, xdot, ydot, rep, 1, nullptr);
    keep = kept_bodies(s0.len, mterm, P.dt, rep);
    const int N1 = 2 * no + 2 * n * m + 3 * m;
    remember_solution(c, N1, m);
    if (static_cast<int>(keep.size()) != m) {
        keep_solutions(c, n, no, m, keep);
        rep.deleted += m - static_cast<int>(keep.size());
        keep_rows(c, S.th0, keep, n);
        keep_rows(c, S.nterm, keep, n);
        s0.keep(keep);
        s05.keep(keep);
        for (std::vector<double>* v : {&mterm, &xdot, &ydot}) {
            std::vector<double> w(keep.size());
            for (size_t r = 0; r < keep.size(); ++r) w[r] = (*v)[static_cast<size_t>(keep[r])];
            *v = std::move(w);
        }
        m = static_cast<int>(keep.size());
    }
    *m_out = m;
    if (m == 0) return;
    Scalars s1;
    advance(c, P, n, m, S.th0, s0, s05, mterm, xdot, ydot, P.dt, h, S.th1, s1);
    STOKES_HIP(hipMemcpyAsync(theta, S.th1.p, static_cast<size_t>(n) * m * sizeof(double),
                              hipMemcpyDeviceToHost, s));
    STOKES_HIP(hipStreamSynchronize(s));
    for (int b = 0; b < m; ++b) {
        len[b] = s1.len[static_cast<size_t>(b)];
        xsm[b] = s1.xsm[static_cast<size_t>(b)];
        ysm[b] = s1.ysm[static_cast<size_t>(b)];
    }
}

void spectral_rows(Context& c, int op, int n, int rows, const double* in, const double* fac,
                   double* out) {
    if (n > kMaxSpecN || n < 2) throw Error("spectral rows take 2..512 points");
    if (op < 0 || op > 3) throw Error("unknown spectral operation");
    StepperBufs& S = c.sp;
    hipStream_t s = c.stream;
    const size_t tot = static_cast<size_t>(n) * rows;
    std::vector<double> h(in, in + tot);
    upload(c, S.th0, h);
    if (op == 3) {  // getxy: fac holds len, xsm, ysm per row; out holds x then y
        std::vector<double> f(fac, fac + 3 * static_cast<size_t>(rows));
        upload(c, S.scal, f);
        S.xx.ensure(tot);
        S.yy.ensure(tot);
        k_getxy<<<rows, kSB, 0, s>>>(S.th0.p, S.scal.p, S.scal.p + rows, S.scal.p + 2 * rows, n,
                                     S.xx.p, S.yy.p);
        STOKES_LAUNCH_CHECK();
        STOKES_HIP(hipMemcpyAsync(out, S.xx.p, tot * sizeof(double), hipMemcpyDeviceToHost, s));
        STOKES_HIP(hipMemcpyAsync(out + tot, S.yy.p, tot * sizeof(double), hipMemcpyDeviceToHost,
                                  s));
    } else {
        std::vector<double> f(static_cast<size_t>(rows), 0.0);
        if (op == 2) f.assign(fac, fac + rows);
        upload(c, S.fac, f);
        S.tmp.ensure(tot);
        k_spec_rows<<<rows, kSB, 0, s>>>(S.th0.p, op, S.fac.p, n, S.tmp.p);
        STOKES_LAUNCH_CHECK();
        STOKES_HIP(hipMemcpyAsync(out, S.tmp.p, tot * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    STOKES_HIP(hipStreamSynchronize(s));
}

}  // namespace stokes
