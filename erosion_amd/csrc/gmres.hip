// Unrestarted GMRES on the device.
//
// Restates DGMRES/DPIGMR as solveBIE configures them
// (ErosionPackage/fortran_code/stokesSolver.f:481-497,
//  ErosionPackage/fortran_code/gmres_routines.f:380-554, 1524-1963):
// no preconditioner, no restart, Arnoldi with a Hessenberg QR by Givens
// rotations (DHEQR), stop on the implicit residual rho = |prod s_i| ||r0||
// relative to ||b|| (ISDGMR, ITOL=0), solution by DHELS.
//
// The reference orthogonalises with modified Gram-Schmidt (DORTH), k
// dependent dot/axpy pairs per iteration.  Here the Krylov basis V lives in
// HBM as (maxl+1) rows of npad doubles and each Gram-Schmidt pass is two
// streaming kernels (all dots at once, then one fused update): classical
// Gram-Schmidt with a second pass when ||w|| drops below ||w_in||/sqrt(2)
// (the Daniel-Gragg-Kaufman-Stewart test).  The Givens update applies the k
// stored rotations with a one-wave scan of the affine carry recurrence
// c_{i+1} = s_i c_i + c_i' h_{i+1} instead of a k-step serial loop.
//
// The host enqueues chunks of iterations and reads the device state once per
// chunk; kernels of iterations after convergence return immediately.
#include <algorithm>
#include <cstring>
#include <chrono>
#include <cmath>

#include "common.hpp"

namespace stokes {

namespace {

constexpr int kBlk = 256;
constexpr int kRows = 4;       // Krylov rows per block in the dot kernel
constexpr int kXcds = 8;       // MI355X: accelerator complex dies, one L2 each
constexpr int kChunkIters = 16;
// DCGS2's second pass runs as one kernel from this many blocks of 512
// elements up (vectors of 2^17 doubles), as split partials below
constexpr int kFusedUpdateBlocks = 256;
// preconditioned solves: implicit target of a cycle as a fraction of the
// tolerance, and the most cycles one solve may run
constexpr double kPcTolFrac = 0.5;
constexpr int kMaxCycles = 4;
// mixed precision: the far field runs in FP32 once the implicit residual is
// below mixed_switch * tol (default 1e4, Context::mixed_switch).  An inexact
// operator whose error is a fraction eps of ||A x|| perturbs the attainable
// residual by about eps * ||r_k|| / ||b|| at iteration k (Simoncini and Szyld
// 2003), so with eps ~ 1e-7 for the FP32 far field the switch costs nothing
// measurable at 1e4 * tol; the true residual is always formed in FP64 and a
// cycle that misses the tolerance restarts.

inline int cdiv_int(int a, int b) { return (a + b - 1) / b; }
inline size_t round_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

// Sum over the block (blockDim.x a multiple of 64), returned to every
// thread: a butterfly within each wave, then the waves' sums through LDS in
// wave order.  Two barriers; sh holds one double per wave and may be reused
// after the call.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int nw = static_cast<int>(blockDim.x) >> 6;
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < nw; ++w) r += sh[w];
    __syncthreads();
    return r;
}

// out = a*x + b*y over n (n even, 16-byte aligned rows); per-block sum of out^2
__global__ __launch_bounds__(kBlk) void k_axpby_norm(double a, const double* __restrict__ x,
                                                     double b, const double* __restrict__ y,
                                                     double* __restrict__ out, size_t n,
                                                     double* __restrict__ npart) {
    __shared__ double sh[kBlk];
    double acc = 0.0;
    for (size_t e = (static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x) * 2; e < n;
         e += static_cast<size_t>(gridDim.x) * kBlk * 2) {
        double2 xv = *reinterpret_cast<const double2*>(x + e);
        double2 yv = y ? *reinterpret_cast<const double2*>(y + e) : make_double2(0.0, 0.0);
        double2 o = make_double2(a * xv.x + b * yv.x, a * xv.y + b * yv.y);
        if (out) *reinterpret_cast<double2*>(out + e) = o;
        acc += o.x * o.x + o.y * o.y;
    }
    double s = block_sum(acc, sh);
    if (threadIdx.x == 0) npart[blockIdx.x] = s;
}

// d[i] = +-ulp(x[i]) / 2 with a sign from a hash of i, zero in the padding:
// the size of the rounding of a density to FP64, whose image A d bounds the
// residual any FP64 density near x can reach
__global__ __launch_bounds__(kBlk) void k_half_ulp(const double* __restrict__ x,
                                                   double* __restrict__ d, int n, int npad) {
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= npad) return;
    double v = 0.0;
    if (i < n) {
        const double a = fabs(x[i]);
        v = 0.5 * (nextafter(a, __builtin_inf()) - a);
        // splitmix64 finaliser of the index: one pseudo-random sign per entry
        unsigned long long h = static_cast<unsigned long long>(i) + 0x9E3779B97F4A7C15ull;
        h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
        h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
        h ^= h >> 31;
        if (h & 1ull) v = -v;
    }
    d[i] = v;
}

// x += y
__global__ __launch_bounds__(kBlk) void k_add(double* __restrict__ x, const double* __restrict__ y,
                                              size_t n) {
    const size_t e = static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x;
    if (e < n) x[e] += y[e];
}

// zero the padding [n, npad) of rows [0, nrows) of a row-major (npad) array
__global__ void k_zero_tails(double* __restrict__ v, int n, int npad) {
    const int i = n + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npad) v[static_cast<size_t>(blockIdx.y) * npad + i] = 0.0;
}

__global__ void k_zero_tail(double* __restrict__ v, int n, int npad) {
    int i = n + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npad) v[i] = 0.0;
}

// Initial state: ||b||, ||r0|| from per-block partial sums.
__global__ void k_init_state(const double* __restrict__ bpart, int nb,
                             const double* __restrict__ rpart, int nr, double tol,
                             GmresState* __restrict__ st) {
    if (threadIdx.x != 0) return;
    double sb = 0.0, sr = 0.0;
    for (int i = 0; i < nb; ++i) sb += bpart[i];
    for (int i = 0; i < nr; ++i) sr += rpart[i];
    GmresState s{};
    s.bnrm = sqrt(sb);
    s.r0nrm = sqrt(sr);
    s.prod = 1.0;
    s.rho = s.r0nrm;
    s.k = 0;
    s.done = (s.r0nrm / s.bnrm <= tol) ? 1 : 0;
    s.reorth = 0;
    s.nreorth = 0;
    *st = s;
}

__global__ void k_scale_row(double* __restrict__ v, size_t n, const GmresState* __restrict__ st,
                            int use_r0) {
    if (st->done) return;
    const double s = use_r0 ? 1.0 / st->r0nrm : st->inv_snormw;
    for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
         e += static_cast<size_t>(gridDim.x) * blockDim.x)
        v[e] *= s;
}

// Gram-Schmidt with delayed reorthogonalisation (DCGS2: Bielich, Langou,
// Thomas, Swirydowicz, Yamazaki and Boman 2022), two passes over the Krylov
// basis per iteration instead of CGS2's four.  Iteration j finds V[0..j-1]
// final (orthonormal Q) and V[j] = u, the previous iteration's vector after
// one projection, normalised.  With w = A u in V[j+1], one pass forms
//   a = Q^T u, alpha = u.u, b = Q^T w, beta = u.w, omega = w.w;
// then r = sqrt(alpha - a.a), q_j = (u - Q a)/r (u's second pass), the
// Arnoldi column j-1 is completed with a (its last vector was u, now q_j),
// and, with g = [b; (beta - a.b)/r] = [Q q_j]^T w,
//   A q_j = (w - [Q q_j] H a)/r  gives column j: (g - H a)/r  (provisional),
//   w' = (w - [Q q_j] g)/r,  ||w'|| = sqrt(omega - g.g)/r,
// so the second pass writes q_j into V[j] and w'/||w'|| into V[j+1].  Column
// j is completed, and its Givens rotation formed, in iteration j + 1: the
// stop test runs one iteration late (one operator application more per
// solve).  H_raw keeps the Arnoldi columns, H their rotated (R) form.

// Partial dots of rows [r0, r0+kRows) of V with u and with w over one chunk:
// hu[chunk][r], hw[chunk][r] (chunk-major, nrows per chunk, so that the sums
// over chunks read consecutive rows in consecutive lanes).  A 1-D grid of bx x
// nchunk blocks (nchunk a multiple of kXcds): workgroups go to the XCDs in
// turn by linear index, so block b runs on XCD b % kXcds, and the bx row
// blocks of chunk c are all given to XCD c % kXcds, whose L2 then fetches
// the chunk of u and w once instead of once per XCD.
__global__ __launch_bounds__(kBlk) void k_dots2(const double* __restrict__ V, size_t npad,
                                                const double* __restrict__ u,
                                                const double* __restrict__ w, int nrows,
                                                size_t chunk, int bx,
                                                const GmresState* __restrict__ st,
                                                double* __restrict__ hu, double* __restrict__ hw) {
    __shared__ double sh[kBlk];
    if (st->done) return;
    const int xcd = static_cast<int>(blockIdx.x) % kXcds, k = static_cast<int>(blockIdx.x) / kXcds;
    const int r0 = (k % bx) * kRows;
    const int c = xcd + kXcds * (k / bx);
    const size_t e0 = static_cast<size_t>(c) * chunk;
    const size_t e1 = std::min(npad, e0 + chunk);
    double au[kRows], aw[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) au[r] = aw[r] = 0.0;
    for (size_t e = e0 + 2 * threadIdx.x; e < e1; e += 2 * kBlk) {
        const double2 uv = *reinterpret_cast<const double2*>(u + e);
        const double2 wv = *reinterpret_cast<const double2*>(w + e);
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
            if (r0 + r < nrows) {
                const double2 vv =
                    *reinterpret_cast<const double2*>(V + static_cast<size_t>(r0 + r) * npad + e);
                au[r] = __builtin_fma(vv.x, uv.x, __builtin_fma(vv.y, uv.y, au[r]));
                aw[r] = __builtin_fma(vv.x, wv.x, __builtin_fma(vv.y, wv.y, aw[r]));
            }
        }
    }
    // the 2 kRows block sums together: a butterfly per wave for each, the
    // waves' values through LDS and added in wave order (as block_sum), one
    // barrier
    constexpr int kW = kBlk / 64;
    static_assert(2 * kRows * kW <= kBlk, "sums fit sh");
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
        for (int off = 32; off > 0; off >>= 1) {
            au[r] += __shfl_xor(au[r], off);
            aw[r] += __shfl_xor(aw[r], off);
        }
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
            sh[(2 * r) * kW + wv] = au[r];
            sh[(2 * r + 1) * kW + wv] = aw[r];
        }
    }
    __syncthreads();
    const int t = threadIdx.x, r = t >> 1;
    if (t < 2 * kRows && r0 + r < nrows) {
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < kW; ++w) v += sh[t * kW + w];
        ((t & 1) ? hw : hu)[static_cast<size_t>(c) * nrows + r0 + r] = v;
    }
}

// wpart[split][e] = sum_{rows of split} h_r V[r][e] for the final solution
// x += V y (h = y).
__global__ __launch_bounds__(kBlk) void k_update(const double* __restrict__ V, size_t npad,
                                                 const double* __restrict__ h, int nrows,
                                                 int rows_per_split, double* __restrict__ wpart) {
    __shared__ double hs[kBlk];
    const size_t e = (static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x) * 2;
    const bool valid = e < npad;
    const int s = blockIdx.y;
    const int ra = s * rows_per_split, rb = min(nrows, ra + rows_per_split);
    double ax = 0.0, ay = 0.0;
    for (int t0 = ra; t0 < rb; t0 += kBlk) {
        const int tn = min(kBlk, rb - t0);
        __syncthreads();
        if (static_cast<int>(threadIdx.x) < tn) hs[threadIdx.x] = h[t0 + threadIdx.x];
        __syncthreads();
        if (!valid) continue;
        for (int q = 0; q < tn; ++q) {
            const double hr = hs[q];
            const double2 v0 =
                *reinterpret_cast<const double2*>(V + static_cast<size_t>(t0 + q) * npad + e);
            ax = __builtin_fma(hr, v0.x, ax);
            ay = __builtin_fma(hr, v0.y, ay);
        }
    }
    if (valid)
        *reinterpret_cast<double2*>(wpart + static_cast<size_t>(s) * npad + e) = make_double2(ax, ay);
}

// The same over two coefficient sets at once (one read of V): the second
// pass of DCGS2, wa[split][e] = sum a_r V[r][e], wb[split][e] = sum b_r V[r][e].
__global__ __launch_bounds__(kBlk) void k_update2(const double* __restrict__ V, size_t npad,
                                                  const double* __restrict__ ca,
                                                  const double* __restrict__ cb, int nrows,
                                                  int rows_per_split,
                                                  const GmresState* __restrict__ st,
                                                  double* __restrict__ wa, double* __restrict__ wb) {
    __shared__ double ha[kBlk], hb[kBlk];
    if (st->done) return;
    const size_t e = (static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x) * 2;
    const bool valid = e < npad;
    const int s = blockIdx.y;
    const int ra = s * rows_per_split, rb = min(nrows, ra + rows_per_split);
    double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
    for (int t0 = ra; t0 < rb; t0 += kBlk) {
        const int tn = min(kBlk, rb - t0);
        __syncthreads();
        if (static_cast<int>(threadIdx.x) < tn) {
            ha[threadIdx.x] = ca[t0 + threadIdx.x];
            hb[threadIdx.x] = cb[t0 + threadIdx.x];
        }
        __syncthreads();
        if (!valid) continue;
#pragma unroll 4
        for (int q = 0; q < tn; ++q) {
            const double2 v0 =
                *reinterpret_cast<const double2*>(V + static_cast<size_t>(t0 + q) * npad + e);
            ax = __builtin_fma(ha[q], v0.x, ax);
            ay = __builtin_fma(ha[q], v0.y, ay);
            bx = __builtin_fma(hb[q], v0.x, bx);
            by = __builtin_fma(hb[q], v0.y, by);
        }
    }
    if (valid) {
        *reinterpret_cast<double2*>(wa + static_cast<size_t>(s) * npad + e) = make_double2(ax, ay);
        *reinterpret_cast<double2*>(wb + static_cast<size_t>(s) * npad + e) = make_double2(bx, by);
    }
}

// DCGS2's second pass in one kernel where the vectors are long (npad of
// 2^17 or more: one block per 512 elements then fills the chip): each thread
// sums a_r V[r][e] and b_r V[r][e] over every row r < nrows (the rows loaded
// 8 at a time, the coefficients staged in LDS) and forms q and w' as
// k_combine2 does, without split partials.
__global__ __launch_bounds__(kBlk) void k_update_combine2(const double* __restrict__ V,
                                                          size_t npad,
                                                          const double* __restrict__ ca,
                                                          const double* __restrict__ cb, int nrows,
                                                          const GmresState* __restrict__ st,
                                                          double* __restrict__ u,
                                                          double* __restrict__ w) {
    __shared__ double ha[kBlk], hb[kBlk];
    if (st->done) return;
    const size_t e = (static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x) * 2;
    const bool valid = e < npad;
    double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
    for (int t0 = 0; t0 < nrows; t0 += kBlk) {
        const int tn = min(kBlk, nrows - t0);
        __syncthreads();
        if (static_cast<int>(threadIdx.x) < tn) {
            ha[threadIdx.x] = ca[t0 + threadIdx.x];
            hb[threadIdx.x] = cb[t0 + threadIdx.x];
        }
        __syncthreads();
        if (!valid) continue;
        for (int q0 = 0; q0 < tn; q0 += 8) {
            double2 vv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                vv[q] = *reinterpret_cast<const double2*>(
                    V + static_cast<size_t>(t0 + min(q0 + q, tn - 1)) * npad + e);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (q0 + q < tn) {
                    ax = __builtin_fma(ha[q0 + q], vv[q].x, ax);
                    ay = __builtin_fma(ha[q0 + q], vv[q].y, ay);
                    bx = __builtin_fma(hb[q0 + q], vv[q].x, bx);
                    by = __builtin_fma(hb[q0 + q], vv[q].y, by);
                }
            }
        }
    }
    if (!valid) return;
    const double rinv = st->rinv, gk = st->gk, sc = st->rinv * st->inv_snormw;
    const double2 uv = *reinterpret_cast<const double2*>(u + e);
    const double2 wv = *reinterpret_cast<const double2*>(w + e);
    const double qx = (uv.x - ax) * rinv, qy = (uv.y - ay) * rinv;
    *reinterpret_cast<double2*>(u + e) = make_double2(qx, qy);
    *reinterpret_cast<double2*>(w + e) =
        make_double2((wv.x - bx - qx * gk) * sc, (wv.y - by - qy * gk) * sc);
}

// w += sum_s wpart[s] (the final solution update)
__global__ __launch_bounds__(kBlk) void k_combine(double* __restrict__ w, size_t npad,
                                                  const double* __restrict__ wpart, int nsplit) {
    const size_t e = (static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x) * 2;
    if (e >= npad) return;
    double2 wv = *reinterpret_cast<const double2*>(w + e);
    // the split partials in order, loaded 8 at a time
    double sx = 0.0, sy = 0.0;
    for (int s0 = 0; s0 < nsplit; s0 += 8) {
        double2 t[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const double2 p = *reinterpret_cast<const double2*>(
                wpart + static_cast<size_t>(min(s0 + q, nsplit - 1)) * npad + e);
            t[q] = s0 + q < nsplit ? p : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            sx += t[q].x;
            sy += t[q].y;
        }
    }
    wv.x += sx;
    wv.y += sy;
    *reinterpret_cast<double2*>(w + e) = wv;
}

// DCGS2's second pass, final step: with sa = sum_s wa[s], sb = sum_s wb[s],
//   q = (u - sa) / r  into u's row,  w' = (w - sb - q g_j) / r / ||w'|| into w's row.
__global__ __launch_bounds__(kBlk) void k_combine2(double* __restrict__ u, double* __restrict__ w,
                                                   size_t npad, const double* __restrict__ wa,
                                                   const double* __restrict__ wb, int nsplit,
                                                   const GmresState* __restrict__ st) {
    if (st->done) return;
    const size_t e = (static_cast<size_t>(blockIdx.x) * kBlk + threadIdx.x) * 2;
    if (e >= npad) return;
    double ax = 0.0, ay = 0.0, bx = 0.0, by = 0.0;
    for (int s = 0; s < nsplit; ++s) {
        const double2 p = *reinterpret_cast<const double2*>(wa + static_cast<size_t>(s) * npad + e);
        const double2 t = *reinterpret_cast<const double2*>(wb + static_cast<size_t>(s) * npad + e);
        ax += p.x;
        ay += p.y;
        bx += t.x;
        by += t.y;
    }
    const double rinv = st->rinv, gk = st->gk, sc = st->rinv * st->inv_snormw;
    const double2 uv = *reinterpret_cast<const double2*>(u + e);
    const double2 wv = *reinterpret_cast<const double2*>(w + e);
    const double qx = (uv.x - ax) * rinv, qy = (uv.y - ay) * rinv;
    *reinterpret_cast<double2*>(u + e) = make_double2(qx, qy);
    *reinterpret_cast<double2*>(w + e) =
        make_double2((wv.x - bx - qx * gk) * sc, (wv.y - by - qy * gk) * sc);
}

// Givens update of Hessenberg column k = st->k (DHEQR, IJOB>1) and the stop
// test: hsum[0..k] the column, sub its subdiagonal entry.  The k stored
// rotations act on (col[i], col[i+1]) in order; the value carried into row i
// obeys carry_{i+1} = s_i carry_i + c_i h_{i+1}, an affine recurrence: each
// lane of the block's first wave composes the maps of its segment of
// ceil(k / 64) rotations, a shuffle scan over the wave composes the
// segments, and each lane then applies its segment from the carry entering
// it.  Every thread of the (kGivT-thread) block calls it.
constexpr int kGivT = 1024;
__device__ void givens_column(double* __restrict__ H, int ldh, double* __restrict__ q,
                              const double* __restrict__ hsum, double sub,
                              GmresState* __restrict__ st, double tol, int maxl) {
    const int k = st->k;
    double* __restrict__ col = H + static_cast<size_t>(k) * ldh;
    const int tid = threadIdx.x;
    if (tid < 64) {
        // segment of rotations for this lane
        const int L = (k + 63) / 64;
        const int i0 = min(k, tid * L), i1 = min(k, i0 + L);
        double A = 1.0, B = 0.0;
        for (int i = i0; i < i1; ++i) {
            const double c = q[2 * i], s = q[2 * i + 1];
            A = s * A;
            B = __builtin_fma(s, B, c * hsum[i + 1]);
        }
        // inclusive scan of the maps x -> A x + B (composition: later after
        // earlier), then the composite of the earlier lanes applied to h_0
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double a2 = __shfl_up(A, off), b2 = __shfl_up(B, off);
            if (tid >= off) {
                B = __builtin_fma(A, b2, B);
                A = A * a2;
            }
        }
        const double Ap = __shfl_up(A, 1), Bp = __shfl_up(B, 1);
        const double h0 = hsum[0];
        double carry = tid == 0 ? h0 : __builtin_fma(Ap, h0, Bp);
        for (int i = i0; i < i1; ++i) {
            const double c = q[2 * i], s = q[2 * i + 1];
            const double hn = hsum[i + 1];
            col[i] = c * carry - s * hn;
            carry = __builtin_fma(s, carry, c * hn);
        }
        // the thread owning the last rotation (or thread 0 when k == 0) finishes
        const bool last = (k == 0) ? (tid == 0) : (i0 < i1 && i1 == k);
        if (last) {
            const double t1 = carry, t2 = sub;
            double c, s;
            if (t2 == 0.0) {
                c = 1.0;
                s = 0.0;
            } else if (fabs(t2) >= fabs(t1)) {
                const double t = t1 / t2;
                s = -1.0 / sqrt(1.0 + t * t);
                c = -s * t;
            } else {
                const double t = t2 / t1;
                c = 1.0 / sqrt(1.0 + t * t);
                s = -c * t;
            }
            q[2 * k] = c;
            q[2 * k + 1] = s;
            const double rkk = c * t1 - s * t2;
            col[k] = rkk;
            col[k + 1] = t2;
            st->k = k + 1;
            if (rkk == 0.0) {  // DHEQR INFO = k: breakdown
                st->done = 3;
            } else {
                st->prod = st->prod * s;
                st->rho = fabs(st->prod * st->r0nrm);
                if (st->rho / st->bnrm <= tol)
                    st->done = 1;
                else if (k + 1 >= maxl)
                    st->done = 2;
            }
        }
    }
    __syncthreads();
}

// DCGS2's step between its passes, one block: the dots from k_dots2's
// chunk partials (a, alpha into ca; b, beta, omega into cb), r, the completed
// Arnoldi column j-1 (Givens, stop test), then unless `last` the provisional
// column j (into H_raw), ||w'||, and the second pass's coefficients (ca = a,
// cb = g) and scalars (st->rinv, st->gk, st->inv_snormw).
__global__ __launch_bounds__(kGivT) void k_dcgs(const double* __restrict__ hu,
                                                const double* __restrict__ hw, int nchunk, int j,
                                                int last, double* __restrict__ Hraw,
                                                double* __restrict__ H, int ldh,
                                                double* __restrict__ q, double* __restrict__ ca,
                                                double* __restrict__ cb,
                                                GmresState* __restrict__ st, double tol, int maxl) {
    constexpr int kWaves = kGivT / 64;
    __shared__ double sA[kGivT];
    __shared__ double sP[kWaves][64];
    __shared__ double sC[kGivT];  // the summed dots, u's rows then w's
    // H a for j <= kFastHa: tiles of 64 rows, each column range in one pass
    constexpr int kHaTiles = 3, kFastHa = 8 * kWaves;
    static_assert(64 * kHaTiles > kFastHa, "rows 0..kFastHa in kHaTiles tiles");
    __shared__ double sPT[kHaTiles][kWaves][64];
    if (st->done) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // the dots pass's rows: 0..j (u's) and, unless last, 0..j+1 (w's); each
    // sum over the chunks in chunk order, loaded 16 at a time (indices
    // clamped, values past nchunk replaced by 0) so that it waits on L2 once
    // per 16; thread t sums u's row t, then w's rows follow.  Up to kGivT
    // sums the block reads them back from LDS (sC), beyond that from ca, cb.
    const int nrows = last ? j + 1 : j + 2;
    const int nsum = (j + 1) + (last ? 0 : j + 2);
    const bool staged = nsum <= kGivT;
    for (int t = tid; t < nsum; t += kGivT) {
        const bool wrow = t > j;
        const int r = wrow ? t - (j + 1) : t;
        const double* __restrict__ hp = (wrow ? hw : hu) + r;
        double v = 0.0;
        for (int c0 = 0; c0 < nchunk; c0 += 16) {
            double tv[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const double x = hp[static_cast<size_t>(min(c0 + e, nchunk - 1)) * nrows];
                tv[e] = c0 + e < nchunk ? x : 0.0;
            }
#pragma unroll
            for (int e = 0; e < 16; ++e) v += tv[e];
        }
        (wrow ? cb : ca)[r] = v;
        if (staged) sC[t] = v;
    }
    __syncthreads();
    auto a_at = [&](int r) { return staged ? sC[r] : ca[r]; };
    auto b_at = [&](int r) { return staged ? sC[j + 1 + r] : cb[r]; };
    double paa = 0.0, pab = 0.0;
    for (int r = tid; r < j; r += kGivT) {
        const double ar = a_at(r);
        paa = __builtin_fma(ar, ar, paa);
        if (!last) pab = __builtin_fma(ar, b_at(r), pab);
    }
    // a.a and a.b: one butterfly per wave for both, the waves' pairs through
    // LDS in wave order
    for (int off = 32; off > 0; off >>= 1) {
        paa += __shfl_xor(paa, off);
        pab += __shfl_xor(pab, off);
    }
    if (lane == 0) {
        sP[0][wave] = paa;
        sP[1][wave] = pab;
    }
    __syncthreads();
    double saa = 0.0, sab = 0.0;
    for (int w = 0; w < kWaves; ++w) {
        saa += sP[0][w];
        sab += sP[1][w];
    }
    __syncthreads();
    const double rr2 = a_at(j) - saa;
    // the last iteration left w' = 0 (its Krylov space is invariant): column
    // j-1 completes with a zero subdiagonal, and its rotation ends the cycle
    const bool invariant = j >= 1 && st->snormw == 0.0;
    if (!(rr2 > 0.0) && !invariant) {  // u lies in the span of Q: breakdown
        if (tid == 0) st->done = 3;
        return;
    }
    const double rn = rr2 > 0.0 ? sqrt(rr2) : 0.0;
    if (j >= 1) {
        // column j-1: its last vector was u, now r q_j + Q a
        double* colr = Hraw + static_cast<size_t>(j - 1) * ldh;
        const double est = st->snormw;
        for (int i = tid; i < j; i += kGivT) colr[i] = __builtin_fma(est, a_at(i), colr[i]);
        if (tid == 0) colr[j] = est * rn;
        __syncthreads();
        givens_column(H, ldh, q, colr, est * rn, st, tol, maxl);
        if (st->done) return;  // block-uniform after givens_column's barrier
    }
    if (last) return;
    // provisional column j: (g - H a) / r, H upper Hessenberg (H[i][c], c >=
    // i - 1).  Rows in tiles of 64 (a lane per row); in each tile wave w takes
    // the columns w, w + 16, ... of the tile's range, 8 loads at a time, each
    // a coalesced column segment, and the waves' partials are added in wave
    // order.
    const double gj = (b_at(j) - sab) / rn;
    double pgg = 0.0;
    double* colj = Hraw + static_cast<size_t>(j) * ldh;
    if (j <= kFastHa) {
        // every tile's columns in one load per (wave, t): all tiles' loads
        // issued before any product, one barrier for all tiles; each row's
        // products and sums as in the loop below, and wave w finishes tile w
        // (g.g summed per tile, not per lane across tiles)
#pragma unroll
        for (int tl = 0; tl < kHaTiles; ++tl) {
            const int i0 = 64 * tl, i = i0 + lane, c0 = max(0, i0 - 1) + wave;
            double hv[8], av[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int cc = max(0, min(c0 + kWaves * t, j - 1));
                const bool use = c0 + kWaves * t < j && cc >= i - 1 && i <= j;
                const double h = Hraw[static_cast<size_t>(cc) * ldh + min(i, j)];
                hv[t] = use ? h : 0.0;
                av[t] = use ? a_at(cc) : 0.0;
            }
            double ha = 0.0;
#pragma unroll
            for (int t = 0; t < 8; ++t) ha = __builtin_fma(hv[t], av[t], ha);
            sPT[tl][wave][lane] = ha;
        }
        __syncthreads();
        const int i = 64 * wave + lane;
        if (wave < kHaTiles && i <= j) {
            double h = 0.0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) h += sPT[wave][w][lane];
            const double gi = i < j ? b_at(i) : gj;
            colj[i] = (gi - h) / rn;
            pgg = __builtin_fma(gi, gi, pgg);
        }
        __syncthreads();
    }
    for (int i0 = 0; i0 <= j && j > kFastHa; i0 += 64) {  // uniform over the block
        const int i = i0 + lane;
        const int cbeg = max(0, i0 - 1);
        double ha = 0.0;
        for (int c0 = cbeg + wave; c0 < j; c0 += 8 * kWaves) {
            double hv[8], av[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                // entries below the subdiagonal are never written: the
                // loaded value itself is dropped there, not multiplied by 0
                const int cc = min(c0 + kWaves * t, j - 1);
                const bool use = c0 + kWaves * t < j && cc >= i - 1 && i <= j;
                const double h = Hraw[static_cast<size_t>(cc) * ldh + min(i, j)];
                hv[t] = use ? h : 0.0;
                av[t] = use ? a_at(cc) : 0.0;
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) ha = __builtin_fma(hv[t], av[t], ha);
        }
        sP[wave][lane] = ha;
        __syncthreads();
        if (wave == 0 && i <= j) {
            double h = 0.0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) h += sP[w][lane];
            const double gi = i < j ? b_at(i) : gj;
            colj[i] = (gi - h) / rn;
            pgg = __builtin_fma(gi, gi, pgg);
        }
        __syncthreads();
    }
    const double sgg = block_sum(pgg, sA);
    if (tid == 0) {
        const double ww = b_at(j + 1) - sgg;
        const double est = ww > 0.0 ? sqrt(ww) / rn : 0.0;
        st->snormw = est;
        st->inv_snormw = est > 0.0 ? 1.0 / est : 0.0;
        st->rinv = 1.0 / rn;
        st->gk = gj;
    }
}

// DHELS: rotate (r0nrm, 0, ..., 0), back-substitute the k x k triangle.
__global__ __launch_bounds__(1024) void k_hels(const double* __restrict__ H, int ldh,
                                               const double* __restrict__ q,
                                               const GmresState* __restrict__ st,
                                               double* __restrict__ y) {
    const int k = st->k;
    if (threadIdx.x == 0) {
        double b = st->r0nrm;
        for (int i = 0; i < k; ++i) {  // (b_i, b_{i+1}) = rot_i (b_i, 0)
            const double c = q[2 * i], s = q[2 * i + 1];
            y[i] = c * b;
            b = s * b;
        }
    }
    __syncthreads();
    for (int kk = k - 1; kk >= 0; --kk) {
        const double* colk = H + static_cast<size_t>(kk) * ldh;
        if (threadIdx.x == 0) y[kk] = y[kk] / colk[kk];
        __syncthreads();
        const double yk = y[kk];
        for (int i = threadIdx.x; i < kk; i += blockDim.x) y[i] = __builtin_fma(-yk, colk[i], y[i]);
        __syncthreads();
    }
}

struct Plan {
    int nrows, nchunk, bx;
    size_t chunk;
};

Plan dots_plan(int nrows, size_t npad) {
    Plan p{};
    p.nrows = nrows;
    p.bx = cdiv_int(nrows, kRows);
    int want = std::max(1, cdiv_int(1024, p.bx));
    int maxc = static_cast<int>(std::max<size_t>(1, npad / (2 * kBlk * 4)));
    // a multiple of kXcds (k_dots2's block mapping); trailing chunks past
    // npad, if any, are empty and their partials 0
    p.nchunk = static_cast<int>(round_up(std::min(std::min(want, maxc), 64), kXcds));
    p.chunk = round_up(cdiv_int(static_cast<int>(npad), p.nchunk), 2 * kBlk);
    return p;
}

}  // namespace

// One DCGS2 step for iteration j: the dot pass, the step between the passes
// and (unless `last`) the second pass.  last: no new column -- only column
// j-1 is completed (after the final iteration of a cycle, V[j] = u).
static void dcgs_step(Context& c, int j, bool last, int ldh, double tol, int maxl) {
    const size_t npad = static_cast<size_t>(c.npad);
    hipStream_t s = c.stream;
    double* u = c.V.p + static_cast<size_t>(j) * npad;
    double* w = last ? u : c.V.p + static_cast<size_t>(j + 1) * npad;
    const int nrows = last ? j + 1 : j + 2;
    Plan p = dots_plan(nrows, npad);
    k_dots2<<<p.bx * p.nchunk, kBlk, 0, s>>>(c.V.p, npad, u, w, nrows, p.chunk, p.bx,
                                                  c.st.p, c.hpart.p, c.hpart2.p);
    STOKES_LAUNCH_CHECK();
    k_dcgs<<<1, kGivT, 0, s>>>(c.hpart.p, c.hpart2.p, p.nchunk, j, last ? 1 : 0, c.Hraw.p, c.H.p,
                               ldh, c.q.p, c.hcur.p, c.hbuf.p, c.st.p, tol, maxl);
    STOKES_LAUNCH_CHECK();
    if (last) return;
    const int ebx = static_cast<int>(npad / (2 * kBlk));
    if (ebx >= kFusedUpdateBlocks) {
        k_update_combine2<<<ebx, kBlk, 0, s>>>(c.V.p, npad, c.hcur.p, c.hbuf.p, j, c.st.p, u, w);
        STOKES_LAUNCH_CHECK();
        return;
    }
    int nsplit = 0;
    const size_t half = c.wpart.cap / 2;
    if (j > 0) {
        nsplit = std::min(j, std::max(1, cdiv_int(1024, ebx)));
        const int rps = cdiv_int(j, nsplit);
        nsplit = cdiv_int(j, rps);
        k_update2<<<dim3(ebx, nsplit), kBlk, 0, s>>>(c.V.p, npad, c.hcur.p, c.hbuf.p, j, rps,
                                                     c.st.p, c.wpart.p, c.wpart.p + half);
        STOKES_LAUNCH_CHECK();
    }
    k_combine2<<<ebx, kBlk, 0, s>>>(u, w, npad, c.wpart.p, c.wpart.p + half, nsplit, c.st.p);
    STOKES_LAUNCH_CHECK();
}

// Everything a captured chunk of iterations bakes in besides its key: the
// allocation epoch (device pointers), sizes, the preconditioner's shape, the
// far-field plans and the near-pair launch shapes.
static std::vector<long long> chunk_signature(const Context& c, int ibary, bool pc, int ldh,
                                              double tol_cycle) {
    const Geometry& g = c.g;
    const NearList& L = c.near;
    long long tol_bits = 0;
    static_assert(sizeof(tol_bits) == sizeof(tol_cycle), "double is 64-bit");
    std::memcpy(&tol_bits, &tol_cycle, sizeof(tol_bits));
    return {static_cast<long long>(g_alloc_epoch.load(std::memory_order_relaxed)),
            g.n, g.m, g.no, g.npts, g.nunk, c.npad, ibary, pc ? 1 : 0, c.pc_nfb, c.pc_nfw,
            farfield_choice(c, false), farfield_choice(c, true), L.nentries, L.nactive,
            L.wall_active, ldh, tol_bits};
}

// Launch one chunk of iterations: replay its graph if one was captured under
// the same signature; else, the first time the chunk comes, launch `run`
// directly, and the second time capture its launches, then replay.
template <typename Run>
static void replay_chunk(Context& c, const std::array<long long, 4>& key,
                         const std::vector<long long>& sig, Run&& run) {
    GraphCache& gc = c.graphs;
    if (sig != gc.sig) {
        gc.clear();
        gc.sig = sig;
    }
    auto it = gc.exec.find(key);
    if (it == gc.exec.end() && gc.seen.insert(key).second) {
        run();
        return;
    }
    if (it == gc.exec.end()) {
        hipGraph_t graph = nullptr;
        STOKES_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        try {
            run();
        } catch (...) {
            // leave the stream out of capture mode before reporting the error
            (void)hipStreamEndCapture(c.stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        STOKES_HIP(hipStreamEndCapture(c.stream, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        STOKES_HIP(e);
        it = gc.exec.emplace(key, exec).first;
        ++gc.captures;
    }
    STOKES_HIP(hipGraphLaunch(it->second, c.stream));
    ++gc.replays;
}

long long gmres_solve(Context& c, int ibary, int ibc, int maxl_in, double tol, double* den_io,
                      bool on_device) {
    auto t_start = std::chrono::steady_clock::now();
    Geometry& g = c.g;
    hipStream_t s = c.stream;
    const int N = g.nunk;
    const size_t npad = round_up(static_cast<size_t>(N), 2 * kBlk);
    c.npad = static_cast<int>(npad);
    int maxl = maxl_in == 0 ? 10 : maxl_in;
    maxl = std::max(1, std::min(maxl, N));
    const int ldh = maxl + 1;
    c.V.ensure(static_cast<size_t>(maxl + 1) * npad);
    c.H.ensure(static_cast<size_t>(ldh) * static_cast<size_t>(maxl));
    c.Hraw.ensure(static_cast<size_t>(ldh) * static_cast<size_t>(maxl));
    c.q.ensure(2 * static_cast<size_t>(maxl));
    c.hbuf.ensure(static_cast<size_t>(maxl + 2 + kRows));
    c.ybuf.ensure(static_cast<size_t>(maxl + 1 + kRows));
    {
        // dot partials of up to maxl + 2 rows, at most 64 chunks each
        const size_t rows = static_cast<size_t>(cdiv_int(maxl + 2, kRows) * kRows);
        c.hpart.ensure(rows * 64);
        c.hpart2.ensure(rows * 64);
    }
    const int ebx = static_cast<int>(npad / (2 * kBlk));
    c.npart.ensure(static_cast<size_t>(std::max(ebx, 1024)));
    // two halves of split partials (DCGS2's two coefficient sets)
    c.wpart.ensure(2 * static_cast<size_t>(std::max(1, cdiv_int(1024, ebx)) + 1) * npad);
    c.rhs.ensure(npad);
    c.xvec.ensure(npad);
    c.rvec.ensure(npad);
    c.st.ensure(1);
    // the operator writes rows [0, N): zero the padding of every Krylov row once
    if (npad > static_cast<size_t>(N)) {
        k_zero_tails<<<dim3(cdiv_int(static_cast<int>(npad) - N, kBlk), maxl + 1), kBlk, 0, s>>>(
            c.V.p, N, static_cast<int>(npad));
        STOKES_LAUNCH_CHECK();
    }
    if (!c.h_st) STOKES_HIP(hipHostMalloc(reinterpret_cast<void**>(&c.h_st), sizeof(GmresState)));
    c.hcur.ensure(static_cast<size_t>(maxl + 2 + kRows));
    DevBuf<double>& bpart = c.bpart;
    bpart.ensure(1024);

    // rhs and initial guess (Julia passes zeros, DensityStress.jl:45)
    STOKES_HIP(hipMemsetAsync(c.rhs.p, 0, npad * sizeof(double), s));
    STOKES_HIP(hipMemsetAsync(c.xvec.p, 0, npad * sizeof(double), s));
    bd_condition(c, ibc, c.rhs.p);
    STOKES_HIP(hipMemcpyAsync(c.xvec.p, den_io, sizeof(double) * N,
                              on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));

    // right preconditioning: GMRES on A P^{-1}, x = x0 + P^{-1} V y
    const bool pc = c.pc_kind != 0;
    if (pc) {
        precond_setup(c);
        c.pc_z.ensure(npad);
        STOKES_HIP(hipMemsetAsync(c.pc_z.p, 0, npad * sizeof(double), s));
    }
    // a guess in the preconditioned variable: u0 as given, x0 = P^{-1} u0;
    // every cycle then adds its V y to u as it adds P^{-1} V y to x
    const bool track_u = c.guess_space == 1;
    c.u_valid = false;
    if (track_u) {
        c.uvec.ensure(npad);
        STOKES_HIP(hipMemcpyAsync(c.uvec.p, c.xvec.p, npad * sizeof(double),
                                  hipMemcpyDeviceToDevice, s));
        if (pc) precond_apply(c, c.uvec.p, c.xvec.p);
    }

    const int nb = static_cast<int>(std::min<size_t>(1024, npad / (2 * kBlk)));
    k_axpby_norm<<<nb, kBlk, 0, s>>>(1.0, c.rhs.p, 0.0, nullptr, nullptr, npad, bpart.p);
    STOKES_LAUNCH_CHECK();
    // The reference's iteration: one GMRES cycle stopped by the implicit
    // residual estimate.  With a preconditioner or the mixed-precision far
    // field the cycle aims at half the tolerance.  Every solve then forms the
    // true residual b - A x of the returned density with the compensated
    // operator and, should it still exceed the tolerance, GMRES restarts from
    // it aiming at half the tolerance (iterative refinement; at most
    // kMaxCycles cycles within maxl iterations in total).
    const bool accel = pc || c.precision != 0;
    // mixed precision switches to the FP32 far field only where that is
    // faster than the FP64 far field in use (the cluster far field can be
    // faster than the direct FP32 sum at 1e5 points); the FP32 far field's
    // set-up (source tiles, the stokeslet logarithm table) is made only when
    // it will run -- its tuning is kept per size class, so this is decided
    // without it after the first solve of a class
    bool fp32_pays = true;
    if (c.precision == 1 || c.use_graphs) prepare_operator(c, c.xvec.p, false);
    if (c.precision == 1) {
        fp32_pays = farfield_fp32_pays(c);
        if (fp32_pays) {
            prepare_operator(c, c.xvec.p, true);
            fp32_pays = farfield_fp32_pays(c);
        }
    }
    // set-up work an operator application would otherwise launch inside a
    // captured chunk (and then not repeat on a later replay)
    if (c.use_graphs && (c.precision == 2 || (c.precision == 1 && fp32_pays)))
        prepare_operator(c, c.xvec.p, true);
    const double tol_cycle = accel ? kPcTolFrac * tol : tol;
    double* V0 = c.V.p;
    long long total_iters = 0, fp32_iters = 0;
    double prev_true = 0.0;
    int status = 0, nreorth = 0, cycles = 0;
    double implicit_relres = 0.0, true_relres = 0.0;
    const bool mixed = c.precision != 0;
    static_assert(kMaxCycles <= kCycleRecord, "the solve record keeps every cycle");
    Context::CycleRecord rec;
    double bnrm_last = 1.0;
    for (int cycle = 0;; ++cycle) {
        const int maxl_cycle = maxl - static_cast<int>(total_iters);
        c.ff_fp32_now = false;  // residuals are formed with the FP64 operator
        if (cycle == 0) {
            // r0 = b - A x0 into V[0] (compensated: x0 may be a warm start)
            apply_operator(c, c.xvec.p, c.rvec.p, ibary, false, true);
            k_zero_tail<<<cdiv_int(static_cast<int>(npad) - N + 1, kBlk), kBlk, 0, s>>>(
                c.rvec.p, N, static_cast<int>(npad));
            STOKES_LAUNCH_CHECK();
            k_axpby_norm<<<nb, kBlk, 0, s>>>(1.0, c.rhs.p, -1.0, c.rvec.p, V0, npad, c.npart.p);
        } else {
            // the previous cycle left r = b - A x in rvec
            k_axpby_norm<<<nb, kBlk, 0, s>>>(1.0, c.rvec.p, 0.0, nullptr, V0, npad, c.npart.p);
        }
        STOKES_LAUNCH_CHECK();
        // a restarted cycle aims at half the tolerance in every variant
        const double tol_now = cycle == 0 ? tol_cycle : kPcTolFrac * tol;
        k_init_state<<<1, 64, 0, s>>>(bpart.p, nb, c.npart.p, nb, tol_now, c.st.p);
        STOKES_LAUNCH_CHECK();
        k_scale_row<<<std::min(1024, ebx * 2), kBlk, 0, s>>>(V0, npad, c.st.p, 1);
        STOKES_LAUNCH_CHECK();

        int k = 0;
        // implicit residual / ||b|| after the last chunk; a restarted cycle
        // starts from the true residual the previous cycle left
        double relres_now = cycle > 0 ? true_relres : 1.0;
        while (k < maxl_cycle) {
            // mixed precision decides per chunk: shorter chunks switch sooner
            const int chunk =
                std::min(mixed && fp32_pays ? kChunkIters / 2 : kChunkIters, maxl_cycle - k);
            // after a mixed cycle missed the tolerance, later cycles stay FP64
            c.ff_fp32_now = c.precision == 2 ||
                            (mixed && fp32_pays && cycle == 0 &&
                             relres_now <= c.mixed_switch * tol);
            if (c.ff_fp32_now) fp32_iters += chunk;
            // iterations k .. k+chunk-1 (each exits early on the device once
            // the state reports convergence)
            auto run_chunk = [&]() {
                // kernel timing samples one iteration in c.time_period
                struct Sampling {
                    Context& c;
                    const bool on;
                    explicit Sampling(Context& cc) : c(cc), on(cc.time_kernels) {}
                    void next() {
                        const bool now = on && c.time_tick++ % c.time_period == 0;
                        c.time_kernels = now;
                        tl_time_weight = now ? c.time_period : 1;
                    }
                    ~Sampling() {
                        c.time_kernels = on;
                        tl_time_weight = 1;
                    }
                } sampling(c);
                for (int j = 0; j < chunk; ++j) {
                    sampling.next();
                    const int kk = k + j;
                    double* vk = c.V.p + static_cast<size_t>(kk) * npad;
                    double* w = c.V.p + static_cast<size_t>(kk + 1) * npad;
                    // V[kk] is normalised (v_0 before the loop, later ones
                    // by the previous iteration's second pass)
                    if (pc) {
                        precond_apply_scaled(c, vk, c.pc_z.p, nullptr, w);
                        apply_operator(c, c.pc_z.p, w, ibary, true);
                    } else {
                        apply_operator(c, vk, w, ibary);
                    }
                    if (c.time_kernels) c.t_gs.start(s);
                    dcgs_step(c, kk, false, ldh, tol_now, maxl_cycle);
                    if (c.time_kernels) {
                        // Krylov rows streamed by the dot and update passes
                        c.t_gs.stop(s, 2.0 * static_cast<double>(kk + 1) *
                                           static_cast<double>(npad) * 8.0);
                    }
                }
            };
            if (c.use_graphs && !c.time_kernels)
                replay_chunk(c, {static_cast<long long>(k), chunk, c.ff_fp32_now ? 1 : 0,
                                 maxl_cycle},
                             chunk_signature(c, ibary, pc, ldh, tol_now), run_chunk);
            else
                run_chunk();
            k += chunk;
            STOKES_HIP(hipMemcpyAsync(c.h_st, c.st.p, sizeof(GmresState), hipMemcpyDeviceToHost, s));
            STOKES_HIP(hipStreamSynchronize(s));
            relres_now = c.h_st->rho / c.h_st->bnrm;
            if (c.h_st->done) {
                if (c.ff_fp32_now) fp32_iters -= chunk - (c.h_st->k - (k - chunk));
                break;
            }
        }
        if (!c.h_st->done) {
            // the cycle's last column is completed by one more dot pass
            dcgs_step(c, k, true, ldh, tol_now, maxl_cycle);
            STOKES_HIP(hipMemcpyAsync(c.h_st, c.st.p, sizeof(GmresState), hipMemcpyDeviceToHost, s));
            STOKES_HIP(hipStreamSynchronize(s));
        }
        c.ff_fp32_now = false;
        GmresState fin = *c.h_st;
        if (cycle == 0) c.last_r0_relres = fin.r0nrm / fin.bnrm;
        const int iters = fin.k;
        status = 0;
        if (fin.done == 3) {
            status = 2;  // breakdown: solution left at x0 (DPIGMR IFLAG=2)
        } else if (iters > 0 && (fin.done == 1 || fin.rho < fin.r0nrm)) {
            // x += V y (DHELS + the sum of DPIGMR label 200)
            k_hels<<<1, 1024, 0, s>>>(c.H.p, ldh, c.q.p, c.st.p, c.ybuf.p);
            STOKES_LAUNCH_CHECK();
            int nsplit = std::min(iters, std::max(1, cdiv_int(1024, ebx)));
            int rps = cdiv_int(iters, nsplit);
            nsplit = cdiv_int(iters, rps);
            k_update<<<dim3(ebx, nsplit), kBlk, 0, s>>>(c.V.p, npad, c.ybuf.p, iters, rps,
                                                        c.wpart.p);
            STOKES_LAUNCH_CHECK();
            if (pc) {
                // u = V y, then x += P^{-1} u
                STOKES_HIP(hipMemsetAsync(c.rvec.p, 0, npad * sizeof(double), s));
                k_combine<<<ebx, kBlk, 0, s>>>(c.rvec.p, npad, c.wpart.p, nsplit);
                STOKES_LAUNCH_CHECK();
                if (track_u) {
                    k_add<<<static_cast<int>(npad / kBlk), kBlk, 0, s>>>(c.uvec.p, c.rvec.p, npad);
                    STOKES_LAUNCH_CHECK();
                }
                precond_apply(c, c.rvec.p, c.pc_z.p);
                k_add<<<static_cast<int>(npad / kBlk), kBlk, 0, s>>>(c.xvec.p, c.pc_z.p, npad);
                STOKES_LAUNCH_CHECK();
            } else {
                k_combine<<<ebx, kBlk, 0, s>>>(c.xvec.p, npad, c.wpart.p, nsplit);
                STOKES_LAUNCH_CHECK();
            }
            status = fin.done == 1 ? 0 : 1;
        } else if (fin.done != 1) {
            status = fin.done == 2 ? 1 : 2;
        }
        // true residual of the returned density, left in rvec
        apply_operator(c, c.xvec.p, c.rvec.p, ibary, false, true);
        k_zero_tail<<<cdiv_int(static_cast<int>(npad) - N + 1, kBlk), kBlk, 0, s>>>(
            c.rvec.p, N, static_cast<int>(npad));
        STOKES_LAUNCH_CHECK();
        k_axpby_norm<<<nb, kBlk, 0, s>>>(1.0, c.rhs.p, -1.0, c.rvec.p, c.rvec.p, npad, c.npart.p);
        STOKES_LAUNCH_CHECK();
        std::vector<double> part(static_cast<size_t>(nb));
        STOKES_HIP(hipMemcpyAsync(part.data(), c.npart.p, sizeof(double) * nb,
                                  hipMemcpyDeviceToHost, s));
        STOKES_HIP(hipStreamSynchronize(s));
        double rr = 0.0;
        for (double v : part) rr += v;
        total_iters += iters;
        nreorth += fin.k;  // every vector had its (delayed) second pass
        implicit_relres = fin.rho / fin.bnrm;
        true_relres = std::sqrt(rr) / fin.bnrm;
        bnrm_last = fin.bnrm;
        ++cycles;
        rec.true_relres[cycle] = true_relres;
        rec.iters[cycle] = static_cast<double>(iters);
        // restart while the true residual misses the tolerance and the last
        // restart still reduced it
        const bool stalled = cycle > 0 && true_relres > 0.5 * prev_true;
        prev_true = true_relres;
        rec.stop = true_relres <= tol     ? 0
                   : status != 0          ? 4
                   : stalled              ? 1
                   : total_iters >= maxl  ? 3
                   : cycle + 1 >= kMaxCycles ? 2
                                          : -1;
        if (rec.stop >= 0) break;
    }
    // a solve that ends above the tolerance is counted where callers see it
    // (erosion_last_solve_info), whatever its status
    c.last_tol_miss = true_relres > tol ? 1 : 0;
    c.tol_misses += c.last_tol_miss;
    c.last_cycles = cycles;
    // a solve above the tolerance measures the FP64 floor of its density:
    // ||A d|| / ||b|| for a half-ulp perturbation d of x (one FP64
    // application to d, whose entries are tiny, so it is formed accurately)
    rec.floor_relres = 0.0;
    if (c.last_tol_miss) {
        c.pc_z.ensure(npad);
        k_half_ulp<<<static_cast<int>(npad / kBlk), kBlk, 0, s>>>(c.xvec.p, c.pc_z.p, N,
                                                                  static_cast<int>(npad));
        STOKES_LAUNCH_CHECK();
        apply_operator(c, c.pc_z.p, c.rvec.p, ibary);
        k_zero_tail<<<cdiv_int(static_cast<int>(npad) - N + 1, kBlk), kBlk, 0, s>>>(
            c.rvec.p, N, static_cast<int>(npad));
        STOKES_LAUNCH_CHECK();
        k_axpby_norm<<<nb, kBlk, 0, s>>>(1.0, c.rvec.p, 0.0, nullptr, nullptr, npad, c.npart.p);
        STOKES_LAUNCH_CHECK();
        std::vector<double> part(static_cast<size_t>(nb));
        STOKES_HIP(hipMemcpyAsync(part.data(), c.npart.p, sizeof(double) * nb,
                                  hipMemcpyDeviceToHost, s));
        STOKES_HIP(hipStreamSynchronize(s));
        double dd = 0.0;
        for (double v : part) dd += v;
        rec.floor_relres = std::sqrt(dd) / bnrm_last;
    }
    c.last_rec = rec;
    if (c.last_tol_miss) {
        c.miss_rec = rec;
        // rounding the density to FP64 alone leaves a residual of this size:
        // the miss is at the floor of an FP64 density for this boundary
        if (rec.floor_relres >= 0.5 * true_relres) ++c.tol_misses_at_floor;
    }
    STOKES_HIP(hipMemcpyAsync(den_io, c.xvec.p, sizeof(double) * N,
                              on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    if (track_u && !pc)  // P = I: u is the density
        STOKES_HIP(hipMemcpyAsync(c.uvec.p, c.xvec.p, npad * sizeof(double),
                                  hipMemcpyDeviceToDevice, s));
    STOKES_HIP(hipStreamSynchronize(s));
    c.u_valid = track_u;
    c.last_iters = total_iters;
    c.last_implicit_relres = implicit_relres;
    c.last_true_relres = true_relres;
    c.last_reorth = nreorth;
    c.last_fp32_iters = fp32_iters;
    c.status = status;
    c.last_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return total_iters;
}

}  // namespace stokes
