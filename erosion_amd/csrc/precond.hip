// Per-curve two-level right preconditioner for the completed double-layer
// operator.  The reference runs GMRES without a preconditioner (solveBIE,
// ErosionPackage/fortran_code/stokesSolver.f:490; its unused msolve_DLP,
// :2211-2302, treats only Fourier modes +-1 of a circle); this is an
// addition with no counterpart there, off unless a caller asks for it.
//
// For every curve c (the wall, each body) let B_c be the block of the
// operator that couples the curve's own unknowns: the trapezoid
// self-interaction, the curvature diagonal and the jump -eta/2, plus for a
// body its stokeslet/rotlet columns and constraint rows, for the wall the
// rank-one term.  With Q_c the orthonormal Fourier modes 0..K of each
// density component sampled at the curve's points (and, for a body, the
// three constraint unknowns),
//     P_c^{-1} r = Q_c (Q_c^T B_c Q_c)^{-1} Q_c^T r - 2 (r - Q_c Q_c^T r),
// the exact inverse on the coarse space and the inverse -2 of the jump on
// the rest (the smooth self-interaction damps high modes).  GMRES runs on
// A P^{-1} (right preconditioning): its residual is the residual of A x = b,
// so the stopping test and the returned density keep their meaning.
//
// Set-up per solve: k_pc_pairs (K-mode projections of the trapezoid
// self-interaction, O(n^2 K) per body), k_pc_coarse (coarse matrix and its
// Gauss-Jordan inverse in LDS, one block per curve).  The wall block depends
// only on the number of wall points and is kept between solves.  Apply:
// k_pc_apply (one block per body, the wall's projection in 256-point
// segments) and k_pc_wall (the wall's coarse solve and expansion).
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace stokes {

namespace {

constexpr int kBlk = 256;
constexpr int kApplyBlk = 256;
constexpr int kQC = 17;                        // basis functions per pair pass
constexpr int kMaxModes = 32;                  // K <= 32 per curve
constexpr int kMaxNF = 2 * kMaxModes + 1;      // functions per component
constexpr int kMaxNQ = 2 * kMaxNF + 3;         // coarse unknowns per curve

constexpr int kCoarseBlk = 1024;
static_assert(kMaxNQ < kApplyBlk, "the coarse apply needs a lane per coarse unknown");

inline int cdiv_int(int a, int b) { return (a + b - 1) / b; }

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// phi[f][i], f = 0 (constant), 2k-1 (cos k theta), 2k (sin k theta), orthonormal
// over the N equispaced parameter points theta_i = 2 pi i / N.
__global__ void k_pc_basis(int N, int K, double* __restrict__ phi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double c0 = 1.0 / sqrt(static_cast<double>(N));
    const double c1 = sqrt(2.0 / static_cast<double>(N));
    phi[i] = c0;
    for (int k = 1; k <= K; ++k) {
        // reduce k i mod N first so the angle is exact in [0, 2 pi)
        const long long ki = (static_cast<long long>(k) * i) % N;
        double sn, cs;
        sincos(kTwoPi * static_cast<double>(ki) / static_cast<double>(N), &sn, &cs);
        phi[static_cast<size_t>(2 * k - 1) * N + i] = c1 * cs;
        phi[static_cast<size_t>(2 * k) * N + i] = c1 * sn;
    }
}

struct CurveView {
    int len, beg, nf;
    const double* phi;
    size_t soff;
};

__device__ __forceinline__ CurveView curve_view(int c, int no, int n, int nfw, int nfb,
                                                const double* phi_w, const double* phi_b) {
    CurveView v;
    if (c == 0) {
        v.len = no;
        v.beg = 0;
        v.nf = nfw;
        v.phi = phi_w;
        v.soff = 0;
    } else {
        v.len = n;
        v.beg = no + (c - 1) * n;
        v.nf = nfb;
        v.phi = phi_b;
        v.soff = static_cast<size_t>(no) * 3 * nfw + static_cast<size_t>(c - 1) * n * 3 * nfb;
    }
    return v;
}

// S[(t nf + f) len + i] = sum_{j != i} w_ij r_a r_b phi_f(j) over the curve's
// own points, t = 0 (xx), 1 (xy), 2 (yy), w_ij = (r.n_j) speed_j (2/len) / rho^4:
// the trapezoid self-interaction of the curve applied to each basis function.
// grid (curves, ceil(len / 256), ceil(nf / kQC)); block 256, one target per
// lane, kQC basis functions per block.
__global__ __launch_bounds__(kBlk) void k_pc_pairs(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ mx,
    const double* __restrict__ my, const double* __restrict__ phi_w,
    const double* __restrict__ phi_b, int nfw, int nfb, int no, int n, int first_curve,
    double* __restrict__ S) {
    __shared__ double sx[kBlk], sy[kBlk], smx[kBlk], smy[kBlk];
    __shared__ double sph[kQC][kBlk];
    const int c = first_curve + static_cast<int>(blockIdx.x);
    const CurveView cv = curve_view(c, no, n, nfw, nfb, phi_w, phi_b);
    if (static_cast<int>(blockIdx.y) * kBlk >= cv.len) return;  // uniform per block
    const int tid = threadIdx.x;
    const int i = blockIdx.y * kBlk + tid;
    const bool valid = i < cv.len;
    // lanes past the curve evaluate a far-away dummy target and store nothing
    const double xi = valid ? x[cv.beg + i] : 1.0e3;
    const double yi = valid ? y[cv.beg + i] : 1.0e3;
    double* Sc = S + cv.soff;
    // blockIdx.z picks the block of kQC basis functions
    const int q0 = blockIdx.z * kQC;
    if (q0 >= cv.nf) return;  // uniform per block
    {
        double axx[kQC], axy[kQC], ayy[kQC];
#pragma unroll
        for (int q = 0; q < kQC; ++q) axx[q] = axy[q] = ayy[q] = 0.0;
        for (int j0 = 0; j0 < cv.len; j0 += kBlk) {
            __syncthreads();
            const int j = j0 + tid;
            if (j < cv.len) {
                sx[tid] = x[cv.beg + j];
                sy[tid] = y[cv.beg + j];
                smx[tid] = mx[cv.beg + j];
                smy[tid] = my[cv.beg + j];
            } else {
                sx[tid] = sy[tid] = smx[tid] = smy[tid] = 0.0;
            }
#pragma unroll
            for (int q = 0; q < kQC; ++q)
                sph[q][tid] = (j < cv.len && q0 + q < cv.nf)
                                  ? cv.phi[static_cast<size_t>(q0 + q) * cv.len + j]
                                  : 0.0;
            __syncthreads();
            const int jn = min(kBlk, cv.len - j0);
            for (int jj = 0; jj < jn; ++jj) {
                const double rx = xi - sx[jj], ry = yi - sy[jj];
                const bool self = (j0 + jj) == i;
                const double rho2 = self ? 1.0 : __builtin_fma(rx, rx, ry * ry);
                const double inv = rcp64(rho2);
                const double a = __builtin_fma(rx, smx[jj], ry * smy[jj]);
                const double w = self ? 0.0 : a * inv * inv;
                const double wxx = w * rx * rx, wxy = w * rx * ry, wyy = w * ry * ry;
#pragma unroll
                for (int q = 0; q < kQC; ++q) {
                    const double ph = sph[q][jj];
                    axx[q] = __builtin_fma(wxx, ph, axx[q]);
                    axy[q] = __builtin_fma(wxy, ph, axy[q]);
                    ayy[q] = __builtin_fma(wyy, ph, ayy[q]);
                }
            }
        }
        if (valid) {
#pragma unroll
            for (int q = 0; q < kQC; ++q) {
                if (q0 + q >= cv.nf) break;
                Sc[static_cast<size_t>(0 * cv.nf + q0 + q) * cv.len + i] = axx[q];
                Sc[static_cast<size_t>(1 * cv.nf + q0 + q) * cv.len + i] = axy[q];
                Sc[static_cast<size_t>(2 * cv.nf + q0 + q) * cv.len + i] = ayy[q];
            }
        }
    }
}

// Coarse matrix G = Q^T B Q of one curve, inverted in place in LDS by
// Gauss-Jordan elimination with partial (row) pivoting; writes G^{-1}
// column-major (transposed) to out + blockIdx.x nq^2.  Coarse unknown order: (x, f) for
// f < nf, then (y, f), then a body's stokeslet x, y and rotlet.
// MAXNQ bounds nq: the LDS matrix of the 16-mode instance (69^2) leaves room
// for several blocks per CU; the 32-mode one (133^2, 140 KB) for one.
template <int MAXNQ>
__global__ __launch_bounds__(kCoarseBlk) void k_pc_coarse(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ cur,
    const double* __restrict__ speed, const double* __restrict__ cx,
    const double* __restrict__ cy, const double* __restrict__ phi_w,
    const double* __restrict__ phi_b, int nfw, int nfb, int no, int n, int first_curve,
    const double* __restrict__ S, double* __restrict__ out, int* __restrict__ bad) {
    constexpr int MAXNF = (MAXNQ - 3) / 2;
    // per-point functions whose projections give a body's stokeslet/rotlet
    // columns (six) and constraint rows (three)
    constexpr int kNX = 9;
    constexpr int kCh = MAXNQ >= 100 ? 32 : 16;  // points per LDS chunk
    constexpr int kLdMax = MAXNF > kNX ? MAXNF : kNX;
    constexpr int kEPT = (3 * MAXNF * MAXNF + kNX * MAXNF + kCoarseBlk - 1) / kCoarseBlk;
    static_assert(kCh * MAXNF + 4 * kCh * kLdMax <= MAXNQ * MAXNQ,
                  "chunk buffers fit in G's memory");
    __shared__ double G[MAXNQ * MAXNQ];
    __shared__ double fcol[MAXNQ];
    __shared__ int perm[MAXNQ];
    __shared__ double uv[4 * MAXNF];
    const int c = first_curve + static_cast<int>(blockIdx.x);
    const bool wall = c == 0;
    const CurveView cv = curve_view(c, no, n, nfw, nfb, phi_w, phi_b);
    const int nf = cv.nf, len = cv.len;
    const int nc = 2 * nf;
    const int nq = wall ? nc : nc + 3;
    const double* Sc = S + cv.soff;
    const double* phi = cv.phi;
    const double dlen = static_cast<double>(len);
    const double cxb = wall ? 0.0 : cx[c - 1], cyb = wall ? 0.0 : cy[c - 1];
    const double k4pi = 0.5 / kTwoPi;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    constexpr int kCoarseWaves = kCoarseBlk / 64;

    // Wall rank-one completion n_i sum_j (n_j . eta_j) speed_j 2 pi / no in
    // the coarse space: U_a[p] = sum_i phi_p(i) n_a(i) and
    // V_b[q] = sum_j speed_j (2 pi / no) n_b(j) phi_q(j), uv = [U_x U_y V_x V_y].
    if (wall) {
        for (int d = wv; d < 4 * nf; d += kCoarseWaves) {
            const bool second = d >= 2 * nf;
            const int rem = second ? d - 2 * nf : d;
            const int a = rem / nf, p = rem - a * nf;
            const double* pp = phi + static_cast<size_t>(p) * len;
            double sum = 0.0;
            for (int i = lane; i < len; i += 64) {
                const double na = a == 0 ? px[i] : py[i];
                sum = second ? __builtin_fma(speed[i] * kTwoPi / dlen * na, pp[i], sum)
                             : __builtin_fma(pp[i], na, sum);
            }
            sum = wave_sum(sum);
            if (lane == 0) uv[d] = sum;
        }
    }

    // Coupling of the curve's own densities, for a, b in {x, y}:
    //   G_ab[p][q] = sum_i phi_p(i) (S_{a+b}[q][i] + L_ab(i) phi_q(i)),
    // L_xx = -f ny^2 - 1/2, L_xy = L_yx = f nx ny, L_yy = -f nx^2 - 1/2 with
    // f = curvature x speed / len.  Both depend on a + b only: three nf x nf
    // products over the curve's points, staged kCh points at a time in LDS
    // (in G's memory, unused until they are done) with each thread's entries
    // in registers.  A body's stokeslet/rotlet columns and constraint rows are
    // the projections sum_i phi_p(i) F_f(i) of nine functions of the point
    // (r = z - c the offset from the body's centre, w = speed / len):
    //   F_0 = (-log|r|^2 / 2 + rx^2 / |r|^2) / 4 pi, F_1 = F_3 = rx ry / (4 pi |r|^2),
    //   F_2 = ry / |r|^2 (stokeslet x, y and rotlet in the x rows),
    //   F_4 = (-log|r|^2 / 2 + ry^2 / |r|^2) / 4 pi, F_5 = -rx / |r|^2 (y rows),
    //   F_6 = -w, F_7 = -y w, F_8 = x w (the constraint rows),
    // staged as a fourth slab and summed in the same pass.
    const int ldm = max(nf, kNX);
    double* Pc = G;            // [kCh][nf]: phi_q at the chunk's points
    double* Mc = G + kCh * nf;  // [4][kCh][ldm]: S_k + L_k phi_q; F_f
    const int nn = nf * nf;
    const int ntot = 3 * nn + (wall ? 0 : kNX * nf);
    double acc[kEPT];
    int poff[kEPT], moff[kEPT];  // entry u: Pc column p, Mc offset k kCh ldm + q
#pragma unroll
    for (int u = 0; u < kEPT; ++u) {
        acc[u] = 0.0;
        const int e = min(tid + u * kCoarseBlk, ntot - 1);
        if (e < 3 * nn) {
            const int k = e / nn, rem = e - k * nn, p = rem / nf, q = rem - p * nf;
            poff[u] = p;
            moff[u] = k * kCh * ldm + q;
        } else {
            const int f = (e - 3 * nn) / nf, p = e - 3 * nn - f * nf;
            poff[u] = p;
            moff[u] = 3 * kCh * ldm + f;
        }
    }
    for (int i0 = 0; i0 < len; i0 += kCh) {
        __syncthreads();
        for (int t = tid; t < kCh * nf; t += kCoarseBlk) {
            const int q = t / kCh, i = t - q * kCh;
            const int pt = i0 + i;
            double ph = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0;
            if (pt < len) {
                const int gi = cv.beg + pt;
                const double f = cur[gi] * speed[gi] / dlen;
                const double nx = px[gi], ny = py[gi];
                ph = phi[static_cast<size_t>(q) * len + pt];
                m0 = __builtin_fma(-f * ny * ny - 0.5, ph, Sc[static_cast<size_t>(q) * len + pt]);
                m1 = __builtin_fma(f * nx * ny, ph, Sc[static_cast<size_t>(nf + q) * len + pt]);
                m2 = __builtin_fma(-f * nx * nx - 0.5, ph,
                                   Sc[static_cast<size_t>(2 * nf + q) * len + pt]);
            }
            Pc[i * nf + q] = ph;
            Mc[(0 * kCh + i) * ldm + q] = m0;
            Mc[(1 * kCh + i) * ldm + q] = m1;
            Mc[(2 * kCh + i) * ldm + q] = m2;
        }
        if (!wall) {
            for (int i = tid; i < kCh; i += kCoarseBlk) {
                const int pt = i0 + i;
                double fv[kNX];
#pragma unroll
                for (int f = 0; f < kNX; ++f) fv[f] = 0.0;
                if (pt < len) {
                    const int gi = cv.beg + pt;
                    const double qx = x[gi] - cxb, qy = y[gi] - cyb;
                    const double r2 = __builtin_fma(qx, qx, qy * qy);
                    const double ir2 = 1.0 / r2;
                    const double lg = log(r2);
                    const double w = speed[gi] / dlen;
                    fv[0] = k4pi * (-0.5 * lg + qx * qx * ir2);
                    fv[1] = k4pi * qx * qy * ir2;
                    fv[2] = qy * ir2;
                    fv[3] = fv[1];
                    fv[4] = k4pi * (-0.5 * lg + qy * qy * ir2);
                    fv[5] = -qx * ir2;
                    fv[6] = -w;
                    fv[7] = -y[gi] * w;
                    fv[8] = x[gi] * w;
                }
#pragma unroll
                for (int f = 0; f < kNX; ++f) Mc[(3 * kCh + i) * ldm + f] = fv[f];
            }
        }
        __syncthreads();
        // entries past ntot repeat the last one; they are never stored
        const int cnt = min(kCh, len - i0);
        for (int i = 0; i < cnt; ++i) {
            const double* pi = Pc + i * nf;
            const double* mi = Mc + i * ldm;
#pragma unroll
            for (int u = 0; u < kEPT; ++u) acc[u] = __builtin_fma(pi[poff[u]], mi[moff[u]], acc[u]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kEPT; ++u) {
        const int e = tid + u * kCoarseBlk;
        if (e < 3 * nn) {
            const int k = e / nn, rem = e - k * nn, p = rem / nf, q = rem - p * nf;
            for (int a = 0; a < 2; ++a) {
                const int b = k - a;
                if (b < 0 || b > 1) continue;
                double g = acc[u];
                if (wall) g = __builtin_fma(uv[a * nf + p], uv[2 * nf + b * nf + q], g);
                G[(a * nf + p) * nq + b * nf + q] = g;
            }
        } else if (e < ntot) {
            const int f = (e - 3 * nn) / nf, p = e - 3 * nn - f * nf;
            const double g = acc[u];
            if (f < 6) {  // stokeslet x, y, rotlet column cc of row (a, p)
                const int a = f / 3, cc = f - 3 * a;
                G[(a * nf + p) * nq + nc + cc] = g;
            } else if (f == 6) {  // -w: constraint row 0 on x, row 1 on y
                G[nc * nq + p] = g;
                G[nc * nq + nf + p] = 0.0;
                G[(nc + 1) * nq + p] = 0.0;
                G[(nc + 1) * nq + nf + p] = g;
            } else {  // -y w on x, x w on y: constraint row 2
                G[(nc + 2) * nq + (f - 7) * nf + p] = g;
            }
        }
    }
    if (!wall && tid < 9) {  // the constraint unknowns' identity block
        const int r = tid / 3, col = tid - 3 * r;
        G[(nc + r) * nq + nc + col] = r == col ? 1.0 : 0.0;
    }

    __syncthreads();

    // largest entry: a pivot below 1e-13 of it (or not finite) marks the
    // block singular, and the caller refuses the preconditioner
    __shared__ double gmax[kCoarseBlk / 64];
    __shared__ int singular;
    {
        double mx = 0.0;
        for (int e = tid; e < nq * nq; e += kCoarseBlk) mx = fmax(mx, fabs(G[e]));
        for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
        if (lane == 0) gmax[wv] = mx;
        if (tid == 0) singular = 0;
        __syncthreads();
    }
    double scale = 0.0;
    for (int w2 = 0; w2 < kCoarseWaves; ++w2) scale = fmax(scale, gmax[w2]);

    for (int k = 0; k < nq; ++k) {
        if (tid < 64) {
            double best = -1.0;
            int bi = k;
            for (int r = k + tid; r < nq; r += 64) {
                const double v = fabs(G[r * nq + k]);
                if (v > best) {
                    best = v;
                    bi = r;
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off);
                const int oi = __shfl_xor(bi, off);
                if (ob > best || (ob == best && oi < bi)) {
                    best = ob;
                    bi = oi;
                }
            }
            if (tid == 0) perm[k] = bi;
        }
        __syncthreads();
        const int p = perm[k];
        if (p != k) {
            for (int col = tid; col < nq; col += kCoarseBlk) {
                const double t = G[k * nq + col];
                G[k * nq + col] = G[p * nq + col];
                G[p * nq + col] = t;
            }
        }
        __syncthreads();
        double piv = G[k * nq + k];
        if (!(fabs(piv) > 1e-13 * scale) || !isfinite(piv)) {
            // block-uniform: every thread reads the same pivot; the
            // elimination goes on with a unit pivot so that no inf or NaN
            // arises, and the flag refuses the result
            if (tid == 0) singular = 1;
            piv = 1.0;
        }
        for (int r = tid; r < nq; r += kCoarseBlk) fcol[r] = r == k ? 0.0 : G[r * nq + k];
        __syncthreads();
        for (int col = tid; col < nq; col += kCoarseBlk)
            G[k * nq + col] = (col == k ? 1.0 : G[k * nq + col]) / piv;
        __syncthreads();
        // thread (tid / 32, tid % 32) walks rows and columns in steps of
        // 32: no division per entry (a runtime divide costs more than the
        // update it indexes)
        for (int r = tid >> 5; r < nq; r += kCoarseBlk >> 5) {
            if (r == k) continue;
            const double f = fcol[r];
            for (int col = tid & 31; col < nq; col += 32)
                G[r * nq + col] = (col == k ? 0.0 : G[r * nq + col]) - f * G[k * nq + col];
        }
        __syncthreads();
    }
    for (int k = nq - 1; k >= 0; --k) {
        const int p = perm[k];
        if (p != k) {
            for (int r = tid; r < nq; r += kCoarseBlk) {
                const double t = G[r * nq + k];
                G[r * nq + k] = G[r * nq + p];
                G[r * nq + p] = t;
            }
        }
        __syncthreads();
    }
    if (tid == 0) bad[blockIdx.x] = singular;
    double* o = out + static_cast<size_t>(blockIdx.x) * nq * nq;
    // transposed: k_pc_apply reads column k of row-lanes coalesced
    for (int e = tid; e < nq * nq; e += kCoarseBlk) {
        const int r = e / nq, col = e - r * nq;
        o[static_cast<size_t>(col) * nq + r] = G[e];
    }
}

// z = P^{-1} r.  A body is one block; the wall, whose points are many, is
// spread over nseg = ceil(no / kApplyBlk) blocks of kApplyBlk points each:
// k_pc_apply projects the segments (blocks 0..nseg-1, partial coefficients
// to wpart) and the bodies (blocks nseg..nseg+m-1, the whole apply), and
// k_pc_wall sums the wall's partial coefficients, solves its coarse system
// and expands each segment.
//
// Projection: a lane per (mode, slice of the points), the basis functions
// generated by rotation from the slice's first point (no table loads), four
// partials per lane meeting in LDS (pc_project).  The inverse coarse matrices
// are stored transposed, so d = G^{-1} c reads them coalesced with one row per
// lane.

// geometry arrays the apply's reductions read (point order)
struct SumGeom {
    const double *x, *y, *px, *py, *speed;
};

// part: the staged point values (rows 0, 1) and the slices' partials (rows
// 2-5) of the projection; the coarse solve's partials (row 0)
struct ApplyLds {
    double part[6][kApplyBlk];
    double red[3][kApplyBlk / 64];
    double cc[kMaxNQ], dd[kMaxNQ];
};

// cc[comp nf + f] = sum over points [i0, i1) of the curve of phi_f(i) r_comp(i)
// (block-wide; every thread calls it).
// With scale, r is first multiplied by sc in place (the Krylov vector's
// normalisation, left to this kernel by GMRES).
// The points' values are staged in LDS, kApplyBlk at a time.  Lane l takes
// mode k = l % (K + 1) over the (l / (K + 1))-th of nsub = kApplyBlk / (K + 1)
// contiguous slices of each staged pass: from e^{i k theta} at the slice's
// first point (the angle reduced mod 2 pi exactly) it steps by
// e^{i k 2 pi / len} to each next point, summing cos and sin times both
// components.  Only the slices' four partials per mode meet in LDS, and they
// are added in slice order.
__device__ __forceinline__ void pc_project(ApplyLds& L, int len, int nf, int i0, int i1,
                                           double* r, int xoff, int yoff, bool scale, double sc,
                                           double* cc) {
    const int tid = threadIdx.x;
    const int K = (nf - 1) / 2;
    const int nmode = K + 1;
    const int npt = i1 - i0;
    double* sx = L.part[0];
    double* sy = L.part[1];
    double* pp = L.part[2];  // [slice][mode][cos x, sin x, cos y, sin y]: rows 2-5
    const int nsub = kApplyBlk / nmode;
    const int k = tid % nmode, sl = tid / nmode;
    double cx = 0.0, sxs = 0.0, cy = 0.0, sys = 0.0;
    for (int p0 = 0; p0 < npt; p0 += kApplyBlk) {  // uniform over the block
        const int pn = min(kApplyBlk, npt - p0);
        __syncthreads();  // the previous pass's reads are done
        if (tid < pn) {
            const int i = i0 + p0 + tid;
            double rx = r[xoff + i], ry = r[yoff + i];
            if (scale) {
                rx *= sc;
                ry *= sc;
                r[xoff + i] = rx;
                r[yoff + i] = ry;
            }
            sx[tid] = rx;
            sy[tid] = ry;
        }
        __syncthreads();
        if (sl < nsub) {
            const int span = (pn + nsub - 1) / nsub;
            const int a = min(pn, sl * span), b = min(pn, a + span);
            if (a < b) {
                const long long ka = (static_cast<long long>(k) * (i0 + p0 + a)) % len;
                double sk, ck, s1, c1t;
                sincospi(2.0 * static_cast<double>(ka) / static_cast<double>(len), &sk, &ck);
                sincospi(2.0 * static_cast<double>(k) / static_cast<double>(len), &s1, &c1t);
                for (int i = a; i < b; ++i) {
                    const double vx = sx[i], vy = sy[i];
                    cx = __builtin_fma(ck, vx, cx);
                    sxs = __builtin_fma(sk, vx, sxs);
                    cy = __builtin_fma(ck, vy, cy);
                    sys = __builtin_fma(sk, vy, sys);
                    const double cn = __builtin_fma(ck, c1t, -sk * s1);
                    sk = __builtin_fma(sk, c1t, ck * s1);
                    ck = cn;
                }
            }
        }
    }
    // slices x modes <= kApplyBlk lanes, four partials each: rows 2-5
    if (sl < nsub) {
        double* o = pp + 4 * (sl * nmode + k);
        o[0] = cx;
        o[1] = sxs;
        o[2] = cy;
        o[3] = sys;
    }
    __syncthreads();
    if (tid < 4 * nmode) {
        const int kk = tid / 4, q = tid % 4;
        double v = 0.0;
        for (int j = 0; j < nsub; ++j) v += pp[4 * (j * nmode + kk) + q];
        const int comp = q / 2;
        const bool sn = (q & 1) != 0;
        if (kk == 0) {
            if (!sn) cc[comp * nf] = v / sqrt(static_cast<double>(len));
        } else {
            cc[comp * nf + (sn ? 2 * kk : 2 * kk - 1)] = v * sqrt(2.0 / static_cast<double>(len));
        }
    }
    __syncthreads();
}

// d = G^{-1} c with G^{-1} stored transposed (Gt[k nq + j]); dd = d + 2 c on
// the density coefficients, and the body's constraint unknowns d[nc..nc+2]
// go straight to zc.  The k range is split over kApplyBlk / nq groups of nq
// lanes, so each lane waits on a few L2 loads, not nq in sequence.
__device__ __forceinline__ void pc_coarse_solve(ApplyLds& L, const double* __restrict__ Gt,
                                                int nq, int nc, double* zc) {
    const int tid = threadIdx.x;
    const int groups = kApplyBlk / nq;  // nq <= kMaxNQ < kApplyBlk
    {
        const int span = (nq + groups - 1) / groups;
        const int j = tid % nq, g = tid / nq;
        if (g < groups) {
            const int k0 = g * span, k1 = min(nq, k0 + span);
            double d = 0.0;
            // batches of kBatch loads issued together (indices clamped into
            // the slice, coefficients past it read as 0), so the chain waits
            // on L2 once per batch; the sum keeps its order
            constexpr int kBatch = 24;
            for (int kb = k0; kb < k1; kb += kBatch) {
                double gv[kBatch], cv[kBatch];
#pragma unroll
                for (int kk = 0; kk < kBatch; ++kk) {
                    const int k = min(kb + kk, k1 - 1);
                    gv[kk] = Gt[static_cast<size_t>(k) * nq + j];
                    cv[kk] = kb + kk < k1 ? L.cc[k] : 0.0;
                }
#pragma unroll
                for (int kk = 0; kk < kBatch; ++kk) d = __builtin_fma(gv[kk], cv[kk], d);
            }
            L.part[0][g * nq + j] = d;
        }
    }
    __syncthreads();
    if (tid < nq) {
        double d = 0.0;
        for (int g = 0; g < groups; ++g) d += L.part[0][g * nq + tid];
        if (tid < nc)
            L.dd[tid] = d + 2.0 * L.cc[tid];
        else
            zc[tid - nc] = d;
    }
    __syncthreads();
}

// z = Q dd - 2 r at the curve's points [i0, i1); the same values also go to
// ex/ey at point index poff + i (the density in point order).
// SUMS 1 (wall): a0 += (px ex + py ey) speed, the rank-one integral's
// integrand; SUMS 2 (body): a0, a1, a2 -= (ex, ey, ex y - ey x) speed, the
// constraint rows' integrands (the terms k_reduce_small sums, same order per
// lane).  Point p = poff + i of the geometry arrays.
template <int SUMS>
__device__ __forceinline__ void pc_expand(const ApplyLds& L, int len, int nf, int i0, int i1,
                                          const double* __restrict__ r, int xoff, int yoff,
                                          double* __restrict__ z, int poff,
                                          double* __restrict__ ex, double* __restrict__ ey,
                                          const SumGeom& sg, double& a0, double& a1, double& a2) {
    const int K = (nf - 1) / 2;
    const double c0 = 1.0 / sqrt(static_cast<double>(len));
    const double c1 = sqrt(2.0 / static_cast<double>(len));
    const double* dd = L.dd;
    for (int i = i0 + static_cast<int>(threadIdx.x); i < i1; i += kApplyBlk) {
        double s1, c1t;
        sincospi(2.0 * static_cast<double>(i) / static_cast<double>(len), &s1, &c1t);
        double zx = __builtin_fma(c0, dd[0], -2.0 * r[xoff + i]);
        double zy = __builtin_fma(c0, dd[nf], -2.0 * r[yoff + i]);
        double ck = 1.0, sk = 0.0;
        for (int k = 1; k <= K; ++k) {
            const double cn = __builtin_fma(ck, c1t, -sk * s1);
            sk = __builtin_fma(sk, c1t, ck * s1);
            ck = cn;
            const double pc = c1 * ck, ps = c1 * sk;
            zx = __builtin_fma(pc, dd[2 * k - 1], __builtin_fma(ps, dd[2 * k], zx));
            zy = __builtin_fma(pc, dd[nf + 2 * k - 1], __builtin_fma(ps, dd[nf + 2 * k], zy));
        }
        z[xoff + i] = zx;
        z[yoff + i] = zy;
        ex[poff + i] = zx;
        ey[poff + i] = zy;
        const int p = poff + i;
        if (SUMS == 1) {
            a0 += (sg.px[p] * zx + sg.py[p] * zy) * sg.speed[p];
        } else if (SUMS == 2) {
            a0 -= zx * sg.speed[p];
            a1 -= zy * sg.speed[p];
            a2 -= (zx * sg.y[p] - zy * sg.x[p]) * sg.speed[p];
        }
    }
}

// Block sums of a0, a1, a2 as k_reduce_small forms them (butterflies within
// each wave, then the waves in order); the result is valid in thread 0.
__device__ __forceinline__ void pc_block_sum3(ApplyLds& L, double& a0, double& a1, double& a2) {
    for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_xor(a0, off);
        a1 += __shfl_xor(a1, off);
        a2 += __shfl_xor(a2, off);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        L.red[0][wave] = a0;
        L.red[1][wave] = a1;
        L.red[2][wave] = a2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kApplyBlk / 64; ++w) {
            a0 += L.red[0][w];
            a1 += L.red[1][w];
            a2 += L.red[2][w];
        }
    }
}

// grid nseg + m: wall segments, then bodies.  wpart[s kMaxNQ + j] receives
// segment s's partial wall coefficients.  With sst (a GMRES state that has not
// stopped), r is first scaled in place by sst->inv_snormw: the new Krylov
// vector's normalisation, folded into its first reader.
// With y (the operator's output A z), a body block also writes the body's
// three constraint rows of y (as k_reduce_small does, stokesSolver.f:1055-1096).
__global__ __launch_bounds__(kApplyBlk) void k_pc_apply(
    int nfw, int nfb, int no, int n, int m, int nseg, const double* __restrict__ Gb,
    const GmresState* __restrict__ sst, double* __restrict__ r, double* __restrict__ wpart,
    double* __restrict__ z, double* __restrict__ ex, double* __restrict__ ey, SumGeom sg,
    double* __restrict__ y) {
    __shared__ ApplyLds L;
    const int tid = threadIdx.x;
    const bool scale = sst != nullptr && !sst->done;
    const double sc = scale ? sst->inv_snormw : 1.0;
    if (static_cast<int>(blockIdx.x) < nseg) {
        const int i0 = blockIdx.x * kApplyBlk, i1 = min(no, i0 + kApplyBlk);
        pc_project(L, no, nfw, i0, i1, r, 0, no, scale, sc,
                   wpart + static_cast<size_t>(blockIdx.x) * kMaxNQ);
        return;
    }
    const int c = static_cast<int>(blockIdx.x) - nseg + 1;  // body c, curve index 1..m
    const int nf = nfb, nc = 2 * nf, nq = nc + 3;
    const int xoff = 2 * no + 2 * (c - 1) * n;
    const int yoff = xoff + n;
    const int coff = 2 * no + 2 * n * m + 3 * (c - 1);
    pc_project(L, n, nf, 0, n, r, xoff, yoff, scale, sc, L.cc);
    if (tid < 3) {
        double v = r[coff + tid];
        if (scale) {
            v *= sc;
            r[coff + tid] = v;
        }
        L.cc[nc + tid] = v;
    }
    __syncthreads();
    pc_coarse_solve(L, Gb + static_cast<size_t>(c - 1) * nq * nq, nq, nc, z + coff);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    if (y == nullptr) {
        pc_expand<0>(L, n, nf, 0, n, r, xoff, yoff, z, no + (c - 1) * n, ex, ey, sg, a0, a1, a2);
        return;
    }
    pc_expand<2>(L, n, nf, 0, n, r, xoff, yoff, z, no + (c - 1) * n, ex, ey, sg, a0, a1, a2);
    pc_block_sum3(L, a0, a1, a2);  // its barrier also orders z[coff..] before the reads below
    if (tid == 0) {
        const double rs = kTwoPi / static_cast<double>(n);
        y[coff + 0] = a0 / kTwoPi * rs + z[coff + 0];
        y[coff + 1] = a1 / kTwoPi * rs + z[coff + 1];
        y[coff + 2] = a2 / kTwoPi * rs + z[coff + 2];
    }
}

// The wall after k_pc_apply: grid nseg, block s expands segment s.  With
// scal, block s also writes scal[s], its segment's part of the unscaled
// rank-one integral (stokesSolver.f:1107-1111; k_assemble adds the parts).
__global__ __launch_bounds__(kApplyBlk) void k_pc_wall(int nfw, int no, int nseg,
                                                       const double* __restrict__ Gw,
                                                       const double* __restrict__ wpart,
                                                       const double* __restrict__ r,
                                                       double* __restrict__ z,
                                                       double* __restrict__ ex,
                                                       double* __restrict__ ey, SumGeom sg,
                                                       double* __restrict__ scal) {
    __shared__ ApplyLds L;
    const int tid = threadIdx.x;
    const int nc = 2 * nfw;
    if (tid < nc) {
        // the segments' partials in segment order, loaded 16 at a time
        // (indices clamped, values past nseg read as 0): one wait on memory
        // per 16 segments instead of a chain of dependent loads
        double s = 0.0;
        for (int g0 = 0; g0 < nseg; g0 += 16) {
            double tv[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const double v = wpart[static_cast<size_t>(min(g0 + e, nseg - 1)) * kMaxNQ + tid];
                tv[e] = g0 + e < nseg ? v : 0.0;
            }
#pragma unroll
            for (int e = 0; e < 16; ++e) s += tv[e];
        }
        L.cc[tid] = s;
    }
    __syncthreads();
    pc_coarse_solve(L, Gw, nc, nc, nullptr);
    const int i0 = blockIdx.x * kApplyBlk, i1 = min(no, i0 + kApplyBlk);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    if (scal == nullptr) {
        pc_expand<0>(L, no, nfw, i0, i1, r, 0, no, z, 0, ex, ey, sg, a0, a1, a2);
        return;
    }
    pc_expand<1>(L, no, nfw, i0, i1, r, 0, no, z, 0, ex, ey, sg, a0, a1, a2);
    pc_block_sum3(L, a0, a1, a2);
    if (tid == 0) scal[blockIdx.x] = a0;
}

}  // namespace

// Coarse inverses of `blocks` curves starting at first_curve, each nq x nq.
// Coarse inverses of `blocks` curves starting at first_curve, each nq x nq;
// bad[i] = 1 where curve first_curve + i's block is singular.
static void coarse_launch(Context& c, int blocks, int nq, int nfw, int nfb, int first_curve,
                          double* out, int* bad) {
    const Geometry& g = c.g;
    constexpr int kSmallNQ = 2 * (2 * 16 + 1) + 3;  // 16 modes
    if (nq <= kSmallNQ)
        k_pc_coarse<kSmallNQ><<<blocks, kCoarseBlk, 0, c.stream>>>(
            g.x.p, g.y.p, g.px.p, g.py.p, g.cur.p, g.speed.p, g.cx.p, g.cy.p, c.pc_phi_w.p,
            c.pc_phi_b.p, nfw, nfb, g.no, g.n, first_curve, c.pc_S.p, out, bad);
    else
        k_pc_coarse<kMaxNQ><<<blocks, kCoarseBlk, 0, c.stream>>>(
            g.x.p, g.y.p, g.px.p, g.py.p, g.cur.p, g.speed.p, g.cx.p, g.cy.p, c.pc_phi_w.p,
            c.pc_phi_b.p, nfw, nfb, g.no, g.n, first_curve, c.pc_S.p, out, bad);
    STOKES_LAUNCH_CHECK();
}

// Refuses the preconditioner when a coarse block came out singular.
static void check_coarse(Context& c, int count, int first_curve) {
    std::vector<int> bad(static_cast<size_t>(count));
    STOKES_HIP(hipMemcpyAsync(bad.data(), c.pc_bad.p, bad.size() * sizeof(int),
                              hipMemcpyDeviceToHost, c.stream));
    STOKES_HIP(hipStreamSynchronize(c.stream));
    for (int i = 0; i < count; ++i)
        if (bad[static_cast<size_t>(i)])
            throw Error("the preconditioner's coarse block of curve " +
                        std::to_string(first_curve + i) +
                        " is singular (a pivot below 1e-13 of its largest entry); solve without "
                        "the preconditioner");
}

void precond_setup(Context& c) {
    Geometry& g = c.g;
    hipStream_t s = c.stream;
    const int n = g.n, m = g.m, no = g.no;
    if (c.pc_body_modes < 1 || c.pc_body_modes > kMaxModes || c.pc_wall_modes < 1 ||
        c.pc_wall_modes > kMaxModes)
        throw Error("preconditioner modes must lie in [1, 32]");
    // a curve of len points holds modes 0..(len-2)/2 without the Nyquist mode
    const int kb = std::min(c.pc_body_modes, (n - 2) / 2);
    const int kw = std::min(c.pc_wall_modes, (no - 2) / 2);
    if (m > 0 && kb < 1)
        throw Error("the preconditioner needs at least 4 points per body");
    if (kw < 1) throw Error("the preconditioner needs at least 4 wall points");
    const int nfb = 2 * kb + 1, nfw = 2 * kw + 1;
    const int nqb = 2 * nfb + 3, nqw = 2 * nfw;
    c.pc_phi_b.ensure(static_cast<size_t>(nfb) * std::max(n, 1));
    c.pc_phi_w.ensure(static_cast<size_t>(nfw) * no);
    c.pc_S.ensure(static_cast<size_t>(no) * 3 * nfw + static_cast<size_t>(m) * n * 3 * nfb);
    // the wall's partial coefficients of the apply (allocated here, before
    // any GMRES chunk may be captured into a graph)
    c.pc_wpart.ensure(static_cast<size_t>(cdiv_int(no, kApplyBlk)) * kMaxNQ);
    c.scal.ensure(std::max<size_t>(8, static_cast<size_t>(cdiv_int(no, kApplyBlk))));
    k_pc_basis<<<cdiv_int(no, kBlk), kBlk, 0, s>>>(no, kw, c.pc_phi_w.p);
    STOKES_LAUNCH_CHECK();
    if (m > 0) {
        k_pc_basis<<<cdiv_int(n, kBlk), kBlk, 0, s>>>(n, kb, c.pc_phi_b.p);
        STOKES_LAUNCH_CHECK();
    }
    // the wall is fixed by its point count, so its block is kept between solves
    if (c.pc_wall_no != no || c.pc_wall_kw != kw) {
        c.pc_Gw.ensure(static_cast<size_t>(nqw) * nqw);
        k_pc_pairs<<<dim3(1, cdiv_int(no, kBlk), cdiv_int(nfw, kQC)), kBlk, 0, s>>>(
            g.x.p, g.y.p, g.mx.p, g.my.p, c.pc_phi_w.p, c.pc_phi_b.p, nfw, nfb, no, n, 0,
            c.pc_S.p);
        STOKES_LAUNCH_CHECK();
        c.pc_bad.ensure(static_cast<size_t>(std::max(m, 1)));
        coarse_launch(c, 1, nqw, nfw, nfb, 0, c.pc_Gw.p, c.pc_bad.p);
        check_coarse(c, 1, 0);
        c.pc_wall_no = no;
        c.pc_wall_kw = kw;
    }
    if (m > 0) {
        c.pc_Gb.ensure(static_cast<size_t>(m) * nqb * nqb);
        k_pc_pairs<<<dim3(m, cdiv_int(n, kBlk), cdiv_int(nfb, kQC)), kBlk, 0, s>>>(
            g.x.p, g.y.p, g.mx.p, g.my.p, c.pc_phi_w.p, c.pc_phi_b.p, nfw, nfb, no, n, 1,
            c.pc_S.p);
        STOKES_LAUNCH_CHECK();
        c.pc_bad.ensure(static_cast<size_t>(m));
        coarse_launch(c, m, nqb, nfw, nfb, 1, c.pc_Gb.p, c.pc_bad.p);
        check_coarse(c, m, 1);
    }
    c.pc_nfb = nfb;
    c.pc_nfw = nfw;
}

int precond_wall_segments(const Geometry& g) { return cdiv_int(g.no, kApplyBlk); }

void precond_apply_scaled(Context& c, double* r, double* z, const GmresState* st, double* y) {
    Geometry& g = c.g;
    if (c.time_kernels) c.t_precond.start(c.stream);
    const int nseg = cdiv_int(g.no, kApplyBlk);
    c.pc_wpart.ensure(static_cast<size_t>(nseg) * kMaxNQ);
    c.ex.ensure(static_cast<size_t>(g.npts));
    c.ey.ensure(static_cast<size_t>(g.npts));
    if (y != nullptr) c.scal.ensure(std::max<size_t>(8, static_cast<size_t>(nseg)));
    const SumGeom sg{g.x.p, g.y.p, g.px.p, g.py.p, g.speed.p};
    k_pc_apply<<<nseg + g.m, kApplyBlk, 0, c.stream>>>(c.pc_nfw, c.pc_nfb, g.no, g.n, g.m, nseg,
                                                      c.pc_Gb.p, st, r, c.pc_wpart.p, z, c.ex.p,
                                                      c.ey.p, sg, y);
    STOKES_LAUNCH_CHECK();
    k_pc_wall<<<nseg, kApplyBlk, 0, c.stream>>>(c.pc_nfw, g.no, nseg, c.pc_Gw.p, c.pc_wpart.p, r,
                                                z, c.ex.p, c.ey.p, sg,
                                                y != nullptr ? c.scal.p : nullptr);
    STOKES_LAUNCH_CHECK();
    // bytes: the vector read and written once
    if (c.time_kernels) c.t_precond.stop(c.stream, 2.0 * 8.0 * static_cast<double>(g.nunk));
    // z's density now sits in c.ex/c.ey and, with y, y's constraint rows and
    // the rank-one parts in c.scal: apply_operator(z, y, from_precond) may
    // reuse them
    c.pc_token_x = y != nullptr ? z : nullptr;
    c.pc_token_y = y;
}

void precond_apply(Context& c, const double* r, double* z) {
    // unscaled: the kernel writes r back only when it scales
    precond_apply_scaled(c, const_cast<double*>(r), z, nullptr);
}

}  // namespace stokes
