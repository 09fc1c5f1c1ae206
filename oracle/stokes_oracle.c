/*
 * stokes_oracle.c -- CPU restatement (plain C, FP64) of the reference's
 * per-step Stokes solve.  TEST INFRASTRUCTURE ONLY: the checker that tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg compare the HIP
 * library against.  The product never links or calls anything in here.
 *
 * Every function names the reference routine it restates
 * (ErosionPackage/fortran_code/stokesSolver.f unless stated).  Loops keep the
 * reference's summation order where it is cheap to do so, so that results
 * agree with the flang build of the reference to rounding level.
 *
 * Parallelism: OpenMP over targets only; every target's sum runs on one
 * thread in a fixed order, so results do not depend on the thread count.
 */
#include "stokes_oracle.h"

#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* pi = 4*datan(1), twopi = 8*datan(1) in the reference; both are the
 * correctly rounded doubles. */
static const double PI = 3.14159265358979323846;
static const double TWOPI = 6.28318530717958647692;

struct orc_geom {
    int n, m, no;
    double *x, *y, *px, *py, *cur, *speed, *cx, *cy;
    double *xo, *yo, *px0, *py0, *cur0, *speed0;
};

static int g_threads = 0;
void orc_set_threads(int nthreads) { g_threads = nthreads; }
int orc_get_threads(void) { return g_threads > 0 ? g_threads : omp_get_max_threads(); }
#define OMP_NT num_threads(orc_get_threads())

static void *xmalloc(size_t bytes) {
    void *p = malloc(bytes ? bytes : 1);
    if (!p) {
        fprintf(stderr, "stokes_oracle: out of memory (%zu bytes)\n", bytes);
        abort();
    }
    return p;
}
static void *xcalloc(size_t count, size_t size) {
    void *p = calloc(count ? count : 1, size ? size : 1);
    if (!p) {
        fprintf(stderr, "stokes_oracle: out of memory\n");
        abort();
    }
    return p;
}

/* ------------------------------------------------------------------ FFT */
/* Unnormalised complex DFT with FFTPACK's sign conventions (dcfft.f):
 * sign=-1 is DCFFTF (forward), sign=+1 is DCFFTB (backward). */
static void dft(int n, cplx *a, int sign) {
    if (n <= 1) return;
    cplx *w = xmalloc(sizeof(cplx) * (size_t)n);
    for (int t = 0; t < n; t++) {
        double ang = TWOPI * (double)t / (double)n;
        w[t] = cos(ang) + (double)sign * I * sin(ang);
    }
    if ((n & (n - 1)) == 0) {
        /* iterative radix-2 */
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                cplx t = a[i];
                a[i] = a[j];
                a[j] = t;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int step = n / len;
            for (int i = 0; i < n; i += len) {
                for (int k = 0; k < len / 2; k++) {
                    cplx u = a[i + k];
                    cplx v = a[i + k + len / 2] * w[k * step];
                    a[i + k] = u + v;
                    a[i + k + len / 2] = u - v;
                }
            }
        }
    } else {
        cplx *out = xmalloc(sizeof(cplx) * (size_t)n);
        for (int k = 0; k < n; k++) {
            cplx s = 0;
            for (int j = 0; j < n; j++) s += a[j] * w[(int)(((long)j * k) % n)];
            out[k] = s;
        }
        memcpy(a, out, sizeof(cplx) * (size_t)n);
        free(out);
    }
    free(w);
}

/* fourierDiff (stokesSolver.f:348-388): spectral derivative in the grid
 * parameter, positive modes 0..n/2-1, Nyquist slot n/2 zeroed, negative modes
 * k-n for k = n/2+1..n-1; forward and backward transforms unnormalised, the
 * 1/n applied with the mode factor. */
static void fourier_diff(int n, cplx *f) {
    dft(n, f, -1);
    for (int k = 0; k < n / 2; k++) f[k] = f[k] * I * (double)k / (double)n;
    f[n / 2] = 0;
    for (int k = n / 2 + 1; k < n; k++) f[k] = f[k] * I * (double)(k - n) / (double)n;
    dft(n, f, +1);
}

void orc_fourier_diff(int n, double *re, double *im) {
    cplx *f = xmalloc(sizeof(cplx) * (size_t)n);
    for (int k = 0; k < n; k++) f[k] = re[k] + I * im[k];
    fourier_diff(n, f);
    for (int k = 0; k < n; k++) {
        re[k] = creal(f[k]);
        im[k] = cimag(f[k]);
    }
    free(f);
}

/* ------------------------------------------------------------- geometry */
/* compute_derivs (stokesSolver.f:417-457). */
static void compute_derivs(int n, int nb, const double *x, const double *y, double *px,
                           double *py, double *cur, double *speed) {
    cplx *z1 = xmalloc(sizeof(cplx) * (size_t)n);
    cplx *z2 = xmalloc(sizeof(cplx) * (size_t)n);
    for (int j = 0; j < nb; j++) {
        for (int k = 0; k < n; k++) z1[k] = x[j * n + k] + I * y[j * n + k];
        fourier_diff(n, z1);
        memcpy(z2, z1, sizeof(cplx) * (size_t)n);
        fourier_diff(n, z2);
        for (int k = 0; k < n; k++) {
            double Dx = creal(z1[k]), Dy = cimag(z1[k]);
            double DDx = creal(z2[k]), DDy = cimag(z2[k]);
            double sp = sqrt(Dx * Dx + Dy * Dy);
            speed[j * n + k] = sp;
            px[j * n + k] = -Dy / sp;
            py[j * n + k] = Dx / sp;
            cur[j * n + k] = (Dy * DDx - Dx * DDy) / (sp * sp * sp);
        }
    }
    free(z1);
    free(z2);
}

/* outer_geometry (stokesSolver.f:169-204): superellipse of order 8, aspect
 * 3:1, normal and curvature negated to point outwards. */
static void outer_geometry(int no, double *x, double *y, double *px, double *py, double *cur,
                           double *speed) {
    double dth = TWOPI / (double)no;
    for (int k = 0; k < no; k++) {
        double th = (double)k * dth;
        double r = pow(pow(cos(th), 8.0) + pow(sin(th), 8.0), -1.0 / 8.0);
        x[k] = 3.0 * r * cos(th);
        y[k] = r * sin(th);
    }
    compute_derivs(no, 1, x, y, px, py, cur, speed);
    for (int k = 0; k < no; k++) {
        px[k] = -px[k];
        py[k] = -py[k];
        cur[k] = -cur[k];
    }
}

orc_geom *orc_geom_create(int n, int m, int no, const double *xx, const double *yy) {
    orc_geom *g = xcalloc(1, sizeof(orc_geom));
    size_t nb = (size_t)n * (size_t)m;
    g->n = n;
    g->m = m;
    g->no = no;
    g->x = xmalloc(sizeof(double) * nb);
    g->y = xmalloc(sizeof(double) * nb);
    g->px = xmalloc(sizeof(double) * nb);
    g->py = xmalloc(sizeof(double) * nb);
    g->cur = xmalloc(sizeof(double) * nb);
    g->speed = xmalloc(sizeof(double) * nb);
    g->cx = xmalloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    g->cy = xmalloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
    g->xo = xmalloc(sizeof(double) * (size_t)no);
    g->yo = xmalloc(sizeof(double) * (size_t)no);
    g->px0 = xmalloc(sizeof(double) * (size_t)no);
    g->py0 = xmalloc(sizeof(double) * (size_t)no);
    g->cur0 = xmalloc(sizeof(double) * (size_t)no);
    g->speed0 = xmalloc(sizeof(double) * (size_t)no);
    /* inner_geometry (stokesSolver.f:136-166): copy, centres = point means */
    memcpy(g->x, xx, sizeof(double) * nb);
    memcpy(g->y, yy, sizeof(double) * nb);
    for (int b = 0; b < m; b++) {
        double sx = 0.0, sy = 0.0;
        for (int k = 0; k < n; k++) {
            sx += g->x[b * n + k];
            sy += g->y[b * n + k];
        }
        g->cx[b] = sx / (double)n;
        g->cy[b] = sy / (double)n;
    }
    if (m > 0) compute_derivs(n, m, g->x, g->y, g->px, g->py, g->cur, g->speed);
    outer_geometry(no, g->xo, g->yo, g->px0, g->py0, g->cur0, g->speed0);
    return g;
}

void orc_geom_free(orc_geom *g) {
    if (!g) return;
    double *arrs[] = {g->x,  g->y,  g->px,  g->py,  g->cur,  g->speed,  g->cx,    g->cy,
                      g->xo, g->yo, g->px0, g->py0, g->cur0, g->speed0};
    for (size_t i = 0; i < sizeof(arrs) / sizeof(arrs[0]); i++) free(arrs[i]);
    free(g);
}

const double *orc_geom_array(const orc_geom *g, int which) {
    switch (which) {
    case 0: return g->px;
    case 1: return g->py;
    case 2: return g->cur;
    case 3: return g->speed;
    case 4: return g->cx;
    case 5: return g->cy;
    case 6: return g->xo;
    case 7: return g->yo;
    case 8: return g->px0;
    case 9: return g->py0;
    case 10: return g->cur0;
    case 11: return g->speed0;
    default: return NULL;
    }
}

/* Overwrite one geometry array with caller values (the same `which` and
 * length as orc_geom_array).  Lets a test evaluate this restatement's
 * operator on exactly another implementation's normals, curvature and
 * speed, so that their difference measures the operator code alone. */
int orc_geom_set_array(orc_geom *g, int which, const double *v) {
    size_t len;
    double *dst = (double *)orc_geom_array(g, which);
    if (!dst || !v) return -1;
    if (which < 4) len = (size_t)g->n * (size_t)g->m;
    else if (which < 6) len = (size_t)g->m;
    else len = (size_t)g->no;
    memcpy(dst, v, len * sizeof(double));
    return 0;
}

/* bd_condition + bgFlow (stokesSolver.f:208-314, 392-413): the live branch
 * sets wall rows to the background flow and every other row to zero. */
void orc_rhs(const orc_geom *g, int ibc, double *rhs) {
    int n = g->n, m = g->m, no = g->no;
    size_t nunk = (size_t)(2 * no + 2 * n * m + 3 * m);
    memset(rhs, 0, sizeof(double) * nunk);
    for (int k = 0; k < no; k++) {
        double yk = g->yo[k];
        if (ibc == 0) {
            rhs[k] = 1.0 * (1.0 + yk) * (1.0 - yk);
            rhs[k + no] = 0.0;
        } else if (ibc == 1) {
            rhs[k] = 1.0;
            rhs[k + no] = 0.0;
        }
    }
}

/* ---------------------------------------------------- trapezoid kernels */
/* One source node of the trapezoid double layer as written in matvec_DLP
 * (stokesSolver.f:1570-1579). */
/* The operator's sums are accumulated in x87 extended precision (64-bit
 * significand): with bodies whose densities are 10^4..10^6 times the
 * right-hand side, the running sums of a target's pair terms are that much
 * larger than the result, and FP64 accumulation alone would leave the operator
 * uncertain at ~1e-8 of ||b||; each pair term is still formed in FP64 as the
 * reference forms it. */
typedef long double acc_t;

static inline void trap_pair(double rx, double ry, double nx, double ny, double dx, double dy,
                             double sp, acc_t *ux, acc_t *uy) {
    double rdotn = rx * nx + ry * ny;
    double rdotden = rx * dx + ry * dy;
    double rho2 = rx * rx + ry * ry;
    *ux += rdotn * rdotden / rho2 / rho2 * rx * sp;
    *uy += rdotn * rdotden / rho2 / rho2 * ry * sp;
}

/* ------------------------------------------------- barycentric quadrature */
/* compute_bdval_ex / compute_bdval_in (stokesSolver.f:5382-5427, 5103-5146):
 * boundary limit of the Cauchy integral of f over the source curve. */
static void compute_bdval(int npts, const double *xs, const double *ys, const cplx *f,
                          int exterior, cplx *bdval) {
    double dth = TWOPI / (double)npts;
    cplx *df = xmalloc(sizeof(cplx) * (size_t)npts);
    cplx *z = xmalloc(sizeof(cplx) * (size_t)npts);
    cplx *dz = xmalloc(sizeof(cplx) * (size_t)npts);
    memcpy(df, f, sizeof(cplx) * (size_t)npts);
    fourier_diff(npts, df);
    for (int k = 0; k < npts; k++) {
        z[k] = xs[k] + I * ys[k];
        dz[k] = z[k] * dth;
    }
    fourier_diff(npts, dz);
#pragma omp parallel for OMP_NT schedule(static)
    for (int j = 0; j < npts; j++) {
        cplx b = 0;
        for (int k = 0; k < npts; k++) {
            if (k != j)
                b = b - (f[k] - f[j]) / (z[k] - z[j]) * dz[k];
            else
                b = b - TWOPI / (double)npts * df[j];
        }
        if (exterior) {
            b = -I * b / TWOPI;
            bdval[j] = -b;
        } else {
            bdval[j] = -I * b / TWOPI - f[j];
        }
    }
    free(df);
    free(z);
    free(dz);
}

/* laplaceInteriorHolomorphic (stokesSolver.f:5149-5230): barycentric value,
 * first and (nder=2) second derivative of the interior Cauchy integral. */
static void holo_interior(int npts, const double *xs, const double *ys, const cplx *bdval,
                          int ntar, const double *xt, const double *yt, int nder, cplx *u,
                          cplx *du, cplx *ddu) {
    double dth = TWOPI / (double)npts;
    cplx *z = xmalloc(sizeof(cplx) * (size_t)npts);
    cplx *dz = xmalloc(sizeof(cplx) * (size_t)npts);
    for (int k = 0; k < npts; k++) {
        z[k] = xs[k] + I * ys[k];
        dz[k] = z[k] * dth;
    }
    fourier_diff(npts, dz);
#pragma omp parallel for OMP_NT schedule(static)
    for (int j = 0; j < ntar; j++) {
        cplx zt = xt[j] + I * yt[j];
        cplx num = 0, den = 0;
        for (int k = 0; k < npts; k++) {
            cplx d1 = z[k] - zt;
            num = num + bdval[k] / d1 * dz[k];
            den = den + 1.0 / d1 * dz[k];
        }
        cplx uj = num / den;
        num = 0;
        for (int k = 0; k < npts; k++) {
            cplx d1 = z[k] - zt;
            cplx d2 = d1 * d1;
            num = num + (bdval[k] - uj) / d2 * dz[k];
        }
        cplx duj = num / den;
        u[j] = uj;
        du[j] = duj;
        if (nder == 2) {
            cplx num1 = 0;
            for (int k = 0; k < npts; k++) {
                cplx d1 = z[k] - zt;
                cplx d3 = d1 * d1 * d1;
                num1 = num1 + (bdval[k] - uj - duj * (z[k] - zt)) / d3 * dz[k];
            }
            ddu[j] = 2.0 * num1 / den;
        }
    }
    free(z);
    free(dz);
}

/* laplaceExteriorHolomorphic (stokesSolver.f:5430-5528).  Keeps the
 * reference's exact forms: value num/(den - 2*pi*i) (:5481); derivatives
 * through the point-mean a with D_a = sum dz/((z-a)(z-x)) (:5507-5523); and
 * the coincident-node branch |z_k - x| <= 1e-10 (:5492-5506) whose
 * accumulators tmp1/tmp2 are cleared once per target, not per node. */
static void holo_exterior(int npts, const double *xs, const double *ys, const cplx *bdval,
                          int ntar, const double *xt, const double *yt, int nder, cplx *u,
                          cplx *du, cplx *ddu) {
    double dth = TWOPI / (double)npts;
    double ax = 0.0, ay = 0.0;
    for (int k = 0; k < npts; k++) {
        ax = ax + xs[k] / (double)npts;
        ay = ay + ys[k] / (double)npts;
    }
    cplx a = ax + I * ay;
    cplx *z = xmalloc(sizeof(cplx) * (size_t)npts);
    cplx *dz = xmalloc(sizeof(cplx) * (size_t)npts);
    for (int k = 0; k < npts; k++) {
        z[k] = xs[k] + I * ys[k];
        dz[k] = z[k] * dth;
    }
    fourier_diff(npts, dz);
#pragma omp parallel for OMP_NT schedule(static)
    for (int j = 0; j < ntar; j++) {
        cplx zt = xt[j] + I * yt[j];
        cplx num = 0, den = 0;
        for (int k = 0; k < npts; k++) {
            cplx d1 = z[k] - zt;
            num = num + bdval[k] / d1 * dz[k];
            den = den + 1.0 / d1 * dz[k];
        }
        cplx uj = num / (-TWOPI * I + den);
        num = 0;
        den = 0;
        cplx tmp1 = 0, tmp2 = 0;
        for (int k = 0; k < npts; k++) {
            cplx d1 = z[k] - zt;
            cplx d2 = d1 * d1;
            if (cabs(z[k] - zt) > 1.0e-10) {
                num = num + (bdval[k] - uj) / d2 * dz[k];
            } else {
                for (int l = 0; l < npts; l++) {
                    if (l != k)
                        tmp1 = tmp1 + (bdval[k] * (z[k] - a) / (z[l] - a) - bdval[l]) * dz[l] /
                                          (z[l] - zt);
                    tmp2 = tmp2 + 1.0 / ((z[l] - a) * (z[l] - zt)) * dz[l];
                }
                num = num + (tmp1 / tmp2 - (z[k] - zt) * bdval[k]) / d2 / (zt - a) * dz[k];
            }
            den = den + 1.0 / ((z[k] - a) * (z[k] - zt)) * dz[k];
        }
        cplx duj = num / (zt - a) / den;
        u[j] = uj;
        du[j] = duj;
        if (nder == 2) {
            cplx num1 = 0;
            for (int k = 0; k < npts; k++) {
                cplx d1 = z[k] - zt;
                cplx d3 = d1 * d1 * d1;
                num1 = num1 + (bdval[k] - uj - duj * (z[k] - zt)) / d3 * dz[k];
            }
            ddu[j] = 2.0 * num1 / (zt - a) / den;
        }
    }
    free(z);
    free(dz);
}

/* StokesExteriorDLP / StokesInteriorDLP (stokesSolver.f:5233-5379,
 * 4937-5100): Stokes double layer (and with nder=2 its gradient) as five
 * Cauchy integrals with densities s.eta, eta_x, eta_y, tau1, tau2. */
static void stokes_dlp(int exterior, int npts, const double *xs, const double *ys,
                       const double *dx, const double *dy, const double *px1,
                       const double *py1, int ntar, const double *xt, const double *yt,
                       int nder, double *ux, double *uy, double *u1x, double *u1y, double *u2x,
                       double *u2y) {
    cplx *f[5];
    for (int i = 0; i < 5; i++) f[i] = xmalloc(sizeof(cplx) * (size_t)npts);
    for (int j = 0; j < npts; j++) {
        cplx outnor = px1[j] + I * py1[j];
        cplx eta = dx[j] + I * dy[j];
        f[3][j] = exterior ? eta * px1[j] / outnor : px1[j] * eta / outnor;
        f[4][j] = exterior ? eta * py1[j] / outnor : py1[j] * eta / outnor;
        f[0][j] = xs[j] * dx[j] + ys[j] * dy[j];
        f[1][j] = dx[j];
        f[2][j] = dy[j];
    }
    cplx *bd = xmalloc(sizeof(cplx) * (size_t)npts);
    cplx *u = xmalloc(sizeof(cplx) * (size_t)(ntar > 0 ? ntar : 1));
    cplx *du = xmalloc(sizeof(cplx) * (size_t)(ntar > 0 ? ntar : 1));
    cplx *ddu = xcalloc((size_t)(ntar > 0 ? ntar : 1), sizeof(cplx));
    for (int j = 0; j < ntar; j++) {
        ux[j] = uy[j] = 0.0;
        if (nder == 2) u1x[j] = u1y[j] = u2x[j] = u2y[j] = 0.0;
    }
    for (int fld = 0; fld < 5; fld++) {
        compute_bdval(npts, xs, ys, f[fld], exterior, bd);
        if (exterior)
            holo_exterior(npts, xs, ys, bd, ntar, xt, yt, nder, u, du, ddu);
        else
            holo_interior(npts, xs, ys, bd, ntar, xt, yt, nder, u, du, ddu);
        for (int j = 0; j < ntar; j++) {
            double rdu = creal(du[j]), idu = cimag(du[j]);
            double rdd = creal(ddu[j]), idd = cimag(ddu[j]);
            switch (fld) {
            case 0:
                ux[j] += rdu;
                uy[j] -= idu;
                if (nder == 2) {
                    u1x[j] += rdd;
                    u1y[j] -= idd;
                    u2x[j] -= idd;
                    u2y[j] -= rdd;
                }
                break;
            case 1:
                ux[j] -= xt[j] * rdu;
                uy[j] += xt[j] * idu;
                if (nder == 2) {
                    u1x[j] = u1x[j] - rdu - xt[j] * rdd;
                    u1y[j] = u1y[j] + xt[j] * idd;
                    u2x[j] = u2x[j] + idu + xt[j] * idd;
                    u2y[j] = u2y[j] + xt[j] * rdd;
                }
                break;
            case 2:
                ux[j] -= yt[j] * rdu;
                uy[j] += yt[j] * idu;
                if (nder == 2) {
                    u1x[j] = u1x[j] - yt[j] * rdd;
                    u1y[j] = u1y[j] - rdu + yt[j] * idd;
                    u2x[j] = u2x[j] + yt[j] * idd;
                    u2y[j] = u2y[j] + idu + yt[j] * rdd;
                }
                break;
            case 3:
                ux[j] += creal(u[j]);
                if (nder == 2) {
                    u1x[j] += rdu;
                    u1y[j] -= idu;
                }
                break;
            case 4:
                uy[j] += creal(u[j]);
                if (nder == 2) {
                    u2x[j] += rdu;
                    u2y[j] -= idu;
                }
                break;
            }
        }
    }
    for (int i = 0; i < 5; i++) free(f[i]);
    free(bd);
    free(u);
    free(du);
    free(ddu);
}

void orc_stokes_exterior_dlp(int npts, const double *xs, const double *ys, const double *dx,
                             const double *dy, const double *px1, const double *py1, int ntar,
                             const double *xt, const double *yt, int nder, double *ux,
                             double *uy, double *u1x, double *u1y, double *u2x, double *u2y) {
    stokes_dlp(1, npts, xs, ys, dx, dy, px1, py1, ntar, xt, yt, nder, ux, uy, u1x, u1y, u2x,
               u2y);
}

void orc_stokes_interior_dlp(int npts, const double *xs, const double *ys, const double *dx,
                             const double *dy, const double *px1, const double *py1, int ntar,
                             const double *xt, const double *yt, int nder, double *ux,
                             double *uy, double *u1x, double *u1y, double *u2x, double *u2y) {
    stokes_dlp(0, npts, xs, ys, dx, dy, px1, py1, ntar, xt, yt, nder, ux, uy, u1x, u1y, u2x,
               u2y);
}

/* ----------------------------------------------------------------- matvec */
/* Completed double-layer operator: matvec_DLP (ibary=0, stokesSolver.f:
 * 1519-1912) and matvec_DLP_fmmbary (ibary=1, stokesSolver.f:570-1124). */
void orc_matvec(const orc_geom *g, int ibary, const double *den, double *vel) {
    const int n = g->n, m = g->m, no = g->no;
    const size_t nunk = (size_t)(2 * no + 2 * n * m + 3 * m);
    const int nbp = n * m;
    const int ntgt = no + nbp;
    acc_t *acc = xcalloc(nunk, sizeof(acc_t));
    const double *wdx = den, *wdy = den + no;

    /* Far field: trapezoid rule over every source node except the target
     * itself, one source curve at a time (stokesSolver.f:1563-1584,
     * 1607-1618, 1650-1672, 1702-1722, 1737-1757). */
#pragma omp parallel for OMP_NT schedule(dynamic, 16)
    for (int t = 0; t < ntgt; t++) {
        double xt, yt;
        int tb = -1, tk = -1;
        if (t < no) {
            xt = g->xo[t];
            yt = g->yo[t];
        } else {
            tb = (t - no) / n;
            tk = (t - no) % n;
            xt = g->x[tb * n + tk];
            yt = g->y[tb * n + tk];
        }
        acc_t ut = 0.0, vt = 0.0;
        acc_t ux = 0.0, uy = 0.0;
        for (int j = 0; j < no; j++) {
            if (t < no && j == t) continue;
            trap_pair(xt - g->xo[j], yt - g->yo[j], g->px0[j], g->py0[j], wdx[j], wdy[j],
                      g->speed0[j], &ux, &uy);
        }
        ut += ux * TWOPI / (double)no / PI;
        vt += uy * TWOPI / (double)no / PI;
        for (int s = 0; s < m; s++) {
            const double *sdx = den + 2 * no + 2 * s * n;
            const double *sdy = sdx + n;
            ux = uy = 0.0;
            for (int j = 0; j < n; j++) {
                if (s == tb && j == tk) continue;
                int q = s * n + j;
                trap_pair(xt - g->x[q], yt - g->y[q], g->px[q], g->py[q], sdx[j], sdy[j],
                          g->speed[q], &ux, &uy);
            }
            ut += ux * TWOPI / (double)n / PI;
            vt += uy * TWOPI / (double)n / PI;
        }
        if (t < no) {
            acc[t] = ut;
            acc[t + no] = vt;
        } else {
            acc[2 * no + 2 * tb * n + tk] = ut;
            acc[2 * no + 2 * tb * n + n + tk] = vt;
        }
    }

    if (ibary == 1 && m > 0) {
        /* Near-pair replacement of the trapezoid rule by barycentric
         * quadrature (stokesSolver.f:687-926). */
        const double alpha1 = 4.0, alpha2 = 4.0;
        double *rtar = xmalloc(sizeof(double) * (size_t)m);
        for (int b = 0; b < m; b++) {
            double s = 0.0;
            for (int k = 0; k < n; k++) s = s + g->speed[b * n + k];
            rtar[b] = s / (double)n;
        }
        /* 'lw' is implicitly INTEGER in the reference (stokesSolver.f:705-709):
         * every partial sum is truncated and the mean is an integer division. */
        int lw = 0;
        for (int k = 0; k < no; k++) lw = (int)((double)lw + g->speed0[k]);
        lw = lw / no;

        int cap = nbp > no ? nbp : no;
        double *xtar = xmalloc(sizeof(double) * (size_t)cap);
        double *ytar = xmalloc(sizeof(double) * (size_t)cap);
        int *itgt = xmalloc(sizeof(int) * (size_t)cap);
        double *ux = xmalloc(sizeof(double) * (size_t)cap);
        double *uy = xmalloc(sizeof(double) * (size_t)cap);
        double *wxt = xmalloc(sizeof(double) * (size_t)cap);
        double *wyt = xmalloc(sizeof(double) * (size_t)cap);
        int *wtgt = xmalloc(sizeof(int) * (size_t)cap);
        int mw = 0;

        for (int it = 0; it < m; it++) {
            double dsum = rtar[it] + alpha2 * (double)lw * TWOPI / (double)no;
            double dsum2 = 1.0 - dsum;
            if (!(fabs(g->cy[it]) >= dsum2)) continue;
            const double *bdx = den + 2 * no + 2 * it * n;
            const double *bdy = bdx + n;
            /* wall -> body points with |y| >= dsum2: remove trapezoid */
            for (int k = 0; k < n; k++) {
                int q = it * n + k;
                if (!(fabs(g->y[q]) >= dsum2)) continue;
                acc_t sx = 0.0, sy = 0.0;
                for (int j = 0; j < no; j++)
                    trap_pair(g->x[q] - g->xo[j], g->y[q] - g->yo[j], g->px0[j], g->py0[j],
                              wdx[j], wdy[j], g->speed0[j], &sx, &sy);
                sx = sx * TWOPI / (double)no / PI;
                sy = sy * TWOPI / (double)no / PI;
                acc[2 * no + 2 * it * n + k] -= sx;
                acc[2 * no + 2 * it * n + n + k] -= sy;
                wxt[mw] = g->x[q];
                wyt[mw] = g->y[q];
                wtgt[mw] = q;
                mw++;
            }
            /* body -> wall points within r of the centre */
            int nw = 0;
            for (int j = 0; j < no; j++) {
                double dx = g->cx[it] - g->xo[j];
                double dy = g->cy[it] - g->yo[j];
                double d2 = dx * dx + dy * dy;
                double r = 1.0 - fabs(g->cy[it]) + 2.0 * rtar[it];
                if (d2 <= r * r) {
                    xtar[nw] = g->xo[j];
                    ytar[nw] = g->yo[j];
                    itgt[nw] = j;
                    nw++;
                }
            }
            for (int k = 0; k < nw; k++) {
                acc_t sx = 0.0, sy = 0.0;
                for (int j = 0; j < n; j++) {
                    int q = it * n + j;
                    trap_pair(xtar[k] - g->x[q], ytar[k] - g->y[q], g->px[q], g->py[q], bdx[j],
                              bdy[j], g->speed[q], &sx, &sy);
                }
                sx = sx * TWOPI / (double)n / PI;
                sy = sy * TWOPI / (double)n / PI;
                acc[itgt[k]] -= sx;
                acc[itgt[k] + no] -= sy;
            }
            if (nw > 0) {
                orc_stokes_exterior_dlp(n, g->x + it * n, g->y + it * n, bdx, bdy,
                                        g->px + it * n, g->py + it * n, nw, xtar, ytar, 1, ux,
                                        uy, NULL, NULL, NULL, NULL);
                for (int k = 0; k < nw; k++) {
                    acc[itgt[k]] += ux[k];
                    acc[itgt[k] + no] += uy[k];
                }
            }
        }
        if (mw > 0) {
            orc_stokes_interior_dlp(no, g->xo, g->yo, wdx, wdy, g->px0, g->py0, mw, wxt, wyt, 1,
                                    ux, uy, NULL, NULL, NULL, NULL);
            for (int k = 0; k < mw; k++) {
                int b = wtgt[k] / n, kk = wtgt[k] % n;
                acc[2 * no + 2 * b * n + kk] += ux[k];
                acc[2 * no + 2 * b * n + n + kk] += uy[k];
            }
        }
        /* body -> body (stokesSolver.f:838-926) */
        for (int is = 0; is < m; is++) {
            const double *sdx = den + 2 * no + 2 * is * n;
            const double *sdy = sdx + n;
            int l = 0;
            for (int it = 0; it < m; it++) {
                if (it == is) continue;
                double dx = g->cx[is] - g->cx[it];
                double dy = g->cy[is] - g->cy[it];
                double d2 = dx * dx + dy * dy;
                double rsum = rtar[is] + rtar[it] + alpha1 * (rtar[is] + rtar[it]) * TWOPI / (double)n;
                if (!(d2 <= rsum * rsum)) continue;
                for (int k = 0; k < n; k++) {
                    int q = it * n + k;
                    dx = g->cx[is] - g->x[q];
                    dy = g->cy[is] - g->y[q];
                    d2 = dx * dx + dy * dy;
                    if (!(d2 <= rsum * rsum)) continue;
                    xtar[l] = g->x[q];
                    ytar[l] = g->y[q];
                    itgt[l] = q;
                    l++;
                    acc_t sx = 0.0, sy = 0.0;
                    for (int j = 0; j < n; j++) {
                        int p = is * n + j;
                        trap_pair(g->x[q] - g->x[p], g->y[q] - g->y[p], g->px[p], g->py[p],
                                  sdx[j], sdy[j], g->speed[p], &sx, &sy);
                    }
                    sx = sx * TWOPI / (double)n / PI;
                    sy = sy * TWOPI / (double)n / PI;
                    acc[2 * no + 2 * it * n + k] -= sx;
                    acc[2 * no + 2 * it * n + n + k] -= sy;
                }
            }
            if (l > 0) {
                orc_stokes_exterior_dlp(n, g->x + is * n, g->y + is * n, sdx, sdy,
                                        g->px + is * n, g->py + is * n, l, xtar, ytar, 1, ux, uy,
                                        NULL, NULL, NULL, NULL);
                for (int k = 0; k < l; k++) {
                    int b = itgt[k] / n, kk = itgt[k] % n;
                    acc[2 * no + 2 * b * n + kk] += ux[k];
                    acc[2 * no + 2 * b * n + n + kk] += uy[k];
                }
            }
        }
        free(rtar);
        free(xtar);
        free(ytar);
        free(itgt);
        free(ux);
        free(uy);
        free(wxt);
        free(wyt);
        free(wtgt);
    }

    /* Diagonal (curvature) term and jump -eta/2 on the wall
     * (stokesSolver.f:932-951) and on every body (:958-987). */
    for (int k = 0; k < no; k++) {
        double tdotden = g->py0[k] * wdx[k] - g->px0[k] * wdy[k];
        acc[k] = acc[k] - g->cur0[k] * tdotden * g->py0[k] * g->speed0[k] / (double)no;
        acc[k + no] = acc[k + no] + g->cur0[k] * tdotden * g->px0[k] * g->speed0[k] / (double)no;
    }
    for (int k = 0; k < no; k++) {
        acc[k] = acc[k] - 0.5 * wdx[k];
        acc[k + no] = acc[k + no] - 0.5 * wdy[k];
    }
    for (int s = 0; s < m; s++) {
        const double *sdx = den + 2 * no + 2 * s * n;
        const double *sdy = sdx + n;
        acc_t *vx = acc + 2 * no + 2 * s * n;
        acc_t *vy = vx + n;
        for (int k = 0; k < n; k++) {
            int q = s * n + k;
            double tdotden = g->py[q] * sdx[k] - g->px[q] * sdy[k];
            vx[k] = vx[k] - g->cur[q] * tdotden * g->py[q] * g->speed[q] / (double)n;
            vy[k] = vy[k] + g->cur[q] * tdotden * g->px[q] * g->speed[q] / (double)n;
        }
        for (int k = 0; k < n; k++) {
            vx[k] = vx[k] - 0.5 * sdx[k];
            vy[k] = vy[k] - 0.5 * sdy[k];
        }
    }

    /* Stokeslets and rotlets at every body centre (stokesSolver.f:992-1049). */
    const double *ext = den + 2 * no + 2 * nbp;
    for (int b = 0; b < m; b++) {
        double sto1 = ext[3 * b], sto2 = ext[3 * b + 1], rot = ext[3 * b + 2];
        for (int t = 0; t < ntgt; t++) {
            double xt, yt;
            acc_t *vx, *vy;
            if (t < no) {
                xt = g->xo[t];
                yt = g->yo[t];
                vx = acc + t;
                vy = acc + t + no;
            } else {
                int tb = (t - no) / n, tk = (t - no) % n;
                xt = g->x[tb * n + tk];
                yt = g->y[tb * n + tk];
                vx = acc + 2 * no + 2 * tb * n + tk;
                vy = vx + n;
            }
            double rx = xt - g->cx[b], ry = yt - g->cy[b];
            double rho2 = rx * rx + ry * ry;
            double rdots = rx * sto1 + ry * sto2;
            double u = 0.5 / TWOPI * (-0.5 * log(rho2) * sto1 + rdots / rho2 * rx);
            double v = 0.5 / TWOPI * (-0.5 * log(rho2) * sto2 + rdots / rho2 * ry);
            u = u + rot * ry / rho2;
            v = v - rot * rx / rho2;
            *vx += u;
            *vy += v;
        }
    }

    /* Constraint rows (stokesSolver.f:1055-1096). */
    for (int b = 0; b < m; b++) {
        const double *sdx = den + 2 * no + 2 * b * n;
        const double *sdy = sdx + n;
        acc_t *c = acc + 2 * no + 2 * nbp + 3 * b;
        for (int k = 0; k < n; k++) {
            int q = b * n + k;
            c[0] = c[0] - sdx[k] * g->speed[q];
            c[1] = c[1] - sdy[k] * g->speed[q];
            c[2] = c[2] - (sdx[k] * g->y[q] - sdy[k] * g->x[q]) * g->speed[q];
        }
        for (int i = 0; i < 3; i++) c[i] = c[i] / TWOPI * (TWOPI / (double)n);
        c[0] += ext[3 * b];
        c[1] += ext[3 * b + 1];
        c[2] += ext[3 * b + 2];
    }

    /* Rank-one completion on the wall rows (stokesSolver.f:1102-1118). */
    acc_t sigdotn = 0.0;
    for (int k = 0; k < no; k++) sigdotn = sigdotn + (g->px0[k] * wdx[k] + g->py0[k] * wdy[k]) * g->speed0[k];
    sigdotn = sigdotn * TWOPI / (double)no;
    for (int k = 0; k < no; k++) {
        acc[k] = acc[k] + sigdotn * g->px0[k];
        acc[k + no] = acc[k + no] + sigdotn * g->py0[k];
    }
    for (size_t i = 0; i < nunk; i++) vel[i] = (double)acc[i];
    free(acc);
}


/* ------------------------------------------------------------------ GMRES */
/* DNRM2 (gmres_routines.f, reference BLAS with scaling). */
static double dnrm2(size_t n, const double *x) {
    if (n < 1) return 0.0;
    if (n == 1) return fabs(x[0]);
    double scale = 0.0, ssq = 1.0;
    for (size_t i = 0; i < n; i++) {
        if (x[i] != 0.0) {
            double a = fabs(x[i]);
            if (scale < a) {
                ssq = 1.0 + ssq * (scale / a) * (scale / a);
                scale = a;
            } else {
                ssq = ssq + (a / scale) * (a / scale);
            }
        }
    }
    return scale * sqrt(ssq);
}
static double ddot(size_t n, const double *a, const double *b) {
    double s = 0.0;
    for (size_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* DHEQR with IJOB>1: apply the previous rotations to the new column and form
 * the next one (gmres_routines.f, DHEQR label 70). H column-major, ld = ldh. */
static int dheqr_update(double *h, int ldh, int nn, double *q) {
    double *col = h + (size_t)(nn - 1) * ldh;
    for (int k = 0; k < nn - 1; k++) {
        double t1 = col[k], t2 = col[k + 1];
        double c = q[2 * k], s = q[2 * k + 1];
        col[k] = c * t1 - s * t2;
        col[k + 1] = s * t1 + c * t2;
    }
    double t1 = col[nn - 1], t2 = col[nn], c, s;
    if (t2 == 0.0) {
        c = 1.0;
        s = 0.0;
    } else if (fabs(t2) >= fabs(t1)) {
        double t = t1 / t2;
        s = -1.0 / sqrt(1.0 + t * t);
        c = -s * t;
    } else {
        double t = t2 / t1;
        c = 1.0 / sqrt(1.0 + t * t);
        s = -c * t;
    }
    q[2 * nn - 2] = c;
    q[2 * nn - 1] = s;
    col[nn - 1] = c * t1 - s * t2;
    return col[nn - 1] == 0.0 ? nn : 0;
}

/* DHELS (gmres_routines.f:1048-): rotate the rhs, back-substitute. */
static void dhels(const double *h, int ldh, int nn, const double *q, double *b) {
    for (int k = 0; k < nn; k++) {
        double c = q[2 * k], s = q[2 * k + 1];
        double t1 = b[k], t2 = b[k + 1];
        b[k] = c * t1 - s * t2;
        b[k + 1] = s * t1 + c * t2;
    }
    for (int kb = 0; kb < nn; kb++) {
        int k = nn - 1 - kb;
        b[k] = b[k] / h[(size_t)k * ldh + k];
        double t = -b[k];
        for (int i = 0; i < k; i++) b[i] += t * h[(size_t)k * ldh + i];
    }
}

/* DGMRES/DPIGMR as solveBIE configures them (JPRE=0, JSCAL=0, KMP=MAXL,
 * NRMAX=0, ITOL=0): gmres_routines.f:380-554, 1524-1963, DORTH 1397-1522. */
int orc_gmres(const orc_geom *g, int ibary, const double *b, double *x, int maxl, double tol,
              double *err, int *ierr) {
    const size_t n = (size_t)(2 * g->no + 2 * g->n * g->m + 3 * g->m);
    if (maxl == 0) maxl = 10;
    if ((size_t)maxl > n) maxl = (int)n;
    const int ldh = maxl + 1;
    double *V = xmalloc(sizeof(double) * n * (size_t)(maxl + 1));
    double *H = xcalloc((size_t)ldh * (size_t)maxl, sizeof(double));
    double *q = xcalloc((size_t)(2 * maxl), sizeof(double));
    double *r = xmalloc(sizeof(double) * n);
    double bnrm = dnrm2(n, b);
    orc_matvec(g, ibary, x, r);
    for (size_t i = 0; i < n; i++) r[i] = b[i] - r[i];
    double r0nrm = dnrm2(n, r);
    memcpy(V, r, sizeof(double) * n);
    *ierr = 0;
    *err = r0nrm / bnrm;
    if (*err <= tol) {
        free(V); free(H); free(q); free(r);
        return 0;
    }
    double tem = 1.0 / r0nrm;
    for (size_t i = 0; i < n; i++) V[i] *= tem;
    double prod = 1.0, rho = r0nrm;
    int ll = 0, lgmr = 0, iflag = 1, converged = 0;
    for (ll = 1; ll <= maxl; ll++) {
        lgmr = ll;
        double *w = V + (size_t)ll * n;
        orc_matvec(g, ibary, V + (size_t)(ll - 1) * n, w);
        /* DORTH */
        double *hcol = H + (size_t)(ll - 1) * ldh;
        double vnrm = dnrm2(n, w);
        for (int i = 0; i < ll; i++) {
            const double *vi = V + (size_t)i * n;
            hcol[i] = ddot(n, vi, w);
            double t = -hcol[i];
            for (size_t k = 0; k < n; k++) w[k] += t * vi[k];
        }
        double snormw = dnrm2(n, w);
        if (!(vnrm + 0.001 * snormw != vnrm)) {
            double sumdsq = 0.0;
            for (int i = 0; i < ll; i++) {
                const double *vi = V + (size_t)i * n;
                double t = -ddot(n, vi, w);
                if (hcol[i] + 0.001 * t == hcol[i]) continue;
                hcol[i] = hcol[i] - t;
                for (size_t k = 0; k < n; k++) w[k] += t * vi[k];
                sumdsq += t * t;
            }
            if (sumdsq != 0.0) {
                double arg = snormw * snormw - sumdsq;
                snormw = sqrt(arg > 0.0 ? arg : 0.0);
            }
        }
        hcol[ll] = snormw;
        int info = dheqr_update(H, ldh, ll, q);
        if (info == ll) {
            iflag = 2;
            break;
        }
        prod = prod * q[2 * ll - 1];
        rho = fabs(prod * r0nrm);
        *err = rho / bnrm;
        if (*err <= tol) {
            converged = 1;
            break;
        }
        if (ll == maxl) break;
        tem = 1.0 / snormw;
        for (size_t k = 0; k < n; k++) w[k] *= tem;
    }
    if (ll > maxl) ll = maxl;
    if (!converged && iflag != 2) iflag = (rho < r0nrm) ? 1 : 2;
    if (converged || iflag == 1) {
        double *y = xcalloc((size_t)(lgmr + 1), sizeof(double));
        y[0] = r0nrm;
        dhels(H, ldh, lgmr, q, y);
        for (int i = 0; i < lgmr; i++) {
            const double *vi = V + (size_t)i * n;
            for (size_t k = 0; k < n; k++) x[k] += y[i] * vi[k];
        }
        free(y);
        *ierr = converged ? 0 : 1;
    } else {
        *ierr = 2;
    }
    free(V);
    free(H);
    free(q);
    free(r);
    return lgmr;
}

void orc_stokessolver(int n, int m, int no, int ibary, int ibc, int maxl, double tol,
                      const double *xx, const double *yy, double *den, int *iter) {
    orc_geom *g = orc_geom_create(n, m, no, xx, yy);
    size_t nunk = (size_t)(2 * no + 2 * n * m + 3 * m);
    double *rhs = xmalloc(sizeof(double) * nunk);
    orc_rhs(g, ibc, rhs);
    double err;
    int ierr;
    *iter = orc_gmres(g, ibary, rhs, den, maxl, tol, &err, &ierr);
    free(rhs);
    orc_geom_free(g);
}

/* ------------------------------------------------------------ shear stress */
/* Odd-even self term of deformation_on_boundary (stokesSolver.f:2744-2835). */
static void self_odd_even(const orc_geom *g, int s, const double *sdx, const double *sdy,
                          int k, int j0, double *e11, double *e12, double *e22) {
    const int n = g->n;
    int qk = s * n + k;
    for (int j = j0; j < n; j += 2) {
        int qj = s * n + j;
        double rx = g->x[qk] - g->x[qj];
        double ry = g->y[qk] - g->y[qj];
        double rho2 = rx * rx + ry * ry;
        double rdotn = rx * g->px[qj] + ry * g->py[qj];
        double routn = rx * g->py[qj] + ry * g->px[qj];
        double sx = sdx[j] - sdx[k];
        double sy = sdy[j] - sdy[k];
        double rdots = rx * sx + ry * sy;
        double routs = rx * sy + ry * sx;
        double rho6 = rho2 * rho2 * rho2;
        double w = g->speed[qj] * TWOPI / (double)n / PI;
        e11[qk] = e11[qk] + 2.0 * 0.5 *
                  (2.0 * rdotn * rdots / rho2 / rho2 + 2.0 * rdots / rho2 / rho2 * rx * g->px[qj] +
                   2.0 * rdotn / rho2 / rho2 * rx * sx - 8.0 * rdotn * rdots / rho6 * rx * rx) * w;
        e12[qk] = e12[qk] + 2.0 * 0.5 *
                  (rdots * routn / rho2 / rho2 + rdotn * routs / rho2 / rho2 -
                   8.0 * rdotn * rdots * rx * ry / rho6) * w;
        e22[qk] = e22[qk] + 2.0 * 0.5 *
                  (2.0 * rdotn * rdots / rho2 / rho2 + 2.0 * rdots / rho2 / rho2 * ry * g->py[qj] +
                   2.0 * rdotn / rho2 / rho2 * ry * sy - 8.0 * rdotn * rdots / rho6 * ry * ry) * w;
    }
}

/* deformation_on_boundary (stokesSolver.f:2618-2998). */
void orc_deformation(const orc_geom *g, const double *den, double *e11, double *e12,
                     double *e22) {
    const int n = g->n, m = g->m, no = g->no;
    const int nbp = n * m;
    memset(e11, 0, sizeof(double) * (size_t)nbp);
    memset(e12, 0, sizeof(double) * (size_t)nbp);
    memset(e22, 0, sizeof(double) * (size_t)nbp);
    double *ux = xmalloc(sizeof(double) * (size_t)n);
    double *uy = xmalloc(sizeof(double) * (size_t)n);
    double *u1x = xmalloc(sizeof(double) * (size_t)n);
    double *u1y = xmalloc(sizeof(double) * (size_t)n);
    double *u2x = xmalloc(sizeof(double) * (size_t)n);
    double *u2y = xmalloc(sizeof(double) * (size_t)n);

    /* wall -> every body, barycentric interior rule (:2707-2727) */
    for (int it = 0; it < m; it++) {
        orc_stokes_interior_dlp(no, g->xo, g->yo, den, den + no, g->px0, g->py0, n,
                                g->x + it * n, g->y + it * n, 2, ux, uy, u1x, u1y, u2x, u2y);
        for (int k = 0; k < n; k++) {
            int q = it * n + k;
            e11[q] = e11[q] + u1x[k];
            e12[q] = e12[q] + 0.5 * (u1y[k] + u2x[k]);
            e22[q] = e22[q] + u2y[k];
        }
    }
    for (int is = 0; is < m; is++) {
        const double *sdx = den + 2 * no + 2 * is * n;
        const double *sdy = sdx + n;
        /* self, odd-even (1-based odd targets use even sources and back) */
#pragma omp parallel for OMP_NT schedule(static)
        for (int k = 0; k < n; k++) self_odd_even(g, is, sdx, sdy, k, (k % 2 == 0) ? 1 : 0, e11, e12, e22);
        /* every other body, barycentric exterior rule (:2893-2925) */
        for (int it = 0; it < m; it++) {
            if (it == is) continue;
            orc_stokes_exterior_dlp(n, g->x + is * n, g->y + is * n, sdx, sdy, g->px + is * n,
                                    g->py + is * n, n, g->x + it * n, g->y + it * n, 2, ux, uy,
                                    u1x, u1y, u2x, u2y);
            for (int k = 0; k < n; k++) {
                int q = it * n + k;
                e11[q] = e11[q] + u1x[k];
                e12[q] = e12[q] + 0.5 * (u1y[k] + u2x[k]);
                e22[q] = e22[q] + u2y[k];
            }
        }
    }
    /* jump along the diagonal (:2933-2958) */
    cplx *zd = xmalloc(sizeof(cplx) * (size_t)n);
    for (int b = 0; b < m; b++) {
        const double *sdx = den + 2 * no + 2 * b * n;
        const double *sdy = sdx + n;
        for (int k = 0; k < n; k++) zd[k] = sdx[k] + I * sdy[k];
        fourier_diff(n, zd);
        for (int k = 0; k < n; k++) {
            int q = b * n + k;
            double tx = -g->py[q], ty = g->px[q];
            double dsdtx = creal(zd[k]) / g->speed[q];
            double dsdty = cimag(zd[k]) / g->speed[q];
            double dd = dsdtx * tx + dsdty * ty;
            e11[q] = e11[q] + 0.5 * dd * (tx * tx - ty * ty);
            e12[q] = e12[q] + 0.5 * dd * 2.0 * tx * ty;
            e22[q] = e22[q] + 0.5 * dd * (-tx * tx + ty * ty);
        }
    }
    free(zd);
    /* stokeslets and rotlets (:2964-2994) */
    const double *ext = den + 2 * no + 2 * nbp;
    for (int b = 0; b < m; b++) {
        double sto1 = ext[3 * b], sto2 = ext[3 * b + 1], rot = ext[3 * b + 2];
        for (int q = 0; q < nbp; q++) {
            double rx = g->x[q] - g->cx[b], ry = g->y[q] - g->cy[b];
            double rho2 = rx * rx + ry * ry;
            double rdots = rx * sto1 + ry * sto2;
            e11[q] = e11[q] + 0.5 / TWOPI * (rdots / rho2 - 2 * rx * rx * rdots / rho2 / rho2);
            e12[q] = e12[q] - 1.0 / TWOPI * (rdots / rho2 / rho2 * rx * ry);
            e22[q] = e22[q] + 0.5 / TWOPI * (rdots / rho2 - 2 * ry * ry * rdots / rho2 / rho2);
            e11[q] = e11[q] - 2.0 * rot * rx * ry / rho2 / rho2;
            e12[q] = e12[q] + rot * (rx * rx - ry * ry) / rho2 / rho2;
            e22[q] = e22[q] + 2.0 * rot * rx * ry / rho2 / rho2;
        }
    }
    free(ux);
    free(uy);
    free(u1x);
    free(u1y);
    free(u2x);
    free(u2y);
}

/* classifyPointsOld (stokesSolver.f:4420-4500): 1 if the target lies on the
 * fluid side of its closest body node (first minimum in body-major order). */
static int classify_point(const orc_geom *g, double xt, double yt) {
    int n = g->n, m = g->m;
    int best = 0;
    double dmin = 1.0e10;
    for (int j = 0; j < m; j++)
        for (int k = 0; k < n; k++) {
            int q = j * n + k;
            double d2 = (xt - g->x[q]) * (xt - g->x[q]) + (yt - g->y[q]) * (yt - g->y[q]);
            if (d2 < dmin) {
                dmin = d2;
                best = q;
            }
        }
    double vdn = (xt - g->x[best]) * g->px[best] + (yt - g->y[best]) * g->py[best];
    return vdn < 0.0 ? 1 : 0;
}

/* computeQoiTargets (stokesSolver.f:3099-3323). */
void orc_qoi_targets(int n, int m, int no, int ibary, const double *xx, const double *yy,
                     const double *den, int ntar, const double *xt, const double *yt, double *u,
                     double *v, double *p, double *vort) {
    orc_geom *g = orc_geom_create(n, m, no, xx, yy);
    for (int j = 0; j < ntar; j++) u[j] = v[j] = p[j] = vort[j] = 0.0;
    double *t6 = xcalloc((size_t)(6 * (ntar > 0 ? ntar : 1)), sizeof(double));
    double *ut = t6, *vt = t6 + ntar, *uxt = t6 + 2 * ntar, *uyt = t6 + 3 * ntar,
           *vxt = t6 + 4 * ntar, *vyt = t6 + 5 * ntar;
    /* one source curve: trapezoid pressure (+ u, v, vorticity if ibary=0) */
    for (int c = 0; c <= m; c++) {
        int len = c == 0 ? no : n;
        const double *xs = c == 0 ? g->xo : g->x + (c - 1) * n;
        const double *ys = c == 0 ? g->yo : g->y + (c - 1) * n;
        const double *nx = c == 0 ? g->px0 : g->px + (c - 1) * n;
        const double *ny = c == 0 ? g->py0 : g->py + (c - 1) * n;
        const double *sp = c == 0 ? g->speed0 : g->speed + (c - 1) * n;
        const double *dx = c == 0 ? den : den + 2 * no + 2 * (c - 1) * n;
        const double *dy = c == 0 ? den + no : dx + n;
        for (int j = 0; j < ntar; j++) {
            for (int k = 0; k < len; k++) {
                double rx = xt[j] - xs[k], ry = yt[j] - ys[k];
                double rho2 = rx * rx + ry * ry;
                double rdotn = rx * nx[k] + ry * ny[k];
                double rdotden = rx * dx[k] + ry * dy[k];
                double dendotn = nx[k] * dx[k] + ny[k] * dy[k];
                double routn = -rx * ny[k] + ry * nx[k];
                double routden = -rx * dy[k] + ry * dx[k];
                double w = sp[k] * TWOPI / (double)len / PI;
                p[j] = p[j] + (2.0 * rdotn * rdotden / rho2 / rho2 - dendotn / rho2) * w;
                if (ibary != 1) {
                    u[j] = u[j] + rdotn / rho2 * rdotden / rho2 * rx * w;
                    v[j] = v[j] + rdotn / rho2 * rdotden / rho2 * ry * w;
                    vort[j] = vort[j] + (rdotden * routn + rdotn * routden) / (rho2 * rho2) * w;
                }
            }
        }
        if (ibary == 1) {
            if (c == 0)
                orc_stokes_interior_dlp(no, xs, ys, dx, dy, nx, ny, ntar, xt, yt, 2, ut, vt, uxt,
                                        uyt, vxt, vyt);
            else
                orc_stokes_exterior_dlp(n, xs, ys, dx, dy, nx, ny, ntar, xt, yt, 2, ut, vt, uxt,
                                        uyt, vxt, vyt);
            for (int j = 0; j < ntar; j++) {
                u[j] += ut[j];
                v[j] += vt[j];
                vort[j] += vxt[j] - uyt[j];
            }
        }
    }
    if (m > 0) {
        const double *ext = den + 2 * no + 2 * n * m;
        for (int b = 0; b < m; b++) {
            double sto1 = ext[3 * b], sto2 = ext[3 * b + 1], rot = ext[3 * b + 2];
            for (int j = 0; j < ntar; j++) {
                double rx = xt[j] - g->cx[b], ry = yt[j] - g->cy[b];
                double rho2 = rx * rx + ry * ry;
                double rdots = rx * sto1 + ry * sto2;
                u[j] = u[j] + 0.5 / TWOPI * (-0.5 * log(rho2) * sto1 + rdots / rho2 * rx);
                v[j] = v[j] + 0.5 / TWOPI * (-0.5 * log(rho2) * sto2 + rdots / rho2 * ry);
                p[j] = p[j] + 1.0 / TWOPI * rdots / rho2;
                vort[j] = vort[j] + (ry * sto1 - rx * sto2) / rho2 / TWOPI;
                u[j] = u[j] + rot * ry / rho2;
                v[j] = v[j] - rot * rx / rho2;
            }
        }
        for (int j = 0; j < ntar; j++)
            if (classify_point(g, xt[j], yt[j]) == 0) u[j] = v[j] = p[j] = vort[j] = 0.0;
    }
    free(t6);
    orc_geom_free(g);
}

/* computeShearStress, ib=1 (stokesSolver.f:3002-3096):
 * traction t = -2 E n, tau = t_x p_y - t_y p_x. */
void orc_shear_stress(int n, int m, int no, const double *xx, const double *yy,
                      const double *den, double *tau) {
    orc_geom *g = orc_geom_create(n, m, no, xx, yy);
    size_t nbp = (size_t)n * (size_t)m;
    double *e11 = xmalloc(sizeof(double) * nbp);
    double *e12 = xmalloc(sizeof(double) * nbp);
    double *e22 = xmalloc(sizeof(double) * nbp);
    orc_deformation(g, den, e11, e12, e22);
    for (size_t q = 0; q < nbp; q++) {
        double tx = -2.0 * e11[q] * g->px[q] - 2.0 * e12[q] * g->py[q];
        double ty = -2.0 * e12[q] * g->px[q] - 2.0 * e22[q] * g->py[q];
        tau[q] = tx * g->py[q] - ty * g->px[q];
    }
    free(e11);
    free(e12);
    free(e22);
    orc_geom_free(g);
}

/* computePressure (stokesSolver.f:4644-4882): pressure on every body point.
 * Kernel of a source node j at a target: (2 (r.n_j)(r.s_j)/rho^4 -
 * (n_j.s_j)/rho^2) speed_j 2pi/N / pi, with s the source density; the wall
 * and other bodies by the trapezoid rule (:4693-4720, :4789-4819), the
 * target's own body by the odd-even rule on s_j = eta_j - eta_k with the
 * factor 2 (:4731-4787), then the jump -(d eta/ds . t) with d/ds =
 * fourierDiff / speed, t = (-p_y, p_x) (:4826-4854), then the stokeslet
 * pressure (r.lambda)/(2pi rho^2) of every body (:4857-4879; rotlets carry
 * none). */
static double pressure_term(double rx, double ry, double nx, double ny, double sx, double sy,
                            double w) {
    double rho2 = rx * rx + ry * ry;
    double rdotn = rx * nx + ry * ny;
    double rdots = rx * sx + ry * sy;
    double sdotn = nx * sx + ny * sy;
    return (2.0 * rdotn * rdots / rho2 / rho2 - sdotn / rho2) * w;
}

void orc_pressure(int n, int m, int no, const double *xx, const double *yy, const double *den,
                  double *pressure) {
    orc_geom *g = orc_geom_create(n, m, no, xx, yy);
    const int nbp = n * m;
#pragma omp parallel for OMP_NT schedule(dynamic, 16)
    for (int q = 0; q < nbp; q++) {
        const int tb = q / n, tk = q % n;
        const double xt = g->x[q], yt = g->y[q];
        double p = 0.0;
        for (int j = 0; j < no; j++)
            p += pressure_term(xt - g->xo[j], yt - g->yo[j], g->px0[j], g->py0[j], den[j],
                               den[j + no], g->speed0[j] * TWOPI / (double)no / PI);
        for (int s = 0; s < m; s++) {
            const double *sdx = den + 2 * no + 2 * s * n;
            const double *sdy = sdx + n;
            if (s == tb) {
                /* odd-even: sources of the other parity, density minus the target's */
                for (int j = (tk % 2 == 0) ? 1 : 0; j < n; j += 2) {
                    const int qs = s * n + j;
                    p += 2.0 * pressure_term(xt - g->x[qs], yt - g->y[qs], g->px[qs], g->py[qs],
                                             sdx[j] - sdx[tk], sdy[j] - sdy[tk],
                                             g->speed[qs] * TWOPI / (double)n / PI);
                }
            } else {
                for (int j = 0; j < n; j++) {
                    const int qs = s * n + j;
                    p += pressure_term(xt - g->x[qs], yt - g->y[qs], g->px[qs], g->py[qs], sdx[j],
                                       sdy[j], g->speed[qs] * TWOPI / (double)n / PI);
                }
            }
        }
        pressure[q] = p;
    }
    cplx *z = xmalloc(sizeof(cplx) * (size_t)n);
    for (int b = 0; b < m; b++) {
        const double *sdx = den + 2 * no + 2 * b * n;
        const double *sdy = sdx + n;
        for (int k = 0; k < n; k++) z[k] = sdx[k] + I * sdy[k];
        fourier_diff(n, z);
        for (int k = 0; k < n; k++) {
            const int q = b * n + k;
            const double tx = -g->py[q], ty = g->px[q];
            const double dsx = creal(z[k]) / g->speed[q], dsy = cimag(z[k]) / g->speed[q];
            pressure[q] = pressure[q] - 1.0 * (dsx * tx + dsy * ty);
        }
    }
    free(z);
    const double *ext = den + 2 * no + 2 * nbp;
    for (int b = 0; b < m; b++) {
        const double sto1 = ext[3 * b], sto2 = ext[3 * b + 1];
        for (int q = 0; q < nbp; q++) {
            const double rx = g->x[q] - g->cx[b], ry = g->y[q] - g->cy[b];
            const double rho2 = rx * rx + ry * ry;
            pressure[q] = pressure[q] + 1.0 / TWOPI * (rx * sto1 + ry * sto2) / rho2;
        }
    }
    orc_geom_free(g);
}
