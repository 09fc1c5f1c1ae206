/*
 * stokes_oracle.h -- CPU restatement (plain C, FP64) of the reference's
 * per-step Stokes solve, ErosionPackage/fortran_code/stokesSolver.f and
 * ErosionPackage/fortran_code/gmres_routines.f.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP library.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library never links, loads or calls it.
 *
 * Layout conventions are the reference's:
 *   xx, yy            body points, body-major (body b at b*n .. b*n+n-1)
 *   den / vel [N_unk] [etax wall No | etay wall No |
 *                      per body (etax n | etay n) | per body (lx, ly, xi)]
 *                     N_unk = 2*No + 2*m*n + 3*m     (stokesSolver.f:631-657)
 *
 * Far field: the reference evaluates the trapezoid double-layer sum with an
 * FMM of tolerance 0.5e-12 (stokesSolver.f:662); this restatement sums it
 * directly in the order of matvec_DLP (stokesSolver.f:1563-1584).
 */
#ifndef STOKES_ORACLE_H
#define STOKES_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_geom orc_geom;

/* Geometry of bodies + fixed outer wall (inner_geometry / outer_geometry,
 * stokesSolver.f:136-204). */
orc_geom *orc_geom_create(int n, int m, int no, const double *xx, const double *yy);
void orc_geom_free(orc_geom *g);
/* which: 0 px 1 py 2 cur 3 speed 4 cx 5 cy 6 xo 7 yo 8 px0 9 py0 10 cur0 11 speed0 */
const double *orc_geom_array(const orc_geom *g, int which);

/* Overwrite array `which` (as orc_geom_array, same length) with v; returns
 * 0 or -1.  Test hook: evaluate on another implementation's geometry. */
int orc_geom_set_array(orc_geom *g, int which, const double *v);

/* fourierDiff (stokesSolver.f:348-388) on a complex vector stored re/im. */
void orc_fourier_diff(int n, double *re, double *im);

/* bd_condition (stokesSolver.f:208-314). */
void orc_rhs(const orc_geom *g, int ibc, double *rhs);

/* Completed double-layer operator.  ibary=1: matvec_DLP_fmmbary
 * (stokesSolver.f:570-1124) with the far field summed directly; ibary=0:
 * matvec_DLP (stokesSolver.f:1519-1912). */
void orc_matvec(const orc_geom *g, int ibary, const double *den, double *vel);

/* StokesExteriorDLP / StokesInteriorDLP (stokesSolver.f:5233-5379,
 * 4937-5100).  px1/py1 are the source curve's own normals. */
void orc_stokes_exterior_dlp(int npts, const double *xs, const double *ys,
                             const double *dx, const double *dy,
                             const double *px1, const double *py1, int ntar,
                             const double *xt, const double *yt, int nder,
                             double *ux, double *uy, double *u1x, double *u1y,
                             double *u2x, double *u2y);
void orc_stokes_interior_dlp(int npts, const double *xs, const double *ys,
                             const double *dx, const double *dy,
                             const double *px1, const double *py1, int ntar,
                             const double *xt, const double *yt, int nder,
                             double *ux, double *uy, double *u1x, double *u1y,
                             double *u2x, double *u2y);

/* DGMRES as configured by solveBIE (stokesSolver.f:481-497): no
 * preconditioner, no restart, MGS (DORTH), Givens (DHEQR), implicit residual
 * stop (ISDGMR ITOL=0).  Returns ITER; *ierr as DGMRES IERR. */
int orc_gmres(const orc_geom *g, int ibary, const double *rhs, double *x,
              int maxl, double tol, double *err, int *ierr);

/* stokesSolver (stokesSolver.f:1-131) with the GMRES tolerance explicit
 * (the reference hard-codes 1e-8 at stokesSolver.f:482). */
void orc_stokessolver(int n, int m, int no, int ibary, int ibc, int maxl,
                      double tol, const double *xx, const double *yy,
                      double *den, int *iter);

/* deformation_on_boundary (stokesSolver.f:2618-2998). */
void orc_deformation(const orc_geom *g, const double *den, double *e11,
                     double *e12, double *e22);

/* computeShearStress with ib=1 (stokesSolver.f:3002-3096). */
void orc_shear_stress(int n, int m, int no, const double *xx, const double *yy,
                      const double *den, double *tau);

/* computeQoiTargets (stokesSolver.f:3099-3323) with classifyPointsOld
 * (:4420-4500): velocity, pressure, vorticity at targets. */
void orc_qoi_targets(int n, int m, int no, int ibary, const double *xx, const double *yy,
                     const double *den, int ntar, const double *xt, const double *yt, double *u,
                     double *v, double *p, double *vort);

/* computePressure (stokesSolver.f:4644-4882): pressure on every body point. */
void orc_pressure(int n, int m, int no, const double *xx, const double *yy, const double *den,
                  double *pressure);

/* Number of OpenMP threads the oracle uses (0 = runtime default). */
void orc_set_threads(int nthreads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
