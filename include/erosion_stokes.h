/*
 * erosion_stokes.h -- C-ABI of libstokes_hip.so, the MI355X (gfx950) Stokes
 * solver of the Erosion package.
 *
 * The two Fortran-convention entry points have exactly the signatures the
 * reference's libstokes.so exports and that ErosionPackage/src/DensityStress.jl
 * binds with ccall: lower-case name with a trailing underscore, every
 * argument by pointer, integers as Julia's Ref{Int} (64-bit).  A caller that
 * swaps the library path (DensityStress.jl:19) needs no other change.
 *
 * Ownership: the caller allocates every array; the library keeps no caller
 * pointer after a call returns.  Device buffers are library-owned and reused
 * across calls.  Calls are synchronous and not reentrant (the reference keeps
 * its state in COMMON blocks and is not reentrant either).
 *
 * Errors: the reference has no error channel.  These entry points never abort
 * the host process; a failure prints one line to stderr, leaves the outputs
 * as described per function and is reported by erosion_last_status() /
 * erosion_last_error().
 *
 * Array layouts (stokesSolver.f:631-657):
 *   xx, yy [npts*nbods]   body points, body-major
 *   den  [2*nouter + 2*npts*nbods + 3*nbods]
 *        [eta_x wall | eta_y wall | per body (eta_x | eta_y) | per body
 *         (stokeslet_x, stokeslet_y, rotlet)]
 *   tau  [npts*nbods]     shear stress, body-major
 */
#ifndef EROSION_STOKES_H
#define EROSION_STOKES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Density of the completed double-layer boundary integral equation.
 * Replaces subroutine stokesSolver (ErosionPackage/fortran_code/stokesSolver.f:1-131),
 * bound at ErosionPackage/src/DensityStress.jl:48-51.
 *   den   in: initial guess (Julia passes zeros); out: solution
 *   nits  out: GMRES iterations
 * GMRES stops when the implicit residual estimate is <= tol * ||b||, tol as
 * set by erosion_set_gmres_tol (default 1e-8, the reference's hard-coded value,
 * stokesSolver.f:482).  ibary=1 selects the near-pair barycentric operator of
 * matvec_DLP_fmmbary, ibary=0 the trapezoid operator of matvec_DLP.  With
 * ifmm=1 the far field is the cluster expansion or the direct sums, whichever
 * the autotuner measures faster (erosion_set_farfield_method); with ifmm=0
 * (and ibary=0) it is summed directly.  The all-pairs-barycentric operator of
 * ifmm=0, ibary=1 (matvec_DLP_bary) is not provided and that combination is
 * reported as an error (den untouched, nits = 0). */
void stokessolver_(const int64_t* npts, const int64_t* nbods, const int64_t* nouter,
                   const int64_t* ifmm, const int64_t* ibary, const int64_t* ibc,
                   const int64_t* maxl, const double* xx, const double* yy, double* den,
                   int64_t* nits);

/* Shear stress on every body point.
 * Replaces subroutine computeShearStress with ib=1
 * (ErosionPackage/fortran_code/stokesSolver.f:3002-3096, deformation_on_boundary
 * :2618-2998), bound at ErosionPackage/src/DensityStress.jl:195-198.
 * ibary=0 (the trapezoid variant deformation_on_boundary_trap, :2305-2614) is
 * not provided and is reported as an error (tau zeroed).  The barycentric
 * rule covers the wall and the bodies near each target body; the other
 * bodies' boundary-value integrals, to which that rule converges away from a
 * curve, come from the cluster expansions (32 or more bodies, far-field
 * method 1). */
void computeshearstress_(const int64_t* npts, const int64_t* nbods, const int64_t* nouter,
                         const double* xx, const double* yy, const double* den,
                         const int64_t* ibary, double* tau);

/* Velocity, pressure and vorticity at target points.
 * Replaces subroutine computeQoiTargets (ErosionPackage/fortran_code/stokesSolver.f:3099-3323,
 * inside/outside test classifyPointsOld :4420-4500), bound at
 * ErosionPackage/src/DensityStress.jl:81-86 and called four times per erosion
 * step when the pressure drop is held fixed (getpdrop, DensityStress.jl:129-154).
 * ibary=1: velocity and vorticity by the barycentric rules from every curve,
 * pressure by the trapezoid rule; ibary=0: trapezoid for all four.  Targets
 * inside a body get zeros.  On error all outputs are zeroed. */
void computeqoitargets_(const int64_t* npts, const int64_t* nbods, const int64_t* nouter,
                        const int64_t* ibary, const double* xx, const double* yy,
                        const double* den, const int64_t* ntargets, const double* xtar,
                        const double* ytar, double* utar, double* vtar, double* ptar,
                        double* vortar);

/* Pressure on every body point.
 * Replaces subroutine computePressure (ErosionPackage/fortran_code/stokesSolver.f:4644-4882),
 * bound at ErosionPackage/src/PostProcess.jl:107-110 (post-processing: drag).
 * The wall and other bodies by the trapezoid rule, the point's own body by
 * the odd-even rule on the density minus its value at the point, the jump
 * term and the stokeslet pressure.  pressure[npts*nbods], body-major; zeroed
 * on error. */
void computepressure_(const int64_t* npts, const int64_t* nbods, const int64_t* nouter,
                      const double* xx, const double* yy, const double* den, double* pressure);

/* ---- Additions with no counterpart in the reference ------------------------
 * The reference fixes the GMRES tolerance in source (stokesSolver.f:482) and
 * has no status channel; these let a caller run at another tolerance and read
 * what happened without changing the two signatures above. */

/* GMRES relative tolerance for stokessolver_; returns 0, or -1 if tol <= 0. */
int erosion_set_gmres_tol(double tol);
double erosion_get_gmres_tol(void);

/* 0 ok; 1 GMRES stopped at maxl without meeting the tolerance; 2 GMRES
 * breakdown; -1 error (see erosion_last_error). */
int erosion_last_status(void);
const char* erosion_last_error(void);

/* Record of the last stokessolver_ call:
 * info[0] iterations, info[1] implicit relative residual, info[2] true relative
 * residual ||b - A x|| / ||b|| of the returned density, info[3] seconds,
 * info[4] re-orthogonalised iterations, info[5] far-field targets per lane,
 * info[6] far-field source splits chosen by the autotuner, info[7] the
 * preconditioner kind, info[8] the iterations whose far field ran in FP32,
 * info[9] 1 if the returned density's true residual is above the tolerance
 * (0 otherwise), info[10] the number of such solves since the library was
 * loaded, info[11] the GMRES cycles run (1 + restarts) and info[12] the
 * relative residual ||b - A x0|| / ||b|| of the initial guess.  The true residual
 * is formed with compensated sums (error-free transformations in the far
 * field and the assembly), and a solve whose true residual misses the
 * tolerance restarts from its density, aiming at half the tolerance.
 * info[13] says why the last solve stopped (0 it met the tolerance, 1 a
 * restart did not halve the true residual, 2 the cycle limit of 4, 3 maxl
 * iterations spent, 4 a cycle ended with GMRES status 1 or 2),
 * info[14..17] the true relative residual after each of its cycles and
 * info[18..21] each cycle's iterations (0 for cycles not run) and info[22],
 * for a solve above the tolerance, the FP64 floor of its density: ||A d|| /
 * ||b|| for a half-ulp perturbation d of the returned density (random signs),
 * the residual that rounding the density to FP64 alone can leave (0 when the
 * tolerance was met); info[23..32] hold the same 10 values for the latest
 * solve that ended above the tolerance, and info[33] counts the solves above
 * the tolerance (of info[10]) whose FP64 floor is at least half the residual
 * they reached.  `info` holds at least 34 doubles.  Returns 0. */
int erosion_last_solve_info(double* info);

/* Initial guess of stokessolver_ (no counterpart in the reference, whose
 * solves start from the den the caller passes, DensityStress.jl:45).
 * space 0 (default): den holds the density x0.  space 1: den holds u0 in the
 * preconditioned variable u = P x of the right preconditioner (x0 = P^{-1}
 * u0; u = x without a preconditioner).  The density out is x in both cases.
 * Returns 0, or -1 for another space. */
int erosion_set_guess_space(int space);
int erosion_get_guess_space(void);
/* After a solve with guess space 1: its final u (n = N_unk values) into u.
 * Returns 0, or -1 when the last solve did not track u or n differs. */
int erosion_last_preconditioned(double* u, int64_t n);

/* One application of the completed operator vel = A den, as matvec_DLP_fmmbary
 * (ibary=1, stokesSolver.f:570-1124, far field summed directly) or matvec_DLP
 * (ibary=0, stokesSolver.f:1519-1912).  Returns 0 or -1. */
int erosion_matvec(int64_t npts, int64_t nbods, int64_t nouter, int64_t ibary,
                   const double* xx, const double* yy, const double* den, double* vel);

/* Test hook (no counterpart in the reference): the geometry the library
 * builds from the boundary (inner_geometry / outer_geometry,
 * stokesSolver.f:136-204) -- out[0..npts) unit normal x, [npts..2npts) unit
 * normal y, then curvature and speed (npts each; wall points first, then the
 * bodies body-major), then the body centres x (nbods) and y (nbods);
 * npts = nouter + npts*nbods here.  Returns 0 or -1. */
int erosion_geometry(int64_t npts, int64_t nbods, int64_t nouter, const double* xx,
                     const double* yy, double* out);

/* Measurement hook: set up the geometry once, then time `reps` operator
 * applications with HIP events on the library's stream.
 * out[0] ms per application, out[1] ms per far-field kernel launch, out[2]
 * ordered source-target pairs per far-field launch, out[3] near-pair entries,
 * out[4] ms per near-pair evaluation (gather) pass, boundary values
 * excluded, out[5] far-field source splits, out[6] algorithmic bytes of one
 * gather pass (per near entry, the source curve's nodes x 14 doubles plus 16
 * bytes written).  `out` holds 7 doubles.  Returns 0 or -1. */
int erosion_time_matvec(int64_t npts, int64_t nbods, int64_t nouter, int64_t ibary,
                        const double* xx, const double* yy, const double* den, int64_t reps,
                        double* out);

/* Measurement hook: enable (n >= 1) or stop (0) recording HIP events around
 * the far-field, near-pair, stress-evaluation, Gram-Schmidt and
 * preconditioner launches of every later call; enabling clears the totals.
 * With n > 1 the launches of a GMRES iteration are timed in one iteration of
 * every n, each standing for n (an event record costs the stream a few
 * microseconds: timing every iteration slows a solve by about a tenth);
 * launches outside the iterations are always timed.  Events are read back
 * lazily, so timing adds no host synchronisation.  Returns 0 or -1. */
int erosion_kernel_timing(int enable);

/* Totals since timing was enabled (sampled launches weighted by their period,
 * so that out[3i] / out[3i+1] is the measured mean duration and the totals
 * estimate all launches), three values per family in the order
 * far field (FP64, direct), near pairs, stress evaluation, Gram-Schmidt,
 * preconditioner, far field (FP32, direct), cluster far field (whole
 * application), its direct sums (k_fmm_direct), its multipole and local
 * expansions (k_fmm_p2m + k_fmm_m2l, on a second stream beside the direct
 * sums): out[3i] milliseconds, out[3i+1] launches, out[3i+2] algorithmic
 * work (ordered source-target pairs; bytes of the near-pair gather pass, as
 * erosion_time_matvec's out[6]; target-node pairs of the barycentric rule;
 * bytes of Krylov basis streamed; bytes of vector read and written; ordered
 * source-target pairs; the Npts (Npts - 1) pairs the application stands for;
 * pairs summed; cluster translations).  `out` holds 27 doubles.  Returns 0 or
 * -1. */
int erosion_kernel_timers(double* out);

/* Precision of the far field in stokessolver_ (no counterpart in the
 * reference, which is FP64 throughout).  mode 0: FP64 (default).  mode 1:
 * mixed -- once GMRES's implicit residual is below 1e4 times the tolerance the
 * far-field pair sums run in FP32 (sources re-centred per 256-point tile,
 * FP32 partial sums of 64 terms accumulated in FP64); the near-pair
 * corrections, local terms, Gram-Schmidt, Givens update and the true
 * residual stay FP64, and a solve whose FP64 true residual misses the
 * tolerance restarts from its result.  mode 2: FP32 far field in every
 * operator application (measurement and tests; erosion_matvec follows it).
 * Returns 0, or -1 for another mode. */
int erosion_set_precision(int mode);
int erosion_get_precision(void);
/* Mixed mode switches the far field to FP32 once the implicit relative
 * residual is at most factor * tol (default 1e4; factor >= 1).  Returns 0 or
 * -1. */
int erosion_set_mixed_switch(double factor);

/* Many independent geometries in one call (no counterpart in the
 * reference).  Geometry i has npts[i], nbods[i], nouter[i] and the arrays
 * xx[i], yy[i] (npts*nbods) and den[i] (initial guess in, density out);
 * tau[i] (npts*nbods) receives its shear stress when tau is not null
 * (ibary must then be 1).  The solves use the current tolerance,
 * preconditioner and precision, run on `nworkers` worker contexts (own
 * stream and buffers each) from as many host threads, and overlap on the
 * GPU.  nits[i] and status[i] (0 ok, 1 not converged, 2 breakdown, -1 error)
 * are per geometry; relres[i] (when relres is not null) receives the true
 * relative residual of geometry i's density.  Returns 0, or -1 if any
 * geometry failed (first failure in erosion_last_error). */
int erosion_solve_batch(int64_t ngeom, const int64_t* npts, const int64_t* nbods,
                        const int64_t* nouter, const double* const* xx, const double* const* yy,
                        double* const* den, double* const* tau, int64_t ibary, int64_t ibc,
                        int64_t maxl, int64_t nworkers, int64_t* nits, int* status,
                        double* relres);

/* Measurement hook: `reps` device-to-device copies of `bytes` bytes on the
 * library's stream (the reference traffic for hardware byte counters);
 * ms = mean milliseconds per copy.  Returns 0 or -1. */
int erosion_measure_copy(int64_t bytes, int64_t reps, double* ms);

/* Measurement hook (no counterpart in the reference): on = 1 runs the
 * cluster far field's multipole and translation kernels on the library's
 * stream ahead of its direct sums instead of on a second stream beside them
 * (same kernels, same results), for hardware-counter passes, which serialize
 * every dispatch; on = 0 (default) restores the two streams.  Returns 0 or
 * -1. */
int erosion_set_fmm_one_stream(int on);

/* Measurement hook: the FP64 vector rate the chip sustains -- `blocks` blocks
 * of 256 threads, each thread `iters` iterations of 8 independent FP64 FMA
 * chains, on the library's stream; ms = milliseconds of the third of three
 * launches.  FMA per launch: 8 iters blocks 256.  Returns 0 or -1. */
int erosion_measure_fp64(int64_t blocks, int64_t iters, double* ms);

/* Right preconditioner of stokessolver_ (no counterpart in the reference,
 * whose solveBIE runs GMRES unpreconditioned, stokesSolver.f:490).
 * kind 0: none (default, the reference's iteration).  kind 1: per-curve
 * two-level preconditioner -- on each curve (wall, every body) the exact
 * inverse of the curve's own diagonal operator block on its Fourier modes
 * 0..K (K = wall_modes for the wall, body_modes for a body, each in [1, 32],
 * lowered to (points - 2)/2 on a curve with fewer points; a body's block
 * includes its stokeslet/rotlet columns and constraint rows), and -2 (the
 * inverse of the jump -1/2) on the remaining modes.  GMRES then iterates on
 * A P^{-1}; its residual is still that of A x = b, so the stopping test and
 * the returned density keep their meaning and only the iteration count
 * changes.  32 body modes and 16 wall modes are the defaults of
 * erosion_get_preconditioner before any call (eroding bodies move their
 * density into higher modes; 16 body modes cost fewer operations per
 * iteration and suit round bodies).
 * Returns 0, or -1 on invalid arguments. */
int erosion_set_preconditioner(int kind, int body_modes, int wall_modes);
/* out[0] kind, out[1] body modes, out[2] wall modes.  Returns 0 or -1. */
int erosion_get_preconditioner(int* out);

/* One erosion step on the device: the theta-L midpoint RK2 of rungekutta2 /
 * timestep! / getderivs / delete_indices (ErosionPackage/src/TimeStepping.jl:
 * 42-202) with getstress! and the fixed-pressure-drop rescaling
 * (DensityStress.jl:23-36, 129-167), replacing the Julia step that calls
 * stokessolver_, computeshearstress_ and computeqoitargets_ per stage.  The
 * boundary, densities and stresses of both stages stay on the GPU.
 *   params[0] dt, [1] epsilon (epsfac/npts), [2] sigma (sigfac/npts),
 *   [3] fixarea (0/1), [4] fixpdrop (0/1), [5] ibc, [6] ibary (must be 1),
 *   [7] maxl, [8] GMRES guess: 0 zeros (the reference), 1 the previous
 *   solution, 2 the linear extrapolation of the last two.  Solutions are
 *   kept on the device across calls in the preconditioned variable u = P x
 *   (the density without a preconditioner); a deletion drops the deleted
 *   bodies' unknowns from them.
 *   theta (npts x nbods, body-major), len, xsm, ysm: the state at t, replaced
 *   by the state at t + dt of the *nbods_out bodies that remain (bodies are
 *   deleted as delete_indices does).
 *   guess: null, or u0 (2 nouter + (2 npts + 3) nbods values, the
 *   preconditioned variable) used as the first stage's GMRES guess when no
 *   kept solution fits the state.
 *   info[0..9]: GMRES iterations of stage 1 and 2, bodies deleted, umax of
 *   stage 1 and 2, largest true relative residual, positive-mterm count,
 *   flux-mismatch count, solves started from kept densities, solves started
 *   from `guess`.
 * GMRES settings (tolerance, preconditioner, precision, graphs) are those of
 * stokessolver_.  At most 512 points per body.  Returns 0 or -1. */
int erosion_rungekutta2(int64_t npts, int64_t nbods, int64_t nouter, const double* params,
                        double* theta, double* len, double* xsm, double* ysm,
                        const double* guess, int64_t* nbods_out, double* info);
/* out[0] solutions kept for guesses (0-2), out[1] their body count, out[2]
 * their length.  Returns 0 or -1. */
int erosion_stepper_history(int64_t* out);
/* Forget the densities erosion_rungekutta2 keeps for its GMRES guesses (a new
 * run).  Returns 0 or -1. */
int erosion_stepper_reset(void);
/* Test hook: the stepper's spectral operations (SpectralMethods.jl:35-67,
 * ThetaLen.jl:92-101) on `rows` rows of npts points: op 0 specdiff, 1
 * specint, 2 expsmooth with fac[r] per row, 3 getxy with fac = (len[rows],
 * xsm[rows], ysm[rows]) writing x rows then y rows (2 x rows x npts values).
 * Returns 0 or -1. */
int erosion_spectral(int op, int64_t npts, int64_t rows, const double* in, const double* fac,
                     double* out);

/* Test and measurement hook: use far-field launch shape `index` (0 ..
 * erosion_farfield_shape_count() - 1: targets per lane, grid size, shared
 * reciprocal) for every size instead of the autotuner's choice; -1 restores
 * the autotuner.  With a direct shape fixed, mixed precision uses the FP32 far
 * field wherever its residual switch allows (no timing decides).  Returns 0,
 * or -1 for an index out of range. */
int erosion_set_farfield_shape(int index);
int erosion_farfield_shape_count(void);
/* Far field of the completed operator (the reference's FMM switch ifmm,
 * stokesSolver.f:530-536, which stokessolver_ still takes): method 1
 * (default) lets the autotuner pick, per boundary size, between the direct
 * all-pairs sums and the cluster expansions of tilefmm.hip (clusters: the
 * bodies and the wall's segments of 64 points; multipole and local
 * expansions of 32 terms between clusters at least three times their radii
 * apart, truncation below 5e-16 of the terms; direct sums for the rest),
 * whichever is faster, on calls with ifmm = 1, and sends the far bodies of
 * the shear stress through the same expansions; method 0 keeps the direct
 * sums and the barycentric rule everywhere.  The last shape index of
 * erosion_set_farfield_shape is the cluster far field.  Returns 0 or -1. */
int erosion_set_farfield_method(int method);

/* HIP graphs for GMRES (no counterpart in the reference).  on = 1: stokessolver_
 * runs each chunk of iterations (8 or 16 iterations of preconditioner,
 * operator, two Gram-Schmidt passes, Givens update) directly the first time
 * it comes, captures it into a HIP graph the second time, and replays that on
 * later solves while every size, launch shape and device buffer it holds is
 * unchanged; any change drops the kept graphs (so a one-off solve captures
 * nothing).  The results are those of on = 0, bit for bit.  Kernel timing
 * (erosion_kernel_timing) bypasses the graphs.  on = 0 (default) launches
 * every kernel directly and frees the kept graphs.  Returns 0, or -1 for
 * another value. */
int erosion_set_graphs(int on);
/* out[0] on/off, out[1] chunks captured, out[2] chunk replays, out[3] graphs
 * kept, counted since the library was loaded.  Returns 0 or -1. */
int erosion_graph_stats(long long* out);

/* Test hook: z = P^{-1} r for the selected preconditioner (kind must be 1) on
 * the given geometry; r and z have the den layout.  Returns 0 or -1. */
int erosion_precond_apply(int64_t npts, int64_t nbods, int64_t nouter, const double* xx,
                          const double* yy, const double* r, double* z);

/* ---- Learned warm start network (no counterpart in the reference, whose
 * GMRES starts from zeros, DensityStress.jl:45) --------------------------
 * A multilayer perceptron: nlayers hidden layers of nhid tanh units and a
 * linear output, h_{l+1} = tanh(W_l h_l + b_l), y = W_L h_L + b_L.  params is
 * one FP32 array, layer by layer, W_l (outputs x inputs, row-major) then b_l.
 * Every product rounds both operands to BF16 (nearest even) and accumulates
 * in FP32 on the matrix cores; bias, tanh, loss and Adam are FP32.  x is
 * nsamples x nin, targets t and outputs y are nsamples x nout, row-major.
 * Loss: mean squared error over samples and outputs.  All return 0 or -1. */
int erosion_mlp_param_count(int nin, int nhid, int nlayers, int nout, int64_t* count);
int erosion_mlp_forward(int nin, int nhid, int nlayers, int nout, const float* params,
                        int64_t nsamples, const float* x, float* y);
/* The forward pass of erosion_mlp_forward with no allocation, for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops that forbid hipMalloc).  No counterpart in the reference.  params,
 * x and y are DEVICE pointers (same layouts); the kernels run on `stream` (a
 * hipStream_t; null = the default stream) and the call returns once they are
 * enqueued.  Hidden activations live in a file-scope device array of 8 MiB
 * per GPU (held until the process ends): samples run in chunks of up to
 * 2^20 / nhid rows (rounded down to a multiple of 32), and a hidden width
 * above 32768 units is refused.  Every call on one GPU shares that array, so
 * calls that overlap on one GPU (other streams, other threads) race.  Errors
 * go to stderr and erosion_noalloc_last_error (this entry point does not
 * create the library context).  Returns 0 or -1. */
int erosion_mlp_forward_noalloc(int nin, int nhid, int nlayers, int nout, const float* params,
                                int64_t nsamples, const float* x, float* y, void* stream);
const char* erosion_noalloc_last_error(void);
/* Gradient of the loss over all nsamples with respect to params (same
 * layout), and the loss. */
int erosion_mlp_gradients(int nin, int nhid, int nlayers, int nout, const float* params,
                          int64_t nsamples, const float* x, const float* t, float* grads,
                          double* loss);
/* Adam (beta1 0.9, beta2 0.999, eps 1e-8) on shuffled mini-batches of
 * `batch` samples (a final partial batch is dropped) for `epochs` passes;
 * params is updated in place; epoch_loss[e] (if not null) is the mean batch
 * loss of epoch e. */
int erosion_mlp_train(int nin, int nhid, int nlayers, int nout, float* params, int64_t nsamples,
                      const float* x, const float* t, int epochs, int batch, float lr,
                      uint64_t seed, double* epoch_loss);

#ifdef __cplusplus
}
#endif
#endif
