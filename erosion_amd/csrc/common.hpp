// Shared declarations of the MI355X Stokes library: device helpers, the
// persistent per-process context, and the host-side entry points of each
// kernel family.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace stokes {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
    explicit Error(const std::string& s) : std::runtime_error(s) {}
};

#define STOKES_HIP(expr)                                                                    \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::stokes::Error(std::string(#expr) + ": " + hipGetErrorString(e_) + " (" + \
                                  __FILE__ + ":" + std::to_string(__LINE__) + ")");         \
    } while (0)

#define STOKES_LAUNCH_CHECK() STOKES_HIP(hipGetLastError())

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.28318530717958647692;
// GMRES cycles whose true residual and iteration count a solve record keeps
constexpr int kCycleRecord = 4;

// ------------------------------------------------------- device helpers
// FP64 reciprocal: hardware estimate refined by two Newton steps.  Used for
// every 1/rho^2 and complex reciprocal on the hot paths; arguments are
// squared distances between distinct boundary points (finite, > 0).
__device__ __forceinline__ double rcp64(double a) {
    double r = __builtin_amdgcn_rcp(a);
    r = __builtin_fma(r, __builtin_fma(-a, r, 1.0), r);
    r = __builtin_fma(r, __builtin_fma(-a, r, 1.0), r);
    return r;
}

struct cplx {
    double r, i;
};
__device__ __forceinline__ cplx cmk(double r, double i) { return cplx{r, i}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return cplx{a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return cplx{a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
    return cplx{__builtin_fma(a.r, b.r, -a.i * b.i), __builtin_fma(a.r, b.i, a.i * b.r)};
}
__device__ __forceinline__ cplx cscale(cplx a, double s) { return cplx{a.r * s, a.i * s}; }
__device__ __forceinline__ cplx cinv(cplx a) {
    double id = rcp64(__builtin_fma(a.r, a.r, a.i * a.i));
    return cplx{a.r * id, -a.i * id};
}
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) { return cmul(a, cinv(b)); }
// acc += a * b
__device__ __forceinline__ void cfma(cplx& acc, cplx a, cplx b) {
    acc.r = __builtin_fma(a.r, b.r, __builtin_fma(-a.i, b.i, acc.r));
    acc.i = __builtin_fma(a.r, b.i, __builtin_fma(a.i, b.r, acc.i));
}

// ------------------------------------------------------------ buffers
// Count of device-buffer allocations in the process.  A captured graph holds
// raw device pointers, so a graph is replayed only while this count is the one
// it was captured under.
inline std::atomic<unsigned long long> g_alloc_epoch{0};

// Owning device allocation; grows only.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    void ensure(size_t count) {
        if (count <= cap && p) return;
        release();
        size_t c = count ? count : 1;
        STOKES_HIP(hipMalloc(reinterpret_cast<void**>(&p), c * sizeof(T)));
        g_alloc_epoch.fetch_add(1, std::memory_order_relaxed);
        cap = c;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// --------------------------------------------------------- geometry
// All boundary points of one solve, wall first (index 0..no-1) then bodies
// body-major (no + b*n + k).  "Curves": curve 0 is the wall, curve 1+b body b.
struct Geometry {
    int n = 0, m = 0, no = 0;
    int npts = 0;  // no + m*n
    int nunk = 0;  // 2*no + 2*m*n + 3*m  (den layout of stokesSolver.f:631-657)
    // per point (npts)
    DevBuf<double> x, y, px, py, cur, speed;
    DevBuf<double> dzr, dzi;  // z'(theta) * dtheta  (fourierDiff(z*dtheta))
    DevBuf<double> mx, my;    // normal * speed * 2/nsrc: trapezoid weight over pi
    DevBuf<double> iar, iai;  // 1/(z - a_curve), a = point mean of the curve
    // per curve (m+1)
    DevBuf<double> ax, ay;
    // per body (m)
    DevBuf<double> cx, cy;
    // spectral derivative tables d(k) = (-1)^k cot(pi k / N) / 2
    DevBuf<double> dtab_n, dtab_no;
    // host copies needed to build near lists
    std::vector<double> h_x, h_y, h_speed, h_cx, h_cy;
    // build inputs and scratch
    DevBuf<double> in_x, in_y, t1x, t1y, t2x, t2y;
    // first and second derivatives of the wall (it depends on no only), kept
    // between builds: [t1x | t1y | t2x | t2y], each of wall_no doubles
    DevBuf<double> wall_d;
    int wall_no = -1;
    // the boundary this geometry was built from; version counts builds
    std::vector<double> key_x, key_y;
    int key_n = -1, key_m = -1, key_no = -1;
    long long version = 0;
};

// Weight of the launches timed now: 1, or P while one GMRES iteration in P
// is timed and stands for the P iterations (one per host thread, as each
// context is driven by one thread at a time).
inline thread_local int tl_time_weight = 1;

// Durations of one kernel family, measured with HIP events recorded around
// each launch on the library stream.  Events come from a pool and are read
// back only when the pool is full or on harvest(), so timing adds no host
// synchronisation to the launch sequence.  Each event record costs the
// stream a few microseconds, so the GMRES iterations may be sampled: every
// timed launch carries its weight, ms, launches and work are weighted sums
// (estimates of the totals over all launches) and ms / launches is the
// measured mean duration.
struct KernelTimer {
    std::vector<hipEvent_t> ev;  // pairs: start 2i, stop 2i+1
    std::vector<int> wt;         // weight of pair i
    size_t used = 0;
    double ms = 0.0;
    long long launches = 0;
    double work = 0.0;  // algorithmic units summed over launches
    void start(hipStream_t s) {
        if (ev.empty()) {
            ev.resize(2 * 512);
            wt.assign(512, 1);
            for (auto& e : ev) STOKES_HIP(hipEventCreate(&e));
        }
        if (used == ev.size() / 2) harvest();
        STOKES_HIP(hipEventRecord(ev[2 * used], s));
    }
    void stop(hipStream_t s, double work_units) {
        STOKES_HIP(hipEventRecord(ev[2 * used + 1], s));
        wt[used] = tl_time_weight;
        ++used;
        work += work_units * tl_time_weight;
    }
    void harvest() {
        if (!used) return;
        STOKES_HIP(hipEventSynchronize(ev[2 * used - 1]));
        for (size_t i = 0; i < used; ++i) {
            float t = 0.f;
            STOKES_HIP(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]));
            ms += static_cast<double>(t) * wt[i];
            launches += wt[i];
        }
        used = 0;
    }
    void reset() {
        harvest();
        ms = 0.0;
        launches = 0;
        work = 0.0;
    }
};

// Near-pair replacement list of matvec_DLP_fmmbary (stokesSolver.f:687-926):
// an entry is (source curve, target point); interior rule iff source is wall.
struct NearList {
    int nentries = 0;
    int nactive = 0;                 // source curves with at least one entry
    int wall_active = 0;             // 1 if the wall is a source (active[0] == 0)
    // algorithmic bytes of one near-evaluation pass: per (target, source curve)
    // entry, the source's nodes x 14 doubles (2 coordinates, 2 dz, 10
    // boundary-value components) plus 16 bytes written
    double gather_bytes = 0.0;
    DevBuf<int> ent_tgt;             // target point per entry (sorted by source)
    DevBuf<int> ent_src;             // source curve per entry
    DevBuf<int> tgt_ptr, tgt_ent;    // CSR: entries of each target point
    DevBuf<int> active;              // active source curves
    DevBuf<double> corr_u, corr_v;   // per-entry correction (bary - trapezoid)
    long long built_version = -1;    // geometry version and ibary of this list
    int built_ibary = -1;
};

// Cluster far field (tilefmm.hip).  Clusters are the wall's segments of 64
// points and the bodies; per cluster its centre and radius, its first point
// and count, the far clusters of each target cluster (multipole/local
// expansions), the source tiles each target cluster sums directly, and the
// expansions of one application.
struct TileFmm {
    long long version = -1;             // geometry version the lists were built for
    int nc = 0, nwall = 0;              // clusters; the first nwall are wall segments
    DevBuf<double> cr;                  // (cx, cy, R) per cluster
    DevBuf<int> cl;                     // (first point, count) per cluster
    DevBuf<int> far_ptr, far_idx;       // CSR: far clusters of each target cluster
    DevBuf<int> tile_ptr, tiles;        // CSR: (first point, count) source tiles per cluster
    DevBuf<double> mp, lc;              // multipoles [nc][3][32], locals [nc][3][32] (complex)
    DevBuf<double> bin;                 // binomials of the translation, matrix-core layout
    DevBuf<int> work;                   // (cluster, tile) per direct-sum block: bodies, wall
    int work_body = 0, work_wall = 0;   // work items of body and wall targets
    int max_tiles = 1;                  // most source tiles of one cluster
    long long far_pairs = 0;            // (target cluster, far cluster) pairs
    long long direct_pairs = 0;         // point pairs summed directly per application
    // shear stress: far bodies of each body (CSR, body indices), the multipoles
    // and locals of its five boundary-value series, and (host) the near
    // bodies of each body, itself included, for the barycentric rule
    DevBuf<int> sfar_ptr, sfar_idx;
    DevBuf<int> nb_ptr, nb_idx;         // near bodies of each cluster (the assembly's stokeslets)
    DevBuf<double> mlog;                // per cluster the log terms (c_F, c_Q) of its stokeslet
    DevBuf<double> logd;                // per far pair log(c_t - c_s) (complex)
    DevBuf<double> smp, slc;            // [m][5][32] complex
    // groups of nearby clusters (parents, cluster indices nc .. nc + npar - 1
    // in cr, mp and mlog): their children (CSR), the two-level far lists
    // (a parent far from a target stands for all its children), and the
    // binomials binom(k, n) of the multipole shift, [32][32]
    int npar = 0;
    DevBuf<int> par_ptr, par_idx;
    DevBuf<int> far2_ptr, far2_idx;
    DevBuf<double> logd2;
    DevBuf<double> binm;
    long long far2_pairs = 0;
    std::vector<int> h_snear_ptr, h_snear_idx;
    // the wall in the shear stress: per body 1 if it is close to the wall
    // (barycentric rule over the whole wall), else the wall comes through
    // the far segments' expansions and the near segments' direct sums
    // (swn: CSR of near segment indices per body); scr holds the expansion
    // sources' (cx, cy, R): bodies, then wall segments
    std::vector<unsigned char> h_swall;
    DevBuf<int> swn_ptr, swn_idx;
    DevBuf<double> scr;
    bool stress_wall = false;
    long long stress_direct_pairs = 0;  // (target point, source node) pairs of the near curves
    // the expansions run on a second stream (default priority) beside the
    // direct sums: fork and join events on the context's stream; one_stream:
    // on the context's stream, before them (hardware-counter passes, which
    // serialize dispatches, then see no waits between streams)
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    bool one_stream = false;
};

// Per-solve GMRES state kept on the device.
struct GmresState {
    double bnrm, r0nrm, prod, rho;
    double snormw, inv_snormw;  // ||w'|| of the last iteration (Gram-Schmidt) and 1/||w'||
    double vnrm;
    double rinv, gk;            // 1/r and g_j of the delayed second pass
    int k;          // Arnoldi columns completed
    int done;       // 1 converged, 2 stopped (maxl), 3 breakdown
    int reorth;     // unused
    int nreorth;    // unused (every vector gets its second pass)
};

// GMRES iteration chunks captured as HIP graphs.  `sig` lists every value the
// captured launches bake in besides the chunk key (sizes, grid shapes, the
// allocation epoch); a different signature drops every graph.
struct GraphCache {
    std::vector<long long> sig;
    std::map<std::array<long long, 4>, hipGraphExec_t> exec;  // (k0, chunk, fp32, maxl)
    // chunks run once directly under the current signature: a chunk is
    // captured the second time it comes, so a one-off solve pays no capture
    std::set<std::array<long long, 4>> seen;
    long long captures = 0, replays = 0;
    void clear() {
        for (auto& kv : exec) (void)hipGraphExecDestroy(kv.second);
        exec.clear();
        seen.clear();
    }
    GraphCache() = default;
    GraphCache(const GraphCache&) = delete;
    GraphCache& operator=(const GraphCache&) = delete;
    ~GraphCache() { clear(); }
};

// Device-resident erosion step (stepper.hip): run parameters, per-step
// report, and the device buffers kept between steps.
struct StepParams {
    double dt = 1e-3, eps = 0.0, sigma = 0.0, fixarea = 0.0;
    int fixpdrop = 1, ibc = 0, ibary = 1, maxl = 8000;
    // GMRES guess, in the preconditioned variable u = P x of the solves: 0
    // zeros, 1 the previous solution, 2 linear extrapolation of the last two
    int warm = 2;
};
struct StepReport {
    long long iters[2] = {0, 0};  // GMRES iterations of the two stages
    int deleted = 0;              // bodies removed by delete_indices
    double umax[2] = {1.0, 1.0};  // fixed-pressure-drop rescaling per stage
    double true_relres_max = 0.0;
    int positive_mterm = 0, flux_warnings = 0, warm_starts = 0, given_guesses = 0;
};
struct StepperBufs {
    DevBuf<double> th0, th05, th1, xx, yy, den, tau, atau, matau, nterm, dscal, scal, fac, tgt,
        qout, tmp, hist0, hist1;
    DevBuf<int> idx;
    // solutions (preconditioned variable) kept for guesses, their body count
    // and length
    int hist_count = 0, hist_m = -1, hist_n = -1;
};

struct Context {
    Geometry g;
    NearList near;
    hipStream_t stream = nullptr;
    // operator work buffers
    DevBuf<double> ex, ey;          // density in point order
    DevBuf<double> part_u, part_v;  // far-field partial sums [split][npts]
    DevBuf<double> part_cu, part_cv;  // their corrections (compensated evaluation)
    DevBuf<double> bd;              // boundary values [5 fields][2][npts]
    DevBuf<double> bdpart;          // wall boundary-value partial sums [split][20][nouter]
    DevBuf<double> fld;             // field densities [5][2][npts]
    DevBuf<double> scal;            // sigdotn etc.
    DevBuf<double> stk_log;           // log |z_i - c_b|^2 [npts][m]
    long long stk_log_version = -1;   // geometry version stk_log was built for
    int ff_splits = 0, ff_tpt = 0;
    std::map<long long, int> ff_choice;      // size class -> far-field shape index
    std::map<long long, float> ff_tuned_ms;  // size class -> tuned launch time
    std::map<long long, int> ff32_choice;    // the same for the FP32 far field
    int ff_force = -1;                 // a fixed far-field shape index (tests), -1 tunes
    // the cluster far field may be chosen: ff_method is the caller's
    // setting (1 allowed, 0 direct sums only), ff_fmm the current call's
    // (stokessolver_'s ifmm = 1 and the setting)
    int ff_method = 1;
    int ff_fmm = 1;
    TileFmm fmm;
    std::map<long long, float> ff32_tuned_ms;
    // FP32 far field: per 256-point tile the FP64 centre, per source the FP32
    // offset from it and the FP32 weighted normal; built once per geometry
    DevBuf<double> ff32_ctr;
    DevBuf<float> ff32_xy;
    long long ff32_version = -1;
    // precision of the far field: 0 FP64; 1 mixed (GMRES switches the far
    // field to FP32 once its residual is small, see gmres.hip); 2 FP32 always
    int precision = 0;
    double mixed_switch = 1.0e4;  // FP32 far field once residual <= mixed_switch * tol
    bool ff_fp32_now = false;  // the far field of the next operator applications
    // space of the initial guess a solve is handed: 0 the density x; 1 the
    // preconditioned variable u = P x (x0 = P^{-1} u0; u = x without a
    // preconditioner).  With guess_space 1 a solve also leaves its final u
    // in uvec (u_valid), the starting point a later, nearby solve can take.
    int guess_space = 0;
    DevBuf<double> uvec;
    bool u_valid = false;
    // GMRES buffers
    int npad = 0;
    int maxl_alloc = 0;
    DevBuf<double> V;       // (maxl+1) x npad
    DevBuf<double> H;       // (maxl+1) x maxl, column-major: the rotated (R) Hessenberg
    DevBuf<double> Hraw;    // the same, Arnoldi columns before rotation
    DevBuf<double> q;       // Givens (c, s) pairs
    DevBuf<double> hbuf, hpart, hpart2, wpart, npart, ybuf, hcur, bpart;
    DevBuf<double> rhs, xvec, rvec;
    DevBuf<GmresState> st;
    GmresState* h_st = nullptr;  // pinned mirror
    // stress / target-evaluation buffers
    DevBuf<double> spart;   // per curve-group partials
    DevBuf<double> tau;
    DevBuf<int> curves_all, iside;
    // barycentric curve lists: CSR over (target block, group), curve 0 the
    // wall and 1 + b body b; a per-(body, curve) mask where the lists of one
    // block serve several bodies.  stress_lists_version: geometry version of
    // the near-curve lists of the shear stress (-1: none)
    DevBuf<int> lst_ptr, lst;
    DevBuf<unsigned char> lst_mask;
    long long stress_lists_version = -1;
    int stress_groups = 0;
    // the wall in the shear stress of bodies close to it, by node ranges:
    // their points (wb_tq), the ranges' partial sums of each pass and the
    // per-target pass results (S0 and u_f; du_f)
    DevBuf<int> wb_tq;
    DevBuf<double> wb_part, wb_u, wb_du;
    int wb_nct = 0;
    DevBuf<double> tgt;     // target coordinates and QoI outputs
    double tol = 1e-8;
    // right preconditioner (precond.hip): kind 0 none (the reference's
    // unpreconditioned GMRES), 1 per-curve two-level Fourier
    int pc_kind = 0, pc_body_modes = 32, pc_wall_modes = 16;
    int pc_nfb = 0, pc_nfw = 0;
    int pc_wall_no = -1, pc_wall_kw = -1;  // key of the kept wall block
    DevBuf<double> pc_phi_b, pc_phi_w, pc_S, pc_Gb, pc_Gw, pc_z, pc_wpart;
    DevBuf<int> pc_bad;  // per-curve flags: coarse block singular
    // the (z, y) of the last precond_apply_scaled with y, cleared by anything
    // else that writes c.ex/c.ey (apply_operator's from_precond checks it)
    const double* pc_token_x = nullptr;
    const double* pc_token_y = nullptr;
    // per-kernel timing with HIP events on `stream` (measurement hooks only);
    // the GMRES iterations' kernels are timed in one iteration of every
    // time_period (time_tick counts the iterations)
    bool time_kernels = false;
    int time_period = 1;
    long long time_tick = 0;
    // replay GMRES iteration chunks as HIP graphs (off while timing kernels)
    int use_graphs = 0;
    GraphCache graphs;
    StepperBufs sp;
    KernelTimer t_farfield, t_near, t_stress_eval, t_gs, t_precond, t_farfield32;
    // cluster far field: the whole application (work: the canonical
    // Npts (Npts - 1) pairs it stands for), its direct sums (k_fmm_direct;
    // work: pairs), and the multipole and local expansions (k_fmm_p2m +
    // k_fmm_m2l on the side stream, beside the direct sums; work: cluster
    // translations)
    KernelTimer t_fmm, t_fmm_direct, t_fmm_expand;
    int status = 0;
    std::string last_error;
    // last solve record
    long long last_iters = 0;
    double last_true_relres = -1.0, last_implicit_relres = -1.0, last_seconds = 0.0;
    int last_reorth = 0;
    long long last_fp32_iters = 0;  // iterations whose far field ran in FP32
    int last_cycles = 0;            // GMRES cycles (restarts + 1) of the last solve
    double last_r0_relres = -1.0;   // ||b - A x0|| / ||b|| of its initial guess
    int last_tol_miss = 0;          // 1: the last solve ended above the tolerance
    long long tol_misses = 0;       // such solves since the process started
    // of those, the ones whose density's FP64 floor (CycleRecord::floor_relres)
    // is at least half the residual reached
    long long tol_misses_at_floor = 0;
    // per-cycle record of a solve (at most kCycleRecord cycles): why it
    // stopped, the true relative residual and the iterations of each cycle;
    // `last` is the latest solve, `miss` the latest one above the tolerance
    struct CycleRecord {
        int stop = 0;  // 0 met the tolerance, 1 a restart did not halve the
                       // true residual, 2 cycle limit, 3 maxl spent, 4 GMRES
                       // status nonzero (maxl or breakdown inside the cycle)
        double true_relres[kCycleRecord] = {};
        double iters[kCycleRecord] = {};
        // measured only above the tolerance: ||A d|| / ||b|| for a random-sign
        // half-ulp perturbation d of the returned density (0 otherwise)
        double floor_relres = 0.0;
    } last_rec, miss_rec;
};

Context& ctx();

// geometry.hip
// xx, yy: body points (n per body, body-major), in host memory or, with
// on_device, in device memory
void build_geometry(Context& c, int n, int m, int no, const double* xx, const double* yy,
                    bool on_device = false);
void build_near_list(Context& c, int ibary);

// operator.hip : y = A x for the completed double-layer operator
// from_precond: x_dev is the output of precond_apply_scaled(..., y_dev), which
// has left x_dev's density in point order in c.ex/c.ey, the body constraint
// rows in y_dev and the wall segments' parts of the rank-one integral in
// c.scal[0 .. precond_wall_segments).  Honoured only when the context's
// token says that call was the last to write c.ex/c.ey with this (x, y);
// otherwise the full set-up runs.
// compensated: a true-residual evaluation, FP64 far field whose sums carry
// error-free corrections (k_farfield_comp) and an assembly in double-double
void apply_operator(Context& c, const double* x_dev, double* y_dev, int ibary,
                    bool from_precond = false, bool compensated = false);
// run the geometry-dependent set-up apply_operator would otherwise do inside
// its first call (far-field tuning, FP32 source tiles, stokeslet logarithms),
// so a captured sequence of operator applications holds none of it; x_dev is
// a density the tuning launches may read (the solve's initial guess)
void prepare_operator(Context& c, const double* x_dev, bool fp32);
// far-field shape index in use for the current size (after prepare_operator)
int farfield_choice(const Context& c, bool fp32);
// direct launch shapes plus the cluster far field (the last index)
int farfield_shape_count();
// true when the tuned FP32 far field is faster than the FP64 far field in use
bool farfield_fp32_pays(const Context& c);

// tilefmm.hip: the cluster far field (ifmm = 1); both split counts are the
// number of source splits whose partial sums it leaves in part_u/part_v
bool tilefmm_available(const Context& c);
void tilefmm_prepare(Context& c);
// lists, part buffers and the side stream of the cluster far field, without
// an application (set-up before a captured chunk)
void tilefmm_ready(Context& c, bool compensated = false);
void tilefmm_apply(Context& c, const double* den, int* wall_nsplit, int* body_nsplit,
                   bool compensated = false);
// the near bodies of every cluster (CSR, body indices): point p belongs to
// cluster p / wall_seg on the wall, nwall + (p - no) / n on a body
struct TileNear {
    const int* ptr;
    const int* idx;
    int nwall, wall_seg;
};
TileNear tilefmm_near_bodies(const Context& c);
// strain rate (E11, E12, E22 at out[{0,1,2} nbp + q]) of the far bodies'
// boundary values in c.bd
void tilefmm_stress(Context& c, double* out);
void gather_density(Context& c, const double* den_dev);
void boundary_values(Context& c, const int* curves_dev, int ncurves, int wall_first);
void bd_condition(Context& c, int ibc, double* rhs_dev);

// gmres.hip
// den: initial guess in, density out (N doubles), in host memory or, with
// on_device, in device memory
long long gmres_solve(Context& c, int ibary, int ibc, int maxl, double tol, double* den,
                      bool on_device = false);

// precond.hip: set up for the current geometry, then z = P^{-1} r
void precond_setup(Context& c);
void precond_apply(Context& c, const double* r_dev, double* z_dev);
// The same with r_dev first scaled in place by st->inv_snormw unless st is
// null or has stopped (GMRES's normalisation of its new Krylov vector).  Both
// also leave z's density in point order in c.ex/c.ey (the operator's gather).
// With y_dev, the apply also does the operator's small reductions of z: the
// body constraint rows of y_dev = A z and the wall segments' parts of the
// rank-one integral (apply_operator(c, z_dev, y_dev, ibary, true) follows).
void precond_apply_scaled(Context& c, double* r_dev, double* z_dev, const GmresState* st,
                          double* y_dev = nullptr);
// blocks over the wall in the apply: ceil(no / 256)
int precond_wall_segments(const Geometry& g);

// batch.hip: independent geometries solved by a pool of worker contexts
struct BatchGeometry {
    int n, m, no;
    const double *xx, *yy;
    double* den;     // in: initial guess; out: density
    double* tau;     // shear stress out, or null
    int64_t* nits;
    int* status;     // 0 ok, 1 not converged, 2 breakdown, -1 error
    double* relres;  // true relative residual of the returned density, or null
};
void solve_batch(Context& settings, int ngeom, const BatchGeometry* geo, int ibary, int ibc,
                 int maxl, int nworkers, std::string* first_error);

// mlp.hip: learned warm start network (FP32 parameters, BF16 products)
size_t mlp_param_count(int nin, int nhid, int nlayers, int nout);
void mlp_forward(Context& c, int nin, int nhid, int nlayers, int nout, const float* params,
                 int nsamples, const float* x, float* y);
// The same forward pass with every pointer in device memory, launched on
// stream s, with no device allocation: hidden activations in a file-scope
// device array of mlp.hip (shared by every call on a GPU).
void mlp_forward_noalloc(int nin, int nhid, int nlayers, int nout, const float* params_dev,
                         int nsamples, const float* x_dev, float* y_dev, hipStream_t s);
double mlp_gradients(Context& c, int nin, int nhid, int nlayers, int nout, const float* params,
                     int nsamples, const float* x, const float* t, float* grads);
void mlp_train(Context& c, int nin, int nhid, int nlayers, int nout, float* params, int nsamples,
               const float* x, const float* t, int epochs, int batch, float lr, uint64_t seed,
               double* epoch_loss);

// stepper.hip: one theta-L RK2 step with the state (theta n x m, len, xsm,
// ysm) in host memory, advanced in place; *m_out bodies remain
// guess0 (host, may be null): stage-1 GMRES guess used when no kept density
// fits the state
void rk2_step(Context& c, const StepParams& P, int n, int no, int m, double* theta, double* len,
              double* xsm, double* ysm, const double* guess0, int* m_out, StepReport& rep);
// op 0 specdiff, 1 specint, 2 expsmooth (fac per row), 3 getxy (fac = len,
// xsm, ysm per row; out = x rows then y rows) on rows of n points
void spectral_rows(Context& c, int op, int n, int rows, const double* in, const double* fac,
                   double* out);

// stress.hip
void shear_stress(Context& c, const double* den_dev, double* tau_dev);
// pressure on every body point (computePressure)
void pressure(Context& c, const double* den_dev, double* out_dev);
// velocity, pressure, vorticity at targets; out[4*t + {0,1,2,3}]
void qoi_targets(Context& c, int ibary, const double* den_dev, int ntar, const double* tx_dev,
                 const double* ty_dev, double* out_dev);

}  // namespace stokes
