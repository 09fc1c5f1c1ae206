// C-ABI entry points of libstokes_hip.so (declared in include/erosion_stokes.h).
#include <cstring>
#include <exception>

#include "../../include/erosion_stokes.h"
#include "common.hpp"

namespace stokes {

Context& ctx() {
    static Context* c = nullptr;
    if (!c) {
        c = new Context();
        STOKES_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    }
    return *c;
}

}  // namespace stokes

using stokes::Context;

namespace {

int record_error(const char* where, const std::string& msg) {
    std::string m = std::string(where) + ": " + msg;
    try {
        Context& c = stokes::ctx();
        c.status = -1;
        c.last_error = m;
    } catch (...) {
    }
    std::fprintf(stderr, "libstokes_hip: %s\n", m.c_str());
    return -1;
}

std::string& noalloc_error() {
    static std::string e;
    return e;
}

void check_sizes(int64_t n, int64_t m, int64_t no) {
    if (n < 2 || m < 0 || no < 4)
        throw stokes::Error("invalid sizes npts=" + std::to_string(n) + " nbods=" +
                            std::to_string(m) + " nouter=" + std::to_string(no));
    const int64_t nunk = 2 * no + 2 * n * m + 3 * m;
    if (nunk > (int64_t{1} << 30)) throw stokes::Error("problem too large for 32-bit indexing");
}

}  // namespace

extern "C" {

void stokessolver_(const int64_t* npts, const int64_t* nbods, const int64_t* nouter,
                   const int64_t* ifmm, const int64_t* ibary, const int64_t* ibc,
                   const int64_t* maxl, const double* xx, const double* yy, double* den,
                   int64_t* nits) {
    try {
        if (!npts || !nbods || !nouter || !ifmm || !ibary || !ibc || !maxl || !den || !nits)
            throw stokes::Error("null argument");
        *nits = 0;
        check_sizes(*npts, *nbods, *nouter);
        if (*nbods > 0 && (!xx || !yy)) throw stokes::Error("null body coordinates");
        if (*ibary == 1 && *ifmm == 0)
            throw stokes::Error(
                "ifmm=0 with ibary=1 (all-pairs barycentric matvec_DLP_bary) is not provided");
        if (*ibary != 0 && *ibary != 1) throw stokes::Error("ibary must be 0 or 1");
        Context& c = stokes::ctx();
        c.status = 0;
        c.last_error.clear();
        // ifmm = 1: the far field may be the cluster expansion scheme
        // (the reference's FMM switch, stokesSolver.f:530-536)
        c.ff_fmm = *ifmm != 0 && c.ff_method == 1 ? 1 : 0;
        const int n = static_cast<int>(*npts), m = static_cast<int>(*nbods),
                  no = static_cast<int>(*nouter);
        stokes::build_geometry(c, n, m, no, xx, yy);
        stokes::build_near_list(c, static_cast<int>(*ibary));
        long long it = stokes::gmres_solve(c, static_cast<int>(*ibary), static_cast<int>(*ibc),
                                           static_cast<int>(*maxl), c.tol, den);
        *nits = it;
        if (c.status == 1)
            std::fprintf(stderr,
                         "libstokes_hip: GMRES stopped after %lld iterations (maxl=%lld) with "
                         "relative residual %.3e (true %.3e)\n",
                         it, static_cast<long long>(*maxl), c.last_implicit_relres,
                         c.last_true_relres);
        else if (c.status == 2)
            std::fprintf(stderr, "libstokes_hip: GMRES breakdown after %lld iterations\n", it);
    } catch (const std::exception& e) {
        record_error("stokessolver_", e.what());
    }
}

void computeshearstress_(const int64_t* npts, const int64_t* nbods, const int64_t* nouter,
                         const double* xx, const double* yy, const double* den,
                         const int64_t* ibary, double* tau) {
    try {
        if (!npts || !nbods || !nouter || !den || !ibary || !tau)
            throw stokes::Error("null argument");
        check_sizes(*npts, *nbods, *nouter);
        const int n = static_cast<int>(*npts), m = static_cast<int>(*nbods),
                  no = static_cast<int>(*nouter);
        std::memset(tau, 0, sizeof(double) * static_cast<size_t>(n) * static_cast<size_t>(m));
        if (m == 0) return;
        if (!xx || !yy) throw stokes::Error("null body coordinates");
        if (*ibary != 1)
            throw stokes::Error("ibary=0 (deformation_on_boundary_trap) is not provided");
        Context& c = stokes::ctx();
        c.status = 0;
        c.last_error.clear();
        stokes::build_geometry(c, n, m, no, xx, yy);
        const size_t nunk = static_cast<size_t>(c.g.nunk);
        const size_t nbp = static_cast<size_t>(n) * static_cast<size_t>(m);
        c.rvec.ensure(std::max(nunk, c.rvec.cap));
        c.tau.ensure(nbp);
        STOKES_HIP(hipMemcpyAsync(c.rvec.p, den, nunk * sizeof(double), hipMemcpyHostToDevice,
                                  c.stream));
        stokes::shear_stress(c, c.rvec.p, c.tau.p);
        STOKES_HIP(hipMemcpyAsync(tau, c.tau.p, nbp * sizeof(double), hipMemcpyDeviceToHost,
                                  c.stream));
        STOKES_HIP(hipStreamSynchronize(c.stream));
    } catch (const std::exception& e) {
        record_error("computeshearstress_", e.what());
    }
}

void computeqoitargets_(const int64_t* npts, const int64_t* nbods, const int64_t* nouter,
                        const int64_t* ibary, const double* xx, const double* yy,
                        const double* den, const int64_t* ntargets, const double* xtar,
                        const double* ytar, double* utar, double* vtar, double* ptar,
                        double* vortar) {
    try {
        if (!npts || !nbods || !nouter || !ibary || !den || !ntargets)
            throw stokes::Error("null argument");
        check_sizes(*npts, *nbods, *nouter);
        const int nt = static_cast<int>(*ntargets);
        if (nt < 0) throw stokes::Error("negative target count");
        if (nt == 0) return;
        if (!xtar || !ytar || !utar || !vtar || !ptar || !vortar)
            throw stokes::Error("null target array");
        for (double* a : {utar, vtar, ptar, vortar})
            std::memset(a, 0, sizeof(double) * static_cast<size_t>(nt));
        const int n = static_cast<int>(*npts), m = static_cast<int>(*nbods),
                  no = static_cast<int>(*nouter);
        if (m > 0 && (!xx || !yy)) throw stokes::Error("null body coordinates");
        if (*ibary != 0 && *ibary != 1) throw stokes::Error("ibary must be 0 or 1");
        Context& c = stokes::ctx();
        c.status = 0;
        c.last_error.clear();
        stokes::build_geometry(c, n, m, no, xx, yy);
        const size_t nunk = static_cast<size_t>(c.g.nunk);
        c.rvec.ensure(std::max(nunk, c.rvec.cap));
        c.tgt.ensure(6 * static_cast<size_t>(nt));
        double* tx = c.tgt.p;
        double* ty = c.tgt.p + nt;
        double* out = c.tgt.p + 2 * nt;
        STOKES_HIP(hipMemcpyAsync(c.rvec.p, den, nunk * sizeof(double), hipMemcpyHostToDevice,
                                  c.stream));
        STOKES_HIP(hipMemcpyAsync(tx, xtar, nt * sizeof(double), hipMemcpyHostToDevice, c.stream));
        STOKES_HIP(hipMemcpyAsync(ty, ytar, nt * sizeof(double), hipMemcpyHostToDevice, c.stream));
        stokes::qoi_targets(c, static_cast<int>(*ibary), c.rvec.p, nt, tx, ty, out);
        std::vector<double> h(4 * static_cast<size_t>(nt));
        STOKES_HIP(hipMemcpyAsync(h.data(), out, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                                  c.stream));
        STOKES_HIP(hipStreamSynchronize(c.stream));
        for (int t = 0; t < nt; ++t) {
            utar[t] = h[4 * static_cast<size_t>(t)];
            vtar[t] = h[4 * static_cast<size_t>(t) + 1];
            ptar[t] = h[4 * static_cast<size_t>(t) + 2];
            vortar[t] = h[4 * static_cast<size_t>(t) + 3];
        }
    } catch (const std::exception& e) {
        record_error("computeqoitargets_", e.what());
    }
}

void computepressure_(const int64_t* npts, const int64_t* nbods, const int64_t* nouter,
                      const double* xx, const double* yy, const double* den, double* pressure) {
    try {
        if (!npts || !nbods || !nouter || !den || !pressure) throw stokes::Error("null argument");
        check_sizes(*npts, *nbods, *nouter);
        const int n = static_cast<int>(*npts), m = static_cast<int>(*nbods),
                  no = static_cast<int>(*nouter);
        const size_t nbp = static_cast<size_t>(n) * static_cast<size_t>(m);
        std::memset(pressure, 0, sizeof(double) * nbp);
        if (m == 0) return;
        if (!xx || !yy) throw stokes::Error("null body coordinates");
        Context& c = stokes::ctx();
        c.status = 0;
        c.last_error.clear();
        stokes::build_geometry(c, n, m, no, xx, yy);
        const size_t nunk = static_cast<size_t>(c.g.nunk);
        c.rvec.ensure(std::max(nunk, c.rvec.cap));
        c.tau.ensure(nbp);
        STOKES_HIP(hipMemcpyAsync(c.rvec.p, den, nunk * sizeof(double), hipMemcpyHostToDevice,
                                  c.stream));
        stokes::pressure(c, c.rvec.p, c.tau.p);
        STOKES_HIP(hipMemcpyAsync(pressure, c.tau.p, nbp * sizeof(double), hipMemcpyDeviceToHost,
                                  c.stream));
        STOKES_HIP(hipStreamSynchronize(c.stream));
    } catch (const std::exception& e) {
        record_error("computepressure_", e.what());
    }
}

int erosion_solve_batch(int64_t ngeom, const int64_t* npts, const int64_t* nbods,
                        const int64_t* nouter, const double* const* xx, const double* const* yy,
                        double* const* den, double* const* tau, int64_t ibary, int64_t ibc,
                        int64_t maxl, int64_t nworkers, int64_t* nits, int* status,
                        double* relres) {
    try {
        if (ngeom < 0 || ngeom > (int64_t{1} << 20)) throw stokes::Error("invalid geometry count");
        if (ngeom == 0) return 0;
        if (!npts || !nbods || !nouter || !xx || !yy || !den || !nits || !status)
            throw stokes::Error("null argument");
        if (ibary != 0 && ibary != 1) throw stokes::Error("ibary must be 0 or 1");
        if (tau && ibary != 1) throw stokes::Error("shear stress needs ibary=1");
        if (nworkers < 1 || nworkers > 64) throw stokes::Error("nworkers must lie in [1, 64]");
        std::vector<stokes::BatchGeometry> geo(static_cast<size_t>(ngeom));
        for (int64_t i = 0; i < ngeom; ++i) {
            check_sizes(npts[i], nbods[i], nouter[i]);
            if (!den[i] || (nbods[i] > 0 && (!xx[i] || !yy[i])))
                throw stokes::Error("null array of geometry " + std::to_string(i));
            geo[static_cast<size_t>(i)] = stokes::BatchGeometry{
                static_cast<int>(npts[i]), static_cast<int>(nbods[i]), static_cast<int>(nouter[i]),
                xx[i], yy[i], den[i], tau ? tau[i] : nullptr, nits + i, status + i,
                relres ? relres + i : nullptr};
            nits[i] = 0;
            status[i] = -1;
        }
        Context& c = stokes::ctx();
        c.status = 0;
        c.last_error.clear();
        std::string err;
        stokes::solve_batch(c, static_cast<int>(ngeom), geo.data(), static_cast<int>(ibary),
                            static_cast<int>(ibc), static_cast<int>(maxl),
                            static_cast<int>(nworkers), &err);
        if (!err.empty()) return record_error("erosion_solve_batch", err);
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_solve_batch", e.what());
    }
}

int erosion_set_gmres_tol(double tol) {
    try {
        if (!(tol > 0.0)) return -1;
        stokes::ctx().tol = tol;
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_set_gmres_tol", e.what());
    }
}

double erosion_get_gmres_tol(void) {
    try {
        return stokes::ctx().tol;
    } catch (...) {
        return -1.0;
    }
}

int erosion_last_status(void) {
    try {
        return stokes::ctx().status;
    } catch (...) {
        return -1;
    }
}

const char* erosion_last_error(void) {
    try {
        return stokes::ctx().last_error.c_str();
    } catch (...) {
        return "library context unavailable";
    }
}

int erosion_last_solve_info(double* info) {
    if (!info) return -1;
    try {
        Context& c = stokes::ctx();
        info[0] = static_cast<double>(c.last_iters);
        info[1] = c.last_implicit_relres;
        info[2] = c.last_true_relres;
        info[3] = c.last_seconds;
        info[4] = static_cast<double>(c.last_reorth);
        info[5] = static_cast<double>(c.ff_tpt);
        info[6] = static_cast<double>(c.ff_splits);
        info[7] = static_cast<double>(c.pc_kind);
        info[8] = static_cast<double>(c.last_fp32_iters);
        info[9] = static_cast<double>(c.last_tol_miss);
        info[10] = static_cast<double>(c.tol_misses);
        info[11] = static_cast<double>(c.last_cycles);
        info[12] = c.last_r0_relres;
        const Context::CycleRecord* recs[2] = {&c.last_rec, &c.miss_rec};
        for (int r = 0; r < 2; ++r) {
            double* o = info + 13 + r * (2 + 2 * stokes::kCycleRecord);
            o[0] = static_cast<double>(recs[r]->stop);
            for (int i = 0; i < stokes::kCycleRecord; ++i) {
                o[1 + i] = recs[r]->true_relres[i];
                o[1 + stokes::kCycleRecord + i] = recs[r]->iters[i];
            }
            o[1 + 2 * stokes::kCycleRecord] = recs[r]->floor_relres;
        }
        info[33] = static_cast<double>(c.tol_misses_at_floor);
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_last_solve_info", e.what());
    }
}

int erosion_matvec(int64_t npts, int64_t nbods, int64_t nouter, int64_t ibary,
                   const double* xx, const double* yy, const double* den, double* vel) {
    try {
        check_sizes(npts, nbods, nouter);
        if (!den || !vel || (nbods > 0 && (!xx || !yy))) throw stokes::Error("null argument");
        if (ibary != 0 && ibary != 1) throw stokes::Error("ibary must be 0 or 1");
        Context& c = stokes::ctx();
        c.status = 0;
        const int n = static_cast<int>(npts), m = static_cast<int>(nbods),
                  no = static_cast<int>(nouter);
        stokes::build_geometry(c, n, m, no, xx, yy);
        stokes::build_near_list(c, static_cast<int>(ibary));
        const size_t nunk = static_cast<size_t>(c.g.nunk);
        c.xvec.ensure(std::max(nunk, c.xvec.cap));
        c.rvec.ensure(std::max(nunk, c.rvec.cap));
        STOKES_HIP(hipMemcpyAsync(c.xvec.p, den, nunk * sizeof(double), hipMemcpyHostToDevice,
                                  c.stream));
        c.ff_fp32_now = c.precision == 2;
        stokes::apply_operator(c, c.xvec.p, c.rvec.p, static_cast<int>(ibary));
        c.ff_fp32_now = false;
        STOKES_HIP(hipMemcpyAsync(vel, c.rvec.p, nunk * sizeof(double), hipMemcpyDeviceToHost,
                                  c.stream));
        STOKES_HIP(hipStreamSynchronize(c.stream));
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_matvec", e.what());
    }
}

int erosion_geometry(int64_t npts, int64_t nbods, int64_t nouter, const double* xx,
                     const double* yy, double* out) {
    try {
        check_sizes(npts, nbods, nouter);
        if (!out || (nbods > 0 && (!xx || !yy))) throw stokes::Error("null argument");
        Context& c = stokes::ctx();
        c.status = 0;
        stokes::build_geometry(c, static_cast<int>(npts), static_cast<int>(nbods),
                               static_cast<int>(nouter), xx, yy);
        const stokes::Geometry& g = c.g;
        const size_t np = static_cast<size_t>(g.npts), m = static_cast<size_t>(g.m);
        const double* src[4] = {g.px.p, g.py.p, g.cur.p, g.speed.p};
        for (int a = 0; a < 4; ++a)
            STOKES_HIP(hipMemcpyAsync(out + a * np, src[a], np * sizeof(double),
                                      hipMemcpyDeviceToHost, c.stream));
        if (m > 0) {
            STOKES_HIP(hipMemcpyAsync(out + 4 * np, g.cx.p, m * sizeof(double),
                                      hipMemcpyDeviceToHost, c.stream));
            STOKES_HIP(hipMemcpyAsync(out + 4 * np + m, g.cy.p, m * sizeof(double),
                                      hipMemcpyDeviceToHost, c.stream));
        }
        STOKES_HIP(hipStreamSynchronize(c.stream));
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_geometry", e.what());
    }
}

int erosion_time_matvec(int64_t npts, int64_t nbods, int64_t nouter, int64_t ibary,
                        const double* xx, const double* yy, const double* den, int64_t reps,
                        double* out) {
    try {
        check_sizes(npts, nbods, nouter);
        if (!den || !out || reps < 1 || (nbods > 0 && (!xx || !yy)))
            throw stokes::Error("invalid argument");
        Context& c = stokes::ctx();
        c.status = 0;
        const int n = static_cast<int>(npts), m = static_cast<int>(nbods),
                  no = static_cast<int>(nouter);
        stokes::build_geometry(c, n, m, no, xx, yy);
        stokes::build_near_list(c, static_cast<int>(ibary));
        const size_t nunk = static_cast<size_t>(c.g.nunk);
        c.xvec.ensure(std::max(nunk, c.xvec.cap));
        c.rvec.ensure(std::max(nunk, c.rvec.cap));
        STOKES_HIP(hipMemcpyAsync(c.xvec.p, den, nunk * sizeof(double), hipMemcpyHostToDevice,
                                  c.stream));
        c.ff_fp32_now = c.precision == 2;
        // the direct far field unless the cluster far field is forced
        // (then its whole application is the far-field time)
        const bool fmm = c.ff_force == stokes::farfield_shape_count() - 1;
        const int fmm_setting = c.ff_fmm;
        c.ff_fmm = fmm ? 1 : 0;
        if (fmm) c.ff_fp32_now = false;
        stokes::apply_operator(c, c.xvec.p, c.rvec.p, static_cast<int>(ibary));  // warm-up
        STOKES_HIP(hipStreamSynchronize(c.stream));
        // whole operator applications, back to back
        hipEvent_t e0, e1;
        STOKES_HIP(hipEventCreate(&e0));
        STOKES_HIP(hipEventCreate(&e1));
        STOKES_HIP(hipEventRecord(e0, c.stream));
        for (int64_t r = 0; r < reps; ++r)
            stokes::apply_operator(c, c.xvec.p, c.rvec.p, static_cast<int>(ibary));
        STOKES_HIP(hipEventRecord(e1, c.stream));
        STOKES_HIP(hipEventSynchronize(e1));
        float tot = 0.f;
        STOKES_HIP(hipEventElapsedTime(&tot, e0, e1));
        STOKES_HIP(hipEventDestroy(e0));
        STOKES_HIP(hipEventDestroy(e1));
        // per-kernel durations
        const bool was = c.time_kernels;
        stokes::KernelTimer& tff = fmm ? c.t_fmm : (c.ff_fp32_now ? c.t_farfield32 : c.t_farfield);
        tff.reset();
        c.t_near.reset();
        c.time_kernels = true;
        for (int64_t r = 0; r < reps; ++r)
            stokes::apply_operator(c, c.xvec.p, c.rvec.p, static_cast<int>(ibary));
        tff.harvest();
        c.t_near.harvest();
        c.time_kernels = was;
        c.ff_fp32_now = false;
        c.ff_fmm = fmm_setting;
        out[0] = static_cast<double>(tot) / static_cast<double>(reps);
        out[1] = tff.ms / static_cast<double>(std::max<long long>(1, tff.launches));
        const double np = static_cast<double>(c.g.npts);
        out[2] = np * (np - 1.0);
        out[3] = static_cast<double>(c.near.nentries);
        out[4] = c.t_near.launches ? c.t_near.ms / static_cast<double>(c.t_near.launches) : 0.0;
        out[5] = static_cast<double>(c.ff_splits);
        out[6] = c.near.gather_bytes;
        return 0;
    } catch (const std::exception& e) {
        Context& c = stokes::ctx();
        c.time_kernels = false;
        c.ff_fp32_now = false;
        c.ff_fmm = 1;
        return record_error("erosion_time_matvec", e.what());
    }
}

int erosion_set_precision(int mode) {
    try {
        if (mode < 0 || mode > 2) return -1;
        stokes::ctx().precision = mode;
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_set_precision", e.what());
    }
}

int erosion_set_mixed_switch(double factor) {
    try {
        if (!(factor >= 1.0)) return -1;
        stokes::ctx().mixed_switch = factor;
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_set_mixed_switch", e.what());
    }
}

int erosion_set_guess_space(int space) {
    try {
        if (space != 0 && space != 1) return -1;
        stokes::ctx().guess_space = space;
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_set_guess_space", e.what());
    }
}

int erosion_get_guess_space(void) {
    try {
        return stokes::ctx().guess_space;
    } catch (...) {
        return -1;
    }
}

int erosion_last_preconditioned(double* u, int64_t n) {
    try {
        Context& c = stokes::ctx();
        if (!u || !c.u_valid || n != c.g.nunk) return -1;
        STOKES_HIP(hipMemcpyAsync(u, c.uvec.p, sizeof(double) * static_cast<size_t>(n),
                                  hipMemcpyDeviceToHost, c.stream));
        STOKES_HIP(hipStreamSynchronize(c.stream));
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_last_preconditioned", e.what());
    }
}

int erosion_get_precision(void) {
    try {
        return stokes::ctx().precision;
    } catch (...) {
        return -1;
    }
}

int erosion_set_preconditioner(int kind, int body_modes, int wall_modes) {
    try {
        if (kind != 0 && kind != 1) return -1;
        if (kind == 1 && (body_modes < 1 || body_modes > 32 || wall_modes < 1 || wall_modes > 32))
            return -1;
        Context& c = stokes::ctx();
        c.pc_kind = kind;
        if (kind == 1) {
            c.pc_body_modes = body_modes;
            if (c.pc_wall_modes != wall_modes) c.pc_wall_no = -1;
            c.pc_wall_modes = wall_modes;
        }
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_set_preconditioner", e.what());
    }
}

int erosion_rungekutta2(int64_t npts, int64_t nbods, int64_t nouter, const double* params,
                        double* theta, double* len, double* xsm, double* ysm,
                        const double* guess, int64_t* nbods_out, double* info) {
    try {
        if (!params || !nbods_out || !info || (nbods > 0 && (!theta || !len || !xsm || !ysm)))
            throw stokes::Error("null argument");
        check_sizes(npts, nbods, nouter);
        stokes::StepParams P;
        P.dt = params[0];
        P.eps = params[1];
        P.sigma = params[2];
        P.fixarea = params[3];
        P.fixpdrop = params[4] != 0.0 ? 1 : 0;
        P.ibc = static_cast<int>(params[5]);
        P.ibary = static_cast<int>(params[6]);
        P.maxl = static_cast<int>(params[7]);
        P.warm = static_cast<int>(params[8]);
        if (!(P.dt > 0.0) || P.warm < 0 || P.warm > 2 || P.maxl < 1)
            throw stokes::Error("invalid step parameters");
        Context& c = stokes::ctx();
        c.status = 0;
        c.last_error.clear();
        stokes::StepReport rep;
        int mo = static_cast<int>(nbods);
        stokes::rk2_step(c, P, static_cast<int>(npts), static_cast<int>(nouter),
                         static_cast<int>(nbods), theta, len, xsm, ysm, guess, &mo, rep);
        *nbods_out = mo;
        info[0] = static_cast<double>(rep.iters[0]);
        info[1] = static_cast<double>(rep.iters[1]);
        info[2] = rep.deleted;
        info[3] = rep.umax[0];
        info[4] = rep.umax[1];
        info[5] = rep.true_relres_max;
        info[6] = rep.positive_mterm;
        info[7] = rep.flux_warnings;
        info[8] = rep.warm_starts;
        info[9] = rep.given_guesses;
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_rungekutta2", e.what());
    }
}

int erosion_stepper_history(int64_t* out) {
    if (!out) return -1;
    try {
        Context& c = stokes::ctx();
        out[0] = c.sp.hist_count;
        out[1] = c.sp.hist_m;
        out[2] = c.sp.hist_n;
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_stepper_history", e.what());
    }
}

int erosion_stepper_reset(void) {
    try {
        Context& c = stokes::ctx();
        c.sp.hist_count = 0;
        c.sp.hist_m = c.sp.hist_n = -1;
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_stepper_reset", e.what());
    }
}

int erosion_spectral(int op, int64_t npts, int64_t rows, const double* in, const double* fac,
                     double* out) {
    try {
        if (!in || !out || ((op == 2 || op == 3) && !fac)) throw stokes::Error("null argument");
        if (rows < 1 || npts < 2) throw stokes::Error("invalid sizes");
        stokes::spectral_rows(stokes::ctx(), op, static_cast<int>(npts), static_cast<int>(rows),
                              in, fac, out);
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_spectral", e.what());
    }
}

int erosion_set_farfield_shape(int index) {
    try {
        if (index < -1 || index >= stokes::farfield_shape_count()) return -1;
        stokes::ctx().ff_force = index;
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_set_farfield_shape", e.what());
    }
}

int erosion_farfield_shape_count(void) { return stokes::farfield_shape_count(); }

int erosion_set_fmm_one_stream(int on) {
    try {
        if (on != 0 && on != 1) return -1;
        Context& c = stokes::ctx();
        STOKES_HIP(hipStreamSynchronize(c.stream));
        if (c.fmm.side) STOKES_HIP(hipStreamSynchronize(c.fmm.side));
        c.fmm.one_stream = on != 0;
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_set_fmm_one_stream", e.what());
    }
}

int erosion_set_farfield_method(int method) {
    try {
        if (method != 0 && method != 1) return -1;
        Context& c = stokes::ctx();
        c.ff_method = method;
        c.ff_fmm = method;
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_set_farfield_method", e.what());
    }
}

int erosion_set_graphs(int on) {
    try {
        if (on != 0 && on != 1) return -1;
        Context& c = stokes::ctx();
        c.use_graphs = on;
        if (!on) c.graphs.clear();
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_set_graphs", e.what());
    }
}

int erosion_graph_stats(long long* out) {
    if (!out) return -1;
    try {
        Context& c = stokes::ctx();
        out[0] = c.use_graphs;
        out[1] = c.graphs.captures;
        out[2] = c.graphs.replays;
        out[3] = static_cast<long long>(c.graphs.exec.size());
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_graph_stats", e.what());
    }
}

int erosion_get_preconditioner(int* out) {
    if (!out) return -1;
    try {
        Context& c = stokes::ctx();
        out[0] = c.pc_kind;
        out[1] = c.pc_body_modes;
        out[2] = c.pc_wall_modes;
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_get_preconditioner", e.what());
    }
}

int erosion_precond_apply(int64_t npts, int64_t nbods, int64_t nouter, const double* xx,
                          const double* yy, const double* r, double* z) {
    try {
        check_sizes(npts, nbods, nouter);
        if (!r || !z || (nbods > 0 && (!xx || !yy))) throw stokes::Error("null argument");
        Context& c = stokes::ctx();
        if (c.pc_kind == 0) throw stokes::Error("no preconditioner selected");
        c.status = 0;
        const int n = static_cast<int>(npts), m = static_cast<int>(nbods),
                  no = static_cast<int>(nouter);
        stokes::build_geometry(c, n, m, no, xx, yy);
        stokes::precond_setup(c);
        const size_t nunk = static_cast<size_t>(c.g.nunk);
        c.xvec.ensure(std::max(nunk, c.xvec.cap));
        c.rvec.ensure(std::max(nunk, c.rvec.cap));
        STOKES_HIP(hipMemcpyAsync(c.xvec.p, r, nunk * sizeof(double), hipMemcpyHostToDevice,
                                  c.stream));
        stokes::precond_apply(c, c.xvec.p, c.rvec.p);
        STOKES_HIP(hipMemcpyAsync(z, c.rvec.p, nunk * sizeof(double), hipMemcpyDeviceToHost,
                                  c.stream));
        STOKES_HIP(hipStreamSynchronize(c.stream));
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_precond_apply", e.what());
    }
}

int erosion_mlp_param_count(int nin, int nhid, int nlayers, int nout, int64_t* count) {
    try {
        if (!count) throw stokes::Error("null argument");
        *count = static_cast<int64_t>(stokes::mlp_param_count(nin, nhid, nlayers, nout));
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_mlp_param_count", e.what());
    }
}

int erosion_mlp_forward(int nin, int nhid, int nlayers, int nout, const float* params,
                        int64_t nsamples, const float* x, float* y) {
    try {
        if (!params || (nsamples > 0 && (!x || !y))) throw stokes::Error("null argument");
        if (nsamples < 0 || nsamples > (int64_t{1} << 30)) throw stokes::Error("invalid sample count");
        Context& c = stokes::ctx();
        stokes::mlp_forward(c, nin, nhid, nlayers, nout, params, static_cast<int>(nsamples), x, y);
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_mlp_forward", e.what());
    }
}

int erosion_mlp_forward_noalloc(int nin, int nhid, int nlayers, int nout, const float* params,
                                int64_t nsamples, const float* x, float* y, void* stream) {
    // no library context here (creating one would create a stream): errors
    // go to stderr and to the message erosion_noalloc_last_error returns
    try {
        if (!params || (nsamples > 0 && (!x || !y))) throw stokes::Error("null argument");
        if (nsamples < 0 || nsamples > (int64_t{1} << 30)) throw stokes::Error("invalid sample count");
        stokes::mlp_forward_noalloc(nin, nhid, nlayers, nout, params, static_cast<int>(nsamples), x,
                                    y, static_cast<hipStream_t>(stream));
        return 0;
    } catch (const std::exception& e) {
        noalloc_error() = std::string("erosion_mlp_forward_noalloc: ") + e.what();
        std::fprintf(stderr, "libstokes_hip: %s\n", noalloc_error().c_str());
        return -1;
    }
}

const char* erosion_noalloc_last_error(void) { return noalloc_error().c_str(); }

int erosion_mlp_gradients(int nin, int nhid, int nlayers, int nout, const float* params,
                          int64_t nsamples, const float* x, const float* t, float* grads,
                          double* loss) {
    try {
        if (!params || !x || !t || !grads || !loss) throw stokes::Error("null argument");
        if (nsamples < 1 || nsamples > (int64_t{1} << 24)) throw stokes::Error("invalid sample count");
        Context& c = stokes::ctx();
        *loss = stokes::mlp_gradients(c, nin, nhid, nlayers, nout, params,
                                      static_cast<int>(nsamples), x, t, grads);
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_mlp_gradients", e.what());
    }
}

int erosion_mlp_train(int nin, int nhid, int nlayers, int nout, float* params, int64_t nsamples,
                      const float* x, const float* t, int epochs, int batch, float lr,
                      uint64_t seed, double* epoch_loss) {
    try {
        if (!params || !x || !t) throw stokes::Error("null argument");
        if (nsamples < 1 || nsamples > (int64_t{1} << 26)) throw stokes::Error("invalid sample count");
        Context& c = stokes::ctx();
        stokes::mlp_train(c, nin, nhid, nlayers, nout, params, static_cast<int>(nsamples), x, t,
                          epochs, batch, lr, seed, epoch_loss);
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_mlp_train", e.what());
    }
}

namespace {
// 8 independent FP64 FMA chains per lane, `iters` iterations; one double out
// per thread (out holds gridDim.x * 256)
__global__ __launch_bounds__(256) void k_fp64_chains(double* __restrict__ out, int iters, double a,
                                                     double b) {
    double x[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) x[q] = threadIdx.x * 1e-9 + q * 1e-9;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = __builtin_fma(x[q], a, b);
    }
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) s += x[q];
    out[static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x] = s;
}
}  // namespace

int erosion_measure_fp64(int64_t blocks, int64_t iters, double* ms) {
    try {
        if (blocks < 1 || blocks > (1 << 20) || iters < 1 || iters > (1 << 24) || !ms)
            throw stokes::Error("invalid argument");
        Context& c = stokes::ctx();
        stokes::DevBuf<double> out;
        out.ensure(static_cast<size_t>(blocks) * 256);
        hipEvent_t e0, e1;
        STOKES_HIP(hipEventCreate(&e0));
        STOKES_HIP(hipEventCreate(&e1));
        float t = 0.f;
        for (int rep = 0; rep < 3; ++rep) {
            STOKES_HIP(hipEventRecord(e0, c.stream));
            k_fp64_chains<<<static_cast<int>(blocks), 256, 0, c.stream>>>(
                out.p, static_cast<int>(iters), 0.999999, 1e-7);
            STOKES_LAUNCH_CHECK();
            STOKES_HIP(hipEventRecord(e1, c.stream));
            STOKES_HIP(hipEventSynchronize(e1));
            STOKES_HIP(hipEventElapsedTime(&t, e0, e1));
        }
        STOKES_HIP(hipEventDestroy(e0));
        STOKES_HIP(hipEventDestroy(e1));
        *ms = static_cast<double>(t);
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_measure_fp64", e.what());
    }
}

int erosion_measure_copy(int64_t bytes, int64_t reps, double* ms) {
    try {
        if (bytes < 1 || reps < 1 || !ms) throw stokes::Error("invalid argument");
        Context& c = stokes::ctx();
        stokes::DevBuf<char> a, b;
        a.ensure(static_cast<size_t>(bytes));
        b.ensure(static_cast<size_t>(bytes));
        STOKES_HIP(hipMemsetAsync(a.p, 1, static_cast<size_t>(bytes), c.stream));
        hipEvent_t e0, e1;
        STOKES_HIP(hipEventCreate(&e0));
        STOKES_HIP(hipEventCreate(&e1));
        STOKES_HIP(hipEventRecord(e0, c.stream));
        for (int64_t r = 0; r < reps; ++r)
            STOKES_HIP(hipMemcpyAsync(b.p, a.p, static_cast<size_t>(bytes),
                                      hipMemcpyDeviceToDevice, c.stream));
        STOKES_HIP(hipEventRecord(e1, c.stream));
        STOKES_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        STOKES_HIP(hipEventElapsedTime(&t, e0, e1));
        STOKES_HIP(hipEventDestroy(e0));
        STOKES_HIP(hipEventDestroy(e1));
        *ms = static_cast<double>(t) / static_cast<double>(reps);
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_measure_copy", e.what());
    }
}

int erosion_kernel_timing(int enable) {
    try {
        if (enable < 0) throw stokes::Error("invalid timing period");
        Context& c = stokes::ctx();
        if (enable) {
            for (stokes::KernelTimer* t : {&c.t_farfield, &c.t_near, &c.t_stress_eval, &c.t_gs,
                                           &c.t_precond, &c.t_farfield32, &c.t_fmm,
                                           &c.t_fmm_direct, &c.t_fmm_expand})
                t->reset();
        } else {
            for (stokes::KernelTimer* t : {&c.t_farfield, &c.t_near, &c.t_stress_eval, &c.t_gs,
                                           &c.t_precond, &c.t_farfield32, &c.t_fmm,
                                           &c.t_fmm_direct, &c.t_fmm_expand})
                t->harvest();
        }
        c.time_kernels = enable != 0;
        c.time_period = enable > 1 ? enable : 1;
        c.time_tick = 0;
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_kernel_timing", e.what());
    }
}

int erosion_kernel_timers(double* out) {
    if (!out) return -1;
    try {
        Context& c = stokes::ctx();
        stokes::KernelTimer* ts[9] = {&c.t_farfield, &c.t_near, &c.t_stress_eval, &c.t_gs,
                                      &c.t_precond, &c.t_farfield32, &c.t_fmm, &c.t_fmm_direct,
                                      &c.t_fmm_expand};
        for (int i = 0; i < 9; ++i) {
            ts[i]->harvest();
            out[3 * i + 0] = ts[i]->ms;
            out[3 * i + 1] = static_cast<double>(ts[i]->launches);
            out[3 * i + 2] = ts[i]->work;
        }
        return 0;
    } catch (const std::exception& e) {
        return record_error("erosion_kernel_timers", e.what());
    }
}

}  // extern "C"
