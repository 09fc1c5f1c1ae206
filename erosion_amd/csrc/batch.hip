// Many independent geometries in one call: a pool of worker contexts, each
// with its own HIP stream and device buffers, driven by host threads that
// take geometries from a shared counter.  Solves of different geometries
// then overlap on the GPU (streams map onto the hardware queues), which
// fills the chip when single geometries are too small to.  No counterpart in
// the reference, which solves one geometry per call.
#include <atomic>
#include <mutex>
#include <thread>

#include "common.hpp"

namespace stokes {

namespace {

std::mutex g_pool_mu;
std::vector<Context*> g_pool;

Context* worker_context(size_t i) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    while (g_pool.size() <= i) {
        Context* c = new Context();
        STOKES_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        g_pool.push_back(c);
    }
    return g_pool[i];
}

}  // namespace

void solve_batch(Context& settings, int ngeom, const BatchGeometry* geo, int ibary, int ibc,
                 int maxl, int nworkers, std::string* first_error) {
    nworkers = std::max(1, std::min(nworkers, ngeom));
    std::vector<Context*> ws;
    for (int w = 0; w < nworkers; ++w) {
        Context* c = worker_context(static_cast<size_t>(w));
        c->tol = settings.tol;
        c->precision = settings.precision;
        c->mixed_switch = settings.mixed_switch;
        if (c->pc_kind != settings.pc_kind || c->pc_body_modes != settings.pc_body_modes ||
            c->pc_wall_modes != settings.pc_wall_modes)
            c->pc_wall_no = -1;  // rebuild the kept wall block with the new modes
        c->pc_kind = settings.pc_kind;
        c->pc_body_modes = settings.pc_body_modes;
        c->pc_wall_modes = settings.pc_wall_modes;
        c->use_graphs = settings.use_graphs;
        c->ff_fmm = settings.ff_fmm;
        c->ff_method = settings.ff_method;
        c->ff_force = settings.ff_force;
        if (!c->use_graphs) c->graphs.clear();
        ws.push_back(c);
    }
    std::atomic<int> next{0};
    std::mutex err_mu;
    auto run = [&](Context& c) {
        for (;;) {
            const int gi = next.fetch_add(1);
            if (gi >= ngeom) break;
            const BatchGeometry& b = geo[gi];
            try {
                c.status = 0;
                c.last_error.clear();
                build_geometry(c, b.n, b.m, b.no, b.xx, b.yy);
                build_near_list(c, ibary);
                *b.nits = gmres_solve(c, ibary, ibc, maxl, c.tol, b.den);
                *b.status = c.status;
                if (b.relres) *b.relres = c.last_true_relres;
                if (b.tau && b.m > 0) {
                    const size_t nunk = static_cast<size_t>(c.g.nunk);
                    const size_t nbp = static_cast<size_t>(b.n) * static_cast<size_t>(b.m);
                    c.rvec.ensure(std::max(nunk, c.rvec.cap));
                    c.tau.ensure(nbp);
                    STOKES_HIP(hipMemcpyAsync(c.rvec.p, b.den, nunk * sizeof(double),
                                              hipMemcpyHostToDevice, c.stream));
                    shear_stress(c, c.rvec.p, c.tau.p);
                    STOKES_HIP(hipMemcpyAsync(b.tau, c.tau.p, nbp * sizeof(double),
                                              hipMemcpyDeviceToHost, c.stream));
                    STOKES_HIP(hipStreamSynchronize(c.stream));
                }
            } catch (const std::exception& e) {
                *b.status = -1;
                std::lock_guard<std::mutex> lk(err_mu);
                if (first_error->empty())
                    *first_error = "geometry " + std::to_string(gi) + ": " + e.what();
            }
        }
    };
    if (nworkers == 1) {
        run(*ws[0]);
        return;
    }
    std::vector<std::thread> th;
    for (int w = 0; w < nworkers; ++w) th.emplace_back(run, std::ref(*ws[static_cast<size_t>(w)]));
    for (auto& t : th) t.join();
}

}  // namespace stokes
