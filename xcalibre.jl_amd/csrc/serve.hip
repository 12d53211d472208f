// Serving engine: many independent steady SIMPLE cases on one device with
// continuous admission.  A case (request) is built on the host when it is
// submitted, waits in a FIFO queue, and is admitted to a free slot as soon as
// its device footprint fits the HBM budget left by the cases in flight.  Its
// mesh and fields stay resident on the device until it has converged (or hit
// its iteration limit); then its fields are copied back and its memory is
// released, which may admit the next waiting case.
//
// Each slot is a host thread with its own context (HIP stream), so the work
// of the cases in flight overlaps on the device.  A case's arithmetic does not
// depend on which other cases are in flight (separate buffers, deterministic
// kernels), so its result is the same alone and in the mix.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "xcal_internal.h"

using namespace xcal;

namespace {

struct Request {
  int64_t id = 0;
  std::shared_ptr<PreparedMesh> mesh;
  xcal_simple_setup setup{};
  std::vector<xcal_bc> bcU, bcp;
  bool turb = false;
  xcal_turbulence_setup tset{};
  std::vector<xcal_bc> bcK, bcW, bcN;
  // initial fields indexed by XCAL_SIMPLE_* (empty: zero start), then results
  std::vector<double> f0[6], f[6];
  int64_t footprint = 0, cells = 0;
  int32_t state = XCAL_REQ_QUEUED, iterations = 0;
  double t_submit = 0, t_start = 0, t_end = 0;
  std::string error;
};

}  // namespace

struct xcal_server {
  xcal_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv;  // queue / budget / completion changes
  std::deque<int64_t> queue;
  std::map<int64_t, std::shared_ptr<Request>> reqs;
  int64_t next_id = 1, in_use = 0, running = 0;
  bool stop = false;
  std::vector<std::thread> workers;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

void run_request(xcal_ctx* ctx, const xcal_server_config& cfg, Request& r) {
  xcal_mesh* m = upload_prepared(ctx, *r.mesh);
  std::unique_ptr<xcal_mesh, int (*)(xcal_mesh*)> mg(m, xcal_mesh_destroy);
  xcal_simple_setup s = r.setup;
  s.bc_U = r.bcU.data();
  s.bc_p = r.bcp.data();
  xcal_simple* sim = nullptr;
  int rc = xcal_simple_create(m, &s, &sim);
  if (rc) throw Error(rc, xcal_last_error());
  std::unique_ptr<xcal_simple, int (*)(xcal_simple*)> sg(sim, xcal_simple_destroy);
  simple_stop_nonfinite(sim, true);
  if (r.turb) {
    xcal_turbulence_setup t = r.tset;
    t.bc_k = r.bcK.data();
    t.bc_omega = r.bcW.data();
    t.bc_nut = r.bcN.data();
    if ((rc = xcal_simple_set_turbulence(sim, &t))) throw Error(rc, xcal_last_error());
  }
  for (int w = 0; w < 6; ++w)
    if (!r.f0[w].empty() && (rc = xcal_simple_set_field(sim, w, r.f0[w].data())))
      throw Error(rc, xcal_last_error());
  std::vector<double> res(4 * (size_t)cfg.max_iterations);
  int32_t done = 0;
  if ((rc = xcal_simple_run(sim, cfg.max_iterations, res.data(), &done, nullptr)))
    throw Error(rc, xcal_last_error());
  r.iterations = done;
  if (simple_diverged(sim)) throw Error(XCAL_ERR_STATE, "diverged: non-finite residual");
  for (int w : {XCAL_SIMPLE_U, XCAL_SIMPLE_P, XCAL_SIMPLE_K, XCAL_SIMPLE_OMEGA, XCAL_SIMPLE_NUT}) {
    if (w >= XCAL_SIMPLE_K && !r.turb) break;
    r.f[w].resize((w == XCAL_SIMPLE_U ? 3 : 1) * (size_t)r.cells);
    if ((rc = xcal_simple_get_field(sim, w, r.f[w].data()))) throw Error(rc, xcal_last_error());
  }
}

void worker(xcal_server* S) {
  xcal_ctx* ctx = nullptr;
  if (xcal_ctx_create(S->cfg.device, &ctx) != 0) ctx = nullptr;
  for (;;) {
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv.wait(lk, [&] {
        if (S->stop) return true;
        if (S->queue.empty()) return false;
        const auto& q = S->reqs[S->queue.front()];
        return q->footprint <= S->cfg.hbm_budget - S->in_use;
      });
      if (S->stop) break;
      r = S->reqs[S->queue.front()];
      S->queue.pop_front();
      S->in_use += r->footprint;
      S->running += 1;
      r->state = XCAL_REQ_RUNNING;
      r->t_start = S->now();
    }
    std::string err;
    try {
      if (!ctx) throw Error(XCAL_ERR_HIP, "slot could not create a device context");
      run_request(ctx, S->cfg, *r);
    } catch (const std::exception& e) {
      err = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->in_use -= r->footprint;
      S->running -= 1;
      r->mesh.reset();
      r->error = err;
      r->state = err.empty() ? XCAL_REQ_DONE : XCAL_REQ_FAILED;
      r->t_end = S->now();
    }
    S->cv.notify_all();
  }
  if (ctx) xcal_ctx_destroy(ctx);
}

}  // namespace

extern "C" {

int xcal_server_create(const xcal_server_config* cfg, xcal_server** out) {
  return guarded([&] {
    XCAL_CHECK(cfg && out, "NULL argument");
    XCAL_CHECK(cfg->slots >= 1 && cfg->slots <= 16, "slots must be 1..16");
    XCAL_CHECK(cfg->hbm_budget > 0, "HBM budget must be positive");
    XCAL_CHECK(cfg->max_iterations >= 1, "max_iterations must be >= 1");
    int ndev = 0;
    XCAL_HIP(hipGetDeviceCount(&ndev));
    XCAL_CHECK(cfg->device >= 0 && cfg->device < ndev, "device index out of range");
    auto S = std::make_unique<xcal_server>();
    S->cfg = *cfg;
    for (int i = 0; i < cfg->slots; ++i) S->workers.emplace_back(worker, S.get());
    *out = S.release();
    return 0;
  });
}

int xcal_server_destroy(xcal_server* S) {
  return guarded([&] {
    if (!S) return 0;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stop = true;
    }
    S->cv.notify_all();
    for (auto& t : S->workers) t.join();  // cases in flight run to completion first
    delete S;
    return 0;
  });
}

int xcal_server_submit(xcal_server* S, const xcal_mesh_host* mesh, const xcal_simple_setup* setup,
                       const xcal_turbulence_setup* turb, const double* U0, const double* p0,
                       const double* k0, const double* omega0, const double* nut0, int64_t* id) {
  return guarded([&] {
    XCAL_CHECK(S && mesh && setup && id, "NULL argument");
    XCAL_CHECK(setup->n_bcs == mesh->n_patches, "one boundary condition per patch is required");
    const bool rans = turb && turb->model != XCAL_TURB_LAMINAR;
    XCAL_CHECK(!rans || (turb->bc_k && turb->bc_omega && turb->bc_nut),
               "k, omega and nut need boundary conditions");
    XCAL_CHECK(rans || !(k0 || omega0 || nut0), "k, omega and nut fields need a turbulence model");
    auto r = std::make_shared<Request>();
    r->t_submit = S->now();
    r->mesh = prepare_mesh(&mesh, 1, false);  // host-side build happens at submission
    r->cells = prepared_cells(*r->mesh);
    r->footprint = prepared_footprint(*r->mesh, rans);
    XCAL_CHECK(r->footprint <= S->cfg.hbm_budget, "case does not fit the server's HBM budget");
    r->setup = *setup;
    r->bcU.assign(setup->bc_U, setup->bc_U + setup->n_bcs);
    r->bcp.assign(setup->bc_p, setup->bc_p + setup->n_bcs);
    if (rans) {
      const int np = mesh->n_patches;
      r->turb = true;
      r->tset = *turb;
      r->bcK.assign(turb->bc_k, turb->bc_k + np);
      r->bcW.assign(turb->bc_omega, turb->bc_omega + np);
      r->bcN.assign(turb->bc_nut, turb->bc_nut + np);
    }
    const double* init[6] = {U0, p0, nullptr, k0, omega0, nut0};
    for (int w = 0; w < 6; ++w)
      if (init[w]) r->f0[w].assign(init[w], init[w] + (w == XCAL_SIMPLE_U ? 3 : 1) * r->cells);
    {
      std::lock_guard<std::mutex> lk(S->mu);
      XCAL_CHECK(!S->stop, "server is shutting down");
      r->id = S->next_id++;
      S->reqs[r->id] = r;
      S->queue.push_back(r->id);
      *id = r->id;
    }
    S->cv.notify_all();
    return 0;
  });
}

int xcal_server_poll(xcal_server* S, int64_t id, xcal_request_status* st) {
  return guarded([&] {
    XCAL_CHECK(S && st, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    XCAL_CHECK(it != S->reqs.end(), "unknown request id");
    const Request& r = *it->second;
    st->state = r.state;
    st->iterations = r.iterations;
    st->cells = r.cells;
    st->footprint = r.footprint;
    st->t_submit = r.t_submit;
    st->t_start = r.t_start;
    st->t_end = r.t_end;
    return 0;
  });
}

int xcal_server_wait(xcal_server* S, int64_t id, double timeout_s) {
  return guarded([&] {
    XCAL_CHECK(S, "NULL argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto finished = [&] {
      if (id > 0) {
        auto it = S->reqs.find(id);
        return it == S->reqs.end() || it->second->state >= XCAL_REQ_DONE;
      }
      return S->queue.empty() && S->running == 0;
    };
    if (timeout_s <= 0) S->cv.wait(lk, finished);
    else if (!S->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), finished))
      return 2;  // timed out
    return 0;
  });
}

int xcal_server_fetch(xcal_server* S, int64_t id, int32_t which, double* host) {
  return guarded([&] {
    XCAL_CHECK(S && host, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    XCAL_CHECK(it != S->reqs.end(), "unknown request id");
    const Request& r = *it->second;
    if (r.state == XCAL_REQ_FAILED) throw Error(XCAL_ERR_STATE, "request failed: " + r.error);
    XCAL_CHECK(r.state == XCAL_REQ_DONE, "request has not finished");
    XCAL_CHECK(which >= 0 && which < 6 && which != XCAL_SIMPLE_MDOTF, "which: U, p, k, omega or nut");
    XCAL_CHECK(!r.f[which].empty(), "the request has no such field (laminar case)");
    std::copy(r.f[which].begin(), r.f[which].end(), host);
    return 0;
  });
}

int xcal_server_release(xcal_server* S, int64_t id) {
  return guarded([&] {
    XCAL_CHECK(S, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    XCAL_CHECK(it != S->reqs.end(), "unknown request id");
    XCAL_CHECK(it->second->state >= XCAL_REQ_DONE, "request has not finished");
    S->reqs.erase(it);
    return 0;
  });
}

}  // extern "C"
