// Serving engine: many independent steady SIMPLE cases on one device with
// continuous batching.  A case (request) is built when it is submitted (its
// matrix connectivity on the device, the rest on the host) and waits in a
// FIFO queue.  Each lane (a scheduler thread with its own stream) keeps its
// cases in flight as one batch mesh (the union of their meshes, each case in
// its own numbering) and steps them together: every kernel launch of a
// SIMPLE iteration covers all of them, and each case's Krylov solves stop on
// their own.  A batch iteration lasts as long as its slowest case's solves,
// so with several lanes a case joins the lane whose cases are closest to it
// in size, and the lanes' launches overlap on the device.  After every iteration each case is tested on its own: a converged
// case (the run!/simple! test), one with a non-finite residual, or one at its
// iteration limit leaves the batch with its fields copied back.  Waiting
// cases join as soon as their device footprint fits the HBM budget left by
// the cases in flight.  When the set changes the batch is rebuilt: the
// continuing cases' fields move on the device, bit for bit, into the new
// batch, and only the joining cases take the initial calculations.  A case's
// arithmetic does not depend on which other cases share the batch (separate
// rows, per-case reductions, deterministic kernels), so its result is the same
// alone and in the mix.
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
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
  int32_t state = XCAL_REQ_QUEUED, iterations = 0, converged = 0;
  int32_t lane = -1;  // the lane that admitted it
  double t_submit = 0, t_start = 0, t_end = 0;
  std::string error;
};

// Two requests can share a batch: the same SIMPLE and turbulence settings
// and boundary conditions (compared bytewise, pointers aside).
bool compatible(const Request& a, const Request& b) {
  xcal_simple_setup sa = a.setup, sb = b.setup;
  sa.bc_U = sa.bc_p = sb.bc_U = sb.bc_p = nullptr;
  if (std::memcmp(&sa, &sb, sizeof sa) != 0 || a.turb != b.turb) return false;
  auto same = [](const std::vector<xcal_bc>& x, const std::vector<xcal_bc>& y) {
    return x.size() == y.size() && std::memcmp(x.data(), y.data(), sizeof(xcal_bc) * x.size()) == 0;
  };
  if (!same(a.bcU, b.bcU) || !same(a.bcp, b.bcp)) return false;
  if (!a.turb) return true;
  xcal_turbulence_setup ta = a.tset, tb = b.tset;
  ta.bc_k = ta.bc_omega = ta.bc_nut = tb.bc_k = tb.bc_omega = tb.bc_nut = nullptr;
  return std::memcmp(&ta, &tb, sizeof ta) == 0 && same(a.bcK, b.bcK) && same(a.bcW, b.bcW) &&
         same(a.bcN, b.bcN);
}

}  // namespace

struct xcal_server {
  xcal_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv;  // queue / budget / completion changes
  std::deque<int64_t> queue;
  std::map<int64_t, std::shared_ptr<Request>> reqs;
  int64_t next_id = 1, in_use = 0, running = 0, rebuilds = 0;
  bool stop = false;
  // XCAL_SERVE_REBUILD=1 at creation: rebuild the batch after every iteration
  // (moving every case's state each time; a test of the state transfer)
  bool rebuild_always = false;
  int nlanes = 1;
  std::vector<std::thread> lanes;  // one scheduler thread per lane
  // mesh ingest at submission: the CSR pattern, slot maps and DILU levels of
  // a request's mesh are built on the device, on this context's stream (one
  // submission at a time)
  xcal_ctx* ingest = nullptr;
  std::mutex ingest_mu;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

using Sim = std::unique_ptr<xcal_simple, int (*)(xcal_simple*)>;
using Mesh = std::unique_ptr<xcal_mesh, int (*)(xcal_mesh*)>;

void check_rc(int rc) {
  if (rc) throw Error(rc, xcal_last_error());
}

// The batch of cases in flight: requests in batch order, their device mesh
// and SIMPLE state.
struct Batch {
  std::vector<std::shared_ptr<Request>> cases;
  Mesh mesh{nullptr, xcal_mesh_destroy};
  Sim sim{nullptr, xcal_simple_destroy};
};

// A new batch of `keep` (cases of `old`, continuing) followed by `join`.
void rebuild(xcal_ctx* ctx, Batch& old, const std::vector<std::shared_ptr<Request>>& keep,
             const std::vector<std::shared_ptr<Request>>& join) {
  Batch nb;
  nb.cases = keep;
  nb.cases.insert(nb.cases.end(), join.begin(), join.end());
  if (nb.cases.empty()) {
    old = Batch();
    return;
  }
  std::vector<const PreparedMesh*> parts;
  for (const auto& r : nb.cases) parts.push_back(r->mesh.get());
  auto P = prepare_union(parts);
  nb.mesh.reset(upload_prepared(ctx, *P));
  xcal_mesh* m = nb.mesh.get();
  Request& r0 = *nb.cases[0];
  xcal_simple_setup s = r0.setup;
  s.bc_U = r0.bcU.data();
  s.bc_p = r0.bcp.data();
  xcal_simple* sim = nullptr;
  check_rc(xcal_simple_create(m, &s, &sim));
  nb.sim.reset(sim);
  if (r0.turb) {
    xcal_turbulence_setup t = r0.tset;
    t.bc_k = r0.bcK.data();
    t.bc_omega = r0.bcW.data();
    t.bc_nut = r0.bcN.data();
    check_rc(xcal_simple_set_turbulence(sim, &t));
  }
  // initial fields in the batch's reference numbering (cases concatenated,
  // vector components in blocks over the whole batch): the joining cases'
  // own, zero for the continuing ones (overwritten below)
  const int64_t n = m->n;
  for (int w : {XCAL_SIMPLE_U, XCAL_SIMPLE_P, XCAL_SIMPLE_K, XCAL_SIMPLE_OMEGA, XCAL_SIMPLE_NUT}) {
    if (w >= XCAL_SIMPLE_K && !r0.turb) break;
    const int nc = w == XCAL_SIMPLE_U ? 3 : 1;
    std::vector<double> h((size_t)nc * n, 0.0);
    bool any = false;
    for (size_t k = keep.size(); k < nb.cases.size(); ++k) {
      const Request& r = *nb.cases[k];
      if (r.f0[w].empty()) continue;
      any = true;
      for (int c = 0; c < nc; ++c)
        std::copy(r.f0[w].begin() + c * r.cells, r.f0[w].begin() + (c + 1) * r.cells,
                  h.begin() + c * n + m->case_cell[k]);
    }
    if (any) check_rc(xcal_simple_set_field(sim, w, h.data()));
  }
  simple_initialise(sim);
  for (size_t k = 0; k < keep.size(); ++k) {
    int ko = -1;
    for (size_t j = 0; j < old.cases.size(); ++j)
      if (old.cases[j] == keep[k]) ko = (int)j;
    XCAL_CHECK(ko >= 0, "internal: continuing case not in the old batch");
    simple_copy_case(sim, (int)k, old.sim.get(), ko);
  }
  XCAL_HIP(hipStreamSynchronize(ctx->stream));
  old.sim.reset();
  old.mesh.reset();
  old = std::move(nb);
}

// A lane whose largest case differs from a request by this size factor or
// more is no better for it than an empty lane.
constexpr double kLaneSpread = 4.0;

// The lane that request q joins (S->mu held): among the empty lanes and the
// lanes whose cases share q's settings, the one whose largest running case
// is closest to q in size (|log| of the cell ratio; an empty lane counts as
// log kLaneSpread), the lowest index on a tie; -1 when none can take it.
int best_lane(const xcal_server* S, const Request& q) {
  const int nl = S->nlanes;
  std::vector<int64_t> big(nl, 0);
  std::vector<const Request*> first(nl, nullptr);
  for (const auto& kv : S->reqs) {
    const Request& r = *kv.second;
    if (r.state != XCAL_REQ_RUNNING || r.lane < 0) continue;
    big[r.lane] = std::max(big[r.lane], r.cells);
    if (!first[r.lane]) first[r.lane] = &r;
  }
  int best = -1;
  double bd = INFINITY;
  for (int l = 0; l < nl; ++l) {
    double d = std::log(kLaneSpread);
    if (first[l]) {
      if (!compatible(*first[l], q)) continue;
      d = std::fabs(std::log((double)q.cells / (double)big[l]));
    }
    if (d < bd) {
      bd = d;
      best = l;
    }
  }
  return best;
}

void scheduler(xcal_server* S, int lane) {
  xcal_ctx* ctx = nullptr;
  const bool ok = xcal_ctx_create(S->cfg.device, &ctx) == 0;
  Batch B;
  std::vector<double> res;
  bool dirty = false;  // cases left the batch since it was built
  for (;;) {
    std::vector<std::shared_ptr<Request>> join;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      auto admissible = [&] {
        if (S->stop || S->queue.empty()) return false;
        const Request& q = *S->reqs[S->queue.front()];
        if (q.footprint > S->cfg.hbm_budget - S->in_use) return false;
        if (S->running >= S->cfg.slots) return false;  // cases in flight over all lanes
        return best_lane(S, q) == lane;  // a lane without a context fails what it takes
      };
      if (B.cases.empty() && !dirty)
        S->cv.wait(lk, [&] { return S->stop || admissible(); });
      if (S->stop && B.cases.empty()) break;
      while (admissible()) {
        auto r = S->reqs[S->queue.front()];
        S->queue.pop_front();
        S->in_use += r->footprint;
        S->running += 1;
        r->state = XCAL_REQ_RUNNING;
        r->lane = lane;
        r->t_start = S->now();
        join.push_back(r);
      }
      // the new queue head may be another lane's
      if (!join.empty() && !S->queue.empty()) S->cv.notify_all();
    }
    std::vector<std::shared_ptr<Request>> failed;
    std::string err;
    if (!join.empty() || dirty) {
      try {
        if (!ok) throw Error(XCAL_ERR_HIP, "server could not create a device context");
        std::vector<std::shared_ptr<Request>> keep;
        for (const auto& r : B.cases)
          if (r->state == XCAL_REQ_RUNNING) keep.push_back(r);
        rebuild(ctx, B, keep, join);
        std::lock_guard<std::mutex> lk(S->mu);
        ++S->rebuilds;
      } catch (const std::exception& e) {  // every case of the failed batch fails
        err = e.what();
        for (const auto& r : B.cases)
          if (r->state == XCAL_REQ_RUNNING) failed.push_back(r);
        failed.insert(failed.end(), join.begin(), join.end());
        B = Batch();
      }
      dirty = false;
    }
    std::vector<std::shared_ptr<Request>> done;
    if (err.empty() && !B.cases.empty()) {
      try {
        simple_step(B.sim.get(), res);
        {
          std::lock_guard<std::mutex> lk(S->mu);  // status readers see whole iterations
          for (const auto& r : B.cases) r->iterations += 1;
        }
        for (size_t k = 0; k < B.cases.size(); ++k) {
          Request& r = *B.cases[k];
          if (simple_case_bad(B.sim.get(), (int)k)) {
            r.error = "diverged: non-finite residual";
            done.push_back(B.cases[k]);
            continue;
          }
          const bool conv = simple_case_converged(B.sim.get(), (int)k);
          if (!conv && r.iterations < S->cfg.max_iterations) continue;
          {
            std::lock_guard<std::mutex> lk(S->mu);
            r.converged = conv ? 1 : 0;
          }
          for (int w : {XCAL_SIMPLE_U, XCAL_SIMPLE_P, XCAL_SIMPLE_K, XCAL_SIMPLE_OMEGA,
                        XCAL_SIMPLE_NUT}) {
            if (w >= XCAL_SIMPLE_K && !r.turb) break;
            r.f[w].resize((w == XCAL_SIMPLE_U ? 3 : 1) * (size_t)r.cells);
            simple_case_field(B.sim.get(), w, (int)k, r.f[w].data());
          }
          done.push_back(B.cases[k]);
        }
      } catch (const std::exception& e) {
        err = e.what();
        for (const auto& r : B.cases) failed.push_back(r);
        done.clear();
        B = Batch();
      }
    }
    {
      std::lock_guard<std::mutex> lk(S->mu);
      const double t = S->now();
      for (const auto& r : failed) r->error = err;
      for (auto* list : {&done, &failed})
        for (const auto& r : *list) {
          if (r->state != XCAL_REQ_RUNNING) continue;
          S->in_use -= r->footprint;
          S->running -= 1;
          r->mesh.reset();
          r->state = r->error.empty() ? XCAL_REQ_DONE : XCAL_REQ_FAILED;
          r->t_end = t;
        }
    }
    if (!done.empty() || !failed.empty()) {
      dirty = !B.cases.empty();
      S->cv.notify_all();
    }
    if (S->rebuild_always && !B.cases.empty()) dirty = true;
  }
  B = Batch();
  if (ctx) xcal_ctx_destroy(ctx);
}

}  // namespace

extern "C" {

int xcal_server_create(const xcal_server_config* cfg, xcal_server** out) {
  return guarded([&] {
    XCAL_CHECK(cfg && out, "NULL argument");
    XCAL_CHECK(cfg->slots >= 1 && cfg->slots <= 256, "slots must be 1..256");
    XCAL_CHECK(cfg->hbm_budget > 0, "HBM budget must be positive");
    XCAL_CHECK(cfg->max_iterations >= 1, "max_iterations must be >= 1");
    XCAL_CHECK(cfg->lanes >= 0 && cfg->lanes <= 16, "lanes must be 0..16");
    int ndev = 0;
    XCAL_HIP(hipGetDeviceCount(&ndev));
    XCAL_CHECK(cfg->device >= 0 && cfg->device < ndev, "device index out of range");
    auto S = std::make_unique<xcal_server>();
    S->cfg = *cfg;
    check_rc(xcal_ctx_create(cfg->device, &S->ingest));
    const char* rb = std::getenv("XCAL_SERVE_REBUILD");
    S->rebuild_always = rb && *rb == '1';
    S->nlanes = std::max(1, (int)cfg->lanes);
    for (int l = 0; l < S->nlanes; ++l) S->lanes.emplace_back(scheduler, S.get(), l);
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
    for (auto& t : S->lanes) t.join();  // the batches in flight run to completion first
    if (S->ingest) xcal_ctx_destroy(S->ingest);
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
    {  // the mesh is built at submission (batches carry no cone form)
      std::lock_guard<std::mutex> lk(S->ingest_mu);
      r->mesh = prepare_mesh(&mesh, 1, false, S->ingest, false);
    }
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
    st->converged = r.converged;
    st->lane = r.lane;
    st->batch_rebuilds = S->rebuilds;
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
