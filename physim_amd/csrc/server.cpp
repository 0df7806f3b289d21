// server.cpp -- request loop for many independent solves on one GPU
// (pgn_server_*), with continuous batching.
//
// Requests (tet meshes and thin shells with their boundary conditions) are
// copied into host staging at submit time and queued.  Each lane is a worker
// thread with its own HIP stream that keeps up to cohorts_per_lane fused
// batches ("cohorts": pgn_batch, one launch per kernel covers all of their
// requests, every request keeps its own Newton state) resident and advances
// them one Newton iteration at a time.  Between iterations it
//   - admits queued requests as a new cohort when their estimated device
//     footprint fits the server's HBM budget (at most max_batch of them):
//     with group_by_size the cheapest queued request (the oldest once it has
//     waited max_wait_s) and requests of at most twice its footprint,
//     otherwise arrival order; a request of at least solo_nodes nodes
//     fills the device alone and runs as a batch of one;
//   - hands every request back the moment its own solve ends (its positions
//     downloaded, its waiter woken), while its cohort-mates go on;
//   - releases a cohort's memory when its last request has ended.
// So a request that arrives while others are on the GPU starts within one
// Newton iteration of some lane, and finished requests never wait for
// their batch-mates.  Lanes overlap on the device through their streams.
// Each result is the one the request gets alone (per-instance Newton,
// block-diagonal linear solve).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "batch.hpp"

using namespace pgn;

namespace pgn {
int api_guard_set(const char* msg, int code);
}

namespace {
using Clock = std::chrono::steady_clock;

struct Request {
  long ticket = 0;
  pgn_batch_inst inst;
  size_t bytes = 0;
  Clock::time_point t_submit, t_start, t_done;
  bool done = false;
  int status = PGN_OK;
  std::string error;
  std::vector<double> x;
  pgn_newton_result res{};
};
}  // namespace

struct pgn_server {
  int device = 0;
  pgn_server_opts opts{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::map<long, std::shared_ptr<Request>> all;
  size_t used = 0;
  long next_ticket = 1;
  bool stop = false;
  int waiters = 0;  // threads inside pgn_server_wait
  std::vector<std::thread> lanes;
  std::vector<pgn_ctx*> ctxs;
  long batches = 0, solved = 0;

  void lane_main(int lane);
  // index in `pending` of the request a free lane starts with, or -1 when
  // none may start now (caller holds mu)
  long pick_head() const;
  std::vector<std::shared_ptr<Request>> admit();
  void finish(const std::shared_ptr<Request>& r, int code, const std::string& err);
};

long pgn_server::pick_head() const {
  if (pending.empty()) return -1;
  auto fits = [&](long i) { return used + pending[i]->bytes <= opts.hbm_budget_bytes; };
  if (!opts.group_by_size) return fits(0) ? 0 : -1;
  const double waited =
      std::chrono::duration<double>(Clock::now() - pending.front()->t_submit).count();
  if (waited >= opts.max_wait_s) return fits(0) ? 0 : -1;  // no starvation
  long best = 0;
  for (long i = 1; i < (long)pending.size(); ++i)
    if (pending[i]->bytes < pending[best]->bytes) best = i;
  return fits(best) ? best : -1;
}

// Takes the requests of a new cohort from `pending` (caller holds mu);
// empty when none may start now.
std::vector<std::shared_ptr<Request>> pgn_server::admit() {
  std::vector<std::shared_ptr<Request>> take;
  const long head = pick_head();
  if (head < 0) return take;
  auto solo = [&](long i) { return pending[i]->inst.n >= opts.solo_nodes; };
  if (solo(head)) {
    // a request that fills the device alone runs as a batch of one (the
    // single-model PCG)
    used += pending[head]->bytes;
    take.push_back(pending[head]);
    pending.erase(pending.begin() + head);
    return take;
  }
  if (!opts.group_by_size) {
    // arrival order while the footprint fits
    while (!pending.empty() && (int)take.size() < opts.max_batch && !solo(0) &&
           used + pending.front()->bytes <= opts.hbm_budget_bytes) {
      used += pending.front()->bytes;
      take.push_back(pending.front());
      pending.pop_front();
    }
    return take;
  }
  // the head, then requests of at most twice its footprint, cheapest first
  // (arrival order among equals)
  const size_t cap = 2 * pending[head]->bytes;
  std::vector<long> order(pending.size());
  for (long i = 0; i < (long)order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](long a, long b) { return pending[a]->bytes < pending[b]->bytes; });
  std::vector<char> taken(pending.size(), 0);
  used += pending[head]->bytes;
  take.push_back(pending[head]);
  taken[head] = 1;
  for (long i : order) {
    if ((int)take.size() >= opts.max_batch) break;
    if (taken[i] || pending[i]->bytes > cap || solo(i)) continue;
    if (used + pending[i]->bytes > opts.hbm_budget_bytes) break;
    used += pending[i]->bytes;
    take.push_back(pending[i]);
    taken[i] = 1;
  }
  std::deque<std::shared_ptr<Request>> rest;
  for (long i = 0; i < (long)pending.size(); ++i)
    if (!taken[i]) rest.push_back(pending[i]);
  pending.swap(rest);
  return take;
}

namespace {
// a fused batch resident on a lane
struct Cohort {
  std::unique_ptr<pgn_batch> B;
  std::vector<std::shared_ptr<Request>> req;
  std::vector<char> ended;
  size_t bytes = 0;
};
}  // namespace

void pgn_server::finish(const std::shared_ptr<Request>& r, int code, const std::string& err) {
  // caller holds mu
  r->status = code;
  r->error = err;
  r->t_done = Clock::now();
  r->done = true;
  r->inst = pgn_batch_inst{};  // free the staged inputs
  if (code == PGN_OK) ++solved;
}

void pgn_server::lane_main(int lane) {
  pgn_ctx* ctx = ctxs[lane];
  std::vector<Cohort> cohorts;
  auto fail_cohort = [&](Cohort& c, int code, const std::string& err) {
    std::lock_guard<std::mutex> lk(mu);
    for (size_t k = 0; k < c.req.size(); ++k)
      if (!c.ended[k]) {
        c.ended[k] = 1;
        finish(c.req[k], code, err);
      }
    used -= c.bytes;
    c.bytes = 0;
  };
  for (;;) {
    // ---- admission (between Newton iterations) ----
    std::vector<std::shared_ptr<Request>> take;
    {
      std::unique_lock<std::mutex> lk(mu);
      if (cohorts.empty())
        cv_work.wait_for(lk, std::chrono::milliseconds(50), [&] {
          return stop || pick_head() >= 0;
        });
      if (stop) {
        lk.unlock();
        for (auto& c : cohorts) fail_cohort(c, PGN_ERR_INVALID, "server destroyed during the solve");
        cv_done.notify_all();
        return;
      }
      if ((int)cohorts.size() < opts.cohorts_per_lane) {
        take = admit();
        if (!take.empty()) ++batches;
      }
    }
    if (!take.empty()) {
      Cohort c;
      c.req = take;
      c.ended.assign(take.size(), 0);
      for (auto& r : take) c.bytes += r->bytes;
      const auto t_start = Clock::now();
      try {
        PGN_HIP(hipSetDevice(device));
        c.B = std::make_unique<pgn_batch>();
        c.B->ctx = ctx;
        c.B->prec = opts.precision;
        for (auto& r : take) {
          r->t_start = t_start;
          c.B->inst.push_back(r->inst);
        }
        c.B->build();
        c.B->newton_begin(opts.newton);
        cohorts.push_back(std::move(c));
      } catch (const Error& e) {
        fail_cohort(c, e.code, e.what());
        cv_done.notify_all();
        cv_work.notify_all();
      } catch (const std::exception& e) {
        fail_cohort(c, PGN_ERR_INVALID, e.what());
        cv_done.notify_all();
        cv_work.notify_all();
      }
    }
    if (cohorts.empty()) continue;
    // ---- one Newton iteration of every resident cohort ----
    for (auto& c : cohorts) {
      std::vector<int> fin;
      std::vector<std::vector<double>> xs;
      std::vector<pgn_newton_result> rs;
      try {
        PGN_HIP(hipSetDevice(device));
        fin = c.B->newton_step();
        xs.resize(fin.size());
        rs.resize(fin.size());
        for (size_t q = 0; q < fin.size(); ++q) {
          const int k = fin[q];
          const long n0 = c.B->hnoff[k], nk = c.B->hnoff[k + 1] - n0;
          xs[q].resize(3 * nk);
          PGN_HIP(hipMemcpyAsync(xs[q].data(), c.B->M->x.p + 3 * n0, 3 * nk * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream));
          c.B->newton_result(k, rs[q]);
        }
        if (!fin.empty()) PGN_HIP(hipStreamSynchronize(ctx->stream));
      } catch (const Error& e) {
        fail_cohort(c, e.code, e.what());
        c.B.reset();
      } catch (const std::exception& e) {
        fail_cohort(c, PGN_ERR_INVALID, e.what());
        c.B.reset();
      }
      if (!fin.empty() && c.B) {
        std::lock_guard<std::mutex> lk(mu);
        for (size_t q = 0; q < fin.size(); ++q) {
          const int k = fin[q];
          auto& r = c.req[k];
          r->res = rs[q];
          r->x = std::move(xs[q]);
          c.ended[k] = 1;
          finish(r, PGN_OK, std::string());
        }
      }
      if (!fin.empty() || !c.B) cv_done.notify_all();
    }
    // ---- release the cohorts whose requests have all ended ----
    bool freed = false;
    for (auto it = cohorts.begin(); it != cohorts.end();) {
      const bool all = std::all_of(it->ended.begin(), it->ended.end(), [](char e) { return e; });
      if (all || !it->B) {
        it->B.reset();  // device memory back
        {
          std::lock_guard<std::mutex> lk(mu);
          used -= it->bytes;
        }
        it = cohorts.erase(it);
        freed = true;
      } else {
        ++it;
      }
    }
    if (freed) cv_work.notify_all();
  }
}

namespace {
template <class F>
int sguard(F&& f) {
  try {
    f();
    return PGN_OK;
  } catch (const Error& e) {
    return api_guard_set(e.what(), e.code);
  } catch (const std::bad_alloc&) {
    return api_guard_set("host allocation failed", PGN_ERR_NOMEM);
  } catch (const std::exception& e) {
    return api_guard_set(e.what(), PGN_ERR_INVALID);
  }
}

long enqueue(pgn_server* S, pgn_batch_inst&& inst) {
  auto r = std::make_shared<Request>();
  r->inst = std::move(inst);
  r->bytes = r->inst.footprint();
  PGN_CHECK(r->bytes <= S->opts.hbm_budget_bytes,
            "request needs more device memory than the server's budget");
  r->t_submit = Clock::now();
  long t;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    t = S->next_ticket++;
    r->ticket = t;
    S->all[t] = r;
    S->pending.push_back(r);
  }
  S->cv_work.notify_all();
  return t;
}

void set_bcs(pgn_batch_inst& I, const unsigned char* fixed, const double* gravity,
             const double* forces, int stages) {
  PGN_CHECK(stages >= 1, "stages must be >= 1");
  if (fixed) I.fixed.assign(fixed, fixed + 3L * I.n);
  if (gravity) {
    I.has_g = true;
    for (int i = 0; i < 3; ++i) {
      I.g_ini[i] = 0.0;
      I.g_end[i] = gravity[i];
    }
    I.g_stages = stages;
  }
  if (forces) {
    I.has_f = true;
    I.f_ini.assign(3L * I.n, 0.0);
    I.f_end.assign(forces, forces + 3L * I.n);
    I.f_stages = stages;
  }
}
}  // namespace

extern "C" {

void pgn_server_default_opts(pgn_server_opts* o) {
  if (!o) return;
  o->hbm_budget_bytes = (size_t)64 << 30;
  o->lanes = 4;
  o->max_batch = 64;
  o->precision = PGN_MIXED;
  pgn_newton_default_opts(&o->newton);
  o->group_by_size = 1;
  o->max_wait_s = 2.0;
  o->cohorts_per_lane = 4;
  o->solo_nodes = 32768;
}

int pgn_server_create(int device, const pgn_server_opts* opts, pgn_server** out) {
  return sguard([&] {
    PGN_CHECK(out, "null argument");
    *out = nullptr;
    auto S = std::make_unique<pgn_server>();
    S->device = device;
    pgn_server_default_opts(&S->opts);
    if (opts) S->opts = *opts;
    PGN_CHECK(S->opts.lanes >= 1 && S->opts.lanes <= 16, "lanes must be 1..16");
    PGN_CHECK(S->opts.max_batch >= 1, "max_batch must be >= 1");
    PGN_CHECK(S->opts.cohorts_per_lane >= 1 && S->opts.cohorts_per_lane <= 64,
              "cohorts_per_lane must be 1..64");
    PGN_CHECK(S->opts.precision == PGN_FP64 || S->opts.precision == PGN_MIXED,
              "unknown precision");
    for (int l = 0; l < S->opts.lanes; ++l) {
      pgn_ctx* c = nullptr;
      const int rc = pgn_ctx_create(device, &c);
      if (rc != PGN_OK) {
        for (pgn_ctx* p : S->ctxs) pgn_ctx_destroy(p);
        throw Error(rc, pgn_last_error());
      }
      S->ctxs.push_back(c);
    }
    pgn_server* raw = S.get();
    for (int l = 0; l < S->opts.lanes; ++l)
      S->lanes.emplace_back([raw, l] { raw->lane_main(l); });
    *out = S.release();
  });
}

void pgn_server_destroy(pgn_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  for (auto& t : S->lanes) t.join();  // running batches finish first
  {
    // requests never started end with an error; waiters wake, read it and
    // leave before the server (its mutex and condition variables) goes away
    std::unique_lock<std::mutex> lk(S->mu);
    for (auto& r : S->pending) {
      r->status = PGN_ERR_INVALID;
      r->error = "server destroyed before the request started";
      r->done = true;
      r->inst = pgn_batch_inst{};
    }
    S->pending.clear();
    S->cv_done.notify_all();
    S->cv_done.wait(lk, [&] { return S->waiters == 0; });
  }
  for (pgn_ctx* c : S->ctxs) pgn_ctx_destroy(c);
  delete S;
}

int pgn_server_submit_tet(pgn_server* S, int n_nodes, const double* X0, int n_tets,
                          const int* T, const unsigned char* material, const double* lame1,
                          const double* lame2, const double* density,
                          const unsigned char* fixed, const double* gravity,
                          const double* forces, int stages, long* ticket) {
  return sguard([&] {
    PGN_CHECK(S && ticket && X0 && T && material && lame1 && lame2 && density, "null argument");
    PGN_CHECK(n_nodes > 0 && n_tets > 0, "empty request");
    for (long q = 0; q < 4L * n_tets; ++q)
      PGN_CHECK(T[q] >= 0 && T[q] < n_nodes, "node id out of range");
    pgn_batch_inst I;
    I.n = n_nodes;
    I.X0.assign(X0, X0 + 3L * n_nodes);
    I.T[0].assign(T, T + 4L * n_tets);
    for (int e = 0; e < n_tets; ++e) PGN_CHECK(material[e] <= 1, "unknown material kind");
    I.kind[0].assign(material, material + n_tets);
    I.l1[0].assign(lame1, lame1 + n_tets);
    I.l2[0].assign(lame2, lame2 + n_tets);
    I.rho[0].assign(density, density + n_tets);
    I.extra[0].assign(n_tets, 0.0);
    I.fixed.assign(3L * n_nodes, 0);
    set_bcs(I, fixed, gravity, forces, stages);
    *ticket = enqueue(S, std::move(I));
  });
}

int pgn_server_submit_shell(pgn_server* S, int n_nodes, const double* X0, int n_tris,
                            const int* F, const double* lame1, const double* lame2,
                            const double* thickness, const double* density, int n_hinges,
                            const int* hinges, const double* bending_k,
                            const unsigned char* fixed, const double* gravity,
                            const double* forces, int stages, long* ticket) {
  return sguard([&] {
    PGN_CHECK(S && ticket && X0, "null argument");
    PGN_CHECK(n_nodes > 0 && n_tris >= 0 && n_hinges >= 0 && n_tris + n_hinges > 0,
              "empty request");
    PGN_CHECK((n_tris == 0 || (F && lame1 && lame2 && thickness && density)) &&
                  (n_hinges == 0 || (hinges && bending_k)),
              "null input array");
    for (long q = 0; q < 3L * n_tris; ++q)
      PGN_CHECK(F[q] >= 0 && F[q] < n_nodes, "node id out of range");
    for (long q = 0; q < 4L * n_hinges; ++q)
      PGN_CHECK(hinges[q] >= 0 && hinges[q] < n_nodes, "node id out of range");
    pgn_batch_inst I;
    I.n = n_nodes;
    I.X0.assign(X0, X0 + 3L * n_nodes);
    for (long t = 0; t < n_tris; ++t) {
      for (int a = 0; a < 3; ++a) I.T[1].push_back(F[3 * t + a]);
      I.T[1].push_back(F[3 * t + 2]);
      I.kind[1].push_back(2);
      I.l1[1].push_back(lame1[t]);
      I.l2[1].push_back(lame2[t]);
      I.rho[1].push_back(density[t]);
      I.extra[1].push_back(thickness[t]);
    }
    for (long h = 0; h < n_hinges; ++h) {
      for (int a = 0; a < 4; ++a) I.T[2].push_back(hinges[4 * h + a]);
      I.kind[2].push_back(3);
      I.l1[2].push_back(0.0);
      I.l2[2].push_back(0.0);
      I.rho[2].push_back(0.0);
      I.extra[2].push_back(bending_k[h]);
    }
    I.fixed.assign(3L * n_nodes, 0);
    set_bcs(I, fixed, gravity, forces, stages);
    *ticket = enqueue(S, std::move(I));
  });
}

int pgn_server_poll(pgn_server* S, long ticket, int* done) {
  return sguard([&] {
    PGN_CHECK(S && done, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    PGN_CHECK(it != S->all.end(), "unknown ticket");
    *done = it->second->done ? 1 : 0;
  });
}

int pgn_server_wait(pgn_server* S, long ticket, double* x, pgn_newton_result* res,
                    double* latency_s, double* queue_s) {
  return sguard([&] {
    PGN_CHECK(S, "null argument");
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      auto it = S->all.find(ticket);
      PGN_CHECK(it != S->all.end(), "unknown ticket");
      r = it->second;
      ++S->waiters;
      S->cv_done.wait(lk, [&] { return r->done; });
      --S->waiters;
      // another wait on the same ticket may have removed it meanwhile
      auto again = S->all.find(ticket);
      if (again != S->all.end() && again->second == r) S->all.erase(again);
      if (S->waiters == 0) S->cv_done.notify_all();
    }
    if (r->status != PGN_OK) throw Error(r->status, "request failed: " + r->error);
    if (x) std::copy(r->x.begin(), r->x.end(), x);
    if (res) *res = r->res;
    if (latency_s) *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    if (queue_s) *queue_s = std::chrono::duration<double>(r->t_start - r->t_submit).count();
  });
}

int pgn_server_stats(pgn_server* S, long* batches, long* solved, size_t* bytes_in_use) {
  return sguard([&] {
    PGN_CHECK(S, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    if (batches) *batches = S->batches;
    if (solved) *solved = S->solved;
    if (bytes_in_use) *bytes_in_use = S->used;
  });
}

}  // extern "C"
