// Serving loop behind the C ABI: a queue of independent Davidson solves on
// one GPU with continuous batching.
//
// `slots` solver threads each own a library context (their own HIP stream),
// so solves of different sizes run concurrently on the device.  A free slot
// takes the oldest queued request, builds its sigma plan under the
// process-wide plan budget (fcg_set_plan_budget: the plan reserves its
// device bytes before allocating anything) and, when the plan does not fit
// beside the resident ones, waits until a running solve releases its plan;
// then it runs fcg_davidson and frees the plan.  Results stay with the
// ticket until the caller polls them.  Each solve is deterministic
// (fixed-order reductions), so a request's energies are the same alone and
// in the mix.  No counterpart in the reference (a Julia host would call
// tps_ci_davidson per request).
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "fcg_internal.h"

namespace {

struct ServeRequest {
  int64_t ticket = 0;
  fcg_ham_t ham = nullptr;
  fcg_ops_t ops = nullptr;
  fcg_space_t sp = nullptr;
  int R = 0;
  int64_t n = 0;
  std::vector<double> v0, evals, evecs;
  fcg_davidson_stats stats{};
  fcg_serve_times times{};
  int rc = FCG_OK;
  bool done = false;
};

}  // namespace

struct fcg_server {
  fcg_serve_opts opts{};
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable queued;    // a request was queued, or stop
  std::condition_variable finished;  // a request finished (pollers)
  std::condition_variable released;  // a plan was freed (admission waits)
  std::deque<std::shared_ptr<ServeRequest>> queue;
  std::map<int64_t, std::shared_ptr<ServeRequest>> tickets;
  std::map<fcg_space_t, std::unique_ptr<std::mutex>> space_mu;  // plan builds per space
  std::vector<std::thread> threads;
  int64_t next_ticket = 1;
  int64_t prev_budget = INT64_MAX;
  bool stopping = false;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  std::mutex &lock_for(fcg_space_t sp) {  // caller holds mu
    auto &m = space_mu[sp];
    if (!m) m.reset(new std::mutex);
    return *m;
  }

  void slot(fcg_ctx_t ctx) {
    const int prec = opts.davidson.precision;
    const int precisions = (1 << FCG_FP64) | (1 << prec);
    for (;;) {
      std::shared_ptr<ServeRequest> rq;
      {
        std::unique_lock<std::mutex> lk(mu);
        queued.wait(lk, [&] { return stopping || !queue.empty(); });
        if (queue.empty()) return;  // stopping and drained
        rq = queue.front();
        queue.pop_front();
      }
      rq->times.started = now();
      // admission by device footprint: retry while the budget is held by others
      fcg_plan_t plan = nullptr;
      int rc;
      for (;;) {
        int64_t need = 0;
        std::mutex *sm;
        {
          std::lock_guard<std::mutex> lk(mu);
          sm = &lock_for(rq->sp);
        }
        {
          std::lock_guard<std::mutex> bl(*sm);
          rc = fcg_plan_create_limited(ctx, rq->ham, rq->ops, rq->sp, opts.nbody, precisions,
                                       INT64_MAX, &need, &plan);
        }
        rq->times.plan_bytes = need;
        if (rc != FCG_ENOMEM || need <= 0) break;
        int64_t budget = 0, held = 0;
        fcg_plan_budget(&budget, &held);
        // can never fit, or the device itself is out of memory with no plan
        // resident that waiting could free: fail the request
        if (need > budget || held == 0) break;
        std::unique_lock<std::mutex> lk(mu);
        released.wait_for(lk, std::chrono::milliseconds(200));
      }
      rq->times.planned = now();
      if (rc == FCG_OK) {
        fcg_davidson_opts o = opts.davidson;
        o.precond_diag = nullptr;
        rq->evals.assign(rq->R, 0.0);
        rq->evecs.assign((size_t)rq->n * rq->R, 0.0);
        rc = fcg_davidson(ctx, plan, rq->R, rq->v0.data(), &o, rq->evals.data(),
                          rq->evecs.data(), &rq->stats);
        fcg_plan_destroy(plan);
      }
      std::vector<double>().swap(rq->v0);
      {
        std::lock_guard<std::mutex> lk(mu);
        rq->rc = rc;
        rq->times.finished = now();
        rq->done = true;
      }
      released.notify_all();
      finished.notify_all();
    }
  }
};

extern "C" int fcg_serve_start(const fcg_serve_opts *opts, fcg_server_t *out) {
  if (!opts || !out || opts->slots < 1 || opts->slots > 64 || opts->nbody < 1 ||
      opts->nbody > 4 || opts->hbm_budget <= 0 || opts->davidson.precision < FCG_FP64 ||
      opts->davidson.precision >= FCG_N_PREC)
    return FCG_EINVAL;
  std::unique_ptr<fcg_server> s(new fcg_server);
  s->opts = *opts;
  s->t0 = std::chrono::steady_clock::now();
  std::vector<fcg_ctx_t> ctxs;
  for (int k = 0; k < opts->slots; k++) {
    fcg_ctx_t c = nullptr;
    if (int rc = fcg_create(opts->device, &c)) {
      for (fcg_ctx_t x : ctxs) fcg_destroy(x);
      return rc;
    }
    ctxs.push_back(c);
  }
  int64_t held = 0;
  fcg_plan_budget(&s->prev_budget, &held);
  fcg_set_plan_budget(opts->hbm_budget);
  for (fcg_ctx_t c : ctxs)
    s->threads.emplace_back([srv = s.get(), c] {
      srv->slot(c);
      fcg_destroy(c);
    });
  *out = s.release();
  return FCG_OK;
}

extern "C" int fcg_serve_submit(fcg_server_t s, fcg_ham_t ham, fcg_ops_t ops, fcg_space_t sp,
                                int R, const double *v0, int64_t *ticket) {
  if (!s || !ham || !ops || !sp || !v0 || !ticket || R < 1 || R > FCG_MAX_R) return FCG_EINVAL;
  auto rq = std::make_shared<ServeRequest>();
  rq->ham = ham;
  rq->ops = ops;
  rq->sp = sp;
  rq->R = R;
  rq->n = fcg_space_nvar(sp);
  if (rq->n < R) return FCG_EINVAL;
  rq->v0.assign(v0, v0 + rq->n * R);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return FCG_EINVAL;
    rq->ticket = s->next_ticket++;
    rq->times.submitted = s->now();
    s->tickets[rq->ticket] = rq;
    s->queue.push_back(rq);
    *ticket = rq->ticket;
  }
  s->queued.notify_one();
  return FCG_OK;
}

extern "C" int fcg_serve_poll(fcg_server_t s, int64_t ticket, int wait, double *evals,
                              double *evecs, fcg_davidson_stats *stats, fcg_serve_times *times) {
  if (!s) return FCG_EINVAL;
  std::shared_ptr<ServeRequest> rq;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return FCG_EINVAL;
    rq = it->second;
    if (wait) s->finished.wait(lk, [&] { return rq->done; });
    if (!rq->done) return FCG_PENDING;
    s->tickets.erase(it);
  }
  if (evals && !rq->evals.empty()) std::memcpy(evals, rq->evals.data(), sizeof(double) * rq->R);
  if (evecs && !rq->evecs.empty())
    std::memcpy(evecs, rq->evecs.data(), sizeof(double) * rq->evecs.size());
  if (stats) *stats = rq->stats;
  if (times) *times = rq->times;
  return rq->rc;
}

extern "C" int fcg_serve_stop(fcg_server_t s) {
  if (!s) return FCG_EINVAL;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->queued.notify_all();
  for (auto &t : s->threads) t.join();
  fcg_set_plan_budget(s->prev_budget);
  delete s;
  return FCG_OK;
}
