// batch.hip -- many independent solves on one device: a batch call and a serving loop.
//
// Every solve owns its operator and Davidson blocks and runs on its own HIP stream from a
// worker thread, so solves of uneven size overlap on the device; all kernels are
// deterministic (no atomics in H.X, Gram reductions in fixed order), so a solve gives the
// same bits alone and in a mix.  The server admits queued requests in arrival order as
// soon as a worker is free and the request's exact device footprint (operator from the
// builder's count pass + Davidson blocks) fits the remaining HBM budget.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "op.h"

namespace cosmo {
int stored_footprint(const System &s, int flags, int64_t *bytes);
void release_thread_streams();

int64_t davidson_footprint(int64_t n, int nume, int mblock, int lim) {
  int ldmax = 4;
  while (ldmax < mblock) ldmax <<= 1;
  const int64_t G = std::min<int64_t>(1024, std::max<int64_t>(1, (n + 31) / 32));
  return 8 * n * (2 * (int64_t)lim + std::max(nume, mblock) + std::max(lim, 16)) + n * ldmax * (8 + 8 + 4 + 4) +
         G * 1024 * 8 + 3 * 8192 * 8;
}

static double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// one solve: build, Davidson, destroy (on the calling thread's stream)
static int solve_one(System *s, int flags, const cosmo_dav_params &p0, int nume, double tol, double *evals,
                     double *resid, cosmo_dav_stats *st) {
  cosmo_op op = nullptr;
  int rc = cosmo_op_create(static_cast<cosmo_system>(s), flags, &op);
  if (rc) return rc;
  cosmo_dav_params p = p0;
  const int64_t n = s->dim;
  p.nume = std::min<int64_t>(nume, n);
  if (p.mblock <= 0 || p.mblock > p.nume) p.mblock = p.nume;
  // a wider basis than the single-solve default keeps small, nearly degenerate spectra
  // from stalling; iteration cap raised accordingly
  int lim = std::max(p.lim, p.nume + std::max(20, p.nume));
  if (lim > n) lim = (int)n;
  p.lim = lim;
  p.maxiter = std::max(p.maxiter, 4000);
  if (tol > 0) p.critr = tol;
  std::vector<double> de(p.nume);
  rc = cosmo_davidson(op, &p, nullptr, evals, de.data(), resid, nullptr, st);
  cosmo_op_destroy(op);
  return rc;
}

}  // namespace cosmo

using namespace cosmo;

struct cosmo_batch_s {
  std::vector<cosmo_system> sys;
  int flags = 0;
  int device = 0;
};

struct Request {
  int64_t id;
  cosmo_system sys;
  int nume;
  double tol;
  double t_submit, t_admit = 0, t_done = 0;
  int64_t bytes = 0;
};

struct cosmo_server_s {
  cosmo_server_params p{};
  int device = 0;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv, cv_done;
  std::deque<Request> queue;
  std::vector<cosmo_result> done;
  std::vector<std::thread> workers;
  int64_t used = 0, next_id = 1, inflight = 0;
  bool stop = false;
};

static void server_worker(cosmo_server_s *S) {
  cosmo_set_device(S->device);
  for (;;) {
    Request r;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      for (;;) {
        if (S->stop) {
          lk.unlock();
          release_thread_streams();
          return;
        }
        // admission: first queued request whose footprint fits (arrival order)
        bool got = false;
        for (auto it = S->queue.begin(); it != S->queue.end(); ++it)
          if (S->used + it->bytes <= S->p.hbm_budget || S->used == 0) {
            r = *it;
            S->queue.erase(it);
            got = true;
            break;
          }
        if (got) break;
        S->cv.wait(lk);
      }
      S->used += r.bytes;
      ++S->inflight;
      r.t_admit = now_s(S->t0);
    }
    cosmo_result res;
    memset(&res, 0, sizeof(res));
    res.id = r.id;
    res.n = cosmo_system_dim(r.sys);
    res.nume = std::min<int64_t>(std::min(r.nume, 64), res.n);
    cosmo_dav_params p;
    cosmo_dav_defaults(&p, std::max(res.nume, 1));
    p.precision = S->p.precision;
    p.lim = res.nume + std::max(20, res.nume);
    p.maxiter = 4000;
    cosmo_dav_stats st;
    int rc = res.n > 0 ? solve_one(r.sys, 0, p, res.nume, r.tol, res.evals, res.resid, &st) : COSMO_EINVAL;
    res.status = rc;
    res.t_submit = r.t_submit;
    res.t_admit = r.t_admit;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      res.t_done = now_s(S->t0);
      S->used -= r.bytes;
      --S->inflight;
      S->done.push_back(res);
    }
    S->cv.notify_all();
    S->cv_done.notify_all();
  }
}

extern "C" {

int cosmo_batch_create(int nsys, const cosmo_system *systems, int flags, cosmo_batch *out) {
  if (nsys <= 0 || !systems || !out) return fail(COSMO_EINVAL, "cosmo_batch_create: bad args");
  auto *b = new cosmo_batch_s();
  b->sys.assign(systems, systems + nsys);
  b->flags = flags;
  b->device = current_device();
  *out = b;
  return COSMO_OK;
}

int cosmo_batch_solve(cosmo_batch b, const cosmo_dav_params *p, const int *nume_per_solve, double *evals,
                      double *resid, cosmo_dav_stats *stats) {
  if (!b || !p || !nume_per_solve || !evals) return fail(COSMO_EINVAL, "cosmo_batch_solve: bad args");
  const int ns = (int)b->sys.size();
  std::vector<int64_t> off(ns + 1, 0);
  for (int i = 0; i < ns; ++i) off[i + 1] = off[i] + nume_per_solve[i];
  std::vector<int> rcs(ns, 0);
  std::vector<cosmo_dav_stats> sts(ns);
  std::atomic<int> next(0);
  const int nthreads = std::min(ns, 16);
  auto t0 = std::chrono::steady_clock::now();
  auto work = [&]() {
    cosmo_set_device(b->device);
    for (;;) {
      int i = next.fetch_add(1);
      if (i >= ns) break;
      const int64_t n = cosmo_system_dim(b->sys[i]);
      const int k = (int)std::min<int64_t>(nume_per_solve[i], n);
      for (int j = k; j < nume_per_solve[i]; ++j) {  // fewer states than asked: pad
        evals[off[i] + j] = 0.0;
        if (resid) resid[off[i] + j] = 0.0;
      }
      std::vector<double> r(std::max(k, 1));
      rcs[i] = k > 0 ? solve_one(b->sys[i], b->flags, *p, k, 0.0, evals + off[i], r.data(), &sts[i]) : COSMO_OK;
      if (resid)
        for (int j = 0; j < k; ++j) resid[off[i] + j] = r[j];
    }
    release_thread_streams();
  };
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) th.emplace_back(work);
  for (auto &t : th) t.join();
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    for (auto &s : sts) {
      stats->iterations = std::max(stats->iterations, s.iterations);
      stats->matvecs += s.matvecs;
      stats->matvecs_fp64 += s.matvecs_fp64;
      stats->restarts += s.restarts;
      stats->ierr |= s.ierr;
      stats->ms_apply += s.ms_apply;
      stats->ms_ortho += s.ms_ortho;
      stats->ms_rr += s.ms_rr;
    }
    stats->ms_total = now_s(t0) * 1e3;
  }
  for (int i = 0; i < ns; ++i)
    if (rcs[i] && rcs[i] != COSMO_ECONV) return rcs[i];
  for (int i = 0; i < ns; ++i)
    if (rcs[i]) return fail(COSMO_ECONV, "a batched solve reached maxiter");
  return COSMO_OK;
}

int cosmo_batch_destroy(cosmo_batch b) {
  delete b;
  return COSMO_OK;
}

int cosmo_server_create(const cosmo_server_params *p, cosmo_server *out) {
  if (!p || !out || p->max_active <= 0 || p->max_active > 16 || p->hbm_budget <= 0)
    return fail(COSMO_EINVAL, "cosmo_server_create: need 1..16 active solves and a positive HBM budget");
  auto *S = new cosmo_server_s();
  S->p = *p;
  S->device = current_device();
  S->t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < p->max_active; ++i) S->workers.emplace_back(server_worker, S);
  *out = S;
  return COSMO_OK;
}

int cosmo_server_submit(cosmo_server S, cosmo_system sys, int nume, double tol, int64_t *id) {
  if (!S || !sys || nume <= 0) return fail(COSMO_EINVAL, "cosmo_server_submit: bad args");
  Request r;
  r.sys = sys;
  r.nume = std::min(nume, 64);
  r.tol = tol;
  int64_t ob = 0;
  int rc = stored_footprint(*sys, 0, &ob);
  if (rc) return rc;
  const int nm = (int)std::min<int64_t>(r.nume, std::max<int64_t>(sys->dim, 1));
  r.bytes = ob + davidson_footprint(sys->dim, nm, nm, nm + std::max(20, nm));
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (r.bytes > S->p.hbm_budget) return fail(COSMO_ENOMEM, "request needs more HBM than the server budget");
    r.id = S->next_id++;
    r.t_submit = now_s(S->t0);
    S->queue.push_back(r);
  }
  if (id) *id = r.id;
  S->cv.notify_all();
  return COSMO_OK;
}

int cosmo_server_poll(cosmo_server S, cosmo_result *results, int cap) {
  if (!S || (cap > 0 && !results)) return fail(COSMO_EINVAL, "cosmo_server_poll: bad args");
  std::lock_guard<std::mutex> lk(S->mu);
  int k = 0;
  while (k < cap && !S->done.empty()) {
    results[k++] = S->done.front();
    S->done.erase(S->done.begin());
  }
  return k;
}

int cosmo_server_drain(cosmo_server S, cosmo_result *results, int cap, int *written) {
  if (!S) return fail(COSMO_EINVAL, "cosmo_server_drain: null server");
  {
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] { return S->queue.empty() && S->inflight == 0; });
  }
  int k = cosmo_server_poll(S, results, cap);
  if (written) *written = k;
  return k < 0 ? k : COSMO_OK;
}

int cosmo_server_destroy(cosmo_server S) {
  if (!S) return COSMO_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    S->queue.clear();
  }
  S->cv.notify_all();
  for (auto &t : S->workers) t.join();
  delete S;
  return COSMO_OK;
}

int cosmo_op_footprint(cosmo_system s, int flags, int64_t *bytes) {
  if (!s || !bytes) return fail(COSMO_EINVAL, "cosmo_op_footprint: null argument");
  if (flags & COSMO_OP_VIRTUAL) {
    *bytes = 8 * s->dim + (int64_t)s->orb.size() * s->orb.size() * s->orb.size() * s->orb.size();
    return COSMO_OK;
  }
  return stored_footprint(*s, flags, bytes);
}

}  // extern "C"
