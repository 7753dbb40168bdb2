// batch.hip -- many independent solves on one device: a batch call and a serving loop.
//
// A batch splits its solves by size.  Those up to the single-launch bound (and within the
// dav_wg.h bounds) run to completion in ONE persistent kernel launch once worker threads have
// built their operators, each solve by a team of workgroups sized to its dimension (one
// workgroup per 2048 rows); the others own their operator and Davidson blocks and run on
// their own HIP streams from the worker threads at the same time (their kernels take CUs as
// the launch's workgroups retire: each of those fills its CU's registers).  All kernels are
// deterministic (no atomics in H.X, reductions in fixed order, team sizes fixed by the
// system), so a solve gives the same bits alone and in a mix on either path.  The server
// admits queued requests in arrival order as
// soon as a worker is free and the request's exact device footprint (operator from the
// builder's count pass + Davidson blocks) fits the remaining HBM budget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "op.h"
#include "dav_wg.h"

namespace cosmo {
int stored_footprint(const System &s, int flags, int64_t *bytes);
void release_thread_streams();

// device bytes of cosmo_davidson's blocks (dav_alloc in davidson.hip): B, AB, R, T in FP64
// and FP32, the H.X staging, the Gram partial sums and the small buffers
int64_t davidson_footprint(int64_t n, int nume, int mblock, int lim) {
  int ldmax = 4;
  while (ldmax < mblock) ldmax <<= 1;
  const int64_t G = std::min<int64_t>(1024, std::max<int64_t>(1, (n + 127) / 128));
  const int64_t cols = 2 * (int64_t)lim + std::max(nume, mblock) + std::max(lim, 16);
  return 12 * n * cols + n * ldmax * (8 + 8 + 4 + 4) + (G + 32) * 2048 * 8 + 8192 * 8 + 256 * 8;
}

static double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Basis limit of a batched or served solve: at least nume + 40 up to 45 (the single-launch
// kernel holds 48).  On the 256-solve sd batch (5 pairs each, FP32 H.V + FP64 refinement)
// lim 25 / 30 / 36 / 45 / 60 took 2.03 / 1.59 / 1.37 / 1.31 / 1.36 s: the wider basis
// restarts less often on the slowly converging, nearly degenerate spectra.
static int batch_lim(int nume) { return std::max(nume + std::max(20, nume), std::min(nume + 40, 45)); }

// Davidson parameters of one batched solve of dimension n: the wider basis above, iteration
// cap raised accordingly
static cosmo_dav_params batch_params(const cosmo_dav_params &p0, int nume, int64_t n, double tol) {
  cosmo_dav_params p = p0;
  p.nume = (int)std::min<int64_t>(nume, n);
  if (p.mblock <= 0 || p.mblock > p.nume) p.mblock = p.nume;
  int lim = std::max(p.lim, batch_lim(p.nume));
  if (lim > n) lim = (int)n;
  p.lim = lim;
  p.maxiter = std::max(p.maxiter, 4000);
  if (tol > 0) p.critr = tol;
  return p;
}

// one solve: build, Davidson, destroy (on the calling thread's stream)
static int solve_one(System *s, int flags, const cosmo_dav_params &p0, int nume, double tol, double *evals,
                     double *resid, cosmo_dav_stats *st) {
  cosmo_op op = nullptr;
  int rc = cosmo_op_create(static_cast<cosmo_system>(s), flags, &op);
  if (rc) return rc;
  const cosmo_dav_params p = batch_params(p0, nume, s->dim, tol);
  std::vector<double> de(p.nume);
  rc = cosmo_davidson(op, &p, nullptr, evals, de.data(), resid, nullptr, st);
  cosmo_op_destroy(op);
  return rc;
}


// ---- single-launch path ---------------------------------------------------------------------
// Solves inside the bounds of dav_wg.h, on stored operators already built on the device: all
// of them in one persistent launch on the calling thread's stream, each by a team of
// workgroups sized to its dimension.
struct SmallJob {
  cosmo_op op;
  cosmo_dav_params p;  // as batch_params gives it
  double *evals, *resid;
  cosmo_dav_stats *st;
  int rc;
};

static bool fits_single_launch(const cosmo_dav_params &p, int64_t n, int flags, int64_t nmax) {
  return n <= nmax && n <= DWG_NMAX && p.nume >= 1 && p.nume <= DWG_MB && p.mblock <= DWG_MB && p.lim <= DWG_LIM &&
         p.nume <= p.lim && !(flags & (COSMO_OP_VIRTUAL | COSMO_OP_FACTORED)) &&
         !((flags & COSMO_OP_NO_FP32) && p.precision == COSMO_PREC_FP32) && p.ninit == 0;
}

// Rows per workgroup of a team (COSMO_DWG_ROWS overrides): a team of G workgroups pays a
// barrier per exchange, about ten per iteration, against rows / G of streaming each.
static int64_t dwg_rows() {
  static const int64_t v = [] {
    const char *e = getenv("COSMO_DWG_ROWS");
    const long long x = e ? atoll(e) : 0;
    return (int64_t)(x > 0 ? x : 2048);
  }();
  return v;
}

// Team size and lanes per row of H.X of a solve: functions of the system (and of a
// device too small for the team), so a solve gives the same bits alone and in any mix.
static void team_shape(int64_t n, int64_t nnz, int cap, int *G, int *lanes) {
  const int64_t rows = dwg_rows();
  int64_t g = (n + rows - 1) / rows;
  g = std::max<int64_t>(1, std::min<int64_t>(g, std::min(DWG_GMAX, cap)));
  *G = (int)g;
  const double avg = n > 0 ? (double)nnz / (double)n : 0.0;
  int L = 1;
  while (L < 64 && L * 16 <= avg) L *= 2;  // about 8 to 16 entries per lane
  *lanes = L;
}

static int run_small(std::vector<SmallJob> &jobs, const cosmo_dav_params &p0) {
  const int ns = (int)jobs.size();
  if (!ns) return COSMO_OK;
  hipStream_t st = lib_stream();
  int dev = 0, ncu = 0, per_cu = 0, wclk_khz = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCHK(hipDeviceGetAttribute(&wclk_khz, hipDeviceAttributeWallClockRate, dev));
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_dav_wg, DWG_THREADS, 0));
  const int cap = std::max(1, per_cu * ncu);  // workgroups resident at once
  // team shapes; seats in order of decreasing team size, then dimension
  std::vector<int> G(ns), L(ns), ord(ns);
  for (int j = 0; j < ns; ++j) {
    team_shape(jobs[j].op->n, jobs[j].op->csr.nnz, cap, &G[j], &L[j]);
    ord[j] = j;
  }
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
    if (G[a] != G[b]) return G[a] > G[b];
    return jobs[a].op->n > jobs[b].op->n;
  });
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  std::vector<size_t> off(ns + 1, 0);
  for (int j = 0; j < ns; ++j) {
    const size_t n = (size_t)jobs[j].op->n;
    off[j + 1] = off[j] + up(sizeof(double) * (n * (3 * (size_t)jobs[j].p.lim + 2 * DWG_MB) +
                                               2 * (size_t)G[j] * DWG_PART));
  }
  const size_t obytes = sizeof(double) * 3 * DWG_MB * (size_t)ns, ibytes = sizeof(int) * 6 * (size_t)ns;
  const size_t cbytes = sizeof(unsigned) * ((size_t)ns + 1);
  char *arena = nullptr, *outs = nullptr;
  DwgSolve *ddesc = nullptr;
  HIPCHK(dmalloc(&arena, off[ns], st));
  HIPCHK(dmalloc(&outs, up(obytes) + up(ibytes) + cbytes, st));
  HIPCHK(dmalloc(&ddesc, sizeof(DwgSolve) * ns, st));
  std::vector<DwgSolve> desc(ns);
  double *dout = (double *)outs;
  int *dinfo = (int *)(outs + up(obytes));
  unsigned *dbar = (unsigned *)(outs + up(obytes) + up(ibytes));  // ns team counters, then the seat counter
  int nseats = 0;
  for (int q = 0; q < ns; ++q) {
    const int j = ord[q];
    const Op &o = *jobs[j].op;
    const cosmo_dav_params &p = jobs[j].p;
    DwgSolve &d = desc[q];
    const int64_t n = o.n;
    double *w = (double *)(arena + off[j]);
    d.n = n;
    d.rowptr = o.csr.rowptr;
    d.cols = o.csr.cols;
    d.v64 = o.csr.v64;
    d.v32 = o.csr.v32;
    d.diag = o.csr.diag;
    d.B = w;
    d.AB = w + n * p.lim;
    d.R = w + 2 * n * p.lim;
    d.X = w + 3 * n * p.lim;
    d.part = d.X + 2 * n * DWG_MB;
    d.bar = dbar + q;
    d.claim = dbar + ns;
    d.G = G[j];
    d.lanes = L[j];
    d.seat0 = nseats;
    nseats += G[j];
    d.nume = p.nume;
    d.lim = p.lim;
    d.mblock = p.mblock;
    d.maxiter = p.maxiter > 0 ? p.maxiter : std::max(p.nume * 40, 200);
    d.critr = p.critr;
    d.evals = dout + (size_t)j * 3 * DWG_MB;
    d.dE = d.evals + DWG_MB;
    d.resid = d.evals + 2 * DWG_MB;
    d.info = dinfo + 6 * j;
  }
  HIPCHK(hipMemcpyAsync(ddesc, desc.data(), sizeof(DwgSolve) * ns, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(dbar, 0, cbytes, st));
  DwgParams prm;
  prm.crite = p0.crite;
  prm.critc = p0.critc;
  prm.ortho = p0.ortho;
  prm.switch_tol = p0.mixed_switch > 0 ? p0.mixed_switch : 1e-5;
  prm.strict = p0.strict;
  prm.mixed = p0.precision == COSMO_PREC_FP32;
  prm.spin_ticks = (uint64_t)std::max(wclk_khz, 1) * 1000ull * 10ull;  // 10 s
  const int grid = std::min(nseats, cap);
  auto t0 = std::chrono::steady_clock::now();
  hipLaunchKernelGGL(k_dav_wg, dim3(grid), dim3(DWG_THREADS), 0, st, (const DwgSolve *)ddesc, ns, nseats,
                     dbar + ns, prm);
  HIPCHK(hipGetLastError());
  std::vector<double> hout(3 * DWG_MB * (size_t)ns);
  std::vector<int> hinfo(6 * (size_t)ns);
  HIPCHK(hipMemcpyAsync(hout.data(), dout, obytes, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(hinfo.data(), dinfo, ibytes, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  dfree(arena);
  dfree(outs);
  dfree(ddesc);
  for (int j = 0; j < ns; ++j) {
    SmallJob &jb = jobs[j];
    const double *o = hout.data() + (size_t)j * 3 * DWG_MB;
    const int *inf = hinfo.data() + 6 * j;
    for (int i = 0; i < jb.p.nume; ++i) {
      jb.evals[i] = o[i];
      jb.resid[i] = o[2 * DWG_MB + i];
    }
    memset(jb.st, 0, sizeof(*jb.st));
    jb.st->iterations = inf[0];
    jb.st->matvecs = inf[1];
    jb.st->matvecs_fp64 = inf[2];
    jb.st->restarts = inf[3];
    jb.st->ierr = inf[4];
    jb.st->ms_total = ms;
    if (inf[4] & DWG_ABORTED)
      jb.rc = fail(COSMO_EHIP, "cosmo_batch_solve: a team barrier of the single launch timed out");
    else jb.rc = inf[5] ? COSMO_OK : COSMO_ECONV;
  }
  return COSMO_OK;
}

}  // namespace cosmo

using namespace cosmo;

struct cosmo_batch_s {
  std::vector<cosmo_system> sys;
  int flags = 0;
  int device = 0;
  // dimension bound of the single-launch path (0: off).  On the 256-solve sd batch (n up to
  // 93,710) bounds of 2048 / 4096 / 8192 / 16384 / 32768 / all took 1.28 / 1.23 / 1.24 /
  // 1.13 / 1.17 / 1.29 s: the largest solves run faster host-driven beside the launch.
  int64_t single_nmax = 16384;
};

struct Request {
  int64_t id;
  cosmo_system sys;
  int nume;
  double tol;
  double t_submit, t_admit = 0, t_done = 0;
  int64_t bytes = 0;
  int flags = 0;  // operator kind: stored, or the product form when H does not fit
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
    int rc = res.n > 0 ? solve_one(r.sys, r.flags, p, res.nume, r.tol, res.evals, res.resid, &st) : COSMO_EINVAL;
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
  std::vector<std::vector<double>> rbuf(ns);
  // the solves that fit the single-launch path, and the rest (worker threads, own streams)
  std::vector<SmallJob> small;
  std::vector<int> small_idx, large;
  for (int i = 0; i < ns; ++i) {
    const int64_t n = cosmo_system_dim(b->sys[i]);
    const int k = (int)std::min<int64_t>(nume_per_solve[i], n);
    for (int j = k; j < nume_per_solve[i]; ++j) {  // fewer states than asked: pad
      evals[off[i] + j] = 0.0;
      if (resid) resid[off[i] + j] = 0.0;
    }
    rbuf[i].assign(std::max(k, 1), 0.0);
    if (k <= 0) continue;
    const cosmo_dav_params q = batch_params(*p, k, n, 0.0);
    if (fits_single_launch(q, n, b->flags, b->single_nmax)) {
      small.push_back({nullptr, q, evals + off[i], rbuf[i].data(), &sts[i], 0});
      small_idx.push_back(i);
    } else {
      large.push_back(i);
    }
  }
  const int nsmall = (int)small.size(), nlarge = (int)large.size();
  std::vector<int> brc(nsmall, 0);
  std::atomic<int> next_s(0), built(0), next_l(0);
  auto t0 = std::chrono::steady_clock::now();
  // workers first build the small solves' operators, then take the large solves
  auto work = [&]() {
    cosmo_set_device(b->device);
    for (;;) {
      const int j = next_s.fetch_add(1);
      if (j >= nsmall) break;
      brc[j] = cosmo_op_create(b->sys[small_idx[j]], b->flags, &small[j].op);
      if (!brc[j] && hipStreamSynchronize(lib_stream()) != hipSuccess)
        brc[j] = fail(COSMO_EHIP, "cosmo_batch_solve: operator build");
      built.fetch_add(1);
    }
    for (;;) {
      const int j = next_l.fetch_add(1);
      if (j >= nlarge) break;
      const int i = large[j];
      const int k = (int)std::min<int64_t>(nume_per_solve[i], cosmo_system_dim(b->sys[i]));
      rcs[i] = solve_one(b->sys[i], b->flags, *p, k, 0.0, evals + off[i], rbuf[i].data(), &sts[i]);
    }
    release_thread_streams();
  };
  const int nthreads = std::max(1, std::min(std::max(nsmall, nlarge), 16));
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) th.emplace_back(work);
  while (built.load() < nsmall) std::this_thread::yield();
  int rc_small = COSMO_OK;
  for (int j = 0; j < nsmall; ++j)
    if (brc[j]) rc_small = brc[j];
  if (!rc_small) rc_small = run_small(small, *p);
  for (auto &t : th) t.join();
  for (int j = 0; j < nsmall; ++j) {
    rcs[small_idx[j]] = rc_small ? rc_small : small[j].rc;
    if (small[j].op) cosmo_op_destroy(small[j].op);
  }
  if (resid)
    for (int i = 0; i < ns; ++i) {
      const int k = (int)std::min<int64_t>(nume_per_solve[i], cosmo_system_dim(b->sys[i]));
      for (int j = 0; j < k; ++j) resid[off[i] + j] = rbuf[i][j];
    }
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
    if (rcs[i]) return fail(COSMO_ECONV, "batched solve " + std::to_string(i) + " stopped above its residual target");
  return COSMO_OK;
}

int cosmo_batch_set_single_launch(cosmo_batch b, int64_t nmax) {
  if (!b || nmax < 0) return fail(COSMO_EINVAL, "cosmo_batch_set_single_launch: bad args");
  b->single_nmax = nmax;
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
  const int64_t dav =
      davidson_footprint(sys->dim, nm, nm, (int)std::min<int64_t>(batch_lim(nm), std::max<int64_t>(sys->dim, 1)));
  r.bytes = ob + dav;
  if (r.bytes > S->p.hbm_budget) {
    // a Hamiltonian too large to store within the budget runs on the product form (no H
    // in HBM) when the system has one
    const int64_t fb = factored_footprint(*sys);
    if (fb >= 0 && fb + dav <= S->p.hbm_budget) {
      r.flags = COSMO_OP_FACTORED;
      r.bytes = fb + dav;
    }
  }
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
  if (flags & COSMO_OP_FACTORED) {
    *bytes = factored_footprint(*s);
    return *bytes < 0 ? fail(COSMO_EUNSUP, "cosmo_op_footprint: no product form for this system") : COSMO_OK;
  }
  if (flags & COSMO_OP_VIRTUAL) {
    *bytes = 8 * s->dim + (int64_t)s->orb.size() * s->orb.size() * s->orb.size() * s->orb.size();
    return COSMO_OK;
  }
  return stored_footprint(*s, flags, bytes);
}

}  // extern "C"
