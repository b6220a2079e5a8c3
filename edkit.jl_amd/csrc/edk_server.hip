// edk_server.hip -- serving loop: Lanczos ground-state and Krylov time-
// evolution jobs on one device, admitted as worker slots and HBM free up.
//
// Each worker thread owns a context (its own HIP stream, buffers from the
// device's block cache), so the jobs it runs overlap on the device with the
// other workers' jobs.  The queue is first-in first-out: its head is admitted when
// a worker is idle and the head's device footprint fits beside the jobs
// already running (a job that is larger than the whole budget runs when
// nothing else does).  A job's arithmetic depends only on its own inputs --
// every reduction has a fixed order -- so its result is the same bit for bit
// alone or in a mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "edk_krylov.h"

struct ServerJob {
  int64_t ticket = 0;
  edk_job job{};
  std::vector<char> psi0;  // host copy of the start vector (empty: unit vector)
  size_t footprint = 0;
  edk_job_result res{};
  std::vector<char> vec;   // result vector when requested
  bool done = false;
};

struct edk_server {
  int device = 0;
  size_t budget = 0;
  std::vector<std::thread> threads;
  std::vector<edk_ctx *> ctxs;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<ServerJob *> queue;
  std::unordered_map<int64_t, ServerJob *> jobs;
  size_t inflight = 0;
  int running = 0;
  bool stop = false;
  int64_t next_ticket = 1;
  int64_t completed = 0, peak_running = 0;
  size_t peak_inflight = 0;
  std::chrono::steady_clock::time_point t0;
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

static size_t job_esize(const edk_job &j) { return j.dtype == EDK_C64 ? 16 : 8; }

// device bytes a job holds at its peak (Krylov basis, work vectors, partials)
static size_t job_footprint(const edk_job &j) {
  const size_t D = (size_t)j.op->basis->D, es = job_esize(j);
  const size_t parts = (size_t)1024 * (EDK_KRYLOV_MAXM + 2) * 16;
  if (j.kind == EDK_JOB_GROUND) {
    // the Lanczos basis grows 64 -> 128 -> ... -> m + 1 columns; while it
    // grows, the old and the new basis coexist
    size_t cols = 0;
    for (size_t cap = std::min<size_t>(j.m + 1, 64); cap < (size_t)j.m + 1;) {
      const size_t ncap = std::min<size_t>(j.m + 1, 2 * cap);
      cols = std::max(cols, cap + ncap);
      cap = ncap;
    }
    cols = std::max<size_t>(cols, std::min<size_t>(j.m + 1, 64));
    // cached allocations round up by at most 1/8
    size_t f = D * es * (cols + 2) / 8 * 9 + parts;
    // the Ritz vector and its product (refinement, or the true-residual check
    // of a partially reorthogonalised ComplexF64 run with a tolerance) and a
    // restart vector
    if (j.refine || j.want_vector || (j.reorth == EDK_REORTH_PARTIAL && j.tol > 0)) f += 3 * D * 16;
    if (j.want_vector) f += D * es;
    return f;
  }
  const int mmax = j.m_max > 0 ? j.m_max : 60;
  return D * es * ((size_t)mmax + 5) + 2 * parts;
}

static void run_job(edk_server *S, edk_ctx *ctx, ServerJob *J) {
  const edk_job &j = J->job;
  edk_job_result &r = J->res;
  const int64_t D = j.op->basis->D;
  const size_t es = job_esize(j);
  void *x0 = nullptr, *out = nullptr;
  int rc = EDK_OK;
  if (edk_dalloc(ctx, &x0, D * es) != hipSuccess) rc = edk_fail(EDK_ERR_OOM, "job start vector");
  if (!rc) {
    if (!J->psi0.empty()) {
      if (hipMemcpyAsync(x0, J->psi0.data(), D * es, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        rc = edk_fail(EDK_ERR_HIP, "job start vector copy");
    } else {
      // productstate: the real unit vector at the representative's position
      // (ToolKit.jl:103-109)
      double one[2] = {1.0, 0.0};
      float onef[2] = {1.0f, 0.0f};
      if (hipMemsetAsync(x0, 0, D * es, ctx->stream) != hipSuccess ||
          hipMemcpyAsync((char *)x0 + (size_t)j.psi0_index * es, es == 16 ? (const void *)one : (const void *)onef,
                         es, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
          hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = edk_fail(EDK_ERR_HIP, "job product state");
    }
  }
  if (!rc && (j.want_vector || j.kind == EDK_JOB_EVOLVE) && edk_dalloc(ctx, &out, D * es) != hipSuccess)
    rc = edk_fail(EDK_ERR_OOM, "job result vector");
  if (!rc && j.kind == EDK_JOB_GROUND) {
    int steps = 0;
    rc = edk_lanczos_run(j.op, ctx, j.dtype, x0, j.m, j.tol, j.refine, &r.E0, &r.resid, &steps,
                         j.want_vector ? out : nullptr, j.reorth);
    r.steps = steps;
    r.matvecs = steps + (j.refine ? 1 : 0);
  } else if (!rc) {
    edk_krylov_params p;
    edk_krylov_default_params(&p);
    if (j.tol > 0) p.tol = j.tol;
    if (j.m_init > 0) p.m_init = j.m_init;
    if (j.m_max > 0) p.m_max = j.m_max;
    edk_krylov *c = nullptr;
    rc = edk_krylov_create_on(j.op, ctx, j.dtype, x0, &p, &c);
    if (!rc) rc = edk_krylov_evolve(c, j.t, out);
    if (!rc) {
      edk_krylov_diag d;
      edk_krylov_diagnostics(c, &d);
      r.matvecs = d.matvecs;
      r.steps = (int32_t)d.restarts;
      KrylovWork K;
      double2 ov;
      double nrm = 0;
      rc = K.init(ctx, D, j.dtype, 1);
      if (!rc) rc = K.dot(x0, out, &ov);
      if (!rc) rc = K.norm(out, &nrm);
      K.release();
      r.overlap_re = ov.x;
      r.overlap_im = ov.y;
      r.norm = nrm;
    }
    if (c) edk_krylov_destroy(c);
  }
  if (!rc && j.want_vector) {
    J->vec.resize(D * es);
    if (hipMemcpyAsync(J->vec.data(), out, D * es, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
      rc = edk_fail(EDK_ERR_HIP, "job result copy");
  }
  (void)hipStreamSynchronize(ctx->stream);
  edk_dfree(ctx, x0);
  edk_dfree(ctx, out);
  (void)hipStreamSynchronize(ctx->stream);
  r.status = rc;
  if (rc) {
    strncpy(r.message, edk_last_error(), sizeof(r.message) - 1);
    r.message[sizeof(r.message) - 1] = 0;
  }
}

static bool can_admit(const edk_server *S) {
  if (S->queue.empty()) return false;
  const ServerJob *h = S->queue.front();
  return S->running == 0 || S->inflight + h->footprint <= S->budget;
}

static void worker_main(edk_server *S, int w) {
  edk_ctx *ctx = S->ctxs[w];
  (void)hipSetDevice(S->device);
  while (true) {
    ServerJob *J = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_work.wait(lk, [&] { return S->stop || can_admit(S); });
      if (S->stop) return;
      J = S->queue.front();
      S->queue.pop_front();
      S->inflight += J->footprint;
      S->running++;
      S->peak_running = std::max<int64_t>(S->peak_running, S->running);
      S->peak_inflight = std::max(S->peak_inflight, S->inflight);
      J->res.t_start = S->now();
    }
    // the head may change: let another idle worker look at the new head
    S->cv_work.notify_one();
    run_job(S, ctx, J);
    {
      std::lock_guard<std::mutex> lk(S->mu);
      J->res.t_end = S->now();
      S->inflight -= J->footprint;
      S->running--;
      S->completed++;
      J->done = true;
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();
  }
}

extern "C" int edk_server_create(int device, int workers, size_t hbm_budget, edk_server **out) {
  if (!out || workers < 1 || workers > 64) return edk_fail(EDK_ERR_ARG, "edk_server_create: workers must be 1..64");
  int n = 0;
  EDK_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return edk_fail(EDK_ERR_ARG, "edk_server_create: no such device");
  EDK_HIP(hipSetDevice(device));
  if (hbm_budget == 0) {
    // 80 % of what is free, counting the blocks cached for reuse as free
    size_t fr = 0, tot = 0;
    EDK_HIP(hipMemGetInfo(&fr, &tot));
    hbm_budget = (fr + edk_cache_bytes(device)) / 10 * 8;
  }
  edk_server *S = new edk_server();
  S->device = device;
  S->budget = hbm_budget;
  S->t0 = std::chrono::steady_clock::now();
  for (int w = 0; w < workers; w++) {
    edk_ctx *c = nullptr;
    int rc = edk_ctx_create(device, nullptr, &c);
    if (rc) {
      for (edk_ctx *q : S->ctxs) edk_ctx_destroy(q);
      delete S;
      return rc;
    }
    c->cache_alloc = true;
    S->ctxs.push_back(c);
  }
  for (int w = 0; w < workers; w++) S->threads.emplace_back(worker_main, S, w);
  *out = S;
  return EDK_OK;
}

extern "C" int edk_server_submit(edk_server *S, const edk_job *job, int64_t *ticket) {
  if (!S || !job || !ticket || !job->op) return edk_fail(EDK_ERR_ARG, "edk_server_submit: null argument");
  const edk_job &j = *job;
  if (j.kind != EDK_JOB_GROUND && j.kind != EDK_JOB_EVOLVE) return edk_fail(EDK_ERR_ARG, "unknown job kind");
  if (j.dtype != EDK_C64 && j.dtype != EDK_C32) return edk_fail(EDK_ERR_ARG, "bad dtype");
  if (j.op->ctx->device != S->device) return edk_fail(EDK_ERR_ARG, "operator lives on another device");
  const int64_t D = j.op->basis->D;
  if (D < 1) return edk_fail(EDK_ERR_ARG, "empty sector");
  if (!j.psi0 && (j.psi0_index < 0 || j.psi0_index >= D)) return edk_fail(EDK_ERR_ARG, "product-state index out of range");
  if (j.kind == EDK_JOB_GROUND && (j.m < 1 || j.m > EDK_KRYLOV_MAXM))
    return edk_fail(EDK_ERR_ARG, "Lanczos steps must be 1..1024");
  if (j.kind == EDK_JOB_EVOLVE && !(j.t >= 0)) return edk_fail(EDK_ERR_ARG, "evolution time must be >= 0");
  if (j.kind == EDK_JOB_GROUND && j.reorth != EDK_REORTH_FULL && j.reorth != EDK_REORTH_PARTIAL)
    return edk_fail(EDK_ERR_ARG, "reorth must be EDK_REORTH_FULL or EDK_REORTH_PARTIAL");
  // the kernels' one-time setup runs here, in the caller's thread
  EDK_HIP(hipSetDevice(S->device));
  (void)edk_matvec_prepare(j.op);
  ServerJob *J = new ServerJob();
  J->job = j;
  if (j.psi0) {
    const size_t bytes = (size_t)D * job_esize(j);
    J->psi0.assign((const char *)j.psi0, (const char *)j.psi0 + bytes);
  }
  J->job.psi0 = nullptr;
  J->footprint = job_footprint(j);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    J->ticket = S->next_ticket++;
    J->res.t_submit = S->now();
    S->jobs[J->ticket] = J;
    S->queue.push_back(J);
    *ticket = J->ticket;
  }
  S->cv_work.notify_one();
  return EDK_OK;
}

extern "C" int edk_server_wait(edk_server *S, int64_t ticket, double timeout_s, edk_job_result *res, void *vec_host) {
  if (!S || !res) return edk_fail(EDK_ERR_ARG, "edk_server_wait: null argument");
  ServerJob *J = nullptr;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(ticket);
    if (it == S->jobs.end()) return edk_fail(EDK_ERR_ARG, "unknown ticket");
    J = it->second;
    auto ready = [&] { return J->done; };
    if (timeout_s < 0) S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
      return edk_fail(EDK_ERR_STATE, "job not finished");
    S->jobs.erase(it);
  }
  *res = J->res;
  if (vec_host && !J->vec.empty()) memcpy(vec_host, J->vec.data(), J->vec.size());
  delete J;
  return EDK_OK;
}

extern "C" double edk_server_clock(edk_server *S) { return S ? S->now() : 0.0; }

extern "C" int edk_server_stats(edk_server *S, edk_server_info *st) {
  if (!S || !st) return edk_fail(EDK_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  st->completed = S->completed;
  st->queued = (int64_t)S->queue.size();
  st->running = S->running;
  st->peak_running = S->peak_running;
  st->inflight_bytes = (int64_t)S->inflight;
  st->peak_inflight_bytes = (int64_t)S->peak_inflight;
  st->budget_bytes = (int64_t)S->budget;
  return EDK_OK;
}

extern "C" int edk_server_destroy(edk_server *S) {
  if (!S) return EDK_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  for (std::thread &t : S->threads) t.join();
  for (edk_ctx *c : S->ctxs) edk_ctx_destroy(c);
  for (auto &kv : S->jobs) delete kv.second;
  delete S;
  return EDK_OK;
}
