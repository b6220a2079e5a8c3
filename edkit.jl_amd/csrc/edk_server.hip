// edk_server.hip -- serving loop: Lanczos ground-state and Krylov time-
// evolution jobs on one device, admitted as worker slots and HBM free up,
// with their matvecs batched step by step across jobs.
//
// Each worker thread owns a context (its own HIP stream, buffers from the
// device's block cache) and runs one job at a time with the ordinary Krylov
// code.  Jobs whose operators share a basis and a two-site family structure
// (the same sector, e.g. XXZ chains of different Delta) form a group.  A
// group member's product H v is not launched by its own thread: the request
// waits in the group's batcher, and once every running member has one
// pending (or a short wait expires) the batcher launches them together --
// the members' vectors gathered into one row-interleaved block, one
// multi-coefficient launch that canonicalises and looks up every target once
// for all columns -- on the group's stream, which the members' streams then
// wait on.  New jobs join a running group as worker slots free up
// (admission prefers queued jobs that join a group, then first-in first-out),
// so a group's block stays full while requests come and go: continuous
// batching at the granularity of one Krylov step.  A column's arithmetic in
// the block launch does not depend on the block's other columns, so each
// result is the same bit for bit alone or in any mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "edk_fast.h"
#include "edk_krylov.h"

struct Batcher;

struct ServerJob {
  int64_t ticket = 0;
  edk_job job{};
  std::vector<char> psi0;  // host copy of the start vector (empty: unit vector)
  size_t footprint = 0;
  edk_job_result res{};
  std::vector<char> vec;   // result vector when requested
  bool done = false;
  bool batchable = false;  // the operator has a multi-coefficient launch
  uint64_t sig = 0;        // its family structure
};

// One pending product of a group member.
struct MvReq {
  edk_op *op = nullptr;
  void *y = nullptr;
  const void *x = nullptr;
  hipEvent_t ready = nullptr;  // recorded on the member's stream after x is written
  hipEvent_t done = nullptr;   // the block launch's completion
  bool issued = false;
  int rc = EDK_OK;
};

// Jobs on one basis with one family structure and dtype.
struct Batcher {
  const edk_basis *basis = nullptr;
  uint64_t sig = 0;
  int dtype = EDK_C64;
  int members = 0;
  std::deque<MvReq *> pending;
  int lane = -1;             // the server's group lane it launches on
  hipStream_t stream = nullptr;  // the lane's stream
  void *X = nullptr;         // D x 8 row-interleaved block (the lane's block cache)
  size_t xbytes = 0;
  hipEvent_t last_done = nullptr;  // completion of the last block launched
  int64_t launches = 0, columns = 0;
};

// A group lane: a context (stream, block cache) and completion events,
// created with the server and kept until it is destroyed, so serving never
// creates or destroys streams or events while other threads launch work.
struct GroupLane {
  edk_ctx *ctx = nullptr;
  std::vector<hipEvent_t> done_ev;
  int ev_next = 0;
  bool busy = false;
};

struct edk_server {
  int device = 0;
  size_t budget = 0;
  edk_server_params params{};
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
  // groups: keyed by (basis, structure, dtype); guarded by bmu
  std::mutex bmu;
  std::condition_variable bcv;
  std::vector<Batcher *> groups;
  std::vector<GroupLane> lanes;  // one per worker
  std::vector<hipEvent_t> ready_ev;  // per worker: its x is written (created with the server)
  std::atomic<size_t> group_bytes{0};  // the groups' blocks, counted in admission
  int64_t launches = 0, columns = 0;
  double wlaunches = 0, wcolumns = 0;  // the same weighted by the sector dimension
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
    // the Lanczos basis grows 64 -> 128 -> 256 -> ... -> m + 1 columns; while
    // it grows, the old and the new basis coexist.  The reservation covers
    // the growth to 256 columns: most solves stop well before 255 steps, and
    // the rare longer one grows into the memory kept outside the admission
    // budget (a fifth of the device)
    const size_t mres = std::min<size_t>(j.m + 1, 256);
    size_t cols = 0;
    for (size_t cap = std::min<size_t>(mres, 64); cap < mres;) {
      const size_t ncap = std::min<size_t>(mres, 2 * cap);
      cols = std::max(cols, cap + ncap);
      cap = ncap;
    }
    cols = std::max<size_t>(cols, std::min<size_t>(mres, 64));
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

// ---------------------------------------------------------------------------
// batching of group members' products
// ---------------------------------------------------------------------------

// Launches up to MC_MAXV pending requests of B as one block (bmu held): the
// group stream waits for each member's x, then gathers and multiplies.
static void flush_group(edk_server *S, Batcher *B) {
  edk_op *ops[MC_MAXV];
  const void *xs[MC_MAXV];
  void *ys[MC_MAXV];
  MvReq *batch[MC_MAXV];
  int n = 0;
  while (n < MC_MAXV && !B->pending.empty()) {
    MvReq *r = B->pending.front();
    B->pending.pop_front();
    batch[n] = r;
    ops[n] = r->op;
    xs[n] = r->x;
    ys[n] = r->y;
    n++;
  }
  if (n == 0) return;
  int rc = EDK_OK;
  for (int c = 0; c < n && !rc; c++)
    if (hipStreamWaitEvent(B->stream, batch[c]->ready, 0) != hipSuccess) rc = edk_fail(EDK_ERR_HIP, "batch wait");
  if (!rc) rc = edk_fast_matvec_mc(ops, n, B->dtype, xs, ys, B->X, B->stream);
  GroupLane &ln = S->lanes[B->lane];
  hipEvent_t ev = ln.done_ev[ln.ev_next];
  ln.ev_next = (ln.ev_next + 1) % (int)ln.done_ev.size();
  if (!rc && hipEventRecord(ev, B->stream) != hipSuccess) rc = edk_fail(EDK_ERR_HIP, "batch event");
  B->last_done = rc ? nullptr : ev;
  B->launches++;
  B->columns += n;
  S->launches++;
  S->columns += n;
  S->wlaunches += (double)B->basis->D;
  S->wcolumns += (double)n * (double)B->basis->D;
  for (int c = 0; c < n; c++) {
    batch[c]->done = ev;
    batch[c]->rc = rc;
    batch[c]->issued = true;
  }
  S->bcv.notify_all();
}

// The worker thread's matvec hook while it runs a group member.
struct GroupHook : edk_mv_hook {
  edk_server *S = nullptr;
  Batcher *B = nullptr;
  hipEvent_t ready = nullptr;
  int launch(edk_op *op, int dtype, void *y, const void *x, hipStream_t st) override {
    uint64_t sig = 0;
    if (!B || dtype != B->dtype || op->basis != B->basis || !edk_fast_signature(op, &sig) || sig != B->sig)
      return -1;  // not this group's product: launched directly
    MvReq r;
    r.op = op;
    r.y = y;
    r.x = x;
    r.ready = ready;
    if (hipEventRecord(ready, st) != hipSuccess) return edk_fail(EDK_ERR_HIP, "member event");
    {
      std::unique_lock<std::mutex> lk(S->bmu);
      B->pending.push_back(&r);
      if ((int)B->pending.size() >= std::min(B->members, MC_MAXV)) flush_group(S, B);
      // A partial block launches once the wait has passed and the group's
      // previous block has finished: while that block still runs, a launch
      // could not start any earlier on the group's stream, so late members
      // keep joining -- the members fall into step with each other.
      const auto wait = std::chrono::duration<double, std::micro>(S->params.batch_wait_us);
      while (!r.issued) {
        if (S->bcv.wait_for(lk, wait) == std::cv_status::timeout && !r.issued &&
            (!B->last_done || hipEventQuery(B->last_done) != hipErrorNotReady))
          flush_group(S, B);
      }
    }
    if (r.rc) return r.rc;
    EDK_HIP(hipStreamWaitEvent(st, r.done, 0));
    return EDK_OK;
  }
};

// group of a batchable job (created on first use, on a free lane; null when
// every lane is taken or its block does not fit); bmu held
static Batcher *group_of(edk_server *S, const ServerJob *J) {
  const edk_basis *b = J->job.op->basis;
  for (Batcher *B : S->groups)
    if (B->basis == b && B->sig == J->sig && B->dtype == J->job.dtype) return B;
  int lane = -1;
  for (size_t q = 0; q < S->lanes.size() && lane < 0; q++)
    if (!S->lanes[q].busy) lane = (int)q;
  if (lane < 0) return nullptr;
  GroupLane &ln = S->lanes[lane];
  const size_t es = J->job.dtype == EDK_C64 ? 16 : 8;
  const size_t bytes = (size_t)b->D * MC_MAXV * es;
  void *X = nullptr;
  if (edk_dalloc(ln.ctx, &X, bytes) != hipSuccess) return nullptr;
  Batcher *B = new Batcher();
  B->basis = b;
  B->sig = J->sig;
  B->dtype = J->job.dtype;
  B->lane = lane;
  B->stream = ln.ctx->stream;
  B->X = X;
  B->xbytes = bytes;
  ln.busy = true;
  S->group_bytes += bytes;
  S->groups.push_back(B);
  return B;
}

// a group with no members gives its lane back (its block to the cache); bmu held
static void release_group(edk_server *S, Batcher *B) {
  GroupLane &ln = S->lanes[B->lane];
  edk_dfree(ln.ctx, B->X);  // waits for the lane's stream
  ln.busy = false;
  S->group_bytes -= B->xbytes;
  S->groups.erase(std::find(S->groups.begin(), S->groups.end(), B));
  delete B;
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
    p.reorth = j.reorth;
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

// index in the queue of the next job to admit, or -1 (mu held).  With
// affinity: first a queued job that joins a running group with a free column;
// else the first job of the queued group whose shared products save the most
// (queued jobs, at most a block's width, times the sector dimension; a job
// alone scores zero), so expensive sectors start as full groups -- unless the
// head has waited longer than EDK_SERVE_HEAD_WAIT seconds, which bounds how
// long a job can be passed over.  The chosen job is admitted when its
// footprint fits beside the running jobs and the groups' blocks (always when
// nothing runs); otherwise admission waits for room.
#define EDK_SERVE_HEAD_WAIT 20.0
static int pick_job(edk_server *S) {
  if (S->queue.empty()) return -1;
  int idx = 0;
  const bool head_late = S->now() - S->queue.front()->res.t_submit > EDK_SERVE_HEAD_WAIT;
  if (S->params.affinity && S->params.batch && !head_late) {
    const int scan = (int)std::min<size_t>(S->queue.size(), 4096);
    int join = -1;
    {
      std::lock_guard<std::mutex> lk(S->bmu);
      for (int q = 0; q < scan && join < 0; q++) {
        const ServerJob *J = S->queue[q];
        if (!J->batchable) continue;
        for (const Batcher *B : S->groups)
          if (B->basis == J->job.op->basis && B->sig == J->sig && B->dtype == J->job.dtype && B->members > 0 &&
              B->members < MC_MAXV) {
            join = q;
            break;
          }
      }
    }
    if (join >= 0) {
      idx = join;
    } else {
      std::map<std::tuple<const void *, uint64_t, int>, std::pair<int, int>> groups;  // -> (first, count)
      double best_s = -1.0;
      for (int q = 0; q < scan; q++) {
        const ServerJob *J = S->queue[q];
        int first = q, n = 1;
        if (J->batchable) {
          auto &g = groups.try_emplace(std::make_tuple((const void *)J->job.op->basis, J->sig, J->job.dtype),
                                       std::make_pair(q, 0))
                        .first->second;
          g.second++;
          first = g.first;
          n = g.second;
        }
        const double score = (n > 1 ? std::min(n, MC_MAXV) : 0) * (double)J->job.op->basis->D;
        if (score > best_s || (score == best_s && first < idx)) idx = first, best_s = score;
      }
    }
  }
  const ServerJob *h = S->queue[idx];
  return (S->running == 0 || S->inflight + S->group_bytes + h->footprint <= S->budget) ? idx : -1;
}

static void worker_main(edk_server *S, int w) {
  edk_ctx *ctx = S->ctxs[w];
  (void)hipSetDevice(S->device);
  GroupHook hook;
  hook.S = S;
  hook.ready = S->ready_ev[w];
  while (true) {
    ServerJob *J = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      int idx = -1;
      S->cv_work.wait(lk, [&] { return S->stop || (idx = pick_job(S)) >= 0; });
      if (S->stop) break;
      J = S->queue[idx];
      S->queue.erase(S->queue.begin() + idx);
      S->inflight += J->footprint;
      S->running++;
      S->peak_running = std::max<int64_t>(S->peak_running, S->running);
      S->peak_inflight = std::max(S->peak_inflight, S->inflight);
      J->res.t_start = S->now();
    }
    // the head may change: let another idle worker look at the queue
    S->cv_work.notify_one();
    Batcher *B = nullptr;
    if (J->batchable && S->params.batch && hook.ready) {
      std::lock_guard<std::mutex> lk(S->bmu);
      B = group_of(S, J);
      if (B) B->members++;
    }
    hook.B = B;
    edk_set_mv_hook(B ? &hook : nullptr);
    run_job(S, ctx, J);
    edk_set_mv_hook(nullptr);
    if (B) {
      std::lock_guard<std::mutex> lk(S->bmu);
      B->members--;
      // the others may be waiting for this member: launch what is pending
      if (!B->pending.empty() && (int)B->pending.size() >= std::min(B->members, MC_MAXV)) flush_group(S, B);
      if (B->members == 0 && B->pending.empty()) release_group(S, B);
    }
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

extern "C" void edk_server_default_params(edk_server_params *p) {
  if (!p) return;
  p->batch = 1;
  p->batch_wait_us = 250.0;
  p->affinity = 1;
}

extern "C" int edk_server_create(int device, int workers, size_t hbm_budget, edk_server **out) {
  edk_server_params p;
  edk_server_default_params(&p);
  return edk_server_create_ex(device, workers, hbm_budget, &p, out);
}

extern "C" int edk_server_create_ex(int device, int workers, size_t hbm_budget, const edk_server_params *params,
                                    edk_server **out) {
  if (!out || !params || workers < 1 || workers > 64)
    return edk_fail(EDK_ERR_ARG, "edk_server_create: workers must be 1..64");
  if (!(params->batch_wait_us > 0) || params->batch_wait_us > 1e6)
    return edk_fail(EDK_ERR_ARG, "edk_server_create: batch_wait_us must be in (0, 1e6]");
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
  S->params = *params;
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
  // group lanes (one per worker: at most that many groups run at once) and
  // the workers' events, all before any worker runs
  S->ready_ev.assign(workers, nullptr);
  for (hipEvent_t &e : S->ready_ev) (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
  S->lanes.resize(workers);
  for (GroupLane &ln : S->lanes) {
    int rc = edk_ctx_create(device, nullptr, &ln.ctx);
    if (!rc) {
      ln.ctx->cache_alloc = true;
      ln.done_ev.assign(16, nullptr);
      for (hipEvent_t &e : ln.done_ev)
        if (!rc && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
          rc = edk_fail(EDK_ERR_HIP, "group lane event");
    }
    if (rc) {
      for (GroupLane &l2 : S->lanes) {
        for (hipEvent_t e : l2.done_ev)
          if (e) (void)hipEventDestroy(e);
        if (l2.ctx) edk_ctx_destroy(l2.ctx);
      }
      for (hipEvent_t e : S->ready_ev)
        if (e) (void)hipEventDestroy(e);
      for (edk_ctx *q : S->ctxs) edk_ctx_destroy(q);
      delete S;
      return rc;
    }
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
  if (j.reorth != EDK_REORTH_FULL && j.reorth != EDK_REORTH_PARTIAL)
    return edk_fail(EDK_ERR_ARG, "reorth must be EDK_REORTH_FULL or EDK_REORTH_PARTIAL");
  // the kernels' one-time setup runs here, in the caller's thread
  EDK_HIP(hipSetDevice(S->device));
  (void)edk_matvec_prepare(j.op);
  ServerJob *J = new ServerJob();
  J->job = j;
  J->batchable = edk_fast_signature(j.op, &J->sig);
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

extern "C" int edk_server_batch_stats(edk_server *S, int64_t *launches, int64_t *columns, double *dim_weighted) {
  if (!S || !launches || !columns || !dim_weighted) return edk_fail(EDK_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(S->bmu);
  *launches = S->launches;
  *columns = S->columns;
  *dim_weighted = S->wlaunches > 0 ? S->wcolumns / S->wlaunches : 0.0;
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
  {
    std::lock_guard<std::mutex> lk(S->bmu);
    while (!S->groups.empty()) release_group(S, S->groups.back());
  }
  for (GroupLane &ln : S->lanes) {
    (void)hipStreamSynchronize(ln.ctx->stream);
    for (hipEvent_t e : ln.done_ev) (void)hipEventDestroy(e);
    edk_ctx_destroy(ln.ctx);
  }
  for (hipEvent_t e : S->ready_ev)
    if (e) (void)hipEventDestroy(e);
  for (edk_ctx *c : S->ctxs) edk_ctx_destroy(c);
  for (auto &kv : S->jobs) delete kv.second;
  delete S;
  return EDK_OK;
}
