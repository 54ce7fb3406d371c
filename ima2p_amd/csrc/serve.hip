// serve.hip -- request loop: independent IM analyses and L-mode queries
// arriving over time share one GPU with continuous batching.
//
// Requests wait in arrival order and are admitted, first fit, while their
// HBM footprint fits the server's budget.  Admitted L-mode queries do not
// run one by one: every step gathers the pending points of all admitted
// queries on the same table (and mode) into one launch of up to max_points
// points, splitting large queries across launches, so a query that arrives
// while another is on the GPU joins the next launch.  An admitted analysis
// gets a context of its own (own stream, own captured sweep graphs) and
// advances a quantum of sweeps per step, overlapping the other analyses and
// the L-mode launches on the device.  Completion is polled through events.
// Building an admitted analysis' context (uploads, set-up, graph capture,
// first quantum) and collecting a finished one (result copies, context
// teardown, whose frees wait for the device) run on a worker thread, so
// the polling loop never waits for the device.  Latency is measured from
// submission to the step that sees the request complete.  Results do not depend on what else is running: an L-mode
// point's sums use the table's fixed chunking, and an analysis' sweep is
// the same however its sweeps are split into runs.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "ctx_impl.h"

using namespace ima;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

enum { RK_LMODE = 0, RK_ANALYSIS = 1 };
// analysis phases: loop-owned (IDLE), worker jobs pending (BUILDING,
// COLLECTING), worker done (BUILT, COLLECTED)
enum { PH_IDLE = 0, PH_BUILDING = 1, PH_BUILT = 2, PH_COLLECTING = 3, PH_COLLECTED = 4 };

struct AnalysisData {  // host copy of an ima_analysis
  int nloci = 0, nchains = 0, npops = 0, npsize = 0, nmig = 0, precision = 0, nsweeps = 0;
  double tmax = 0;
  uint64_t seed = 0;
  std::vector<int32_t> model, n, S, totsites, mult, samppop, plist, npresent, popdown, pope, join;
  std::vector<int32_t> theta_off, theta_wp, mig_off, mig_wp, up0, up1, down, root, A, pop, nmigev;
  std::vector<int32_t> mig_edge, mig_to;
  std::vector<int8_t> seq;
  std::vector<double> pi, qmax, mmax, time, mig_time, u, kappa, beta, tvals;
};

struct Req {
  int64_t id = 0;
  int kind = 0;
  int state = IMA_REQ_WAITING;
  double t_submit = 0, t_done = 0;
  size_t bytes = 0;
  std::string err;
  // L mode
  int table = -1, mode = 0, modeltype = 0, r0 = 0, r1 = 0, np = 0, npoints = 0;
  int dispatched = 0, finished = 0;
  std::vector<double> x, out, ess;
  // analysis.  `phase` hands the request between the polling loop and the
  // worker: the worker owns ctx, ev and the results while a job of its is
  // pending (phase BUILDING or COLLECTING), the loop owns them otherwise.
  std::unique_ptr<AnalysisData> a;
  ima_ctx *ctx = nullptr;
  int remaining = 0;
  hipEvent_t ev = nullptr, ev0 = nullptr;  // quantum end / start
  int q_last = 0;                           // sweeps in the quantum in flight
  double sweep_ms = -1.0;                   // measured GPU time per sweep
  bool inflight = false;
  std::atomic<int> phase{0};
  int wrc = 0;  // the worker's result code
  std::vector<double> pdg, probg, beta;
  uint64_t stats[8] = {};
};

struct Batch {  // one L-mode launch over points of several queries
  bool inflight = false;
  int table = -1, mode = 0, modeltype = 0, r0 = 0, r1 = 0, np = 0, npoints = 0;
  std::vector<std::array<int64_t, 3>> segs;  // (request, first point, count)
  double *hx = nullptr, *hout = nullptr, *hess = nullptr;  // pinned host staging
  double *dx = nullptr, *dout = nullptr, *dess = nullptr, *dwork = nullptr;
  size_t cap_points = 0, work_bytes = 0;
  hipEvent_t ev = nullptr;
};

}  // namespace

struct ima_server {
  int device = 0;
  size_t budget = 0, used = 0;
  ima_ctx *lctx = nullptr;
  std::map<int64_t, std::unique_ptr<Req>> reqs;
  std::deque<int64_t> queue;   // waiting, in arrival order
  std::vector<int64_t> active;  // admitted, not finished
  Batch batch;
  int64_t next_id = 1;
  int quantum = 25;  // most sweeps per quantum
  // Quanta are sized to about this much GPU time: streams share the
  // device's few hardware queues, so work queued by one analysis can hold
  // up an L-mode launch behind it for as long as one quantum runs.
  double quantum_ms = 4.0;
  int max_points = 8192;
  int max_analyses = 8;
  std::vector<std::pair<int, double>> lat;  // (kind, seconds) of finished requests
  std::string err;
  // worker thread: builds and collects analyses
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Req *> jobs;
  bool stop = false;
  // bytes of finished analyses whose contexts the worker is still freeing
  std::atomic<long long> freeing{0};
};

namespace {

int sfail(ima_server *s, int rc, const std::string &msg) {
  s->err = msg;
  return rc;
}

size_t lmode_bytes(const LTable &t, int npoints) {
  const int np = t.npsize + t.nmig;
  return (size_t)npoints * (size_t)(np * 8 + 16) + jointp_work_bytes(t, npoints);
}

// Per-context device memory beside the data: captured graphs, kernel
// argument buffers, events and the small fixed buffers of the sweep.
constexpr size_t kCtxOverhead = 16u << 20;

size_t analysis_bytes(const AnalysisData &a) {
  const int P = a.npops, R = prior_rec_len(P);
  size_t b = kCtxOverhead;
  // each genealogy twice: the chain's slot and the proposal slot of the
  // genealogy move, each with room for twice its migrations plus 32
  for (int c = 0; c < a.nchains; ++c)
    for (int l = 0; l < a.nloci; ++l) {
      const size_t n = (size_t)a.n[l];
      b += 2 * ((2 * n - 1) * (5 * 4 + 2 * 8) + (n - 1) * sizeof(ProgOp) + 32 * 16);
    }
  for (int l = 0; l < a.nloci; ++l) b += (size_t)a.n[l] * (size_t)((a.S[l] + 63) / 64 * 64 + 64);
  b += (size_t)a.nchains * a.nloci * ((size_t)R * 16 + 128);
  b += a.mig_time.size() * 4 * 16;
  return b;
}

void finish(ima_server *s, Req &r, int state) {
  r.state = state;
  r.t_done = now_s();
  if (state == IMA_REQ_DONE) s->lat.push_back({r.kind, r.t_done - r.t_submit});
  s->used -= std::min(s->used, r.bytes);
  s->active.erase(std::remove(s->active.begin(), s->active.end(), r.id), s->active.end());
}

void drop_ctx(Req &r) {
  if (r.ev) hipEventDestroy(r.ev);
  if (r.ev0) hipEventDestroy(r.ev0);
  r.ev = r.ev0 = nullptr;
  if (r.ctx) ima_ctx_destroy(r.ctx);
  r.ctx = nullptr;
}

#define STRY(call)                            \
  do {                                        \
    const int rc_ = (call);                   \
    if (rc_ != IMA_OK) return rc_;            \
  } while (0)

// Build the analysis' context, upload everything and set up the sweep.
int build_analysis(ima_server *s, Req &r) {
  AnalysisData &a = *r.a;
  STRY(ima_ctx_create(s->device, &r.ctx));
  ima_ctx *c = r.ctx;
  int mid = -1;
  STRY(ima_model_add(c, a.npops, a.plist.data(), a.npresent.data(), a.popdown.data(),
                     a.pope.data(), a.join.data(), a.npsize, a.theta_off.data(),
                     a.theta_wp.empty() ? nullptr : a.theta_wp.data(), a.nmig,
                     a.mig_off.data(), a.mig_wp.empty() ? nullptr : a.mig_wp.data(),
                     a.qmax.empty() ? nullptr : a.qmax.data(),
                     a.mmax.empty() ? nullptr : a.mmax.data(), &mid));
  std::vector<int32_t> lid(a.nloci);
  size_t soff = 0, moff = 0;
  for (int l = 0; l < a.nloci; ++l) {
    const int n = a.n[l];
    if (a.model[l] == IMA_MODEL_HKY) {
      STRY(ima_locus_add_hky(c, n, a.S[l], a.seq.data() + soff, a.mult.data() + moff,
                             a.totsites[l], a.pi.data() + 4 * l, &lid[l]));
      soff += (size_t)n * a.S[l];
      moff += (size_t)a.S[l];
    } else if (a.model[l] == IMA_MODEL_IS) {
      STRY(ima_locus_add_is(c, n, a.S[l], a.seq.data() + soff, &lid[l]));
      soff += (size_t)n * a.S[l];
    } else {
      STRY(ima_locus_add_sw(c, n, &lid[l]));
    }
    STRY(ima_locus_set_model(c, lid[l], mid, a.samppop.data() + (size_t)l * a.npops, 1.0));
  }
  const int N = a.nchains * a.nloci;
  std::vector<int32_t> of(N);
  for (int i = 0; i < N; ++i) of[i] = lid[i % a.nloci];
  int first = 0;
  STRY(ima_slots_alloc(c, N, of.data(), &first));
  STRY(ima_slots_write(c, first, N, a.up0.data(), a.up1.data(), a.down.data(), a.time.data(),
                       a.root.data(), a.A.empty() ? nullptr : a.A.data()));
  STRY(ima_slots_write_pops(c, first, N, a.pop.data(), a.nmigev.data(),
                            a.mig_edge.empty() ? nullptr : a.mig_edge.data(),
                            a.mig_time.empty() ? nullptr : a.mig_time.data(),
                            a.mig_to.empty() ? nullptr : a.mig_to.data()));
  for (int l = 0; l < a.nloci; ++l)
    if (a.model[l] == IMA_MODEL_IS) {
      double sl = 0.0;
      STRY(ima_locus_sumlogk_from_slot(c, lid[l], first + l, &sl));
    }
  STRY(ima_mc3_setup(c, a.nchains, a.nloci, first, a.u.data(), a.kappa.data(), a.beta.data(),
                     a.seed, a.precision, std::max(1, a.nchains / 10), 0.2));
  STRY(ima_mc3_set_prior(c, a.tvals.empty() ? nullptr : a.tvals.data(), a.tmax));
  if (hipEventCreate(&r.ev) != hipSuccess || hipEventCreate(&r.ev0) != hipSuccess)
    return IMA_ERR_HIP;
  return IMA_OK;
}

// A served analysis' sweep: the reference's step, as the benchmark runs it
// (updategenealogy's move, RY1 or NW on alternate sweeps, the mutation-rate
// rounds every fifth sweep, the swap).
constexpr int kServedSweep = IMA_MC3_GRAPH | IMA_MC3_TOPOLOGY | IMA_MC3_SCALE | IMA_MC3_RY1 |
                             IMA_MC3_NW | IMA_MC3_UPDATE_U;

// The first quantum is one sweep, to measure; later ones hold about
// quantum_ms of GPU time at the measured rate.
int run_quantum(ima_server *s, Req &r) {
  int q = 1;
  if (r.sweep_ms > 0)
    q = (int)std::max(1.0, std::min((double)s->quantum, s->quantum_ms / r.sweep_ms));
  q = std::min(q, r.remaining);
  if (hipEventRecord(r.ev0, r.ctx->stream) != hipSuccess) return IMA_ERR_HIP;
  STRY(mc3_run_async(r.ctx, q, kServedSweep));
  if (r.sweep_ms <= 0) {
    // the first quantum made the move's proposal slots: charge what the
    // context holds now
    const size_t real = ctx_device_bytes(r.ctx) + kCtxOverhead;
    s->used = s->used - std::min(s->used, r.bytes) + real;
    r.bytes = real;
  }
  if (hipEventRecord(r.ev, r.ctx->stream) != hipSuccess) return IMA_ERR_HIP;
  r.q_last = q;
  r.remaining -= q;
  r.inflight = true;
  return IMA_OK;
}

int collect_analysis(Req &r) {
  const AnalysisData &a = *r.a;
  r.pdg.assign((size_t)a.nchains * a.nloci, 0.0);
  r.probg.assign(a.nchains, 0.0);
  r.beta.assign(a.nchains, 0.0);
  uint64_t sw = 0;
  STRY(ima_mc3_read(r.ctx, r.pdg.data(), r.beta.data(), r.stats, &sw));
  std::vector<double> tv((size_t)a.nchains * std::max(a.npops - 1, 1));
  STRY(ima_mc3_read_prior(r.ctx, r.probg.data(), tv.data()));
  return IMA_OK;
}

// Hand r to the worker for `phase` (PH_BUILDING or PH_COLLECTING).
void post(ima_server *s, Req &r, int phase) {
  r.phase.store(phase, std::memory_order_release);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->jobs.push_back(&r);
  }
  s->cv.notify_one();
}

// The worker: builds admitted analyses (and launches their first quantum,
// capturing the sweep graphs) and collects finished ones.  A collected
// request is published before its context is torn down, so the frees'
// wait for the device does not delay the completion; the bytes stay
// counted in `freeing` until they are back.
void worker_main(ima_server *s) {
  hipSetDevice(s->device);
  for (;;) {
    Req *r = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] { return s->stop || !s->jobs.empty(); });
      if (s->stop) return;
      r = s->jobs.front();
      s->jobs.pop_front();
    }
    if (r->phase.load(std::memory_order_acquire) == PH_BUILDING) {
      int rc = build_analysis(s, *r);
      if (rc == IMA_OK && r->remaining > 0) rc = run_quantum(s, *r);
      if (rc != IMA_OK) {
        r->err = r->ctx ? ima_ctx_last_error(r->ctx) : "analysis set-up failed";
        drop_ctx(*r);
      }
      r->wrc = rc;
      r->phase.store(PH_BUILT, std::memory_order_release);
      continue;
    }
    // PH_COLLECTING: wrc != IMA_OK marks a failed run, torn down unread
    ima_ctx *ctx = r->ctx;
    hipEvent_t ev = r->ev, ev0 = r->ev0;
    const long long bytes = (long long)r->bytes;
    int rc = r->wrc;
    if (rc == IMA_OK) {
      rc = collect_analysis(*r);
      if (rc != IMA_OK) r->err = ima_ctx_last_error(ctx);
    }
    r->ctx = nullptr;
    r->ev = r->ev0 = nullptr;
    r->wrc = rc;
    s->freeing.fetch_add(bytes);
    r->phase.store(PH_COLLECTED, std::memory_order_release);
    // r may be released from here on
    if (ev) hipEventDestroy(ev);
    if (ev0) hipEventDestroy(ev0);
    if (ctx) ima_ctx_destroy(ctx);
    s->freeing.fetch_sub(bytes);
  }
}

int ensure_batch_buffers(ima_server *s, int npoints, int np, size_t work) {
  Batch &b = s->batch;
  (void)np;  // at most 16 parameters per point
  if ((size_t)npoints > b.cap_points) {
    const size_t cap = std::max<size_t>(npoints, (size_t)s->max_points);
    if (b.hx) hipHostFree(b.hx);
    if (b.hout) hipHostFree(b.hout);
    if (b.hess) hipHostFree(b.hess);
    if (b.dx) hipFree(b.dx);
    if (b.dout) hipFree(b.dout);
    if (b.dess) hipFree(b.dess);
    b.hx = b.hout = b.hess = b.dx = b.dout = b.dess = nullptr;
    if (hipHostMalloc(&b.hx, cap * 16 * 8) != hipSuccess ||
        hipHostMalloc(&b.hout, cap * 8) != hipSuccess ||
        hipHostMalloc(&b.hess, cap * 8) != hipSuccess ||
        hipMalloc(&b.dx, cap * 16 * 8) != hipSuccess || hipMalloc(&b.dout, cap * 8) != hipSuccess ||
        hipMalloc(&b.dess, cap * 8) != hipSuccess)
      return IMA_ERR_HIP;
    b.cap_points = cap;
  }
  if (work > b.work_bytes) {
    if (b.dwork) hipFree(b.dwork);
    b.dwork = nullptr;
    if (hipMalloc(&b.dwork, work) != hipSuccess) return IMA_ERR_HIP;
    b.work_bytes = work;
  }
  if (!b.ev && hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) != hipSuccess)
    return IMA_ERR_HIP;
  return IMA_OK;
}

// Gather the pending points of admitted queries that share the oldest
// pending query's (table, mode, model) into one launch.
int launch_batch(ima_server *s) {
  Batch &b = s->batch;
  Req *head = nullptr;
  for (int64_t id : s->active) {
    Req &r = *s->reqs[id];
    if (r.kind == RK_LMODE && r.dispatched < r.npoints) {
      head = &r;
      break;
    }
  }
  if (!head) return IMA_OK;
  b.table = head->table;
  b.mode = head->mode;
  b.modeltype = head->modeltype;
  b.r0 = head->r0;
  b.r1 = head->r1;
  b.np = head->np;
  b.segs.clear();
  int total = 0;
  for (int64_t id : s->active) {
    if (total >= s->max_points) break;
    Req &r = *s->reqs[id];
    if (r.kind != RK_LMODE || r.dispatched >= r.npoints || r.table != b.table ||
        r.mode != b.mode || r.modeltype != b.modeltype || r.r0 != b.r0 || r.r1 != b.r1)
      continue;
    const int take = std::min(r.npoints - r.dispatched, s->max_points - total);
    b.segs.push_back({r.id, r.dispatched, take});
    r.dispatched += take;
    total += take;
  }
  b.npoints = total;
  const LTable t = ctx_ltable(s->lctx, b.table);
  const size_t work = jointp_work_bytes(t, total);
  STRY(ensure_batch_buffers(s, total, b.np, work));
  int off = 0;
  for (const auto &sg : b.segs) {
    const Req &r = *s->reqs[sg[0]];
    std::copy(r.x.begin() + sg[1] * b.np, r.x.begin() + (sg[1] + sg[2]) * b.np,
              b.hx + (size_t)off * b.np);
    off += (int)sg[2];
  }
  hipStream_t st = s->lctx->stream;
  if (hipMemcpyAsync(b.dx, b.hx, (size_t)total * b.np * 8, hipMemcpyHostToDevice, st) !=
      hipSuccess)
    return IMA_ERR_HIP;
  if (launch_jointp(t, b.dx, total, b.modeltype, b.r0, b.r1, b.mode, b.dwork, b.work_bytes,
                    b.dout, b.dess, st) != hipSuccess)
    return IMA_ERR_HIP;
  if (hipMemcpyAsync(b.hout, b.dout, (size_t)total * 8, hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipMemcpyAsync(b.hess, b.dess, (size_t)total * 8, hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipEventRecord(b.ev, st) != hipSuccess)
    return IMA_ERR_HIP;
  b.inflight = true;
  return IMA_OK;
}

}  // namespace

extern "C" {

int ima_server_create(int device, double budget_bytes, int quantum, ima_server **out) {
  if (!out) return IMA_ERR_ARG;
  *out = nullptr;
  auto s = std::make_unique<ima_server>();
  s->device = device;
  const int rc = ima_ctx_create(device, &s->lctx);
  if (rc != IMA_OK) return rc;
  if (budget_bytes > 0) {
    s->budget = (size_t)budget_bytes;
  } else {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return IMA_ERR_HIP;
    s->budget = fr / 10 * 8;
  }
  if (quantum > 0) s->quantum = quantum;
  s->worker = std::thread(worker_main, s.get());
  *out = s.release();
  return IMA_OK;
}

void ima_server_destroy(ima_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->worker.joinable()) s->worker.join();
  hipSetDevice(s->device);
  for (auto &kv : s->reqs) {
    Req &r = *kv.second;
    if (r.ctx) hipStreamSynchronize(r.ctx->stream);
    drop_ctx(r);
  }
  if (s->lctx) hipStreamSynchronize(s->lctx->stream);
  Batch &b = s->batch;
  if (b.hx) hipHostFree(b.hx);
  if (b.hout) hipHostFree(b.hout);
  if (b.hess) hipHostFree(b.hess);
  if (b.dx) hipFree(b.dx);
  if (b.dout) hipFree(b.dout);
  if (b.dess) hipFree(b.dess);
  if (b.dwork) hipFree(b.dwork);
  if (b.ev) hipEventDestroy(b.ev);
  if (s->lctx) ima_ctx_destroy(s->lctx);
  delete s;
}

const char *ima_server_last_error(const ima_server *s) { return s ? s->err.c_str() : ""; }

int ima_server_table(ima_server *s, const float *rows, int G, int K, int npsize, int nmig,
                     int npops, int *table) {
  if (!s || !table) return IMA_ERR_ARG;
  const int rc = ima_lmode_table(s->lctx, rows, G, K, npsize, nmig, npops, table);
  if (rc != IMA_OK) return sfail(s, rc, ima_ctx_last_error(s->lctx));
  s->used += (size_t)G * K * 4;
  // size the shared launch buffers for a full launch on this table now, so
  // that no step has to grow them (a free waits for the whole device)
  // (buffers in use by a launch grow later, when it has finished)
  const LTable t = ctx_ltable(s->lctx, *table);
  if (!s->batch.inflight) {
    const int brc = ensure_batch_buffers(s, s->max_points, t.npsize + t.nmig,
                                         jointp_work_bytes(t, s->max_points));
    if (brc != IMA_OK) return sfail(s, brc, "L-mode launch buffers");
  }
  return IMA_OK;
}

int ima_server_submit_lmode(ima_server *s, int table, const double *x, int npoints,
                            int modeltype, int r0, int r1, int mode, int64_t *req) {
  if (!s || !x || npoints < 1 || !req) return IMA_ERR_ARG;
  const LTable t = ctx_ltable(s->lctx, table);
  if (!t.rows) return sfail(s, IMA_ERR_ARG, "no such table");
  const int np = t.npsize + t.nmig;
  if (r0 < 0 || r1 > np || r0 >= r1 || modeltype < 0 || modeltype > 2 ||
      (mode != IMA_LMODE_REF_COMPAT && mode != IMA_LMODE_EXACT))
    return sfail(s, IMA_ERR_ARG, "bad query");
  auto r = std::make_unique<Req>();
  r->id = s->next_id++;
  r->kind = RK_LMODE;
  r->t_submit = now_s();
  r->table = table;
  r->mode = mode;
  r->modeltype = modeltype;
  r->r0 = r0;
  r->r1 = r1;
  r->np = np;
  r->npoints = npoints;
  r->x.assign(x, x + (size_t)npoints * np);
  r->out.assign(npoints, 0.0);
  r->ess.assign(npoints, 0.0);
  r->bytes = lmode_bytes(t, npoints);
  if (r->bytes > s->budget) return sfail(s, IMA_ERR_SIZE, "query exceeds the memory budget");
  *req = r->id;
  s->queue.push_back(r->id);
  s->reqs[r->id] = std::move(r);
  return IMA_OK;
}

int ima_server_submit_analysis(ima_server *s, const ima_analysis *d, int64_t *req) {
  if (!s || !d || !req || d->nloci < 1 || d->nchains < 1 || d->npops < 1 || d->nsweeps < 0)
    return IMA_ERR_ARG;
  auto a = std::make_unique<AnalysisData>();
  const int L = d->nloci, C = d->nchains, P = d->npops, T = 2 * P - 1, N = C * L;
  a->nloci = L;
  a->nchains = C;
  a->npops = P;
  a->npsize = d->npsize;
  a->nmig = d->nmig;
  a->precision = d->precision;
  a->nsweeps = d->nsweeps;
  a->tmax = d->tmax;
  a->seed = d->seed;
  auto cp = [](auto &v, const auto *p, size_t n) {
    if (p && n) v.assign(p, p + n);
  };
  cp(a->model, d->model, L);
  cp(a->n, d->n, L);
  cp(a->S, d->S, L);
  cp(a->totsites, d->totsites, L);
  cp(a->pi, d->pi, 4 * (size_t)L);
  cp(a->samppop, d->samppop, (size_t)L * P);
  size_t nseq = 0, nmult = 0, nnodes = 0;
  for (int l = 0; l < L; ++l) {
    if (d->model[l] == IMA_MODEL_HKY || d->model[l] == IMA_MODEL_IS)
      nseq += (size_t)d->n[l] * d->S[l];
    if (d->model[l] == IMA_MODEL_HKY) nmult += (size_t)d->S[l];
  }
  for (int i = 0; i < N; ++i) nnodes += 2 * (size_t)d->n[i % L] - 1;
  cp(a->seq, d->seq, nseq);
  cp(a->mult, d->mult, nmult);
  cp(a->plist, d->plist, (size_t)P * P);
  cp(a->npresent, d->npresent, P);
  cp(a->popdown, d->popdown, T);
  cp(a->pope, d->pope, T);
  cp(a->join, d->join, 3 * (size_t)P);
  cp(a->theta_off, d->theta_off, (size_t)d->npsize + 1);
  cp(a->theta_wp, d->theta_wp, 2 * (size_t)d->theta_off[d->npsize]);
  cp(a->mig_off, d->mig_off, (size_t)d->nmig + 1);
  cp(a->mig_wp, d->mig_wp, 3 * (size_t)d->mig_off[d->nmig]);
  cp(a->qmax, d->qmax, d->npsize);
  cp(a->mmax, d->mmax, d->nmig);
  cp(a->up0, d->up0, nnodes);
  cp(a->up1, d->up1, nnodes);
  cp(a->down, d->down, nnodes);
  cp(a->time, d->time, nnodes);
  cp(a->pop, d->pop, nnodes);
  cp(a->A, d->A, nnodes);
  cp(a->root, d->root, N);
  cp(a->nmigev, d->nmigev, N);
  size_t nm = 0;
  for (int i = 0; i < N; ++i) nm += (size_t)d->nmigev[i];
  cp(a->mig_edge, d->mig_edge, nm);
  cp(a->mig_time, d->mig_time, nm);
  cp(a->mig_to, d->mig_to, nm);
  cp(a->u, d->u, N);
  cp(a->kappa, d->kappa, N);
  cp(a->beta, d->beta, C);
  cp(a->tvals, d->tvals, (size_t)C * (P > 1 ? P - 1 : 0));
  auto r = std::make_unique<Req>();
  r->id = s->next_id++;
  r->kind = RK_ANALYSIS;
  r->t_submit = now_s();
  r->bytes = analysis_bytes(*a);
  r->remaining = a->nsweeps;
  r->a = std::move(a);
  if (r->bytes > s->budget) return sfail(s, IMA_ERR_SIZE, "analysis exceeds the memory budget");
  *req = r->id;
  s->queue.push_back(r->id);
  s->reqs[r->id] = std::move(r);
  return IMA_OK;
}

int ima_server_step(ima_server *s, int *pending) {
  if (!s) return IMA_ERR_ARG;
  hipSetDevice(s->device);
  // 1. completions
  Batch &b = s->batch;
  if (b.inflight) {
    const hipError_t q = hipEventQuery(b.ev);
    if (q == hipSuccess) {
      b.inflight = false;
      int off = 0;
      for (const auto &sg : b.segs) {
        Req &r = *s->reqs[sg[0]];
        std::copy(b.hout + off, b.hout + off + sg[2], r.out.begin() + sg[1]);
        std::copy(b.hess + off, b.hess + off + sg[2], r.ess.begin() + sg[1]);
        off += (int)sg[2];
        r.finished += (int)sg[2];
        if (r.finished == r.npoints) finish(s, r, IMA_REQ_DONE);
      }
    } else if (q != hipErrorNotReady) {
      return sfail(s, IMA_ERR_HIP, hipGetErrorString(q));
    }
  }
  const std::vector<int64_t> act = s->active;
  for (int64_t id : act) {
    Req &r = *s->reqs[id];
    if (r.kind != RK_ANALYSIS) continue;
    const int ph = r.phase.load(std::memory_order_acquire);
    if (ph == PH_BUILDING || ph == PH_COLLECTING) continue;
    if (ph == PH_COLLECTED) {
      finish(s, r, r.wrc == IMA_OK ? IMA_REQ_DONE : IMA_REQ_FAILED);
      continue;
    }
    if (ph == PH_BUILT) {
      r.phase.store(PH_IDLE, std::memory_order_relaxed);
      if (r.wrc != IMA_OK) {
        finish(s, r, IMA_REQ_FAILED);
        continue;
      }
      // charge what the context holds rather than the estimate
      const size_t real = ctx_device_bytes(r.ctx) + kCtxOverhead;
      s->used = s->used - std::min(s->used, r.bytes) + real;
      r.bytes = real;
      if (!r.inflight) {  // no sweeps asked for
        post(s, r, PH_COLLECTING);
        continue;
      }
    }
    if (!r.inflight) continue;
    const hipError_t q = hipEventQuery(r.ev);
    if (q == hipErrorNotReady) continue;
    r.inflight = false;
    int rc = q == hipSuccess ? IMA_OK : IMA_ERR_HIP;
    if (rc != IMA_OK) r.err = hipGetErrorString(q);
    float ms = 0.0f;
    if (rc == IMA_OK && r.q_last > 0 && hipEventElapsedTime(&ms, r.ev0, r.ev) == hipSuccess) {
      const double per = (double)ms / r.q_last;
      r.sweep_ms = r.sweep_ms > 0 ? 0.5 * (r.sweep_ms + per) : per;
    }
    if (rc == IMA_OK && r.remaining > 0) {
      rc = run_quantum(s, r);
      if (rc == IMA_OK) continue;
      r.err = ima_ctx_last_error(r.ctx);
    }
    r.wrc = rc;
    post(s, r, PH_COLLECTING);
  }
  // 2. admission, first fit in arrival order
  int running = 0;
  for (int64_t id : s->active) running += s->reqs[id]->kind == RK_ANALYSIS;
  const size_t freeing = (size_t)std::max(0LL, s->freeing.load());
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    Req &r = *s->reqs[*it];
    const bool room = s->used + freeing + r.bytes <= s->budget &&
                      (r.kind != RK_ANALYSIS || running < s->max_analyses);
    if (!room) {
      ++it;
      continue;
    }
    it = s->queue.erase(it);
    s->used += r.bytes;
    r.state = IMA_REQ_RUNNING;
    s->active.push_back(r.id);
    if (r.kind == RK_ANALYSIS) {
      ++running;
      post(s, r, PH_BUILDING);
    }
  }
  // 3. the next L-mode launch
  if (!b.inflight) {
    const int rc = launch_batch(s);
    if (rc != IMA_OK) return sfail(s, rc, "L-mode launch failed");
  }
  if (pending) *pending = (int)(s->queue.size() + s->active.size());
  return IMA_OK;
}

int ima_server_status(ima_server *s, int64_t req, int *state, double *latency_s,
                      double *bytes) {
  if (!s) return IMA_ERR_ARG;
  auto it = s->reqs.find(req);
  if (it == s->reqs.end()) return sfail(s, IMA_ERR_ARG, "no such request");
  const Req &r = *it->second;
  if (state) *state = r.state;
  if (latency_s) *latency_s = r.state >= IMA_REQ_DONE ? r.t_done - r.t_submit : -1.0;
  if (bytes) *bytes = (double)r.bytes;
  return IMA_OK;
}

int ima_server_lmode_result(ima_server *s, int64_t req, double *out, double *ess) {
  if (!s) return IMA_ERR_ARG;
  auto it = s->reqs.find(req);
  if (it == s->reqs.end() || it->second->kind != RK_LMODE)
    return sfail(s, IMA_ERR_ARG, "no such L-mode request");
  const Req &r = *it->second;
  if (r.state != IMA_REQ_DONE) return sfail(s, IMA_ERR_STATE, "request not finished");
  if (out) std::copy(r.out.begin(), r.out.end(), out);
  if (ess) std::copy(r.ess.begin(), r.ess.end(), ess);
  return IMA_OK;
}

int ima_server_analysis_result(ima_server *s, int64_t req, double *pdg, double *probg,
                               double *beta, uint64_t *stats) {
  if (!s) return IMA_ERR_ARG;
  auto it = s->reqs.find(req);
  if (it == s->reqs.end() || it->second->kind != RK_ANALYSIS)
    return sfail(s, IMA_ERR_ARG, "no such analysis");
  const Req &r = *it->second;
  if (r.state == IMA_REQ_FAILED) return sfail(s, IMA_ERR_STATE, "analysis failed: " + r.err);
  if (r.state != IMA_REQ_DONE) return sfail(s, IMA_ERR_STATE, "request not finished");
  if (pdg) std::copy(r.pdg.begin(), r.pdg.end(), pdg);
  if (probg) std::copy(r.probg.begin(), r.probg.end(), probg);
  if (beta) std::copy(r.beta.begin(), r.beta.end(), beta);
  if (stats) std::copy(r.stats, r.stats + 8, stats);
  return IMA_OK;
}

int ima_server_release(ima_server *s, int64_t req) {
  if (!s) return IMA_ERR_ARG;
  auto it = s->reqs.find(req);
  if (it == s->reqs.end()) return sfail(s, IMA_ERR_ARG, "no such request");
  if (it->second->state < IMA_REQ_DONE) return sfail(s, IMA_ERR_STATE, "request not finished");
  s->reqs.erase(it);
  return IMA_OK;
}

int ima_server_latency(ima_server *s, int kind, int *count, double *p50, double *p99,
                       double *mean) {
  if (!s) return IMA_ERR_ARG;
  std::vector<double> v;
  for (const auto &x : s->lat)
    if (kind < 0 || x.first == kind) v.push_back(x.second);
  std::sort(v.begin(), v.end());
  if (count) *count = (int)v.size();
  auto q = [&](double f) {
    if (v.empty()) return 0.0;
    const size_t i = std::min(v.size() - 1, (size_t)(f * (double)(v.size() - 1) + 0.5));
    return v[i];
  };
  if (p50) *p50 = q(0.50);
  if (p99) *p99 = q(0.99);
  if (mean) {
    double t = 0.0;
    for (double x : v) t += x;
    *mean = v.empty() ? 0.0 : t / (double)v.size();
  }
  return IMA_OK;
}

int ima_server_usage(ima_server *s, double *used_bytes, double *budget_bytes, int *waiting,
                     int *running) {
  if (!s) return IMA_ERR_ARG;
  if (used_bytes) *used_bytes = (double)s->used;
  if (budget_bytes) *budget_bytes = (double)s->budget;
  if (waiting) *waiting = (int)s->queue.size();
  if (running) *running = (int)s->active.size();
  return IMA_OK;
}

}  // extern "C"
