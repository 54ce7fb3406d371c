// ctx.hip -- context, device-resident data and the extern "C" entry points.
#include <math.h>
#include <cmath>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ima_internal.h"
#include "ctx_impl.h"

using namespace ima;

static void mc3_release(ima_ctx *c);
static void mc3_drop_graphs(ima_ctx *c);

namespace {

int fail(ima_ctx *c, int rc, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return rc;
}

#define HIPCHK(c, call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail((c), IMA_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_),  \
                  __FILE__, __LINE__);                                               \
  } while (0)

// Grow b to hold n elements; with keep, the old contents move over on
// stream st, after the work already queued there.
template <typename T>
hipError_t reserve(DevBuf<T> &b, size_t n, bool keep = false, hipStream_t st = nullptr) {
  if (n <= b.n && b.p) return hipSuccess;
  size_t cap = std::max<size_t>(n, b.n * 2);
  if (cap == 0) cap = 1;
  T *p = nullptr;
  hipError_t e = hipMalloc(&p, cap * sizeof(T));
  if (e != hipSuccess) return e;
  if (keep && b.p && b.n) {
    e = hipMemcpyAsync(p, b.p, b.n * sizeof(T), hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  if (b.p) hipFree(b.p);
  b.p = p;
  b.n = cap;
  return hipSuccess;
}

template <typename T>
void release(DevBuf<T> &b) {
  if (b.p) hipFree(b.p);
  b.p = nullptr;
  b.n = 0;
}

// A copy ordered after the context's queued work that returns once done.
// Every host-synchronous copy of the library goes through the context's
// own stream: nothing runs on the null stream, so one context's copies
// never wait for another context's kernels.
hipError_t copy_now(ima_ctx *c, void *dst, const void *src, size_t bytes, hipMemcpyKind k) {
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, k, c->stream);
  return e != hipSuccess ? e : hipStreamSynchronize(c->stream);
}

int upload_loci(ima_ctx *c) {
  if (!c->loci_dirty) return IMA_OK;
  HIPCHK(c, reserve(c->d_loci, c->loci.size() + 1));
  HIPCHK(c, reserve(c->d_seq, c->seq.size() + 1));
  HIPCHK(c, reserve(c->d_mult, c->mult.size() + 1));
  HIPCHK(c, reserve(c->d_mask, c->mask.size() + 1));
  if (!c->loci.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_loci.p, c->loci.data(), c->loci.size() * sizeof(LocusMeta),
                             hipMemcpyHostToDevice, c->stream));
  if (!c->seq.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_seq.p, c->seq.data(), c->seq.size(), hipMemcpyHostToDevice,
                             c->stream));
  if (!c->mult.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_mult.p, c->mult.data(), c->mult.size() * 4,
                             hipMemcpyHostToDevice, c->stream));
  if (!c->mask.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_mask.p, c->mask.data(), c->mask.size() * 8,
                             hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->loci_dirty = false;
  return IMA_OK;
}

int upload_slot_meta(ima_ctx *c) {
  if (!c->slots_meta_dirty) return IMA_OK;
  HIPCHK(c, reserve(c->d_slots, c->slots.size() + 1));
  if (!c->slots.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_slots.p, c->slots.data(), c->slots.size() * sizeof(SlotMeta),
                             hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->slots_meta_dirty = false;
  return IMA_OK;
}

// The metadata of slots [first, first + count) after a write: only those
// records go to the device when its copy is otherwise current (a per-call
// writer such as IMa2p's per-locus path changes one slot at a time), else
// the whole table is marked for upload.  Enqueued on the context's stream.
hipError_t upload_slot_meta_range(ima_ctx *c, int first, int count) {
  if (c->slots_meta_dirty || !c->d_slots.p || c->d_slots.n < c->slots.size()) {
    c->slots_meta_dirty = true;
    return hipSuccess;
  }
  return hipMemcpyAsync(c->d_slots.p + first, c->slots.data() + first,
                        (size_t)count * sizeof(SlotMeta), hipMemcpyHostToDevice, c->stream);
}

DeviceView view(ima_ctx *c) {
  DeviceView v;
  v.loci = c->d_loci.p;
  v.slots = c->d_slots.p;
  v.up0 = c->d_up0.p;
  v.up1 = c->d_up1.p;
  v.down = c->d_down.p;
  v.time = c->d_time.p;
  v.A = c->d_A.p;
  v.prog = c->d_prog.p;
  v.seq = c->d_seq.p;
  v.mult = c->d_mult.p;
  v.mask = c->d_mask.p;
  v.pop = c->d_pop.p;
  v.mig_first = c->d_mig_first.p;
  v.mig_count = c->d_mig_count.p;
  v.mig_edge = c->d_mig_edge.p;
  v.mig_to = c->d_mig_to.p;
  v.mig_time = c->d_mig_time.p;
  v.models = c->d_models.p;
  v.mint = c->d_mint.p;
  v.mdbl = c->d_mdbl.p;
  v.logfact = c->d_logfact.p;
  return v;
}

// log(k!) for k < need, by the running sum of log(i) (setlogfact).
void grow_logfact(ima_ctx *c, size_t need) {
  if (c->logfact.empty()) c->logfact.push_back(0.0);
  while (c->logfact.size() < need) {
    const double k = (double)c->logfact.size();
    c->logfact.push_back(c->logfact.back() + log(k));
  }
}

int upload_models(ima_ctx *c) {
  if (c->models_dirty) {
    HIPCHK(c, reserve(c->d_models, c->models.size() + 1));
    HIPCHK(c, reserve(c->d_mint, c->mint.size() + 1));
    HIPCHK(c, reserve(c->d_mdbl, c->mdbl.size() + 1));
    if (!c->models.empty()) {
      HIPCHK(c, hipMemcpyAsync(c->d_models.p, c->models.data(),
                               c->models.size() * sizeof(ModelMeta), hipMemcpyHostToDevice,
                               c->stream));
      HIPCHK(c, hipMemcpyAsync(c->d_mint.p, c->mint.data(), c->mint.size() * 4,
                               hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipMemcpyAsync(c->d_mdbl.p, c->mdbl.data(), c->mdbl.size() * 8,
                               hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->models_dirty = false;
  }
  if (c->logfact.size() > c->logfact_dev) {
    HIPCHK(c, reserve(c->d_logfact, c->logfact.size()));
    HIPCHK(c, copy_now(c, c->d_logfact.p, c->logfact.data(), c->logfact.size() * 8,
                        hipMemcpyHostToDevice));
    c->logfact_dev = c->logfact.size();
  }
  return IMA_OK;
}

// Slot need of the subtree below edge e (0 for a tip): Strahler-style.
int need_of(const int32_t *u0, const int32_t *u1, int e, std::vector<int> &need) {
  if (u0[e] < 0) return need[e] = 0;
  const int a = need_of(u0, u1, u0[e], need), b = need_of(u0, u1, u1[e], need);
  return need[e] = (a == b) ? a + 1 : std::max(a, b);
}

// Postorder stack program: the daughter with the larger need goes first, so
// its result sits below the other's (or is the only one on the stack).
void emit(const int32_t *u0, const int32_t *u1, int e, const std::vector<int> &need,
          std::vector<ProgOp> &out) {
  int a = u0[e], b = u1[e];
  if (need[a] < need[b]) std::swap(a, b);
  const bool ai = u0[a] >= 0, bi = u0[b] >= 0;
  if (ai) emit(u0, u1, a, need, out);
  if (bi) emit(u0, u1, b, need, out);
  ProgOp op;
  op.a = a;
  op.b = b;
  op.type = (ai && bi) ? OP_SS : (ai ? OP_ST : OP_TT);
  op.node = e;
  out.push_back(op);
}

// Check that the arrays of one slot form a binary genealogy rooted at root.
bool valid_tree(int n, const int32_t *u0, const int32_t *u1, const int32_t *dn,
                const double *tm, int root, std::string &why) {
  const int ne = 2 * n - 1;
  if (root < n || root >= ne) return why = "root out of range", false;
  if (dn[root] != -1) return why = "root edge has a parent", false;
  for (int e = 0; e < ne; ++e) {
    if (e < n) {
      if (u0[e] != -1 || u1[e] != -1) return why = "tip with daughters", false;
    } else {
      if (u0[e] < 0 || u0[e] >= ne || u1[e] < 0 || u1[e] >= ne || u0[e] == u1[e])
        return why = "internal edge with bad daughters", false;
      if (dn[u0[e]] != e || dn[u1[e]] != e) return why = "up/down links disagree", false;
    }
    if (e != root && (dn[e] < n || dn[e] >= ne)) return why = "bad parent link", false;
    if (!(tm[e] == tm[e]) || fabs(tm[e]) > 1e300) return why = "non-finite time", false;
  }
  // every edge reachable from the root exactly once
  std::vector<int> stack{root};
  std::vector<char> seen(ne, 0);
  int cnt = 0;
  while (!stack.empty()) {
    const int e = stack.back();
    stack.pop_back();
    if (seen[e]) return why = "cycle", false;
    seen[e] = 1;
    ++cnt;
    if (u0[e] >= 0) {
      stack.push_back(u0[e]);
      stack.push_back(u1[e]);
    }
  }
  if (cnt != ne) return why = "disconnected genealogy", false;
  return true;
}

int timing_begin(ima_ctx *c, int k) {
  if (c->timing.on) HIPCHK(c, hipEventRecord(c->timing.ev0[k], c->stream));
  return IMA_OK;
}
int timing_end(ima_ctx *c, int k) {
  if (c->timing.on) {
    HIPCHK(c, hipEventRecord(c->timing.ev1[k], c->stream));
  }
  return IMA_OK;
}
int timing_collect(ima_ctx *c, int k, int count = 1) {
  if (c->timing.on) {
    HIPCHK(c, hipEventSynchronize(c->timing.ev1[k]));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->timing.ev0[k], c->timing.ev1[k]));
    c->timing.ms[k] += ms;
    c->timing.launches[k] += count;
  }
  return IMA_OK;
}

}  // namespace

extern "C" {

int ima_abi_version(void) { return IMA_ABI_VERSION; }

int ima_device_count(int *count) {
  if (!count) return IMA_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return IMA_OK;
}

int ima_device_memory(int device, double *free_bytes, double *total_bytes) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return IMA_ERR_NODEV;
  if (device < 0 || device >= n) return IMA_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return IMA_ERR_HIP;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return IMA_ERR_HIP;
  if (free_bytes) *free_bytes = (double)fr;
  if (total_bytes) *total_bytes = (double)tot;
  return IMA_OK;
}

// Work area of ima_lmode_jointp_noalloc: the per (point, chunk) partial
// results of jointp.  One copy per device, held until the process ends and
// shared by every context on that device.  Host code never passes this
// name to a kernel: each device's address comes from hipGetSymbolAddress,
// taken once per device when its first context is created.
constexpr size_t kScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// The device address of `scratch` on `device` (which is current), looked up
// on the first call for that device.
static hipError_t scratch_address(int device, void **ptr) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  return hipSuccess;
}

int ima_ctx_create(int device, ima_ctx **out) {
  if (!out) return IMA_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return IMA_ERR_NODEV;
  if (device < 0 || device >= n) return IMA_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return IMA_ERR_HIP;
  std::unique_ptr<ima_ctx> c(new ima_ctx);
  c->device = device;
  if (scratch_address(device, &c->scratch) != hipSuccess) return IMA_ERR_HIP;
  if (hipStreamCreate(&c->stream) != hipSuccess) return IMA_ERR_HIP;
  if (hipStreamCreate(&c->side) != hipSuccess) return IMA_ERR_HIP;
  for (int k = 0; k < kTimingClasses; ++k) {
    if (hipEventCreate(&c->timing.ev0[k]) != hipSuccess) return IMA_ERR_HIP;
    if (hipEventCreate(&c->timing.ev1[k]) != hipSuccess) return IMA_ERR_HIP;
  }
  // Every launch keeps its dynamic LDS within 64 KiB even where the device
  // allows more per workgroup (160 KiB on gfx950): group sizes shrink to fit
  // and larger requests fail with IMA_ERR_SIZE before any launch.
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) ==
          hipSuccess &&
      lds > 0)
    c->max_lds = std::min(lds, kMaxLaunchLds);
  *out = c.release();
  return IMA_OK;
}

void ima_ctx_destroy(ima_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  mc3_release(c);
  release(c->d_loci);
  release(c->d_seq);
  release(c->d_mult);
  release(c->d_mask);
  release(c->d_slots);
  release(c->d_up0);
  release(c->d_up1);
  release(c->d_down);
  release(c->d_A);
  release(c->d_time);
  release(c->d_prog);
  release(c->d_items);
  release(c->d_idx);
  release(c->d_out);
  release(c->d_status);
  release(c->d_aux);
  release(c->d_aux64);
  release(c->d_units);
  release(c->d_red);
  release(c->d_partial);
  release(c->d_lx);
  release(c->d_lwork);
  release(c->d_lout);
  release(c->d_less);
  release(c->d_S);
  release(c->d_beta);
  release(c->d_pairs);
  release(c->d_pop);
  release(c->d_mig_edge);
  release(c->d_mig_to);
  release(c->d_mig_first);
  release(c->d_mig_count);
  release(c->d_mig_time);
  release(c->d_nw_pop);
  release(c->d_nw_mf);
  release(c->d_nw_mc);
  release(c->d_nw_me);
  release(c->d_nw_mto);
  release(c->d_nw_mt);
  release(c->d_nw_nbk);
  release(c->d_nw_st);
  release(c->d_nw_period);
  release(c->d_nw_w);
  release(c->d_nw_newt);
  release(c->d_mv_off);
  release(c->d_mv_cnt);
  release(c->d_models);
  release(c->d_mint);
  release(c->d_mdbl);
  release(c->d_logfact);
  release(c->d_rec);
  release(c->d_tv);
  release(c->d_probg);
  release(c->d_qint);
  release(c->d_mintg);
  release(c->d_pslots);
  release(c->d_prow);
  release(c->d_goff);
  release(c->d_pstatus);
  release(c->d_swmoves);
  for (auto &t : c->tables)
    if (t.rows) hipFree(t.rows);
  for (int k = 0; k < kTimingClasses; ++k) {
    hipEventDestroy(c->timing.ev0[k]);
    hipEventDestroy(c->timing.ev1[k]);
  }
  hipStreamDestroy(c->stream);
  if (c->side) hipStreamDestroy(c->side);
  delete c;
}

const char *ima_ctx_last_error(const ima_ctx *c) { return c ? c->err.c_str() : "no context"; }

// ------------------------------------------------------------------ loci --

int ima_locus_add_hky(ima_ctx *c, int n, int S, const int8_t *seq, const int32_t *mult,
                      int totsites, const double pi[4], int *locus) {
  if (!c || !seq || !mult || !pi || !locus) return IMA_ERR_ARG;
  if (n < 2 || n > kMaxGenes) return fail(c, IMA_ERR_SIZE, "HKY locus n=%d outside [2,%d]", n, kMaxGenes);
  if (S < 1 || totsites < 1) return fail(c, IMA_ERR_ARG, "HKY locus with no sites");
  for (size_t i = 0; i < (size_t)n * S; ++i)
    if (seq[i] < 0 || seq[i] > 3) return fail(c, IMA_ERR_ARG, "HKY base outside 0..3");
  LocusMeta m{};
  m.pmodel = -1;
  m.hval = 1.0;
  m.model = IMA_MODEL_HKY;
  m.n = n;
  m.S = S;
  m.totsites = totsites;
  // rows of 64-pattern blocks, 64-byte aligned, zero padded: one block of
  // one tip is 16 aligned dwords
  m.Spad = (S + 2 * kWave - 1) / (2 * kWave) * (2 * kWave);  // blocks of up to 128 patterns
  c->seq.resize((c->seq.size() + 63) / 64 * 64, 0);
  m.seq_off = (int64_t)c->seq.size();
  m.mult_off = (int64_t)c->mult.size();
  for (int i = 0; i < 4; ++i) m.pi[i] = pi[i];
  c->seq.resize(c->seq.size() + (size_t)n * m.Spad, 0);
  for (int t = 0; t < n; ++t)
    std::copy(seq + (size_t)t * S, seq + (size_t)(t + 1) * S,
              c->seq.begin() + m.seq_off + (size_t)t * m.Spad);
  c->mult.insert(c->mult.end(), mult, mult + S);
  *locus = (int)c->loci.size();
  c->loci.push_back(m);
  c->scratch_slot.push_back(-1);
  c->loci_dirty = true;
  return IMA_OK;
}

int ima_locus_add_is(ima_ctx *c, int n, int S, const int8_t *seq, int *locus) {
  if (!c || (!seq && S > 0) || !locus) return IMA_ERR_ARG;
  if (n < 2 || n > kMaxGenes) return fail(c, IMA_ERR_SIZE, "IS locus n=%d outside [2,%d]", n, kMaxGenes);
  if (S < 0) return IMA_ERR_ARG;
  LocusMeta m{};
  m.pmodel = -1;
  m.hval = 1.0;
  m.model = IMA_MODEL_IS;
  m.n = n;
  m.S = S;
  m.W = (n + 63) / 64;
  m.mask_off = (int64_t)c->mask.size();
  for (int s = 0; s < S; ++s) {
    uint64_t w[4] = {0, 0, 0, 0};
    for (int t = 0; t < n; ++t) {
      const int8_t v = seq[(size_t)t * S + s];
      if (v != 0 && v != 1) return fail(c, IMA_ERR_ARG, "IS state outside 0/1");
      if (v) w[t >> 6] |= 1ull << (t & 63);
    }
    for (int k = 0; k < m.W; ++k) c->mask.push_back(w[k]);
  }
  *locus = (int)c->loci.size();
  c->loci.push_back(m);
  c->scratch_slot.push_back(-1);
  c->loci_dirty = true;
  return IMA_OK;
}

int ima_locus_add_sw(ima_ctx *c, int n, int *locus) {
  if (!c || !locus) return IMA_ERR_ARG;
  if (n < 2 || n > kMaxGenes) return fail(c, IMA_ERR_SIZE, "SW locus n=%d outside [2,%d]", n, kMaxGenes);
  LocusMeta m{};
  m.pmodel = -1;
  m.hval = 1.0;
  m.model = IMA_MODEL_SW;
  m.n = n;
  *locus = (int)c->loci.size();
  c->loci.push_back(m);
  c->scratch_slot.push_back(-1);
  c->loci_dirty = true;
  return IMA_OK;
}

int ima_locus_set_sumlogk(ima_ctx *c, int locus, double sumlogk) {
  if (!c || locus < 0 || locus >= (int)c->loci.size()) return IMA_ERR_ARG;
  c->loci[locus].sumlogk = sumlogk;
  c->loci_dirty = true;
  return IMA_OK;
}

// ----------------------------------------------------------------- slots --

int ima_slots_alloc(ima_ctx *c, int count, const int32_t *locus_of, int *first) {
  if (!c || count < 0 || (!locus_of && count > 0) || !first) return IMA_ERR_ARG;
  hipSetDevice(c->device);
  *first = (int)c->slots.size();
  for (int i = 0; i < count; ++i) {
    const int l = locus_of[i];
    if (l < 0 || l >= (int)c->loci.size()) return fail(c, IMA_ERR_ARG, "slot locus %d out of range", l);
    SlotMeta s{};
    s.locus = l;
    s.n = c->loci[l].n;
    s.root = -1;
    s.nslots = 1;
    s.node_off = (int64_t)c->nodes;
    s.prog_off = (int64_t)c->progs;
    c->nodes += 2 * (size_t)s.n - 1;
    c->progs += (size_t)s.n - 1;
    c->slots.push_back(s);
  }
  c->up0.resize(c->nodes, -1);
  c->pop.resize(c->nodes, 0);
  c->mig_first.resize(c->nodes, 0);
  c->mig_count.resize(c->nodes, 0);
  c->up1.resize(c->nodes, -1);
  c->down.resize(c->nodes, -1);
  c->A.resize(c->nodes, 0);
  c->time.resize(c->nodes, 0.0);
  c->prog.resize(c->progs);
  HIPCHK(c, reserve(c->d_up0, c->nodes, true, c->stream));
  HIPCHK(c, reserve(c->d_pop, c->nodes, true, c->stream));
  HIPCHK(c, reserve(c->d_up1, c->nodes, true, c->stream));
  HIPCHK(c, reserve(c->d_down, c->nodes, true, c->stream));
  HIPCHK(c, reserve(c->d_A, c->nodes, true, c->stream));
  HIPCHK(c, reserve(c->d_time, c->nodes, true, c->stream));
  HIPCHK(c, reserve(c->d_prog, c->progs + 1, true, c->stream));
  c->slots_meta_dirty = true;
  return IMA_OK;
}

int ima_slots_write(ima_ctx *c, int first, int count, const int32_t *up0, const int32_t *up1,
                    const int32_t *down, const double *time, const int32_t *root,
                    const int32_t *A) {
  if (!c || !up0 || !up1 || !down || !time || !root) return IMA_ERR_ARG;
  if (first < 0 || count < 0 || first + count > (int)c->slots.size())
    return fail(c, IMA_ERR_ARG, "slot range [%d,%d) out of range", first, first + count);
  if (count == 0) return IMA_OK;
  hipSetDevice(c->device);
  size_t src = 0;
  std::vector<int> need;
  std::vector<ProgOp> ops;
  for (int i = 0; i < count; ++i) {
    SlotMeta &s = c->slots[first + i];
    const int n = s.n, ne = 2 * n - 1;
    std::string why;
    if (!valid_tree(n, up0 + src, up1 + src, down + src, time + src, root[i], why))
      return fail(c, IMA_ERR_ARG, "slot %d: %s", first + i, why.c_str());
    need.assign(ne, 0);
    const int r = need_of(up0 + src, up1 + src, root[i], need);
    if (c->mc3.live && first + i >= c->mc3.first &&
        first + i < c->mc3.first + c->mc3.C * c->mc3.L && std::max(1, r) > c->mc3.hmax_slots)
      return fail(c, IMA_ERR_STATE, "slot %d: program depth %d exceeds the running sweep's %d",
                  first + i, r, c->mc3.hmax_slots);
    std::copy(up0 + src, up0 + src + ne, c->up0.begin() + s.node_off);
    std::copy(up1 + src, up1 + src + ne, c->up1.begin() + s.node_off);
    std::copy(down + src, down + src + ne, c->down.begin() + s.node_off);
    std::copy(time + src, time + src + ne, c->time.begin() + s.node_off);
    if (A) std::copy(A + src, A + src + ne, c->A.begin() + s.node_off);
    s.root = root[i];
    ops.clear();
    emit(up0 + src, up1 + src, root[i], need, ops);
    std::copy(ops.begin(), ops.end(), c->prog.begin() + s.prog_off);
    s.nslots = std::max(1, r);
    src += ne;
  }
  const SlotMeta &s0 = c->slots[first];
  const SlotMeta &s1 = c->slots[first + count - 1];
  const size_t nb = s1.node_off + 2 * (size_t)s1.n - 1 - s0.node_off;
  const size_t pb = s1.prog_off + (size_t)s1.n - 1 - s0.prog_off;
  HIPCHK(c, hipMemcpyAsync(c->d_up0.p + s0.node_off, c->up0.data() + s0.node_off, nb * 4,
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_up1.p + s0.node_off, c->up1.data() + s0.node_off, nb * 4,
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_down.p + s0.node_off, c->down.data() + s0.node_off, nb * 4,
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_time.p + s0.node_off, c->time.data() + s0.node_off, nb * 8,
                           hipMemcpyHostToDevice, c->stream));
  if (A)
    HIPCHK(c, hipMemcpyAsync(c->d_A.p + s0.node_off, c->A.data() + s0.node_off, nb * 4,
                             hipMemcpyHostToDevice, c->stream));
  if (pb)
    HIPCHK(c, hipMemcpyAsync(c->d_prog.p + s0.prog_off, c->prog.data() + s0.prog_off,
                             pb * sizeof(ProgOp), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, upload_slot_meta_range(c, first, count));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return IMA_OK;
}

int ima_slots_read(ima_ctx *c, int first, int count, int32_t *up0, int32_t *up1,
                   int32_t *down, double *time, int32_t *root, int32_t *A) {
  if (!c || first < 0 || count < 0 || first + count > (int)c->slots.size()) return IMA_ERR_ARG;
  if (count == 0) return IMA_OK;
  const SlotMeta &s0 = c->slots[first];
  const SlotMeta &s1 = c->slots[first + count - 1];
  const size_t nb = s1.node_off + 2 * (size_t)s1.n - 1 - s0.node_off;
  if (up0) HIPCHK(c, copy_now(c, up0, c->d_up0.p + s0.node_off, nb * 4, hipMemcpyDeviceToHost));
  if (up1) HIPCHK(c, copy_now(c, up1, c->d_up1.p + s0.node_off, nb * 4, hipMemcpyDeviceToHost));
  if (down) HIPCHK(c, copy_now(c, down, c->d_down.p + s0.node_off, nb * 4, hipMemcpyDeviceToHost));
  if (time) HIPCHK(c, copy_now(c, time, c->d_time.p + s0.node_off, nb * 8, hipMemcpyDeviceToHost));
  if (A) HIPCHK(c, copy_now(c, A, c->d_A.p + s0.node_off, nb * 4, hipMemcpyDeviceToHost));
  if (root)
    for (int i = 0; i < count; ++i) root[i] = c->slots[first + i].root;
  return IMA_OK;
}

int ima_slots_read_pops(ima_ctx *c, int first, int count, int32_t *pop) {
  if (!c || !pop || first < 0 || count < 0 || first + count > (int)c->slots.size())
    return IMA_ERR_ARG;
  if (count == 0) return IMA_OK;
  if (!c->d_pop.p) return fail(c, IMA_ERR_STATE, "no population labels were written");
  const SlotMeta &s0 = c->slots[first];
  const SlotMeta &s1 = c->slots[first + count - 1];
  const size_t nb = s1.node_off + 2 * (size_t)s1.n - 1 - s0.node_off;
  HIPCHK(c, copy_now(c, pop, c->d_pop.p + s0.node_off, nb * 4, hipMemcpyDeviceToHost));
  return IMA_OK;
}

int ima_slots_count(ima_ctx *c, int *count) {
  if (!c || !count) return IMA_ERR_ARG;
  *count = (int)c->slots.size();
  return IMA_OK;
}

int ima_slot_info(ima_ctx *c, int slot, int *locus, int *n) {
  if (!c || slot < 0 || slot >= (int)c->slots.size()) return IMA_ERR_ARG;
  if (locus) *locus = c->slots[slot].locus;
  if (n) *n = c->slots[slot].n;
  return IMA_OK;
}

// --------------------------------------------------------------- scoring --

// Orders HKY work units for the L2 of each XCD.  The dispatcher deals a
// launch's workgroups round-robin to the 8 XCDs (workgroup i to XCD i mod
// 8), so placing every unit of a locus at positions of one residue class
// mod 8 keeps the locus' tip states, and the genealogy of an item that
// spans several pattern blocks, in one XCD's L2: each is read from HBM once
// instead of once per XCD.  Loci are dealt to the 8 classes greedily,
// largest unit count first, to the class with the fewest units; the
// classes then fill positions in rounds, and once one runs out the rest
// continue in round-robin order.  Only the order changes: a unit's partial
// index is part of the unit.
static std::vector<int4> xcd_order(const std::vector<int4> &units,
                                   const std::vector<int> &locus_of_unit) {
  constexpr int kXcds = 8;
  std::map<int, std::vector<int>> by_locus;
  for (size_t i = 0; i < units.size(); ++i) by_locus[locus_of_unit[i]].push_back((int)i);
  std::vector<std::pair<int, int>> order;  // (count, locus)
  for (const auto &kv : by_locus) order.push_back({(int)kv.second.size(), kv.first});
  std::stable_sort(order.begin(), order.end(),
                   [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
                     return x.first > y.first;
                   });
  std::vector<std::vector<int>> bins(kXcds);
  for (const auto &cl : order) {
    int b = 0;
    for (int x = 1; x < kXcds; ++x)
      if (bins[x].size() < bins[b].size()) b = x;
    const std::vector<int> &u = by_locus[cl.second];
    bins[b].insert(bins[b].end(), u.begin(), u.end());
  }
  std::vector<int4> out;
  out.reserve(units.size());
  for (size_t k = 0; out.size() < units.size(); ++k)
    for (int x = 0; x < kXcds; ++x)
      if (k < bins[x].size()) out.push_back(units[bins[x][k]]);
  return out;
}

// Patterns per lane of HKY site blocks (1: 64, 2: 128) for HKY items in
// slots[0..n): the context's setting, else 64.  Two patterns per lane share
// the program, coefficient and tip reads, but each wavefront's program
// takes longer and holds twice the LDS; on the uneven batch (150 HKY loci
// of 200-3000 bp) 128-pattern blocks made the sweep 19% slower, so 64 is
// the default and ima_mc3_tune times both on the chains' sweeps.
static int hky_auto_ppl(const ima_ctx *c, const std::vector<int> &slots) {
  (void)slots;
  return c->site_block == 128 ? 2 : 1;
}

static int score_impl(ima_ctx *c, const ima_item *items, int nitems, int precision,
                      double *loglik, int32_t *status, int reps = 1) {
  if (!c || (!items && nitems > 0) || !loglik || nitems < 0 || reps < 1) return IMA_ERR_ARG;
  if (precision != IMA_PREC_FP64 && precision != IMA_PREC_FP32) return IMA_ERR_ARG;
  if (nitems == 0) return IMA_OK;
  hipSetDevice(c->device);
  int rc;
  if ((rc = upload_loci(c)) != IMA_OK) return rc;
  if ((rc = upload_slot_meta(c)) != IMA_OK) return rc;
  std::vector<int32_t> idx_h(nitems), idx_i, idx_s;
  std::vector<int32_t> hky;
  int hmax_n = 2, hmax_slots = 1, imax_n = 2, imax_w = 1;
  hky.reserve(nitems);
  for (int i = 0; i < nitems; ++i) {
    const int sl = items[i].slot;
    if (sl < 0 || sl >= (int)c->slots.size()) return fail(c, IMA_ERR_ARG, "item %d: slot %d out of range", i, sl);
    const SlotMeta &s = c->slots[sl];
    if (s.root < 0) return fail(c, IMA_ERR_STATE, "item %d: slot %d holds no genealogy", i, sl);
    const LocusMeta &m = c->loci[s.locus];
    if (m.model == IMA_MODEL_HKY) {
      if (!(items[i].u > 0) || !(items[i].kappa > 0))
        return fail(c, IMA_ERR_ARG, "item %d: HKY needs u > 0 and kappa > 0", i);
      hky.push_back(i);
      hmax_n = std::max(hmax_n, s.n);
      hmax_slots = std::max(hmax_slots, s.nslots);
    } else if (m.model == IMA_MODEL_IS) {
      if (!(items[i].u > 0)) return fail(c, IMA_ERR_ARG, "item %d: IS needs u > 0", i);
      idx_i.push_back(i);
      imax_n = std::max(imax_n, s.n);
      imax_w = std::max(imax_w, m.W);
    } else {
      idx_s.push_back(i);
    }
  }
  std::vector<int> hslots;
  for (int i : hky) hslots.push_back(items[i].slot);
  int ppl = hky_auto_ppl(c, hslots);
  if (ppl == 2 && c->site_block != 128 &&
      hky_lds_bytes(precision, hmax_n, hmax_slots, 2) > (size_t)c->max_lds)
    ppl = 1;
  const size_t hlds = hky_lds_bytes(precision, hmax_n, hmax_slots, ppl);
  if (!hky.empty() && hmax_slots > 8)
    return fail(c, IMA_ERR_SIZE, "a genealogy needs %d partial slots (limit 8)", hmax_slots);
  if (!hky.empty() && hlds > (size_t)c->max_lds)
    return fail(c, IMA_ERR_SIZE, "HKY batch needs %zu B of LDS (limit %d)", hlds, c->max_lds);
  HIPCHK(c, reserve(c->d_items, nitems));
  HIPCHK(c, reserve(c->d_idx, nitems));
  HIPCHK(c, reserve(c->d_out, nitems));
  HIPCHK(c, reserve(c->d_status, nitems));
  size_t k = 0;
  for (int i : hky) idx_h[k++] = i;
  for (int i : idx_i) idx_h[k++] = i;
  for (int i : idx_s) idx_h[k++] = i;
  HIPCHK(c, hipMemcpyAsync(c->d_items.p, items, nitems * sizeof(ima_item), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_idx.p, idx_h.data(), nitems * 4, hipMemcpyHostToDevice,
                           c->stream));
  const DeviceView v = view(c);
  const int kh = precision == IMA_PREC_FP32 ? K_HKY32 : K_HKY64;
  HkyPlan plan;
  if (!hky.empty()) {
    std::vector<int4> units, red;
    int nblk = 0;
    for (int i : hky)
      hky_plan_add(units, red, i, c->loci[c->slots[items[i].slot].locus].S, &nblk, ppl);
    {
      std::vector<int> lu(units.size());
      for (size_t k = 0; k < units.size(); ++k) lu[k] = c->slots[items[units[k].x].slot].locus;
      units = xcd_order(units, lu);
    }
    HIPCHK(c, reserve(c->d_units, units.size()));
    HIPCHK(c, reserve(c->d_red, red.size() + 1));
    HIPCHK(c, reserve(c->d_partial, (size_t)nblk + 1));
    HIPCHK(c, hipMemcpyAsync(c->d_units.p, units.data(), units.size() * sizeof(int4),
                             hipMemcpyHostToDevice, c->stream));
    if (!red.empty())
      HIPCHK(c, hipMemcpyAsync(c->d_red.p, red.data(), red.size() * sizeof(int4),
                               hipMemcpyHostToDevice, c->stream));
    plan.units = c->d_units.p;
    plan.nunits = (int)units.size();
    plan.ppl = ppl;
    plan.red = c->d_red.p;
    plan.nred = (int)red.size();
    plan.partial = c->d_partial.p;
  }
  if (!hky.empty()) {
    if ((rc = timing_begin(c, kh)) != IMA_OK) return rc;
    // repeated passes write the same block partials; only the last merges them
    HkyPlan pass = plan;
    pass.nred = 0;
    for (int r = 0; r < reps; ++r)
      HIPCHK(c, launch_hky(v, c->d_items.p, r == reps - 1 ? plan : pass, precision, hmax_n,
                           hmax_slots, c->d_out.p, c->d_status.p, c->stream));
    if ((rc = timing_end(c, kh)) != IMA_OK) return rc;
  }
  if (!idx_i.empty()) {
    if ((rc = timing_begin(c, K_IS)) != IMA_OK) return rc;
    for (int r = 0; r < reps; ++r)
      HIPCHK(c, launch_is(v, c->d_items.p, c->d_idx.p + hky.size(), (int)idx_i.size(), imax_n,
                          imax_w, c->d_out.p, c->d_status.p, nullptr, nullptr, c->stream));
    if ((rc = timing_end(c, K_IS)) != IMA_OK) return rc;
  }
  if (!idx_s.empty()) {
    if ((rc = timing_begin(c, K_SW)) != IMA_OK) return rc;
    for (int r = 0; r < reps; ++r)
      HIPCHK(c, launch_sw(v, c->d_items.p, c->d_idx.p + hky.size() + idx_i.size(),
                          (int)idx_s.size(), c->d_out.p, c->d_status.p, nullptr, c->stream));
    if ((rc = timing_end(c, K_SW)) != IMA_OK) return rc;
  }
  HIPCHK(c, hipMemcpyAsync(loglik, c->d_out.p, nitems * 8, hipMemcpyDeviceToHost, c->stream));
  if (status)
    HIPCHK(c, hipMemcpyAsync(status, c->d_status.p, nitems * 4, hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!hky.empty() && (rc = timing_collect(c, kh, reps)) != IMA_OK) return rc;
  if (!idx_i.empty() && (rc = timing_collect(c, K_IS, reps)) != IMA_OK) return rc;
  if (!idx_s.empty() && (rc = timing_collect(c, K_SW, reps)) != IMA_OK) return rc;
  return IMA_OK;
}

int ima_score(ima_ctx *c, const ima_item *items, int nitems, int precision, double *loglik,
              int32_t *status) {
  return score_impl(c, items, nitems, precision, loglik, status);
}

int ima_score_repeat(ima_ctx *c, const ima_item *items, int nitems, int precision, int reps,
                     double *loglik, int32_t *status) {
  return score_impl(c, items, nitems, precision, loglik, status, reps);
}

int ima_is_site_nodes(ima_ctx *c, int slot, int32_t *site_node) {
  if (!c || !site_node || slot < 0 || slot >= (int)c->slots.size()) return IMA_ERR_ARG;
  const SlotMeta &s = c->slots[slot];
  const LocusMeta &m = c->loci[s.locus];
  if (m.model != IMA_MODEL_IS) return fail(c, IMA_ERR_ARG, "slot %d is not an IS locus", slot);
  if (s.root < 0) return fail(c, IMA_ERR_STATE, "slot %d holds no genealogy", slot);
  hipSetDevice(c->device);
  int rc;
  if ((rc = upload_loci(c)) != IMA_OK) return rc;
  if ((rc = upload_slot_meta(c)) != IMA_OK) return rc;
  ima_item it{slot, 0, 1.0, 1.0, 1.0};
  const int32_t zero = 0;
  HIPCHK(c, reserve(c->d_items, 1));
  HIPCHK(c, reserve(c->d_idx, 1));
  HIPCHK(c, reserve(c->d_out, 1));
  HIPCHK(c, reserve(c->d_status, 1));
  HIPCHK(c, reserve(c->d_aux, (size_t)std::max(m.S, 1)));
  HIPCHK(c, hipMemcpyAsync(c->d_items.p, &it, sizeof it, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_idx.p, &zero, 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, launch_is(view(c), c->d_items.p, c->d_idx.p, 1, s.n, m.W, c->d_out.p,
                      c->d_status.p, c->d_aux.p, nullptr, c->stream));
  if (m.S > 0)
    HIPCHK(c, hipMemcpyAsync(site_node, c->d_aux.p, (size_t)m.S * 4, hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return IMA_OK;
}

int ima_locus_sumlogk_from_slot(ima_ctx *c, int locus, int slot, double *sumlogk) {
  if (!c || !sumlogk || locus < 0 || locus >= (int)c->loci.size() || slot < 0 ||
      slot >= (int)c->slots.size())
    return IMA_ERR_ARG;
  const SlotMeta &s = c->slots[slot];
  if (s.locus != locus) return fail(c, IMA_ERR_ARG, "slot %d belongs to locus %d", slot, s.locus);
  const LocusMeta &m = c->loci[locus];
  if (m.model != IMA_MODEL_IS) return fail(c, IMA_ERR_ARG, "locus %d is not IS", locus);
  if (s.root < 0) return fail(c, IMA_ERR_STATE, "slot %d holds no genealogy", slot);
  hipSetDevice(c->device);
  int rc;
  if ((rc = upload_loci(c)) != IMA_OK) return rc;
  if ((rc = upload_slot_meta(c)) != IMA_OK) return rc;
  const int ne = 2 * s.n - 1;
  ima_item it{slot, 0, 1.0, 1.0, 1.0};
  const int32_t zero = 0;
  HIPCHK(c, reserve(c->d_items, 1));
  HIPCHK(c, reserve(c->d_idx, 1));
  HIPCHK(c, reserve(c->d_out, 1));
  HIPCHK(c, reserve(c->d_status, 1));
  HIPCHK(c, reserve(c->d_aux, (size_t)ne));
  HIPCHK(c, hipMemcpyAsync(c->d_items.p, &it, sizeof it, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_idx.p, &zero, 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, launch_is(view(c), c->d_items.p, c->d_idx.p, 1, s.n, m.W, c->d_out.p,
                      c->d_status.p, nullptr, c->d_aux.p, c->stream));
  std::vector<int32_t> cnt(ne);
  int32_t st = 0;
  HIPCHK(c, hipMemcpyAsync(cnt.data(), c->d_aux.p, (size_t)ne * 4, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(&st, c->d_status.p, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // calc_prob_data.cpp:654-658, 705-717: an incompatible genealogy leaves
  // the constant at zero; otherwise sum over internal nodes of log(k!).
  double sum = 0.0;
  if (st == IMA_STATUS_OK) {
    for (int i = s.n; i < ne; ++i) {
      double fact = 1.0;
      for (int j = 1; j <= cnt[i]; ++j) fact *= (double)j;
      sum += log(fact);
    }
  }
  c->loci[locus].sumlogk = sum;
  c->loci_dirty = true;
  *sumlogk = sum;
  if (st == IMA_STATUS_IS_NOSITE)
    return fail(c, IMA_ERR_ARG, "a site has no 0/1 node on this genealogy");
  return IMA_OK;
}

int ima_sw_dlike(ima_ctx *c, const ima_item *item, double *dlikeA) {
  if (!c || !item || !dlikeA || item->slot < 0 || item->slot >= (int)c->slots.size())
    return IMA_ERR_ARG;
  const SlotMeta &s = c->slots[item->slot];
  if (c->loci[s.locus].model != IMA_MODEL_SW) return fail(c, IMA_ERR_ARG, "not an SW locus");
  if (s.root < 0) return fail(c, IMA_ERR_STATE, "slot holds no genealogy");
  hipSetDevice(c->device);
  int rc;
  if ((rc = upload_loci(c)) != IMA_OK) return rc;
  if ((rc = upload_slot_meta(c)) != IMA_OK) return rc;
  const int ne = 2 * s.n - 1;
  const int32_t zero = 0;
  HIPCHK(c, reserve(c->d_items, 1));
  HIPCHK(c, reserve(c->d_idx, 1));
  HIPCHK(c, reserve(c->d_out, 1));
  HIPCHK(c, reserve(c->d_status, 1));
  HIPCHK(c, reserve(c->d_aux64, (size_t)ne));
  HIPCHK(c, hipMemcpyAsync(c->d_items.p, item, sizeof *item, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_idx.p, &zero, 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, launch_sw(view(c), c->d_items.p, c->d_idx.p, 1, c->d_out.p, c->d_status.p,
                      c->d_aux64.p, c->stream));
  HIPCHK(c, hipMemcpyAsync(dlikeA, c->d_aux64.p, (size_t)ne * 8, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return IMA_OK;
}

int ima_sw_update_A(ima_ctx *c, const ima_sw_move *moves, int nmoves, int calc, double *dlike,
                    double *Aterm, int32_t *newA, double *edge_dlikeA) {
  if (!c || nmoves < 0 || (nmoves > 0 && !moves)) return IMA_ERR_ARG;
  if (nmoves == 0) return IMA_OK;
  std::vector<int32_t> seen;
  for (int i = 0; i < nmoves; ++i) {
    const ima_sw_move &m = moves[i];
    if (m.slot < 0 || m.slot >= (int)c->slots.size())
      return fail(c, IMA_ERR_ARG, "move %d: slot %d out of range", i, m.slot);
    const SlotMeta &s = c->slots[m.slot];
    if (c->loci[s.locus].model != IMA_MODEL_SW)
      return fail(c, IMA_ERR_ARG, "move %d: slot %d is not a stepwise locus", i, m.slot);
    if (s.root < 0) return fail(c, IMA_ERR_STATE, "move %d: slot holds no genealogy", i);
    const int ne = 2 * s.n - 1;
    const int32_t ed[4] = {m.edge, m.downedge, m.sisedge, m.newsisedge};
    for (int k = 0; k < 4; ++k)
      if (ed[k] < 0 || ed[k] >= ne) return fail(c, IMA_ERR_ARG, "move %d: edge out of range", i);
    if (m.edge == m.downedge || m.newsisedge == m.downedge || m.edge == m.newsisedge ||
        m.sisedge == m.downedge || m.sisedge == m.edge)
      return fail(c, IMA_ERR_ARG, "move %d: edges must be distinct", i);
    if (!(m.uni > 0.0 && m.uni <= 1.0) || m.minA > m.maxA || !(m.u > 0.0))
      return fail(c, IMA_ERR_ARG, "move %d: uni in (0, 1], minA <= maxA, u > 0", i);
    seen.push_back(m.slot);
  }
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
    return fail(c, IMA_ERR_ARG, "two moves on one slot");
  hipSetDevice(c->device);
  int rc;
  if ((rc = upload_loci(c)) != IMA_OK) return rc;
  if ((rc = upload_slot_meta(c)) != IMA_OK) return rc;
  HIPCHK(c, reserve(c->d_swmoves, (size_t)nmoves));
  HIPCHK(c, reserve(c->d_aux64, (size_t)nmoves * 6));
  HIPCHK(c, reserve(c->d_aux, (size_t)nmoves * 2));
  HIPCHK(c, hipMemcpyAsync(c->d_swmoves.p, moves, (size_t)nmoves * sizeof(ima_sw_move),
                           hipMemcpyHostToDevice, c->stream));
  // pass 0: structure of every move; nothing is written unless all pass
  HIPCHK(c, launch_sw_update(view(c), c->d_A.p, c->d_swmoves.p, nmoves, 0, calc, c->d_aux64.p,
                             c->d_aux.p, c->stream));
  std::vector<int32_t> iout((size_t)nmoves * 2);
  HIPCHK(c, copy_now(c, iout.data(), c->d_aux.p, iout.size() * 4, hipMemcpyDeviceToHost));
  for (int i = 0; i < nmoves; ++i)
    if (iout[2 * i + 1])
      return fail(c, IMA_ERR_ARG, "move %d: edge %d and newsisedge %d are not under downedge %d",
                  i, moves[i].edge, moves[i].newsisedge, moves[i].downedge);
  HIPCHK(c, launch_sw_update(view(c), c->d_A.p, c->d_swmoves.p, nmoves, 1, calc, c->d_aux64.p,
                             c->d_aux.p, c->stream));
  std::vector<double> out((size_t)nmoves * 6);
  HIPCHK(c, hipMemcpyAsync(out.data(), c->d_aux64.p, out.size() * 8, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, copy_now(c, iout.data(), c->d_aux.p, iout.size() * 4, hipMemcpyDeviceToHost));
  for (int i = 0; i < nmoves; ++i) {
    if (dlike) dlike[i] = out[6 * i];
    if (Aterm) Aterm[i] = out[6 * i + 1];
    if (newA) newA[i] = iout[2 * i];
    if (edge_dlikeA)
      for (int k = 0; k < 4; ++k) edge_dlikeA[4 * i + k] = out[6 * i + 2 + k];
  }
  return IMA_OK;
}

int ima_likelihood_hky(ima_ctx *c, int locus, const int32_t *up0, const int32_t *up1,
                       const int32_t *down, const double *time, int root, double u,
                       double kappa, int e1, int e2, int e3, int e4, double *loglik) {
  (void)e1;
  (void)e2;
  (void)e3;
  (void)e4;
  if (!c || !loglik || locus < 0 || locus >= (int)c->loci.size()) return IMA_ERR_ARG;
  if (c->loci[locus].model != IMA_MODEL_HKY) return fail(c, IMA_ERR_ARG, "locus %d is not HKY", locus);
  int rc;
  if (c->scratch_slot[locus] < 0) {
    int first = 0;
    const int32_t l = locus;
    if ((rc = ima_slots_alloc(c, 1, &l, &first)) != IMA_OK) return rc;
    c->scratch_slot[locus] = first;
  }
  const int slot = c->scratch_slot[locus];
  const int32_t r = root;
  if ((rc = ima_slots_write(c, slot, 1, up0, up1, down, time, &r, nullptr)) != IMA_OK) return rc;
  ima_item it{slot, 0, u, kappa, 1.0};
  return score_impl(c, &it, 1, IMA_PREC_FP64, loglik, nullptr);
}

// ------------------------------------------------------------------ swap --

int ima_swap(ima_ctx *c, int nchains, const double *S, double *beta, int swaptries,
             uint64_t seed, uint64_t counter, int *swaps_out, int32_t *pairs_out) {
  if (!c || !S || !beta || nchains < 2 || swaptries < 0) return IMA_ERR_ARG;
  hipSetDevice(c->device);
  HIPCHK(c, reserve(c->d_S, nchains));
  HIPCHK(c, reserve(c->d_beta, nchains));
  HIPCHK(c, reserve(c->d_pairs, 3 * (size_t)swaptries + 1));
  HIPCHK(c, reserve(c->d_aux, 1));
  HIPCHK(c, hipMemcpyAsync(c->d_S.p, S, nchains * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_beta.p, beta, nchains * 8, hipMemcpyHostToDevice, c->stream));
  int rc;
  if ((rc = timing_begin(c, K_SWAP)) != IMA_OK) return rc;
  HIPCHK(c, launch_swap(nchains, c->d_S.p, c->d_beta.p, swaptries, seed, counter, c->d_aux.p,
                        c->d_pairs.p, c->stream));
  if ((rc = timing_end(c, K_SWAP)) != IMA_OK) return rc;
  int32_t nacc = 0;
  HIPCHK(c, hipMemcpyAsync(beta, c->d_beta.p, nchains * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(&nacc, c->d_aux.p, 4, hipMemcpyDeviceToHost, c->stream));
  if (pairs_out && swaptries)
    HIPCHK(c, hipMemcpyAsync(pairs_out, c->d_pairs.p, 3 * (size_t)swaptries * 4,
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if ((rc = timing_collect(c, K_SWAP)) != IMA_OK) return rc;
  if (swaps_out) *swaps_out = nacc;
  return IMA_OK;
}

// ---------------------------------------------------------------- L mode --

int ima_lmode_table(ima_ctx *c, const float *table, int G, int K, int npsize, int nmig,
                    int npops, int *table_id) {
  if (!c || !table || !table_id || G < 1 || npsize < 0 || nmig < 0) return IMA_ERR_ARG;
  if (npsize + nmig > 16 || npsize + nmig < 1) return fail(c, IMA_ERR_SIZE, "1..16 parameters supported");
  if (K < 4 * npsize + 3 * nmig + 2) return fail(c, IMA_ERR_ARG, "K=%d too small for the layout", K);
  if ((size_t)8 * K * 8 > (size_t)c->max_lds) return fail(c, IMA_ERR_SIZE, "row too long");
  hipSetDevice(c->device);
  LTableHost t;
  HIPCHK(c, hipMalloc(&t.rows, (size_t)G * K * sizeof(float)));
  HIPCHK(c, hipMemcpyAsync(t.rows, table, (size_t)G * K * sizeof(float), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  t.live = true;
  t.G = G;
  t.cap = G;
  t.K = K;
  t.npsize = npsize;
  t.nmig = nmig;
  t.npops = npops;
  *table_id = (int)c->tables.size();
  c->tables.push_back(t);
  return IMA_OK;
}

int ima_lmode_table_reserve(ima_ctx *c, int capacity, int npsize, int nmig, int npops,
                            int *table_id) {
  if (!c || !table_id || capacity < 1 || npsize < 0 || nmig < 0 || npops < 1) return IMA_ERR_ARG;
  if (npsize + nmig > 16 || npsize + nmig < 1) return fail(c, IMA_ERR_SIZE, "1..16 parameters supported");
  const int K = 4 * npsize + 3 * nmig + 2 + (npops - 1);
  if ((size_t)8 * K * 8 > (size_t)c->max_lds) return fail(c, IMA_ERR_SIZE, "row too long");
  hipSetDevice(c->device);
  LTableHost t;
  HIPCHK(c, hipMalloc(&t.rows, (size_t)capacity * K * sizeof(float)));
  t.live = true;
  t.G = 0;
  t.cap = capacity;
  t.K = K;
  t.npsize = npsize;
  t.nmig = nmig;
  t.npops = npops;
  *table_id = (int)c->tables.size();
  c->tables.push_back(t);
  return IMA_OK;
}

int ima_lmode_table_info(ima_ctx *c, int id, int *G, int *K, int *npsize, int *nmig,
                         int *npops) {
  if (!c || id < 0 || id >= (int)c->tables.size() || !c->tables[id].live) return IMA_ERR_ARG;
  const LTableHost &h = c->tables[id];
  if (G) *G = h.G;
  if (K) *K = h.K;
  if (npsize) *npsize = h.npsize;
  if (nmig) *nmig = h.nmig;
  if (npops) *npops = h.npops;
  return IMA_OK;
}

int ima_lmode_table_read(ima_ctx *c, int id, int first, int count, float *rows) {
  if (!c || !rows || id < 0 || id >= (int)c->tables.size() || !c->tables[id].live) return IMA_ERR_ARG;
  const LTableHost &h = c->tables[id];
  if (first < 0 || count < 0 || first + count > h.G)
    return fail(c, IMA_ERR_ARG, "rows [%d,%d) outside the table's %d", first, first + count, h.G);
  if (count == 0) return IMA_OK;
  hipSetDevice(c->device);
  HIPCHK(c, copy_now(c, rows, h.rows + (size_t)first * h.K, (size_t)count * h.K * sizeof(float),
                     hipMemcpyDeviceToHost));
  return IMA_OK;
}

int ima_lmode_table_free(ima_ctx *c, int id) {
  if (!c || id < 0 || id >= (int)c->tables.size() || !c->tables[id].live) return IMA_ERR_ARG;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  hipFree(c->tables[id].rows);
  c->tables[id].rows = nullptr;
  c->tables[id].live = false;
  return IMA_OK;
}

static LTable ltable(const LTableHost &h) {
  LTable t;
  t.rows = h.rows;
  t.G = h.G;
  t.K = h.K;
  t.npsize = h.npsize;
  t.nmig = h.nmig;
  t.npops = h.npops;
  return t;
}

int ima_lmode_jointp(ima_ctx *c, int id, const double *x, int npoints, int modeltype,
                     int r0, int r1, int mode, double *out, double *ess) {
  if (!c || !x || !out || npoints < 0 || id < 0 || id >= (int)c->tables.size() ||
      !c->tables[id].live)
    return IMA_ERR_ARG;
  const LTableHost &h = c->tables[id];
  const int np = h.npsize + h.nmig;
  if (h.G < 1) return fail(c, IMA_ERR_STATE, "table %d holds no genealogies", id);
  if (r0 < 0 || r1 > np || r0 >= r1) return fail(c, IMA_ERR_ARG, "bad parameter range");
  if (modeltype < 0 || modeltype > 2 || (mode != IMA_LMODE_EXACT && mode != IMA_LMODE_REF_COMPAT))
    return IMA_ERR_ARG;
  if (npoints == 0) return IMA_OK;
  for (size_t i = 0; i < (size_t)npoints * np; ++i)
    if (!(x[i] > 0)) return fail(c, IMA_ERR_ARG, "parameter values must be > 0");
  hipSetDevice(c->device);
  const LTable t = ltable(h);
  const size_t wb = jointp_work_bytes(t, npoints);
  HIPCHK(c, reserve(c->d_lx, (size_t)npoints * np));
  HIPCHK(c, reserve(c->d_lwork, wb / 8 + 1));
  HIPCHK(c, reserve(c->d_lout, (size_t)npoints));
  HIPCHK(c, reserve(c->d_less, (size_t)npoints));
  HIPCHK(c, hipMemcpyAsync(c->d_lx.p, x, (size_t)npoints * np * 8, hipMemcpyHostToDevice,
                           c->stream));
  int rc;
  if ((rc = timing_begin(c, K_LMODE)) != IMA_OK) return rc;
  HIPCHK(c, launch_jointp(t, c->d_lx.p, npoints, modeltype, r0, r1, mode, c->d_lwork.p,
                          c->d_lwork.n * 8, c->d_lout.p, c->d_less.p, c->stream));
  if ((rc = timing_end(c, K_LMODE)) != IMA_OK) return rc;
  HIPCHK(c, hipMemcpyAsync(out, c->d_lout.p, (size_t)npoints * 8, hipMemcpyDeviceToHost,
                           c->stream));
  if (ess)
    HIPCHK(c, hipMemcpyAsync(ess, c->d_less.p, (size_t)npoints * 8, hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return timing_collect(c, K_LMODE);
}

int ima_lmode_jointp_noalloc(ima_ctx *c, int id, const double *x, int npoints, int modeltype,
                             int r0, int r1, int mode, double *out, double *ess) {
  if (!c || !x || !out || npoints < 0 || id < 0 || id >= (int)c->tables.size() ||
      !c->tables[id].live)
    return IMA_ERR_ARG;
  const LTableHost &h = c->tables[id];
  const int np = h.npsize + h.nmig;
  if (h.G < 1) return fail(c, IMA_ERR_STATE, "table %d holds no genealogies", id);
  if (r0 < 0 || r1 > np || r0 >= r1) return fail(c, IMA_ERR_ARG, "bad parameter range");
  if (modeltype < 0 || modeltype > 2 || (mode != IMA_LMODE_EXACT && mode != IMA_LMODE_REF_COMPAT))
    return IMA_ERR_ARG;
  if (!c->scratch) return fail(c, IMA_ERR_STATE, "no scratch area on device %d", c->device);
  if (npoints == 0) return IMA_OK;
  const LTable t = ltable(h);
  const size_t wb = jointp_work_bytes(t, npoints);
  if (wb > kScratchBytes)
    return fail(c, IMA_ERR_SIZE,
                "%d points on %d genealogies need %zu B of work area; the limit is %zu B "
                "(fewer points per call)", npoints, h.G, wb, kScratchBytes);
  hipSetDevice(c->device);
  // every (point, chunk) partial is written before it is read: nothing to clear
  HIPCHK(c, launch_jointp(t, x, npoints, modeltype, r0, r1, mode,
                          static_cast<double *>(c->scratch), kScratchBytes, out, ess, c->stream));
  return IMA_OK;
}

int ima_lmode_popmig(ima_ctx *c, int id, int theta_i, int mig_i, double qmax, double mprior,
                     int expo, int kind, int prob_or_like, int first, int last, const double *x,
                     int nx, double *out) {
  if (!c || !x || !out || nx < 0 || id < 0 || id >= (int)c->tables.size() ||
      !c->tables[id].live)
    return IMA_ERR_ARG;
  const LTableHost &h = c->tables[id];
  if (h.G < 1) return fail(c, IMA_ERR_STATE, "table %d holds no genealogies", id);
  if (theta_i < 0 || theta_i >= h.npsize || mig_i < 0 || mig_i >= h.nmig)
    return fail(c, IMA_ERR_ARG, "no population size %d / migration rate %d in the table", theta_i,
                mig_i);
  if (!(qmax > 0) || !(mprior > 0)) return fail(c, IMA_ERR_ARG, "prior bounds must be > 0");
  if (kind == IMA_POPMIG_CALC) {
    for (int i = 0; i < nx; ++i)
      if (!(x[i] > 0)) return fail(c, IMA_ERR_ARG, "2NM values must be > 0");
  } else if (kind == IMA_POPMIG_MARGIN) {
    if (first < 0 || last > h.G || first >= last)
      return fail(c, IMA_ERR_ARG, "genealogy range [%d, %d) outside the table's %d", first, last,
                  h.G);
  } else {
    return IMA_ERR_ARG;
  }
  if (nx == 0) return IMA_OK;
  hipSetDevice(c->device);
  grow_logfact(c, 16384);
  int rc;
  if ((rc = upload_models(c)) != IMA_OK) return rc;
  const LTable t = ltable(h);
  HIPCHK(c, reserve(c->d_lx, (size_t)nx));
  HIPCHK(c, reserve(c->d_lwork, popmig_work_bytes(t, nx) / 8 + 1));
  HIPCHK(c, reserve(c->d_lout, (size_t)nx));
  HIPCHK(c, hipMemcpyAsync(c->d_lx.p, x, (size_t)nx * 8, hipMemcpyHostToDevice, c->stream));
  if ((rc = timing_begin(c, K_LMODE)) != IMA_OK) return rc;
  HIPCHK(c, launch_popmig(t, theta_i, mig_i, qmax, mprior, expo, kind, prob_or_like, first, last,
                          c->d_lx.p, nx, c->d_logfact.p, (int)c->logfact.size(), c->d_lwork.p,
                          c->d_lout.p, c->stream));
  if ((rc = timing_end(c, K_LMODE)) != IMA_OK) return rc;
  HIPCHK(c, hipMemcpyAsync(out, c->d_lout.p, (size_t)nx * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return timing_collect(c, K_LMODE);
}

int ima_lmode_greater_than(ima_ctx *c, int id, int kind, int n, const double *prior_max,
                           double *out) {
  if (!c || !prior_max || !out || id < 0 || id >= (int)c->tables.size() || !c->tables[id].live ||
      (kind != 0 && kind != 1))
    return IMA_ERR_ARG;
  const LTableHost &h = c->tables[id];
  if (h.G < 1) return fail(c, IMA_ERR_STATE, "table %d holds no genealogies", id);
  if (n != (kind ? h.nmig : h.npsize))
    return fail(c, IMA_ERR_ARG, "the table has %d %s, not %d", kind ? h.nmig : h.npsize,
                kind ? "migration rates" : "population sizes", n);
  std::vector<int2> pairs;
  std::vector<double> pm;
  for (int i = 0; i < n * n; ++i) out[i] = -1.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      if (i == j || prior_max[i] != prior_max[j]) continue;
      if (!(prior_max[i] > 0)) return fail(c, IMA_ERR_ARG, "prior bounds must be > 0");
      if (kind && !(prior_max[i] > 1e-7)) continue;  // MINPARAMVAL
      pairs.push_back(make_int2(i, j));
      pm.push_back(prior_max[i]);
    }
  if (pairs.empty()) return IMA_OK;
  hipSetDevice(c->device);
  grow_logfact(c, 16384);
  int rc;
  if ((rc = upload_models(c)) != IMA_OK) return rc;
  const LTable t = ltable(h);
  const int np = (int)pairs.size();
  const size_t ntrees = std::min(h.G, 20000);
  HIPCHK(c, reserve(c->d_lx, (size_t)np * 2));
  HIPCHK(c, reserve(c->d_lwork, (size_t)np * ntrees + (size_t)np));
  HIPCHK(c, reserve(c->d_lout, (size_t)np));
  std::vector<double> stage(2 * (size_t)np);
  memcpy(stage.data(), pairs.data(), (size_t)np * sizeof(int2));
  memcpy(stage.data() + np, pm.data(), (size_t)np * 8);
  HIPCHK(c, hipMemcpyAsync(c->d_lx.p, stage.data(), stage.size() * 8, hipMemcpyHostToDevice,
                           c->stream));
  if ((rc = timing_begin(c, K_LMODE)) != IMA_OK) return rc;
  HIPCHK(c, launch_greater_than(t, kind, reinterpret_cast<const int2 *>(c->d_lx.p),
                                c->d_lx.p + np, np, c->d_logfact.p, (int)c->logfact.size(),
                                c->d_lwork.p, c->d_lout.p, c->stream));
  if ((rc = timing_end(c, K_LMODE)) != IMA_OK) return rc;
  std::vector<double> res(np);
  HIPCHK(c, hipMemcpyAsync(res.data(), c->d_lout.p, (size_t)np * 8, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < np; ++k) out[pairs[k].x * n + pairs[k].y] = res[k];
  return timing_collect(c, K_LMODE);
}

int ima_lmode_margincalc(ima_ctx *c, int id, int param, const double *x, int nx, int logi,
                         double yadjust, int expo, double prior_mean, double *out) {
  if (!c || !x || !out || nx < 0 || id < 0 || id >= (int)c->tables.size() ||
      !c->tables[id].live)
    return IMA_ERR_ARG;
  const LTableHost &h = c->tables[id];
  if (h.G < 1) return fail(c, IMA_ERR_STATE, "table %d holds no genealogies", id);
  if (param < 0 || param >= h.npsize + h.nmig) return fail(c, IMA_ERR_ARG, "bad parameter index");
  if (expo && !(prior_mean > 0)) return fail(c, IMA_ERR_ARG, "prior mean must be > 0");
  if (nx == 0) return IMA_OK;
  for (int i = 0; i < nx; ++i)
    if (!(x[i] > 0)) return fail(c, IMA_ERR_ARG, "x must be > 0");
  hipSetDevice(c->device);
  const LTable t = ltable(h);
  const size_t wb = margincalc_work_bytes(t, nx);
  HIPCHK(c, reserve(c->d_lx, (size_t)nx));
  HIPCHK(c, reserve(c->d_lwork, wb / 8 + 1));
  HIPCHK(c, reserve(c->d_lout, (size_t)nx));
  HIPCHK(c, hipMemcpyAsync(c->d_lx.p, x, (size_t)nx * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, launch_margincalc(t, param, c->d_lx.p, nx, logi, yadjust, expo, prior_mean,
                              c->d_lwork.p, c->d_lout.p, c->stream));
  HIPCHK(c, hipMemcpyAsync(out, c->d_lout.p, (size_t)nx * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return IMA_OK;
}

// ------------------------------------------------------------------- MC3 --

// The captured sweeps hold the sweep's configuration; drop them when it
// changes (they are captured again at the next run).
static void mc3_drop_graphs(ima_ctx *c) {
  auto &m = c->mc3;
  for (int k = 0; k < 2; ++k) {
    if (m.exec[k]) hipGraphExecDestroy(m.exec[k]);
    if (m.graph[k]) hipGraphDestroy(m.graph[k]);
    if (m.bexec[k]) hipGraphExecDestroy(m.bexec[k]);
    if (m.bgraph[k]) hipGraphDestroy(m.bgraph[k]);
    m.exec[k] = m.bexec[k] = nullptr;
    m.graph[k] = m.bgraph[k] = nullptr;
    for (int p = 0; p < 2; ++p) {
      if (m.uexec[k][p]) hipGraphExecDestroy(m.uexec[k][p]);
      if (m.ugraph[k][p]) hipGraphDestroy(m.ugraph[k][p]);
      m.uexec[k][p] = nullptr;
      m.ugraph[k][p] = nullptr;
    }
  }
}

static void mc3_release(ima_ctx *c) {
  auto &m = c->mc3;
  mc3_drop_graphs(c);
  for (int k = 0; k < 4; ++k) {
    if (m.ev[k]) {
      hipEventDestroy(m.ev[k]);
      m.ev[k] = nullptr;
    }
    if (m.fj[k]) {
      hipEventDestroy(m.fj[k]);
      m.fj[k] = nullptr;
    }
  }
  release(m.pdg);
  release(m.newpdg);
  release(m.beta);
  release(m.S);
  release(m.rec_t);
  release(m.tbackup);
  release(m.logs);
  release(m.status);
  release(m.rec_v);
  release(m.decision);
  release(m.idx);
  release(m.node_item);
  release(m.iblk);
  release(m.done);
  release(m.acc);
  release(m.cacc);
  release(m.counter);
  release(m.stats);
  release(m.items);
  release(m.units);
  release(m.red);
  release(m.partial);
  release(m.tv);
  release(m.tvbackup);
  release(m.wrec);
  release(m.wnew);
  release(m.wall);
  release(m.qterm);
  release(m.qf);
  release(m.probg);
  release(m.mtbackup);
  release(m.qc);
  release(m.tv_row);
  release(m.tw_slots);
  release(m.wstatus);
  release(m.uitems);
  release(m.ustep);
  release(m.ulist0);
  release(m.ulist1);
  release(m.uuni);
  release(m.uorder);
  release(m.uloff);
  release(m.ulast);
  release(m.ucnt);
  release(m.uilist0);
  release(m.uilist1);
  release(m.uslist0);
  release(m.uslist1);
  release(m.hast);
  release(m.uhost);
  release(m.idec);
  release(m.ptw);
  release(m.pitems);
  release(m.thast);
  release(m.titems);
  release(m.tptw);
  release(m.arange);
  release(m.nwk);
  release(m.nwst);
  release(m.nwnbk);
  release(m.nwt);
  release(m.nwmig);
  m.dev.nw = 0;
  m.topo_first = -1;
  m.gview_set = false;
  m.udev = URateDev{};
  m.dev.prior = -1;
  m.live = false;
}

// The MC3 chains' HKY scoring plan with site blocks of 64 * ppl patterns:
// units in XCD order, block-partial merges, each item's first partial and
// block count (iblk).  Captured sweeps embed the plan, so they are dropped.
static int mc3_build_plan(ima_ctx *c, int ppl) {
  auto &m = c->mc3;
  const int N = m.C * m.L;
  std::vector<int4> units, red;
  int nblk = 0;
  for (int i = 0; i < N; ++i) {
    const LocusMeta &lm = c->loci[c->slots[m.first + i].locus];
    if (lm.model == IMA_MODEL_HKY) hky_plan_add(units, red, i, lm.S, &nblk, ppl);
  }
  {
    std::vector<int> lu(units.size());
    for (size_t k = 0; k < units.size(); ++k) lu[k] = c->slots[m.first + units[k].x].locus;
    units = xcd_order(units, lu);
  }
  HIPCHK(c, reserve(m.units, units.size() + 1));
  HIPCHK(c, reserve(m.red, red.size() + 1));
  HIPCHK(c, reserve(m.partial, (size_t)nblk + 1));
  if (!units.empty())
    HIPCHK(c, hipMemcpyAsync(m.units.p, units.data(), units.size() * sizeof(int4),
                             hipMemcpyHostToDevice, c->stream));
  if (!red.empty())
    HIPCHK(c, hipMemcpyAsync(m.red.p, red.data(), red.size() * sizeof(int4),
                             hipMemcpyHostToDevice, c->stream));
  m.plan.units = m.units.p;
  m.plan.nunits = (int)units.size();
  m.plan.red = m.red.p;
  m.plan.nred = (int)red.size();
  m.plan.partial = m.partial.p;
  m.plan.ppl = ppl;
  std::vector<int2> iblk(N, make_int2(0, 0));
  for (const int4 &u : units)
    if (u.y == 0) iblk[u.x] = make_int2(u.z, u.w);
  HIPCHK(c, reserve(m.iblk, N));
  HIPCHK(c, hipMemcpyAsync(m.iblk.p, iblk.data(), N * sizeof(int2), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  m.dev.partial = m.partial.p;
  m.dev.iblk = m.iblk.p;
  mc3_drop_graphs(c);
  return IMA_OK;
}

int ima_mc3_setup(ima_ctx *c, int nchains, int nloci, int first_slot, const double *u,
                  const double *kappa, const double *beta, uint64_t seed, int precision,
                  int swaptries, double scale_width) {
  if (!c || !u || !beta || nchains < 1 || nloci < 1 || swaptries < 0) return IMA_ERR_ARG;
  if (precision != IMA_PREC_FP64 && precision != IMA_PREC_FP32) return IMA_ERR_ARG;
  const int N = nchains * nloci;
  if (first_slot < 0 || first_slot + N > (int)c->slots.size())
    return fail(c, IMA_ERR_ARG, "MC3 slots [%d,%d) out of range", first_slot, first_slot + N);
  hipSetDevice(c->device);
  mc3_release(c);
  auto &m = c->mc3;
  m.C = nchains;
  m.L = nloci;
  m.first = first_slot;
  m.precision = precision;
  m.swaptries = swaptries;
  std::vector<ima_item> items(N);
  std::vector<int32_t> hky, is, sw;
  m.hmax_n = 2;
  m.hmax_slots = 1;
  m.imax_n = 2;
  m.imax_w = 1;
  for (int i = 0; i < N; ++i) {
    const SlotMeta &s = c->slots[first_slot + i];
    if (s.root < 0) return fail(c, IMA_ERR_STATE, "MC3 slot %d holds no genealogy", first_slot + i);
    items[i].slot = first_slot + i;
    items[i].flags = 0;
    items[i].u = u[i];
    items[i].kappa = kappa ? kappa[i] : 1.0;
    items[i].tr = 1.0;
    const LocusMeta &lm = c->loci[s.locus];
    if (lm.model == IMA_MODEL_HKY) {
      hky.push_back(i);
      m.hmax_n = std::max(m.hmax_n, s.n);
      m.hmax_slots = std::max(m.hmax_slots, s.nslots);
    } else if (lm.model == IMA_MODEL_IS) {
      is.push_back(i);
      m.imax_n = std::max(m.imax_n, s.n);
      m.imax_w = std::max(m.imax_w, lm.W);
    } else {
      sw.push_back(i);
    }
  }
  // time-slide proposals keep every topology, so the stack depth the
  // programs need is fixed for the run unless topology moves are enabled
  // (mc3_topo_mode raises it to the bound of any topology)
  if (!hky.empty() && hky_lds_bytes(precision, m.hmax_n, m.hmax_slots, 1) > (size_t)c->max_lds)
    return fail(c, IMA_ERR_SIZE, "HKY loci of %d gene copies need %zu B of LDS (limit %d)",
                m.hmax_n, hky_lds_bytes(precision, m.hmax_n, m.hmax_slots, 1), c->max_lds);
  m.nh = (int)hky.size();
  m.ni = (int)is.size();
  m.ns = (int)sw.size();
  std::vector<int32_t> idx;
  idx.insert(idx.end(), hky.begin(), hky.end());
  idx.insert(idx.end(), is.begin(), is.end());
  idx.insert(idx.end(), sw.begin(), sw.end());
  int rc;
  if ((rc = upload_loci(c)) != IMA_OK) return rc;
  if ((rc = upload_slot_meta(c)) != IMA_OK) return rc;
  HIPCHK(c, reserve(m.pdg, N));
  HIPCHK(c, reserve(m.newpdg, N));
  HIPCHK(c, reserve(m.status, N));
  HIPCHK(c, reserve(m.beta, nchains));
  HIPCHK(c, reserve(m.S, nchains));
  HIPCHK(c, reserve(m.rec_v, N));
  HIPCHK(c, reserve(m.rec_t, N));
  HIPCHK(c, reserve(m.tbackup, c->nodes + 1));
  HIPCHK(c, reserve(m.decision, nchains));
  HIPCHK(c, reserve(m.logs, nchains));
  HIPCHK(c, reserve(m.counter, 1));
  HIPCHK(c, reserve(m.stats, 8));
  HIPCHK(c, reserve(m.items, N));
  HIPCHK(c, reserve(m.idx, N));
  // node -> item map over the contiguous node range of the MC3 slots
  const SlotMeta &s0 = c->slots[first_slot], &s1 = c->slots[first_slot + N - 1];
  const int64_t node_first = s0.node_off;
  const int64_t nnodes = s1.node_off + 2 * (int64_t)s1.n - 1 - node_first;
  std::vector<int32_t> node_item((size_t)nnodes);
  for (int i = 0; i < N; ++i) {
    const SlotMeta &s = c->slots[first_slot + i];
    for (int k = 0; k < 2 * s.n - 1; ++k) node_item[(size_t)(s.node_off - node_first + k)] = i;
  }
  HIPCHK(c, reserve(m.node_item, (size_t)nnodes));
  HIPCHK(c, reserve(m.acc, N));
  HIPCHK(c, reserve(m.cacc, nchains));
  HIPCHK(c, hipMemcpyAsync(m.node_item.p, node_item.data(), (size_t)nnodes * 4,
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(m.acc.p, 0, N * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(m.cacc.p, 0, nchains * 4, c->stream));
  m.scale_sweeps = 0;
  m.host_sweeps = 0;
  HIPCHK(c, hipMemcpyAsync(m.items.p, items.data(), N * sizeof(ima_item), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(m.idx.p, idx.data(), N * 4, hipMemcpyHostToDevice, c->stream));
  {
    std::vector<int> hslots;
    for (int i : hky) hslots.push_back(first_slot + i);
    int ppl = hky_auto_ppl(c, hslots);
    if (ppl == 2 && hky_lds_bytes(precision, m.hmax_n, m.hmax_slots, 2) > (size_t)c->max_lds)
      ppl = 1;
    int rc;
    if ((rc = mc3_build_plan(c, ppl)) != IMA_OK) return rc;
    // the rate rounds' unit lists hold up to every unit at 64 patterns
    size_t units64 = 0;
    for (int i : hky) units64 += (c->loci[c->slots[first_slot + i].locus].S + kWave - 1) / kWave;
    HIPCHK(c, reserve(m.done, 1));
    HIPCHK(c, hipMemsetAsync(m.done.p, 0, sizeof(unsigned), c->stream));
    // mutation-rate rounds (update_mc_params.cpp:24-429): log(UMAX)/L
    // window, reflection at 3 log(UMAX) (initialize.cpp:1448-1450,
    // update_mc_params.cpp:52-56), kappa window 2 up to KAPPAMAX
    // (initialize.cpp:1499-1501)
    const int J = nloci > 2 ? nloci : 1;
    HIPCHK(c, reserve(m.uitems, N));
    HIPCHK(c, reserve(m.ustep, (size_t)nchains * J));
    HIPCHK(c, reserve(m.uuni, (size_t)nchains * J));
    HIPCHK(c, reserve(m.uorder, (size_t)nchains * J));
    HIPCHK(c, reserve(m.uloff, (size_t)nchains * (J + 2)));
    HIPCHK(c, reserve(m.ulast, N));
    HIPCHK(c, reserve(m.ulist0, units64 + 1));
    HIPCHK(c, reserve(m.ulist1, units64 + 1));
    HIPCHK(c, reserve(m.uilist0, is.size() + 1));
    HIPCHK(c, reserve(m.uilist1, is.size() + 1));
    HIPCHK(c, reserve(m.uslist0, sw.size() + 1));
    HIPCHK(c, reserve(m.uslist1, sw.size() + 1));
    HIPCHK(c, reserve(m.ucnt, 6));
    HIPCHK(c, hipMemsetAsync(m.ucnt.p, 0, 6 * sizeof(int32_t), c->stream));
    if (urate_plan_lds(nloci, J) > 65536)
      return fail(c, IMA_ERR_ARG, "%d loci per chain: the mutation-rate plan needs more LDS",
                  nloci);
    URateDev &u = m.udev;
    u.J = J;
    u.uitems = m.uitems.p;
    u.step = m.ustep.p;
    u.uni = m.uuni.p;
    u.order = m.uorder.p;
    u.loff = m.uloff.p;
    u.last = m.ulast.p;
    u.list[0] = m.ulist0.p;
    u.list[1] = m.ulist1.p;
    u.ilist[0] = m.uilist0.p;
    u.ilist[1] = m.uilist1.p;
    u.slist[0] = m.uslist0.p;
    u.slist[1] = m.uslist1.p;
    u.cnt = m.ucnt.p;
    u.win = log(10000.0) / nloci;
    u.maxr = 3.0 * log(10000.0);
    u.kwin = 2.0;
    u.kmax = 100.0;
  }
  HIPCHK(c, hipMemcpyAsync(m.beta.p, beta, nchains * 8, hipMemcpyHostToDevice, c->stream));
  m.beta0.assign(beta, beta + nchains);
  const uint64_t zero = 0;
  const unsigned long long zeros[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  HIPCHK(c, hipMemcpyAsync(m.counter.p, &zero, 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(m.stats.p, zeros, sizeof zeros, hipMemcpyHostToDevice, c->stream));
  MC3Dev &d = m.dev;
  d.C = nchains;
  d.group = nchains;
  d.L = nloci;
  d.first_slot = first_slot;
  d.pdg = m.pdg.p;
  d.newpdg = m.newpdg.p;
  d.status = m.status.p;
  d.beta = m.beta.p;
  d.S = m.S.p;
  d.rec_v = m.rec_v.p;
  d.rec_t = m.rec_t.p;
  d.tbackup = m.tbackup.p;
  d.decision = m.decision.p;
  d.logs = m.logs.p;
  d.counter = m.counter.p;
  d.stats = m.stats.p;
  d.acc = m.acc.p;
  d.cacc = m.cacc.p;
  d.node_first = node_first;
  d.nnodes = nnodes;
  d.node_item = m.node_item.p;
  d.iblk = m.iblk.p;
  d.partial = m.partial.p;
  d.done = m.done.p;
  HIPCHK(c, reserve(m.tbackup, (size_t)nnodes + 1));
  d.tbackup = m.tbackup.p;
  d.seed = seed;
  d.scale_width = scale_width;
  d.ry1 = 0;
  d.nw = 0;
  d.prior = -1;
  d.P = 1;
  d.R = 0;
  d.nparam = 0;
  d.pchunk = 1;
  d.mints = 0;
  d.host = 0;
  d.gbeta = 1.0;
  d.hast = nullptr;
  d.uhost = nullptr;
  d.idec = nullptr;
  d.prop_first = 0;
  d.topo = 0;
  d.thast = nullptr;
  d.arange = nullptr;
  d.pitems = nullptr;
  d.ptw = nullptr;
  d.mig_pool = 0;
  d.maxn = 2;
  for (int i = 0; i < N; ++i) d.maxn = std::max(d.maxn, c->slots[first_slot + i].n);
  m.topo_first = -1;
  for (int k = 0; k < 4; ++k) {
    HIPCHK(c, hipEventCreate(&m.ev[k]));
    HIPCHK(c, hipEventCreateWithFlags(&m.fj[k], hipEventDisableTiming));
  }
  // initial likelihood of every genealogy
  const DeviceView v = view(c);
  if (m.nh)
    HIPCHK(c, launch_hky(v, m.items.p, m.plan, precision, m.hmax_n, m.hmax_slots, m.pdg.p,
                         m.status.p, c->stream));
  if (m.ni)
    HIPCHK(c, launch_is(v, m.items.p, m.idx.p + m.nh, m.ni, m.imax_n, m.imax_w, m.pdg.p,
                        m.status.p, nullptr, nullptr, c->stream));
  if (m.ns)
    HIPCHK(c, launch_sw(v, m.items.p, m.idx.p + m.nh + m.ni, m.ns, m.pdg.p, m.status.p,
                        nullptr, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  m.live = true;
  return IMA_OK;
}

static int mc3_enqueue(ima_ctx *c, bool with_scale, hipEvent_t *score_ev) {
  auto &m = c->mc3;
  HIPCHK(c, enqueue_mc3_sweep(view(c), m.dev, with_scale, m.items.p, m.idx.p, m.plan, m.nh,
                              m.ni, m.ns,
                              m.precision, m.hmax_n, m.hmax_slots, m.imax_n, m.imax_w,
                              m.swaptries, score_ev, c->stream, c->side, m.fj));
  return IMA_OK;
}

}  // extern "C"

// The mutation-rate round after the sweep whose counter is ctr - 1: the
// host draws each chain's partners from the same streams as the device to
// know how many levels to launch (the deepest chain's).
static int mc3_urate_round(ima_ctx *c, uint64_t ctr, bool serial) {
  auto &m = c->mc3;
  const int L = m.L, J = m.udev.J;
  int nlev = 1;
  if (serial) {
    nlev = J;
  } else {
    std::vector<int32_t> &last = m.ulevel_scratch;
    last.assign((size_t)L, 0);
    for (int ch = 0; ch < m.C; ++ch) {
      std::fill(last.begin(), last.end(), 0);
      for (int s = 0; s < J; ++s) {
        PhiloxStream r = urate_rng(m.dev.seed, ch, L, s, ctr);
        const int k = urate_partner(r, s, L);
        const int lev = 1 + std::max(last[s], k >= 0 ? last[k] : 0);
        last[s] = lev;
        if (k >= 0) last[k] = lev;
        nlev = std::max(nlev, lev);
      }
    }
  }
  HIPCHK(c, enqueue_urate_round(view(c), m.dev, m.udev, m.items.p, m.idx.p, m.plan, m.nh, m.ni,
                                m.ns, m.precision, m.hmax_n, m.hmax_slots, m.imax_n, m.imax_w,
                                ctr, nlev, serial ? 1 : 0, c->stream));
  return IMA_OK;
}

// Enqueues nsweeps sweeps on the context's stream without waiting (timed
// runs synchronise per sweep).  The sweep counter's host copy decides which
// sweeps rescale and which end with a mutation-rate round, so nothing is
// read back from the device.
// Turns the device's own genealogy move on or off (IMA_MC3_TOPOLOGY).  On
// first use the proposal slots are allocated (one per chain and locus) and
// the stepwise loci' allele ranges set from chain 0's gene copies
// (L[li].minA / maxA, build_gtree.cpp:707-722).
static int grow_mig_caps(ima_ctx *c, const int32_t *slots, const int32_t *caps, int n);
// Room for new migrations a slot gets beyond twice its count, when a move
// that draws migration paths is first used.
constexpr int kNwRoom = 32;

// The move with migration rewrites migration times that the rescaling
// backup covers: the backup follows the event arrays' size.
static int mc3_topo_mig_buffers(ima_ctx *c) {
  auto &m = c->mc3;
  HIPCHK(c, reserve(m.mtbackup, c->migs + 1));
  if (m.dev.mtbackup != m.mtbackup.p) {
    m.dev.mtbackup = m.mtbackup.p;
    mc3_drop_graphs(c);
  }
  return IMA_OK;
}

static int mc3_topo_mode(ima_ctx *c, bool on) {
  auto &m = c->mc3;
  MC3Dev &d = m.dev;
  if (!on) {
    if (d.topo) {
      d.topo = 0;
      d.hast = nullptr;
      mc3_drop_graphs(c);
    }
    return IMA_OK;
  }
  if (d.topo) return d.mig_pool > 0 ? mc3_topo_mig_buffers(c) : IMA_OK;
  {
    // any topology may come up: the pruning stack must hold floor(log2 n)
    int bound = 1;
    while ((2 << bound) <= m.hmax_n) ++bound;
    if (m.nh && hky_lds_bytes(m.precision, m.hmax_n, std::max(m.hmax_slots, bound),
                              m.plan.ppl) > (size_t)c->max_lds)
      return fail(c, IMA_ERR_SIZE, "topology moves on %d gene copies need more LDS than %d B",
                  m.hmax_n, c->max_lds);
  }
  const int N = m.C * m.L;
  int rc;
  if (m.topo_first < 0) {
    std::vector<int32_t> loci(N);
    for (int i = 0; i < N; ++i) loci[i] = c->slots[m.first + i].locus;
    int first = 0;
    if ((rc = ima_slots_alloc(c, N, loci.data(), &first)) != IMA_OK) return rc;
    m.topo_first = first;
    HIPCHK(c, reserve(c->d_mig_first, c->nodes, true, c->stream));
    HIPCHK(c, reserve(c->d_mig_count, c->nodes, true, c->stream));
  }
  HIPCHK(c, reserve(m.thast, N));
  HIPCHK(c, reserve(m.titems, N));
  HIPCHK(c, reserve(m.tptw, N));
  std::vector<int2> ar(m.L, make_int2(1, 1000));
  for (int l = 0; l < m.L; ++l) {
    const SlotMeta &s = c->slots[m.first + l];
    if (c->loci[s.locus].model != IMA_MODEL_SW) continue;
    int mn = 1000, mx = 0;
    for (int e = 0; e < s.n; ++e) {
      mn = std::min(mn, c->A[s.node_off + e]);
      mx = std::max(mx, c->A[s.node_off + e]);
    }
    ar[l] = make_int2(std::max(1, (mx + mn) / 2 - (mx - mn)), std::min(1000, (mx + mn) / 2 + (mx - mn)));
  }
  HIPCHK(c, reserve(m.arange, m.L));
  HIPCHK(c, hipMemcpyAsync(m.arange.p, ar.data(), ar.size() * sizeof(int2), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // any topology may come up: the pruning stack must hold floor(log2 n)
  int bound = 1;
  while ((2 << bound) <= m.hmax_n) ++bound;
  m.hmax_slots = std::max(m.hmax_slots, bound);
  d.mig_pool = 0;
  if (d.prior >= 0 && c->models[d.prior].nmig > 0 && d.P > 1) {
    // the move draws migration paths: every chain slot and its proposal
    // slot get the same room (twice the chain's migrations plus kNwRoom),
    // the prior's event bound and log-factorial table grow to match, and
    // each wavefront holds an LDS pool of four times the largest room
    std::vector<int32_t> sl(2 * (size_t)N), cp(2 * (size_t)N);
    for (int i = 0; i < N; ++i) {
      const SlotMeta &cs = c->slots[m.first + i];
      const int cap = std::max(cs.mig_cap, 2 * cs.nmig + kNwRoom);
      sl[i] = m.first + i;
      sl[N + i] = m.topo_first + i;
      cp[i] = cp[N + i] = cap;
    }
    if ((rc = grow_mig_caps(c, sl.data(), cp.data(), 2 * N)) != IMA_OK) return rc;
    int emax = d.emax, maxcap = 0;
    size_t lfn = 4;
    for (int ch = 0; ch < m.C; ++ch) {
      size_t tot = 4;
      for (int l = 0; l < m.L; ++l) {
        const SlotMeta &cs = c->slots[m.first + ch * m.L + l];
        emax = std::max(emax, (cs.n - 1) + cs.mig_cap + d.P - 1);
        maxcap = std::max(maxcap, cs.mig_cap);
        tot += (size_t)cs.n - 1 + (size_t)cs.mig_cap;
      }
      lfn = std::max(lfn, tot);
    }
    if (treeweight_lds_bytes(d.P, emax) > (size_t)c->max_lds)
      return fail(c, IMA_ERR_SIZE, "room for the move's migrations: %d prior events exceed LDS",
                  emax);
    const int pool = 4 * std::max(maxcap, 16);
    // + 1 KiB: the kernel's static LDS (model maps, split times, flags)
    if (topo_mig_lds_bytes(d.maxn, pool) + 1024 > (size_t)c->max_lds)
      return fail(c, IMA_ERR_SIZE, "the move with migration on %d gene copies and %d events "
                  "needs more LDS than %d B", d.maxn, maxcap, c->max_lds);
    grow_logfact(c, lfn);
    d.emax = emax;
    if (d.nw) d.nwa.emax = emax;
    d.mig_pool = pool;
    if ((rc = mc3_topo_mig_buffers(c)) != IMA_OK) return rc;
  }
  HIPCHK(c, reserve(m.idec, N));
  d.idec = m.idec.p;  // the sweep's decisions, read when it copies accepted proposals
  d.topo = 1;
  d.thast = m.thast.p;
  d.hast = m.thast.p;
  d.arange = m.arange.p;
  d.pitems = m.titems.p;
  d.ptw = m.tptw.p;
  d.prop_first = m.topo_first;
  mc3_drop_graphs(c);
  return IMA_OK;
}

// Gives slots[i] room for at least caps[i] migration events: the events of
// each slot that needs more move to a new region at the end of the event
// arrays (host mirror and device).
static int grow_mig_caps(ima_ctx *c, const int32_t *slots, const int32_t *caps, int n) {
  std::vector<int64_t> from, to;
  std::vector<int32_t> cnt;
  for (int i = 0; i < n; ++i) {
    SlotMeta &d = c->slots[slots[i]];
    if (d.mig_cap >= caps[i]) continue;
    from.push_back(d.mig_off);
    to.push_back((int64_t)c->migs);
    cnt.push_back(d.nmig);
    d.mig_off = (int64_t)c->migs;
    d.mig_cap = caps[i];
    c->migs += (size_t)caps[i];
  }
  if (from.empty()) return IMA_OK;
  c->mig_edge.resize(c->migs, 0);
  c->mig_to.resize(c->migs, 0);
  c->mig_time.resize(c->migs, 0.0);
  for (size_t i = 0; i < from.size(); ++i) {
    std::copy(c->mig_edge.begin() + from[i], c->mig_edge.begin() + from[i] + cnt[i],
              c->mig_edge.begin() + to[i]);
    std::copy(c->mig_to.begin() + from[i], c->mig_to.begin() + from[i] + cnt[i],
              c->mig_to.begin() + to[i]);
    std::copy(c->mig_time.begin() + from[i], c->mig_time.begin() + from[i] + cnt[i],
              c->mig_time.begin() + to[i]);
  }
  HIPCHK(c, reserve(c->d_mig_edge, c->migs + 1, true, c->stream));
  HIPCHK(c, reserve(c->d_mig_to, c->migs + 1, true, c->stream));
  HIPCHK(c, reserve(c->d_mig_time, c->migs + 1, true, c->stream));
  // the device's events (their times may have moved since the host wrote
  // them) follow to the new regions, one launch for every slot
  const size_t m = from.size();
  HIPCHK(c, reserve(c->d_mv_off, 2 * m));
  HIPCHK(c, reserve(c->d_mv_cnt, m));
  HIPCHK(c, hipMemcpyAsync(c->d_mv_off.p, from.data(), m * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_mv_off.p + m, to.data(), m * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_mv_cnt.p, cnt.data(), m * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, launch_move_events((int)m, c->d_mv_off.p, c->d_mv_off.p + m, c->d_mv_cnt.p,
                               c->d_mig_edge.p, c->d_mig_to.p, c->d_mig_time.p, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));  // the host arrays above are freed on return
  c->slots_meta_dirty = true;
  return IMA_OK;
}

// The host mirror of the slots' populations, per-edge indexes and events,
// uploaded (the NW update reads them on the device).
static int nw_upload_slots(ima_ctx *c, const int32_t *slots, int nslots) {
  HIPCHK(c, reserve(c->d_pop, c->nodes, true, c->stream));
  for (int i = 0; i < nslots; ++i) {
    const SlotMeta &s = c->slots[slots[i]];
    const size_t ne = 2 * (size_t)s.n - 1;
    HIPCHK(c, hipMemcpyAsync(c->d_pop.p + s.node_off, c->pop.data() + s.node_off, ne * 4,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_mig_first.p + s.node_off, c->mig_first.data() + s.node_off,
                             ne * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_mig_count.p + s.node_off, c->mig_count.data() + s.node_off,
                             ne * 4, hipMemcpyHostToDevice, c->stream));
    if (s.nmig) {
      HIPCHK(c, hipMemcpyAsync(c->d_mig_edge.p + s.mig_off, c->mig_edge.data() + s.mig_off,
                               (size_t)s.nmig * 4, hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipMemcpyAsync(c->d_mig_to.p + s.mig_off, c->mig_to.data() + s.mig_off,
                               (size_t)s.nmig * 4, hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipMemcpyAsync(c->d_mig_time.p + s.mig_off, c->mig_time.data() + s.mig_off,
                               (size_t)s.nmig * 8, hipMemcpyHostToDevice, c->stream));
    }
  }
  return IMA_OK;
}

// Largest event count of a genealogy whose weights fit treeweight's LDS.
static int nw_event_limit(ima_ctx *c, int P) {
  const size_t fixed = treeweight_lds_bytes(P, 0);
  if ((size_t)c->max_lds <= fixed) return 1;
  return (int)(((size_t)c->max_lds - fixed) / 24);
}

// The NW update's backups, indexed like the node and migration arrays.
static int nw_reserve_backups(ima_ctx *c) {
  HIPCHK(c, reserve(c->d_nw_pop, c->nodes + 1));
  HIPCHK(c, reserve(c->d_nw_mf, c->nodes + 1));
  HIPCHK(c, reserve(c->d_nw_mc, c->nodes + 1));
  HIPCHK(c, reserve(c->d_nw_me, c->migs + 1));
  HIPCHK(c, reserve(c->d_nw_mto, c->migs + 1));
  HIPCHK(c, reserve(c->d_nw_mt, c->migs + 1));
  // per-edge migration indexes on the device for every slot
  if (c->d_mig_first.n < c->nodes || c->d_mig_count.n < c->nodes) {
    HIPCHK(c, reserve(c->d_mig_first, c->nodes, false, c->stream));
    HIPCHK(c, reserve(c->d_mig_count, c->nodes, false, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_mig_first.p, c->mig_first.data(), c->nodes * 4,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_mig_count.p, c->mig_count.data(), c->nodes * 4,
                             hipMemcpyHostToDevice, c->stream));
  }
  return IMA_OK;
}

// The host mirror of slots whose populations and migrations the device
// rewrote: slot records, populations, per-edge indexes and events.
static int nw_sync_slots(ima_ctx *c, const int32_t *slots, int nslots) {
  for (int i = 0; i < nslots; ++i) {
    const int sl = slots[i];
    HIPCHK(c, copy_now(c, &c->slots[sl], c->d_slots.p + sl, sizeof(SlotMeta),
                       hipMemcpyDeviceToHost));
    const SlotMeta &s = c->slots[sl];
    const size_t ne = 2 * (size_t)s.n - 1;
    HIPCHK(c, copy_now(c, c->pop.data() + s.node_off, c->d_pop.p + s.node_off, ne * 4,
                       hipMemcpyDeviceToHost));
    HIPCHK(c, copy_now(c, c->mig_first.data() + s.node_off, c->d_mig_first.p + s.node_off,
                       ne * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, copy_now(c, c->mig_count.data() + s.node_off, c->d_mig_count.p + s.node_off,
                       ne * 4, hipMemcpyDeviceToHost));
    if (s.nmig) {
      HIPCHK(c, copy_now(c, c->mig_edge.data() + s.mig_off, c->d_mig_edge.p + s.mig_off,
                         (size_t)s.nmig * 4, hipMemcpyDeviceToHost));
      HIPCHK(c, copy_now(c, c->mig_to.data() + s.mig_off, c->d_mig_to.p + s.mig_off,
                         (size_t)s.nmig * 4, hipMemcpyDeviceToHost));
      HIPCHK(c, copy_now(c, c->mig_time.data() + s.mig_off, c->d_mig_time.p + s.mig_off,
                         (size_t)s.nmig * 8, hipMemcpyDeviceToHost));
    }
  }
  return IMA_OK;
}

// Before every run with the NW update: the migration-time backup and the
// NW backups cover the current node and event arrays (slots allocated or
// moved since may have grown them); a buffer that moved drops the graphs.
static int mc3_nw_buffers(ima_ctx *c) {
  auto &m = c->mc3;
  MC3Dev &d = m.dev;
  HIPCHK(c, reserve(m.mtbackup, c->migs + 1));
  int rc;
  if ((rc = nw_reserve_backups(c)) != IMA_OK) return rc;
  NwArgs &a = d.nwa;
  const bool moved = d.mtbackup != m.mtbackup.p || a.pop_bk != c->d_nw_pop.p ||
                     a.mf_bk != c->d_nw_mf.p || a.mc_bk != c->d_nw_mc.p ||
                     a.me_bk != c->d_nw_me.p || a.mto_bk != c->d_nw_mto.p ||
                     a.mt_bk != c->d_nw_mt.p || a.tv != m.tvbackup.p || a.row != m.tv_row.p ||
                     a.counter != m.counter.p;
  d.mtbackup = m.mtbackup.p;
  a.pop_bk = c->d_nw_pop.p;
  a.mf_bk = c->d_nw_mf.p;
  a.mc_bk = c->d_nw_mc.p;
  a.me_bk = c->d_nw_me.p;
  a.mto_bk = c->d_nw_mto.p;
  a.mt_bk = c->d_nw_mt.p;
  a.tv = m.tvbackup.p;
  a.row = m.tv_row.p;
  a.counter = m.counter.p;
  if (moved) mc3_drop_graphs(c);
  return IMA_OK;
}

// changet_NW as the other time update (IMA_MC3_NW).  On first use every
// chain slot gets room for new migrations (kNwRoom beyond twice its
// count), and the event bound of the genealogy-prior kernels, the
// log-factorial table and the backups grow to match.
static int mc3_nw_mode(ima_ctx *c, bool on) {
  auto &m = c->mc3;
  MC3Dev &d = m.dev;
  if (!on) {
    if (d.nw) {
      d.nw = 0;
      mc3_drop_graphs(c);
    }
    return IMA_OK;
  }
  if (d.nw) return mc3_nw_buffers(c);
  // as in the reference, NW is not drawn without migration (ima_main_mpi.cpp:
  // 1808-1814), nor without split times to move
  if (d.prior < 0 || d.P < 2 || c->models[d.prior].nmig == 0) return IMA_OK;
  const int N = m.C * m.L, P = d.P;
  int emax = d.emax, maxn = 2, rc;
  size_t lfn = 4;
  {
    std::vector<int32_t> sl(N), cp(N);
    for (int i = 0; i < N; ++i) {
      sl[i] = m.first + i;
      cp[i] = 2 * c->slots[m.first + i].nmig + kNwRoom;
    }
    if ((rc = grow_mig_caps(c, sl.data(), cp.data(), N)) != IMA_OK) return rc;
  }
  for (int ch = 0; ch < m.C; ++ch) {
    size_t tot = 4;
    for (int l = 0; l < m.L; ++l) {
      const int sl = m.first + ch * m.L + l;
      const SlotMeta &s = c->slots[sl];
      emax = std::max(emax, (s.n - 1) + s.mig_cap + P - 1);
      maxn = std::max(maxn, s.n);
      tot += (size_t)s.n - 1 + (size_t)s.mig_cap;
    }
    lfn = std::max(lfn, tot);
  }
  if (treeweight_lds_bytes(P, emax) > (size_t)c->max_lds)
    return fail(c, IMA_ERR_SIZE, "room for NW's migrations: %d prior events exceed LDS", emax);
  if (nw_lds_bytes(maxn, 0, 0) > (size_t)c->max_lds)
    return fail(c, IMA_ERR_SIZE, "the NW update of %d gene copies needs more LDS than %d B", maxn,
                c->max_lds);
  int maxcap = 0;
  for (int i = 0; i < N; ++i) maxcap = std::max(maxcap, c->slots[m.first + i].mig_cap);
  grow_logfact(c, lfn);
  HIPCHK(c, reserve(m.nwk, m.C));
  HIPCHK(c, reserve(m.nwt, m.C));
  HIPCHK(c, reserve(m.nwst, N));
  HIPCHK(c, reserve(m.nwnbk, N));
  HIPCHK(c, reserve(m.nwmig, N));
  // no chain has an NW move pending
  HIPCHK(c, hipMemsetAsync(m.nwk.p, 0xff, (size_t)m.C * 4, c->stream));
  d.emax = emax;
  d.nwk = m.nwk.p;
  d.nwt = m.nwt.p;
  NwArgs &a = d.nwa;
  a = NwArgs{};
  a.count = N;
  a.first_slot = m.first;
  a.row = m.tv_row.p;
  a.tv = m.tvbackup.p;
  a.tw = P - 1;
  a.period = m.nwk.p;
  a.newt = m.nwt.p;
  a.seed = d.seed ^ 0x4E57A1C3B2D5E6F7ull;
  a.counter = m.counter.p;
  a.model = d.prior;
  a.emax = emax;
  a.max_n = maxn;
  a.max_cap = maxcap;
  a.stage = nw_lds_bytes(maxn, maxcap, 1) <= (size_t)c->max_lds ? 1 : 0;
  a.nmig_bk = m.nwnbk.p;
  a.migweight = m.nwmig.p;
  a.status = m.nwst.p;
  d.nw = 1;
  mc3_drop_graphs(c);
  return mc3_nw_buffers(c);
}

// After a run with the NW update: the host mirror of the chains' slot
// records, populations, per-edge migration indexes and events follows the
// device.
static int mc3_nw_sync(ima_ctx *c) {
  auto &m = c->mc3;
  const int N = m.C * m.L;
  const SlotMeta &s0 = c->slots[m.first];
  const SlotMeta &s1 = c->slots[m.first + N - 1];
  const size_t nb = s1.node_off + 2 * (size_t)s1.n - 1 - s0.node_off;
  const size_t o = s0.node_off;
  HIPCHK(c, copy_now(c, c->slots.data() + m.first, c->d_slots.p + m.first, N * sizeof(SlotMeta),
                     hipMemcpyDeviceToHost));
  HIPCHK(c, copy_now(c, c->pop.data() + o, c->d_pop.p + o, nb * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, copy_now(c, c->mig_first.data() + o, c->d_mig_first.p + o, nb * 4,
                     hipMemcpyDeviceToHost));
  HIPCHK(c, copy_now(c, c->mig_count.data() + o, c->d_mig_count.p + o, nb * 4,
                     hipMemcpyDeviceToHost));
  if (c->migs) {
    HIPCHK(c, copy_now(c, c->mig_edge.data(), c->d_mig_edge.p, c->migs * 4,
                       hipMemcpyDeviceToHost));
    HIPCHK(c, copy_now(c, c->mig_to.data(), c->d_mig_to.p, c->migs * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, copy_now(c, c->mig_time.data(), c->d_mig_time.p, c->migs * 8,
                       hipMemcpyDeviceToHost));
  }
  return IMA_OK;
}

// After a run with the device's own genealogy move: the host mirror of the
// chains' slots follows the device (topologies, times, populations,
// alleles, programs and slot records).
static int mc3_topo_sync(ima_ctx *c) {
  auto &m = c->mc3;
  const int N = m.C * m.L;
  const SlotMeta &s0 = c->slots[m.first];
  const SlotMeta &s1 = c->slots[m.first + N - 1];
  const size_t nb = s1.node_off + 2 * (size_t)s1.n - 1 - s0.node_off;
  const size_t pb = s1.prog_off + (size_t)s1.n - 1 - s0.prog_off;
  const size_t o = s0.node_off;
  HIPCHK(c, copy_now(c, c->slots.data() + m.first, c->d_slots.p + m.first, N * sizeof(SlotMeta),
                     hipMemcpyDeviceToHost));
  HIPCHK(c, copy_now(c, c->up0.data() + o, c->d_up0.p + o, nb * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, copy_now(c, c->up1.data() + o, c->d_up1.p + o, nb * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, copy_now(c, c->down.data() + o, c->d_down.p + o, nb * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, copy_now(c, c->time.data() + o, c->d_time.p + o, nb * 8, hipMemcpyDeviceToHost));
  HIPCHK(c, copy_now(c, c->pop.data() + o, c->d_pop.p + o, nb * 4, hipMemcpyDeviceToHost));
  HIPCHK(c, copy_now(c, c->A.data() + o, c->d_A.p + o, nb * 4, hipMemcpyDeviceToHost));
  if (pb)
    HIPCHK(c, copy_now(c, c->prog.data() + s0.prog_off, c->d_prog.p + s0.prog_off,
                       pb * sizeof(ProgOp), hipMemcpyDeviceToHost));
  if (m.dev.mig_pool > 0) return mc3_nw_sync(c);  // the migrations the moves drew
  return IMA_OK;
}

int mc3_run_async(ima_ctx *c, int nsweeps, int flags) {
  if (!c || nsweeps < 0) return IMA_ERR_ARG;
  auto &m = c->mc3;
  if (!m.live) return fail(c, IMA_ERR_STATE, "ima_mc3_setup has not been called");
  hipSetDevice(c->device);
  int rc;
  const bool use_graph = (flags & IMA_MC3_GRAPH) != 0;
  const bool scale = (flags & IMA_MC3_SCALE) != 0;
  const bool timed = (flags & IMA_MC3_TIMED) != 0;
  const bool userial = (flags & IMA_MC3_U_SERIAL) != 0;
  // changeu needs two loci; changekappa a lone HKY locus (every item HKY)
  const bool urate = (flags & IMA_MC3_UPDATE_U) && (m.L >= 2 || m.nh == m.C * m.L);
  auto u_due = [&](uint64_t q) { return urate && q % 5 == 4; };
  const int ry1 = (flags & IMA_MC3_RY1) ? 1 : 0;
  if (ry1 != m.dev.ry1) {
    mc3_drop_graphs(c);  // the graphs hold the sweep's parameters
    m.dev.ry1 = ry1;
  }
  // the device's own genealogy move: proposals are made at every sweep's
  // start, nothing is left pending between sweeps
  const bool topo = (flags & IMA_MC3_TOPOLOGY) != 0;
  if ((rc = mc3_topo_mode(c, topo)) != IMA_OK) return rc;
  // NW needs the split-time update, so the rescaling sweeps
  if ((rc = mc3_nw_mode(c, (flags & IMA_MC3_NW) && (flags & IMA_MC3_SCALE))) != IMA_OK)
    return rc;
  if ((rc = upload_slot_meta(c)) != IMA_OK) return rc;
  if ((rc = upload_models(c)) != IMA_OK) return rc;
  {
    // captured graphs hold the device buffers of their capture time
    const DeviceView cur = view(c);
    if (m.gview_set && memcmp(&cur, &m.gview, sizeof cur) != 0) mc3_drop_graphs(c);
    m.gview = cur;
    m.gview_set = true;
  }
  const int bk = scale ? 1 : 0;
  // n sweeps from one of the given parity; odd sweeps rescale when sc
  auto capture = [&](hipGraphExec_t *exec, hipGraph_t *graph, int n, int parity,
                     bool sc) -> int {
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int r = IMA_OK;
    for (int s = 0; s < n && r == IMA_OK; ++s)
      r = mc3_enqueue(c, sc && ((parity + s) & 1), nullptr);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (r != IMA_OK) return r;
    HIPCHK(c, e);
    *graph = g;
    HIPCHK(c, hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return IMA_OK;
  };
  if (use_graph && !timed) {
    for (int k = 0; k < 2; ++k)
      if (!m.exec[k] && (rc = capture(&m.exec[k], &m.graph[k], 1, k, true)) != IMA_OK) return rc;
    if (!urate && !m.bexec[bk] && nsweeps >= kSweepBlock &&
        (rc = capture(&m.bexec[bk], &m.bgraph[bk], kSweepBlock, 0, scale)) != IMA_OK)
      return rc;
    if (urate && nsweeps >= kURateBlock)
      for (int p = 0; p < 2; ++p)
        if (!m.uexec[bk][p] &&
            (rc = capture(&m.uexec[bk][p], &m.ugraph[bk][p], kURateBlock, p, scale)) != IMA_OK)
          return rc;
  }
  const uint64_t done = m.host_sweeps;
  bool pending = false;  // the next sweep's genealogy proposals are in place
  for (int s = 0; s < nsweeps;) {
    const uint64_t q = done + s;
    if (!pending && !topo) {
      HIPCHK(c, enqueue_mc3_propose(view(c), m.dev, c->stream));
      pending = true;
    }
    int ran = 0;
    if (use_graph && !timed) {
      if (!urate && m.bexec[bk] && (q & 1) == 0 && nsweeps - s >= kSweepBlock) {
        // a block of sweeps in one launch
        HIPCHK(c, hipGraphLaunch(m.bexec[bk], c->stream));
        m.scale_sweeps += scale ? kSweepBlock / 2 : 0;
        ran = kSweepBlock;
      } else if (urate && m.uexec[bk][q & 1] && nsweeps - s >= kURateBlock) {
        bool clear = true;  // no round after any but the block's last sweep
        for (int i = 0; i + 1 < kURateBlock; ++i) clear = clear && !u_due(q + i);
        if (clear) {
          HIPCHK(c, hipGraphLaunch(m.uexec[bk][q & 1], c->stream));
          for (int i = 0; i < kURateBlock; ++i) m.scale_sweeps += (scale && ((q + i) & 1)) ? 1 : 0;
          ran = kURateBlock;
        }
      }
    }
    if (!ran) {
      ran = 1;
      const bool with_scale = scale && (q & 1);
      m.scale_sweeps += with_scale ? 1 : 0;
      if (timed) {
        // events bracket the whole sweep (and its mutation-rate round) and
        // each batched scoring launch
        Timing &T = c->timing;
        HIPCHK(c, hipEventRecord(T.ev0[K_MC3], c->stream));
        if ((rc = mc3_enqueue(c, with_scale, m.ev)) != IMA_OK) return rc;
        if (u_due(q)) {
          if (pending) HIPCHK(c, enqueue_mc3_revert(view(c), m.dev, c->stream));
          pending = false;
          if ((rc = mc3_urate_round(c, q + 1, userial)) != IMA_OK) return rc;
        }
        HIPCHK(c, hipEventRecord(T.ev1[K_MC3], c->stream));
        HIPCHK(c, hipEventSynchronize(T.ev1[K_MC3]));
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, T.ev0[K_MC3], T.ev1[K_MC3]));
        T.ms[K_MC3] += ms;
        T.launches[K_MC3] += 1;
        const int ks = m.precision == IMA_PREC_FP32 ? K_HKY32 : K_HKY64;
        HIPCHK(c, hipEventElapsedTime(&ms, m.ev[0], m.ev[1]));
        T.ms[ks] += ms;
        T.launches[ks] += 1;
        if (with_scale) {
          HIPCHK(c, hipEventElapsedTime(&ms, m.ev[2], m.ev[3]));
          T.ms[ks] += ms;
          T.launches[ks] += 1;
        }
        // the accept step: from the end of the first scoring pass to the
        // start of the second (rescaling sweeps) or to the sweep's end
        // (others, without a mutation-rate round)
        if (with_scale || !u_due(q)) {
          HIPCHK(c, hipEventElapsedTime(&ms, m.ev[1], with_scale ? m.ev[2] : T.ev1[K_MC3]));
          T.ms[K_ACCEPT] += ms;
          T.launches[K_ACCEPT] += 1;
        }
      } else if (use_graph) {
        HIPCHK(c, hipGraphLaunch(m.exec[with_scale ? 1 : 0], c->stream));
      } else {
        if ((rc = mc3_enqueue(c, with_scale, nullptr)) != IMA_OK) return rc;
      }
    }
    s += ran;
    const uint64_t qlast = done + s - 1;
    if (u_due(qlast) && (pending || topo)) {
      // the round scores the committed genealogies: take the next sweep's
      // proposals back, and make them again (the same ones) afterwards
      if (pending) HIPCHK(c, enqueue_mc3_revert(view(c), m.dev, c->stream));
      pending = false;
      if ((rc = mc3_urate_round(c, qlast + 1, userial)) != IMA_OK) return rc;
    }
  }
  if (pending) HIPCHK(c, enqueue_mc3_revert(view(c), m.dev, c->stream));
  m.host_sweeps += (uint64_t)nsweeps;
  return IMA_OK;
}

extern "C" {

int ima_mc3_run(ima_ctx *c, int nsweeps, int flags) {
  int rc = mc3_run_async(c, nsweeps, flags);
  if (rc != IMA_OK) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->mc3.dev.topo) return mc3_topo_sync(c);
  if (c->mc3.dev.nw) return mc3_nw_sync(c);
  return IMA_OK;
}

int ima_mc3_read_rates(ima_ctx *c, double *u, double *kappa) {
  if (!c) return IMA_ERR_ARG;
  auto &m = c->mc3;
  if (!m.live) return fail(c, IMA_ERR_STATE, "ima_mc3_setup has not been called");
  hipSetDevice(c->device);
  const size_t N = (size_t)m.C * m.L;
  std::vector<ima_item> it(N);
  HIPCHK(c, copy_now(c, it.data(), m.items.p, N * sizeof(ima_item), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < N; ++i) {
    if (u) u[i] = it[i].u;
    if (kappa) kappa[i] = it[i].kappa;
  }
  return IMA_OK;
}

int ima_mc3_read(ima_ctx *c, double *pdg, double *beta, uint64_t *stats, uint64_t *sweeps) {
  if (!c) return IMA_ERR_ARG;
  auto &m = c->mc3;
  if (!m.live) return fail(c, IMA_ERR_STATE, "ima_mc3_setup has not been called");
  hipSetDevice(c->device);
  if (pdg)
    HIPCHK(c, copy_now(c, pdg, m.pdg.p, (size_t)m.C * m.L * 8, hipMemcpyDeviceToHost));
  if (beta) HIPCHK(c, copy_now(c, beta, m.beta.p, (size_t)m.C * 8, hipMemcpyDeviceToHost));
  if (stats) {
    HIPCHK(c, copy_now(c, stats, m.stats.p, 8 * 8, hipMemcpyDeviceToHost));
    std::vector<uint32_t> acc((size_t)m.C * m.L), cacc(m.C);
    HIPCHK(c, copy_now(c, acc.data(), m.acc.p, acc.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(c, copy_now(c, cacc.data(), m.cacc.p, cacc.size() * 4, hipMemcpyDeviceToHost));
    uint64_t a = 0, ca = 0;
    for (uint32_t x : acc) a += x;
    for (uint32_t x : cacc) ca += x;
    stats[1] = a;
    stats[2] = (uint64_t)m.scale_sweeps * m.C;
    stats[3] = ca;
  }
  if (sweeps) HIPCHK(c, copy_now(c, sweeps, m.counter.p, 8, hipMemcpyDeviceToHost));
  return IMA_OK;
}

// Dynamic LDS of the sweep's per-chain prior kernels (accept, rescale
// commit) at D loci per group, with the swap's static staging.
static size_t prior_kernel_lds(int R, int np, int D, int mints) {
  return mc3_prior_lds_bytes(R, np, D, mints) + 2 * (size_t)kSwapLds * 8 + 2048;
}

// Largest accept group size 1..kSpecDepth whose LDS fits max_lds; 0 if none.
static int prior_group_plan(int R, int np, int mints, int max_lds) {
  for (int D = kSpecDepth; D >= 1; --D)
    if (prior_kernel_lds(R, np, D, mints) <= (size_t)max_lds) return D;
  return 0;
}

int ima_mc3_prior_plan(int npops, int npsize, int ntheta_wp, int nmig, int nmig_wp,
                       int *loci, int64_t *lds_bytes) {
  if (npops < 1 || npops > kMaxPops || npsize < 0 || nmig < 0 || ntheta_wp < 0 || nmig_wp < 0 ||
      !loci)
    return IMA_ERR_ARG;
  const int P = npops, T = 2 * P - 1, R = prior_rec_len(P), np = npsize + nmig;
  // the model's int block as ima_model_add lays it out
  const int mints = P * P + P + 2 * T + 3 * P + P * T + (npsize + 1) + 2 * ntheta_wp +
                    (nmig + 1) + 3 * nmig_wp;
  const int D = prior_group_plan(R, np, mints, kMaxLaunchLds);
  *loci = D;
  if (lds_bytes) *lds_bytes = (int64_t)prior_kernel_lds(R, np, D > 0 ? D : 1, mints);
  return D > 0 ? IMA_OK : IMA_ERR_SIZE;
}

int ima_mc3_set_prior(ima_ctx *c, const double *tvals, double tmax) {
  if (!c) return IMA_ERR_ARG;
  auto &m = c->mc3;
  if (!m.live) return fail(c, IMA_ERR_STATE, "ima_mc3_setup has not been called");
  if (!(tmax > 0)) return fail(c, IMA_ERR_ARG, "tmax must be > 0");
  hipSetDevice(c->device);
  const int N = m.C * m.L;
  int model = -1, emax = 1;
  for (int i = 0; i < N; ++i) {
    const SlotMeta &s = c->slots[m.first + i];
    const LocusMeta &lm = c->loci[s.locus];
    if (lm.pmodel < 0)
      return fail(c, IMA_ERR_STATE, "MC3 slot %d: locus has no population model", m.first + i);
    if (model < 0) model = lm.pmodel;
    if (lm.pmodel != model) return fail(c, IMA_ERR_ARG, "MC3 loci must share one population model");
  }
  const ModelMeta &mm = c->models[model];
  const int P = mm.npops, tw = P > 1 ? P - 1 : 1, R = prior_rec_len(P);
  const int np = mm.npsize + mm.nmig;
  if (P > 1 && !tvals) return IMA_ERR_ARG;
  for (int ch = 0; ch < m.C && P > 1; ++ch)
    for (int k = 0; k < P - 1; ++k) {
      const double t = tvals[ch * (P - 1) + k];
      if (!(t > 0) || t > tmax || (k > 0 && t < tvals[ch * (P - 1) + k - 1]))
        return fail(c, IMA_ERR_ARG, "chain %d: split times must be increasing in (0, tmax]", ch);
    }
  size_t lfn = 4;
  for (int ch = 0; ch < m.C; ++ch) {
    size_t tot = 4;
    for (int l = 0; l < m.L; ++l) {
      const SlotMeta &s = c->slots[m.first + ch * m.L + l];
      tot += (size_t)s.n - 1 + (size_t)s.nmig;
      emax = std::max(emax, (s.n - 1) + s.nmig + P - 1);
    }
    lfn = std::max(lfn, tot);
  }
  if (treeweight_lds_bytes(P, emax) > (size_t)c->max_lds)
    return fail(c, IMA_ERR_SIZE, "a genealogy has %d prior events", emax);
  // per-chain kernels: loci per speculative group, 2^D - 1 scenarios per parameter
  const int pchunk = prior_group_plan(R, np, mm.nints, c->max_lds);
  if (pchunk < 1) return fail(c, IMA_ERR_SIZE, "prior state of %d populations exceeds LDS", P);
  grow_logfact(c, lfn);
  int rc;
  if ((rc = upload_loci(c)) != IMA_OK) return rc;
  if ((rc = upload_slot_meta(c)) != IMA_OK) return rc;
  if ((rc = upload_models(c)) != IMA_OK) return rc;
  std::vector<int32_t> row(N), sl(N);
  for (int i = 0; i < N; ++i) {
    row[i] = i / m.L;
    sl[i] = m.first + i;
  }
  std::vector<double> tv((size_t)m.C * tw, 0.0);
  if (P > 1) std::copy(tvals, tvals + (size_t)m.C * tw, tv.begin());
  HIPCHK(c, reserve(m.tv, tv.size()));
  HIPCHK(c, reserve(m.tvbackup, tv.size()));
  HIPCHK(c, reserve(m.wrec, (size_t)N * R));
  HIPCHK(c, reserve(m.wnew, (size_t)N * R));
  HIPCHK(c, reserve(m.wall, (size_t)m.C * R));
  HIPCHK(c, reserve(m.qterm, (size_t)m.C * std::max(np, 1)));
  HIPCHK(c, reserve(m.qf, (size_t)m.C * std::max(np, 1)));
  HIPCHK(c, reserve(m.qc, (size_t)m.C * std::max(np, 1)));
  HIPCHK(c, reserve(m.probg, m.C));
  HIPCHK(c, reserve(m.mtbackup, c->migs + 1));
  HIPCHK(c, reserve(m.tv_row, N));
  HIPCHK(c, reserve(m.tw_slots, N));
  HIPCHK(c, reserve(m.wstatus, N));
  HIPCHK(c, hipMemcpyAsync(m.tv.p, tv.data(), tv.size() * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(m.tv_row.p, row.data(), N * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(m.tw_slots.p, sl.data(), N * 4, hipMemcpyHostToDevice, c->stream));
  MC3Dev &d = m.dev;
  d.prior = model;
  d.P = P;
  d.R = R;
  d.nparam = np;
  d.emax = emax;
  d.nw = 0;  // NW's room for new migrations is made again on its next use
  if (d.topo) {
    // the genealogy move is set up again for this model on its next use
    d.topo = 0;
    d.mig_pool = 0;
    d.hast = nullptr;
  }
  d.tmax = tmax;
  d.tv = m.tv.p;
  d.tvbackup = m.tvbackup.p;
  d.wrec = m.wrec.p;
  d.wnew = m.wnew.p;
  d.wall = m.wall.p;
  d.qterm = m.qterm.p;
  d.qf = m.qf.p;
  d.qc = m.qc.p;
  d.probg = m.probg.p;
  d.mtbackup = m.mtbackup.p;
  d.tv_row = m.tv_row.p;
  d.tw_slots = m.tw_slots.p;
  d.wstatus = m.wstatus.p;
  d.pchunk = pchunk;
  d.mints = mm.nints;
  HIPCHK(c, enqueue_mc3_prior_init(view(c), d, c->stream));
  std::vector<int32_t> st(N);
  HIPCHK(c, hipMemcpyAsync(st.data(), m.wstatus.p, N * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < N; ++i)
    if (st[i] != IMA_STATUS_OK) {
      d.prior = -1;
      return fail(c, IMA_ERR_ARG, "MC3 slot %d: population labels or migrations do not fit the model",
                  m.first + i);
    }
  // graphs captured before hold the sweep without the prior
  mc3_drop_graphs(c);
  return IMA_OK;
}

int ima_mc3_set_analyses(ima_ctx *c, int chains_per_analysis) {
  if (!c) return IMA_ERR_ARG;
  auto &m = c->mc3;
  if (!m.live) return fail(c, IMA_ERR_STATE, "ima_mc3_setup has not been called");
  if (chains_per_analysis < 1 || m.C % chains_per_analysis != 0)
    return fail(c, IMA_ERR_ARG, "%d chains do not split into analyses of %d", m.C,
                chains_per_analysis);
  m.dev.group = chains_per_analysis;
  mc3_drop_graphs(c);
  return IMA_OK;
}

int ima_mc3_set_prior_group(ima_ctx *c, int loci) {
  if (!c) return IMA_ERR_ARG;
  auto &m = c->mc3;
  if (!m.live || m.dev.prior < 0) return fail(c, IMA_ERR_STATE, "the sweep has no prior");
  if (loci < 1 || loci > kSpecDepth)
    return fail(c, IMA_ERR_ARG, "loci per group must be in 1..%d", kSpecDepth);
  if (prior_kernel_lds(m.dev.R, m.dev.nparam, loci, m.dev.mints) > (size_t)c->max_lds)
    return fail(c, IMA_ERR_SIZE, "%d loci per group exceed LDS", loci);
  m.dev.pchunk = loci;
  mc3_drop_graphs(c);
  return IMA_OK;
}

// The sweep's persistent device state (what one run leaves for the next),
// copied to and from one device buffer on the context's stream.
namespace {
struct Region {
  void *p;
  size_t bytes;
};
std::vector<Region> mc3_state_regions(ima_ctx *c) {
  auto &m = c->mc3;
  const MC3Dev &d = m.dev;
  const size_t N = (size_t)m.C * m.L;
  std::vector<Region> r;
  auto add = [&](void *p, size_t bytes) {
    if (p && bytes) r.push_back({p, bytes});
  };
  add(c->d_time.p + d.node_first, (size_t)d.nnodes * 8);
  add(c->d_mig_time.p, c->migs * 8);
  add(m.pdg.p, N * 8);
  add(m.beta.p, (size_t)m.C * 8);
  add(m.counter.p, 8);
  add(m.stats.p, 8 * 8);
  add(m.acc.p, N * 4);
  add(m.cacc.p, (size_t)m.C * 4);
  add(m.items.p, N * sizeof(ima_item));  // u and kappa, moved by the rate rounds
  if (d.topo) {
    // the device's own genealogy move changes topologies, labels and alleles
    const int64_t nf = d.node_first;
    add(c->d_up0.p + nf, (size_t)d.nnodes * 4);
    add(c->d_up1.p + nf, (size_t)d.nnodes * 4);
    add(c->d_down.p + nf, (size_t)d.nnodes * 4);
    add(c->d_pop.p + nf, (size_t)d.nnodes * 4);
    add(c->d_A.p + nf, (size_t)d.nnodes * 4);
    const SlotMeta &s0 = c->slots[m.first], &s1 = c->slots[m.first + N - 1];
    add(c->d_prog.p + s0.prog_off, (s1.prog_off + s1.n - 1 - s0.prog_off) * sizeof(ProgOp));
    add(c->d_slots.p + m.first, N * sizeof(SlotMeta));
  }
  if (d.nw || d.mig_pool > 0) {
    // NW and the move with migration rewrite populations, per-edge indexes,
    // events and the slots' event counts
    const int64_t nf = d.node_first;
    if (!d.topo) {
      add(c->d_pop.p + nf, (size_t)d.nnodes * 4);
      add(c->d_slots.p + m.first, N * sizeof(SlotMeta));
    }
    add(c->d_mig_first.p + nf, (size_t)d.nnodes * 4);
    add(c->d_mig_count.p + nf, (size_t)d.nnodes * 4);
    add(c->d_mig_edge.p, c->migs * 4);
    add(c->d_mig_to.p, c->migs * 4);
  }
  if (d.prior >= 0) {
    const size_t np = (size_t)std::max(d.nparam, 1);
    add(m.tv.p, (size_t)m.C * mc3_tw(d.P) * 8);
    add(m.wrec.p, N * d.R * 8);
    add(m.wall.p, (size_t)m.C * d.R * 8);
    add(m.qterm.p, (size_t)m.C * np * 8);
    add(m.qf.p, (size_t)m.C * np * 8);
    add(m.qc.p, (size_t)m.C * np * 4);
    add(m.probg.p, (size_t)m.C * 8);
  }
  return r;
}
int mc3_state_copy(ima_ctx *c, char *buf, bool save) {
  size_t off = 0;
  for (const Region &g : mc3_state_regions(c)) {
    HIPCHK(c, hipMemcpyAsync(save ? buf + off : g.p, save ? g.p : buf + off, g.bytes,
                             hipMemcpyDeviceToDevice, c->stream));
    off += (g.bytes + 255) & ~(size_t)255;
  }
  return IMA_OK;
}
size_t mc3_state_bytes(ima_ctx *c) {
  size_t off = 0;
  for (const Region &g : mc3_state_regions(c)) off += (g.bytes + 255) & ~(size_t)255;
  return off;
}
}  // namespace

int ima_mc3_tune(ima_ctx *c, int sweeps, int flags, int *best, double *ms_per_sweep) {
  if (!c || sweeps < 1) return IMA_ERR_ARG;
  // the sweeps the caller will run, replayed from captured graphs
  const int run_flags = (flags & ~(IMA_MC3_TIMED | IMA_MC3_U_SERIAL)) | IMA_MC3_GRAPH;
  auto &m = c->mc3;
  if (!m.live) return fail(c, IMA_ERR_STATE, "ima_mc3_setup has not been called");
  hipSetDevice(c->device);
  for (int k = 0; k < kSpecDepth; ++k)
    if (ms_per_sweep) ms_per_sweep[k] = -1.0;
  if (m.dev.prior < 0) {  // nothing to tune: the accept step has no groups
    if (best) *best = 0;
    return IMA_OK;
  }
  // the state saved and restored is that of the sweeps to be timed
  int rc0 = mc3_topo_mode(c, (run_flags & IMA_MC3_TOPOLOGY) != 0);
  if (rc0 != IMA_OK) return rc0;
  rc0 = mc3_nw_mode(c, (run_flags & IMA_MC3_NW) && (run_flags & IMA_MC3_SCALE));
  if (rc0 != IMA_OK) return rc0;
  // slot records (room for migrations, moved events) on the device first
  if ((rc0 = upload_slot_meta(c)) != IMA_OK) return rc0;
  const size_t bytes = mc3_state_bytes(c);
  char *buf = nullptr;
  HIPCHK(c, hipMalloc(&buf, bytes));
  const uint64_t hs = m.host_sweeps;
  const int64_t ss = m.scale_sweeps;
  const int keep = m.dev.pchunk;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = mc3_state_copy(c, buf, true);
  if (rc == IMA_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(c, IMA_ERR_HIP, "event creation failed");
  int pick = keep;
  float pick_ms = 0.0f;
  for (int D = 1; D <= kSpecDepth && rc == IMA_OK; ++D) {
    if (ima_mc3_set_prior_group(c, D) != IMA_OK) break;  // larger groups do not fit LDS
    float ms = 0.0f;
    for (int pass = 0; pass < 2 && rc == IMA_OK; ++pass) {  // capture + warm, then timed
      rc = mc3_state_copy(c, buf, false);
      m.host_sweeps = hs;
      m.scale_sweeps = ss;
      const int n = pass == 0 ? 2 : sweeps;
      if (rc == IMA_OK && pass == 1 && hipEventRecord(e0, c->stream) != hipSuccess)
        rc = fail(c, IMA_ERR_HIP, "event record failed");
      if (rc == IMA_OK) rc = mc3_run_async(c, n, run_flags);
      if (rc == IMA_OK && pass == 1) {
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
          rc = fail(c, IMA_ERR_HIP, "timing failed");
      }
    }
    if (rc != IMA_OK) break;
    if (ms_per_sweep) ms_per_sweep[D - 1] = ms / sweeps;
    if (D == 1 || ms < pick_ms) {
      pick = D;
      pick_ms = ms;
    }
  }
  if (rc == IMA_OK) rc = ima_mc3_set_prior_group(c, pick);
  // HKY site blocks of 64 or 128 patterns (one or two per lane), timed at
  // the chosen group size unless the context fixes the block size
  if (rc == IMA_OK && m.nh && c->site_block == 0 &&
      hky_lds_bytes(m.precision, m.hmax_n, m.hmax_slots, 2) <= (size_t)c->max_lds) {
    const int ppl0 = m.plan.ppl;
    int best_ppl = ppl0;
    float best_ms = pick_ms;
    for (int ppl = 1; ppl <= 2 && rc == IMA_OK; ++ppl) {
      if (ppl == ppl0) continue;
      rc = mc3_build_plan(c, ppl);
      float ms = 0.0f;
      for (int pass = 0; pass < 2 && rc == IMA_OK; ++pass) {
        rc = mc3_state_copy(c, buf, false);
        m.host_sweeps = hs;
        m.scale_sweeps = ss;
        const int n = pass == 0 ? 2 : sweeps;
        if (rc == IMA_OK && pass == 1 && hipEventRecord(e0, c->stream) != hipSuccess)
          rc = fail(c, IMA_ERR_HIP, "event record failed");
        if (rc == IMA_OK) rc = mc3_run_async(c, n, run_flags);
        if (rc == IMA_OK && pass == 1) {
          if (hipEventRecord(e1, c->stream) != hipSuccess ||
              hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(c, IMA_ERR_HIP, "timing failed");
        }
      }
      if (rc == IMA_OK && ms < best_ms) {
        best_ms = ms;
        best_ppl = ppl;
      }
    }
    if (rc == IMA_OK && m.plan.ppl != best_ppl) rc = mc3_build_plan(c, best_ppl);
  }
  // the chains go back to where they were, with the fastest shapes
  if (rc == IMA_OK) rc = mc3_state_copy(c, buf, false);
  m.host_sweeps = hs;
  m.scale_sweeps = ss;
  if (hipStreamSynchronize(c->stream) != hipSuccess && rc == IMA_OK)
    rc = fail(c, IMA_ERR_HIP, "synchronize failed");
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipFree(buf);
  if (rc == IMA_OK && best) *best = pick;
  return rc;
}

int ima_set_site_block(ima_ctx *c, int patterns) {
  if (!c || (patterns != 0 && patterns != 64 && patterns != 128)) return IMA_ERR_ARG;
  c->site_block = patterns;
  return IMA_OK;
}

int ima_mc3_site_block(ima_ctx *c, int *patterns) {
  if (!c || !patterns) return IMA_ERR_ARG;
  if (!c->mc3.live) return fail(c, IMA_ERR_STATE, "ima_mc3_setup has not been called");
  *patterns = kWave * c->mc3.plan.ppl;
  return IMA_OK;
}

// Decisions for a batch of proposals written by the host (IMa2p's own
// updategenealogy proposals): score, decide each chain's loci in order,
// copy the accepted ones into the chains' slots.
int ima_mc3_decide(ima_ctx *c, int prop_first, const double *hastings, const double *uniforms,
                   double gbeta, int32_t *accepted) {
  if (!c) return IMA_ERR_ARG;
  auto &m = c->mc3;
  if (!m.live) return fail(c, IMA_ERR_STATE, "ima_mc3_setup has not been called");
  const int N = m.C * m.L;
  if (prop_first < 0 || prop_first + N > (int)c->slots.size())
    return fail(c, IMA_ERR_ARG, "proposal slots [%d,%d) out of range", prop_first, prop_first + N);
  if (prop_first < m.first + N && m.first < prop_first + N)
    return fail(c, IMA_ERR_ARG, "proposal slots overlap the chains' slots");
  if (!(gbeta >= 0.0) || !std::isfinite(gbeta)) return fail(c, IMA_ERR_ARG, "gbeta must be >= 0");
  hipSetDevice(c->device);
  const bool prior = m.dev.prior >= 0;
  const int P = m.dev.P;
  int hmax_slots = m.hmax_slots, emax = m.dev.emax;
  size_t lfn = 4;
  bool grew = false;
  for (int ch = 0; ch < m.C; ++ch) {
    size_t tot = 4;
    for (int l = 0; l < m.L; ++l) {
      const int i = ch * m.L + l;
      SlotMeta &d = c->slots[m.first + i];
      const SlotMeta &s = c->slots[prop_first + i];
      if (s.locus != d.locus)
        return fail(c, IMA_ERR_ARG, "proposal slot %d is not of its chain slot's locus", prop_first + i);
      if (s.root < 0) return fail(c, IMA_ERR_STATE, "proposal slot %d holds no genealogy", prop_first + i);
      if (c->loci[s.locus].model == IMA_MODEL_HKY) hmax_slots = std::max(hmax_slots, s.nslots);
      if (prior) {
        emax = std::max(emax, (s.n - 1) + s.nmig + P - 1);
        tot += (size_t)s.n - 1 + (size_t)std::max(s.nmig, d.nmig);
      }
      {
        if (s.nmig > d.mig_cap) {
          // the chain's slot must hold the proposal's migrations: move its
          // events to a larger region (the device copy keeps their times)
          const int cap = s.nmig + s.nmig / 2 + 4;
          const int64_t old = d.mig_off;
          d.mig_off = (int64_t)c->migs;
          d.mig_cap = cap;
          c->migs += (size_t)cap;
          c->mig_edge.resize(c->migs, 0);
          c->mig_to.resize(c->migs, 0);
          c->mig_time.resize(c->migs, 0.0);
          std::copy(c->mig_edge.begin() + old, c->mig_edge.begin() + old + d.nmig,
                    c->mig_edge.begin() + d.mig_off);
          std::copy(c->mig_to.begin() + old, c->mig_to.begin() + old + d.nmig,
                    c->mig_to.begin() + d.mig_off);
          HIPCHK(c, reserve(c->d_mig_edge, c->migs + 1, true, c->stream));
          HIPCHK(c, reserve(c->d_mig_to, c->migs + 1, true, c->stream));
          HIPCHK(c, reserve(c->d_mig_time, c->migs + 1, true, c->stream));
          if (d.nmig) {
            HIPCHK(c, hipMemcpyAsync(c->d_mig_edge.p + d.mig_off, c->d_mig_edge.p + old,
                                     (size_t)d.nmig * 4, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(c->d_mig_to.p + d.mig_off, c->d_mig_to.p + old,
                                     (size_t)d.nmig * 4, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(c->d_mig_time.p + d.mig_off, c->d_mig_time.p + old,
                                     (size_t)d.nmig * 8, hipMemcpyDeviceToDevice, c->stream));
          }
          c->slots_meta_dirty = true;
          grew = true;
        }
      }
    }
    lfn = std::max(lfn, tot);
  }
  if (prior && treeweight_lds_bytes(P, emax) > (size_t)c->max_lds)
    return fail(c, IMA_ERR_SIZE, "a proposal has %d prior events", emax);
  // per-edge migration indexes of every slot on the device (slots allocated
  // after the last ima_slots_write_pops have none there yet)
  if (c->d_mig_first.n < c->nodes || c->d_mig_count.n < c->nodes) {
    HIPCHK(c, reserve(c->d_mig_first, c->nodes, false, c->stream));
    HIPCHK(c, reserve(c->d_mig_count, c->nodes, false, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_mig_first.p, c->mig_first.data(), c->nodes * 4,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_mig_count.p, c->mig_count.data(), c->nodes * 4,
                             hipMemcpyHostToDevice, c->stream));
  }
  if (grew) {
    HIPCHK(c, reserve(m.mtbackup, c->migs + 1));
    m.dev.mtbackup = m.mtbackup.p;
  }
  if (emax != m.dev.emax || hmax_slots != m.hmax_slots || grew) {
    // the chains' slots may now hold these genealogies: later sweeps use
    // the larger bounds
    m.dev.emax = emax;
    m.hmax_slots = hmax_slots;
    mc3_drop_graphs(c);
  }
  grow_logfact(c, lfn);
  int rc;
  if ((rc = upload_slot_meta(c)) != IMA_OK) return rc;
  if ((rc = upload_models(c)) != IMA_OK) return rc;
  HIPCHK(c, reserve(m.idec, N));
  HIPCHK(c, reserve(m.ptw, N));
  HIPCHK(c, reserve(m.pitems, N));
  MC3Dev d = m.dev;
  d.host = 1;
  d.topo = 0;
  d.gbeta = gbeta;
  d.idec = m.idec.p;
  d.prop_first = prop_first;
  d.hast = nullptr;
  d.uhost = nullptr;
  if (hastings) {
    for (int i = 0; i < N; ++i)
      if (std::isnan(hastings[i])) return fail(c, IMA_ERR_ARG, "hastings[%d] is NaN", i);
    HIPCHK(c, reserve(m.hast, N));
    HIPCHK(c, hipMemcpyAsync(m.hast.p, hastings, (size_t)N * 8, hipMemcpyHostToDevice, c->stream));
    d.hast = m.hast.p;
  }
  if (uniforms) {
    for (int i = 0; i < N; ++i)
      if (!(uniforms[i] >= 0.0 && uniforms[i] <= 1.0))
        return fail(c, IMA_ERR_ARG, "uniforms[%d] outside [0, 1]", i);
    HIPCHK(c, reserve(m.uhost, N));
    HIPCHK(c, hipMemcpyAsync(m.uhost.p, uniforms, (size_t)N * 8, hipMemcpyHostToDevice, c->stream));
    d.uhost = m.uhost.p;
  }
  const DeviceView v = view(c);
  HIPCHK(c, enqueue_mc3_prop_items(d, m.items.p, m.pitems.p, m.ptw.p, c->stream));
  HkyPlan p = m.plan;
  p.nred = 0;  // the decision kernel merges pattern blocks itself
  HIPCHK(c, enqueue_mc3_score(v, d, m.pitems.p, m.ptw.p, m.idx.p, p, m.nh, m.ni, m.ns,
                              m.precision, m.hmax_n, hmax_slots, m.imax_n, m.imax_w, nullptr, 0,
                              c->stream, c->side, m.fj));
  HIPCHK(c, enqueue_mc3_decide(v, d, c->stream));
  std::vector<int32_t> dec(N);
  HIPCHK(c, copy_now(c, dec.data(), m.idec.p, (size_t)N * 4, hipMemcpyDeviceToHost));
  m.host_sweeps += 1;  // the device counter advanced with the decisions
  // the host mirror of the chains' slots follows the device copy
  for (int i = 0; i < N; ++i) {
    if (!dec[i]) continue;
    SlotMeta &dst = c->slots[m.first + i];
    const SlotMeta &src = c->slots[prop_first + i];
    const int64_t so = src.node_off, dn = dst.node_off, ne = 2 * (int64_t)src.n - 1;
    std::copy(c->up0.begin() + so, c->up0.begin() + so + ne, c->up0.begin() + dn);
    std::copy(c->up1.begin() + so, c->up1.begin() + so + ne, c->up1.begin() + dn);
    std::copy(c->down.begin() + so, c->down.begin() + so + ne, c->down.begin() + dn);
    std::copy(c->time.begin() + so, c->time.begin() + so + ne, c->time.begin() + dn);
    std::copy(c->A.begin() + so, c->A.begin() + so + ne, c->A.begin() + dn);
    std::copy(c->pop.begin() + so, c->pop.begin() + so + ne, c->pop.begin() + dn);
    std::copy(c->mig_first.begin() + so, c->mig_first.begin() + so + ne, c->mig_first.begin() + dn);
    std::copy(c->mig_count.begin() + so, c->mig_count.begin() + so + ne, c->mig_count.begin() + dn);
    std::copy(c->prog.begin() + src.prog_off, c->prog.begin() + src.prog_off + src.n - 1,
              c->prog.begin() + dst.prog_off);
    if (src.nmig) {
      std::copy(c->mig_edge.begin() + src.mig_off, c->mig_edge.begin() + src.mig_off + src.nmig,
                c->mig_edge.begin() + dst.mig_off);
      std::copy(c->mig_to.begin() + src.mig_off, c->mig_to.begin() + src.mig_off + src.nmig,
                c->mig_to.begin() + dst.mig_off);
      std::copy(c->mig_time.begin() + src.mig_off, c->mig_time.begin() + src.mig_off + src.nmig,
                c->mig_time.begin() + dst.mig_off);
    }
    dst.root = src.root;
    dst.nslots = src.nslots;
    dst.nmig = src.nmig;
  }
  if (accepted) std::copy(dec.begin(), dec.end(), accepted);
  return IMA_OK;
}

int ima_mc3_record(ima_ctx *c, int table_id) {
  if (!c) return IMA_ERR_ARG;
  auto &m = c->mc3;
  if (!m.live || m.dev.prior < 0) return fail(c, IMA_ERR_STATE, "the sweep has no prior");
  if (table_id < 0 || table_id >= (int)c->tables.size() || !c->tables[table_id].live)
    return fail(c, IMA_ERR_ARG, "table %d unknown", table_id);
  LTableHost &h = c->tables[table_id];
  const ModelMeta &mm = c->models[m.dev.prior];
  if (h.npsize != mm.npsize || h.nmig != mm.nmig || h.npops != mm.npops)
    return fail(c, IMA_ERR_ARG, "table %d does not have the sweep's parameters", table_id);
  if (h.G >= h.cap)
    return fail(c, IMA_ERR_SIZE, "table %d is full (%d genealogies)", table_id, h.cap);
  if (std::find(m.beta0.begin(), m.beta0.end(), 1.0) == m.beta0.end())
    return fail(c, IMA_ERR_STATE, "no chain has beta = 1 (no cold chain to record)");
  hipSetDevice(c->device);
  HIPCHK(c, enqueue_mc3_record(view(c), m.dev, h.rows + (size_t)h.G * h.K, c->stream));
  h.G += 1;
  return IMA_OK;
}

int ima_mc3_read_weights(ima_ctx *c, double *wall, double *qint) {
  if (!c) return IMA_ERR_ARG;
  auto &m = c->mc3;
  if (!m.live || m.dev.prior < 0) return fail(c, IMA_ERR_STATE, "the sweep has no prior");
  hipSetDevice(c->device);
  if (wall)
    HIPCHK(c, copy_now(c, wall, m.wall.p, (size_t)m.C * m.dev.R * 8, hipMemcpyDeviceToHost));
  if (qint && m.dev.nparam)
    HIPCHK(c, copy_now(c, qint, m.qterm.p, (size_t)m.C * m.dev.nparam * 8, hipMemcpyDeviceToHost));
  return IMA_OK;
}

int ima_mc3_prior_group(ima_ctx *c, int *loci) {
  if (!c || !loci) return IMA_ERR_ARG;
  if (!c->mc3.live || c->mc3.dev.prior < 0)
    return fail(c, IMA_ERR_STATE, "the sweep has no prior");
  *loci = c->mc3.dev.pchunk;
  return IMA_OK;
}

int ima_mc3_read_prior(ima_ctx *c, double *probg, double *tvals) {
  if (!c) return IMA_ERR_ARG;
  auto &m = c->mc3;
  if (!m.live || m.dev.prior < 0) return fail(c, IMA_ERR_STATE, "the sweep has no prior");
  hipSetDevice(c->device);
  if (probg) HIPCHK(c, copy_now(c, probg, m.probg.p, (size_t)m.C * 8, hipMemcpyDeviceToHost));
  if (tvals && m.dev.P > 1)
    HIPCHK(c, copy_now(c, tvals, m.tv.p, (size_t)m.C * (m.dev.P - 1) * 8, hipMemcpyDeviceToHost));
  return IMA_OK;
}

int ima_slots_read_mig(ima_ctx *c, int slot, double *mig_time) {
  if (!c || !mig_time || slot < 0 || slot >= (int)c->slots.size()) return IMA_ERR_ARG;
  const SlotMeta &s = c->slots[slot];
  if (s.nmig)
    HIPCHK(c, copy_now(c, mig_time, c->d_mig_time.p + s.mig_off, (size_t)s.nmig * 8,
                        hipMemcpyDeviceToHost));
  return IMA_OK;
}

int ima_slots_read_mig_events(ima_ctx *c, int slot, int32_t *count, int32_t *mig_edge,
                              double *mig_time, int32_t *mig_to) {
  if (!c || !count || slot < 0 || slot >= (int)c->slots.size()) return IMA_ERR_ARG;
  const SlotMeta &s = c->slots[slot];
  *count = s.nmig;
  if (!s.nmig || (!mig_edge && !mig_time && !mig_to)) return IMA_OK;
  if (mig_edge)
    HIPCHK(c, copy_now(c, mig_edge, c->d_mig_edge.p + s.mig_off, (size_t)s.nmig * 4,
                        hipMemcpyDeviceToHost));
  if (mig_time)
    HIPCHK(c, copy_now(c, mig_time, c->d_mig_time.p + s.mig_off, (size_t)s.nmig * 8,
                        hipMemcpyDeviceToHost));
  if (mig_to)
    HIPCHK(c, copy_now(c, mig_to, c->d_mig_to.p + s.mig_off, (size_t)s.nmig * 4,
                        hipMemcpyDeviceToHost));
  return IMA_OK;
}

// -------------------------------------------------------- genealogy prior --

int ima_model_add(ima_ctx *c, int npops, const int32_t *plist, const int32_t *npresent,
                  const int32_t *popdown, const int32_t *pope, const int32_t *join,
                  int npsize, const int32_t *theta_off, const int32_t *theta_wp, int nmig,
                  const int32_t *mig_off, const int32_t *mig_wp, const double *qmax,
                  const double *mmax, int *model) {
  if (!c || !plist || !npresent || !popdown || !pope || !join || !theta_off || !mig_off ||
      !model || npsize < 0 || nmig < 0 || (npsize && (!theta_wp || !qmax)) ||
      (nmig && (!mig_wp || !mmax)))
    return IMA_ERR_ARG;
  const int P = npops, T = 2 * npops - 1;
  if (P < 1 || P > kMaxPops) return fail(c, IMA_ERR_SIZE, "npops %d outside 1..%d", P, kMaxPops);
  for (int k = 0; k < P; ++k) {
    if (npresent[k] < 1 || npresent[k] > P)
      return fail(c, IMA_ERR_ARG, "npresent[%d] = %d", k, npresent[k]);
    for (int i = 0; i < npresent[k]; ++i)
      if (plist[k * P + i] < 0 || plist[k * P + i] >= T)
        return fail(c, IMA_ERR_ARG, "plist[%d][%d] = %d", k, i, plist[k * P + i]);
  }
  for (int q = 0; q < T; ++q)
    if (popdown[q] < -1 || popdown[q] >= T || pope[q] < -1 || pope[q] > P)
      return fail(c, IMA_ERR_ARG, "population %d: popdown/pope out of range", q);
  for (int k = 0; k + 1 < P; ++k)
    for (int x = 0; x < 3; ++x)
      if (join[3 * k + x] < 0 || join[3 * k + x] >= T)
        return fail(c, IMA_ERR_ARG, "join[%d] out of range", k);
  if (theta_off[0] != 0 || mig_off[0] != 0) return fail(c, IMA_ERR_ARG, "parameter maps must start at 0");
  for (int t = 0; t < npsize; ++t) {
    if (theta_off[t + 1] < theta_off[t]) return fail(c, IMA_ERR_ARG, "theta_off not monotone");
    if (!(qmax[t] > 0)) return fail(c, IMA_ERR_ARG, "qmax[%d] must be > 0", t);
    for (int w = theta_off[t]; w < theta_off[t + 1]; ++w) {
      const int k = theta_wp[2 * w], i = theta_wp[2 * w + 1];
      if (k < 0 || k >= P || i < 0 || i >= npresent[k])
        return fail(c, IMA_ERR_ARG, "theta parameter %d: bad entry (%d, %d)", t, k, i);
    }
  }
  for (int m = 0; m < nmig; ++m) {
    if (mig_off[m + 1] < mig_off[m]) return fail(c, IMA_ERR_ARG, "mig_off not monotone");
    if (!(mmax[m] > 0)) return fail(c, IMA_ERR_ARG, "mmax[%d] must be > 0", m);
    for (int w = mig_off[m]; w < mig_off[m + 1]; ++w) {
      const int k = mig_wp[3 * w], i = mig_wp[3 * w + 1], j = mig_wp[3 * w + 2];
      if (k < 0 || k >= P || i < 0 || i >= npresent[k] || j < 0 || j >= npresent[k] || i == j)
        return fail(c, IMA_ERR_ARG, "migration parameter %d: bad entry (%d, %d, %d)", m, k, i, j);
    }
  }
  ModelMeta mm{};
  mm.npops = P;
  mm.npsize = npsize;
  mm.nmig = nmig;
  mm.ioff = (int64_t)c->mint.size();
  mm.doff = (int64_t)c->mdbl.size();
  std::vector<int32_t> &I = c->mint;
  I.insert(I.end(), plist, plist + P * P);
  I.insert(I.end(), npresent, npresent + P);
  I.insert(I.end(), popdown, popdown + T);
  I.insert(I.end(), pope, pope + T);
  for (int k = 0; k < P; ++k)
    for (int x = 0; x < 3; ++x) I.push_back(k + 1 < P ? join[3 * k + x] : -1);
  for (int k = 0; k < P; ++k)
    for (int q = 0; q < T; ++q) {
      int pos = -1;
      for (int i = 0; i < npresent[k]; ++i)
        if (plist[k * P + i] == q) pos = i;
      I.push_back(pos);
    }
  I.insert(I.end(), theta_off, theta_off + npsize + 1);
  I.insert(I.end(), theta_wp, theta_wp + 2 * theta_off[npsize]);
  I.insert(I.end(), mig_off, mig_off + nmig + 1);
  I.insert(I.end(), mig_wp, mig_wp + 3 * mig_off[nmig]);
  c->mdbl.insert(c->mdbl.end(), qmax, qmax + npsize);
  c->mdbl.insert(c->mdbl.end(), mmax, mmax + nmig);
  mm.nints = (int32_t)((int64_t)I.size() - mm.ioff);
  *model = (int)c->models.size();
  c->models.push_back(mm);
  c->models_dirty = true;
  return IMA_OK;
}

int ima_locus_set_model(ima_ctx *c, int locus, int model, const int32_t *samppop, double hval) {
  if (!c || !samppop || locus < 0 || locus >= (int)c->loci.size() || model < 0 ||
      model >= (int)c->models.size() || !(hval > 0))
    return IMA_ERR_ARG;
  LocusMeta &lm = c->loci[locus];
  const int P = c->models[model].npops;
  int tot = 0;
  for (int p = 0; p < P; ++p) {
    if (samppop[p] < 0) return fail(c, IMA_ERR_ARG, "samppop[%d] < 0", p);
    tot += samppop[p];
  }
  if (tot != lm.n) return fail(c, IMA_ERR_ARG, "samppop sums to %d, locus has %d gene copies", tot, lm.n);
  lm.pmodel = model;
  for (int p = 0; p < kMaxPops; ++p) lm.samppop[p] = p < P ? samppop[p] : 0;
  lm.hval = hval;
  c->loci_dirty = true;
  return IMA_OK;
}

int ima_slots_write_pops(ima_ctx *c, int first, int count, const int32_t *pop,
                         const int32_t *nmig, const int32_t *mig_edge, const double *mig_time,
                         const int32_t *mig_to) {
  if (!c || !pop || !nmig) return IMA_ERR_ARG;
  if (first < 0 || count < 0 || first + count > (int)c->slots.size())
    return fail(c, IMA_ERR_ARG, "slot range [%d,%d) out of range", first, first + count);
  if (count == 0) return IMA_OK;
  hipSetDevice(c->device);
  size_t src = 0, msrc = 0;
  for (int i = 0; i < count; ++i) {
    SlotMeta &s = c->slots[first + i];
    const LocusMeta &lm = c->loci[s.locus];
    if (lm.pmodel < 0) return fail(c, IMA_ERR_STATE, "slot %d: its locus has no population model", first + i);
    const int T = 2 * c->models[lm.pmodel].npops - 1, ne = 2 * s.n - 1;
    if (nmig[i] < 0) return fail(c, IMA_ERR_ARG, "slot %d: nmig < 0", first + i);
    if (nmig[i] && (!mig_edge || !mig_time || !mig_to)) return IMA_ERR_ARG;
    for (int e = 0; e < ne; ++e)
      if (pop[src + e] < 0 || pop[src + e] >= T)
        return fail(c, IMA_ERR_ARG, "slot %d: edge %d in population %d", first + i, e, pop[src + e]);
    for (int j = 0; j < nmig[i]; ++j) {
      const size_t q = msrc + j;
      if (mig_edge[q] < 0 || mig_edge[q] >= ne || mig_to[q] < 0 || mig_to[q] >= T ||
          !(mig_time[q] >= 0) || !std::isfinite(mig_time[q]))
        return fail(c, IMA_ERR_ARG, "slot %d: migration %d out of range", first + i, j);
      if (j > 0 && (mig_edge[q] < mig_edge[q - 1] ||
                    (mig_edge[q] == mig_edge[q - 1] && mig_time[q] < mig_time[q - 1])))
        return fail(c, IMA_ERR_ARG, "slot %d: migrations not sorted by edge then time", first + i);
    }
    std::copy(pop + src, pop + src + ne, c->pop.begin() + s.node_off);
    if (nmig[i] > s.mig_cap) {
      s.mig_cap = nmig[i] + nmig[i] / 2 + 4;
      s.mig_off = (int64_t)c->migs;
      c->migs += (size_t)s.mig_cap;
      c->mig_edge.resize(c->migs, 0);
      c->mig_to.resize(c->migs, 0);
      c->mig_time.resize(c->migs, 0.0);
    }
    s.nmig = nmig[i];
    std::fill(c->mig_count.begin() + s.node_off, c->mig_count.begin() + s.node_off + ne, 0);
    std::fill(c->mig_first.begin() + s.node_off, c->mig_first.begin() + s.node_off + ne, 0);
    for (int j = nmig[i] - 1; j >= 0; --j) {
      const int e = mig_edge[msrc + j];
      c->mig_first[s.node_off + e] = j;
      c->mig_count[s.node_off + e] += 1;
    }
    if (nmig[i]) {
      std::copy(mig_edge + msrc, mig_edge + msrc + nmig[i], c->mig_edge.begin() + s.mig_off);
      std::copy(mig_to + msrc, mig_to + msrc + nmig[i], c->mig_to.begin() + s.mig_off);
      std::copy(mig_time + msrc, mig_time + msrc + nmig[i], c->mig_time.begin() + s.mig_off);
    }
    src += ne;
    msrc += nmig[i];
  }
  const SlotMeta &s0 = c->slots[first];
  const SlotMeta &s1 = c->slots[first + count - 1];
  const size_t nb = s1.node_off + 2 * (size_t)s1.n - 1 - s0.node_off;
  HIPCHK(c, reserve(c->d_pop, c->nodes, true, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_pop.p + s0.node_off, c->pop.data() + s0.node_off, nb * 4,
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, reserve(c->d_mig_first, c->nodes, true, c->stream));
  HIPCHK(c, reserve(c->d_mig_count, c->nodes, true, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_mig_first.p + s0.node_off, c->mig_first.data() + s0.node_off,
                           nb * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_mig_count.p + s0.node_off, c->mig_count.data() + s0.node_off,
                           nb * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, reserve(c->d_mig_edge, c->migs + 1, true, c->stream));
  HIPCHK(c, reserve(c->d_mig_to, c->migs + 1, true, c->stream));
  HIPCHK(c, reserve(c->d_mig_time, c->migs + 1, true, c->stream));
  if (c->migs) {
    HIPCHK(c, hipMemcpyAsync(c->d_mig_edge.p, c->mig_edge.data(), c->migs * 4,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_mig_to.p, c->mig_to.data(), c->migs * 4,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_mig_time.p, c->mig_time.data(), c->migs * 8,
                             hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, upload_slot_meta_range(c, first, count));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return IMA_OK;
}

// Checks the slots of one prior call: returns their shared model, its P,
// the largest event count of a genealogy (*emax) and the log-factorial
// table length the groups' summed counts need (*lf_need).
static int prior_prepare(ima_ctx *c, const int32_t *slots, int nslots, int *model, int *P, int *emax,
                  size_t *lf_need, const int32_t *group_off, int ngroups) {
  *model = -1;
  *emax = 1;
  for (int i = 0; i < nslots; ++i) {
    const int sl = slots[i];
    if (sl < 0 || sl >= (int)c->slots.size()) return fail(c, IMA_ERR_ARG, "slot %d out of range", sl);
    const SlotMeta &s = c->slots[sl];
    if (s.root < 0) return fail(c, IMA_ERR_STATE, "slot %d holds no genealogy", sl);
    const LocusMeta &lm = c->loci[s.locus];
    if (lm.pmodel < 0) return fail(c, IMA_ERR_STATE, "slot %d: locus has no population model", sl);
    if (*model < 0) *model = lm.pmodel;
    if (lm.pmodel != *model) return fail(c, IMA_ERR_ARG, "slots of one call must share a model");
    *emax = std::max(*emax, (s.n - 1) + s.nmig + c->models[*model].npops - 1);
  }
  *P = *model >= 0 ? c->models[*model].npops : 1;
  const size_t lds = treeweight_lds_bytes(*P, *emax);
  if (lds > (size_t)c->max_lds)
    return fail(c, IMA_ERR_SIZE, "a genealogy has %d prior events (LDS %zu B)", *emax, lds);
  *lf_need = 4;
  for (int g = 0; g < ngroups; ++g) {
    size_t tot = 4;
    for (int i = group_off[g]; i < group_off[g + 1]; ++i) {
      const SlotMeta &s = c->slots[slots[i]];
      tot += (size_t)s.n - 1 + (size_t)s.nmig;
    }
    *lf_need = std::max(*lf_need, tot);
  }
  return IMA_OK;
}

int ima_slots_reserve_migs(ima_ctx *c, int first, int count, int cap) {
  if (!c || first < 0 || count < 0 || first + count > (int)c->slots.size() || cap < 0)
    return IMA_ERR_ARG;
  hipSetDevice(c->device);
  std::vector<int32_t> sl(count), cp(count, cap);
  for (int i = 0; i < count; ++i) sl[i] = first + i;
  return grow_mig_caps(c, sl.data(), cp.data(), count);
}

int ima_t_nw_update(ima_ctx *c, const int32_t *slots, int nslots, const double *tvals,
                    int period, double newt, uint64_t seed, uint64_t counter, double *migweight,
                    int32_t *updated) {
  if (!c || (!slots && nslots > 0) || nslots < 0 || !tvals) return IMA_ERR_ARG;
  if (nslots == 0) return IMA_OK;
  hipSetDevice(c->device);
  int model, P, emax, rc;
  size_t lfn;
  const int32_t goff[2] = {0, nslots};
  if ((rc = prior_prepare(c, slots, nslots, &model, &P, &emax, &lfn, goff, 1)) != IMA_OK) return rc;
  if (P < 2) return fail(c, IMA_ERR_ARG, "the NW update needs two or more populations");
  if (c->mc3.live)
    for (int i = 0; i < nslots; ++i)
      if (slots[i] >= c->mc3.first && slots[i] < c->mc3.first + c->mc3.C * c->mc3.L)
        return fail(c, IMA_ERR_STATE, "slot %d belongs to the device sweep's chains", slots[i]);
  if (period < 0 || period > P - 2) return fail(c, IMA_ERR_ARG, "split %d out of range", period);
  {
    const double t_u = period == 0 ? 0.0 : tvals[period - 1];
    const double t_d = period == P - 2 ? kTimeMax : tvals[period + 1];
    if (!(newt > t_u && newt < t_d) || !(tvals[period] > t_u && tvals[period] < t_d))
      return fail(c, IMA_ERR_ARG, "split times %g -> %g outside (%g, %g)", tvals[period], newt,
                  t_u, t_d);
  }
  int max_n = 2, max_cap = 0;
  for (int i = 0; i < nslots; ++i) {
    max_n = std::max(max_n, c->slots[slots[i]].n);
    max_cap = std::max(max_cap, c->slots[slots[i]].mig_cap);
  }
  if (nw_lds_bytes(max_n, 0, 0) > (size_t)c->max_lds)
    return fail(c, IMA_ERR_SIZE, "the NW update of %d gene copies needs more LDS than %d B", max_n,
                c->max_lds);
  if ((rc = upload_loci(c)) != IMA_OK) return rc;
  if ((rc = upload_slot_meta(c)) != IMA_OK) return rc;
  if ((rc = upload_models(c)) != IMA_OK) return rc;
  if ((rc = nw_reserve_backups(c)) != IMA_OK) return rc;
  if ((rc = nw_upload_slots(c, slots, nslots)) != IMA_OK) return rc;
  const int tw = P - 1;
  HIPCHK(c, reserve(c->d_pslots, nslots));
  HIPCHK(c, reserve(c->d_tv, (size_t)tw));
  HIPCHK(c, reserve(c->d_nw_period, 1));
  HIPCHK(c, reserve(c->d_nw_newt, 1));
  HIPCHK(c, reserve(c->d_nw_nbk, nslots));
  HIPCHK(c, reserve(c->d_nw_st, nslots));
  HIPCHK(c, reserve(c->d_nw_w, nslots));
  HIPCHK(c, hipMemcpyAsync(c->d_pslots.p, slots, nslots * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_tv.p, tvals, (size_t)tw * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_nw_period.p, &period, 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_nw_newt.p, &newt, 8, hipMemcpyHostToDevice, c->stream));
  NwArgs a;
  a.count = nslots;
  a.slot = c->d_pslots.p;
  a.tv = c->d_tv.p;
  a.tw = tw;
  a.period = c->d_nw_period.p;
  a.newt = c->d_nw_newt.p;
  a.seed = seed;
  a.ctr = counter;
  a.model = model;
  a.emax = nw_event_limit(c, P);
  a.max_n = max_n;
  a.max_cap = max_cap;
  a.stage = nw_lds_bytes(max_n, max_cap, 1) <= (size_t)c->max_lds ? 1 : 0;
  a.pop_bk = c->d_nw_pop.p;
  a.mf_bk = c->d_nw_mf.p;
  a.mc_bk = c->d_nw_mc.p;
  a.me_bk = c->d_nw_me.p;
  a.mto_bk = c->d_nw_mto.p;
  a.mt_bk = c->d_nw_mt.p;
  a.nmig_bk = c->d_nw_nbk.p;
  a.migweight = c->d_nw_w.p;
  a.status = c->d_nw_st.p;
  HIPCHK(c, launch_nw_update(view(c), a, c->stream));
  std::vector<double> w(nslots);
  std::vector<int32_t> st(nslots);
  HIPCHK(c, hipMemcpyAsync(w.data(), c->d_nw_w.p, nslots * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(st.data(), c->d_nw_st.p, nslots * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<int32_t> changed;
  for (int i = 0; i < nslots; ++i) {
    if (migweight) migweight[i] = w[i];
    if (updated) updated[i] = st[i];
    if (st[i] == 1) changed.push_back(slots[i]);
  }
  return nw_sync_slots(c, changed.data(), (int)changed.size());
}

int ima_treeweight(ima_ctx *c, const int32_t *slots, int nslots, const double *tvals,
                   int32_t *cc, double *fc, double *hcc, int32_t *mc, double *fm,
                   double *length) {
  if (!c || (!slots && nslots > 0) || nslots < 0) return IMA_ERR_ARG;
  if (nslots == 0) return IMA_OK;
  hipSetDevice(c->device);
  int model, P, emax, rc;
  size_t lfn;
  const int32_t goff[2] = {0, nslots};
  if ((rc = prior_prepare(c, slots, nslots, &model, &P, &emax, &lfn, goff, 1)) != IMA_OK) return rc;
  if (P > 1 && !tvals) return IMA_ERR_ARG;
  if ((rc = upload_loci(c)) != IMA_OK) return rc;
  if ((rc = upload_slot_meta(c)) != IMA_OK) return rc;
  if ((rc = upload_models(c)) != IMA_OK) return rc;
  const int R = prior_rec_len(P), tw = P > 1 ? P - 1 : 1;
  HIPCHK(c, reserve(c->d_pslots, nslots));
  HIPCHK(c, reserve(c->d_tv, (size_t)nslots * tw));
  HIPCHK(c, reserve(c->d_rec, (size_t)nslots * R));
  HIPCHK(c, reserve(c->d_pstatus, nslots));
  HIPCHK(c, hipMemcpyAsync(c->d_pslots.p, slots, nslots * 4, hipMemcpyHostToDevice, c->stream));
  if (P > 1)
    HIPCHK(c, hipMemcpyAsync(c->d_tv.p, tvals, (size_t)nslots * tw * 8, hipMemcpyHostToDevice,
                             c->stream));
  HIPCHK(c, launch_treeweight(view(c), c->d_pslots.p, nslots, c->d_tv.p, nullptr, P, emax,
                              c->d_rec.p, c->d_pstatus.p, c->stream));
  std::vector<double> rec((size_t)nslots * R);
  std::vector<int32_t> st(nslots);
  HIPCHK(c, hipMemcpyAsync(rec.data(), c->d_rec.p, rec.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(st.data(), c->d_pstatus.p, nslots * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < nslots; ++i)
    if (st[i] != IMA_STATUS_OK)
      return fail(c, IMA_ERR_ARG, "slot %d: population labels or migrations do not fit the model",
                  slots[i]);
  const int P2 = P * P, P3 = P * P * P;
  for (int i = 0; i < nslots; ++i) {
    const double *r = rec.data() + (size_t)i * R;
    for (int x = 0; x < P2; ++x) {
      if (cc) cc[(size_t)i * P2 + x] = (int32_t)r[x];
      if (fc) fc[(size_t)i * P2 + x] = r[P2 + x];
      if (hcc) hcc[(size_t)i * P2 + x] = r[2 * P2 + x];
    }
    for (int x = 0; x < P3; ++x) {
      if (mc) mc[(size_t)i * P3 + x] = (int32_t)r[3 * P2 + x];
      if (fm) fm[(size_t)i * P3 + x] = r[3 * P2 + P3 + x];
    }
    if (length) length[i] = r[R - 1];
  }
  return IMA_OK;
}

int ima_integrate_tree_prob(ima_ctx *c, int model, const double *wsum, int ngroups,
                            double *probg, double *qint, double *mint) {
  if (!c || !wsum || !probg || ngroups < 0) return IMA_ERR_ARG;
  if (model < 0 || model >= (int)c->models.size()) return fail(c, IMA_ERR_ARG, "model %d unknown", model);
  if (ngroups == 0) return IMA_OK;
  hipSetDevice(c->device);
  const ModelMeta &mm = c->models[model];
  const int P = mm.npops, R = prior_rec_len(P), P2 = P * P, P3 = P2 * P;
  // counts (cc, mc) are whole numbers >= 0, weights finite: the log-factorial
  // table must reach the largest parameter count
  size_t lfn = 4;
  for (int g = 0; g < ngroups; ++g) {
    const double *w = wsum + (size_t)g * R;
    double tot = 0.0;
    for (int r = 0; r < R; ++r) {
      if (!std::isfinite(w[r])) return fail(c, IMA_ERR_ARG, "group %d: entry %d not finite", g, r);
      const bool count = r < P2 || (r >= 3 * P2 && r < 3 * P2 + P3);
      if (count) {
        if (w[r] < 0 || w[r] != std::floor(w[r]))
          return fail(c, IMA_ERR_ARG, "group %d: count entry %d is not a whole number", g, r);
        tot += w[r];
      }
    }
    if (tot > 1e8) return fail(c, IMA_ERR_SIZE, "group %d: %g events", g, tot);
    lfn = std::max(lfn, (size_t)tot + 4);
  }
  grow_logfact(c, lfn);
  int rc;
  if ((rc = upload_models(c)) != IMA_OK) return rc;
  std::vector<int32_t> goff(ngroups + 1);
  for (int g = 0; g <= ngroups; ++g) goff[g] = g;
  HIPCHK(c, reserve(c->d_rec, (size_t)ngroups * R));
  HIPCHK(c, reserve(c->d_goff, ngroups + 1));
  HIPCHK(c, reserve(c->d_probg, ngroups));
  HIPCHK(c, reserve(c->d_qint, (size_t)ngroups * std::max(1, mm.npsize)));
  HIPCHK(c, reserve(c->d_mintg, (size_t)ngroups * std::max(1, mm.nmig)));
  HIPCHK(c, hipMemcpyAsync(c->d_rec.p, wsum, (size_t)ngroups * R * 8, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_goff.p, goff.data(), (ngroups + 1) * 4, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, launch_prior(view(c), model, P, mm.npsize + mm.nmig, c->d_rec.p, c->d_goff.p, ngroups,
                         c->d_probg.p, c->d_qint.p, c->d_mintg.p, c->stream));
  HIPCHK(c, hipMemcpyAsync(probg, c->d_probg.p, ngroups * 8, hipMemcpyDeviceToHost, c->stream));
  if (qint && mm.npsize)
    HIPCHK(c, hipMemcpyAsync(qint, c->d_qint.p, (size_t)ngroups * mm.npsize * 8,
                             hipMemcpyDeviceToHost, c->stream));
  if (mint && mm.nmig)
    HIPCHK(c, hipMemcpyAsync(mint, c->d_mintg.p, (size_t)ngroups * mm.nmig * 8,
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return IMA_OK;
}

int ima_prior(ima_ctx *c, const int32_t *slots, const int32_t *group_off, int ngroups,
              const double *tvals, double *probg, double *qint, double *mint) {
  if (!c || !group_off || ngroups < 0 || !probg) return IMA_ERR_ARG;
  if (ngroups == 0) return IMA_OK;
  if (group_off[0] != 0) return fail(c, IMA_ERR_ARG, "group_off[0] must be 0");
  for (int g = 0; g < ngroups; ++g)
    if (group_off[g + 1] <= group_off[g]) return fail(c, IMA_ERR_ARG, "group %d is empty", g);
  const int nslots = group_off[ngroups];
  if (!slots) return IMA_ERR_ARG;
  hipSetDevice(c->device);
  int model, P, emax, rc;
  size_t lfn;
  if ((rc = prior_prepare(c, slots, nslots, &model, &P, &emax, &lfn, group_off, ngroups)) != IMA_OK)
    return rc;
  if (P > 1 && !tvals) return IMA_ERR_ARG;
  grow_logfact(c, lfn);
  if ((rc = upload_loci(c)) != IMA_OK) return rc;
  if ((rc = upload_slot_meta(c)) != IMA_OK) return rc;
  if ((rc = upload_models(c)) != IMA_OK) return rc;
  const ModelMeta &mm = c->models[model];
  const int R = prior_rec_len(P), tw = P > 1 ? P - 1 : 1;
  std::vector<int32_t> row(nslots);
  for (int g = 0; g < ngroups; ++g)
    for (int i = group_off[g]; i < group_off[g + 1]; ++i) row[i] = g;
  HIPCHK(c, reserve(c->d_pslots, nslots));
  HIPCHK(c, reserve(c->d_prow, nslots));
  HIPCHK(c, reserve(c->d_goff, ngroups + 1));
  HIPCHK(c, reserve(c->d_tv, (size_t)ngroups * tw));
  HIPCHK(c, reserve(c->d_rec, (size_t)nslots * R));
  HIPCHK(c, reserve(c->d_pstatus, nslots));
  HIPCHK(c, reserve(c->d_probg, ngroups));
  HIPCHK(c, reserve(c->d_qint, (size_t)ngroups * std::max(1, mm.npsize)));
  HIPCHK(c, reserve(c->d_mintg, (size_t)ngroups * std::max(1, mm.nmig)));
  HIPCHK(c, hipMemcpyAsync(c->d_pslots.p, slots, nslots * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_prow.p, row.data(), nslots * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_goff.p, group_off, (ngroups + 1) * 4, hipMemcpyHostToDevice,
                           c->stream));
  if (P > 1)
    HIPCHK(c, hipMemcpyAsync(c->d_tv.p, tvals, (size_t)ngroups * tw * 8, hipMemcpyHostToDevice,
                             c->stream));
  const DeviceView v = view(c);
  HIPCHK(c, launch_treeweight(v, c->d_pslots.p, nslots, c->d_tv.p, c->d_prow.p, P, emax,
                              c->d_rec.p, c->d_pstatus.p, c->stream));
  int trc;
  if ((trc = timing_begin(c, K_PRIOR)) != IMA_OK) return trc;
  HIPCHK(c, launch_prior(v, model, P, mm.npsize + mm.nmig, c->d_rec.p, c->d_goff.p, ngroups,
                         c->d_probg.p, c->d_qint.p, c->d_mintg.p, c->stream));
  if ((trc = timing_end(c, K_PRIOR)) != IMA_OK) return trc;
  std::vector<int32_t> st(nslots);
  HIPCHK(c, hipMemcpyAsync(st.data(), c->d_pstatus.p, nslots * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(probg, c->d_probg.p, ngroups * 8, hipMemcpyDeviceToHost, c->stream));
  if (qint && mm.npsize)
    HIPCHK(c, hipMemcpyAsync(qint, c->d_qint.p, (size_t)ngroups * mm.npsize * 8,
                             hipMemcpyDeviceToHost, c->stream));
  if (mint && mm.nmig)
    HIPCHK(c, hipMemcpyAsync(mint, c->d_mintg.p, (size_t)ngroups * mm.nmig * 8,
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if ((trc = timing_collect(c, K_PRIOR)) != IMA_OK) return trc;
  for (int i = 0; i < nslots; ++i)
    if (st[i] != IMA_STATUS_OK)
      return fail(c, IMA_ERR_ARG, "slot %d: population labels or migrations do not fit the model",
                  slots[i]);
  return IMA_OK;
}

// ---------------------------------------------------------------- timing --

int ima_timing_enable(ima_ctx *c, int on) {
  if (!c) return IMA_ERR_ARG;
  c->timing.on = on != 0;
  return IMA_OK;
}

int ima_timing_get(ima_ctx *c, int k, int64_t *launches, double *total_ms) {
  if (!c || k < 0 || k >= kTimingClasses) return IMA_ERR_ARG;
  if (launches) *launches = c->timing.launches[k];
  if (total_ms) *total_ms = c->timing.ms[k];
  return IMA_OK;
}

int ima_timing_reset(ima_ctx *c) {
  if (!c) return IMA_ERR_ARG;
  for (int k = 0; k < kTimingClasses; ++k) {
    c->timing.launches[k] = 0;
    c->timing.ms[k] = 0;
  }
  return IMA_OK;
}

}  // extern "C"

// ------------------------------------------------ helpers for serve.hip --

LTable ctx_ltable(ima_ctx *c, int table_id) {
  if (!c || table_id < 0 || table_id >= (int)c->tables.size() || !c->tables[table_id].live)
    return LTable{};
  return ltable(c->tables[table_id]);
}

int ctx_set_error(ima_ctx *c, int rc, const char *msg) { return fail(c, rc, "%s", msg); }

size_t ctx_device_bytes(const ima_ctx *c) {
  size_t b = 0;
  auto add = [&](size_t n, size_t sz) { b += n * sz; };
  add(c->d_loci.n, sizeof(LocusMeta));
  add(c->d_seq.n, 1);
  add(c->d_mult.n, 4);
  add(c->d_mask.n, 8);
  add(c->d_slots.n, sizeof(SlotMeta));
  add(c->d_up0.n + c->d_up1.n + c->d_down.n + c->d_A.n + c->d_pop.n, 4);
  add(c->d_time.n, 8);
  add(c->d_prog.n, sizeof(ProgOp));
  add(c->d_mig_edge.n + c->d_mig_to.n + c->d_mig_first.n + c->d_mig_count.n, 4);
  add(c->d_mig_time.n, 8);
  add(c->d_items.n, sizeof(ima_item));
  add(c->d_out.n + c->d_partial.n + c->d_rec.n + c->d_lwork.n, 8);
  const auto &m = c->mc3;
  add(m.pdg.n + m.newpdg.n + m.tbackup.n + m.partial.n + m.wrec.n + m.wnew.n + m.wall.n +
          m.mtbackup.n,
      8);
  add(m.units.n + m.red.n, sizeof(int4));
  for (const auto &t : c->tables)
    if (t.live) add((size_t)t.G * t.K, 4);
  return b;
}
