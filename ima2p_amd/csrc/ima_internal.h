// ima_internal.h -- shared definitions of the HIP engine (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/ima_gpu.h"
#include "philox.h"

namespace ima {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kMaxGenes = IMA_MAX_GENES;
constexpr int kMaxEdges = 2 * kMaxGenes - 1;
constexpr int kTipSlot = 0xFFFF;   // program operand that is a tip
constexpr double kLn2 = 0.69314718055994530941723212146;
constexpr int kMaxPops = 10;       // sampled populations (the reference's MAXPOPS)
constexpr int kMaxLaunchLds = 65536;  // dynamic LDS any launch may take, in bytes

// Per-locus metadata resident on the device.
struct LocusMeta {
  int32_t model;     // ima_model
  int32_t n;         // gene copies
  int32_t S;         // HKY patterns / IS segregating sites
  int32_t totsites;  // HKY sites after gap removal
  int64_t seq_off;   // HKY: int8 [n][S] bases; IS: unused
  int64_t mult_off;  // HKY: int32 [S]
  int64_t mask_off;  // IS: uint64 [S][W] tips carrying state 1
  int32_t W;         // IS: 64-bit words per tip mask
  int32_t Spad;      // HKY: row stride of seq (S rounded up to 128, zero padded)
  double pi[4];
  double sumlogk;
  int32_t pmodel;               // population model of the prior, -1 if none
  int32_t samppop[kMaxPops];    // gene copies sampled per population
  double hval;                  // inheritance scalar h
};

// Per-slot metadata: where the genealogy and its postorder program live.
struct SlotMeta {
  int32_t locus;
  int32_t n;
  int32_t root;
  int32_t nslots;    // LDS partial slots the program needs
  int64_t node_off;  // into up0/up1/down/time/A arrays (2n-1 entries)
  int64_t prog_off;  // into prog (n-1 ops)
  int64_t mig_off;   // into mig_edge/mig_time/mig_to (events sorted by edge, time)
  int32_t nmig;      // migration events of the genealogy
  int32_t mig_cap;   // events the slot's region holds
};

// Population model of the genealogy prior.  Periods k = 0..P-1 lie between
// split times; populations 0..2P-2 (sampled ones first, then ancestral
// P + k created by the split closing period k).  The model's ints live in
// one block at ioff:
//   plist[P][P] (-1 padded) | npresent[P] | popdown[2P-1] | pope[2P-1] |
//   join[P][3] = (a, b, new) | pos[P][2P-1] (position in plist or -1) |
//   theta_off[npsize+1] | theta_wp[][2] = (k, i) |
//   mig_off[nmig+1] | mig_wp[][3] = (k, i, j)
// and qmax[npsize] | mmax[nmig] in the double block at doff.
struct ModelMeta {
  int32_t npops, npsize, nmig, nints;  // nints: length of the int block
  int64_t ioff, doff;
};
struct ModelView {
  const int32_t *plist, *npresent, *popdown, *pope, *join, *pos;
  const int32_t *theta_off, *theta_wp, *mig_off, *mig_wp;
  const double *qmax, *mmax;
};
// Doubles per genealogy weight record:
//   cc[P][P] | fc[P][P] | hcc[P][P] | mc[P][P][P] | fm[P][P][P] | length
__host__ __device__ inline int prior_rec_len(int P) { return 3 * P * P + 2 * P * P * P + 1; }

// Dynamic LDS budget of the combined likelihood-and-weights launch (the
// static tables add a few KB).
constexpr size_t kTwLdsBudget = 61440;

// Orders one wavefront's LDS writes before its later LDS reads (lanes read
// what other lanes of the same wavefront wrote).
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One postorder operation of the HKY pruning program: compute the partial
// at the top of edge `node` from daughters a and b.  The program is a stack
// machine: OP_TT combines two tips and pushes, OP_ST combines the stack top
// (daughter a) with tip b in place, OP_SS pops a (second) and b (top) and
// pushes their combination.  Depth never exceeds the Strahler number.
enum { OP_TT = 0, OP_ST = 1, OP_SS = 2 };
struct ProgOp {
  int32_t a;       // edge of daughter a
  int32_t b;       // edge of daughter b
  int32_t type;    // OP_TT / OP_ST / OP_SS
  int32_t node;    // edge whose top the result belongs to
};

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
};

// Launch plumbing shared between translation units.
enum KClass { K_HKY64 = 0, K_HKY32 = 1, K_IS = 2, K_SW = 3, K_LMODE = 4,
              K_SWAP = 5, K_MC3 = 6, K_ACCEPT = 7, K_PRIOR = 8, kTimingClasses = 9 };

struct Timing {
  bool on = false;
  hipEvent_t ev0[kTimingClasses]{}, ev1[kTimingClasses]{};
  int64_t launches[kTimingClasses]{};
  double ms[kTimingClasses]{};
};

struct DeviceView {
  const LocusMeta *loci;
  const SlotMeta *slots;
  const int32_t *up0, *up1, *down;
  const double *time;
  const int32_t *A;
  const ProgOp *prog;
  const int8_t *seq;
  const int32_t *mult;
  const uint64_t *mask;
  const int32_t *pop;        // population of each edge at its lower end
  const int32_t *mig_edge, *mig_to;
  const double *mig_time;
  // per edge (indexed like time[]): its first migration within the slot's
  // events and how many it has (events are sorted by edge, then time)
  const int32_t *mig_first, *mig_count;
  const ModelMeta *models;
  const int32_t *mint;
  const double *mdbl;
  const double *logfact;     // log(k!) by running sums of log(i)
};

__host__ __device__ inline ModelView model_view(const int32_t *mi, const double *mdb,
                                                const ModelMeta &m) {
  const int P = m.npops, T = 2 * P - 1;
  ModelView r;
  const int32_t *b = mi + m.ioff;
  r.plist = b;
  r.npresent = r.plist + P * P;
  r.popdown = r.npresent + P;
  r.pope = r.popdown + T;
  r.join = r.pope + T;
  r.pos = r.join + 3 * P;
  r.theta_off = r.pos + P * T;
  r.theta_wp = r.theta_off + m.npsize + 1;
  r.mig_off = r.theta_wp + 2 * r.theta_off[m.npsize];
  r.mig_wp = r.mig_off + m.nmig + 1;
  r.qmax = mdb + m.doff;
  r.mmax = r.qmax + m.npsize;
  return r;
}
__device__ inline ModelView model_view(const DeviceView &v, const ModelMeta &m) {
  return model_view(v.mint, v.mdbl, m);
}

// HKY work plan: one wavefront per (item, 64-pattern block).  units[k] =
// {item, block, first partial, blocks of the item}; items with more than
// one block are merged from `partial` in block order by a second kernel
// (red[k] = {item, first partial, blocks, 0}).
struct HkyPlan {
  const int4 *units = nullptr;
  int nunits = 0;
  int ppl = 1;  // patterns per lane: site blocks of 64 * ppl patterns
  const int4 *red = nullptr;
  int nred = 0;
  double *partial = nullptr;
};
// Host side: append the units of HKY item ii with S patterns.
void hky_plan_add(std::vector<int4> &units, std::vector<int4> &red, int ii, int S,
                  int *nblocks_total, int ppl);

// Kernel launchers (kern_hky.hip, kern_lik.hip).
size_t hky_lds_bytes(int precision, int max_n, int max_slots, int ppl);
hipError_t launch_hky(const DeviceView &v, const ima_item *items, const HkyPlan &plan,
                      int precision, int max_n, int max_slots, double *out,
                      int32_t *status, hipStream_t st);
// Genealogy weights run by the combined scoring launch: entries [0, n) of
// slots / tv_row / rec_out / status (treeweight_kernel's arguments), emax
// events at most per genealogy.
struct TwLaunch {
  const int32_t *slots;
  int n, emax;
  const double *tvals;
  const int32_t *tv_row;
  double *rec_out;
  int32_t *status;
};
// HKY units (no merge step) and treeweight in one launch, when its LDS
// (hky_tw_lds_bytes) is within kTwLdsBudget.
size_t hky_tw_lds_bytes(int precision, int max_n, int max_slots, int ppl, int P, int emax);
hipError_t launch_hky_tw(const DeviceView &v, const ima_item *items, const HkyPlan &plan,
                         int precision, int max_n, int max_slots, double *out, int32_t *status,
                         const TwLaunch &tw, int P, hipStream_t st);
// With cnt, item_idx is a list of which only the first *cnt entries are
// scored (nitems is the list's capacity, the grid), and block 0 clears
// *cnt_next, the list filled for the next launch.
hipError_t launch_is(const DeviceView &v, const ima_item *items,
                     const int32_t *item_idx, int nitems, int max_n, int max_w,
                     double *out, int32_t *status, int32_t *site_node,
                     int32_t *counts, hipStream_t st, const int32_t *cnt = nullptr,
                     int32_t *cnt_next = nullptr);
hipError_t launch_sw_update(const DeviceView &v, int32_t *A, const ima_sw_move *moves,
                            int nmoves, int pass, int calc, double *out, int32_t *iout,
                            hipStream_t st);
hipError_t launch_sw(const DeviceView &v, const ima_item *items,
                     const int32_t *item_idx, int nitems, double *out,
                     int32_t *status, double *dlike, hipStream_t st,
                     const int32_t *cnt = nullptr, int32_t *cnt_next = nullptr);

// L mode (kern_lmode.hip).
struct LTable {
  const float *rows = nullptr;  // [G][K]
  int G = 0, K = 0, npsize = 0, nmig = 0, npops = 0;
};
hipError_t launch_jointp(const LTable &t, const double *x, int npoints,
                         int modeltype, int r0, int r1, int mode,
                         double *work, size_t work_bytes, double *out,
                         double *ess, hipStream_t st);
size_t jointp_work_bytes(const LTable &t, int npoints);
hipError_t launch_greater_than(const LTable &t, int kind, const int2 *pairs, const double *pmax,
                               int npairs, const double *lf, int nlf, double *per, double *out,
                               hipStream_t st);
size_t popmig_work_bytes(const LTable &t, int nx);
hipError_t launch_popmig(const LTable &t, int theta_i, int mig_i, double qmax, double mprior,
                         int expo, int kind, int prob_or_like, int first, int last,
                         const double *x, int nx, const double *lf, int nlf, double *work,
                         double *out, hipStream_t st);
hipError_t launch_margincalc(const LTable &t, int param, const double *x,
                             int nx, int logi, double yadjust, int expo,
                             double prior_mean, double *work, double *out,
                             hipStream_t st);
size_t margincalc_work_bytes(const LTable &t, int nx);
hipError_t upload_pow10_table(hipStream_t st);

// Genealogy prior (kern_prior.hip).
size_t treeweight_lds_bytes(int P, int emax);
hipError_t launch_treeweight(const DeviceView &v, const int32_t *slots, int nslots,
                             const double *tvals, const int32_t *tv_row, int P, int emax,
                             double *rec_out, int32_t *status, hipStream_t st);
hipError_t launch_prior(const DeviceView &v, int model, int P, int nparams, const double *recs,
                        const int32_t *group_off, int ngroups, double *probg, double *qint,
                        double *mint, hipStream_t st);

// Device-resident MC3 state (kern_mc3.hip).
// Arguments of the NW migration update (kern_nw.hip): genealogy i is slot
// slot[i] (or first_slot + i) of chain row row[i] (or 0); the row's split
// period[row] (-1: no update) moves from tv[row * tw + period] to
// newt[row].  Edges draw from Philox streams keyed by seed, the
// genealogy's key id (keyid[i] or i) and the edge, at counter ctr (or the
// sweep counter's stream 5).  Backups are indexed like the node and
// migration arrays.
struct NwArgs {
  int count = 0;
  const int32_t *slot = nullptr;
  int first_slot = 0;
  const int32_t *row = nullptr;
  const double *tv = nullptr;
  int tw = 1;
  const int32_t *period = nullptr;
  const double *newt = nullptr;
  const int32_t *keyid = nullptr;
  uint64_t seed = 0, ctr = 0;
  const uint64_t *counter = nullptr;
  int model = 0;
  int emax = 0;   // event limit of the genealogy-prior kernels
  int max_n = 2;  // largest gene-copy count (sizes the LDS)
  int max_cap = 0;  // largest event count of a genealogy before the update
  int stage = 0;    // 1: the genealogies are staged in LDS (it has room)
  int32_t *pop_bk = nullptr, *mf_bk = nullptr, *mc_bk = nullptr;
  int32_t *me_bk = nullptr, *mto_bk = nullptr;
  double *mt_bk = nullptr;
  int32_t *nmig_bk = nullptr;  // [count]
  double *migweight = nullptr;  // [count]
  int32_t *status = nullptr;    // [count]: 0 untouched, 1 updated, 2 left (no room)
};
size_t nw_lds_bytes(int max_n, int max_cap, int stage);
hipError_t launch_nw_update(const DeviceView &v, const NwArgs &a, hipStream_t st);
hipError_t launch_nw_restore(const DeviceView &v, const NwArgs &a, const int32_t *decision,
                             hipStream_t st);
// Copies cnt[i] events of every slot i from offset from[i] to to[i]
// (regions that do not overlap).
hipError_t launch_move_events(int n, const int64_t *from, const int64_t *to, const int32_t *cnt,
                              int32_t *me, int32_t *mto, double *mt, hipStream_t st);

struct MC3Dev {
  int C, L;
  int group;  // chains per uncoupled analysis (swaps stay inside one); C for one analysis
  int first_slot;
  double *pdg, *newpdg;  // [C*L]
  int32_t *status;       // [C*L]
  double *beta, *S;      // [C]
  int32_t *rec_v;        // [C*L]
  double *rec_t;         // [C*L]
  double *tbackup;       // node times backup (indexed like time[])
  int32_t *decision;     // [C]
  double *logs;          // [C] log s of the rescaling
  uint64_t *counter;     // sweep counter
  unsigned long long *stats;  // [8]: [0] proposals, [4] swap tries, [5] swaps
  uint32_t *acc;         // [C*L] accepted genealogy proposals per item
  uint32_t *cacc;        // [C] accepted rescalings per chain
  int64_t node_first;    // first node of the MC3 slots in the node arrays
  int64_t nnodes;        // nodes covered by the MC3 slots
  int32_t *node_item;    // [nnodes] item owning each node
  int2 *iblk;            // [C*L] HKY pattern blocks: {first partial, blocks} (blocks <= 1: direct)
  double *partial;       // HKY per-block partial log-likelihoods
  unsigned *done;        // chains finished in the current sweep (last one swaps)
  uint64_t seed;
  double scale_width;
  int ry1;               // 1: the time update is the rubber-band move of one split time
  // genealogy prior (prior < 0: the sweep scores the data likelihood only)
  int prior;             // population model id
  int P, R, nparam;      // populations, weight record length, parameters
  int emax;              // largest event count of a genealogy
  double tmax;           // split times: uniform prior on [0, tmax]
  double *tv;            // [C][TW] split times of each chain (TW = max(P-1, 1))
  double *tvbackup;      // [C][TW]
  double *wrec;          // [C*L][R] committed genealogy weights
  double *wnew;          // [C*L][R] weights of the proposals
  double *wall;          // [C][R] all-loci sums (allgweight)
  double *qterm;         // [C][nparam] integrals of the committed sums
  int32_t *qc;           // [C][nparam] counts those integrals were taken at
  double *qf;            // [C][nparam] weight sums those integrals were taken at
  double *probg;         // [C]
  double *mtbackup;      // migration times backup (indexed like mig_time)
  int32_t *tv_row;       // [C*L] chain of each item
  int32_t *tw_slots;     // [C*L] slot of each item
  int32_t *wstatus;      // [C*L] weights status of the proposals
  int pchunk;            // loci decided per speculative group by the accept kernel
  int mints;             // length of the model's int block (staged in LDS)
  // decisions for proposals written by the host (ima_mc3_decide): no device
  // proposal is pending, a rejection leaves the chain's slot as it is, the
  // decision is log U - hast < beta (dprobg + gbeta dpdg)
  int host;              // 1 inside ima_mc3_decide's launches
  double gbeta;          // multiplier of dpdg (1 outside thermodynamic integration)
  const double *hast;    // [C*L] Hastings terms, or null
  const double *uhost;   // [C*L] the decisions' uniforms, or null (Philox)
  int32_t *idec;         // [C*L] decisions of the last ima_mc3_decide
  int prop_first;        // first proposal slot (slot prop_first + it)
  // the reference's genealogy move (topo = 1): proposals are made on the
  // device in the proposal slots, with Hastings terms thast; accepted ones
  // are copied into the chains' slots by the accept step
  int topo;
  double *thast;         // [C*L]
  const int2 *arange;    // [L] allele range (minA, maxA) of each stepwise locus
  ima_item *pitems;      // [C*L] the proposals' scoring items
  int32_t *ptw;          // [C*L] the proposals' slots
  int maxn;              // largest gene-copy count of the chains' loci
  int mig_pool;          // > 0: the genealogy move draws migration paths, with an LDS
                         // event pool of this many entries per genealogy
  // changet_NW as the other time update (nw = 1): a chain that draws it
  // this sweep has its split nwk[c] (else -1) moving to nwt[c]; the
  // per-genealogy update (kern_nw.hip) reads nwa
  int nw;
  int32_t *nwk;          // [C]
  double *nwt;           // [C]
  NwArgs nwa;
};

// Dynamic LDS of the per-chain prior kernels (kern_mc3.hip), the same on
// host and device: doubles all[R] hold[R] | qt qf nt nf [np] | npd pdg U
// [LC] | dold dnew [LC][cap] | qmax/mmax, log Gamma cache [np], then ints
// qc nc [np] | ok nch dec [LC] | didx [LC][cap] | aff [LC][aw] | the
// model's int block [mints] | counts of the log Gamma cache [np].
// Dynamic LDS of the per-chain prior kernels.  The accept kernel decides
// the loci of a chain in groups of D (1..kSpecDepth): for every parameter
// it evaluates, in one wavefront, the integral under each of the 2^D - 1
// (locus, earlier decisions) scenarios of the group, one lane each.
constexpr int kSpecDepth = 6;
struct PriorLds {
  int R, np, LC, W, aw, mints;  // LC = D loci per group, W = 2^D scenario slots
  __host__ __device__ PriorLds(int R_, int np_, int LC_, int mints_)
      : R(R_), np(np_), LC(LC_), W(1 << LC_), aw((np_ + 31) / 32), mints(mints_) {}
  // per-locus rows come in three buffers: one group is decided while the
  // previous one is committed and the next one's records are read
  __host__ __device__ size_t doubles() const {
    return (size_t)R + 5 * np + 9 * LC + (size_t)6 * LC * R + (size_t)3 * np * W + W;
  }
  __host__ __device__ size_t ints() const {
    return (size_t)2 * np + (size_t)3 * LC * (2 + R + aw) + (size_t)np * W + mints + R;
  }
  __host__ __device__ size_t bytes() const { return doubles() * 8 + ints() * 4; }
};
__host__ __device__ inline int mc3_tw(int P) { return P > 1 ? P - 1 : 1; }

hipError_t enqueue_mc3_sweep(const DeviceView &v, const MC3Dev &m, bool with_scale,
                             const ima_item *items, const int32_t *idx, const HkyPlan &plan,
                             int nh, int ni, int ns, int precision, int hmax_n, int hmax_slots, int imax_n,
                             int imax_w, int swaptries, hipEvent_t *score_ev,
                             hipStream_t st, hipStream_t side, hipEvent_t *fj);

hipError_t enqueue_mc3_propose(const DeviceView &v, const MC3Dev &m, hipStream_t st);
// One scoring pass of the sweep: the items' likelihoods into m.newpdg /
// m.status and (with the prior) the genealogy weights of the slots
// tw_slots[0..C*L) into m.wnew / m.wstatus.
hipError_t enqueue_mc3_score(const DeviceView &v, const MC3Dev &m, const ima_item *items,
                             const int32_t *tw_slots, const int32_t *idx, const HkyPlan &plan,
                             int nh, int ni, int ns, int precision, int hmax_n, int hmax_slots,
                             int imax_n, int imax_w, hipEvent_t *score_ev, int pass,
                             hipStream_t st, hipStream_t side, hipEvent_t *fj);
// ima_mc3_decide's launches after the scoring pass (m.host = 1): the
// decisions of every chain's loci in order, the accepted proposals copied
// into the chains' slots, the sweep counter advanced.
hipError_t enqueue_mc3_decide(const DeviceView &v, const MC3Dev &m, hipStream_t st);
// The cold chain's genealogy-table row (savegsampinf) into row[0..K).
hipError_t enqueue_mc3_record(const DeviceView &v, const MC3Dev &m, float *row, hipStream_t st);
// The reference's genealogy move for every item into its proposal slot
// (kern_topo.hip), dynamic LDS topo_lds_bytes(max_n).
size_t topo_lds_bytes(int max_n);
size_t topo_mig_lds_bytes(int max_n, int pool);
hipError_t enqueue_mc3_propose_topo(const DeviceView &v, const MC3Dev &m, int max_n,
                                    hipStream_t st);
// Items and weight slots of the proposals: items[it] with slot
// prop_first + it, tw_slots[it] = prop_first + it.
hipError_t enqueue_mc3_prop_items(const MC3Dev &m, const ima_item *items, ima_item *pitems,
                                  int32_t *ptw, hipStream_t st);
hipError_t enqueue_mc3_prior_init(const DeviceView &v, const MC3Dev &m, hipStream_t st);
size_t mc3_prior_lds_bytes(int R, int nparam, int pchunk, int mints);
constexpr int kSweepBlock = 8;  // sweeps captured in one graph
constexpr int kURateBlock = 4;  // sweeps per graph when mutation-rate rounds interleave

// Mutation-rate round (changeu, update_mc_params.cpp:24-382; changekappa
// :387-429 for a lone HKY locus).  Each chain runs J steps; step s moves
// the rates of loci s and k_s (k_s drawn first in the step's own stream).
// Steps are grouped into levels: a step's level is one more than the
// highest level of the earlier steps touching either of its loci, so the
// steps of one level touch disjoint loci and each reads the rates the
// sequential loop would have left.
struct URateDev {
  int J;                // steps per chain
  ima_item *uitems;     // [C*L] the items with the proposed rates
  int4 *step;           // [C*J] {k (-1: kappa only), level, 0, 0}
  double4 *uni;         // [C*J] uniforms {window, kappa of s, kappa of k, accept}
  int32_t *order;       // [C*J] each chain's steps sorted by level
  int32_t *loff;        // [C*(J+2)] start of each level in order[] (levels from 1)
  int32_t *last;        // [C*L] plan scratch: highest level touching each locus
  int4 *list[2];        // HKY units of a level's proposals, by level parity
  int32_t *ilist[2];    // IS items of a level's proposals, by level parity
  int32_t *slist[2];    // SW items of a level's proposals, by level parity
  int32_t *cnt;         // [6] entries in list[0..1], ilist[0..1], slist[0..1]
  double win, maxr;     // window and reflection bound of log(u_j / u_k)
  double kwin, kmax;    // kappa window and upper bound
};

// Stream of step s of chain c in the round whose sweep counter is ctr (the
// same on host and device: the host counts the levels from it).
__host__ __device__ inline PhiloxStream urate_rng(uint64_t seed, int c, int L, int s,
                                                  uint64_t ctr) {
  const uint64_t item = (uint64_t)c * (uint64_t)L + (uint64_t)s;
  return PhiloxStream(seed ^ (0x9E3779B97F4A7C15ull * (item + 1)), ((ctr << 3) | 5u) << 16);
}

// k of step s (j = s): uniform over the other loci; with two loci only
// step 0 runs and k = 1; one locus: -1 (kappa only).
__host__ __device__ inline int urate_partner(PhiloxStream &r, int s, int L) {
  if (L == 1) return -1;
  if (L == 2) return 1;
  int k;
  do {
    k = (int)(r.uniform() * L);
  } while (k == s || k < 0 || k >= L);
  return k;
}

// LDS of a round's plan: per-locus levels, level offsets, and per step its
// partner, level and position (L + 4J + 2 ints).
size_t urate_plan_lds(int L, int J);
hipError_t enqueue_urate_round(const DeviceView &v, const MC3Dev &m, const URateDev &u,
                               const ima_item *items, const int32_t *idx, const HkyPlan &plan,
                               int nh, int ni, int ns, int precision, int hmax_n,
                               int hmax_slots, int imax_n, int imax_w, uint64_t ctr,
                               int nlev, int serial, hipStream_t st);
hipError_t launch_hky_list(const DeviceView &v, const ima_item *items, const int4 *list,
                           const int32_t *cnt, int32_t *cnt_next, int grid, int precision,
                           int max_n, int max_slots, int ppl, double *partial, double *out,
                           int32_t *status, hipStream_t st);
constexpr double kTimeMax = 1000000.0;  // the reference's TIMEMAX (imamp.hpp:129)
constexpr int kSwapLds = 256;  // chains whose S and beta the swap stages in LDS
hipError_t enqueue_mc3_revert(const DeviceView &v, const MC3Dev &m, hipStream_t st);

// Swap (kern_swap.hip).
hipError_t launch_swap(int nchains, const double *S, double *beta,
                       int swaptries, uint64_t seed, uint64_t counter,
                       int32_t *nacc, int32_t *pairs, hipStream_t st);

}  // namespace ima
