// ima_internal.h -- shared definitions of the HIP engine (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/ima_gpu.h"

namespace ima {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kMaxGenes = IMA_MAX_GENES;
constexpr int kMaxEdges = 2 * kMaxGenes - 1;
constexpr int kTipSlot = 0xFFFF;   // program operand that is a tip
constexpr double kLn2 = 0.69314718055994530941723212146;

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
  int32_t Spad;      // HKY: row stride of seq (S rounded up to 64, zero padded)
  double pi[4];
  double sumlogk;
};

// Per-slot metadata: where the genealogy and its postorder program live.
struct SlotMeta {
  int32_t locus;
  int32_t n;
  int32_t root;
  int32_t nslots;    // LDS partial slots the program needs
  int64_t node_off;  // into up0/up1/down/time/A arrays (2n-1 entries)
  int64_t prog_off;  // into prog (n-1 ops)
};

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
struct Timing {
  bool on = false;
  hipEvent_t ev0[8]{}, ev1[8]{};
  int64_t launches[8]{};
  double ms[8]{};
};

enum KClass { K_HKY64 = 0, K_HKY32 = 1, K_IS = 2, K_SW = 3, K_LMODE = 4,
              K_SWAP = 5, K_MC3 = 6 };

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
};

// HKY work plan: one wavefront per (item, 64-pattern block).  units[k] =
// {item, block, first partial, blocks of the item}; items with more than
// one block are merged from `partial` in block order by a second kernel
// (red[k] = {item, first partial, blocks, 0}).
struct HkyPlan {
  const int4 *units = nullptr;
  int nunits = 0;
  const int4 *red = nullptr;
  int nred = 0;
  double *partial = nullptr;
};
// Host side: append the units of HKY item ii with S patterns.
void hky_plan_add(std::vector<int4> &units, std::vector<int4> &red, int ii, int S,
                  int *nblocks_total);

// Kernel launchers (kern_hky.hip, kern_lik.hip).
size_t hky_lds_bytes(int precision, int max_n);
hipError_t launch_hky(const DeviceView &v, const ima_item *items, const HkyPlan &plan,
                      int precision, int max_n, int max_slots, double *out,
                      int32_t *status, hipStream_t st);
hipError_t launch_is(const DeviceView &v, const ima_item *items,
                     const int32_t *item_idx, int nitems, int max_n, int max_w,
                     double *out, int32_t *status, int32_t *site_node,
                     int32_t *counts, hipStream_t st);
hipError_t launch_sw(const DeviceView &v, const ima_item *items,
                     const int32_t *item_idx, int nitems, double *out,
                     int32_t *status, double *dlike, hipStream_t st);

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
hipError_t launch_margincalc(const LTable &t, int param, const double *x,
                             int nx, int logi, double yadjust, int expo,
                             double prior_mean, double *work, double *out,
                             hipStream_t st);
size_t margincalc_work_bytes(const LTable &t, int nx);
hipError_t upload_pow10_table(hipStream_t st);

// Device-resident MC3 state (kern_mc3.hip).
struct MC3Dev {
  int C, L;
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
};

hipError_t enqueue_mc3_sweep(const DeviceView &v, const MC3Dev &m, bool with_scale,
                             const ima_item *items, const int32_t *idx, const HkyPlan &plan,
                             int nh, int ni, int ns, int precision, int hmax_n, int hmax_slots, int imax_n,
                             int imax_w, int swaptries, hipEvent_t *score_ev,
                             hipStream_t st);

hipError_t enqueue_mc3_propose(const DeviceView &v, const MC3Dev &m, hipStream_t st);
hipError_t enqueue_mc3_revert(const DeviceView &v, const MC3Dev &m, hipStream_t st);

// Swap (kern_swap.hip).
hipError_t launch_swap(int nchains, const double *S, double *beta,
                       int swaptries, uint64_t seed, uint64_t counter,
                       int32_t *nacc, int32_t *pairs, hipStream_t st);

}  // namespace ima
