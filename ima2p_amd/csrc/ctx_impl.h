// ctx_impl.h -- the context behind the opaque ima_ctx handle (private to
// ctx.hip and serve.hip).
#pragma once

#include "ima_internal.h"

namespace ima {

struct LTableHost {
  bool live = false;
  float *rows = nullptr;
  int G = 0, K = 0, npsize = 0, nmig = 0, npops = 0;
  int cap = 0;  // rows the device buffer holds (G <= cap)
};

}  // namespace ima

using ima::LTableHost;
using ima::DevBuf;
using ima::LocusMeta;
using ima::SlotMeta;
using ima::ProgOp;
using ima::ModelMeta;
using ima::MC3Dev;
using ima::URateDev;
using ima::HkyPlan;
using ima::Timing;

struct ima_ctx {
  int device = 0;
  void *scratch = nullptr;  // this device's address of the static work area
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // genealogy weights beside the likelihood launch
  std::string err;

  // loci (host mirror + device copy)
  std::vector<LocusMeta> loci;
  std::vector<int8_t> seq;
  std::vector<int32_t> mult;
  std::vector<uint64_t> mask;
  bool loci_dirty = true;
  DevBuf<LocusMeta> d_loci;
  DevBuf<int8_t> d_seq;
  DevBuf<int32_t> d_mult;
  DevBuf<uint64_t> d_mask;

  // genealogy slots
  std::vector<SlotMeta> slots;
  size_t nodes = 0, progs = 0;
  std::vector<int32_t> up0, up1, down, A;
  std::vector<double> time;
  std::vector<ProgOp> prog;
  DevBuf<SlotMeta> d_slots;
  DevBuf<int32_t> d_up0, d_up1, d_down, d_A;
  DevBuf<double> d_time;
  DevBuf<ProgOp> d_prog;
  bool slots_meta_dirty = true;
  std::vector<int> scratch_slot;  // per locus, for ima_likelihood_hky
  // population labels and migration events of the genealogies (prior)
  std::vector<int32_t> pop, mig_edge, mig_to, mig_first, mig_count;
  std::vector<double> mig_time;
  size_t migs = 0;
  DevBuf<int32_t> d_pop, d_mig_edge, d_mig_to, d_mig_first, d_mig_count;
  DevBuf<double> d_mig_time;
  // the NW update's backups (indexed like the node and migration arrays)
  // and per-call arrays
  DevBuf<int32_t> d_nw_pop, d_nw_mf, d_nw_mc, d_nw_me, d_nw_mto;
  DevBuf<double> d_nw_mt;
  DevBuf<int32_t> d_nw_nbk, d_nw_st, d_nw_period;
  DevBuf<double> d_nw_w, d_nw_newt;
  DevBuf<int64_t> d_mv_off;  // event regions moved to larger ones: from | to
  DevBuf<int32_t> d_mv_cnt;

  // population models of the genealogy prior
  std::vector<ModelMeta> models;
  std::vector<int32_t> mint;
  std::vector<double> mdbl, logfact;
  bool models_dirty = true;
  size_t logfact_dev = 0;
  DevBuf<ModelMeta> d_models;
  DevBuf<int32_t> d_mint;
  DevBuf<double> d_mdbl, d_logfact;
  DevBuf<double> d_rec, d_tv, d_probg, d_qint, d_mintg;
  DevBuf<int32_t> d_pslots, d_prow, d_goff, d_pstatus;

  // scoring scratch
  DevBuf<ima_item> d_items;
  DevBuf<int32_t> d_idx;
  DevBuf<double> d_out;
  DevBuf<int32_t> d_status;
  DevBuf<int32_t> d_aux;
  DevBuf<double> d_aux64;
  DevBuf<ima_sw_move> d_swmoves;  // ima_sw_update_A's moves
  DevBuf<int4> d_units, d_red;
  DevBuf<double> d_partial;

  // L mode
  std::vector<LTableHost> tables;
  DevBuf<double> d_lx, d_lwork, d_lout, d_less;

  // swap
  DevBuf<double> d_S, d_beta;
  DevBuf<int32_t> d_pairs;

  Timing timing;
  int max_lds = 65536;
  int site_block = 0;  // HKY site-block size: 0 automatic, 64 or 128 patterns

  // device-resident MC3 sweep
  struct {
    bool live = false;
    int C = 0, L = 0, first = 0, precision = IMA_PREC_FP32, swaptries = 1;
    int nh = 0, ni = 0, ns = 0, hmax_n = 2, hmax_slots = 1, imax_n = 2, imax_w = 1;
    MC3Dev dev{};
    DevBuf<double> pdg, newpdg, beta, S, rec_t, tbackup, logs;
    DevBuf<int32_t> status, rec_v, decision, idx, node_item;
    DevBuf<int2> iblk;
    DevBuf<unsigned> done;
    DevBuf<uint32_t> acc, cacc;
    int64_t scale_sweeps = 0;
    uint64_t host_sweeps = 0;  // sweeps enqueued since setup (the device counter's copy)
    DevBuf<uint64_t> counter;
    DevBuf<unsigned long long> stats;
    DevBuf<ima_item> items;
    DevBuf<int4> units, red;
    DevBuf<double> partial;
    // genealogy prior
    DevBuf<double> tv, tvbackup, wrec, wnew, wall, qterm, qf, probg, mtbackup;
    DevBuf<int32_t> qc, tv_row, tw_slots, wstatus;
    HkyPlan plan;
    hipGraphExec_t exec[2]{nullptr, nullptr};  // one sweep without / with rescale
    hipGraph_t graph[2]{nullptr, nullptr};
    // kSweepBlock sweeps in one graph (from an even sweep index), without /
    // with the alternating rescale
    hipGraphExec_t bexec[2]{nullptr, nullptr};
    hipGraph_t bgraph[2]{nullptr, nullptr};
    hipEvent_t ev[4]{};
    hipEvent_t fj[4]{};  // fork / join of the side stream, per scoring pass
    // mutation-rate rounds: kURateBlock sweeps in one graph, by rescale and
    // by the parity of the first sweep
    hipGraphExec_t uexec[2][2]{{nullptr, nullptr}, {nullptr, nullptr}};
    hipGraph_t ugraph[2][2]{{nullptr, nullptr}, {nullptr, nullptr}};
    URateDev udev{};
    DevBuf<ima_item> uitems;
    DevBuf<int4> ustep, ulist0, ulist1;
    DevBuf<double4> uuni;
    DevBuf<int32_t> uorder, uloff, ulast, ucnt, uilist0, uilist1, uslist0, uslist1;
    std::vector<int32_t> ulevel_scratch;  // host copy of the level bookkeeping
    // ima_mc3_decide: Hastings terms, uniforms, decisions, proposal items
    DevBuf<double> hast, uhost;
    DevBuf<int32_t> idec, ptw;
    DevBuf<ima_item> pitems;
    std::vector<double> beta0;  // the ladder set up (swaps permute it)
    // the reference's genealogy move (IMA_MC3_TOPOLOGY): proposal slots
    // made once, their Hastings terms and items, stepwise allele ranges
    int topo_first = -1;
    DevBuf<double> thast;
    DevBuf<ima_item> titems;
    DevBuf<int32_t> tptw;
    DevBuf<int2> arange;
    // changet_NW as the other time update (IMA_MC3_NW)
    DevBuf<int32_t> nwk, nwst, nwnbk;
    DevBuf<double> nwt, nwmig;
    // device buffers the captured graphs point at: a change (a buffer grown
    // by later slot allocations or writes) drops the graphs
    ima::DeviceView gview{};
    bool gview_set = false;
  } mc3;
};


// Helpers shared by ctx.hip and serve.hip.
int mc3_run_async(ima_ctx *c, int nsweeps, int flags);
ima::LTable ctx_ltable(ima_ctx *c, int table_id);
int ctx_set_error(ima_ctx *c, int rc, const char *msg);
size_t ctx_device_bytes(const ima_ctx *c);
