// kern_mc3.hip -- device-resident Metropolis-coupled sweep.
//
// One sweep = for every chain and locus one genealogy proposal scored in a
// single batched likelihood launch and accepted or rejected
// (updategenealogy's scoring, update_gtree.cpp:827-961); on alternate
// sweeps a whole-chain time rescaling whose acceptance needs a full
// re-score of every locus (the role of changet_RY1, update_t_RY.cpp:343-388);
// then the heated-chain swap (swapchains.cpp:527-662).  Nothing returns to
// the host inside a sweep, so a sweep is captured once into a hipGraph and
// replayed.
//
// Proposals.  The genealogy proposal moves the time of one coalescence
// (chosen uniformly) by a normal step reflected into the interval bounded
// by its daughters' coalescences and its parent's; with the genealogy prior
// on, also by the period it lies in (split times) and by the migrations on
// its daughter edges (below) and its parent edge (above), so every
// population label and migration stays valid.  The step scale depends only
// on the bounds, so the move is symmetric.  Topology and migrations are
// unchanged.  The rescaling multiplies every coalescence time of a chain's
// genealogies -- and with the prior every migration time and split time --
// by s = exp(d (U - 1/2)), with Hastings term (number of scaled times) log s.
// Neither is the reference's proposal (slider/addmigration); they exercise
// the same scoring, acceptance, commit/rollback and swap path.
//
// Acceptance (update_gtree.cpp:919-927): U < min(1, exp(beta (dprobg +
// dpdg))).  Without the prior dprobg = 0 and the loci of a chain are
// independent given its parameters, so they are decided in parallel.  With
// the prior, loci are coupled through the chain's all-loci weight sums
// (allgweight), so one wavefront per chain takes the loci in order: the sums
// become max(0, (all - old) + new) as sum_subtract_treeinfo forms them
// (ginfo.cpp:249-285), a parameter's integral is recomputed only when its
// (count, weight) changed (integrate_tree_prob's reuse, update_gtree_common
// .cpp:1997-2006), and a rejection restores the held sums.
//
// Kernels per sweep: the batched likelihood (plus the genealogy weights of
// the proposals with the prior on), then one 1024-thread workgroup per
// chain that accepts or rejects every locus of its chain (merging the
// likelihood's pattern-block partials in block order), applies the
// rescaling or computes the chain sum, and draws the next sweep's
// genealogy proposals.  The workgroup that finishes last (device-scope
// fence + counter) performs the swap.  A rescaling sweep adds a second
// scoring pass and a per-chain commit kernel.  Every reduction has a fixed
// order, so results do not depend on scheduling.
#include <float.h>

#include "ima_internal.h"
#include "philox.h"
#include "prior_math.h"

namespace ima {

// Every (key, sweep, phase) stream gets its own range of 2^16 counter
// blocks: a stream advances its counter by one per four 32-bit words, so
// streams whose starting counters were adjacent would share blocks (and
// hand the same uniform to, say, a proposal and its acceptance test).
constexpr int kStreamBits = 16;
__device__ inline uint64_t stream_counter(uint64_t sweep, int phase) {
  return ((sweep << 3) | (uint64_t)phase) << kStreamBits;
}

__device__ inline PhiloxStream item_rng(const MC3Dev &m, int item, uint64_t sweep, int phase) {
  return PhiloxStream(m.seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(item + 1)),
                      stream_counter(sweep, phase));
}

__device__ inline double normal01(PhiloxStream &r) {
  const double u1 = 1.0 - r.uniform(), u2 = r.uniform();  // u1 in (0, 1]
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Move one coalescence time of item `it` for sweep `sweep`; remembers the
// edge and old time for rejection.
__device__ void propose_item(const DeviceView &v, const MC3Dev &m, int it, uint64_t sweep) {
  const SlotMeta sm = v.slots[m.first_slot + it];
  const int n = sm.n;
  const int32_t *up0 = v.up0 + sm.node_off, *up1 = v.up1 + sm.node_off;
  const int32_t *dn = v.down + sm.node_off;
  double *tm = const_cast<double *>(v.time) + sm.node_off;
  PhiloxStream r = item_rng(m, it, sweep, 0);
  int e = n + (int)(r.uniform() * (n - 1));
  if (e > 2 * n - 2) e = 2 * n - 2;
  const int a = up0[e], b = up1[e];
  const double T = tm[a];
  const double la = (a >= n) ? tm[up0[a]] : 0.0;
  const double lb = (b >= n) ? tm[up0[b]] : 0.0;
  double lo = fmax(la, lb);
  double hi = (dn[e] == -1) ? DBL_MAX : tm[e];
  if (m.prior >= 0) {
    // stay inside the period and between the migrations around the node
    const double *tv = m.tv + (int64_t)(it / m.L) * mc3_tw(m.P);
    int k = 0;
    while (k < m.P - 1 && tv[k] <= T) k++;
    if (k > 0) lo = fmax(lo, tv[k - 1]);
    if (k < m.P - 1) hi = fmin(hi, tv[k]);
    // an edge's migrations are in time order: the last ones on the two
    // daughter edges and the first one on the node's own edge bound it
    const double *mt = v.mig_time + sm.mig_off;
    const int32_t *mf = v.mig_first + sm.node_off, *mc = v.mig_count + sm.node_off;
    const int ca = mc[a], cb = mc[b], ce = mc[e];
    if (ca) lo = fmax(lo, mt[mf[a] + ca - 1]);
    if (cb) lo = fmax(lo, mt[mf[b] + cb - 1]);
    if (ce) hi = fmin(hi, mt[mf[e]]);
  }
  const bool open = hi == DBL_MAX;
  const double sigma = open ? 0.5 * fmax(lo, 0.05) : 0.5 * (hi - lo);
  double Tn = T + sigma * normal01(r);
  if (open) {
    Tn = lo + fabs(Tn - lo);
  } else {
    const double w = hi - lo;
    double y = fmod(Tn - lo, 2.0 * w);
    if (y < 0) y += 2.0 * w;
    if (y > w) y = 2.0 * w - y;
    Tn = lo + y;
  }
  if (!(Tn > lo) || !(Tn < hi)) Tn = T;  // keep every branch strictly positive
  m.rec_v[it] = e;
  m.rec_t[it] = T;
  tm[a] = Tn;
  tm[b] = Tn;
}

__device__ inline void restore_item(const DeviceView &v, const MC3Dev &m, int it) {
  const SlotMeta sm = v.slots[m.first_slot + it];
  const int e = m.rec_v[it];
  double *tm = const_cast<double *>(v.time) + sm.node_off;
  const int32_t *up0 = v.up0 + sm.node_off, *up1 = v.up1 + sm.node_off;
  tm[up0[e]] = m.rec_t[it];
  tm[up1[e]] = m.rec_t[it];
}

// An accepted proposal (slot prop_first + it) replaces the chain's genealogy
// (slot first_slot + it): edges, times, populations, alleles, program,
// migrations, and the slot's root, stack depth and migration count.
// Called by threads lane = 0 .. nl - 1 of one workgroup (nl >= 1); the
// fields of the chain slot's record it writes are not the ones it reads.
__device__ void copy_proposal(const DeviceView &v, const MC3Dev &m, int it, int lane, int nl) {
  SlotMeta *slots = const_cast<SlotMeta *>(v.slots);
  const SlotMeta src = slots[m.prop_first + it];
  SlotMeta &dst = slots[m.first_slot + it];
  const int64_t so = src.node_off, d = dst.node_off;
  const int ne = 2 * src.n - 1;
  int32_t *up0 = const_cast<int32_t *>(v.up0), *up1 = const_cast<int32_t *>(v.up1);
  int32_t *dn = const_cast<int32_t *>(v.down), *pop = const_cast<int32_t *>(v.pop);
  int32_t *A = const_cast<int32_t *>(v.A);
  int32_t *mf = const_cast<int32_t *>(v.mig_first), *mc = const_cast<int32_t *>(v.mig_count);
  double *tm = const_cast<double *>(v.time);
  for (int e = lane; e < ne; e += nl) {
    up0[d + e] = up0[so + e];
    up1[d + e] = up1[so + e];
    dn[d + e] = dn[so + e];
    tm[d + e] = tm[so + e];
    if (pop) pop[d + e] = pop[so + e];
    if (A) A[d + e] = A[so + e];
    if (mf) mf[d + e] = mf[so + e];
    if (mc) mc[d + e] = mc[so + e];
  }
  ProgOp *prog = const_cast<ProgOp *>(v.prog);
  for (int k = lane; k < src.n - 1; k += nl) prog[dst.prog_off + k] = prog[src.prog_off + k];
  int32_t *me = const_cast<int32_t *>(v.mig_edge), *mto = const_cast<int32_t *>(v.mig_to);
  double *mt = const_cast<double *>(v.mig_time);
  for (int k = lane; k < src.nmig; k += nl) {
    me[dst.mig_off + k] = me[src.mig_off + k];
    mto[dst.mig_off + k] = mto[src.mig_off + k];
    mt[dst.mig_off + k] = mt[src.mig_off + k];
  }
  if (lane == 0) {
    dst.root = src.root;
    dst.nslots = src.nslots;
    dst.nmig = src.nmig;
  }
}

// Proposals of the current sweep for every item (first sweep of a run).
__global__ void mc3_propose(DeviceView v, MC3Dev m) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it < m.C * m.L) propose_item(v, m, it, *m.counter);
}

// Undo the pending proposals at the end of a run, so slot contents and
// stored likelihoods agree between runs.
__global__ void mc3_revert(DeviceView v, MC3Dev m) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it < m.C * m.L) restore_item(v, m, it);
}

// Likelihood of item `it` from the batched launch: HKY items spanning
// several pattern blocks are merged here in block order.
__device__ inline double new_loglik(const MC3Dev &m, int it) {
  // both loads in flight together: most items are one block
  const double direct = m.newpdg[it];
  const int2 ib = m.iblk[it];
  if (ib.y <= 1) return direct;
  double s = 0.0;
  for (int b = 0; b < ib.y; ++b) s += m.partial[ib.x + b];
  return s;
}

// Deterministic block reduction (fixed order) of one double per thread.
__device__ inline double block_sum(double x, double *sh) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  x = wsum(x);
  __syncthreads();
  if (lane == 0) sh[w] = x;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < (int)(blockDim.x / kWave); ++k) s += sh[k];
  __syncthreads();
  return s;
}

__device__ inline void chain_nodes(const DeviceView &v, const MC3Dev &m, int c, int64_t *j0,
                                   int64_t *j1) {
  const SlotMeta s0 = v.slots[m.first_slot + c * m.L];
  const SlotMeta s1 = v.slots[m.first_slot + c * m.L + m.L - 1];
  *j0 = s0.node_off - m.node_first;
  *j1 = s1.node_off + 2 * (int64_t)s1.n - 1 - m.node_first;
}

__device__ inline double chain_log_scale(const MC3Dev &m, int c, uint64_t sweep) {
  PhiloxStream r = item_rng(m, m.C * m.L + c, sweep, 2);
  return m.scale_width * (r.uniform() - 0.5);
}

// swapchains (swapchains.cpp:527-653) on one lane: partner within SWAPDIST
// = 7 unless fewer than 2*SWAPDIST+3 chains, w = exp((b_i - b_j)(S_j - S_i))
// (swapweight :13-34), accept when w >= 1 or w > U, swap the betas.
// S and beta are staged in LDS (up to kSwapLds chains) first.
__device__ void swap_chains(const MC3Dev &m, int swaptries, uint64_t sweep, double *Ssh,
                            double *Bsh) {
  const bool staged = m.C <= kSwapLds;
  if (staged)
    for (int c = threadIdx.x; c < m.C; c += blockDim.x) {
      Ssh[c] = m.S[c];
      Bsh[c] = m.beta[c];
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    // each uncoupled analysis (m.group chains) swaps among its own chains,
    // from a stream of its own; one analysis keeps the first stream
    const int SWAPDIST = 7, nchains = m.group, ngroups = m.C / m.group;
    unsigned long long acc = 0;
    for (int g = 0; g < ngroups; ++g) {
    PhiloxStream rng(m.seed ^ 0x5A5A5A5A5A5A5A5Aull ^ (0x9E3779B97F4A7C15ull * (uint64_t)g),
                     stream_counter(sweep, 4));
    const double *S = (staged ? Ssh : m.S) + g * nchains;
    double *beta = (staged ? Bsh : m.beta) + g * nchains;
    for (int i = 0; i < swaptries && nchains > 1; ++i) {
      int ci, cj, cjmin, cjrange;
      do {
        ci = (int)(rng.uniform() * nchains);
      } while (ci < 0 || ci >= nchains);
      if (nchains < 2 * SWAPDIST + 3) {
        cjmin = 0;
        cjrange = nchains;
      } else {
        cjmin = max(0, ci - SWAPDIST);
        cjrange = min(nchains, ci + SWAPDIST) - cjmin;
      }
      do {
        cj = cjmin + (int)(rng.uniform() * cjrange);
      } while (cj == ci || cj < 0 || cj >= nchains);
      const double wgt = exp((beta[ci] - beta[cj]) * (S[cj] - S[ci]));
      if (wgt >= 1.0 || wgt > rng.uniform()) {
        const double b = beta[ci];
        beta[ci] = beta[cj];
        beta[cj] = b;
        ++acc;
      }
    }
    }
    m.stats[4] += (unsigned long long)(nchains > 1 ? swaptries * ngroups : 0);
    m.stats[5] += acc;
    m.stats[0] += (unsigned long long)(m.C * m.L);
    *m.counter = sweep + 1;
    *m.done = 0u;
  }
  __syncthreads();
  if (staged)
    for (int c = threadIdx.x; c < m.C; c += blockDim.x) m.beta[c] = Bsh[c];
}

// Chain-sum epilogue shared by the per-chain kernels: publish S_c, count
// finished chains; the last chain runs the swap.
__device__ void finish_chain(const MC3Dev &m, int c, double S, int swaptries, uint64_t sweep,
                             double *Ssh, double *Bsh) {
  __shared__ int last;
  if (threadIdx.x == 0) {
    m.S[c] = S;
    __threadfence();
    last = (atomicAdd(m.done, 1u) == (unsigned)(m.C - 1));
  }
  __syncthreads();
  if (last) {
    __threadfence();
    swap_chains(m, swaptries, sweep, Ssh, Bsh);
  }
}

// Intra-wavefront ordering of LDS traffic (the accept loop runs in one wave).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane l's value of x, for every lane (l uniform).
__device__ inline double readlane_d(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Entry r of the all-loci sums after replacing one genealogy's weights
// (sum_subtract_treeinfo, ginfo.cpp:249-285): counts add the difference,
// fc / fm subtract then add and stay >= 0, hcc subtracts then adds.
__device__ inline double wsum_update(int r, int P, double all, double old, double nw) {
  const int P2 = P * P, P3 = P2 * P;
  if (r < P2) return all + (nw - old);
  if (r < 2 * P2) return fmax(0.0, (all - old) + nw);
  if (r < 3 * P2) return (all - old) + nw;
  if (r < 3 * P2 + P3) return all + (nw - old);
  if (r < 3 * P2 + 2 * P3) return fmax(0.0, (all - old) + nw);
  return (all - old) + nw;
}

// Pointers into the dynamic LDS of the per-chain prior kernels (PriorLds).
//   all[R]           the chain's all-loci sums
//   qt/qf/qc[np]     committed integral, weight, count of every parameter
//   nt/nf/nc[np]     the same for the locus being decided
//   npd/pdg/U/ok/dec per locus of the group
//   dold/dnew/flag   [D][R] the group's committed and proposed records and
//                    which entries differ
//   aff[D][aw]       parameters whose entries a locus changed
//   sterm/sf/sc      [np][W] per (parameter, scenario) integral, weight, count
//   su               [np][W] the term a scenario's decision uses
//   pns              [W] a scenario's prior total (sum of su over parameters)
//   fmk              [R] bit i: locus i of the group changed the entry
struct PriorSh {
  double *all, *qt, *qf, *nt, *nf, *npd, *pdg, *U, *dold, *dnew, *mdl, *sterm, *sf, *su, *pns;
  int *qc, *nc, *ok, *dec, *flag, *sc, *mi, *fmk;
  uint32_t *aff;
};
__device__ inline PriorSh prior_carve(double *dyn, const PriorLds &Lz) {
  PriorSh p;
  const int L2 = 3 * Lz.LC;  // three buffers of per-locus rows
  const size_t DR = (size_t)L2 * Lz.R, NW = (size_t)Lz.np * Lz.W;
  p.all = dyn;
  p.qt = p.all + Lz.R;
  p.qf = p.qt + Lz.np;
  p.nt = p.qf + Lz.np;
  p.nf = p.nt + Lz.np;
  p.mdl = p.nf + Lz.np;
  p.npd = p.mdl + Lz.np;
  p.pdg = p.npd + L2;
  p.U = p.pdg + L2;
  p.dold = p.U + L2;
  p.dnew = p.dold + DR;
  p.sterm = p.dnew + DR;
  p.sf = p.sterm + NW;
  p.su = p.sf + NW;
  p.pns = p.su + NW;
  p.qc = reinterpret_cast<int *>(p.pns + Lz.W);
  p.nc = p.qc + Lz.np;
  p.ok = p.nc + Lz.np;
  p.dec = p.ok + L2;
  p.flag = p.dec + L2;
  p.sc = p.flag + DR;
  p.aff = reinterpret_cast<uint32_t *>(p.sc + NW);
  p.mi = reinterpret_cast<int *>(p.aff + (size_t)L2 * Lz.aw);
  p.fmk = p.mi + Lz.mints;
  return p;
}

// All-loci sums of chain c from the records w (loci in order), every
// parameter's (count, weight, integral) and probg = sum of the integrals;
// called by the whole workgroup.
__device__ double chain_prior_fresh(const DeviceView &v, const MC3Dev &m, int c,
                                    const double *w, const PriorSh &sh) {
  __shared__ double pg;
  const int R = m.R, np = m.nparam;
  const ModelMeta mm = v.models[m.prior];
  const ModelView md = model_view(v, mm);
  // the chain's records go through LDS a block of loci at a time (read
  // coalesced by the whole workgroup, in the accept step's record rows and
  // scenario arrays: dold .. pns are contiguous), each entry summed over
  // the loci in order
  const PriorLds Lz(R, np, m.pchunk, m.mints);
  const int chunk = (int)((6 * (size_t)Lz.LC * R + 3 * (size_t)np * Lz.W + Lz.W) / R);
  double *stage = sh.dold;
  const double *base = w + (int64_t)c * m.L * R;
  for (int l0 = 0; l0 < m.L; l0 += chunk) {
    const int nl = min(chunk, m.L - l0);
    const int cnt = nl * R;
    __syncthreads();
    for (int x = threadIdx.x; x < cnt; x += blockDim.x) stage[x] = base[(int64_t)l0 * R + x];
    __syncthreads();
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
      double acc = l0 == 0 ? 0.0 : sh.all[r];
      for (int i = 0; i < nl; ++i) acc += stage[i * R + r];
      sh.all[r] = acc;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    int cc;
    double f, h;
    prior_param_sums(md, m.P, mm.npsize, t, sh.all, &cc, &f, &h);
    sh.nc[t] = cc;
    sh.nf[t] = f;
    sh.nt[t] = prior_param_term(md, mm.npsize, t, cc, f, h, v.logfact);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double q = 0.0, mg = 0.0;
    for (int t = 0; t < mm.npsize; ++t) q += sh.nt[t];
    for (int t = mm.npsize; t < np; ++t) mg += sh.nt[t];
    pg = q + mg;
  }
  __syncthreads();
  return pg;
}

// Prior state of every chain from the committed weight records.
__global__ __launch_bounds__(256) void mc3_prior_init(DeviceView v, MC3Dev m) {
  extern __shared__ double dyn[];
  const int c = blockIdx.x, R = m.R, np = m.nparam;
  const PriorSh sh = prior_carve(dyn, PriorLds(R, np, m.pchunk, m.mints));
  const double pg = chain_prior_fresh(v, m, c, m.wrec, sh);
  for (int r = threadIdx.x; r < R; r += blockDim.x) m.wall[(int64_t)c * R + r] = sh.all[r];
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    m.qterm[c * np + t] = sh.nt[t];
    m.qf[c * np + t] = sh.nf[t];
    m.qc[c * np + t] = sh.nc[t];
  }
  if (threadIdx.x == 0) m.probg[c] = pg;
}

// True when parameter t adds up a record entry flagged in flag[0..R).
__device__ inline bool param_hit(const ModelView &md, int P, int npsize, int t,
                                 const int *flag) {
  const int P2 = P * P, P3 = P2 * P;
  if (t < npsize) {
    for (int w = md.theta_off[t]; w < md.theta_off[t + 1]; ++w) {
      const int e = md.theta_wp[2 * w] * P + md.theta_wp[2 * w + 1];
      if (flag[e] | flag[P2 + e] | flag[2 * P2 + e]) return true;
    }
  } else {
    const int mm = t - npsize;
    for (int w = md.mig_off[mm]; w < md.mig_off[mm + 1]; ++w) {
      const int e = (md.mig_wp[3 * w] * P + md.mig_wp[3 * w + 1]) * P + md.mig_wp[3 * w + 2];
      if (flag[3 * P2 + e] | flag[3 * P2 + P3 + e]) return true;
    }
  }
  return false;
}

// Loci of chain c in order, each accepted with beta (dprobg + dpdg), where
// a locus' prior term depends on the all-loci sums the earlier decisions
// left.  The loci are taken in groups of D:
//  1. one wavefront per locus reads its likelihood, uniform and records,
//     flags the entries its proposal changed and marks the parameters they
//     feed (for group g+2 while group g is being decided);
//  2. one wavefront per parameter evaluates, one lane per scenario, the
//     parameter's sums and integral for every locus j of the group under
//     every combination S of the decisions on loci 0..j-1 (2^D - 1 lanes),
//     applying sum_subtract_treeinfo's rules in locus order;
//  3. the first wavefront walks the group in order, picking each locus'
//     terms from the scenario its actual earlier decisions select (the
//     committed term when count and weight are unchanged,
//     update_gtree_common.cpp:1997-2006), deciding, and applying the
//     accepted changes to the sums in locus order;
//  4. the group's records are committed or restored while the next group
//     runs step 2.
// The decisions and sums are those of deciding the loci one at a time.

// Step 1 for the loci [l0, l0 + nl) of chain c into buffer rows b0.., by
// wavefronts w0.. of the workgroup.
__device__ inline void prior_fetch(const MC3Dev &m, const ModelView &md, int npsize,
                                   const PriorSh &sh, const PriorLds &Lz, int c, int l0,
                                   int nl, int b0, int w0, uint64_t sweep) {
  const int R = m.R, np = m.nparam, P = m.P;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int nwaves = blockDim.x / kWave;
  for (int x = wave - w0; x < nl; x += nwaves - w0) {
    if (x < 0) break;
    const int it = c * m.L + l0 + x, row = b0 + x;
    const double *wo = m.wrec + (int64_t)it * R, *wn = m.wnew + (int64_t)it * R;
    for (int r = lane; r < R; r += kWave) {
      const double o = wo[r], w = wn[r];
      sh.dold[row * R + r] = o;
      sh.dnew[row * R + r] = w;
      sh.flag[row * R + r] = o != w ? 1 : 0;
    }
    if (lane == 0) {
      sh.ok[row] = m.status[it] == IMA_STATUS_OK && m.wstatus[it] == IMA_STATUS_OK;
      sh.npd[row] = new_loglik(m, it);
      sh.pdg[row] = m.pdg[it];
      PhiloxStream rr = item_rng(m, it, sweep, 1);
      // the decision U < min(1, e^x) is taken as log U < x (and U < 1);
      // +inf stands for U = 1.  A host proposal's Hastings term h moves to
      // this side: log U - h < beta (dprobg + gbeta dpdg)
      const double U = m.uhost ? m.uhost[it] : rr.uniform();
      sh.U[row] = (U < 1.0 ? log(U) : INFINITY) - (m.hast ? m.hast[it] : 0.0);
    }
    wave_sync();
    for (int t0 = 0; t0 < np; t0 += kWave) {
      const int t = t0 + lane;
      const bool hit = t < np && param_hit(md, P, npsize, t, sh.flag + row * R);
      const unsigned long long bal = __ballot(hit);
      if (lane == 0) {
        sh.aff[row * Lz.aw + t0 / 32] = (uint32_t)bal;
        if (t0 / 32 + 1 < Lz.aw) sh.aff[row * Lz.aw + t0 / 32 + 1] = (uint32_t)(bal >> 32);
      }
    }
  }
}

// Step 4 for the loci [l0, l0 + nl) held in buffer rows b0.., by the
// workgroup's wavefronts from the last one down.
__device__ inline void prior_commit(const DeviceView &v, const MC3Dev &m, const PriorSh &sh,
                                    int c, int l0, int nl, int b0) {
  const int R = m.R;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int nwaves = blockDim.x / kWave;
  for (int x = nwaves - 1 - wave; x < nl; x += nwaves) {
    const int it = c * m.L + l0 + x, row = b0 + x;
    const bool acc = sh.dec[row] != 0;
    if (acc) {
      double *wc = m.wrec + (int64_t)it * R;
      for (int r = lane; r < R; r += kWave)
        if (sh.flag[row * R + r]) wc[r] = sh.dnew[row * R + r];
    }
    if (lane == 0) {
      if (acc)
        m.pdg[it] = sh.npd[row];
      else if (!m.host && !m.topo)
        restore_item(v, m, it);
      if (m.host || m.topo) m.idec[it] = acc ? 1 : 0;
      m.acc[it] += acc ? 1u : 0u;
    }
  }
}

__device__ void prior_accept(const DeviceView &v, const MC3Dev &m, int c, uint64_t sweep,
                             double *dyn) {
  const int R = m.R, np = m.nparam, P = m.P;
  const PriorLds Lz(R, np, m.pchunk, m.mints);
  const int D = Lz.LC, W = Lz.W;
  const PriorSh sh = prior_carve(dyn, Lz);
  // the model's maps live in LDS, so the decision loop reads no global memory
  ModelMeta mm = v.models[m.prior];
  for (int x = threadIdx.x; x < m.mints; x += blockDim.x) sh.mi[x] = v.mint[mm.ioff + x];
  for (int t = threadIdx.x; t < np; t += blockDim.x) sh.mdl[t] = v.mdbl[mm.doff + t];
  mm.ioff = 0;
  mm.doff = 0;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const int nwaves = blockDim.x / kWave;
  for (int r = threadIdx.x; r < R; r += blockDim.x) sh.all[r] = m.wall[(int64_t)c * R + r];
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    sh.qt[t] = m.qterm[c * np + t];
    sh.qf[t] = m.qf[c * np + t];
    sh.qc[t] = m.qc[c * np + t];
  }
  __syncthreads();
  const ModelView md = model_view(sh.mi, sh.mdl, mm);
  double probg = m.probg[c];
  const double beta = m.beta[c];
  // the first two groups; later ones are fetched two ahead while wave 0
  // decides (their rows are those of the group committed just before)
  prior_fetch(m, md, mm.npsize, sh, Lz, c, 0, min(D, m.L), 0, 0, sweep);
  if (D < m.L) prior_fetch(m, md, mm.npsize, sh, Lz, c, D, min(D, m.L - D), D, 0, sweep);
  __syncthreads();
  // which loci of the group change each entry (group 0 here, group g + 1
  // beside the sums' update of group g)
  auto entry_masks = [&](int r, int b, int n) {
    int mk = 0;
    for (int j = 0; j < n; ++j) mk |= (sh.flag[(b + j) * R + r] ? 1 : 0) << j;
    sh.fmk[r] = mk;
  };
  for (int r = threadIdx.x; r < R; r += blockDim.x) entry_masks(r, 0, min(D, m.L));
  __syncthreads();
  // np <= 64: each scenario's prior total is formed before the decisions
  // (one lane per scenario), and a decision reads it from LDS
  const bool regs = np <= kWave;
  int pl0 = 0, pnl = 0, pb0 = 0;  // the previous group, still to commit
  for (int l0 = 0, g = 0; l0 < m.L; l0 += D, ++g) {
    const int nl = min(D, m.L - l0), b0 = (g % 3) * D, nb0 = ((g + 2) % 3) * D;
    const int nnl = l0 + 2 * D < m.L ? min(D, m.L - l0 - 2 * D) : 0;  // group g + 2
    // 2. one wavefront per parameter, one lane per (locus j, decisions S)
    for (int t = wave; t < np; t += nwaves) {
      const int s = lane;
      if (s < W - 1) {
        const int j = 31 - __clz(s + 1);
        const int S = s + 1 - (1 << j);
        if (j < nl && sh.ok[b0 + j] &&
            ((sh.aff[(b0 + j) * Lz.aw + t / 32] >> (t % 32)) & 1u)) {
          // the accepted loci of S and locus j, in order, that changed entry e
          const uint32_t sel = (uint32_t)S | (1u << j);
          auto get = [&](int e) {
            double val = sh.all[e];
            for (uint32_t bits = (uint32_t)sh.fmk[e] & sel; bits; bits &= bits - 1) {
              const int row = b0 + (__ffs(bits) - 1);
              val = wsum_update(e, P, val, sh.dold[row * R + e], sh.dnew[row * R + e]);
            }
            return val;
          };
          int cc;
          double f, h;
          prior_param_sums_of(md, P, mm.npsize, t, get, &cc, &f, &h);
          sh.sc[t * W + s] = cc;
          sh.sf[t * W + s] = f;
          sh.sterm[t * W + s] = prior_param_term(md, mm.npsize, t, cc, f, h, v.logfact);
        }
      }
      if (regs) {
        // the term scenario (j, S) uses for parameter t: from the group's
        // committed (term, weight, count), each accepted locus i of S that
        // marks t, in order, then locus j, takes its scenario's values,
        // keeping the committed term when count and weight are unchanged
        // (update_gtree_common.cpp:1997-2006)
        wave_sync();  // the scenario values of this wavefront's lanes
        if (s < W - 1) {
          const int j = 31 - __clz(s + 1);
          const int S = s + 1 - (1 << j);
          if (j < nl) {
            double term = sh.qt[t], f = sh.qf[t];
            int cc = sh.qc[t];
            for (int i = 0; i <= j; ++i) {
              if (i < j && !((S >> i) & 1)) continue;
              if (!((sh.aff[(b0 + i) * Lz.aw + t / 32] >> (t % 32)) & 1u)) continue;
              const int si = i == j ? s : (1 << i) - 1 + (S & ((1 << i) - 1));
              const int c2 = sh.sc[t * W + si];
              const double f2 = sh.sf[t * W + si];
              if (!(c2 == cc && f2 == f)) term = sh.sterm[t * W + si];
              cc = c2;
              f = f2;
            }
            sh.su[t * W + s] = term;
          }
        }
      }
    }
    // 4. for the previous group (whose rows group g + 2 takes next)
    prior_commit(v, m, sh, c, pl0, pnl, pb0);
    __syncthreads();
    if (wave == 0) {
      // 3. decisions in order
      int Smask = 0;
      if (regs) {
        // each scenario's prior total, the parameters in order (population
        // sizes, then migration rates)
        if (lane < W - 1 && 31 - __clz(lane + 1) < nl) {
          double q = 0.0, mg = 0.0;
          for (int t = 0; t < mm.npsize; ++t) q += sh.su[t * W + lane];
          for (int t = mm.npsize; t < np; ++t) mg += sh.su[t * W + lane];
          sh.pns[lane] = q + mg;
        }
        double xU = 0.0, xd = 0.0;  // lane j: U and the likelihood change of locus j
        int okm = 0;
        for (int j = 0; j < nl; ++j) okm |= (sh.ok[b0 + j] ? 1 : 0) << j;
        if (lane < nl) {
          xU = sh.U[b0 + lane];  // log U
          xd = m.gbeta * (sh.npd[b0 + lane] - sh.pdg[b0 + lane]);
        }
        wave_sync();
        for (int j = 0; j < nl; ++j) {
          if (!((okm >> j) & 1)) continue;
          const double pn = sh.pns[(1 << j) - 1 + Smask];
          const double x = beta * ((pn - probg) + readlane_d(xd, j));
          const double lu = readlane_d(xU, j);
          const bool acc = x != x ? lu != INFINITY : lu < x;
          if (acc) {
            probg = pn;
            Smask |= 1 << j;
          }
        }
        if (lane < nl) sh.dec[b0 + lane] = (Smask >> lane) & 1;
      } else {
        for (int j = 0; j < nl; ++j) {
          const int row = b0 + j;
          if (!sh.ok[row]) {
            if (lane == 0) sh.dec[row] = 0;
            continue;
          }
          const int s = (1 << j) - 1 + Smask;
          for (int t = lane; t < np; t += kWave) {
            double term = sh.qt[t], f = sh.qf[t];
            int cc = sh.qc[t];
            if ((sh.aff[row * Lz.aw + t / 32] >> (t % 32)) & 1u) {
              const int c2 = sh.sc[t * W + s];
              const double f2 = sh.sf[t * W + s];
              if (!(c2 == cc && f2 == f)) term = sh.sterm[t * W + s];
              cc = c2;
              f = f2;
            }
            sh.nt[t] = term;
            sh.nc[t] = cc;
            sh.nf[t] = f;
          }
          wave_sync();
          double q = 0.0, mg = 0.0;
          for (int t = 0; t < mm.npsize; ++t) q += sh.nt[t];
          for (int t = mm.npsize; t < np; ++t) mg += sh.nt[t];
          const double pn = q + mg;
          const double x = beta * ((pn - probg) + m.gbeta * (sh.npd[row] - sh.pdg[row]));
          const bool acc = x != x ? sh.U[row] != INFINITY : sh.U[row] < x;
          if (acc) {
            probg = pn;
            Smask |= 1 << j;
            for (int t = lane; t < np; t += kWave) {
              sh.qt[t] = sh.nt[t];
              sh.qf[t] = sh.nf[t];
              sh.qc[t] = sh.nc[t];
            }
          }
          if (lane == 0) sh.dec[row] = acc ? 1 : 0;
          wave_sync();
        }
      }
    } else if (nnl > 0) {
      // 1. group g + 2
      prior_fetch(m, md, mm.npsize, sh, Lz, c, l0 + 2 * D, nnl, nb0, 1, sweep);
    }
    __syncthreads();
    // the accepted changes into the sums, in locus order per entry; the
    // next group's entry masks
    const int n1 = l0 + D < m.L ? min(D, m.L - l0 - D) : 0, b1 = ((g + 1) % 3) * D;
    int accm = 0;  // the group's accepted loci
    for (int j = 0; j < nl; ++j) accm |= (sh.dec[b0 + j] ? 1 : 0) << j;
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
      double a = sh.all[r];
      for (uint32_t bits = (uint32_t)(sh.fmk[r] & accm); bits; bits &= bits - 1) {
        const int row = b0 + (__ffs(bits) - 1);
        a = wsum_update(r, P, a, sh.dold[row * R + r], sh.dnew[row * R + r]);
      }
      sh.all[r] = a;
      entry_masks(r, b1, n1);
    }
    if (regs && wave == nwaves - 1 && lane < np) {
      // the committed (term, weight, count) of parameter t after the group:
      // the scenario rule over the accepted loci
      const int t = lane, Smask = accm;
      double term = sh.qt[t], f = sh.qf[t];
      int cc = sh.qc[t];
      for (int i = 0; i < nl; ++i) {
        if (!((Smask >> i) & 1)) continue;
        if (!((sh.aff[(b0 + i) * Lz.aw + t / 32] >> (t % 32)) & 1u)) continue;
        const int si = (1 << i) - 1 + (Smask & ((1 << i) - 1));
        const int c2 = sh.sc[t * W + si];
        const double f2 = sh.sf[t * W + si];
        if (!(c2 == cc && f2 == f)) term = sh.sterm[t * W + si];
        cc = c2;
        f = f2;
      }
      sh.qt[t] = term;
      sh.qf[t] = f;
      sh.qc[t] = cc;
    }
    __syncthreads();
    pl0 = l0;
    pnl = nl;
    pb0 = b0;
  }
  prior_commit(v, m, sh, c, pl0, pnl, pb0);
  __syncthreads();
  if (threadIdx.x == 0) m.probg[c] = probg;
  for (int r = threadIdx.x; r < R; r += blockDim.x) m.wall[(int64_t)c * R + r] = sh.all[r];
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    m.qterm[c * np + t] = sh.qt[t];
    m.qf[c * np + t] = sh.qf[t];
    m.qc[c * np + t] = sh.qc[t];
  }
  __syncthreads();
}

// changet_RY1's move of chain c (update_t_RY.cpp:223-330, after Rannala
// and Yang 2003): split k drawn uniformly (randposint), its new time newt
// from getnewt's window (update_gtree_common.cpp:2501-2518): width
// (t_d' - t_u') / (log(L + 1) (P - k)) around the old time, reflected at
// the prior bounds t_u' = max(0, t_u), t_d' = min(tmax, t_d), where t_u and
// t_d are the neighbouring split times (0 and TIMEMAX at the ends).  Every
// node time (on each edge below the root) and migration time in (t_u,
// oldt] is stretched by beforesplit, in (oldt, t_d) by aftersplit
// (update_t_RY.cpp:59-82); the moved events are counted (node times appear
// on both daughter edges, so their counts are halved) and the Hastings
// term (ecd + emd) log t_d_hterm + (ecu + emu) log t_u_hterm goes to
// m.logs[c].  Times are backed up for the restore as by the rescale.
// The split time move shared by changet_RY1 and changet_NW: split k drawn
// uniformly (randposint), its new time from getnewt's window
// (update_gtree_common.cpp:2501-2518): width (t_d' - t_u') / (log(L + 1)
// (P - k)) around the old time, reflected at the prior bounds t_u' =
// max(0, t_u), t_d' = min(tmax, t_d), where t_u and t_d are the
// neighbouring split times (0 and TIMEMAX at the ends).  Thread 0 only;
// ry = {k, oldt, newt, t_u, t_d}.
__device__ void draw_split_move(const MC3Dev &m, int c, uint64_t sweep, double *ry) {
  const int tw = mc3_tw(m.P), ns = m.P - 1;
  PhiloxStream r = item_rng(m, m.C * m.L + c, sweep, 2);
  int k = (int)(r.uniform() * ns);
  if (k > ns - 1) k = ns - 1;
  const double *tv = m.tv + (int64_t)c * tw;
  const double oldt = tv[k];
  const double t_u = k == 0 ? 0.0 : tv[k - 1];
  const double t_d = k == ns - 1 ? kTimeMax : tv[k + 1];
  const double tdp = fmin(m.tmax, t_d), tup = fmax(0.0, t_u);
  const double twin = (tdp - tup) / (log((double)m.L + 1.0) * (double)(m.P - k));
  double newt = (oldt - twin / 2) + r.uniform() * twin;
  if (newt >= tdp)
    newt = 2.0 * tdp - newt;
  else if (newt <= tup)
    newt = 2.0 * tup - newt;
  ry[0] = k;
  ry[1] = oldt;
  ry[2] = newt;
  ry[3] = t_u;
  ry[4] = t_d;
}

__device__ void ry1_stretch(const DeviceView &v, const MC3Dev &m, int c, uint64_t sweep,
                            double *red) {
  __shared__ double ry[5];  // k, oldt, newt, t_u, t_d
  const int tw = mc3_tw(m.P), ns = m.P - 1;
  if (threadIdx.x == 0) draw_split_move(m, c, sweep, ry);
  __syncthreads();
  const int k = (int)ry[0];
  const double oldt = ry[1], newt = ry[2], t_u = ry[3], t_d = ry[4];
  const bool first = k == 0, last = k == ns - 1;
  auto before = [&](double t) {
    return first ? t * newt / oldt : t_u + (t - t_u) * (newt - t_u) / (oldt - t_u);
  };
  auto after = [&](double t) {
    return last ? t + newt - oldt : t_d - (t_d - newt) * (t_d - t) / (t_d - oldt);
  };
  double nu = 0.0, nd = 0.0, mu = 0.0, md = 0.0;  // events moved up / down
  int64_t j0, j1;
  chain_nodes(v, m, c, &j0, &j1);
  double *tm = const_cast<double *>(v.time) + m.node_first;
  const int32_t *dn = v.down + m.node_first;
  for (int64_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    const double t = tm[j];
    m.tbackup[j] = t;
    if (dn[j] == -1) continue;
    if (t <= oldt && t > t_u) {
      tm[j] = before(t);
      nu += 1.0;
    } else if (t > oldt && t < t_d) {
      tm[j] = after(t);
      nd += 1.0;
    }
  }
  double *mt = const_cast<double *>(v.mig_time);
  for (int l = threadIdx.x; l < m.L; l += blockDim.x) {
    const SlotMeta sm = v.slots[m.first_slot + c * m.L + l];
    for (int x = 0; x < sm.nmig; ++x) {
      const double t = mt[sm.mig_off + x];
      m.mtbackup[sm.mig_off + x] = t;
      if (t <= oldt && t > t_u) {
        mt[sm.mig_off + x] = before(t);
        mu += 1.0;
      } else if (t > oldt && t < t_d) {
        mt[sm.mig_off + x] = after(t);
        md += 1.0;
      }
    }
  }
  for (int x = threadIdx.x; x < tw; x += blockDim.x) {
    m.tvbackup[c * tw + x] = m.tv[c * tw + x];
    if (x == k) m.tv[c * tw + x] = newt;
  }
  nu = block_sum(nu, red);
  nd = block_sum(nd, red);
  mu = block_sum(mu, red);
  md = block_sum(md, red);
  if (threadIdx.x == 0) {
    const double ecu = (double)((long long)nu / 2), ecd = (double)((long long)nd / 2);
    const double t_u_hterm = (newt - t_u) / (oldt - t_u);
    const double t_d_hterm = last ? 1.0 : (t_d - newt) / (t_d - oldt);
    m.logs[c] = (ecd + md) * log(t_d_hterm) + (ecu + mu) * log(t_u_hterm);
  }
}

// changet_NW's move of chain c (update_t_NW.cpp:920-1032): the split and
// its new time as for RY1; the split times are backed up and the new one
// set (the genealogies' weights are scored at it), and the chain's split
// and new time are left for the per-genealogy migration update
// (kern_nw.hip), which reads the old split times from the backup.
__device__ void nw_start(const MC3Dev &m, int c, uint64_t sweep) {
  __shared__ double ry[5];
  const int tw = mc3_tw(m.P);
  if (threadIdx.x == 0) {
    draw_split_move(m, c, sweep, ry);
    m.nwk[c] = (int)ry[0];
    m.nwt[c] = ry[2];
  }
  __syncthreads();
  const int k = (int)ry[0];
  for (int x = threadIdx.x; x < tw; x += blockDim.x) {
    m.tvbackup[c * tw + x] = m.tv[c * tw + x];
    if (x == k) m.tv[c * tw + x] = ry[2];
  }
}

// One workgroup per chain.  Accept or reject every locus' genealogy
// proposal (update_gtree.cpp:919-927; rejection restores the moved time).
// Rescaling sweep: back up and rescale every coalescence time of the chain
// (and with the prior its migration and split times).  Otherwise: chain
// sum, swap by the last chain, next sweep's proposals.
// The start of a chain's time update, after its loci are decided: NW or
// RY1 as qupdate draws them, or the global rescale (one workgroup per
// chain; sh holds a double per wavefront).
__device__ void time_update_start(const DeviceView &v, const MC3Dev &m, int c, uint64_t sweep,
                                  double *sh) {
  __syncthreads();  // restored times visible to the whole workgroup
  if (m.nw) {
    // qupdate's time update: RY1 or NW with equal probability when the
    // model has migration (ima_main_mpi.cpp:1867-1911)
    __shared__ int pick;
    if (threadIdx.x == 0) {
      PhiloxStream r = item_rng(m, m.C * m.L + c, sweep, 7);
      pick = r.uniform() < 0.5;
      if (!pick) m.nwk[c] = -1;
    }
    __syncthreads();
    if (pick) {
      nw_start(m, c, sweep);
      return;
    }
  }
  if (m.ry1 && m.prior >= 0 && m.P > 1) {
    ry1_stretch(v, m, c, sweep, sh);
    return;
  }
  const double ls = chain_log_scale(m, c, sweep);
  const double s = exp(ls);
  if (threadIdx.x == 0) m.logs[c] = ls;
  int64_t j0, j1;
  chain_nodes(v, m, c, &j0, &j1);
  double *tm = const_cast<double *>(v.time) + m.node_first;
  const int32_t *dn = v.down + m.node_first;
  for (int64_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    const double t = tm[j];
    m.tbackup[j] = t;
    if (dn[j] != -1) tm[j] = t * s;
  }
  if (m.prior >= 0) {
    double *mt = const_cast<double *>(v.mig_time);
    for (int l = threadIdx.x; l < m.L; l += blockDim.x) {
      const SlotMeta sm = v.slots[m.first_slot + c * m.L + l];
      for (int j = 0; j < sm.nmig; ++j) {
        const double t = mt[sm.mig_off + j];
        m.mtbackup[sm.mig_off + j] = t;
        mt[sm.mig_off + j] = t * s;
      }
    }
    const int tw = mc3_tw(m.P);
    for (int k = threadIdx.x; k < tw; k += blockDim.x) {
      const double t = m.tv[c * tw + k];
      m.tvbackup[c * tw + k] = t;
      if (k < m.P - 1) m.tv[c * tw + k] = t * s;
    }
  }
}

__global__ __launch_bounds__(1024) void mc3_accept(DeviceView v, MC3Dev m, int with_scale,
                                                   int swaptries) {
  __shared__ double sh[1024 / kWave];
  __shared__ double Ssh[kSwapLds], Bsh[kSwapLds];
  extern __shared__ double dyn[];
  const int c = blockIdx.x;
  const uint64_t sweep = *m.counter;
  double part = 0.0;
  if (m.prior < 0) {
    for (int l = threadIdx.x; l < m.L; l += blockDim.x) {
      const int it = c * m.L + l;
      PhiloxStream r = item_rng(m, it, sweep, 1);
      const double U = r.uniform();
      const double np = new_loglik(m, it);
      bool acc = false;
      const double h = m.topo ? m.thast[it] : 0.0;
      if (m.status[it] == IMA_STATUS_OK)
        acc = U < fmin(1.0, exp(m.beta[c] * (np - m.pdg[it]) + h));
      double p = m.pdg[it];
      if (acc) {
        p = np;
        m.pdg[it] = p;
      } else if (!m.topo) {
        restore_item(v, m, it);
      }
      if (m.topo) m.idec[it] = acc ? 1 : 0;
      m.acc[it] += acc ? 1u : 0u;
      part += p;
      if (!with_scale && !m.topo) propose_item(v, m, it, sweep + 1);
    }
  } else {
    prior_accept(v, m, c, sweep, dyn);
    for (int l = threadIdx.x; l < m.L; l += blockDim.x) {
      part += m.pdg[c * m.L + l];
      if (!with_scale && !m.topo) propose_item(v, m, c * m.L + l, sweep + 1);
    }
    if (threadIdx.x == 0) part += m.probg[c];
  }
  if (!with_scale) {
    finish_chain(m, c, block_sum(part, sh), swaptries, sweep, Ssh, Bsh);
    return;
  }
  // with the device's genealogy move the accepted proposals are copied
  // into the chains' slots by mc3_copy_accepted (every item at once), and
  // mc3_time_start begins the time update after it
  if (!m.topo) time_update_start(v, m, c, sweep, sh);
}

// Time-update sweeps with the device's genealogy move: the update's start,
// once the accepted proposals are in the chains' slots.
__global__ __launch_bounds__(1024) void mc3_time_start(DeviceView v, MC3Dev m) {
  __shared__ double sh[1024 / kWave];
  time_update_start(v, m, blockIdx.x, *m.counter, sh);
}

// One workgroup per chain: log acceptance of the rescaling
// = sum_l beta (new - old) + (scaled times) log s [+ beta dprobg]; commit
// or restore; chain sum; swap by the last chain; next sweep's proposals.
__global__ __launch_bounds__(1024) void mc3_scale_finish(DeviceView v, MC3Dev m, int swaptries) {
  __shared__ double sh[1024 / kWave];
  __shared__ double Ssh[kSwapLds], Bsh[kSwapLds];
  __shared__ int okall, decision;
  extern __shared__ double dyn[];
  const int c = blockIdx.x;
  const uint64_t sweep = *m.counter;
  const bool prior = m.prior >= 0;
  // rubber band: m.logs[c] holds the whole Hastings term
  const bool ry = m.ry1 && prior && m.P > 1;
  // changet_NW this sweep: node times (so the data likelihood) did not
  // move; the Hastings term is the genealogies' migweight sum, and
  // exp(beta dprobg + migweight) decides (update_t_NW.cpp:983-1000)
  const bool nwc = m.nw && m.nwk[c] >= 0;
  if (threadIdx.x == 0) okall = 1;
  __syncthreads();
  double a = 0.0;
  for (int l = threadIdx.x; l < m.L; l += blockDim.x) {
    const int it = c * m.L + l;
    if (prior && m.wstatus[it] != IMA_STATUS_OK) okall = 0;
    if (nwc) {
      if (m.nwa.status[it] == 2) okall = 0;
      a += m.nwa.migweight[it];
      continue;
    }
    if (m.status[it] != IMA_STATUS_OK) okall = 0;
    const SlotMeta sm = v.slots[m.first_slot + it];
    const int scaled = ry ? 0 : (sm.n - 1) + (prior ? sm.nmig : 0);
    a += m.beta[c] * (new_loglik(m, it) - m.pdg[it]) + scaled * m.logs[c];
  }
  a = block_sum(a, sh);
  const int R = m.R, np = m.nparam, tw = mc3_tw(m.P);
  const PriorSh psh = prior_carve(dyn, PriorLds(R, np, m.pchunk, m.mints));
  double *all = psh.all, *nt = psh.nt, *nf = psh.nf;
  int *nc = psh.nc;
  double pn = 0.0;
  if (prior) {
    pn = chain_prior_fresh(v, m, c, m.wnew, psh);
    a += m.beta[c] * (pn - m.probg[c]) + (nwc ? 0.0 : (ry ? 1 : m.P - 1) * m.logs[c]);
    if (threadIdx.x == 0)
      for (int k = 0; k < m.P - 1; ++k)
        if (m.tv[c * tw + k] > m.tmax) okall = 0;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    PhiloxStream r = item_rng(m, m.C * m.L + m.C + c, sweep, 3);
    const double U = r.uniform();
    decision = okall && (U < fmin(1.0, exp(a)));
    m.decision[c] = decision;
    m.cacc[c] += (uint32_t)decision;
  }
  __syncthreads();
  double part = 0.0;
  if (decision) {
    for (int l = threadIdx.x; l < m.L; l += blockDim.x) {
      const int it = c * m.L + l;
      const double np_ = nwc ? m.pdg[it] : new_loglik(m, it);
      m.pdg[it] = np_;
      part += np_;
    }
    if (prior) {
      const int64_t base = (int64_t)c * m.L * R, cnt = (int64_t)m.L * R;
      for (int64_t x = threadIdx.x; x < cnt; x += blockDim.x) m.wrec[base + x] = m.wnew[base + x];
      for (int r = threadIdx.x; r < R; r += blockDim.x) m.wall[(int64_t)c * R + r] = all[r];
      for (int t = threadIdx.x; t < np; t += blockDim.x) {
        m.qterm[c * np + t] = nt[t];
        m.qf[c * np + t] = nf[t];
        m.qc[c * np + t] = nc[t];
      }
      if (threadIdx.x == 0) {
        m.probg[c] = pn;
        part += pn;
      }
    }
  } else if (nwc) {
    // the genealogies the update changed get their populations and
    // migrations back
    for (int l = threadIdx.x; l < m.L; l += blockDim.x) part += m.pdg[c * m.L + l];
    const NwArgs &na = m.nwa;
    int32_t *pop = const_cast<int32_t *>(v.pop), *mf = const_cast<int32_t *>(v.mig_first);
    int32_t *mc = const_cast<int32_t *>(v.mig_count), *me = const_cast<int32_t *>(v.mig_edge);
    int32_t *mto = const_cast<int32_t *>(v.mig_to);
    double *mt = const_cast<double *>(v.mig_time);
    // one wavefront per locus, its lanes over edges and events
    const int wv = threadIdx.x / kWave, lane = threadIdx.x % kWave, nwv = blockDim.x / kWave;
    for (int l = wv; l < m.L; l += nwv) {
      const int it = c * m.L + l;
      if (na.status[it] != 1) continue;
      const int s = m.first_slot + it;
      const SlotMeta sm = v.slots[s];
      const int64_t no = sm.node_off, mo = sm.mig_off;
      const int ne = 2 * sm.n - 1, nm = na.nmig_bk[it];
      for (int e = lane; e < ne; e += kWave) {
        pop[no + e] = na.pop_bk[no + e];
        mf[no + e] = na.mf_bk[no + e];
        mc[no + e] = na.mc_bk[no + e];
      }
      for (int x = lane; x < nm; x += kWave) {
        me[mo + x] = na.me_bk[mo + x];
        mto[mo + x] = na.mto_bk[mo + x];
        mt[mo + x] = na.mt_bk[mo + x];
      }
      // the wavefront's lanes read the slot record above
      if (lane == 0) const_cast<SlotMeta *>(v.slots)[s].nmig = nm;
    }
    for (int k = threadIdx.x; k < tw; k += blockDim.x) m.tv[c * tw + k] = m.tvbackup[c * tw + k];
    if (threadIdx.x == 0) part += m.probg[c];
  } else {
    for (int l = threadIdx.x; l < m.L; l += blockDim.x) part += m.pdg[c * m.L + l];
    int64_t j0, j1;
    chain_nodes(v, m, c, &j0, &j1);
    double *tm = const_cast<double *>(v.time) + m.node_first;
    for (int64_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) tm[j] = m.tbackup[j];
    if (prior) {
      double *mt = const_cast<double *>(v.mig_time);
      for (int l = threadIdx.x; l < m.L; l += blockDim.x) {
        const SlotMeta sm = v.slots[m.first_slot + c * m.L + l];
        for (int j = 0; j < sm.nmig; ++j) mt[sm.mig_off + j] = m.mtbackup[sm.mig_off + j];
      }
      for (int k = threadIdx.x; k < tw; k += blockDim.x) m.tv[c * tw + k] = m.tvbackup[c * tw + k];
      if (threadIdx.x == 0) part += m.probg[c];
    }
  }
  __syncthreads();  // committed times visible before the next proposals
  if (!m.topo)
    for (int l = threadIdx.x; l < m.L; l += blockDim.x) propose_item(v, m, c * m.L + l, sweep + 1);
  finish_chain(m, c, block_sum(part, sh), swaptries, sweep, Ssh, Bsh);
}

// ---- the genealogy table's rows (savegsampinf, ginfo.cpp:319-377) ----------

// One row for the cold chain (the last chain whose beta is exactly 1.0,
// whichiscoldchain ima_main_mpi.cpp:3023-3033) in the layout L mode reads
// (ginfo.cpp:288-316): per population-size parameter the count (an int sum
// of cc), then fc and hcc accumulated in float over the parameter's wp
// positions in order; per migration parameter the count of mc and fm in
// float; qintegrate and mintegrate; pdg (the chain's loci summed in order);
// probg; the split times.  Everything is stored as float.
__global__ void mc3_record_row(DeviceView v, MC3Dev m, float *row) {
  __shared__ int z;
  if (threadIdx.x == 0) {
    int w = -1;
    for (int c = 0; c < m.C; ++c)
      if (m.beta[c] == 1.0) w = c;
    z = w;
  }
  __syncthreads();
  if (z < 0) return;  // the host checked that the ladder holds beta = 1
  const ModelMeta mm = v.models[m.prior];
  const ModelView md = model_view(v, mm);
  const int P = m.P, P2 = P * P, P3 = P2 * P, R = m.R;
  const int nps = mm.npsize, nmg = mm.nmig, np = m.nparam;
  const double *all = m.wall + (int64_t)z * R;
  const int ccp = 0, fcp = nps, hccp = 2 * nps, mcp = 3 * nps, fmp = 3 * nps + nmg;
  const int qip = 3 * nps + 2 * nmg, mip = 4 * nps + 2 * nmg, pdgp = 4 * nps + 3 * nmg;
  for (int i = threadIdx.x; i < nps; i += blockDim.x) {
    int c = 0;
    float f = 0.0f, hc = 0.0f;
    for (int w = md.theta_off[i]; w < md.theta_off[i + 1]; ++w) {
      const int e = md.theta_wp[2 * w] * P + md.theta_wp[2 * w + 1];
      c += (int)all[e];
      f += (float)all[P2 + e];
      hc += (float)all[2 * P2 + e];
    }
    row[ccp + i] = (float)c;
    row[fcp + i] = f;
    row[hccp + i] = hc;
    row[qip + i] = (float)m.qterm[(int64_t)z * np + i];
  }
  for (int i = threadIdx.x; i < nmg; i += blockDim.x) {
    int c = 0;
    float f = 0.0f;
    for (int w = md.mig_off[i]; w < md.mig_off[i + 1]; ++w) {
      const int e = (md.mig_wp[3 * w] * P + md.mig_wp[3 * w + 1]) * P + md.mig_wp[3 * w + 2];
      c += (int)all[3 * P2 + e];
      f += (float)all[3 * P2 + P3 + e];
    }
    row[mcp + i] = (float)c;
    row[fmp + i] = f;
    row[mip + i] = (float)m.qterm[(int64_t)z * np + nps + i];
  }
  if (threadIdx.x == 0) {
    double pd = 0.0;
    for (int l = 0; l < m.L; ++l) pd += m.pdg[(int64_t)z * m.L + l];
    row[pdgp] = (float)pd;
    row[pdgp + 1] = (float)m.probg[z];
    for (int k = 0; k < P - 1; ++k) row[pdgp + 2 + k] = (float)m.tv[(int64_t)z * mc3_tw(P) + k];
  }
}

hipError_t enqueue_mc3_record(const DeviceView &v, const MC3Dev &m, float *row, hipStream_t st) {
  hipLaunchKernelGGL(mc3_record_row, dim3(1), dim3(kWave), 0, st, v, m, row);
  return hipGetLastError();
}

// ---- decisions for proposals written by the host (ima_mc3_decide) -------

// One workgroup per chain: every locus' proposal (in slot prop_first + it,
// scored into newpdg / wnew) accepted or rejected in order as
// updategenealogy decides (update_gtree.cpp:900-961):
// U < min(1, exp(beta (dprobg + gbeta dpdg) + h)), h the host's
// migweight + slideweight + Atermsum; an incompatible infinite-sites
// genealogy or a zero stepwise term is rejected.  Decisions go to idec[];
// the chain sum to S[c].
__global__ __launch_bounds__(1024) void mc3_decide(DeviceView v, MC3Dev m) {
  __shared__ double sh[1024 / kWave];
  extern __shared__ double dyn[];
  const int c = blockIdx.x;
  const uint64_t sweep = *m.counter;
  double part = 0.0;
  if (m.prior < 0) {
    for (int l = threadIdx.x; l < m.L; l += blockDim.x) {
      const int it = c * m.L + l;
      PhiloxStream r = item_rng(m, it, sweep, 1);
      const double U = m.uhost ? m.uhost[it] : r.uniform();
      const double np = new_loglik(m, it);
      const double h = m.hast ? m.hast[it] : 0.0;
      bool acc = false;
      if (m.status[it] == IMA_STATUS_OK)
        acc = U < fmin(1.0, exp(m.beta[c] * (m.gbeta * (np - m.pdg[it])) + h));
      if (acc) m.pdg[it] = np;
      m.idec[it] = acc ? 1 : 0;
      m.acc[it] += acc ? 1u : 0u;
      part += m.pdg[it];
    }
  } else {
    prior_accept(v, m, c, sweep, dyn);
    for (int l = threadIdx.x; l < m.L; l += blockDim.x) part += m.pdg[c * m.L + l];
    if (threadIdx.x == 0) part += m.probg[c];
  }
  part = block_sum(part, sh);
  if (threadIdx.x == 0) m.S[c] = part;
}

// One workgroup per item: an accepted proposal replaces the chain's
// genealogy (edges, times, populations, alleles, program, migrations and
// the slot's root, stack depth and migration count).  The host checked
// that the chain's slot holds the proposal's migrations.
__global__ __launch_bounds__(256) void mc3_copy_accepted(DeviceView v, MC3Dev m) {
  const int it = blockIdx.x;
  if (m.idec[it]) copy_proposal(v, m, it, threadIdx.x, blockDim.x);
}

__global__ void mc3_host_finish(MC3Dev m) {
  *m.counter += 1;
  m.stats[0] += (unsigned long long)(m.C * m.L);
}

__global__ void mc3_prop_items(MC3Dev m, const ima_item *items, ima_item *pitems, int32_t *ptw) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= m.C * m.L) return;
  ima_item x = items[it];
  x.slot = m.prop_first + it;
  pitems[it] = x;
  ptw[it] = m.prop_first + it;
}

hipError_t enqueue_mc3_prop_items(const MC3Dev &m, const ima_item *items, ima_item *pitems,
                                  int32_t *ptw, hipStream_t st) {
  const int N = m.C * m.L;
  hipLaunchKernelGGL(mc3_prop_items, dim3((N + 255) / 256), dim3(256), 0, st, m, items, pitems,
                     ptw);
  return hipGetLastError();
}

hipError_t enqueue_mc3_decide(const DeviceView &v, const MC3Dev &m, hipStream_t st) {
  const size_t dl = m.prior >= 0 ? PriorLds(m.R, m.nparam, m.pchunk, m.mints).bytes() : 0;
  hipLaunchKernelGGL(mc3_decide, dim3(m.C), dim3(1024), dl, st, v, m);
  hipLaunchKernelGGL(mc3_copy_accepted, dim3(m.C * m.L), dim3(256), 0, st, v, m);
  hipLaunchKernelGGL(mc3_host_finish, dim3(1), dim3(1), 0, st, m);
  return hipGetLastError();
}

// One scoring pass (see ima_internal.h).  The proposals' genealogy weights
// run in the HKY launch (IS and SW launches beside it on the side stream);
// without HKY items (or for genealogies too large for that launch's LDS)
// they run on the side stream beside the likelihood launches.  The caller's
// next launch on st waits for both streams.
hipError_t enqueue_mc3_score(const DeviceView &v, const MC3Dev &m, const ima_item *items,
                             const int32_t *tw_slots, const int32_t *idx, const HkyPlan &p,
                             int nh, int ni, int ns, int precision, int hmax_n, int hmax_slots,
                             int imax_n, int imax_w, hipEvent_t *score_ev, int pass,
                             hipStream_t st, hipStream_t side, hipEvent_t *fj) {
  const bool prior = m.prior >= 0;
  const TwLaunch tw{tw_slots, m.C * m.L, m.emax, m.tv, m.tv_row, m.wnew, m.wstatus};
  hipError_t e = hipSuccess;
  if (prior && nh &&
      hky_tw_lds_bytes(precision, hmax_n, hmax_slots, p.ppl, m.P, m.emax) <= kTwLdsBudget) {
    // IS and SW items, if any, are scored on the side stream beside it
    const bool fork = ni || ns;
    if (score_ev) e = hipEventRecord(score_ev[2 * pass], st);
    if (e == hipSuccess && fork) e = hipEventRecord(fj[2 * pass], st);
    if (e == hipSuccess && fork) e = hipStreamWaitEvent(side, fj[2 * pass], 0);
    if (e == hipSuccess && ni)
      e = launch_is(v, items, idx + nh, ni, imax_n, imax_w, m.newpdg, m.status, nullptr,
                    nullptr, side);
    if (e == hipSuccess && ns)
      e = launch_sw(v, items, idx + nh + ni, ns, m.newpdg, m.status, nullptr, side);
    if (e == hipSuccess && fork) e = hipEventRecord(fj[2 * pass + 1], side);
    if (e == hipSuccess)
      e = launch_hky_tw(v, items, p, precision, hmax_n, hmax_slots, m.newpdg, m.status, tw,
                        m.P, st);
    if (e == hipSuccess && fork) e = hipStreamWaitEvent(st, fj[2 * pass + 1], 0);
    if (e == hipSuccess && score_ev) e = hipEventRecord(score_ev[2 * pass + 1], st);
    return e;
  }
  if (prior) {
    e = hipEventRecord(fj[2 * pass], st);
    if (e == hipSuccess) e = hipStreamWaitEvent(side, fj[2 * pass], 0);
    if (e == hipSuccess)
      e = launch_treeweight(v, tw_slots, m.C * m.L, m.tv, m.tv_row, m.P, m.emax, m.wnew,
                            m.wstatus, side);
    if (e == hipSuccess) e = hipEventRecord(fj[2 * pass + 1], side);
  }
  if (e == hipSuccess && score_ev) e = hipEventRecord(score_ev[2 * pass], st);
  if (e == hipSuccess && nh)
    e = launch_hky(v, items, p, precision, hmax_n, hmax_slots, m.newpdg, m.status, st);
  if (e == hipSuccess && ni)
    e = launch_is(v, items, idx + nh, ni, imax_n, imax_w, m.newpdg, m.status, nullptr, nullptr, st);
  if (e == hipSuccess && ns)
    e = launch_sw(v, items, idx + nh + ni, ns, m.newpdg, m.status, nullptr, st);
  if (e == hipSuccess && score_ev) e = hipEventRecord(score_ev[2 * pass + 1], st);
  if (e == hipSuccess && prior) e = hipStreamWaitEvent(st, fj[2 * pass + 1], 0);
  return e;
}

// Enqueue the kernels of one sweep (no host synchronisation).  The sweep's
// proposals are already in place (mc3_propose or the previous sweep).
hipError_t enqueue_mc3_sweep(const DeviceView &v, const MC3Dev &m, bool with_scale,
                             const ima_item *items, const int32_t *idx, const HkyPlan &plan,
                             int nh, int ni, int ns, int precision, int hmax_n, int hmax_slots,
                             int imax_n, int imax_w, int swaptries, hipEvent_t *score_ev,
                             hipStream_t st, hipStream_t side, hipEvent_t *fj) {
  // the per-chain kernels merge pattern blocks themselves
  HkyPlan p = plan;
  p.nred = 0;
  const bool prior = m.prior >= 0;
  const size_t dl = prior ? PriorLds(m.R, m.nparam, m.pchunk, m.mints).bytes() : 0;
  // score_ev (optional): events recorded around the two scoring passes.
  // The proposals' genealogy weights run in the HKY launch (IS and SW
  // launches beside it on the side stream); without HKY items (or for
  // genealogies too large for that launch's LDS) they run on the side
  // stream beside the likelihood launches.  The accept kernel waits for
  // both streams
  auto score = [&](int pass) -> hipError_t {
    // topology moves score the proposal slots; the second (time-update)
    // pass scores the chains' own
    const bool prop = m.topo && pass == 0;
    return enqueue_mc3_score(v, m, prop ? m.pitems : items, prop ? m.ptw : m.tw_slots, idx, p,
                             nh, ni, ns, precision, hmax_n, hmax_slots, imax_n, imax_w, score_ev,
                             pass, st, side, fj);
  };
  hipError_t e = hipSuccess;
  if (m.topo) {
    e = enqueue_mc3_prop_items(m, items, m.pitems, m.ptw, st);
    if (e == hipSuccess) e = enqueue_mc3_propose_topo(v, m, m.maxn, st);
    if (e != hipSuccess) return e;
  }
  e = score(0);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mc3_accept, dim3(m.C), dim3(1024), dl, st, v, m, with_scale ? 1 : 0,
                     swaptries);
  if (m.topo) {
    hipLaunchKernelGGL(mc3_copy_accepted, dim3(m.C * m.L), dim3(kWave), 0, st, v, m);
    if (with_scale) hipLaunchKernelGGL(mc3_time_start, dim3(m.C), dim3(1024), 0, st, v, m);
  }
  if (with_scale && m.nw) {
    e = launch_nw_update(v, m.nwa, st);
    if (e != hipSuccess) return e;
  }
  if (with_scale) {
    e = score(1);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mc3_scale_finish, dim3(m.C), dim3(1024), dl, st, v, m, swaptries);
  }
  return hipGetLastError();
}

hipError_t enqueue_mc3_prior_init(const DeviceView &v, const MC3Dev &m, hipStream_t st) {
  hipError_t e = launch_treeweight(v, m.tw_slots, m.C * m.L, m.tv, m.tv_row, m.P, m.emax,
                                   m.wrec, m.wstatus, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mc3_prior_init, dim3(m.C), dim3(256),
                     PriorLds(m.R, m.nparam, m.pchunk, m.mints).bytes(), st, v, m);
  return hipGetLastError();
}

size_t mc3_prior_lds_bytes(int R, int np, int pchunk, int mints) {
  return PriorLds(R, np, pchunk, mints).bytes();
}

hipError_t enqueue_mc3_propose(const DeviceView &v, const MC3Dev &m, hipStream_t st) {
  const int N = m.C * m.L;
  hipLaunchKernelGGL(mc3_propose, dim3((N + 255) / 256), dim3(256), 0, st, v, m);
  return hipGetLastError();
}

hipError_t enqueue_mc3_revert(const DeviceView &v, const MC3Dev &m, hipStream_t st) {
  const int N = m.C * m.L;
  hipLaunchKernelGGL(mc3_revert, dim3((N + 255) / 256), dim3(256), 0, st, v, m);
  return hipGetLastError();
}

}  // namespace ima
