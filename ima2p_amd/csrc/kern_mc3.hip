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
    const int32_t *me = v.mig_edge + sm.mig_off;
    const double *mt = v.mig_time + sm.mig_off;
    for (int j = 0; j < sm.nmig; ++j) {
      const int x = me[j];
      if (x == a || x == b) lo = fmax(lo, mt[j]);
      else if (x == e) hi = fmin(hi, mt[j]);
    }
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
  const int2 ib = m.iblk[it];
  if (ib.y <= 1) return m.newpdg[it];
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
    PhiloxStream rng(m.seed ^ 0x5A5A5A5A5A5A5A5Aull, stream_counter(sweep, 4));
    const int SWAPDIST = 7, nchains = m.C;
    const double *S = staged ? Ssh : m.S;
    double *beta = staged ? Bsh : m.beta;
    unsigned long long acc = 0;
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
    m.stats[4] += (unsigned long long)(nchains > 1 ? swaptries : 0);
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
struct PriorSh {
  double *all, *hold, *qt, *qf, *nt, *nf, *npd, *pdg, *U, *dold, *dnew, *mdl, *lga;
  int *qc, *nc, *ok, *nch, *dec, *didx, *mi, *lgc;
  uint32_t *aff;
};
__device__ inline PriorSh prior_carve(double *dyn, const PriorLds &Lz) {
  PriorSh p;
  p.all = dyn;
  p.hold = p.all + Lz.R;
  p.qt = p.hold + Lz.R;
  p.qf = p.qt + Lz.np;
  p.nt = p.qf + Lz.np;
  p.nf = p.nt + Lz.np;
  p.npd = p.nf + Lz.np;
  p.pdg = p.npd + Lz.LC;
  p.U = p.pdg + Lz.LC;
  p.dold = p.U + Lz.LC;
  p.dnew = p.dold + (size_t)Lz.LC * Lz.cap;
  p.mdl = p.dnew + (size_t)Lz.LC * Lz.cap;
  p.lga = p.mdl + Lz.np;
  p.qc = reinterpret_cast<int *>(p.lga + Lz.np);
  p.nc = p.qc + Lz.np;
  p.ok = p.nc + Lz.np;
  p.nch = p.ok + Lz.LC;
  p.dec = p.nch + Lz.LC;
  p.didx = p.dec + Lz.LC;
  p.aff = reinterpret_cast<uint32_t *>(p.didx + (size_t)Lz.LC * Lz.cap);
  p.mi = reinterpret_cast<int *>(p.aff + (size_t)Lz.LC * Lz.aw);
  p.lgc = p.mi + Lz.mints;
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
  for (int r = threadIdx.x; r < R; r += blockDim.x)
    sh.all[r] = ordered_rec_sum(w + (int64_t)c * m.L * R, R, m.L, r);
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

// True when parameter t adds up one of the record entries in idx[0..n).
__device__ inline bool param_touches(const ModelView &md, int P, int npsize, int t,
                                     const int *idx, int n) {
  const int P2 = P * P, P3 = P2 * P;
  if (t < npsize) {
    for (int w = md.theta_off[t]; w < md.theta_off[t + 1]; ++w) {
      const int e = md.theta_wp[2 * w] * P + md.theta_wp[2 * w + 1];
      for (int i = 0; i < n; ++i) {
        const int r = idx[i];
        if (r == e || r == P2 + e || r == 2 * P2 + e) return true;
      }
    }
  } else {
    const int mm = t - npsize;
    for (int w = md.mig_off[mm]; w < md.mig_off[mm + 1]; ++w) {
      const int e = (md.mig_wp[3 * w] * P + md.mig_wp[3 * w + 1]) * P + md.mig_wp[3 * w + 2];
      for (int i = 0; i < n; ++i) {
        const int r = idx[i];
        if (r == 3 * P2 + e || r == 3 * P2 + P3 + e) return true;
      }
    }
  }
  return false;
}

// Loci of chain c in order, each accepted with beta (dprobg + dpdg).  The
// work that does not depend on earlier decisions is done in parallel, a
// chunk of loci at a time: each wavefront reads one locus' likelihood,
// acceptance uniform and the weight entries its proposal changed (up to
// kPriorCap, else the locus takes the full record from memory) and marks
// the parameters those entries feed.  The first wavefront then decides the
// chunk's loci in order from LDS only, recomputing just the marked
// parameters' integrals; commits and restores are applied in parallel.
__device__ void prior_accept(const DeviceView &v, const MC3Dev &m, int c, uint64_t sweep,
                             double *dyn) {
  const int R = m.R, np = m.nparam, P = m.P;
  const PriorLds Lz(R, np, m.pchunk, m.mints);
  const PriorSh sh = prior_carve(dyn, Lz);
  // the model's maps and each parameter's log Gamma live in LDS, so the
  // in-order loop below reads no global memory
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
    const int cq = m.qc[c * np + t];
    sh.qc[t] = cq;
    sh.lgc[t] = cq;
    sh.lga[t] = prior_param_lga(mm.npsize, t, cq, v.logfact);
  }
  __syncthreads();
  const ModelView md = model_view(sh.mi, sh.mdl, mm);
  double probg = m.probg[c];
  const double beta = m.beta[c];
  for (int l0 = 0; l0 < m.L; l0 += Lz.LC) {
    const int nl = min(Lz.LC, m.L - l0);
    // ---- parallel: one wavefront per locus of the chunk
    for (int x = wave; x < nl; x += nwaves) {
      const int it = c * m.L + l0 + x;
      const double *wo = m.wrec + (int64_t)it * R, *wn = m.wnew + (int64_t)it * R;
      int count = 0;
      for (int r0 = 0; r0 < R; r0 += kWave) {
        const int r = r0 + lane;
        double o = 0.0, w = 0.0;
        if (r < R) {
          o = wo[r];
          w = wn[r];
        }
        const bool diff = r < R && o != w;
        const unsigned long long bal = __ballot(diff);
        const int pos = count + __popcll(bal & ((1ull << lane) - 1ull));
        if (diff && pos < Lz.cap) {
          sh.didx[x * Lz.cap + pos] = r;
          sh.dold[x * Lz.cap + pos] = o;
          sh.dnew[x * Lz.cap + pos] = w;
        }
        count += __popcll(bal);
      }
      if (lane == 0) {
        sh.ok[x] = m.status[it] == IMA_STATUS_OK && m.wstatus[it] == IMA_STATUS_OK;
        sh.npd[x] = new_loglik(m, it);
        sh.pdg[x] = m.pdg[it];
        PhiloxStream rr = item_rng(m, it, sweep, 1);
        sh.U[x] = rr.uniform();
        sh.nch[x] = count;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const bool dense = count > Lz.cap;
      for (int t0 = 0; t0 < np; t0 += kWave) {
        const int t = t0 + lane;
        const bool hit = t < np && (dense || param_touches(md, P, mm.npsize, t,
                                                          sh.didx + x * Lz.cap, count));
        const unsigned long long bal = __ballot(hit);
        if (lane == 0) {
          sh.aff[x * Lz.aw + t0 / 32] = (uint32_t)bal;
          if (t0 / 32 + 1 < Lz.aw) sh.aff[x * Lz.aw + t0 / 32 + 1] = (uint32_t)(bal >> 32);
        }
      }
    }
    __syncthreads();
    // ---- in order: wave 0 updates the sums and decides; the marked
    // parameters' integrals are spread one per wavefront, so their
    // different series run side by side rather than diverging in one wave
    double hv = 0.0;
    for (int x = 0; x < nl; ++x) {
      const int it = c * m.L + l0 + x;
      const bool ok = sh.ok[x] != 0;
      const int cnt = sh.nch[x];
      const bool dense = cnt > Lz.cap;
      if (wave == 0 && ok) {
        if (!dense) {
          if (lane < cnt) {
            const int r = sh.didx[x * Lz.cap + lane];
            hv = sh.all[r];
            sh.all[r] = wsum_update(r, P, hv, sh.dold[x * Lz.cap + lane],
                                    sh.dnew[x * Lz.cap + lane]);
          }
        } else {
          const double *wo = m.wrec + (int64_t)it * R, *wn = m.wnew + (int64_t)it * R;
          for (int r = lane; r < R; r += kWave) {
            const double a = sh.all[r];
            sh.hold[r] = a;
            sh.all[r] = wsum_update(r, P, a, wo[r], wn[r]);
          }
        }
      }
      __syncthreads();
      if (ok) {
        for (int t = wave; t < np; t += nwaves) {
          if ((sh.aff[x * Lz.aw + t / 32] >> (t % 32)) & 1u) {
            int cc;
            double f, h;
            prior_param_sums(md, P, mm.npsize, t, sh.all, &cc, &f, &h);
            double term = sh.qt[t];
            if (!(cc == sh.qc[t] && f == sh.qf[t])) {
              const double lga = cc == sh.lgc[t] ? sh.lga[t]
                                                  : prior_param_lga(mm.npsize, t, cc, v.logfact);
              term = prior_param_term_lga(md, mm.npsize, t, cc, f, h, lga);
            }
            if (lane == 0) {
              sh.nc[t] = cc;
              sh.nf[t] = f;
              sh.nt[t] = term;
            }
          } else if (lane == 0) {
            sh.nc[t] = sh.qc[t];
            sh.nf[t] = sh.qf[t];
            sh.nt[t] = sh.qt[t];
          }
        }
      }
      __syncthreads();
      if (wave == 0) {
        if (!ok) {
          if (lane == 0) sh.dec[x] = 0;
        } else {
          double q = 0.0, mg = 0.0;
          for (int t = 0; t < mm.npsize; ++t) q += sh.nt[t];
          for (int t = mm.npsize; t < np; ++t) mg += sh.nt[t];
          const double pn = q + mg;
          const bool acc =
              sh.U[x] < fmin(1.0, exp(beta * ((pn - probg) + (sh.npd[x] - sh.pdg[x]))));
          if (acc) {
            probg = pn;
            for (int t = lane; t < np; t += kWave) {
              sh.qt[t] = sh.nt[t];
              sh.qf[t] = sh.nf[t];
              sh.qc[t] = sh.nc[t];
            }
          } else if (!dense) {
            if (lane < cnt) sh.all[sh.didx[x * Lz.cap + lane]] = hv;
          } else {
            for (int r = lane; r < R; r += kWave) sh.all[r] = sh.hold[r];
          }
          if (lane == 0) sh.dec[x] = acc ? 1 : 0;
        }
      }
    }
    __syncthreads();
    // ---- parallel: commit or restore the chunk's loci
    for (int x = wave; x < nl; x += nwaves) {
      const int it = c * m.L + l0 + x;
      const bool acc = sh.dec[x] != 0;
      if (acc) {
        double *wc = m.wrec + (int64_t)it * R;
        const double *wn = m.wnew + (int64_t)it * R;
        if (sh.nch[x] > Lz.cap) {
          for (int r = lane; r < R; r += kWave) wc[r] = wn[r];
        } else if (lane < sh.nch[x]) {
          wc[sh.didx[x * Lz.cap + lane]] = sh.dnew[x * Lz.cap + lane];
        }
      }
      if (lane == 0) {
        if (acc)
          m.pdg[it] = sh.npd[x];
        else
          restore_item(v, m, it);
        m.acc[it] += acc ? 1u : 0u;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) m.probg[c] = probg;
  for (int r = threadIdx.x; r < R; r += blockDim.x) m.wall[(int64_t)c * R + r] = sh.all[r];
  for (int t = threadIdx.x; t < np; t += blockDim.x) {
    m.qterm[c * np + t] = sh.qt[t];
    m.qf[c * np + t] = sh.qf[t];
    m.qc[c * np + t] = sh.qc[t];
  }
  __syncthreads();
}

// One workgroup per chain.  Accept or reject every locus' genealogy
// proposal (update_gtree.cpp:919-927; rejection restores the moved time).
// Rescaling sweep: back up and rescale every coalescence time of the chain
// (and with the prior its migration and split times).  Otherwise: chain
// sum, swap by the last chain, next sweep's proposals.
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
      if (m.status[it] == IMA_STATUS_OK) acc = U < fmin(1.0, exp(m.beta[c] * (np - m.pdg[it])));
      double p = m.pdg[it];
      if (acc) {
        p = np;
        m.pdg[it] = p;
      } else {
        restore_item(v, m, it);
      }
      m.acc[it] += acc ? 1u : 0u;
      part += p;
      if (!with_scale) propose_item(v, m, it, sweep + 1);
    }
  } else {
    prior_accept(v, m, c, sweep, dyn);
    for (int l = threadIdx.x; l < m.L; l += blockDim.x) {
      part += m.pdg[c * m.L + l];
      if (!with_scale) propose_item(v, m, c * m.L + l, sweep + 1);
    }
    if (threadIdx.x == 0) part += m.probg[c];
  }
  if (!with_scale) {
    finish_chain(m, c, block_sum(part, sh), swaptries, sweep, Ssh, Bsh);
    return;
  }
  __syncthreads();  // restored times visible to the whole workgroup
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
  if (threadIdx.x == 0) okall = 1;
  __syncthreads();
  double a = 0.0;
  for (int l = threadIdx.x; l < m.L; l += blockDim.x) {
    const int it = c * m.L + l;
    if (m.status[it] != IMA_STATUS_OK) okall = 0;
    if (prior && m.wstatus[it] != IMA_STATUS_OK) okall = 0;
    const SlotMeta sm = v.slots[m.first_slot + it];
    const int scaled = (sm.n - 1) + (prior ? sm.nmig : 0);
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
    a += m.beta[c] * (pn - m.probg[c]) + (m.P - 1) * m.logs[c];
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
      const double np_ = new_loglik(m, it);
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
  for (int l = threadIdx.x; l < m.L; l += blockDim.x) propose_item(v, m, c * m.L + l, sweep + 1);
  finish_chain(m, c, block_sum(part, sh), swaptries, sweep, Ssh, Bsh);
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
  // score_ev (optional): events recorded around the two scoring passes
  // the proposals' genealogy weights run on the side stream beside the
  // likelihood launch; the accept kernel waits for both
  auto score = [&](int pass) -> hipError_t {
    hipError_t e = hipSuccess;
    if (prior) {
      e = hipEventRecord(fj[2 * pass], st);
      if (e == hipSuccess) e = hipStreamWaitEvent(side, fj[2 * pass], 0);
      if (e == hipSuccess)
        e = launch_treeweight(v, m.tw_slots, m.C * m.L, m.tv, m.tv_row, m.P, m.emax, m.wnew,
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
  };
  hipError_t e = score(0);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mc3_accept, dim3(m.C), dim3(1024), dl, st, v, m, with_scale ? 1 : 0,
                     swaptries);
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
