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
// by its daughters' coalescences and its parent's; the step scale depends
// only on those bounds, so the move is symmetric.  Topology is unchanged.
// The rescaling multiplies every coalescence time of a chain's genealogies
// by s = exp(d (U - 1/2)), with Hastings term (n-1) log s per genealogy.
// Neither is the reference's proposal (slider/addmigration); they exercise
// the same scoring, acceptance, commit/rollback and swap path.
//
// Kernels per sweep: the batched likelihood, then one 1024-thread workgroup
// per chain that accepts or rejects every locus of its chain (merging the
// likelihood's pattern-block partials in block order), applies the
// rescaling or computes the chain sum, and draws the next sweep's
// genealogy proposals.  The workgroup that finishes last (device-scope
// fence + counter) performs the swap.  A rescaling sweep adds a second
// likelihood launch and a per-chain commit kernel.  Every reduction has a
// fixed order, so results do not depend on scheduling.
#include <float.h>

#include "ima_internal.h"
#include "philox.h"

namespace ima {

__device__ inline PhiloxStream item_rng(const MC3Dev &m, int item, uint64_t sweep, int phase) {
  return PhiloxStream(m.seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(item + 1)),
                      sweep * 8 + (uint64_t)phase);
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
  const double lo = fmax(la, lb);
  const bool root = dn[e] == -1;
  const double hi = root ? DBL_MAX : tm[e];
  const double sigma = root ? 0.5 * fmax(lo, 0.05) : 0.5 * (hi - lo);
  double Tn = T + sigma * normal01(r);
  if (root) {
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
constexpr int kSwapLds = 2048;
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
    PhiloxStream rng(m.seed ^ 0x5A5A5A5A5A5A5A5Aull, sweep);
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

// One workgroup per chain.  Accept or reject every locus' genealogy
// proposal (update_gtree.cpp:919-927; rejection restores the moved time).
// Rescaling sweep: back up and rescale every coalescence time of the chain.
// Otherwise: chain sum, swap by the last chain, next sweep's proposals.
__global__ __launch_bounds__(1024) void mc3_accept(DeviceView v, MC3Dev m, int with_scale,
                                                   int swaptries) {
  __shared__ double sh[1024 / kWave];
  __shared__ double Ssh[kSwapLds], Bsh[kSwapLds];
  const int c = blockIdx.x;
  const uint64_t sweep = *m.counter;
  double part = 0.0;
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
}

// One workgroup per chain: log acceptance of the rescaling
// = sum_l beta (new - old) + (n_l - 1) log s; commit or restore; chain sum;
// swap by the last chain; next sweep's proposals.
__global__ __launch_bounds__(1024) void mc3_scale_finish(DeviceView v, MC3Dev m, int swaptries) {
  __shared__ double sh[1024 / kWave];
  __shared__ double Ssh[kSwapLds], Bsh[kSwapLds];
  __shared__ int okall, decision;
  const int c = blockIdx.x;
  const uint64_t sweep = *m.counter;
  if (threadIdx.x == 0) okall = 1;
  __syncthreads();
  double a = 0.0;
  for (int l = threadIdx.x; l < m.L; l += blockDim.x) {
    const int it = c * m.L + l;
    if (m.status[it] != IMA_STATUS_OK) okall = 0;
    const int n = v.slots[m.first_slot + it].n;
    a += m.beta[c] * (new_loglik(m, it) - m.pdg[it]) + (n - 1) * m.logs[c];
  }
  a = block_sum(a, sh);
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
      const double np = new_loglik(m, it);
      m.pdg[it] = np;
      part += np;
    }
  } else {
    for (int l = threadIdx.x; l < m.L; l += blockDim.x) part += m.pdg[c * m.L + l];
    int64_t j0, j1;
    chain_nodes(v, m, c, &j0, &j1);
    double *tm = const_cast<double *>(v.time) + m.node_first;
    for (int64_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) tm[j] = m.tbackup[j];
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
                             hipStream_t st) {
  // the per-chain kernels merge pattern blocks themselves
  HkyPlan p = plan;
  p.nred = 0;
  // score_ev (optional): events recorded around the two scoring passes
  auto score = [&](int pass) -> hipError_t {
    hipError_t e = hipSuccess;
    if (score_ev) e = hipEventRecord(score_ev[2 * pass], st);
    if (e == hipSuccess && nh)
      e = launch_hky(v, items, p, precision, hmax_n, hmax_slots, m.newpdg, m.status, st);
    if (e == hipSuccess && ni)
      e = launch_is(v, items, idx + nh, ni, imax_n, imax_w, m.newpdg, m.status, nullptr, nullptr, st);
    if (e == hipSuccess && ns)
      e = launch_sw(v, items, idx + nh + ni, ns, m.newpdg, m.status, nullptr, st);
    if (e == hipSuccess && score_ev) e = hipEventRecord(score_ev[2 * pass + 1], st);
    return e;
  };
  hipError_t e = score(0);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mc3_accept, dim3(m.C), dim3(1024), 0, st, v, m, with_scale ? 1 : 0,
                     swaptries);
  if (with_scale) {
    e = score(1);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mc3_scale_finish, dim3(m.C), dim3(1024), 0, st, v, m, swaptries);
  }
  return hipGetLastError();
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
