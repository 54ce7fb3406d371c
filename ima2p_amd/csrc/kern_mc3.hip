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
// Work mapping: per-item kernels use one thread per (chain, locus);
// per-node kernels one thread per genealogy node (node_item maps a node to
// its item); per-chain reductions one wavefront per chain.
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
__device__ __forceinline__ int wand(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v &= __shfl_xor(v, off, kWave);
  return v;
}

__global__ void mc3_slide_propose(DeviceView v, MC3Dev m) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= m.C * m.L) return;
  const uint64_t sweep = *m.counter;
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

__global__ void mc3_slide_accept(DeviceView v, MC3Dev m) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= m.C * m.L) return;
  const uint64_t sweep = *m.counter;
  const int c = it / m.L;
  PhiloxStream r = item_rng(m, it, sweep, 1);
  const double U = r.uniform();
  bool acc = false;
  if (m.status[it] == IMA_STATUS_OK) {
    const double mh = exp(m.beta[c] * (m.newpdg[it] - m.pdg[it]));
    acc = U < fmin(1.0, mh);  // update_gtree.cpp:919-927
  }
  if (acc) {
    m.pdg[it] = m.newpdg[it];
  } else {
    const SlotMeta sm = v.slots[m.first_slot + it];
    const int e = m.rec_v[it];
    double *tm = const_cast<double *>(v.time) + sm.node_off;
    const int32_t *up0 = v.up0 + sm.node_off, *up1 = v.up1 + sm.node_off;
    tm[up0[e]] = m.rec_t[it];
    tm[up1[e]] = m.rec_t[it];
  }
  m.acc[it] += acc ? 1u : 0u;
}

// One log scale factor per chain.
__global__ void mc3_scale_draw(MC3Dev m) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.C) return;
  PhiloxStream r = item_rng(m, m.C * m.L + c, *m.counter, 2);
  m.logs[c] = m.scale_width * (r.uniform() - 0.5);
}

// Node-parallel rescaling of every coalescence time (root edge excluded).
__global__ void mc3_scale_apply(DeviceView v, MC3Dev m) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m.nnodes) return;
  const int64_t g = m.node_first + j;
  const double t = v.time[g];
  m.tbackup[j] = t;
  if (v.down[g] != -1) {
    const int c = m.node_item[j] / m.L;
    const_cast<double *>(v.time)[g] = t * exp(m.logs[c]);
  }
}

// One wavefront per chain: log acceptance = sum_l beta (new - old) +
// (n_l - 1) log s.
__global__ __launch_bounds__(kWave) void mc3_scale_decide(DeviceView v, MC3Dev m) {
  const int c = blockIdx.x, lane = threadIdx.x;
  double a = 0.0;
  int ok = 1;
  for (int l = lane; l < m.L; l += kWave) {
    const int it = c * m.L + l;
    ok &= (m.status[it] == IMA_STATUS_OK);
    const int n = v.slots[m.first_slot + it].n;
    a += m.beta[c] * (m.newpdg[it] - m.pdg[it]) + (n - 1) * m.logs[c];
  }
  a = wsum(a);
  ok = wand(ok);
  if (lane == 0) {
    PhiloxStream r = item_rng(m, m.C * m.L + m.C + c, *m.counter, 3);
    const double U = r.uniform();
    const int acc = ok && (U < fmin(1.0, exp(a)));
    m.decision[c] = acc;
    m.cacc[c] += (uint32_t)acc;
  }
}

__global__ void mc3_scale_commit_items(MC3Dev m) {
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= m.C * m.L) return;
  if (m.decision[it / m.L]) m.pdg[it] = m.newpdg[it];
}

__global__ void mc3_scale_commit_nodes(DeviceView v, MC3Dev m) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m.nnodes) return;
  if (!m.decision[m.node_item[j] / m.L])
    const_cast<double *>(v.time)[m.node_first + j] = m.tbackup[j];
}

// Chain sums S_c = sum_l pdg (swapweight, swapchains.cpp:13-34) with one
// wavefront per chain, then the swap attempts of swapchains on one lane,
// then advance the sweep counter.
__global__ __launch_bounds__(256) void mc3_swap(MC3Dev m, int swaptries) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  for (int c = w; c < m.C; c += blockDim.x / kWave) {
    double s = 0.0;
    for (int l = lane; l < m.L; l += kWave) s += m.pdg[c * m.L + l];
    s = wsum(s);
    if (lane == 0) m.S[c] = s;
  }
  __syncthreads();
  __threadfence_block();
  if (threadIdx.x != 0) return;
  const uint64_t sweep = *m.counter;
  PhiloxStream rng(m.seed ^ 0x5A5A5A5A5A5A5A5Aull, sweep);
  const int SWAPDIST = 7, nchains = m.C;
  const double *S = m.S;
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
    const double wgt = exp((m.beta[ci] - m.beta[cj]) * (S[cj] - S[ci]));
    if (wgt >= 1.0 || wgt > rng.uniform()) {
      const double b = m.beta[ci];
      m.beta[ci] = m.beta[cj];
      m.beta[cj] = b;
      ++acc;
    }
  }
  m.stats[4] += (unsigned long long)(nchains > 1 ? swaptries : 0);
  m.stats[5] += acc;
  m.stats[0] += (unsigned long long)(m.C * m.L);
  *m.counter = sweep + 1;
}

// Enqueue the kernels of one sweep (no host synchronisation).
hipError_t enqueue_mc3_sweep(const DeviceView &v, const MC3Dev &m, bool with_scale,
                             const ima_item *items, const int32_t *idx, const HkyPlan &plan,
                             int nh, int ni, int ns, int precision, int hmax_n, int hmax_slots, int imax_n,
                             int imax_w, int swaptries, hipEvent_t *score_ev,
                             hipStream_t st) {
  const int N = m.C * m.L;
  const int tb = 256, gb = (N + tb - 1) / tb;
  const int gn = (int)((m.nnodes + tb - 1) / tb);
  // score_ev (optional): events recorded around the two scoring passes
  auto score = [&](int pass) -> hipError_t {
    hipError_t e = hipSuccess;
    if (score_ev) e = hipEventRecord(score_ev[2 * pass], st);
    if (e == hipSuccess && nh)
      e = launch_hky(v, items, plan, precision, hmax_n, hmax_slots, m.newpdg, m.status, st);
    if (e == hipSuccess && ni)
      e = launch_is(v, items, idx + nh, ni, imax_n, imax_w, m.newpdg, m.status, nullptr, nullptr, st);
    if (e == hipSuccess && ns)
      e = launch_sw(v, items, idx + nh + ni, ns, m.newpdg, m.status, nullptr, st);
    if (e == hipSuccess && score_ev) e = hipEventRecord(score_ev[2 * pass + 1], st);
    return e;
  };
  hipLaunchKernelGGL(mc3_slide_propose, dim3(gb), dim3(tb), 0, st, v, m);
  hipError_t e = score(0);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mc3_slide_accept, dim3(gb), dim3(tb), 0, st, v, m);
  if (with_scale) {
    hipLaunchKernelGGL(mc3_scale_draw, dim3((m.C + 63) / 64), dim3(64), 0, st, m);
    hipLaunchKernelGGL(mc3_scale_apply, dim3(gn), dim3(tb), 0, st, v, m);
    e = score(1);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mc3_scale_decide, dim3(m.C), dim3(kWave), 0, st, v, m);
    hipLaunchKernelGGL(mc3_scale_commit_items, dim3(gb), dim3(tb), 0, st, m);
    hipLaunchKernelGGL(mc3_scale_commit_nodes, dim3(gn), dim3(tb), 0, st, v, m);
  }
  hipLaunchKernelGGL(mc3_swap, dim3(1), dim3(256), 0, st, m, swaptries);
  return hipGetLastError();
}

}  // namespace ima
