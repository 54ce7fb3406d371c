// kern_urate.hip -- the mutation-rate round of the device sweep.
//
// changeu (update_mc_params.cpp:24-382), called from qupdate every fifth
// step for every chain (ima_main_mpi.cpp:1917-1945): for j = 0 .. L-1
// (only j = 0 with two loci) pick k != j uniformly, move r = log(u_j/u_k)
// by a uniform window of log(UMAX)/L, reflect it at +-3 log(UMAX), set
// u_j *= d and u_k /= d with d = exp((r' - r)/2), move each HKY locus'
// kappa (window 2, reflected at 0 and KAPPAMAX), re-score both loci and
// accept with U < min(1, exp(beta sum(new - old))).  One HKY locus alone:
// changekappa (:387-429), the kappa move with the same re-scoring.
//
// The sequential j loop becomes levels: urate_plan draws every step's
// partner and uniforms (each step has its own stream) and gives a step
// the level one above the highest level of the earlier steps touching
// either of its loci.  The steps of one level touch disjoint loci and
// read the rates the sequential loop would have left, so running a
// level's steps together gives the sequential loop's result.  Per level:
// scoring launches over the listed HKY units and the listed IS and SW
// items, and one urate_step launch that decides the level and proposes the
// next.

#include <float.h>

#include "ima_internal.h"

namespace ima {
namespace {

__device__ inline bool item_is_hky(const DeviceView &v, const MC3Dev &m, int it) {
  return v.loci[v.slots[m.first_slot + it].locus].model == IMA_MODEL_HKY;
}

// Queue item it for the level of list parity lp: its HKY pattern-block
// units, or the item itself on the IS or SW list.
__device__ inline void queue_units(const MC3Dev &m, const URateDev &u, int it, int lp);
__device__ inline void queue_item(const DeviceView &v, const MC3Dev &m, const URateDev &u,
                                  int it, int lp) {
  const int model = v.loci[v.slots[m.first_slot + it].locus].model;
  if (model == IMA_MODEL_HKY) {
    queue_units(m, u, it, lp);
  } else if (model == IMA_MODEL_IS) {
    u.ilist[lp][atomicAdd(&u.cnt[2 + lp], 1)] = it;
  } else {
    u.slist[lp][atomicAdd(&u.cnt[4 + lp], 1)] = it;
  }
}

// Log-likelihood of item it from this level's scoring: the direct value, or
// its pattern blocks' partial sums added in block order.
__device__ inline double level_loglik(const MC3Dev &m, int it) {
  const double direct = m.newpdg[it];
  const int2 ib = m.iblk[it];
  if (ib.y <= 1) return direct;
  double s = 0.0;
  for (int b = 0; b < ib.y; ++b) s += m.partial[ib.x + b];
  return s;
}

// kappa move (update_mc_params.cpp:258-273, :394-404)
__device__ inline double kappa_move(double k, double U, const URateDev &u) {
  double nk;
  if (U > 0.5) {
    nk = k + (2.0 * U - 1.0) * u.kwin;
    if (nk > u.kmax) nk = 2.0 * u.kmax - nk;
  } else {
    nk = k - u.kwin * U * 2.0;
    if (nk < 0) nk = -nk;
  }
  return nk;
}

// Queue item it's HKY pattern-block units for the level of list parity lp.
__device__ inline void queue_units(const MC3Dev &m, const URateDev &u, int it, int lp) {
  const int2 ib = m.iblk[it];
  const int nb = ib.y <= 1 ? 1 : ib.y;
  const int pos = atomicAdd(&u.cnt[lp], nb);
  for (int b = 0; b < nb; ++b) u.list[lp][pos + b] = make_int4(it, b, ib.x, nb);
}

// Proposal of step s of chain c from the committed rates (update_mc_params
// .cpp:99-104 and 201-217 for the rates, 258-273 for kappa), written into
// the round's items; HKY items' units are queued for scoring.
__device__ void urate_propose(const DeviceView &v, const MC3Dev &m, const URateDev &u,
                              const ima_item *items, int c, int s, int lp) {
  const int4 st = u.step[c * u.J + s];
  const double4 U = u.uni[c * u.J + s];
  const int ij = c * m.L + s, k = st.x;
  ima_item a = items[ij];
  if (k >= 0) {
    const int ik = c * m.L + k;
    ima_item b = items[ik];
    const double r = log(a.u / b.u);
    double nr = U.x > 0.5 ? r + (2.0 * U.x - 1.0) * u.win : r - u.win * U.x * 2.0;
    if (nr > u.maxr)
      nr = 2.0 * u.maxr - nr;
    else if (nr < -u.maxr)
      nr = 2.0 * (-u.maxr) - nr;
    const double d = exp((nr - r) / 2);
    a.u *= d;
    b.u /= d;
    if (item_is_hky(v, m, ik)) b.kappa = kappa_move(b.kappa, U.z, u);
    u.uitems[ik] = b;
    queue_item(v, m, u, ik, lp);
  }
  if (item_is_hky(v, m, ij)) a.kappa = kappa_move(a.kappa, U.y, u);
  u.uitems[ij] = a;
  queue_item(v, m, u, ij, lp);
}

// Decision of step s of chain c (update_mc_params.cpp:250-287 and 406-409):
// commit the proposed rates and likelihoods, or put the committed rates
// back into the round's items.
__device__ void urate_decide(const MC3Dev &m, const URateDev &u, ima_item *items, int c,
                             int s) {
  const int4 st = u.step[c * u.J + s];
  const double Ua = u.uni[c * u.J + s].w;
  const int ij = c * m.L + s, k = st.x, ik = k >= 0 ? c * m.L + k : -1;
  bool ok = m.status[ij] == IMA_STATUS_OK;
  const double nj = level_loglik(m, ij);
  double sum = 0.0;
  sum += nj - m.pdg[ij];
  double nk = 0.0;
  if (ik >= 0) {
    ok = ok && m.status[ik] == IMA_STATUS_OK;
    nk = level_loglik(m, ik);
    sum += nk - m.pdg[ik];
  }
  const double mh = exp(m.beta[c] * sum);
  // changeu takes U < min(1, mh); changekappa mh >= 1 or mh > U
  const bool acc = ok && (ik >= 0 ? Ua < fmin(1.0, mh) : (mh >= 1.0 || mh > Ua));
  if (acc) {
    items[ij] = u.uitems[ij];
    m.pdg[ij] = nj;
    if (ik >= 0) {
      items[ik] = u.uitems[ik];
      m.pdg[ik] = nk;
    }
  } else {
    u.uitems[ij] = items[ij];
    if (ik >= 0) u.uitems[ik] = items[ik];
  }
  atomicAdd(&m.stats[6], 1ull);
  if (acc) atomicAdd(&m.stats[7], 1ull);
}

// One workgroup per chain: copy the chain's items, draw every step's
// partner and uniforms in the reference's order (partner, window, kappa of
// j, kappa of k, acceptance; a kappa uniform only for an HKY locus), one
// lane per step since each step has its own stream; then one lane gives
// each step its level (bookkeeping in LDS), the steps are sorted by level
// and level 1 is proposed.  serial != 0 puts step s at level s + 1 (the
// sequential loop itself).  Dynamic LDS: urate_plan_lds(L, J).
__global__ __launch_bounds__(256) void urate_plan(DeviceView v, MC3Dev m, URateDev u,
                                                  const ima_item *items, uint64_t ctr,
                                                  int serial) {
  extern __shared__ int32_t pl[];
  const int c = blockIdx.x, L = m.L, J = u.J;
  int32_t *last = pl, *off = pl + L, *kk = off + (J + 2), *lv = kk + J, *ord = lv + J;
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    u.uitems[c * L + i] = items[c * L + i];
    last[i] = 0;
  }
  for (int l = threadIdx.x; l <= J + 1; l += blockDim.x) off[l] = 0;
  for (int s = threadIdx.x; s < J; s += blockDim.x) {
    PhiloxStream r = urate_rng(m.seed, c, L, s, ctr);
    const int k = urate_partner(r, s, L);
    double4 U;
    U.x = k >= 0 ? r.uniform() : 0.0;
    U.y = item_is_hky(v, m, c * L + s) ? r.uniform() : 0.0;
    U.z = (k >= 0 && item_is_hky(v, m, c * L + k)) ? r.uniform() : 0.0;
    U.w = r.uniform();
    u.uni[c * J + s] = U;
    kk[s] = k;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int s = 0; s < J; ++s) {
      const int k = kk[s];
      const int lev = serial ? s + 1 : 1 + max(last[s], k >= 0 ? last[k] : 0);
      last[s] = lev;
      if (k >= 0) last[k] = lev;
      lv[s] = lev;
      off[lev] += 1;
    }
    // off[l] <- first position of level l; last[l - 1] is level l's cursor
    int run = 0;
    for (int l = 1; l <= J + 1; ++l) {
      const int t = off[l];
      off[l] = run;
      if (l <= J) last[l - 1] = run;
      run += t;
    }
    for (int s = 0; s < J; ++s) ord[last[lv[s] - 1]++] = s;
  }
  __syncthreads();
  for (int s = threadIdx.x; s < J; s += blockDim.x) {
    u.step[c * J + s] = make_int4(kk[s], lv[s], 0, 0);
    u.order[(int64_t)c * J + s] = ord[s];
  }
  for (int l = threadIdx.x; l <= J + 1; l += blockDim.x) u.loff[(int64_t)c * (J + 2) + l] = off[l];
  __syncthreads();  // the steps' records visible to the proposals
  for (int i = off[1] + threadIdx.x; i < off[2]; i += blockDim.x)
    urate_propose(v, m, u, items, c, ord[i], 1);
}

// One workgroup per chain: decide level lev, then (unless it is the last
// level) propose level lev + 1 from the rates just committed.
__global__ __launch_bounds__(256) void urate_step(DeviceView v, MC3Dev m, URateDev u,
                                                  ima_item *items, int lev, int last) {
  const int c = blockIdx.x, J = u.J;
  const int32_t *off = u.loff + (int64_t)c * (J + 2);
  const int32_t *ord = u.order + (int64_t)c * J;
  for (int i = off[lev] + threadIdx.x; i < off[lev + 1]; i += blockDim.x)
    urate_decide(m, u, items, c, ord[i]);
  if (last) {
    // the next round's plan queues level 1 on list 1 (nothing queues now)
    if (c == 0 && threadIdx.x < 6) u.cnt[threadIdx.x] = 0;
    return;
  }
  __syncthreads();  // the level's commits visible to the next level's proposals
  for (int i = off[lev + 1] + threadIdx.x; i < off[lev + 2]; i += blockDim.x)
    urate_propose(v, m, u, items, c, ord[i], (lev + 1) & 1);
}

}  // namespace

size_t urate_plan_lds(int L, int J) { return 4 * ((size_t)L + 4 * (size_t)J + 2); }

hipError_t enqueue_urate_round(const DeviceView &v, const MC3Dev &m, const URateDev &u,
                               const ima_item *items, const int32_t *idx, const HkyPlan &plan,
                               int nh, int ni, int ns, int precision, int hmax_n,
                               int hmax_slots, int imax_n, int imax_w, uint64_t ctr,
                               int nlev, int serial, hipStream_t st) {
  if (nlev < 1 || nlev > u.J) return hipErrorInvalidValue;
  const size_t lds = urate_plan_lds(m.L, u.J);
  if (lds > 65536) return hipErrorInvalidValue;
  ima_item *it = const_cast<ima_item *>(items);
  // u.cnt is zero here: set up so, and left so by every round's last step
  hipLaunchKernelGGL(urate_plan, dim3(m.C), dim3(256), lds, st, v, m, u, items, ctr, serial);
  for (int lev = 1; lev <= nlev; ++lev) {
    hipError_t e = hipSuccess;
    if (nh)
      e = launch_hky_list(v, u.uitems, u.list[lev & 1], u.cnt + (lev & 1),
                          u.cnt + ((lev + 1) & 1), plan.nunits, precision, hmax_n, hmax_slots,
                          plan.ppl, m.partial, m.newpdg, m.status, st);
    if (e == hipSuccess && ni)
      e = launch_is(v, u.uitems, u.ilist[lev & 1], ni, imax_n, imax_w, m.newpdg, m.status,
                    nullptr, nullptr, st, u.cnt + 2 + (lev & 1), u.cnt + 2 + ((lev + 1) & 1));
    if (e == hipSuccess && ns)
      e = launch_sw(v, u.uitems, u.slist[lev & 1], ns, m.newpdg, m.status, nullptr, st,
                    u.cnt + 4 + (lev & 1), u.cnt + 4 + ((lev + 1) & 1));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(urate_step, dim3(m.C), dim3(256), 0, st, v, m, u, it, lev,
                       lev == nlev ? 1 : 0);
  }
  return hipGetLastError();
}

}  // namespace ima
