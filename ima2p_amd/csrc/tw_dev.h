// tw_dev.h -- treeweight (update_gtree_common.cpp:1680-1931) for one
// genealogy per 64-lane workgroup, shared by the weights kernel and the
// sweep's combined likelihood-and-weights launch.
//
// Every coalescence (at the time of its node, in the population of the
// node's edge), migration (in the population the lineage is in just before
// it, after any splits it has passed) and split time below the root is an
// event; the events are put in time order by a rank sort in LDS (rank =
// events earlier, ties broken by index).  The sweep over the sorted events
// is uniform across the wave: lane p follows population p -- its lineage
// count, and for the current period its coalescence and migration counts,
// sum n(n-1) dt / (2h) and sum n dt -- accumulating in event order exactly
// as the sequential sweep does, so the weights match it bit for bit
// (floating-point contraction is off in the function).  Split k merges the
// counts of join[k] through lane shuffles.
//
// Weight record per genealogy (doubles, P = sampled populations = periods,
// positions i, j index plist[k]):
//   cc[k][i] | fc[k][i] | hcc[k][i] | mc[k][i][j] | fm[k][i][j] | length
// at offsets 0, P^2, 2P^2, 3P^2, 3P^2 + P^3, 3P^2 + 2P^3.
#pragma once

#include "ima_internal.h"

namespace ima {
namespace twdev {

enum : int { EV_COAL = 0, EV_MIG = 1, EV_SPLIT = 2 };

__device__ inline int ev_pack(int type, int pop, int topop) {
  return type | (pop << 4) | ((topop + 1) << 12);
}

__device__ inline double bcast_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ inline int period_of(const double *tv, int nper, double t) {
  int k = 0;
  while (k < nper - 1 && tv[k] <= t) k++;
  return k;
}

// The weight record of genealogy i (slot slots[i], split times row
// tv_row[i] of tvals, or row i) into rec_out[i], with status[i]; called by
// one 64-lane workgroup with lds (dynamic LDS) of treeweight_lds_bytes.
__device__ __forceinline__ void treeweight_one(const DeviceView &v, const int32_t *slots, int i,
                                               const double *tvals, const int32_t *tv_row,
                                               double *rec_out, int32_t *status, double *lds) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & (kWave - 1);
  const SlotMeta sm = v.slots[slots[i]];
  const LocusMeta &lm = v.loci[sm.locus];
  const ModelMeta mm = v.models[lm.pmodel];
  const ModelView md = model_view(v, mm);
  const int P = mm.npops, nper = P, T = 2 * P - 1;
  const int R = prior_rec_len(P);
  const double *tvg = tvals + (int64_t)(tv_row ? tv_row[i] : i) * (P > 1 ? P - 1 : 1);
  const int n = sm.n;
  const int32_t *up0 = v.up0 + sm.node_off;
  const double *tm = v.time + sm.node_off;
  const int32_t *pop = v.pop + sm.node_off;
  const int32_t *me = v.mig_edge + sm.mig_off;
  const int32_t *mt_to = v.mig_to + sm.mig_off;
  const double *mtime = v.mig_time + sm.mig_off;
  const int nmig = sm.nmig;

  // the model's small tables, read inside the event loops, live in LDS
  __shared__ int s_pos[kMaxPops * (2 * kMaxPops - 1)], s_join[3 * kMaxPops];
  __shared__ int s_pope[2 * kMaxPops - 1], s_popdown[2 * kMaxPops - 1];
  __shared__ int s_npresent[kMaxPops];
  __shared__ double s_tv[kMaxPops];
  __shared__ int bad;
  __shared__ int s_spos[kMaxPops];  // sorted position of split k
  for (int x = lane; x < P * T; x += kWave) s_pos[x] = md.pos[x];
  for (int x = lane; x < 3 * P; x += kWave) s_join[x] = md.join[x];
  for (int x = lane; x < T; x += kWave) {
    s_pope[x] = md.pope[x];
    s_popdown[x] = md.popdown[x];
  }
  for (int x = lane; x < P; x += kWave) s_npresent[x] = md.npresent[x];
  for (int x = lane; x < P - 1; x += kWave) s_tv[x] = tvg[x];
  if (lane == 0) bad = 0;
  __syncthreads();
  const double *tv = s_tv;

  const double roottime = tm[up0[sm.root]];
  const int kroot = period_of(tv, nper, roottime);
  const int E = (n - 1) + nmig + kroot;

  double *rec = lds;                       // [R]
  double *et = rec + R;                    // [E] unsorted times
  double *st = et + E;                     // [E] sorted times
  double *tfc = st + E;                    // [T][64] fc terms of a chunk
  double *tfm = tfc + T * kWave;           // [T][64] fm terms
  double *tln = tfm + T * kWave;           // [64] length terms
  int *ek = (int *)(tln + kWave);          // [E] unsorted codes
  int *sk = ek + E;                        // [E] sorted codes
  for (int r = lane; r < R; r += kWave) rec[r] = 0.0;

  // events
  for (int e = lane; e < n - 1; e += kWave) {
    const int edge = n + e;
    et[e] = tm[up0[edge]];
    ek[e] = ev_pack(EV_COAL, pop[edge], -1);
  }
  for (int j = lane; j < nmig; j += kWave) {
    const int edge = me[j];
    const double t = mtime[j];
    int now = (j > 0 && me[j - 1] == edge) ? mt_to[j - 1] : pop[edge];
    const int k = period_of(tv, nper, t);
    while (now >= 0 && now < T && s_pope[now] != -1 && s_pope[now] <= k) now = s_popdown[now];
    et[n - 1 + j] = t;
    ek[n - 1 + j] = ev_pack(EV_MIG, now, mt_to[j]);
  }
  for (int k = lane; k < kroot; k += kWave) {
    et[n - 1 + nmig + k] = tv[k];
    ek[n - 1 + nmig + k] = ev_pack(EV_SPLIT, k, -1);
  }
  __syncthreads();
  for (int j = lane; j < E; j += kWave) {
    const double t = et[j];
    int r = 0;
    for (int q = 0; q < E; ++q) {
      const double u = et[q];
      r += (u < t) || (u == t && q < j);
    }
    st[r] = t;
    sk[r] = ek[j];
    if (j >= n - 1 + nmig) s_spos[j - (n - 1 + nmig)] = r;
  }
  __syncthreads();

  // walk: the sorted events in chunks of up to 64 (lane l = event j0 + l),
  // never past the split that ends the current period.  A population's
  // lineage count before each event of the chunk is its count at the
  // chunk's start plus the migrations into it, minus the coalescences and
  // migrations out of it, among the chunk's earlier events (lane ballots);
  // every event's terms n(n-1) dt / (2h), n dt and (all lineages) dt then
  // come out in parallel, and lane p adds population p's terms in event
  // order, the sequence of additions of the event-by-event sweep.  Lanes
  // add the terms of absent populations too (and fm in the last period):
  // those sums are never written, and every sum restarts at a split.
  const int p = lane;
  const bool mine = p < T;
  int cnt = (p < P) ? lm.samppop[p] : 0;  // lane p: population p's count
  int nsum = n, k = 0, cc = 0, mybad = 0;
  double last = 0.0, len = 0.0, fcacc = 0.0, fmacc = 0.0;
  const double h2term = 1 / (2 * lm.hval);
  int ip = mine ? s_pos[p] : -1;
  const uint64_t below = (1ull << lane) - 1ull;  // lanes before this one
  int j0 = 0;
  while (j0 < E) {
    const int jsplit = k < kroot ? s_spos[k] : E;  // split k ends the period
    // split times out of order would put split k behind the walk: flagged,
    // and the walk still advances
    if (jsplit < j0) mybad = 1;
    const int nc = min(kWave, min(max(jsplit + 1, j0 + 1), E) - j0);
    const int j = j0 + lane;
    const bool in = lane < nc;
    const double t = in ? st[j] : 0.0;
    const int code = in ? sk[j] : 0;
    const double tp = __shfl_up(t, 1, kWave);
    const double dt = t - (lane == 0 ? last : tp);
    const int type = code & 15, a = (code >> 4) & 255, b = (code >> 12) - 1;
    const bool ev = in && j != jsplit;  // a coalescence or migration
    const bool coal = ev && type == EV_COAL, mig = ev && type == EV_MIG;
    // period positions of the event's populations; the bad cases of the
    // sequential sweep (a lineage in an absent population, a migration to
    // one)
    const int ipa = (a < T) ? s_pos[k * T + a] : -1;
    const int ib = (b >= 0 && b < T) ? s_pos[k * T + b] : -1;
    if ((coal && ipa < 0) || (mig && (ipa < 0 || ib < 0))) mybad = 1;
    if (mig && ipa >= 0 && ib >= 0) atomicAdd(&rec[3 * P * P + (k * P + ipa) * P + ib], 1.0);
    const uint64_t mcoal = __ballot(coal);
    const int nbefore = nsum - (int)__popcll(mcoal & below);
    tln[lane] = nbefore * dt;
    for (int q = 0; q < T; ++q) {
      const uint64_t mfrom = __ballot((coal || mig) && a == q);
      const uint64_t mto = __ballot(mig && b == q);
      const int c0 = __builtin_amdgcn_readlane(cnt, q);
      const int c = c0 + (int)__popcll(mto & below) - (int)__popcll(mfrom & below);
      tfc[q * kWave + lane] = ((double)c * ((double)c - 1)) * dt * h2term;
      tfm[q * kWave + lane] = c * dt;
      if (lane == q) {
        cnt = c0 + (int)__popcll(mto) - (int)__popcll(mfrom);
        cc += (int)__popcll(mfrom & mcoal);
      }
    }
    wave_lds_fence();
    {
      // lanes past T add population 0's terms: never written
      const double *fc = tfc + (mine ? p : 0) * kWave, *fm = tfm + (mine ? p : 0) * kWave;
      for (int l = 0; l < nc; ++l) {
        len += tln[l];
        fcacc += fc[l];
        fmacc += fm[l];
      }
    }
    nsum -= (int)__popcll(mcoal);
    last = bcast_d(t, nc - 1);
    j0 += nc;
    wave_lds_fence();  // the term buffers are refilled by the next chunk
    if (j0 == jsplit + 1) {
      // split k: the new population takes the two merged counts
      const int ja = s_join[3 * k], jb_ = s_join[3 * k + 1], jn = s_join[3 * k + 2];
      const int na = __shfl(cnt, ja < 0 ? 0 : ja, kWave);
      const int nbb = __shfl(cnt, jb_ < 0 ? 0 : jb_, kWave);
      if (ip >= 0) {
        rec[k * P + ip] = (double)cc;
        rec[P * P + k * P + ip] = fcacc;
        if (k < nper - 1)
          for (int q = 0; q < s_npresent[k]; ++q)
            if (q != ip) rec[3 * P * P + P * P * P + (k * P + ip) * P + q] = fmacc;
      }
      if (p == jn) cnt = na + nbb;
      if (p == ja || p == jb_) cnt = 0;
      cc = 0;
      fcacc = fmacc = 0.0;
      k++;
      ip = mine ? s_pos[k * T + p] : -1;
    }
  }
  if (ip >= 0) {
    rec[k * P + ip] = (double)cc;
    rec[P * P + k * P + ip] = fcacc;
    if (k < nper - 1)
      for (int jj = 0; jj < s_npresent[k]; ++jj)
        if (jj != ip) rec[3 * P * P + P * P * P + (k * P + ip) * P + jj] = fmacc;
  }
  if (mybad) bad = 1;
  __syncthreads();
  const double hlog = log(lm.hval);
  if (hlog != 0.0)
    for (int r = lane; r < P * P; r += kWave) rec[2 * P * P + r] = hlog * rec[r];
  if (lane == 0) rec[R - 1] = len;
  __syncthreads();
  double *out = rec_out + (int64_t)i * R;
  for (int r = lane; r < R; r += kWave) out[r] = rec[r];
  if (lane == 0 && status) status[i] = bad ? IMA_STATUS_PRIOR_BAD : IMA_STATUS_OK;
}

}  // namespace twdev
}  // namespace ima
