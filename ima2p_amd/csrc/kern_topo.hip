// kern_topo.hip -- the reference's genealogy move on the device, for
// population models without migration (updategenealogy, update_gtree.cpp:
// 723-826, with slider_nomigration :79-253, joinsisdown :375-454 and
// splitsisdown :457-528), plus the stepwise allele of the moved node
// (finishSWupdateA, update_gtree_common.cpp:2219-2362).
//
// One wavefront per (chain, locus): the chain's genealogy is read into LDS,
// lane 0 makes the move there, and the moved genealogy, its postorder
// program and its Hastings term go to the item's proposal slot.  The sweep
// then scores every proposal in one batched launch and decides each
// chain's loci in order (the accept step with Hastings terms); accepted
// proposals are copied into the chains' slots.
//
// The move: an edge other than the root's is picked uniformly
// (randposint(numlines)); its parent node is freed by joining the sister
// and parent edges; the edge's top slides by a normal distance with
// standard deviation min(20, root time / 3), up toward the tips (reflecting
// at the edge's own daughters' node, or at the time the edge's and the
// sister's populations have merged), down toward the root, choosing a
// branch at random at every node passed; the freed node is put back where
// the slide stops, with the population of the branch it splits at that
// time.  When the root moved the Hastings term is
// log N(d; 0, s_new) - log N(d; 0, s_old).  At a stepwise locus the freed
// node's allele is redrawn: a geometric step (parameter from the allele
// differences to its new neighbours, at most 0.95) up or down from its old
// value, kept inside the locus' allele range, with the Hastings term of the
// reverse draw from its old neighbours.
#include <float.h>

#include "ima_internal.h"
#include "mig_dev.h"
#include "philox.h"

namespace ima {
namespace {

constexpr int kMaxSlideSteps = 1 << 16;  // nodes a slide may pass (never reached in practice)
constexpr double kSlideStdvMax = 20.0;    // SLIDESTDVMAX

__device__ inline PhiloxStream topo_rng(const MC3Dev &m, int it, uint64_t sweep) {
  return PhiloxStream(m.seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(it + 1)),
                      ((sweep << 3) | 6ull) << 16);
}

__device__ inline int randposint(PhiloxStream &r, int n) {
  int k;
  do k = (int)(r.uniform() * n);
  while (k < 0 || k >= n);
  return k;
}

__device__ inline double normal01(PhiloxStream &r) {
  const double u1 = 1.0 - r.uniform(), u2 = r.uniform();  // u1 in (0, 1]
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// log of the normal density N(x; 0, s)
__device__ inline double log_normprob(double s, double x) {
  return -0.91893853320467274178 - log(s) - 0.5 * (x / s) * (x / s);
}

// findperiod (update_gtree_common.cpp:1626-1633)
__device__ inline int period_of(const double *tv, int P, double t) {
  int k = 0;
  while (k < P - 1 && tv[k] <= t) k++;
  return k;
}

// populations merged at the start of period p >= 1: the split closing
// period p - 1 (the reference's droppops[p])
__device__ inline bool dropped(const ModelView &md, int p, int q) {
  return md.join[3 * (p - 1)] == q || md.join[3 * (p - 1) + 1] == q;
}

// findjointime (update_gtree.cpp:35-75): when populations a (an edge's,
// from time ta) and b (from tb) have become one
__device__ double join_time(const ModelView &md, const double *tv, int P, int a, int b, double ta,
                            double tb) {
  int pa = period_of(tv, P, ta), pb = period_of(tv, P, tb);
  while (pa < pb) {
    pa++;
    if (dropped(md, pa, a)) a = md.popdown[a];
  }
  while (pb < pa) {
    pb++;
    if (dropped(md, pb, b)) b = md.popdown[b];
  }
  while (a != b && pa < P - 1) {
    pa++;
    if (dropped(md, pa, a)) a = md.popdown[a];
    if (dropped(md, pa, b)) b = md.popdown[b];
  }
  return pa == 0 ? 0.0 : tv[pa - 1];
}

__device__ inline double upper_time(const int32_t *u0, const double *tm, int n, int e) {
  return e < n ? 0.0 : tm[u0[e]];
}

// finishSWupdateA's proposal for the freed node's allele (update_gtree_common
// .cpp:2219-2362) and its Hastings term; the likelihood is scored in full
// afterwards.
__device__ __forceinline__ double sw_allele(PhiloxStream &r, const MC3Dev &m, int it, int32_t *A,
                            const int32_t *dn, int edge, int sis, int freed, int copyA0,
                            int copyA1, int old_down_of_freed, int hold_down_A) {
  const int oldA = A[freed];
  const int e3[3] = {edge, sis, dn[edge]};
  int j = 0, wsum = 0;
  for (int k = 0; k < 3; ++k) {
    if (dn[e3[k]] == -1) continue;
    wsum += abs((k < 2 ? A[e3[k]] : A[dn[e3[k]]]) - oldA);
    ++j;
  }
  double gnew = j / (double)(wsum + j);
  if (gnew > 0.95) gnew = 0.95;
  const double uni = 1.0 - r.uniform();  // geometric() by inversion, in (0, 1]
  int dA = (int)ceil(log(uni) / log(1.0 - gnew)) - 1;
  if (r.next32() & 1u) dA = -dA;
  const int2 rng_ = m.arange[it % m.L];
  const int newA = dA >= 0 ? min(rng_.y, oldA + dA) : max(rng_.x, oldA + dA);
  A[freed] = newA;
  dA = newA - oldA;
  int rsum = abs(copyA0 - newA) + abs(copyA1 - newA), rr = 2;
  if (old_down_of_freed >= 0) {
    rsum += abs(hold_down_A - newA);
    ++rr;
  }
  double gold = rr / (double)(rsum + rr);
  if (gold > 0.95) gold = 0.95;
  return (abs(dA) * log(1.0 - gold) + log(gold)) - (abs(dA) * log(1.0 - gnew) + log(gnew));
}

// The postorder program of a topology, as the host's emit builds it: each
// internal node's operation after those of its daughters' subtrees, the
// daughter with the larger Strahler need (the LDS partial slots its
// subtree uses; ties: up0) first.  The wavefront's lanes build it together
// in rounds that run up the tree and back down (as many as its height):
// need[e] becomes (internal nodes at or below e) << 8 | (slots e needs),
// stk[e] the position of the first operation of e's subtree, so e's own
// operation sits at stk[e] + (internal nodes at or below e) - 1.  Every
// lane calls it after the genealogy's last LDS write and gets the slot
// count; down[] must be the parents of up0[] and up1[].
__device__ __forceinline__ int wave_emit_program(int n, int ne, int root, const int32_t *u0, const int32_t *u1,
                                 const int32_t *dn, int32_t *need, int32_t *stk, ProgOp *prog,
                                 int lane) {
  for (int e = lane; e < ne; e += kWave) need[e] = e < n ? 0 : -1;
  wave_lds_fence();
  for (int round = 0; round < ne; ++round) {
    bool pend = false;
    for (int e = n + lane; e < ne; e += kWave) {
      if (need[e] >= 0) continue;
      const int va = need[u0[e]], vb = need[u1[e]];
      if (va < 0 || vb < 0) {
        pend = true;
        continue;
      }
      const int na = va & 255, nb = vb & 255;
      need[e] = (((va >> 8) + (vb >> 8) + 1) << 8) | (na == nb ? na + 1 : max(na, nb));
    }
    wave_lds_fence();
    if (!__ballot(pend)) break;
  }
  for (int e = lane; e < ne; e += kWave) stk[e] = e == root ? 0 : -1;
  wave_lds_fence();
  for (int round = 0; round < ne; ++round) {
    bool pend = false;
    for (int e = n + lane; e < ne; e += kWave) {
      const int p = dn[e];
      if (stk[e] >= 0 || p < 0) continue;
      const int sp = stk[p];
      if (sp < 0) {
        pend = true;
        continue;
      }
      int a = u0[p];
      const int b = u1[p];
      if ((need[a] & 255) < (need[b] & 255)) a = b;
      stk[e] = e == a ? sp : sp + (a >= n ? need[a] >> 8 : 0);
    }
    wave_lds_fence();
    if (!__ballot(pend)) break;
  }
  for (int e = n + lane; e < ne; e += kWave) {
    int a = u0[e], b = u1[e];
    if ((need[a] & 255) < (need[b] & 255)) {
      const int t = a;
      a = b;
      b = t;
    }
    const bool ai = a >= n, bi = b >= n;
    ProgOp op;
    op.a = a;
    op.b = b;
    op.type = (ai && bi) ? OP_SS : (ai ? OP_ST : OP_TT);
    op.node = e;
    const int k = stk[e] + (need[e] >> 8) - 1;
    if (k >= 0 && k < n - 1) prog[k] = op;
  }
  return max(1, need[root] & 255);
}

}  // namespace

// One wavefront per item; dynamic LDS topo_lds_bytes(max_n).
__global__ __launch_bounds__(kWave) void mc3_propose_topo(DeviceView v, MC3Dev m) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int it = blockIdx.x, lane = threadIdx.x;
  const uint64_t sweep = *m.counter;
  SlotMeta *slots = const_cast<SlotMeta *>(v.slots);
  const SlotMeta cs = slots[m.first_slot + it];
  const SlotMeta ps = slots[m.prop_first + it];
  const int n = cs.n, ne = 2 * n - 1;
  const LocusMeta lm = v.loci[cs.locus];
  const bool sw = lm.model == IMA_MODEL_SW;
  double *tm = reinterpret_cast<double *>(smem);
  int32_t *u0 = reinterpret_cast<int32_t *>(tm + ne);
  int32_t *u1 = u0 + ne, *dn = u1 + ne, *pop = dn + ne, *A = pop + ne, *need = A + ne,
          *stk = need + ne;
  __shared__ int s_root;
  for (int e = lane; e < ne; e += kWave) {
    tm[e] = v.time[cs.node_off + e];
    u0[e] = v.up0[cs.node_off + e];
    u1[e] = v.up1[cs.node_off + e];
    dn[e] = v.down[cs.node_off + e];
    pop[e] = (v.pop && m.prior >= 0) ? v.pop[cs.node_off + e] : 0;
    A[e] = v.A ? v.A[cs.node_off + e] : 0;
  }
  wave_lds_fence();
  if (lane == 0) {
    PhiloxStream r = topo_rng(m, it, sweep);
    const int P = m.prior >= 0 ? m.P : 1;
    const double *tv = m.prior >= 0 ? m.tv + (int64_t)(it / m.L) * mc3_tw(m.P) : nullptr;
    ModelView md{};
    if (m.prior >= 0) md = model_view(v, v.models[m.prior]);
    int root = cs.root;
    double hast = 0.0;
    bool ok = true;
    // pick the edge, its parent (freed) and sister
    int edge;
    do edge = randposint(r, ne);
    while (dn[edge] == -1);
    const int freed = dn[edge];
    const int oldsis = u0[freed] == edge ? u1[freed] : u0[freed];
    const int old_down_of_freed = dn[freed];
    const int copyA0 = A[edge], copyA1 = A[oldsis];
    const int hold_down_A = old_down_of_freed >= 0 ? A[old_down_of_freed] : 0;
    const double roottime0 = tm[u0[root]];
    const double sd0 = fmin(kSlideStdvMax, roottime0 / 3.0);
    double slide = sd0 * normal01(r);
    const double hold = slide;
    bool rootmove = false;
    double roottime = roottime0;
    // joinsisdown: the sister takes the freed node's place
    {
      const int down = freed;
      tm[oldsis] = tm[down];
      dn[oldsis] = dn[down];
      const int dd = dn[oldsis];
      if (dd != -1) {
        if (u0[dd] == down) u0[dd] = oldsis;
        else u1[dd] = oldsis;
      } else {
        rootmove = true;
        root = oldsis;
        dn[oldsis] = -1;
        tm[oldsis] = kTimeMax;
        roottime = upper_time(u0, tm, n, oldsis);
      }
    }
    // slider_nomigration: slide the edge's top along the remaining genealogy
    int sis = oldsis;
    double tp = tm[edge];
    int steps = 0;
    for (; steps < kMaxSlideSteps; ++steps) {
      if (slide < 0) {  // up, toward the tips
        slide = -slide;
        const double eup = upper_time(u0, tm, n, edge), sup = upper_time(u0, tm, n, sis);
        const double pj = (P > 1 && pop[edge] != pop[sis])
                              ? join_time(md, tv, P, pop[edge], pop[sis], eup, sup)
                              : 0.0;
        if (pj > eup && pj > sup) {
          if (slide < tp - pj) {
            tp -= slide;
            break;
          }
          slide -= tp - pj;  // reflect: continue down
          tp = pj;
        } else if (sup == 0.0 || eup >= sup) {
          if (slide < tp - eup) {
            tp -= slide;
            break;
          }
          slide -= tp - eup;
          tp = eup;
        } else {
          if (slide < tp - sup) {
            tp -= slide;
            break;
          }
          slide -= tp - sup;  // a node: up one of its two branches
          tp = sup;
          sis = (r.next32() & 1u) ? u0[sis] : u1[sis];
          slide = -slide;
        }
      } else {  // down, toward the root
        if (dn[sis] == -1 || tp + slide < tm[sis]) {
          tp += slide;
          if (tp >= kTimeMax) tp = kTimeMax;
          break;
        }
        slide -= tm[sis] - tp;
        tp = tm[sis];
        if (r.next32() & 1u) {
          sis = dn[sis];
        } else {
          const int d = dn[sis];
          sis = u0[d] == sis ? u1[d] : u0[d];
          slide = -slide;
        }
      }
    }
    if (steps == kMaxSlideSteps) ok = false;
    // splitsisdown: the freed node goes back at tp on the new sister
    {
      const int newsis = sis, down = freed;
      tm[edge] = tp;
      tm[down] = tm[newsis];
      tm[newsis] = tp;
      const int dd = dn[newsis];
      if (dd != -1) {
        if (u0[dd] == newsis) u0[dd] = down;
        else u1[dd] = down;
      } else {
        root = down;
        roottime = tp;
        rootmove = true;
      }
      dn[down] = dd;
      int nowpop = pop[newsis];
      if (P > 1) {
        const int j = period_of(tv, P, tp);
        while (md.pope[nowpop] != -1 && md.pope[nowpop] <= j) nowpop = md.popdown[nowpop];
      }
      pop[down] = nowpop;
      dn[newsis] = down;
      dn[edge] = down;
      u0[down] = newsis;
      u1[down] = edge;
    }
    if (rootmove) {
      const double sd1 = fmin(kSlideStdvMax, roottime / 3.0);
      hast += log_normprob(sd1, hold) - log_normprob(sd0, hold);
    }
    if (sw)
      hast += sw_allele(r, m, it, A, dn, edge, sis, freed, copyA0, copyA1, old_down_of_freed,
                        hold_down_A);
    SlotMeta &pm = slots[m.prop_first + it];
    pm.root = root;
    pm.nmig = 0;
    s_root = root;
    m.thast[it] = ok ? hast : -INFINITY;  // a runaway slide is rejected
  }
  wave_lds_fence();
  {
    ProgOp *prog = const_cast<ProgOp *>(v.prog) + ps.prog_off;
    const int nslots = wave_emit_program(n, ne, s_root, u0, u1, dn, need, stk, prog, lane);
    if (lane == 0) slots[m.prop_first + it].nslots = nslots;
  }
  const int64_t po = ps.node_off;
  int32_t *gu0 = const_cast<int32_t *>(v.up0), *gu1 = const_cast<int32_t *>(v.up1);
  int32_t *gdn = const_cast<int32_t *>(v.down);
  double *gtm = const_cast<double *>(v.time);
  for (int e = lane; e < ne; e += kWave) {
    gu0[po + e] = u0[e];
    gu1[po + e] = u1[e];
    gdn[po + e] = dn[e];
    gtm[po + e] = tm[e];
    if (v.pop) const_cast<int32_t *>(v.pop)[po + e] = pop[e];
    if (v.A) const_cast<int32_t *>(v.A)[po + e] = A[e];
    if (v.mig_first) const_cast<int32_t *>(v.mig_first)[po + e] = 0;
    if (v.mig_count) const_cast<int32_t *>(v.mig_count)[po + e] = 0;
  }
}

// ---- the move for models with migration --------------------------------
//
// updategenealogy with slider (update_gtree.cpp:256-372) and addmigration
// (:589-665): the edge's lower end slides freely (up to the edge's own
// daughters' node), joinsisdown and splitsisdown carry the sister's and
// the parent's migrations, and the edge's migrations -- and the new
// sister's when the edge now ends at the root -- are drawn again period by
// period (mwork_single_edge, mwork_two_edges, simmpath, getmprob,
// update_gtree_common.cpp:330-1583).  Lane 0 makes the move in LDS; each
// edge's migration list is a (first, count) range of an append-only LDS
// pool, so the old lists stay readable for the reverse path probability.

namespace {

constexpr double kMigCloseFrac = 0.9;  // MIGCLOSEFRAC

// edgemiginfo: one edge part's periods, times and counts, and its path
struct DEmi {
  int edgeid, pop, temppop, fpop, b, e, mpall, first, count, capn;
  double upt, dnt, mtall;
  double mtav[kMaxPops];
  int mp[kMaxPops];
};

struct MigCtx {
  ModelView md;
  int P, lastp, rootpop;
  const double *tv;
  double *pt;  // pool times
  int32_t *pp;  // pool destinations
  int pool_cap, pool_end;
  bool full;
  __device__ __forceinline__ double tvx(int k) const { return k < P - 1 ? tv[k] : kTimeMax; }
  __device__ __forceinline__ void settle(int *p, int period) const {
    while (md.pope[*p] <= period && md.pope[*p] != -1) *p = md.popdown[*p];
  }
  __device__ __forceinline__ int findperiod(double t) const {
    int k = 0;
    while (k < lastp && tv[k] <= t) k++;
    return k;
  }
  __device__ __forceinline__ int picktopop(PhiloxStream &r, int period, int nowpop, int notother) const {
    const int32_t *pl = md.plist + period * P;
    const int np = md.npresent[period];
    int q, tries = 0;
    do q = mig_randposint(r, np);
    while ((pl[q] == nowpop || pl[q] == notother) && ++tries < kMaxTries);
    return pl[q];
  }
};

__device__ __forceinline__ void emi_reset(DEmi &m) {
  m.edgeid = -1;
  m.b = m.e = -1;
  m.pop = m.temppop = m.fpop = -1;
  m.upt = m.dnt = -1.0;
  for (int k = 0; k < kMaxPops; ++k) {
    m.mtav[k] = 0.0;
    m.mp[k] = 0;
  }
  m.mtall = 0.0;
  m.mpall = 0;
  m.first = m.count = m.capn = 0;
}

// fillmiginfoperiods (:1101-1162)
__device__ __forceinline__ void emi_periods(const MigCtx &c, DEmi &m) {
  m.b = 0;
  while (m.upt > c.tvx(m.b)) m.b++;
  m.e = m.b;
  while (m.dnt > c.tvx(m.e)) m.e++;
  if (m.e == m.b) {
    m.mtav[m.b] = m.b == c.lastp ? 0.0 : m.dnt - m.upt;
  } else {
    m.mtav[m.b] = c.tvx(m.b) - m.upt;
    m.mtav[m.e] = m.e == c.lastp ? 0.0 : m.dnt - c.tvx(m.e - 1);
    for (int i = m.b + 1; i < m.e; ++i) m.mtav[i] = c.tvx(i) - c.tvx(i - 1);
  }
  m.mtall = m.b < c.lastp ? fmin(c.tvx(c.lastp - 1), m.dnt) - m.upt : 0.0;
}

// fillmiginfo (:1170-1249): an edge as it is, its list a range of the pool
__device__ __forceinline__ void emi_old(const MigCtx &c, DEmi &m, int edge, const int32_t *u0, const int32_t *dn,
                        const double *tm, const int32_t *pop, const int32_t *lf,
                        const int32_t *lc, int n) {
  emi_reset(m);
  m.edgeid = edge;
  m.pop = pop[edge];
  m.dnt = tm[edge];
  m.upt = edge < n ? 0.0 : tm[u0[edge]];
  m.fpop = pop[dn[edge]];
  emi_periods(c, m);
  m.first = lf[edge];
  m.count = lc[edge];
  int j = m.b;
  for (int i = 0; i < m.count; ++i) {
    while (c.pt[m.first + i] > c.tvx(j)) j++;
    m.mp[j]++;
    m.mpall++;
  }
}

// simmpath (:330-441): numm events appended to m's pool range; returns the
// population reached
__device__ __forceinline__ int emi_simmpath(PhiloxStream &r, MigCtx &c, DEmi &m, int period, int numm,
                            double timein, double upt, int pop, int constrainpop) {
#pragma clang fp contract(off)
  if (m.count + numm > m.capn) {
    c.full = true;
    numm = m.capn - m.count;
    if (numm <= 0) return pop;
  }
  double *t = c.pt + m.first + m.count;
  int32_t *q = c.pp + m.first + m.count;
  int tries = 0;
  bool dup;
  do {
    for (int i = 0; i < numm; ++i) t[i] = upt + r.uniform() * timein;
    for (int a = 1; a < numm; ++a) {
      const double x = t[a];
      int b = a - 1;
      while (b >= 0 && t[b] > x) {
        t[b + 1] = t[b];
        b--;
      }
      t[b + 1] = x;
    }
    dup = false;
    for (int i = 0; i + 1 < numm; ++i)
      if (t[i] == t[i + 1]) dup = true;
  } while (dup && ++tries < kMaxTries);
  int last = pop;
  for (int i = 0; i < numm; ++i) {
    int d;
    if (constrainpop >= 0 && i >= numm - 2)
      d = i == numm - 2 ? c.picktopop(r, period, last, constrainpop) : constrainpop;
    else
      d = c.picktopop(r, period, last, -1);
    q[i] = d;
    last = d;
  }
  m.count += numm;
  return last;
}

// mwork_single_edge (:1354-1429)
__device__ __forceinline__ int emi_single(PhiloxStream &r, MigCtx &c, DEmi &m, const DEmi &old, int lastmig) {
#pragma clang fp contract(off)
  if (lastmig < m.b) return 0;
  int mpall = 0, k;
  double tstart = m.upt;
  for (k = m.b; k <= lastmig; ++k) {
    c.settle(&m.temppop, k);
    const double rate = calcmrate(old.mp[k], old.mtav[k]) * m.mtav[k];
    int constrain = -1;
    if (k < m.e) {
      m.mp[k] = poisson_cond(r, rate, -1);
    } else if (c.P - k == 2) {
      m.mp[k] = poisson_cond(r, rate, m.temppop == m.fpop ? 0 : 1);
    } else {
      m.mp[k] = poisson_cond(r, rate, m.temppop == m.fpop ? 3 : 2);
      constrain = m.fpop;
    }
    if (m.mp[k] > 0) {
      m.temppop = emi_simmpath(r, c, m, k, m.mp[k], m.mtav[k], tstart, m.temppop, constrain);
      mpall += m.mp[k];
    }
    tstart = c.tvx(k);
  }
  if (m.mtav[k] > 0 && c.md.pope[m.temppop] == k) m.temppop = c.md.popdown[m.temppop];
  return mpall;
}

// mwork_two_edges (:1431-1583)
__device__ __forceinline__ void emi_two(PhiloxStream &r, MigCtx &c, DEmi &em, DEmi &sm, const DEmi &oe,
                        const DEmi &os, int lastmig) {
#pragma clang fp contract(off)
  DEmi *mm[2] = {&em, &sm};
  const DEmi *om[2] = {&oe, &os};
  double tstart[2] = {em.upt, sm.upt};
  const int b = min(em.b, sm.b);
  const int lastk = em.e == c.lastp ? lastmig : lastmig - 1;
  int k;
  for (k = b; k <= lastk; ++k)
    for (int ii = 0; ii < 2; ++ii) {
      DEmi &m = *mm[ii];
      if (m.b > k) continue;
      c.settle(&m.temppop, k);
      const double rate = calcmrate(om[ii]->mp[k], om[ii]->mtav[k]) * m.mtav[k];
      m.mp[k] = poisson_cond(r, rate, -1);
      if (m.mp[k] > 0) {
        m.temppop = emi_simmpath(r, c, m, k, m.mp[k], m.mtav[k], tstart[ii], m.temppop, -1);
        m.mpall += m.mp[k];
      }
      tstart[ii] = c.tvx(k);
    }
  if (k == c.lastp) {
    em.fpop = sm.fpop = c.rootpop;
    return;
  }
  for (int ii = 0; ii < 2; ++ii)
    if (c.md.pope[mm[ii]->temppop] == k) mm[ii]->temppop = c.md.popdown[mm[ii]->temppop];
  int f;
  if (em.temppop == sm.temppop)
    f = r.uniform() < kMigCloseFrac ? em.temppop : c.picktopop(r, k, em.temppop, -1);
  else if (c.P - k == 2)
    f = r.uniform() < 0.5 ? em.temppop : sm.temppop;
  else if (r.uniform() < kMigCloseFrac)
    f = r.uniform() < 0.5 ? em.temppop : sm.temppop;
  else
    f = c.picktopop(r, k, em.temppop, sm.temppop);
  em.fpop = sm.fpop = f;
  for (int ii = 0; ii < 2; ++ii) {
    DEmi &m = *mm[ii];
    const double rate = calcmrate(om[ii]->mp[k], om[ii]->mtav[k]) * m.mtav[k];
    int constrain = -1;
    if (c.P - k == 2) {
      m.mp[k] = poisson_cond(r, rate, m.temppop == m.fpop ? 0 : 1);
    } else {
      m.mp[k] = poisson_cond(r, rate, m.temppop == m.fpop ? 3 : 2);
      constrain = m.fpop;
    }
    if (m.mp[k] > 0) {
      m.temppop = emi_simmpath(r, c, m, k, m.mp[k], m.mtav[k], tstart[ii], m.temppop, constrain);
      m.mpall += m.mp[k];
    }
    if (m.mtav[k] > 0 && c.md.pope[m.temppop] == k) m.temppop = c.md.popdown[m.temppop];
  }
}

// one edge part of getmprob: its free periods, then its last period
__device__ __forceinline__ double emi_prob(const MigCtx &c, const DEmi &m, const DEmi &old, int lastmig) {
#pragma clang fp contract(off)
  double tempp = 0.0;
  int cm = 0;
  for (int k = m.b; k <= m.e; ++k)
    if (k < lastmig || (k == lastmig && m.e == c.lastp)) {
      const double rate = calcmrate(old.mp[k], old.mtav[k]) * m.mtav[k];
      tempp += m.mp[k] * log(rate / (m.mtav[k] * (c.P - (k + 1)))) - rate;
      cm += m.mp[k];
    }
  if (m.e < c.lastp) {
    const int e = m.e;
    const double rate = calcmrate(old.mp[e], old.mtav[e]) * m.mtav[e];
    if (e == c.lastp - 1) {
      tempp += (m.mp[e] & 1) ? m.mp[e] * log(rate / m.mtav[e]) - logsinh_(rate)
                             : m.mp[e] * log(rate / m.mtav[e]) - logcosh_(rate);
    } else {
      int p0 = cm == 0 ? m.pop : c.pp[m.first + cm - 1];
      while (c.md.pope[p0] <= e && c.md.pope[p0] != -1) p0 = c.md.popdown[p0];
      const int popc = c.P - e - 1;
      const double d = p0 == m.fpop ? log(1 - rate * exp(-rate)) : log(1 - exp(-rate));
      double nn;
      if (m.mp[e] == 0) {
        nn = -rate;
      } else if (m.mp[e] == 1) {
        nn = log(rate / m.mtav[e]) - rate;
      } else {
        const int l2 = m.mp[e] == 2 ? p0 : c.pp[m.first + m.mpall - 3];
        const double pathc = l2 == m.fpop ? -log((double)popc) : -log((double)popc - 1);
        nn = m.mp[e] * log(rate / m.mtav[e]) - rate + (2 - m.mp[e]) * log((double)popc) + pathc;
      }
      tempp += nn - d;
    }
  }
  return tempp;
}

// getmprob (:851-1099)
__device__ __forceinline__ double emi_getmprob(const MigCtx &c, const DEmi &em, const DEmi &sm, const DEmi &oe,
                               const DEmi &os) {
#pragma clang fp contract(off)
  const int lastmig = min(em.e, c.lastp - 1);
  if (sm.mtall <= 0) return emi_prob(c, em, oe, lastmig);
  double tempp = 0.0;
  if (em.mtall > 0 && em.e < c.lastp) {
    int pp2[2];
    const DEmi *mm[2] = {&em, &sm};
    for (int ii = 0; ii < 2; ++ii) {
      const DEmi &m = *mm[ii];
      pp2[ii] = m.pop;
      if (m.mpall > 0 && m.e > 0) {
        const double t = c.tvx(m.e - 1);
        int j = -1;
        while (j + 1 < m.count && c.pt[m.first + j + 1] < t) j++;
        if (j >= 0) pp2[ii] = c.pp[m.first + j];
      }
      if (m.e > 0)
        while (c.md.pope[pp2[ii]] <= m.e && c.md.pope[pp2[ii]] != -1)
          pp2[ii] = c.md.popdown[pp2[ii]];
    }
    if (pp2[0] == pp2[1])
      tempp = pp2[0] == em.fpop ? log(kMigCloseFrac)
                                : log((1.0 - kMigCloseFrac) / (double)(c.P - em.e - 1));
    else if (c.P - em.e == 2)
      tempp = log(0.5);
    else if (em.fpop == pp2[0] || em.fpop == pp2[1])
      tempp = log(0.5 * kMigCloseFrac);
    else
      tempp = log((1.0 - kMigCloseFrac) / (double)(c.P - em.e - 2));
  }
  if (em.mtall > 0) tempp += emi_prob(c, em, oe, lastmig);
  if (sm.mtall > 0) tempp += emi_prob(c, sm, os, lastmig);
  return tempp;
}

}  // namespace

// One wavefront per item; dynamic LDS topo_mig_lds_bytes(max_n, pool).
__global__ __launch_bounds__(kWave) void mc3_propose_topo_mig(DeviceView v, MC3Dev m) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned char smem[];
  const int it = blockIdx.x, lane = threadIdx.x;
  const uint64_t sweep = *m.counter;
  SlotMeta *slots = const_cast<SlotMeta *>(v.slots);
  const SlotMeta cs = slots[m.first_slot + it];
  const SlotMeta ps = slots[m.prop_first + it];
  const int n = cs.n, ne = 2 * n - 1, E = 2 * m.maxn - 1, POOL = m.mig_pool;
  const LocusMeta lm = v.loci[cs.locus];
  const bool sw = lm.model == IMA_MODEL_SW;
  double *tm = reinterpret_cast<double *>(smem);
  double *pt = tm + E;
  int32_t *u0 = reinterpret_cast<int32_t *>(pt + POOL);
  int32_t *u1 = u0 + E, *dn = u1 + E, *pop = dn + E, *A = pop + E, *need = A + E, *stk = need + E,
          *lf = stk + E, *lc = lf + E, *nf = lc + E, *pp = nf + E;
  DEmi *em = reinterpret_cast<DEmi *>(smem + ((((size_t)E * 8 + (size_t)POOL * 8 +
                                                (size_t)E * 10 * 4 + (size_t)POOL * 4) + 15) &
                                               ~(size_t)15));
  __shared__ int s_bad, s_root;
  // the population model's plist | npresent | popdown | pope and the
  // chain's split times, read by every step of the move
  __shared__ int32_t s_mi[kMaxPops * kMaxPops + 5 * kMaxPops];
  __shared__ double s_tv[kMaxPops];
  const ModelMeta mm = v.models[m.prior];
  const int PM = mm.npops;
  for (int x = lane; x < PM * PM + 5 * PM - 2; x += kWave) s_mi[x] = v.mint[mm.ioff + x];
  for (int x = lane; x < m.P - 1; x += kWave)
    s_tv[x] = m.tv[(int64_t)(it / m.L) * mc3_tw(m.P) + x];
  const int nm0 = min(cs.nmig, POOL);
  for (int e = lane; e < ne; e += kWave) {
    tm[e] = v.time[cs.node_off + e];
    u0[e] = v.up0[cs.node_off + e];
    u1[e] = v.up1[cs.node_off + e];
    dn[e] = v.down[cs.node_off + e];
    pop[e] = v.pop[cs.node_off + e];
    A[e] = v.A ? v.A[cs.node_off + e] : 0;
    lf[e] = v.mig_first[cs.node_off + e];
    lc[e] = v.mig_count[cs.node_off + e];
  }
  for (int x = lane; x < nm0; x += kWave) {
    pt[x] = v.mig_time[cs.mig_off + x];
    pp[x] = v.mig_to[cs.mig_off + x];
  }
  wave_lds_fence();
  if (lane == 0 && cs.nmig > POOL) {
    // more events than the pool holds (the room made at set-up bounds
    // them): the genealogy is proposed unchanged and rejected
    slots[m.prop_first + it].root = cs.root;
    s_root = cs.root;
    s_bad = 1;
    m.thast[it] = 0.0;
  } else if (lane == 0) {
    PhiloxStream r = topo_rng(m, it, sweep);
    MigCtx c;
    c.md = ModelView{};
    c.md.plist = s_mi;
    c.md.npresent = s_mi + PM * PM;
    c.md.popdown = c.md.npresent + PM;
    c.md.pope = c.md.popdown + 2 * PM - 1;
    c.P = m.P;
    c.lastp = m.P - 1;
    c.rootpop = 2 * m.P - 2;
    c.tv = s_tv;
    c.pt = pt;
    c.pp = pp;
    c.pool_cap = POOL;
    c.pool_end = nm0;
    c.full = false;
    DEmi &oe = em[0], &os = em[1], &nw = em[2], &ns = em[3];
    int root = cs.root;
    double hast = 0.0;
    bool ok = true;
    int edge;
    do edge = randposint(r, ne);
    while (dn[edge] == -1);
    const int freed = dn[edge];
    const int oldsis = u0[freed] == edge ? u1[freed] : u0[freed];
    const int old_down_of_freed = dn[freed];
    const int copyA0 = A[edge], copyA1 = A[oldsis];
    const int hold_down_A = old_down_of_freed >= 0 ? A[old_down_of_freed] : 0;
    emi_old(c, oe, edge, u0, dn, tm, pop, lf, lc, n);
    emi_reset(os);
    if (freed == root) emi_old(c, os, oldsis, u0, dn, tm, pop, lf, lc, n);
    lc[edge] = 0;
    double roottime = tm[u0[root]];
    const double sd0 = fmin(kSlideStdvMax, roottime / 3.0);
    const double hold = mig_normdev(r, 0.0, sd0);
    double dist = hold;
    bool rootmove = false;
    // joinsisdown: the sister's list then the freed edge's, as one new range
    {
      const int cs_ = lc[oldsis], cf = lc[freed];
      if (c.pool_end + cs_ + cf > POOL) {
        c.full = true;
      } else {
        const int f0 = c.pool_end;
        for (int i = 0; i < cs_; ++i) {
          pt[f0 + i] = pt[lf[oldsis] + i];
          pp[f0 + i] = pp[lf[oldsis] + i];
        }
        for (int i = 0; i < cf; ++i) {
          pt[f0 + cs_ + i] = pt[lf[freed] + i];
          pp[f0 + cs_ + i] = pp[lf[freed] + i];
        }
        lf[oldsis] = f0;
        lc[oldsis] = cs_ + cf;
        c.pool_end += cs_ + cf;
      }
      tm[oldsis] = tm[freed];
      dn[oldsis] = dn[freed];
      const int dd = dn[oldsis];
      if (dd != -1) {
        if (u0[dd] == freed) u0[dd] = oldsis;
        else u1[dd] = oldsis;
      } else {
        rootmove = true;
        root = oldsis;
        tm[oldsis] = kTimeMax;
        lc[oldsis] = 0;
        roottime = upper_time(u0, tm, n, oldsis);
      }
      lc[freed] = 0;
    }
    // slider: up to the edge's own daughters' node, down toward the root
    int sis = oldsis;
    double tp = tm[edge];
    const double uplimit = upper_time(u0, tm, n, edge);
    int steps = 0;
    for (; steps < kMaxSlideSteps && dist != 0.0; ++steps) {
      if (dist < 0) {
        const double d = -dist;
        const int su = u0[sis];
        if (su == -1 || uplimit >= tm[su]) {
          if (d < tp - uplimit) {
            tp -= d;
            dist = 0.0;
          } else {
            dist = d - (tp - uplimit);
            tp = uplimit;
          }
        } else if (d < tp - tm[su]) {
          tp -= d;
          dist = 0.0;
        } else {
          const double left = d - (tp - tm[su]);
          tp = tm[su];
          sis = r.uniform() < 0.5 ? u0[sis] : u1[sis];
          dist = -left;
        }
      } else {
        if (dn[sis] == -1 || tp + dist < tm[sis]) {
          tp += dist;
          if (tp >= kTimeMax) tp = kTimeMax;
          dist = 0.0;
        } else {
          const double left = dist - (tm[sis] - tp);
          tp = tm[sis];
          if (r.uniform() < 0.5) {
            sis = dn[sis];
            dist = left;
          } else {
            const int d = dn[sis];
            sis = u0[d] == sis ? u1[d] : u0[d];
            dist = -left;
          }
        }
      }
    }
    if (steps == kMaxSlideSteps) ok = false;
    // splitsisdown: the new sister's list split at tp
    {
      tm[edge] = tp;
      tm[freed] = tm[sis];
      tm[sis] = tp;
      const int dd = dn[sis];
      if (dd != -1) {
        if (u0[dd] == sis) u0[dd] = freed;
        else u1[dd] = freed;
      } else {
        root = freed;
        roottime = tp;
        rootmove = true;
      }
      dn[freed] = dd;
      int i = 0;
      while (i < lc[sis] && pt[lf[sis] + i] < tp) i++;
      int nowpop = i > 0 ? pp[lf[sis] + i - 1] : pop[sis];
      c.settle(&nowpop, c.findperiod(tp));
      pop[freed] = nowpop;
      if (dd != -1) {
        lf[freed] = lf[sis] + i;
        lc[freed] = lc[sis] - i;
      } else {
        lc[freed] = 0;
      }
      lc[sis] = i;
      dn[sis] = freed;
      dn[edge] = freed;
      u0[freed] = sis;
      u1[freed] = edge;
    }
    if (rootmove)
      hast += log_normprob(fmin(kSlideStdvMax, roottime / 3.0), hold) - log_normprob(sd0, hold);
    // addmigration: new paths in two pool ranges of equal room
    emi_reset(nw);
    emi_reset(ns);
    const int room = (POOL - c.pool_end) / 2;
    if (room <= 0) c.full = true;
    nw.edgeid = edge;
    nw.upt = upper_time(u0, tm, n, edge);
    nw.pop = nw.temppop = pop[edge];
    nw.dnt = tm[edge];
    emi_periods(c, nw);
    nw.first = c.pool_end;
    nw.capn = max(room, 0);
    const int lastmig = min(nw.e, c.lastp - 1);
    int newsis = -1;
    if (dn[edge] == root) {
      nw.fpop = -1;
      newsis = u0[root] == edge ? u1[root] : u0[root];
      ns.edgeid = newsis;
      ns.upt = upper_time(u0, tm, n, newsis);
      ns.fpop = -1;
      ns.pop = ns.temppop = pop[newsis];
      ns.dnt = tm[newsis];
      emi_periods(c, ns);
      ns.first = c.pool_end + max(room, 0);
      ns.capn = max(room, 0);
    } else {
      nw.fpop = pop[dn[edge]];
    }
    if (ns.edgeid == -1)
      nw.mpall = emi_single(r, c, nw, oe, lastmig);
    else if (nw.mtall <= 0)
      ns.mpall = emi_single(r, c, ns, os, lastmig);
    else
      emi_two(r, c, nw, ns, oe, os, lastmig);
    if (nw.fpop == -1) nw.fpop = ns.fpop = c.rootpop;
    if (dn[edge] == root) pop[root] = nw.fpop;
    if (!c.full) {  // a path cut short by the pool is rejected below
      const double denom = emi_getmprob(c, nw, ns, oe, os);
      const double num = emi_getmprob(c, oe, os, nw, ns);
      hast += num - denom;
    }
    // copynewmig_to_gtree
    lf[edge] = nw.first;
    lc[edge] = nw.count;
    if (ns.edgeid >= 0) {
      lf[newsis] = ns.first;
      lc[newsis] = ns.count;
      int p = ns.count > 0 ? pp[ns.first + ns.count - 1] : pop[newsis];
      c.settle(&p, c.findperiod(tm[newsis]));
      pop[dn[newsis]] = p;
    }
    if (sw)
      hast += sw_allele(r, m, it, A, dn, edge, sis, freed, copyA0, copyA1, old_down_of_freed,
                        hold_down_A);
    slots[m.prop_first + it].root = root;
    s_root = root;
    s_bad = (!ok || c.full) ? 1 : 0;
    m.thast[it] = hast;
  }
  wave_lds_fence();
  // the new lists, packed in order of edge, then time
  int carry = 0;
  for (int base = 0; base < ne; base += kWave) {
    const int e = base + lane;
    const int x = e < ne ? lc[e] : 0;
    int incl = x;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int y = __shfl_up(incl, off, kWave);
      if (lane >= off) incl += y;
    }
    if (e < ne) nf[e] = carry + incl - x;
    carry += __shfl(incl, kWave - 1, kWave);
  }
  // an accepted proposal's events are copied into the chain's slot
  const int limit = min(min(ps.mig_cap, cs.mig_cap), m.emax - (n - 1) - (m.P - 1));
  const bool bad = s_bad || carry > limit;
  const int64_t po = ps.node_off, pmo = ps.mig_off;
  int32_t *gu0 = const_cast<int32_t *>(v.up0), *gu1 = const_cast<int32_t *>(v.up1);
  int32_t *gdn = const_cast<int32_t *>(v.down);
  double *gtm = const_cast<double *>(v.time);
  int32_t *gme = const_cast<int32_t *>(v.mig_edge), *gmto = const_cast<int32_t *>(v.mig_to);
  double *gmt = const_cast<double *>(v.mig_time);
  for (int e = lane; e < ne; e += kWave) {
    gu0[po + e] = u0[e];
    gu1[po + e] = u1[e];
    gdn[po + e] = dn[e];
    gtm[po + e] = tm[e];
    const_cast<int32_t *>(v.pop)[po + e] = pop[e];
    if (v.A) const_cast<int32_t *>(v.A)[po + e] = A[e];
    const_cast<int32_t *>(v.mig_first)[po + e] = bad ? 0 : nf[e];
    const_cast<int32_t *>(v.mig_count)[po + e] = bad ? 0 : lc[e];
    if (!bad)
      for (int j = 0; j < lc[e]; ++j) {
        gme[pmo + nf[e] + j] = e;
        gmt[pmo + nf[e] + j] = pt[lf[e] + j];
        gmto[pmo + nf[e] + j] = pp[lf[e] + j];
      }
  }
  ProgOp *prog = const_cast<ProgOp *>(v.prog) + ps.prog_off;
  const int nslots = wave_emit_program(n, ne, s_root, u0, u1, dn, need, stk, prog, lane);
  if (lane == 0) {
    slots[m.prop_first + it].nslots = nslots;
    slots[m.prop_first + it].nmig = bad ? 0 : carry;
    if (bad) m.thast[it] = -INFINITY;  // no room for the new migrations: rejected
  }
}

size_t topo_mig_lds_bytes(int max_n, int pool) {
  const size_t E = 2 * (size_t)max_n - 1;
  return ((E * 8 + (size_t)pool * 8 + E * 10 * 4 + (size_t)pool * 4 + 15) & ~(size_t)15) +
         4 * sizeof(DEmi);
}

size_t topo_lds_bytes(int max_n) {
  const size_t ne = 2 * (size_t)max_n - 1;
  return ne * 8 + ne * 7 * 4;
}

hipError_t enqueue_mc3_propose_topo(const DeviceView &v, const MC3Dev &m, int max_n,
                                    hipStream_t st) {
  if (m.mig_pool > 0)
    hipLaunchKernelGGL(mc3_propose_topo_mig, dim3(m.C * m.L), dim3(kWave),
                       topo_mig_lds_bytes(max_n, m.mig_pool), st, v, m);
  else
    hipLaunchKernelGGL(mc3_propose_topo, dim3(m.C * m.L), dim3(kWave), topo_lds_bytes(max_n), st,
                       v, m);
  return hipGetLastError();
}

}  // namespace ima
