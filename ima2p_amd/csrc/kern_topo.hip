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
    if (sw) {
      // the freed node's allele (finishSWupdateA's proposal and its
      // Hastings term; the likelihood is scored in full afterwards)
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
      hast += (abs(dA) * log(1.0 - gold) + log(gold)) - (abs(dA) * log(1.0 - gnew) + log(gnew));
    }
    // the postorder program: Strahler need bottom-up, then the daughter
    // with the larger need first
    for (int e = 0; e < ne; ++e) need[e] = e < n ? 0 : -1;
    int sp = 0;
    stk[sp++] = root;
    while (sp > 0) {
      const int e = stk[sp - 1];
      if (need[e] >= 0) {
        --sp;
        continue;
      }
      const int a = u0[e], b = u1[e];
      if (need[a] >= 0 && need[b] >= 0) {
        need[e] = need[a] == need[b] ? need[a] + 1 : max(need[a], need[b]);
        --sp;
      } else {
        if (need[a] < 0) stk[sp++] = a;
        if (need[b] < 0) stk[sp++] = b;
      }
    }
    ProgOp *prog = const_cast<ProgOp *>(v.prog) + ps.prog_off;
    int nop = 0;
    sp = 0;
    stk[sp++] = root;
    // entries: e (descend) or ~e (emit e's operation)
    while (sp > 0) {
      const int x = stk[--sp];
      const int e = x >= 0 ? x : ~x;
      int a = u0[e], b = u1[e];
      if (need[a] < need[b]) {
        const int t = a;
        a = b;
        b = t;
      }
      const bool ai = a >= n, bi = b >= n;
      if (x >= 0) {
        stk[sp++] = ~e;
        if (bi) stk[sp++] = b;
        if (ai) stk[sp++] = a;
      } else {
        ProgOp op;
        op.a = a;
        op.b = b;
        op.type = (ai && bi) ? OP_SS : (ai ? OP_ST : OP_TT);
        op.node = e;
        prog[nop++] = op;
      }
    }
    SlotMeta &pm = slots[m.prop_first + it];
    pm.root = root;
    pm.nslots = max(1, need[root]);
    pm.nmig = 0;
    m.thast[it] = ok ? hast : -INFINITY;  // a runaway slide is rejected
  }
  wave_lds_fence();
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

size_t topo_lds_bytes(int max_n) {
  const size_t ne = 2 * (size_t)max_n - 1;
  return ne * 8 + ne * 7 * 4;
}

hipError_t enqueue_mc3_propose_topo(const DeviceView &v, const MC3Dev &m, int max_n,
                                    hipStream_t st) {
  hipLaunchKernelGGL(mc3_propose_topo, dim3(m.C * m.L), dim3(kWave), topo_lds_bytes(max_n), st, v,
                     m);
  return hipGetLastError();
}

}  // namespace ima
