// kern_nw.hip -- the migration update of changet_NW, one wavefront per
// genealogy.
//
// changet_NW (update_t_NW.cpp:920-1032, after Nielsen and Wakeley 2001)
// moves split time k from oldt to newt without moving any node.  Every
// edge that overlaps (tu, td] = (min, max of the two times] gets the
// migrations it has inside that interval erased and drawn again, ending in
// a population that fits the new split time, and the genealogy contributes
// the log ratio of the reverse and forward path probabilities to the
// move's Hastings term (update_mig_tNW :392-792, addmigration_NW :238-333,
// simmpath_tNW :144-236, getmprob_NW :335-372; nowedgepop,
// update_gtree_common.cpp:1638-1654; calcmrate :463-482; picktopop and
// picktopop2 :1321-1351; the conditioned Poisson draws of poisson(),
// utilities.cpp:517-596).
//
// The reference walks a list of updates in edge order, one edge at a time.
// Here the lanes of a wavefront take the edges: (A) each edge finds whether
// it overlaps the interval and, for the edge that owns an update (an edge
// alone, or the smaller of two sister edges ending inside the interval),
// the population at the bottom of the interval before and after; (B) the
// population at the top of each updated edge's part of the interval; (C)
// the old migration count there, the rate, the new count; then a prefix
// sum over the edges' new counts places every edge's new migration list,
// (D) each lane writes its edges' lists (migrations before the interval,
// the drawn path, those after it), populations and per-edge indexes, and
// the path probabilities; (E) the lanes sum the Hastings term.  Per-edge
// state lives in LDS, and so does the genealogy when there is room.  Every
// edge draws from its own Philox stream (first half of its counter range
// for the population choice, second half for the count, times and
// destinations), so the result does not depend on how edges map to lanes.
//
// The genealogy's old populations and migrations are first copied to the
// backup arrays (the update reads its old lists from there while it
// rewrites them in place, and a rejected move restores them).  A genealogy
// whose new migrations would exceed its slot's capacity or the event limit
// of the genealogy-prior kernels is left as it is, with status 2.
#include <float.h>

#include "ima_internal.h"
#include "mig_dev.h"
#include "philox.h"

namespace ima {
namespace {

constexpr double kMigSimFrac = 0.999;  // MIGSIMFRAC
constexpr double kLog2Half = 0.34657359027997265470861606073;
constexpr uint64_t kEdgeMix = 0xC2B2AE3D27D4EB4Full;
constexpr uint64_t kSecondHalf = 1ull << 15;

__device__ double getmprob_nw(int period, int lastp, double mrate, double mtime, int mcount,
                              int uppop, int dpop, int cm2pop, int numpops) {
#pragma clang fp contract(off)
  if (period == lastp) return 0;
  if (period == lastp - 1)
    return (mcount & 1) ? mcount * log(mrate / mtime) - logsinh_(mrate)
                        : mcount * log(mrate / mtime) - logcosh_(mrate);
  const double logs = uppop == dpop ? log(1 - mrate * exp(-mrate)) : log(1 - exp(-mrate));
  if (mcount == 0) return -mrate - logs;
  if (mcount == 1) return log(mrate / mtime) - mrate - logs;
  const double lognp = log((double)numpops - 1);
  const double logb = cm2pop == dpop ? -lognp : -log((double)numpops - 2);
  return mcount * log(mrate / mtime) + (2 - mcount) * lognp + logb - mrate - logs;
}

// One genealogy's read-only view for the update.
struct NwGene {
  int n, ne, root;
  const int32_t *up0, *up1, *down, *pop, *mf, *mc, *mto;
  const double *time, *mt;
  const int32_t *popdown, *pope;
  const double *tv;  // split times before the update
  __device__ double uptime(int e) const { return e < n ? 0.0 : time[up0[e]]; }
  // nowedgepop: population of edge e at time t (at a split time, the one
  // above it)
  __device__ int nowpop(int e, double t) const {
    int p = pop[e];
    const int f = mf[e], c = mc[e];
    for (int j = 0; j < c; ++j) {
      if (!(mt[f + j] < t)) break;
      p = mto[f + j];
    }
    while (p != -1 && pope[p] >= 1 && t > tv[pope[p] - 1]) p = popdown[p];
    return p;
  }
};

// LDS layout per wavefront: ints own, db, da, upa, upb, mcount, mnew,
// cm2a, cm2b, ncnt, nfirst; doubles lpf, lpfr, mtime, mrate, mrate_r, gmf,
// gmr -- each [E], E = 2 max_n - 1.
struct NwLds {
  int32_t *own, *db, *da, *upa, *upb, *mcount, *mnew, *cm2a, *cm2b, *ncnt, *nfirst;
  double *lpf, *lpfr, *mtime, *mrate, *mrate_r, *gmf, *gmr;
};
__host__ __device__ inline size_t nw_state_bytes(int E) { return (size_t)E * (7 * 8 + 11 * 4); }

__device__ inline NwLds nw_carve(char *base, int E) {
  NwLds s;
  double *d = reinterpret_cast<double *>(base);
  s.lpf = d;
  s.lpfr = d + E;
  s.mtime = d + 2 * E;
  s.mrate = d + 3 * E;
  s.mrate_r = d + 4 * E;
  s.gmf = d + 5 * E;
  s.gmr = d + 6 * E;
  int32_t *i = reinterpret_cast<int32_t *>(d + 7 * E);
  s.own = i;
  s.db = i + E;
  s.da = i + 2 * E;
  s.upa = i + 3 * E;
  s.upb = i + 4 * E;
  s.mcount = i + 5 * E;
  s.mnew = i + 6 * E;
  s.cm2a = i + 7 * E;
  s.cm2b = i + 8 * E;
  s.ncnt = i + 9 * E;
  s.nfirst = i + 10 * E;
  return s;
}

__global__ __launch_bounds__(kWave) void nw_update_kernel(DeviceView v, NwArgs a) {
#pragma clang fp contract(off)
  extern __shared__ char nw_lds[];
  __shared__ int err;
  const int it = blockIdx.x;
  const int lane = threadIdx.x;
  if (it >= a.count) return;
  const int s = a.slot ? a.slot[it] : a.first_slot + it;
  const SlotMeta sm = v.slots[s];
  const int row = a.row ? a.row[it] : 0;
  const int k = a.period[row];
  if (k < 0 || sm.n > a.max_n) {
    if (lane == 0) {
      a.status[it] = k < 0 ? 0 : 2;
      a.migweight[it] = 0.0;
    }
    return;
  }
  const ModelMeta mm = v.models[a.model];
  const ModelView md = model_view(v, mm);
  const int P = mm.npops, lastp = P - 1;
  const double *tvr = a.tv + (int64_t)row * a.tw;
  const double oldt = tvr[k], newt = a.newt[row];
  const double tu = fmin(oldt, newt), td = fmax(oldt, newt);
  const int64_t no = sm.node_off, mo = sm.mig_off;
  NwGene g;
  g.n = sm.n;
  g.ne = 2 * sm.n - 1;
  g.root = sm.root;
  g.up0 = v.up0 + no;
  g.up1 = v.up1 + no;
  g.down = v.down + no;
  g.pop = a.pop_bk + no;  // the old values, from the backup
  g.mf = a.mf_bk + no;
  g.mc = a.mc_bk + no;
  g.mto = a.mto_bk + mo;
  g.time = v.time + no;
  g.mt = a.mt_bk + mo;
  g.popdown = md.popdown;
  g.pope = md.pope;
  g.tv = tvr;
  const int ne = g.ne, root = g.root;
  if (!(g.time[g.up0[root]] > tu)) {  // nothing of the genealogy lies above tu
    if (lane == 0) {
      a.status[it] = 0;
      a.migweight[it] = 0.0;
    }
    return;
  }
  const int E = 2 * a.max_n - 1;
  const NwLds L = nw_carve(nw_lds, E);
  // with room in LDS the genealogy is staged there: the phases below walk
  // edges and migration lists through chains of dependent reads
  const bool stage = a.stage && sm.nmig <= a.max_cap;
  double *s_time = reinterpret_cast<double *>(nw_lds + ((nw_state_bytes(E) + 7) & ~(size_t)7));
  double *s_mt = s_time + E;  // old event times, then the new list's
  int32_t *s_up0 = reinterpret_cast<int32_t *>(s_mt + 2 * a.max_cap);
  int32_t *s_up1 = s_up0 + E, *s_down = s_up1 + E, *s_pop = s_down + E, *s_mf = s_pop + E;
  int32_t *s_mc = s_mf + E, *s_mto = s_mc + E;  // old destinations, then the new list's edges
                                                 // and destinations
  __shared__ int32_t s_pope[2 * kMaxPops - 1], s_popdown[2 * kMaxPops - 1];
  __shared__ double s_tv[kMaxPops];
  // backups: populations, per-edge indexes and migrations as they are
  {
    int32_t *pop = const_cast<int32_t *>(v.pop) + no;
    const int32_t *mf = v.mig_first + no, *mc = v.mig_count + no;
    for (int e = lane; e < ne; e += kWave) {
      const int32_t p = pop[e], f = mf[e], c = mc[e];
      a.pop_bk[no + e] = p;
      a.mf_bk[no + e] = f;
      a.mc_bk[no + e] = c;
      if (stage) {
        s_pop[e] = p;
        s_mf[e] = f;
        s_mc[e] = c;
        s_up0[e] = g.up0[e];
        s_up1[e] = g.up1[e];
        s_down[e] = g.down[e];
        s_time[e] = g.time[e];
      }
    }
    for (int x = lane; x < sm.nmig; x += kWave) {
      const int32_t ed = v.mig_edge[mo + x], to = v.mig_to[mo + x];
      const double t = v.mig_time[mo + x];
      a.me_bk[mo + x] = ed;
      a.mto_bk[mo + x] = to;
      a.mt_bk[mo + x] = t;
      if (stage) {
        s_mt[x] = t;
        s_mto[x] = to;
      }
    }
    for (int x = lane; x < 2 * P - 1; x += kWave) {
      s_pope[x] = md.pope[x];
      s_popdown[x] = md.popdown[x];
    }
    for (int x = lane; x < P - 1; x += kWave) s_tv[x] = tvr[x];
    if (lane == 0) {
      a.nmig_bk[it] = sm.nmig;
      err = 0;
    }
  }
  __syncthreads();
  g.popdown = s_popdown;
  g.pope = s_pope;
  g.tv = s_tv;
  if (stage) {
    g.up0 = s_up0;
    g.up1 = s_up1;
    g.down = s_down;
    g.pop = s_pop;
    g.mf = s_mf;
    g.mc = s_mc;
    g.mto = s_mto;
    g.time = s_time;
    g.mt = s_mt;
  }
  const int period_a = newt < oldt ? k + 1 : k, period_b = newt < oldt ? k : k + 1;
  const int dropa = md.join[3 * k], dropb = md.join[3 * k + 1], addp = md.join[3 * k + 2];
  const uint64_t key = a.seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)((a.keyid ? a.keyid[it] : it) + 1));
  const uint64_t ctr0 = a.counter ? ((((*a.counter) << 3) | 5ull) << 16) : a.ctr;
  // (A) edges to update and the owners' end populations (:424-609)
  for (int e = lane; e < ne; e += kWave) {
    const double ut = g.uptime(e);
    int o = -1;
    if (g.time[e] > tu && ut <= td) {
      if (g.time[e] > td) {
        o = e;
      } else {
        const int d = g.down[e], sis = g.up0[d] == e ? g.up1[d] : g.up0[d];
        o = min(e, sis);
      }
    }
    L.own[e] = o;
    L.db[e] = L.da[e] = -1;
    L.lpf[e] = L.lpfr[e] = 0.0;
    if (o != e) continue;
    PhiloxStream r(key ^ (kEdgeMix * (uint64_t)(e + 1)), ctr0);
    int db, da = -1;
    double lpf = 0.0, lpfr = 0.0;
    if (g.time[e] > td) {
      db = g.nowpop(e, td);
      if (oldt < newt) {
        if (db == addp) {
          const int c0 = g.nowpop(e, tu);
          if (ut < tu && (c0 == dropa || c0 == dropb)) {
            if (r.uniform() < kMigSimFrac) {
              da = c0;
              lpf = log(kMigSimFrac);
            } else {
              da = c0 == dropa ? dropb : dropa;
              lpf = log(1.0 - kMigSimFrac);
            }
          } else {
            lpf = -kLog2;
            da = r.uniform() < 0.5 ? dropa : dropb;
          }
        } else {
          da = db;
        }
      } else if (db == dropa || db == dropb) {
        da = addp;
        const int c0 = g.nowpop(e, tu);
        if (ut < tu && (c0 == dropa || c0 == dropb))
          lpfr = c0 == db ? log(kMigSimFrac) : log(1.0 - kMigSimFrac);
        else
          lpfr = -kLog2;
      } else {
        da = db;
      }
    } else {
      const int d = g.down[e], sis = g.up0[d] == e ? g.up1[d] : g.up0[d];
      const double us = g.uptime(sis);
      db = g.nowpop(e, g.time[e]);
      if (oldt < newt) {
        if (db == addp) {
          const int c0 = g.nowpop(e, tu), c1 = g.nowpop(sis, tu);
          if (ut < tu && us < tu && c0 == c1 && (c0 == dropa || c0 == dropb)) {
            if (r.uniform() < kMigSimFrac) {
              da = c0;
              lpf = log(kMigSimFrac) / 2.0;
            } else {
              da = c0 == dropa ? dropb : dropa;
              lpf = log(1.0 - kMigSimFrac) / 2.0;
            }
          } else {
            lpf = -kLog2Half;
            da = r.uniform() < 0.5 ? dropa : dropb;
          }
        } else {
          da = db;
        }
      } else if (db == dropa || db == dropb) {
        const int c0 = g.nowpop(e, tu), c1 = g.nowpop(sis, tu);
        da = addp;
        if (ut < tu && us < tu && c0 == c1 && (c0 == dropa || c0 == dropb))
          lpfr = c0 == db ? log(kMigSimFrac) / 2.0 : log(1.0 - kMigSimFrac) / 2.0;
        else
          lpfr = -kLog2Half;
      } else {
        da = db;
      }
    }
    L.db[e] = db;
    L.da[e] = da;
    L.lpf[e] = lpf;
    L.lpfr[e] = lpfr;
  }
  __syncthreads();
  // (B) populations at the top of each updated edge's part (:612-652)
  for (int e = lane; e < ne; e += kWave) {
    L.upa[e] = L.upb[e] = -1;
    if (L.own[e] == -1) continue;
    const double ut = g.uptime(e);
    if (ut < tu) {
      if (newt < oldt) {
        const int b = g.nowpop(e, tu);
        L.upb[e] = b;
        L.upa[e] = (b == dropa || b == dropb) ? addp : b;
      } else {
        const int b = g.nowpop(e, tu * (1 + DBL_EPSILON));
        L.upb[e] = b;
        L.upa[e] = b == addp ? g.nowpop(e, tu) : b;
      }
    } else {
      const int o = e < g.n ? -1 : L.own[g.up0[e]];
      if (o < 0) {
        err = 1;  // a node exactly at tu: the reference has no update entry for it
        continue;
      }
      L.upb[e] = L.db[o];
      L.upa[e] = L.da[o];
    }
  }
  __syncthreads();
  // (C) old counts, rates and new counts (:655-731)
  const int npopsa = P - period_a, npopsb = P - period_b;
  const bool keep_after = !(newt < td && period_a == lastp);
  for (int e = lane; e < ne; e += kWave) {
    const int c = g.mc[e], f = g.mf[e];
    L.cm2a[e] = L.cm2b[e] = -1;
    L.mnew[e] = L.mcount[e] = 0;
    L.mrate[e] = L.mrate_r[e] = 0.0;
    if (L.own[e] == -1 || e == root || err) {
      L.ncnt[e] = c;
      continue;
    }
    const int o = L.own[e];
    const double ut = g.uptime(e);
    const double mtime = fmin(td, g.time[e]) - fmax(tu, ut);
    if (!(mtime > 0)) err = 1;  // the reference asserts a positive length
    if (npopsa >= 2) {
      // both ends of the path must be populations of the period (a
      // genealogy whose labels fit its split times has them there)
      const int32_t *pl = md.plist + period_a * P;
      bool ua = false, ud = false;
      for (int q = 0; q < md.npresent[period_a]; ++q) {
        ua = ua || pl[q] == L.upa[e];
        ud = ud || pl[q] == L.da[o];
      }
      if (!ua || !ud) err = 1;
    }
    if (err) {
      L.ncnt[e] = c;
      continue;
    }
    int kk = 0, mi = 0, mstart = -1;
    while (kk < c && g.mt[f + kk] < td) {
      if (g.mt[f + kk] > tu) {
        if (mi == 0) mstart = kk;
        mi++;
      }
      kk++;
    }
    if (kk >= 2 && mstart >= 0 && kk - mstart >= 2) L.cm2b[e] = kk == 2 ? L.upb[e] : g.mto[f + kk - 3];
    const double mrate = period_a < lastp ? calcmrate(mi, mtime) * mtime : 0.0;
    PhiloxStream r(key ^ (kEdgeMix * (uint64_t)(e + 1)), ctr0 + kSecondHalf);
    int mnew = 0;
    if (npopsa == 2)
      mnew = poisson_cond(r, mrate, L.upa[e] == L.da[o] ? 0 : 1);
    else if (npopsa > 2)
      mnew = poisson_cond(r, mrate, L.upa[e] == L.da[o] ? 3 : 2);
    int before = 0;
    if (ut < tu)
      while (before < c && g.mt[f + before] < tu) before++;
    int after0 = before;
    while (after0 < c && g.mt[f + after0] < td) after0++;
    L.mtime[e] = mtime;
    L.mcount[e] = mi;
    L.mrate[e] = mrate;
    L.mnew[e] = mnew;
    L.mrate_r[e] = period_b < lastp ? calcmrate(mnew, mtime) * mtime : 0.0;
    L.ncnt[e] = before + (keep_after ? mnew + (c - after0) : 0);
  }
  __syncthreads();
  // new list positions: exclusive prefix sum of the new counts
  int carry = 0;
  for (int base = 0; base < ne; base += kWave) {
    const int e = base + lane;
    const int x = e < ne ? L.ncnt[e] : 0;
    int incl = x;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int y = __shfl_up(incl, off, kWave);
      if (lane >= off) incl += y;
    }
    if (e < ne) L.nfirst[e] = carry + incl - x;
    carry += __shfl(incl, kWave - 1, kWave);
  }
  const int total = carry;
  const int limit = min(sm.mig_cap, a.emax - (sm.n - 1) - (P - 1));
  __syncthreads();
  if (err || total > limit) {
    if (lane == 0) {
      a.status[it] = 2;
      a.migweight[it] = 0.0;
    }
    return;
  }
  // (D) the new lists, populations and indexes, and the path
  // probabilities (addmigration_NW, simmpath_tNW)
  {
    int32_t *pop = const_cast<int32_t *>(v.pop) + no;
    int32_t *mf = const_cast<int32_t *>(v.mig_first) + no;
    int32_t *mcnt = const_cast<int32_t *>(v.mig_count) + no;
    // with staging the new lists are built in LDS (the path's times are
    // sorted there) and copied out in one coalesced pass below
    double *n_mt = s_mt + a.max_cap;
    int32_t *n_me = s_mto + a.max_cap, *n_mto = n_me + a.max_cap;
    int32_t *me = stage ? n_me : const_cast<int32_t *>(v.mig_edge) + mo;
    int32_t *mto = stage ? n_mto : const_cast<int32_t *>(v.mig_to) + mo;
    double *mt = stage ? n_mt : const_cast<double *>(v.mig_time) + mo;
    const int32_t *pl = md.plist + period_a * P;
    const int npl = md.npresent[period_a];
    for (int e = lane; e < ne; e += kWave) {
      const int c = g.mc[e], f = g.mf[e], nf = L.nfirst[e];
      mf[e] = nf;
      mcnt[e] = L.ncnt[e];
      L.gmf[e] = L.gmr[e] = 0.0;
      if (L.own[e] != -1 && !(g.uptime(e) < tu)) pop[e] = L.upa[e];
      if (L.own[e] == -1 || e == root) {
        for (int j = 0; j < c; ++j) {
          me[nf + j] = e;
          mt[nf + j] = g.mt[f + j];
          mto[nf + j] = g.mto[f + j];
        }
        continue;
      }
      const int o = L.own[e];
      const double ut = g.uptime(e);
      int before = 0;
      if (ut < tu)
        while (before < c && g.mt[f + before] < tu) before++;
      int after0 = before;
      while (after0 < c && g.mt[f + after0] < td) after0++;
      int w = nf;
      for (int j = 0; j < before; ++j, ++w) {
        me[w] = e;
        mt[w] = g.mt[f + j];
        mto[w] = g.mto[f + j];
      }
      const int add = keep_after ? L.mnew[e] : 0;
      if (add > 0) {
        // the same stream as the count: replay the count, then the path
        PhiloxStream r(key ^ (kEdgeMix * (uint64_t)(e + 1)), ctr0 + kSecondHalf);
        if (npopsa == 2)
          (void)poisson_cond(r, L.mrate[e], L.upa[e] == L.da[o] ? 0 : 1);
        else
          (void)poisson_cond(r, L.mrate[e], L.upa[e] == L.da[o] ? 3 : 2);
        const double upt = fmax(tu, ut), tin = L.mtime[e];
        double *t = mt + w;
        bool dup;
        int tries = 0;
        do {
          for (int j = 0; j < add; ++j) t[j] = upt + r.uniform() * tin;
          for (int x = 1; x < add; ++x) {  // insertion sort
            const double y = t[x];
            int b = x - 1;
            while (b >= 0 && t[b] > y) {
              t[b + 1] = t[b];
              b--;
            }
            t[b + 1] = y;
          }
          dup = false;
          for (int j = 0; j + 1 < add; ++j)
            if (t[j] == t[j + 1]) dup = true;
        } while (dup && ++tries < kMaxTries);
        int last = L.upa[e];
        if (add >= 2) {
          for (int j = 0; j < add - 2; ++j) {
            int q, tries = 0;
            do q = mig_randposint(r, npl);
            while (pl[q] == last && ++tries < kMaxTries);
            mto[w + j] = pl[q];
            last = pl[q];
          }
          L.cm2a[e] = last;
          int q, tries = 0;
          do q = mig_randposint(r, npl);
          while ((pl[q] == last || pl[q] == L.da[o]) && ++tries < kMaxTries);
          mto[w + add - 2] = pl[q];
        }
        mto[w + add - 1] = L.da[o];
        for (int j = 0; j < add; ++j) me[w + j] = e;
        w += add;
      }
      if (keep_after)
        for (int j = after0; j < c; ++j, ++w) {
          me[w] = e;
          mt[w] = g.mt[f + j];
          mto[w] = g.mto[f + j];
        }
      // this edge's terms of the forward and reverse path probabilities
      if (L.mrate[e] > 0)
        L.gmf[e] = L.lpf[o] + getmprob_nw(period_a, lastp, L.mrate[e], L.mtime[e], L.mnew[e],
                                          L.upa[e], L.da[o], L.cm2a[e], npopsa);
      if (L.mrate_r[e] > 0)
        L.gmr[e] = L.lpfr[o] + getmprob_nw(period_b, lastp, L.mrate_r[e], L.mtime[e],
                                           L.mcount[e], L.upb[e], L.db[o], L.cm2b[e], npopsb);
    }
  }
  if (stage) {
    __syncthreads();  // the new lists are complete in LDS
    int32_t *me = const_cast<int32_t *>(v.mig_edge) + mo;
    int32_t *mto = const_cast<int32_t *>(v.mig_to) + mo;
    double *mt = const_cast<double *>(v.mig_time) + mo;
    const double *n_mt = s_mt + a.max_cap;
    const int32_t *n_me = s_mto + a.max_cap, *n_mto = n_me + a.max_cap;
    for (int x = lane; x < total; x += kWave) {
      me[x] = n_me[x];
      mt[x] = n_mt[x];
      mto[x] = n_mto[x];
    }
  }
  // (E) the Hastings term: each lane sums its edges' terms, then the lanes
  // in a fixed tree order (the reference adds them edge by edge; the sums
  // agree to rounding)
  double denom = 0.0, num = 0.0;
  for (int e = lane; e < ne; e += kWave) {
    denom += L.gmf[e];
    num += L.gmr[e];
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    denom += __shfl_xor(denom, off, kWave);
    num += __shfl_xor(num, off, kWave);
  }
  if (lane == 0) {
    a.migweight[it] = num - denom;
    a.status[it] = 1;
    const_cast<SlotMeta *>(v.slots)[s].nmig = total;
  }
}

// Restores the genealogies a rejected move updated (status 1): one
// wavefront per genealogy.
__global__ __launch_bounds__(kWave) void nw_restore_kernel(DeviceView v, NwArgs a,
                                                           const int32_t *decision) {
  const int it = blockIdx.x, lane = threadIdx.x;
  if (it >= a.count || a.status[it] != 1) return;
  const int row = a.row ? a.row[it] : 0;
  if (decision && decision[row]) return;
  const int s = a.slot ? a.slot[it] : a.first_slot + it;
  const SlotMeta sm = v.slots[s];
  const int64_t no = sm.node_off, mo = sm.mig_off;
  const int ne = 2 * sm.n - 1, nm = a.nmig_bk[it];
  int32_t *pop = const_cast<int32_t *>(v.pop) + no;
  int32_t *mf = const_cast<int32_t *>(v.mig_first) + no;
  int32_t *mc = const_cast<int32_t *>(v.mig_count) + no;
  for (int e = lane; e < ne; e += kWave) {
    pop[e] = a.pop_bk[no + e];
    mf[e] = a.mf_bk[no + e];
    mc[e] = a.mc_bk[no + e];
  }
  int32_t *me = const_cast<int32_t *>(v.mig_edge) + mo;
  int32_t *mto = const_cast<int32_t *>(v.mig_to) + mo;
  double *mt = const_cast<double *>(v.mig_time) + mo;
  for (int x = lane; x < nm; x += kWave) {
    me[x] = a.me_bk[mo + x];
    mto[x] = a.mto_bk[mo + x];
    mt[x] = a.mt_bk[mo + x];
  }
  if (lane == 0) const_cast<SlotMeta *>(v.slots)[s].nmig = nm;
}

__global__ __launch_bounds__(kWave) void move_events_kernel(int n, const int64_t *from,
                                                            const int64_t *to, const int32_t *cnt,
                                                            int32_t *me, int32_t *mto, double *mt) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const int64_t f = from[i], t = to[i];
  for (int x = threadIdx.x; x < cnt[i]; x += kWave) {
    me[t + x] = me[f + x];
    mto[t + x] = mto[f + x];
    mt[t + x] = mt[f + x];
  }
}

}  // namespace

hipError_t launch_move_events(int n, const int64_t *from, const int64_t *to, const int32_t *cnt,
                              int32_t *me, int32_t *mto, double *mt, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(move_events_kernel, dim3(n), dim3(kWave), 0, st, n, from, to, cnt, me, mto,
                     mt);
  return hipGetLastError();
}

size_t nw_lds_bytes(int max_n, int max_cap, int stage) {
  const int E = 2 * max_n - 1;
  const size_t base = nw_state_bytes(E);
  if (!stage) return base;
  // genealogy (edges), old and new event lists
  return ((base + 7) & ~(size_t)7) + (size_t)E * 8 + (size_t)max_cap * 16 + (size_t)E * 6 * 4 +
         (size_t)max_cap * 12;
}

hipError_t launch_nw_update(const DeviceView &v, const NwArgs &a, hipStream_t st) {
  if (a.count <= 0) return hipSuccess;
  hipLaunchKernelGGL(nw_update_kernel, dim3(a.count), dim3(kWave),
                     nw_lds_bytes(a.max_n, a.max_cap, a.stage), st, v, a);
  return hipGetLastError();
}

hipError_t launch_nw_restore(const DeviceView &v, const NwArgs &a, const int32_t *decision,
                             hipStream_t st) {
  if (a.count <= 0) return hipSuccess;
  hipLaunchKernelGGL(nw_restore_kernel, dim3(a.count), dim3(kWave), 0, st, v, a, decision);
  return hipGetLastError();
}

}  // namespace ima
