/*
 * ora_nw.c -- the migration part of changet_NW for one genealogy.
 *
 * TEST INFRASTRUCTURE ONLY (see ima_oracle.h): the tests use it as the
 * checker of the engine's device update and the oracle's own sweep calls
 * it; nothing in the product calls it.
 *
 * changet_NW (update_t_NW.cpp:920-1032, after Nielsen and Wakeley 2001)
 * moves split time `period` from oldt to newt without moving any node:
 * every edge that overlaps (tu, td] = (min, max of oldt and newt] gets its
 * migrations in that interval erased and re-simulated, so that its
 * population labels fit the new split time, and the proposal's Hastings
 * term is returned.  This file restates update_mig_tNW (:392-792) with
 * addmigration_NW (:238-333), simmpath_tNW (:144-236), getmprob_NW
 * (:335-372), nowedgepop (update_gtree_common.cpp:1638-1654), calcmrate
 * (:463-482), picktopop/picktopop2 (:1321-1351) and the conditioned
 * Poisson draws of poisson() (utilities.cpp:517-596), on the packed
 * layout of the engine's slots: per edge, the population at its tipward
 * end; migrations packed in order of edge, then time (destination
 * population in the coalescent direction).
 *
 * Randomness: every edge draws from its own Philox4x32-10 stream (key
 * derived from `key` and the edge, 64-bit counter from `ctr`), the first
 * half of the counter range for the choice of its final population, the
 * second for its migration count, times and destinations.  The engine's
 * device update uses the same streams, so both produce the same genealogy
 * from the same (key, ctr).  The reference draws from one sequential
 * generator; a random bit (bitran) is here U < 1/2; normdev's polar method
 * does not cache its second deviate; poisson's arithmetic is double
 * (the reference's long double).
 *
 * Quirks kept because they define the reference's proposal: the odd count
 * for small rates (< 0.25) is 1, 3 or else 5; rates of 100 or more use a
 * normal deviate with standard deviation equal to the rate; counts above
 * ADDMIGMAX = 1000 are cut to 999 (odd) or 1000; the population before
 * the second-to-last old migration of an edge is read from the migration
 * before it even when that one lies before the interval.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ima_oracle.h"

/* ---- Philox4x32-10 (Salmon et al. 2011), the engine's stream layout ---- */
typedef struct {
  uint32_t k0, k1;
  uint64_t ctr;
  uint32_t b[4];
  int left;
} ph_t;

static void ph_init(ph_t *r, uint64_t key, uint64_t ctr) {
  r->k0 = (uint32_t)key;
  r->k1 = (uint32_t)(key >> 32);
  r->ctr = ctr;
  r->left = 0;
}

static uint32_t ph_next32(ph_t *r) {
  if (r->left == 0) {
    uint32_t v[4] = {(uint32_t)r->ctr, (uint32_t)(r->ctr >> 32), 0x5EED0001u, 0u};
    uint32_t k0 = r->k0, k1 = r->k1;
    for (int i = 0; i < 10; i++) {
      const uint64_t p0 = (uint64_t)0xD2511F53u * v[0], p1 = (uint64_t)0xCD9E8D57u * v[2];
      const uint32_t n0 = (uint32_t)(p1 >> 32) ^ v[1] ^ k0, n1 = (uint32_t)p1;
      const uint32_t n2 = (uint32_t)(p0 >> 32) ^ v[3] ^ k1, n3 = (uint32_t)p0;
      v[0] = n0;
      v[1] = n1;
      v[2] = n2;
      v[3] = n3;
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    memcpy(r->b, v, sizeof v);
    r->left = 4;
    r->ctr++;
  }
  return r->b[4 - (r->left--)];
}

static double ph_unif(ph_t *r) {
  const uint64_t hi = ph_next32(r), lo = ph_next32(r);
  return (double)(((hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
}

/* ---- the reference's helpers ---------------------------------------- */
#define MIGSIMFRAC 0.999
#define NW_LOG2 0.69314718055994530941723212146
#define NW_LOG2HALF 0.34657359027997265470861606073
#define ADDMIGMAX 1000
#define NW_TIMEMAX 1000000.0
#define NW_EDGE_MIX 0xC2B2AE3D27D4EB4Full
#define NW_MAX_TRIES (1 << 20) /* bound of every rejection loop (the engine's) */

static int randposint(ph_t *r, int n) { return (int)floor(ph_unif(r) * n); }

/* calcmrate, update_gtree_common.cpp:463-482 */
static double calcmrate(int mc, double mt) {
  if (mt <= 0.0) return 1.0;
  if (mc == 0) return mt < 1 ? 0.1 : 0.1 / mt;
  return mt < 1 ? (double)mc : (double)mc / mt;
}

static double logcosh_(double x) { return x < 100 ? log(cosh(x)) : x - NW_LOG2; }
static double logsinh_(double x) { return x < 100 ? log(sinh(x)) : x - NW_LOG2; }

/* normdev (utilities.cpp:471-500): polar method, mean + stdev * deviate */
static double normdev(ph_t *r, double mean, double stdev) {
  double v1, v2, rsq;
  do {
    v1 = 2.0 * ph_unif(r) - 1.0;
    v2 = 2.0 * ph_unif(r) - 1.0;
    rsq = v1 * v1 + v2 * v2;
  } while (rsq >= 1.0 || rsq == 0.0);
  return v2 * sqrt(-2.0 * log(rsq) / rsq) * stdev + mean;
}

/* poisson (utilities.cpp:517-596): condition 0 even, 1 odd, 2 not 0,
 * 3 not 1 */
static int poisson_c(ph_t *r, double param, int cond) {
  if (param < 0.25 && cond == 1) {
    const double U = ph_unif(r);
    if (U < param / sinh(param)) return 1;
    if (U < param * (6 + param * param) / (6 * sinh(param))) return 3;
    return 5;
  }
  if (param < 0.25 && cond == 2) {
    double raised = exp(-param);
    double rcheck = raised = param * raised / (1 - raised);
    const double U = ph_unif(r);
    int i = 1;
    while (U > rcheck) {
      raised *= param / (i + 1);
      rcheck += raised;
      i++;
    }
    return i;
  }
  int i, stop, tries = 0;
  do {
    if (param >= 100.0) {
      const long v = (long)(normdev(r, param, param) + 0.5);
      i = v > 0 ? (int)(v > 2000000000L ? 2000000000L : v) : 0;
    } else {
      const double raised = exp(-param);
      double U = ph_unif(r);
      for (i = 0; U > raised && i < NW_MAX_TRIES; i++) U *= ph_unif(r);
    }
    stop = cond == 0 ? !(i & 1) : cond == 1 ? (i & 1) : cond == 2 ? i != 0 : i != 1;
    if (!stop && ++tries >= NW_MAX_TRIES) { /* a bound no valid rate reaches */
      i = cond == 1 || cond == 2 ? 1 : 0;
      stop = 1;
    }
  } while (!stop);
  if (i > ADDMIGMAX) i = cond == 1 ? ADDMIGMAX - 1 : ADDMIGMAX;
  return i;
}

/* getmprob_NW (update_t_NW.cpp:335-372) */
double ora_getmprob_nw(int period, int lastp, double mrate, double mtime, int mcount, int uppop,
                       int dpop, int cm2pop, int numpops) {
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

int ora_poisson_cond(double param, int cond, uint64_t key, uint64_t ctr) {
  ph_t r;
  ph_init(&r, key, ctr);
  return poisson_c(&r, param, cond);
}

typedef struct {
  int n, ne, root;
  const int *up0, *up1, *down, *pop;
  const double *time;
  const int *first, *cnt; /* per edge: its old migrations */
  const double *mt;
  const int *mp;
  const int *popdown, *pope;
  const double *tv; /* split times before the update */
} nwg_t;

static double uptime_of(const nwg_t *g, int e) { return e < g->n ? 0.0 : g->time[g->up0[e]]; }

/* nowedgepop: population of edge e at time t (at a split time, the one
 * above it) */
static int nowpop(const nwg_t *g, int e, double t) {
  int p = g->pop[e];
  for (int j = 0; j < g->cnt[e]; j++) {
    if (!(g->mt[g->first[e] + j] < t)) break;
    p = g->mp[g->first[e] + j];
  }
  while (p != -1 && g->pope[p] >= 1 && t > g->tv[g->pope[p] - 1]) p = g->popdown[p];
  return p;
}

int ora_update_mig_tnw(int n, const int *up0, const int *up1, const int *down,
                       const double *time, int root, int *pop, int *nmig, int cap,
                       int *mig_edge, double *mig_time, int *mig_to, int P, const int *plist,
                       const int *npresent, const int *popdown, const int *pope,
                       const int *join, const double *tv, int period, double newt,
                       uint64_t key, uint64_t ctr, double *migweight, int *updated) {
  const int ne = 2 * n - 1, lastp = P - 1;
  *migweight = 0.0;
  *updated = 0;
  if (P < 2 || period < 0 || period > P - 2 || n < 2) return -1;
  const double oldt = tv[period];
  const double tu = fmin(oldt, newt), td = fmax(oldt, newt);
  const double roottime = time[up0[root]];
  if (!(roottime > tu)) return 0; /* nothing of the genealogy lies above tu */
  const int period_a = newt < oldt ? period + 1 : period;
  const int period_b = newt < oldt ? period : period + 1;
  const int dropa = join[3 * period], dropb = join[3 * period + 1], addp = join[3 * period + 2];
  const int nm0 = *nmig;
  int *first = (int *)calloc(ne, sizeof(int)), *cnt = (int *)calloc(ne, sizeof(int));
  for (int j = nm0 - 1; j >= 0; j--) {
    first[mig_edge[j]] = j;
    cnt[mig_edge[j]]++;
  }
  double *omt = (double *)malloc(sizeof(double) * (nm0 + 1));
  int *omp = (int *)malloc(sizeof(int) * (nm0 + 1));
  memcpy(omt, mig_time, sizeof(double) * nm0);
  memcpy(omp, mig_to, sizeof(int) * nm0);
  int *opop = (int *)malloc(sizeof(int) * ne);
  memcpy(opop, pop, sizeof(int) * ne);
  nwg_t g = {n, ne, root, up0, up1, down, opop, time, first, cnt, omt, omp, popdown, pope, tv};

  int *own = (int *)malloc(sizeof(int) * ne), *sis = (int *)malloc(sizeof(int) * ne);
  int *db = (int *)malloc(sizeof(int) * ne), *da = (int *)malloc(sizeof(int) * ne);
  int *upa = (int *)malloc(sizeof(int) * ne), *upb = (int *)malloc(sizeof(int) * ne);
  int *mcount = (int *)calloc(ne, sizeof(int)), *mnew = (int *)calloc(ne, sizeof(int));
  int *cm2a = (int *)malloc(sizeof(int) * ne), *cm2b = (int *)malloc(sizeof(int) * ne);
  int *ncnt = (int *)calloc(ne, sizeof(int)), *npop = (int *)malloc(sizeof(int) * ne);
  double *lpf = (double *)calloc(ne, sizeof(double)), *lpfr = (double *)calloc(ne, sizeof(double));
  double *mtime = (double *)calloc(ne, sizeof(double)), *mrate = (double *)calloc(ne, sizeof(double));
  double *mrate_r = (double *)calloc(ne, sizeof(double));
  double *newmt = (double *)malloc(sizeof(double) * (cap + 1));
  int *newmp = (int *)malloc(sizeof(int) * (cap + 1)), *newme = (int *)malloc(sizeof(int) * (cap + 1));
  int rc = 0;
  memcpy(npop, opop, sizeof(int) * ne);
  for (int e = 0; e < ne; e++) {
    own[e] = sis[e] = -1;
    cm2a[e] = cm2b[e] = -1;
    upa[e] = upb[e] = db[e] = da[e] = -1;
  }
  /* the edges to update, in index order: an edge whose lower end lies
   * past td alone, two sister edges ending at a node inside (tu, td]
   * together (update_t_NW.cpp:424-609) */
  for (int e = 0; e < ne; e++) {
    const double ut = uptime_of(&g, e);
    if (!(time[e] > tu && ut <= td) || own[e] != -1) continue;
    ph_t r;
    ph_init(&r, key ^ (NW_EDGE_MIX * (uint64_t)(e + 1)), ctr);
    own[e] = e;
    if (time[e] > td) {
      db[e] = nowpop(&g, e, td);
      if (oldt < newt) {
        if (db[e] == addp) {
          lpfr[e] = 0.0;
          const int c0 = nowpop(&g, e, tu);
          if (ut < tu && (c0 == dropa || c0 == dropb)) {
            if (ph_unif(&r) < MIGSIMFRAC) {
              da[e] = c0;
              lpf[e] = log(MIGSIMFRAC);
            } else {
              da[e] = c0 == dropa ? dropb : dropa;
              lpf[e] = log(1.0 - MIGSIMFRAC);
            }
          } else {
            lpf[e] = -NW_LOG2;
            da[e] = ph_unif(&r) < 0.5 ? dropa : dropb;
          }
        } else {
          da[e] = db[e];
        }
      } else if (db[e] == dropa || db[e] == dropb) {
        da[e] = addp;
        const int c0 = nowpop(&g, e, tu);
        if (ut < tu && (c0 == dropa || c0 == dropb))
          lpfr[e] = c0 == db[e] ? log(MIGSIMFRAC) : log(1.0 - MIGSIMFRAC);
        else
          lpfr[e] = -NW_LOG2;
      } else {
        da[e] = db[e];
      }
    } else {
      const int d = down[e], s = up0[d] == e ? up1[d] : up0[d];
      sis[e] = s;
      own[s] = e;
      const double us = uptime_of(&g, s);
      db[e] = nowpop(&g, e, time[e]);
      if (oldt < newt) {
        if (db[e] == addp) {
          const int c0 = nowpop(&g, e, tu), c1 = nowpop(&g, s, tu);
          if (ut < tu && us < tu && c0 == c1 && (c0 == dropa || c0 == dropb)) {
            if (ph_unif(&r) < MIGSIMFRAC) {
              da[e] = c0;
              lpf[e] = log(MIGSIMFRAC) / 2.0;
            } else {
              da[e] = c0 == dropa ? dropb : dropa;
              lpf[e] = log(1.0 - MIGSIMFRAC) / 2.0;
            }
          } else {
            lpf[e] = -NW_LOG2HALF;
            da[e] = ph_unif(&r) < 0.5 ? dropa : dropb;
          }
        } else {
          da[e] = db[e];
        }
      } else if (db[e] == dropa || db[e] == dropb) {
        const int c0 = nowpop(&g, e, tu), c1 = nowpop(&g, s, tu);
        da[e] = addp;
        if (ut < tu && us < tu && c0 == c1 && (c0 == dropa || c0 == dropb))
          lpfr[e] = c0 == db[e] ? log(MIGSIMFRAC) / 2.0 : log(1.0 - MIGSIMFRAC) / 2.0;
        else
          lpfr[e] = -NW_LOG2HALF;
      } else {
        da[e] = db[e];
      }
    }
  }
  /* populations at the top of each updated edge's part of the interval,
   * before (upb) and after (upa) the update (:612-652) */
  for (int e = 0; e < ne; e++) {
    if (own[e] == -1) continue;
    const double ut = uptime_of(&g, e);
    if (ut < tu) {
      if (newt < oldt) {
        upb[e] = nowpop(&g, e, tu);
        upa[e] = (upb[e] == dropa || upb[e] == dropb) ? addp : upb[e];
      } else {
        upb[e] = nowpop(&g, e, tu * (1 + DBL_EPSILON));
        upa[e] = upb[e] == addp ? nowpop(&g, e, tu) : upb[e];
      }
    } else {
      const int o = e < n ? -1 : own[up0[e]];
      if (o < 0) { /* a node exactly at tu: the reference reads no update entry */
        rc = 2;
        goto done;
      }
      upb[e] = db[o];
      upa[e] = da[o];
      npop[e] = upa[e];
    }
  }
  /* counts, rates, new counts and paths (:655-731, addmigration_NW) */
  const int npopsa = P - period_a, npopsb = P - period_b;
  int tot = 0;
  for (int e = 0; e < ne; e++) {
    if (own[e] == -1 || e == root) {
      ncnt[e] = cnt[e];
      for (int j = 0; j < cnt[e]; j++) {
        if (tot >= cap) {
          rc = 1;
          goto done;
        }
        newme[tot] = e;
        newmt[tot] = omt[first[e] + j];
        newmp[tot] = omp[first[e] + j];
        tot++;
      }
      continue;
    }
    const int o = own[e];
    const double ut = uptime_of(&g, e);
    mtime[e] = fmin(td, time[e]) - fmax(tu, ut);
    if (!(mtime[e] > 0)) { /* the reference asserts a positive length */
      rc = 2;
      goto done;
    }
    if (npopsa >= 2) { /* both path ends must be populations of the period */
      int ua = 0, ud = 0;
      for (int q = 0; q < npresent[period_a]; q++) {
        ua = ua || plist[period_a * P + q] == upa[e];
        ud = ud || plist[period_a * P + q] == da[o];
      }
      if (!ua || !ud) {
        rc = 2;
        goto done;
      }
    }
    int k = 0, mi = 0, mstart = -1;
    while (k < cnt[e] && omt[first[e] + k] < td) {
      if (omt[first[e] + k] > tu) {
        if (mi == 0) mstart = k;
        mi++;
      }
      k++;
    }
    if (k >= 2 && mstart >= 0 && k - mstart >= 2)
      cm2b[e] = k == 2 ? upb[e] : omp[first[e] + k - 3];
    mcount[e] = mi;
    mrate[e] = period_a < lastp ? calcmrate(mi, mtime[e]) * mtime[e] : 0.0;
    ph_t r;
    ph_init(&r, key ^ (NW_EDGE_MIX * (uint64_t)(e + 1)), ctr + (1ull << 15));
    if (npopsa == 1)
      mnew[e] = 0;
    else if (npopsa == 2)
      mnew[e] = poisson_c(&r, mrate[e], upa[e] == da[o] ? 0 : 1);
    else
      mnew[e] = poisson_c(&r, mrate[e], upa[e] == da[o] ? 3 : 2);
    mrate_r[e] = period_b < lastp ? calcmrate(mnew[e], mtime[e]) * mtime[e] : 0.0;
    /* the new list: migrations before tu, the simulated ones, those after td */
    int before = 0;
    if (ut < tu)
      while (before < cnt[e] && omt[first[e] + before] < tu) before++;
    int after0 = before;
    while (after0 < cnt[e] && omt[first[e] + after0] < td) after0++;
    const int keep_after = !(newt < td && period_a == lastp);
    const int add = keep_after ? mnew[e] : 0;
    const int held = keep_after ? cnt[e] - after0 : 0;
    if (tot + before + add + held > cap) {
      rc = 1;
      goto done;
    }
    for (int j = 0; j < before; j++, tot++) {
      newme[tot] = e;
      newmt[tot] = omt[first[e] + j];
      newmp[tot] = omp[first[e] + j];
    }
    if (add > 0) {
      /* simmpath_tNW: times uniform on the interval part, sorted, drawn
       * again if two coincide; destinations picked from the period's
       * populations, the last two fixed so the path ends in da */
      const double upt = fmax(tu, ut), tin = mtime[e];
      double *t = newmt + tot;
      int dup, tries = 0;
      do {
        for (int j = 0; j < add; j++) t[j] = upt + ph_unif(&r) * tin;
        for (int a = 1; a < add; a++) { /* insertion sort */
          const double x = t[a];
          int b = a - 1;
          while (b >= 0 && t[b] > x) {
            t[b + 1] = t[b];
            b--;
          }
          t[b + 1] = x;
        }
        dup = 0;
        for (int j = 0; j + 1 < add; j++)
          if (t[j] == t[j + 1]) dup = 1;
      } while (dup && ++tries < NW_MAX_TRIES);
      const int *pl = plist + period_a * P, npl = npresent[period_a];
      int last = upa[e];
      if (add >= 2) {
        for (int j = 0; j < add - 2; j++) {
          int q, tries = 0;
          do q = randposint(&r, npl);
          while (pl[q] == last && ++tries < NW_MAX_TRIES);
          newmp[tot + j] = pl[q];
          last = pl[q];
        }
        cm2a[e] = last;
        int q, tries = 0;
        do q = randposint(&r, npl);
        while ((pl[q] == last || pl[q] == da[o]) && ++tries < NW_MAX_TRIES);
        newmp[tot + add - 2] = pl[q];
      }
      newmp[tot + add - 1] = da[o];
      for (int j = 0; j < add; j++) newme[tot + j] = e;
      tot += add;
    }
    for (int j = 0; j < held; j++, tot++) {
      newme[tot] = e;
      newmt[tot] = omt[first[e] + after0 + j];
      newmp[tot] = omp[first[e] + after0 + j];
    }
    ncnt[e] = before + add + held;
  }
  {
    /* the Hastings term, in the order of the reference's update list */
    double denom = 0.0, num = 0.0;
    for (int e = 0; e < ne; e++) {
      if (own[e] != e) continue;
      for (int i = 0; i < 2; i++) {
        const int x = i == 0 ? e : sis[e];
        if (x < 0 || x == root) continue;
        if (mrate[x] > 0)
          denom += lpf[e] + ora_getmprob_nw(period_a, lastp, mrate[x], mtime[x], mnew[x], upa[x], da[e],
                                     cm2a[x], npopsa);
        if (mrate_r[x] > 0)
          num += lpfr[e] + ora_getmprob_nw(period_b, lastp, mrate_r[x], mtime[x], mcount[x], upb[x],
                                    db[e], cm2b[x], npopsb);
      }
    }
    *migweight = num - denom;
  }
  memcpy(pop, npop, sizeof(int) * ne);
  memcpy(mig_edge, newme, sizeof(int) * tot);
  memcpy(mig_time, newmt, sizeof(double) * tot);
  memcpy(mig_to, newmp, sizeof(int) * tot);
  *nmig = tot;
  *updated = 1;
done:
  free(npop);
  free(first);
  free(cnt);
  free(omt);
  free(omp);
  free(opop);
  free(own);
  free(sis);
  free(db);
  free(da);
  free(upa);
  free(upb);
  free(mcount);
  free(mnew);
  free(cm2a);
  free(cm2b);
  free(ncnt);
  free(lpf);
  free(lpfr);
  free(mtime);
  free(mrate);
  free(mrate_r);
  free(newmt);
  free(newmp);
  free(newme);
  if (rc) *migweight = 0.0;
  return rc;
}
