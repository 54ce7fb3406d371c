/*
 * ora_mc3.c -- the CPU baseline's Metropolis-coupled sweep, in C.
 *
 * TEST INFRASTRUCTURE ONLY (see ima_oracle.h): the bench's cpu_baseline
 * leg times this loop on the host cores; nothing in the product calls it.
 *
 * The control structure is the reference's (qupdate, ima_main_mpi.cpp:
 * 1788-2047; updategenealogy's scoring, update_gtree.cpp:827-961): chains
 * in turn, each locus proposed, scored -- incremental likelihoodHKY on the
 * edges that moved, treeweight, the all-loci weight sums updated as
 * sum_subtract_treeinfo does (ginfo.cpp:249-285), integrals recomputed for
 * parameters whose (count, weight) changed (update_gtree_common.cpp:
 * 1997-2006) -- and accepted with beta (dprobg + dpdg) before the next; on
 * alternate sweeps every genealogy of a chain, its migrations and split
 * times are rescaled and fully re-scored; a swap closes the sweep
 * (swapchains.cpp:527-662).  The proposals are the device sweep's
 * (kern_mc3.hip): a reflected normal slide of one coalescence inside its
 * period and between the migrations around it, and a global rescale by
 * exp(d (U - 1/2)) -- or, with ry1, the reference's rubber-band update of
 * one split time (changet_RY1, update_t_RY.cpp:223-517).  With topo the
 * genealogy update is instead the reference's own move: slider_nomigration
 * / joinsisdown / splitsisdown for models without migration, slider with
 * addmigration's new migration paths (ora_mig.c) for models with it.  With
 * nw, time updates are changet_NW (ora_nw.c) with probability 1/2 (nw 1)
 * or always (nw 2).  Randomness:
 * xoshiro256** seeded by splitmix64.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ima_oracle.h"

typedef struct {
  uint64_t s[4];
} rng_t;

static uint64_t splitmix(uint64_t *x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t xoshiro(rng_t *r) {
  uint64_t *s = r->s;
  const uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return res;
}
static double unif(rng_t *r) { return (double)(xoshiro(r) >> 11) * (1.0 / 9007199254740992.0); }
static double norm01(rng_t *r) {
  const double u1 = 1.0 - unif(r), u2 = unif(r);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

typedef struct { /* one genealogy of one chain */
  int n, root;
  int *up0, *up1, *down, *pop;
  double *time;
  int nmig, mig_cap, *mig_edge, *mig_to;
  double *mig_time;
  ora_hky_state st;
  double u, kappa; /* mutation-rate scalar and kappa of this chain's locus */
  double pdg;
  double *rec; /* weight record (layout of the device's records) */
} gene_t;

typedef struct {
  /* loci */
  int L;
  const int *S, *totsites, *seq_off, *mult_off, *samppop;
  const int *seq, *mult;
  const double *pi;
  /* model */
  int P, npsize, nmig;
  const int *plist, *npresent, *popdown, *pope, *join;
  const int *theta_off, *theta_wp, *mig_off, *mig_wp;
  const double *qmax, *mmax;
  double tmax, scale_width;
  int ry1;
  int topo; /* the genealogy update is the reference's slider move */
  int nw;   /* 1: time updates pick RY1 or NW; 2: NW alone, no genealogy update */
} prob_t;

static int rec_len(int P) { return 3 * P * P + 2 * P * P * P + 1; }

static ora_tree tree_of(const gene_t *g) {
  ora_tree t = {g->n, g->up0, g->up1, g->down, g->time, g->root};
  return t;
}

static int weights(const prob_t *pb, int l, const gene_t *g, const double *tv, double *rec) {
  const int P = pb->P, P2 = P * P, P3 = P2 * P;
  int cc[100], mc[1000];
  double fc[100], hcc[100], fm[1000], len;
  ora_tree t = tree_of(g);
  const int st = ora_treeweight(&t, g->pop, g->nmig, g->mig_edge, g->mig_time, g->mig_to, P, P,
                                pb->plist, pb->npresent, pb->popdown, pb->pope, pb->join, tv,
                                pb->samppop + l * P, 1.0, cc, fc, hcc, mc, fm, &len);
  for (int i = 0; i < P2; i++) {
    rec[i] = cc[i];
    rec[P2 + i] = fc[i];
    rec[2 * P2 + i] = hcc[i];
  }
  for (int i = 0; i < P3; i++) {
    rec[3 * P2 + i] = mc[i];
    rec[3 * P2 + P3 + i] = fm[i];
  }
  rec[3 * P2 + 2 * P3] = len;
  return st;
}

/* integrals of every parameter at the sums `all`; terms whose (count,
 * weight) equal the cached ones are reused */
static double prior_terms(const prob_t *pb, const double *all, const int *qc, const double *qf,
                          const double *qt, int *nc, double *nf, double *nt) {
  const int P = pb->P, P2 = P * P, P3 = P2 * P;
  double q = 0.0, m = 0.0;
  for (int t = 0; t < pb->npsize + pb->nmig; t++) {
    int c = 0;
    double f = 0.0, h = 0.0;
    if (t < pb->npsize) {
      for (int w = pb->theta_off[t]; w < pb->theta_off[t + 1]; w++) {
        const int k = pb->theta_wp[2 * w], i = pb->theta_wp[2 * w + 1];
        c += (int)all[k * P + i];
        f += all[P2 + k * P + i];
        h += all[2 * P2 + k * P + i];
      }
    } else {
      const int mm = t - pb->npsize;
      for (int w = pb->mig_off[mm]; w < pb->mig_off[mm + 1]; w++) {
        const int k = pb->mig_wp[3 * w], i = pb->mig_wp[3 * w + 1], j = pb->mig_wp[3 * w + 2];
        c += (int)all[3 * P2 + (k * P + i) * P + j];
        f += all[3 * P2 + P3 + (k * P + i) * P + j];
      }
    }
    nc[t] = c;
    nf[t] = f;
    if (qc && qc[t] == c && qf[t] == f)
      nt[t] = qt[t];
    else if (t < pb->npsize)
      nt[t] = ora_integrate_coalescent_term(c, f, h, pb->qmax[t], 0.0);
    else
      nt[t] = ora_integrate_migration_term(c, f, pb->mmax[t - pb->npsize], 0.0);
    if (t < pb->npsize)
      q += nt[t];
    else
      m += nt[t];
  }
  return q + m;
}

static double hky(const prob_t *pb, int l, gene_t *g, int e) {
  ora_tree t = tree_of(g);
  return ora_likelihoodHKY(&t, pb->S[l], pb->seq + pb->seq_off[l], pb->mult + pb->mult_off[l],
                           pb->totsites[l], pb->pi + 4 * l, g->u, g->kappa, e, e, e, e,
                           &g->st);
}

static void hky_settle(gene_t *g, int accept) {
  ora_tree t = tree_of(g);
  if (accept) {
    ora_copyfraclike(&t, &g->st);
    ora_storescalefactors(&t, &g->st);
  } else {
    ora_restorescalefactors(&t, &g->st);
  }
}

typedef struct {
  gene_t *g;  /* [L] */
  double *tv; /* [P-1] */
  double *all, *hold, *nrec;
  double *tback, *mback; /* node and migration times before a rescale */
  int *qc, *nc;
  double *qf, *qt, *nf, *nt;
  double probg, beta;
} chain_t;

static void slide(const prob_t *pb, chain_t *ch, int l, rng_t *r, long *acc) {
  gene_t *g = &ch->g[l];
  const int n = g->n, R = rec_len(pb->P);
  int e = n + (int)(unif(r) * (n - 1));
  if (e > 2 * n - 2) e = 2 * n - 2;
  const int a = g->up0[e], b = g->up1[e];
  const double T = g->time[a];
  double lo = fmax(a >= n ? g->time[g->up0[a]] : 0.0, b >= n ? g->time[g->up0[b]] : 0.0);
  double hi = g->down[e] == -1 ? INFINITY : g->time[e];
  int k = 0;
  while (k < pb->P - 1 && ch->tv[k] <= T) k++;
  if (k > 0) lo = fmax(lo, ch->tv[k - 1]);
  if (k < pb->P - 1) hi = fmin(hi, ch->tv[k]);
  for (int j = 0; j < g->nmig; j++) {
    if (g->mig_edge[j] == a || g->mig_edge[j] == b) lo = fmax(lo, g->mig_time[j]);
    else if (g->mig_edge[j] == e) hi = fmin(hi, g->mig_time[j]);
  }
  const int open = isinf(hi);
  const double sigma = open ? 0.5 * fmax(lo, 0.05) : 0.5 * (hi - lo);
  double Tn = T + sigma * norm01(r);
  if (open) {
    Tn = lo + fabs(Tn - lo);
  } else {
    const double w = hi - lo;
    double y = fmod(Tn - lo, 2.0 * w);
    if (y < 0) y += 2.0 * w;
    if (y > w) y = 2.0 * w - y;
    Tn = lo + y;
  }
  if (!(Tn > lo) || !(Tn < hi)) Tn = T;
  g->time[a] = g->time[b] = Tn;
  const double np = hky(pb, l, g, e);
  double *nrec = ch->nrec;
  weights(pb, l, g, ch->tv, nrec);
  const int P2 = pb->P * pb->P, P3 = P2 * pb->P;
  for (int i = 0; i < R; i++) {
    const double x = ch->all[i], o = g->rec[i], w = nrec[i];
    ch->hold[i] = x;
    if ((i >= P2 && i < 2 * P2) || (i >= 3 * P2 + P3 && i < 3 * P2 + 2 * P3))
      ch->all[i] = fmax(0.0, (x - o) + w);
    else if (i < P2 || (i >= 3 * P2 && i < 3 * P2 + P3))
      ch->all[i] = x + (w - o);
    else
      ch->all[i] = (x - o) + w;
  }
  const double pn = prior_terms(pb, ch->all, ch->qc, ch->qf, ch->qt, ch->nc, ch->nf, ch->nt);
  const double U = unif(r);
  const int ok = U < fmin(1.0, exp(ch->beta * ((pn - ch->probg) + (np - g->pdg))));
  if (ok) {
    g->pdg = np;
    ch->probg = pn;
    memcpy(g->rec, nrec, sizeof(double) * R);
    const int np_ = pb->npsize + pb->nmig;
    memcpy(ch->qc, ch->nc, sizeof(int) * np_);
    memcpy(ch->qf, ch->nf, sizeof(double) * np_);
    memcpy(ch->qt, ch->nt, sizeof(double) * np_);
    (*acc)++;
  } else {
    g->time[a] = g->time[b] = T;
    memcpy(ch->all, ch->hold, sizeof(double) * R);
  }
  hky_settle(g, ok);
}

/* ---- updategenealogy's own move, models without migration ------------
 * update_gtree.cpp:723-826 with slider_nomigration (:79-253), joinsisdown
 * (:375-454), splitsisdown (:457-528) and findjointime (:35-75); the
 * Hastings term of a root move is log N(d; 0, s_new) - log N(d; 0, s_old)
 * with s = min(SLIDESTDVMAX = 20, root time / 3) (:780-812).  HKY loci,
 * full re-scoring, the same acceptance as slide() plus the Hastings term. */

static int o_period(const prob_t *pb, const double *tv, double t) {
  int k = 0; /* findperiod, update_gtree_common.cpp:1626-1633 */
  while (k < pb->P - 1 && tv[k] <= t) k++;
  return k;
}

static int o_dropped(const prob_t *pb, int p, int q) {
  /* droppops[p]: the populations merged at the split that opens period p */
  return pb->join[3 * (p - 1)] == q || pb->join[3 * (p - 1) + 1] == q;
}

static double o_jointime(const prob_t *pb, const double *tv, int slidepop, int sispop,
                         double edgeuptime, double sisuptime) {
  int ep = o_period(pb, tv, edgeuptime), sp = o_period(pb, tv, sisuptime);
  while (ep < sp) {
    ep++;
    if (o_dropped(pb, ep, slidepop)) slidepop = pb->popdown[slidepop];
  }
  while (sp < ep) {
    sp++;
    if (o_dropped(pb, sp, sispop)) sispop = pb->popdown[sispop];
  }
  while (slidepop != sispop) {
    ep++;
    if (o_dropped(pb, ep, slidepop)) slidepop = pb->popdown[slidepop];
    if (o_dropped(pb, ep, sispop)) sispop = pb->popdown[sispop];
  }
  return ep == 0 ? 0.0 : tv[ep - 1];
}

static double o_uptime(const gene_t *g, int e) { return e < g->n ? 0.0 : g->time[g->up0[e]]; }

/* slider_nomigration, the recursion as a loop: *sis and *tp move, *d is the
 * signed remaining distance (negative: up toward the tips) */
static void o_slider(const prob_t *pb, const gene_t *g, const double *tv, rng_t *r, int edge,
                     int *sis, double *tp, double d) {
  for (;;) {
    if (d < 0) {
      d = -d;
      const double eu = o_uptime(g, edge), su = o_uptime(g, *sis);
      const int spop = g->pop[edge], ppop = g->pop[*sis];
      const double pj = spop != ppop ? o_jointime(pb, tv, spop, ppop, eu, su) : 0.0;
      if (pj > eu && pj > su) {
        if (d < *tp - pj) { *tp -= d; return; }
        d -= *tp - pj;
        *tp = pj;
      } else if (su == 0 || eu >= su) {
        if (d < *tp - eu) { *tp -= d; return; }
        d -= *tp - eu;
        *tp = eu;
      } else {
        if (d < *tp - su) { *tp -= d; return; }
        d -= *tp - su;
        *tp = su;
        *sis = (xoshiro(r) >> 63) ? g->up0[*sis] : g->up1[*sis];
        d = -d;
      }
    } else {
      if (g->down[*sis] == -1 || *tp + d < g->time[*sis]) {
        *tp += d;
        if (*tp >= 1000000.0) *tp = 1000000.0; /* TIMEMAX */
        return;
      }
      d -= g->time[*sis] - *tp;
      *tp = g->time[*sis];
      if (xoshiro(r) >> 63) {
        *sis = g->down[*sis];
      } else {
        const int dd = g->down[*sis];
        *sis = g->up0[dd] == *sis ? g->up1[dd] : g->up0[dd];
        d = -d;
      }
    }
  }
}

static double o_lognorm(double s, double x) {
  return -0.91893853320467274178 - log(s) - 0.5 * (x / s) * (x / s);
}

static void topo_move(const prob_t *pb, chain_t *ch, int l, rng_t *r, long *acc) {
  gene_t *g = &ch->g[l];
  const int n = g->n, ne = 2 * n - 1, R = rec_len(pb->P);
  int *bk = (int *)malloc(sizeof(int) * 4 * ne);
  double *bt = (double *)malloc(sizeof(double) * ne);
  memcpy(bk, g->up0, sizeof(int) * ne);
  memcpy(bk + ne, g->up1, sizeof(int) * ne);
  memcpy(bk + 2 * ne, g->down, sizeof(int) * ne);
  memcpy(bk + 3 * ne, g->pop, sizeof(int) * ne);
  memcpy(bt, g->time, sizeof(double) * ne);
  const int broot = g->root;
  int edge;
  do edge = (int)(unif(r) * ne);
  while (edge < 0 || edge >= ne || g->down[edge] == -1);
  const int freed = g->down[edge];
  const int oldsis = g->up0[freed] == edge ? g->up1[freed] : g->up0[freed];
  const double sd0 = fmin(20.0, g->time[g->up0[g->root]] / 3.0);
  const double hold = sd0 * norm01(r);
  int rootmove = 0;
  double roottime = g->time[g->up0[g->root]];
  /* joinsisdown */
  g->time[oldsis] = g->time[freed];
  g->down[oldsis] = g->down[freed];
  const int dd0 = g->down[oldsis];
  if (dd0 != -1) {
    if (g->up0[dd0] == freed) g->up0[dd0] = oldsis;
    else g->up1[dd0] = oldsis;
  } else {
    rootmove = 1;
    g->root = oldsis;
    g->time[oldsis] = 1000000.0;
    roottime = o_uptime(g, oldsis);
  }
  /* slide */
  int sis = oldsis;
  double tp = g->time[edge];
  o_slider(pb, g, ch->tv, r, edge, &sis, &tp, hold);
  /* splitsisdown */
  g->time[edge] = tp;
  g->time[freed] = g->time[sis];
  g->time[sis] = tp;
  const int dd = g->down[sis];
  if (dd != -1) {
    if (g->up0[dd] == sis) g->up0[dd] = freed;
    else g->up1[dd] = freed;
  } else {
    g->root = freed;
    roottime = tp;
    rootmove = 1;
  }
  g->down[freed] = dd;
  int nowpop = g->pop[sis];
  const int j = o_period(pb, ch->tv, tp);
  while (pb->pope[nowpop] != -1 && pb->pope[nowpop] <= j) nowpop = pb->popdown[nowpop];
  g->pop[freed] = nowpop;
  g->down[sis] = freed;
  g->down[edge] = freed;
  g->up0[freed] = sis;
  g->up1[freed] = edge;
  double hast = 0.0;
  if (rootmove) hast = o_lognorm(fmin(20.0, roottime / 3.0), hold) - o_lognorm(sd0, hold);
  /* score and decide as slide() does */
  const double np = hky(pb, l, g, -1);
  double *nrec = ch->nrec;
  weights(pb, l, g, ch->tv, nrec);
  const int P2 = pb->P * pb->P, P3 = P2 * pb->P;
  for (int i = 0; i < R; i++) {
    const double x = ch->all[i], o = g->rec[i], w = nrec[i];
    ch->hold[i] = x;
    if ((i >= P2 && i < 2 * P2) || (i >= 3 * P2 + P3 && i < 3 * P2 + 2 * P3))
      ch->all[i] = fmax(0.0, (x - o) + w);
    else if (i < P2 || (i >= 3 * P2 && i < 3 * P2 + P3))
      ch->all[i] = x + (w - o);
    else
      ch->all[i] = (x - o) + w;
  }
  const double pn = prior_terms(pb, ch->all, ch->qc, ch->qf, ch->qt, ch->nc, ch->nf, ch->nt);
  const double U = unif(r);
  const int ok = U < fmin(1.0, exp(ch->beta * ((pn - ch->probg) + (np - g->pdg)) + hast));
  if (ok) {
    g->pdg = np;
    ch->probg = pn;
    memcpy(g->rec, nrec, sizeof(double) * R);
    const int np_ = pb->npsize + pb->nmig;
    memcpy(ch->qc, ch->nc, sizeof(int) * np_);
    memcpy(ch->qf, ch->nf, sizeof(double) * np_);
    memcpy(ch->qt, ch->nt, sizeof(double) * np_);
    (*acc)++;
  } else {
    memcpy(g->up0, bk, sizeof(int) * ne);
    memcpy(g->up1, bk + ne, sizeof(int) * ne);
    memcpy(g->down, bk + 2 * ne, sizeof(int) * ne);
    memcpy(g->pop, bk + 3 * ne, sizeof(int) * ne);
    memcpy(g->time, bt, sizeof(double) * ne);
    g->root = broot;
    memcpy(ch->all, ch->hold, sizeof(double) * R);
  }
  hky_settle(g, ok);
  free(bk);
  free(bt);
}

/* the reference's move for a model with migration (ora_gene_move_mig):
 * the whole genealogy, its populations and migrations are held for a
 * rejection; decided as topo_move with the Hastings term migweight +
 * slideweight.  A move whose migrations exceed the genealogy's room is
 * rejected. */
static double u_of(void *s) { return unif((rng_t *)s); }

static void topo_move_mig(const prob_t *pb, chain_t *ch, int l, rng_t *r, long *acc) {
  gene_t *g = &ch->g[l];
  const int n = g->n, ne = 2 * n - 1, R = rec_len(pb->P);
  int *bk = (int *)malloc(sizeof(int) * 4 * ne);
  double *bt = (double *)malloc(sizeof(double) * ne);
  memcpy(bk, g->up0, sizeof(int) * ne);
  memcpy(bk + ne, g->up1, sizeof(int) * ne);
  memcpy(bk + 2 * ne, g->down, sizeof(int) * ne);
  memcpy(bk + 3 * ne, g->pop, sizeof(int) * ne);
  memcpy(bt, g->time, sizeof(double) * ne);
  const int broot = g->root, bnm = g->nmig;
  int *bme = (int *)malloc(sizeof(int) * (bnm + 1)), *bmo = (int *)malloc(sizeof(int) * (bnm + 1));
  double *bmt = (double *)malloc(sizeof(double) * (bnm + 1));
  memcpy(bme, g->mig_edge, sizeof(int) * bnm);
  memcpy(bmo, g->mig_to, sizeof(int) * bnm);
  memcpy(bmt, g->mig_time, sizeof(double) * bnm);
  double hast = 0.0;
  int rm = 0;
  const int rc = ora_gene_move_mig(n, g->up0, g->up1, g->down, g->time, &g->root, g->pop, &g->nmig,
                                   g->mig_cap, g->mig_edge, g->mig_time, g->mig_to, pb->P,
                                   pb->plist, pb->npresent, pb->popdown, pb->pope, ch->tv, u_of, r,
                                   &hast, &rm);
  int ok = 0, scored = 0;
  double np = g->pdg, *nrec = ch->nrec;
  if (rc == 0 && weights(pb, l, g, ch->tv, nrec) == 0) {
    np = hky(pb, l, g, -1);
    scored = 1;
    const int P2 = pb->P * pb->P, P3 = P2 * pb->P;
    for (int i = 0; i < R; i++) {
      const double x = ch->all[i], o = g->rec[i], w = nrec[i];
      ch->hold[i] = x;
      if ((i >= P2 && i < 2 * P2) || (i >= 3 * P2 + P3 && i < 3 * P2 + 2 * P3))
        ch->all[i] = fmax(0.0, (x - o) + w);
      else if (i < P2 || (i >= 3 * P2 && i < 3 * P2 + P3))
        ch->all[i] = x + (w - o);
      else
        ch->all[i] = (x - o) + w;
    }
    const double pn = prior_terms(pb, ch->all, ch->qc, ch->qf, ch->qt, ch->nc, ch->nf, ch->nt);
    const double U = unif(r);
    ok = U < fmin(1.0, exp(ch->beta * ((pn - ch->probg) + (np - g->pdg)) + hast));
    if (ok) {
      g->pdg = np;
      ch->probg = pn;
      memcpy(g->rec, nrec, sizeof(double) * R);
      const int np_ = pb->npsize + pb->nmig;
      memcpy(ch->qc, ch->nc, sizeof(int) * np_);
      memcpy(ch->qf, ch->nf, sizeof(double) * np_);
      memcpy(ch->qt, ch->nt, sizeof(double) * np_);
      (*acc)++;
    } else {
      memcpy(ch->all, ch->hold, sizeof(double) * R);
    }
  }
  if (!ok) {
    memcpy(g->up0, bk, sizeof(int) * ne);
    memcpy(g->up1, bk + ne, sizeof(int) * ne);
    memcpy(g->down, bk + 2 * ne, sizeof(int) * ne);
    memcpy(g->pop, bk + 3 * ne, sizeof(int) * ne);
    memcpy(g->time, bt, sizeof(double) * ne);
    g->root = broot;
    g->nmig = bnm;
    memcpy(g->mig_edge, bme, sizeof(int) * bnm);
    memcpy(g->mig_to, bmo, sizeof(int) * bnm);
    memcpy(g->mig_time, bmt, sizeof(double) * bnm);
  }
  if (scored) hky_settle(g, ok);
  free(bk);
  free(bt);
  free(bme);
  free(bmo);
  free(bmt);
}

static void rescale(const prob_t *pb, chain_t *ch, rng_t *r, long *acc) {
  const int L = pb->L, R = rec_len(pb->P), np_ = pb->npsize + pb->nmig;
  const double ls = pb->scale_width * (unif(r) - 0.5), s = exp(ls);
  double a = 0.0;
  double *newp = (double *)malloc(sizeof(double) * L);
  double *recs = (double *)malloc(sizeof(double) * L * R);
  double tv_new[16];
  int ok = 1;
  for (int k = 0; k < pb->P - 1; k++) {
    tv_new[k] = ch->tv[k] * s;
    if (tv_new[k] > pb->tmax) ok = 0;
  }
  size_t tb = 0, mb = 0;
  for (int l = 0; l < L; l++) {
    gene_t *g = &ch->g[l];
    for (int i = 0; i < 2 * g->n - 1; i++) {
      ch->tback[tb++] = g->time[i];
      if (g->down[i] != -1) g->time[i] *= s;
    }
    for (int j = 0; j < g->nmig; j++) {
      ch->mback[mb++] = g->mig_time[j];
      g->mig_time[j] *= s;
    }
    newp[l] = hky(pb, l, g, -1);
    weights(pb, l, g, tv_new, recs + (size_t)l * R);
    a += ch->beta * (newp[l] - g->pdg) + (g->n - 1 + g->nmig) * ls;
  }
  double *fresh = (double *)calloc(R, sizeof(double));
  for (int l = 0; l < L; l++)
    for (int i = 0; i < R; i++) fresh[i] += recs[(size_t)l * R + i];
  const double pn = prior_terms(pb, fresh, NULL, NULL, NULL, ch->nc, ch->nf, ch->nt);
  a += ch->beta * (pn - ch->probg) + (pb->P - 1) * ls;
  const int dec = ok && unif(r) < fmin(1.0, exp(a));
  tb = mb = 0;
  for (int l = 0; l < L; l++) {
    gene_t *g = &ch->g[l];
    if (dec) {
      g->pdg = newp[l];
      memcpy(g->rec, recs + (size_t)l * R, sizeof(double) * R);
    } else {
      for (int i = 0; i < 2 * g->n - 1; i++) g->time[i] = ch->tback[tb++];
      for (int j = 0; j < g->nmig; j++) g->mig_time[j] = ch->mback[mb++];
    }
    hky_settle(g, dec);
  }
  if (dec) {
    memcpy(ch->all, fresh, sizeof(double) * R);
    memcpy(ch->qc, ch->nc, sizeof(int) * np_);
    memcpy(ch->qf, ch->nf, sizeof(double) * np_);
    memcpy(ch->qt, ch->nt, sizeof(double) * np_);
    ch->probg = pn;
    for (int k = 0; k < pb->P - 1; k++) ch->tv[k] = tv_new[k];
    (*acc)++;
  }
  free(newp);
  free(recs);
  free(fresh);
}

/* changet_RY1 (update_t_RY.cpp:223-517): split k = randposint(P - 1);
 * getnewt (update_gtree_common.cpp:2501-2518): a window of width
 * (t_d' - t_u') / (log(L + 1) (P - k)) centred on the old time, reflected
 * at t_u' = max(0, t_u) and t_d' = min(tmax, t_d), with t_u, t_d the
 * neighbouring split times (0, and TIMEMAX = 1e6 after the last split).
 * Node times of edges below the root and migration times in (t_u, oldt]
 * follow beforesplit, those in (oldt, t_d) aftersplit (update_t_RY.cpp:
 * 59-82).  Node times sit on both daughter edges, so edge counts are
 * halved; the Hastings term is (ecd + emd) log t_d_hterm + (ecu + emu)
 * log t_u_hterm (:259-266, :397-406).  Every locus is re-scored; accepted
 * when log U < min(1, beta (dpdg + dprobg) + Hastings). */
static void ry1(const prob_t *pb, chain_t *ch, rng_t *r, long *acc) {
  const int L = pb->L, R = rec_len(pb->P), np_ = pb->npsize + pb->nmig, ns = pb->P - 1;
  int k = (int)(unif(r) * ns);
  if (k > ns - 1) k = ns - 1;
  const double oldt = ch->tv[k];
  const double t_u = k == 0 ? 0.0 : ch->tv[k - 1];
  const double t_d = k == ns - 1 ? 1000000.0 : ch->tv[k + 1];
  const double tdp = fmin(pb->tmax, t_d), tup = fmax(0.0, t_u);
  const double twin = (tdp - tup) / (log((double)L + 1.0) * (double)(pb->P - k));
  double newt = (oldt - twin / 2) + unif(r) * twin;
  if (newt >= tdp)
    newt = 2.0 * tdp - newt;
  else if (newt <= tup)
    newt = 2.0 * tup - newt;
  double tv_new[16];
  for (int x = 0; x < ns; x++) tv_new[x] = ch->tv[x];
  tv_new[k] = newt;
  long ecu = 0, ecd = 0, emu = 0, emd = 0;
#define RY1_MOVE(t, cu, cd)                                                    \
  do {                                                                         \
    if ((t) <= oldt && (t) > t_u) {                                            \
      (t) = k == 0 ? (t) * newt / oldt                                         \
                   : t_u + ((t) - t_u) * (newt - t_u) / (oldt - t_u);          \
      (cu)++;                                                                  \
    } else if ((t) > oldt && (t) < t_d) {                                      \
      (t) = k == ns - 1 ? (t) + newt - oldt                                    \
                        : t_d - (t_d - newt) * (t_d - (t)) / (t_d - oldt);     \
      (cd)++;                                                                  \
    }                                                                          \
  } while (0)
  double a = 0.0;
  double *newp = (double *)malloc(sizeof(double) * L);
  double *recs = (double *)malloc(sizeof(double) * L * R);
  size_t tb = 0, mb = 0;
  for (int l = 0; l < L; l++) {
    gene_t *g = &ch->g[l];
    for (int i = 0; i < 2 * g->n - 1; i++) {
      ch->tback[tb++] = g->time[i];
      if (g->down[i] != -1) RY1_MOVE(g->time[i], ecu, ecd);
    }
    for (int j = 0; j < g->nmig; j++) {
      ch->mback[mb++] = g->mig_time[j];
      RY1_MOVE(g->mig_time[j], emu, emd);
    }
    newp[l] = hky(pb, l, g, -1);
    weights(pb, l, g, tv_new, recs + (size_t)l * R);
    a += ch->beta * (newp[l] - g->pdg);
  }
#undef RY1_MOVE
  ecu /= 2;
  ecd /= 2;
  double *fresh = (double *)calloc(R, sizeof(double));
  for (int l = 0; l < L; l++)
    for (int i = 0; i < R; i++) fresh[i] += recs[(size_t)l * R + i];
  const double pn = prior_terms(pb, fresh, NULL, NULL, NULL, ch->nc, ch->nf, ch->nt);
  const double t_u_hterm = (newt - t_u) / (oldt - t_u);
  const double t_d_hterm = k == ns - 1 ? 1.0 : (t_d - newt) / (t_d - oldt);
  a += ch->beta * (pn - ch->probg) + (double)(ecd + emd) * log(t_d_hterm) +
       (double)(ecu + emu) * log(t_u_hterm);
  const double lu = log(unif(r));
  const int dec = lu < fmin(1.0, a);
  tb = mb = 0;
  for (int l = 0; l < L; l++) {
    gene_t *g = &ch->g[l];
    if (dec) {
      g->pdg = newp[l];
      memcpy(g->rec, recs + (size_t)l * R, sizeof(double) * R);
    } else {
      for (int i = 0; i < 2 * g->n - 1; i++) g->time[i] = ch->tback[tb++];
      for (int j = 0; j < g->nmig; j++) g->mig_time[j] = ch->mback[mb++];
    }
    hky_settle(g, dec);
  }
  if (dec) {
    memcpy(ch->all, fresh, sizeof(double) * R);
    memcpy(ch->qc, ch->nc, sizeof(int) * np_);
    memcpy(ch->qf, ch->nf, sizeof(double) * np_);
    memcpy(ch->qt, ch->nt, sizeof(double) * np_);
    ch->probg = pn;
    ch->tv[k] = newt;
    (*acc)++;
  }
  free(newp);
  free(recs);
  free(fresh);
}

/* changet_NW (update_t_NW.cpp:920-1032): split k and newt drawn as for
 * RY1 (getnewt); every genealogy whose root lies above min(oldt, newt)
 * gets its migrations in the interval between the two times re-simulated
 * (ora_update_mig_tnw, one fresh 64-bit stream key per locus); node times
 * and so the data likelihood do not change.  Accepted when U < min(1,
 * exp(beta dprobg + migweight)); a genealogy whose new migrations exceed
 * its capacity, or that treeweight rejects, rejects the move. */
static void nw(const prob_t *pb, chain_t *ch, rng_t *r, long *acc) {
  const int L = pb->L, R = rec_len(pb->P), np_ = pb->npsize + pb->nmig, ns = pb->P - 1;
  int k = (int)(unif(r) * ns);
  if (k > ns - 1) k = ns - 1;
  const double oldt = ch->tv[k];
  const double t_u = k == 0 ? 0.0 : ch->tv[k - 1];
  const double t_d = k == ns - 1 ? 1000000.0 : ch->tv[k + 1];
  const double tdp = fmin(pb->tmax, t_d), tup = fmax(0.0, t_u);
  const double twin = (tdp - tup) / (log((double)L + 1.0) * (double)(pb->P - k));
  double newt = (oldt - twin / 2) + unif(r) * twin;
  if (newt >= tdp)
    newt = 2.0 * tdp - newt;
  else if (newt <= tup)
    newt = 2.0 * tup - newt;
  double tv_new[16];
  for (int x = 0; x < ns; x++) tv_new[x] = ch->tv[x];
  tv_new[k] = newt;
  double *recs = (double *)malloc(sizeof(double) * L * R);
  int **bpop = (int **)calloc(L, sizeof(int *)), **bme = (int **)calloc(L, sizeof(int *));
  int **bmo = (int **)calloc(L, sizeof(int *)), *bnm = (int *)calloc(L, sizeof(int));
  double **bmt = (double **)calloc(L, sizeof(double *));
  double mw = 0.0;
  int ok = 1;
  for (int l = 0; l < L; l++) {
    gene_t *g = &ch->g[l];
    const int ne = 2 * g->n - 1, nm = g->nmig;
    bpop[l] = (int *)malloc(sizeof(int) * ne);
    memcpy(bpop[l], g->pop, sizeof(int) * ne);
    bnm[l] = nm;
    bme[l] = (int *)malloc(sizeof(int) * (nm + 1));
    bmo[l] = (int *)malloc(sizeof(int) * (nm + 1));
    bmt[l] = (double *)malloc(sizeof(double) * (nm + 1));
    memcpy(bme[l], g->mig_edge, sizeof(int) * nm);
    memcpy(bmo[l], g->mig_to, sizeof(int) * nm);
    memcpy(bmt[l], g->mig_time, sizeof(double) * nm);
    const uint64_t key = xoshiro(r);
    double w = 0.0;
    int upd = 0;
    if (ora_update_mig_tnw(g->n, g->up0, g->up1, g->down, g->time, g->root, g->pop, &g->nmig,
                           g->mig_cap, g->mig_edge, g->mig_time, g->mig_to, pb->P, pb->plist,
                           pb->npresent, pb->popdown, pb->pope, pb->join, ch->tv, k, newt, key,
                           0, &w, &upd) != 0)
      ok = 0;
    mw += w;
    if (weights(pb, l, g, tv_new, recs + (size_t)l * R) != 0) ok = 0;
  }
  double *fresh = (double *)calloc(R, sizeof(double));
  for (int l = 0; l < L; l++)
    for (int i = 0; i < R; i++) fresh[i] += recs[(size_t)l * R + i];
  const double pn = prior_terms(pb, fresh, NULL, NULL, NULL, ch->nc, ch->nf, ch->nt);
  const double a = ch->beta * (pn - ch->probg) + mw;
  const double U = unif(r);
  const int dec = ok && U < fmin(1.0, exp(a));
  for (int l = 0; l < L; l++) {
    gene_t *g = &ch->g[l];
    if (dec) {
      memcpy(g->rec, recs + (size_t)l * R, sizeof(double) * R);
    } else {
      memcpy(g->pop, bpop[l], sizeof(int) * (2 * g->n - 1));
      g->nmig = bnm[l];
      memcpy(g->mig_edge, bme[l], sizeof(int) * bnm[l]);
      memcpy(g->mig_to, bmo[l], sizeof(int) * bnm[l]);
      memcpy(g->mig_time, bmt[l], sizeof(double) * bnm[l]);
    }
    free(bpop[l]);
    free(bme[l]);
    free(bmo[l]);
    free(bmt[l]);
  }
  if (dec) {
    memcpy(ch->all, fresh, sizeof(double) * R);
    memcpy(ch->qc, ch->nc, sizeof(int) * np_);
    memcpy(ch->qf, ch->nf, sizeof(double) * np_);
    memcpy(ch->qt, ch->nt, sizeof(double) * np_);
    ch->probg = pn;
    ch->tv[k] = newt;
    (*acc)++;
  }
  free(bpop);
  free(bme);
  free(bmo);
  free(bmt);
  free(bnm);
  free(recs);
  free(fresh);
}

/* kappa move of an HKY locus (update_mc_params.cpp:258-273, :394-404) */
static double kappa_move(double k, double U) {
  const double win = 2.0, kmax = 100.0; /* initialize.cpp:1499-1501, KAPPAMAX */
  double nk;
  if (U > 0.5) {
    nk = k + (2.0 * U - 1.0) * win;
    if (nk > kmax) nk = 2.0 * kmax - nk;
  } else {
    nk = k - win * U * 2.0;
    if (nk < 0) nk = -nk;
  }
  return nk;
}

/* changeu (update_mc_params.cpp:24-382) over one chain as qupdate runs it
 * (ima_main_mpi.cpp:1917-1945): for j = 0 .. L-1 (only j = 0 when L = 2)
 * k is uniform over the other loci, log(u_j/u_k) moves by a uniform window
 * of log(UMAX)/L (initialize.cpp:1448-1450) reflected at 3 log(UMAX)
 * (:52), u_j *= d and u_k /= d with d = exp(half the move), each locus'
 * kappa moves, both loci are re-scored and the pair is accepted with
 * U < min(1, exp(beta sum(new - old))).  One locus: changekappa (:387-429),
 * accepted when the ratio is >= 1 or > U.  Every locus here is HKY. */
static void change_rates(const prob_t *pb, chain_t *ch, rng_t *r, long *tries, long *acc) {
  const int L = pb->L;
  if (L == 1) {
    gene_t *g = &ch->g[0];
    const double ok = g->kappa;
    g->kappa = kappa_move(ok, unif(r));
    const double np = hky(pb, 0, g, -1);
    const double mh = exp(ch->beta * (np - g->pdg));
    (*tries)++;
    if (mh >= 1.0 || mh > unif(r)) {
      g->pdg = np;
      hky_settle(g, 1);
      (*acc)++;
    } else {
      g->kappa = ok;
      hky_settle(g, 0);
    }
    return;
  }
  const double UMAX = 10000.0;
  const double win = log(UMAX) / L, maxr = 3.0 * log(UMAX);
  const int J = L > 2 ? L : 1;
  for (int j = 0; j < J; j++) {
    int k = 1;
    if (L > 2) {
      do k = (int)(unif(r) * L);
      while (k == j || k < 0 || k >= L);
    }
    gene_t *gl[2] = {&ch->g[j], &ch->g[k]};
    const int li[2] = {j, k};
    const double ou[2] = {gl[0]->u, gl[1]->u}, ok[2] = {gl[0]->kappa, gl[1]->kappa};
    const double rr = log(ou[0] / ou[1]);
    const double U = unif(r);
    double nr = U > 0.5 ? rr + (2.0 * U - 1.0) * win : rr - win * U * 2.0;
    if (nr > maxr)
      nr = 2.0 * maxr - nr;
    else if (nr < -maxr)
      nr = 2.0 * (-maxr) - nr;
    const double d = exp((nr - rr) / 2);
    gl[0]->u *= d;
    gl[1]->u /= d;
    double np[2], sum = 0.0;
    for (int i = 0; i < 2; i++) {
      gl[i]->kappa = kappa_move(gl[i]->kappa, unif(r));
      np[i] = hky(pb, li[i], gl[i], -1);
      sum += np[i] - gl[i]->pdg;
    }
    const double mh = exp(ch->beta * sum);
    (*tries)++;
    const int a = unif(r) < fmin(1.0, mh);
    for (int i = 0; i < 2; i++) {
      if (a) {
        gl[i]->pdg = np[i];
      } else {
        gl[i]->u = ou[i];
        gl[i]->kappa = ok[i];
      }
      hky_settle(gl[i], a);
    }
    *acc += a;
  }
}

/*
 * Runs sweeps over nchains chains x L loci for about `seconds` of wall
 * time, or exactly max_sweeps sweeps when max_sweeps > 0, and returns the
 * number of sweeps done; stats
 * (may be NULL, 6 entries) receives genealogy proposals, accepts, rescale
 * proposals, accepts, mutation-rate steps and accepts.  u and kappa are
 * per (chain, locus), read and updated in place; urate adds changeu after
 * every sweep s with s % 5 == 4.  Genealogies are packed chain-major, locus-minor:
 * tree arrays of 2n_l-1 entries each (up0, up1, down, pop, time), roots,
 * migration counts and their (edge, time, destination) packed in order.
 * tvs[c][0..P-2] are chain c's split times; times, migration times and
 * split times are updated in place.  ry1 selects the rubber-band time
 * update (two or more populations) in place of the global rescale.  nw
 * (models with migration): 1, each time update is NW with probability 1/2;
 * 2, every sweep is one NW update per chain and nothing else.  With nw
 * each genealogy's migrations sit in a region of mig_cap[c * L + l]
 * entries (regions in the same order), and nmig and pop are written back.
 */
long ora_mc3_bench(int nchains, int L, const int *n_l, const int *S, const int *seq,
                   const int *seq_off, const int *mult, const int *mult_off,
                   const int *totsites, const double *pi, double *u, double *kappa,
                   const int *samppop, int P, const int *plist, const int *npresent,
                   const int *popdown, const int *pope, const int *join, int npsize,
                   const int *theta_off, const int *theta_wp, int nmigp, const int *mig_off,
                   const int *mig_wp, const double *qmax, const double *mmax, double tmax,
                   const int *up0, const int *up1, const int *down, int *pop,
                   double *times, const int *roots, int *nmig, int *mig_edge,
                   double *mig_time, int *mig_to, double *tvs,
                   const double *beta0, double scale_width, uint64_t seed, double seconds,
                   long max_sweeps, long *stats, int ry1_update, int urate_update,
                   int topo_update, int *roots_out, int nw_update, const int *mig_cap) {
  prob_t pb = {L, S, totsites, seq_off, mult_off, samppop, seq, mult, pi,
               P, npsize, nmigp, plist, npresent, popdown, pope, join,
               theta_off, theta_wp, mig_off, mig_wp, qmax, mmax, tmax, scale_width,
               ry1_update && P > 1, topo_update, (P > 1 && nmigp > 0) ? nw_update : 0};
  /* new migrations (NW, the move with migration paths) need room */
  if ((pb.nw || (topo_update && nmigp > 0)) && !mig_cap) return -1;
  const int R = rec_len(P), np_ = npsize + nmigp;
  if (P > 10 || P - 1 > 16) return -1;
  chain_t *ch = (chain_t *)calloc(nchains, sizeof(chain_t));
  size_t off = 0, moff = 0;
  for (int c = 0; c < nchains; c++) {
    chain_t *h = &ch[c];
    h->g = (gene_t *)calloc(L, sizeof(gene_t));
    h->tv = (double *)malloc(sizeof(double) * (P > 1 ? P - 1 : 1));
    for (int k = 0; k < P - 1; k++) h->tv[k] = tvs[c * (P - 1) + k];
    h->all = (double *)calloc(R, sizeof(double));
    h->hold = (double *)calloc(R, sizeof(double));
    h->nrec = (double *)calloc(R, sizeof(double));
    h->qc = (int *)calloc(np_ + 1, sizeof(int));
    h->nc = (int *)calloc(np_ + 1, sizeof(int));
    h->qf = (double *)calloc(np_ + 1, sizeof(double));
    h->qt = (double *)calloc(np_ + 1, sizeof(double));
    h->nf = (double *)calloc(np_ + 1, sizeof(double));
    h->nt = (double *)calloc(np_ + 1, sizeof(double));
    h->beta = beta0[c];
    for (int l = 0; l < L; l++) {
      gene_t *g = &h->g[l];
      const int n = n_l[l], ne = 2 * n - 1;
      g->n = n;
      g->root = roots[c * L + l];
      g->up0 = (int *)(up0 + off);
      g->up1 = (int *)(up1 + off);
      g->down = (int *)(down + off);
      g->pop = (int *)(pop + off);
      g->time = times + off;
      g->nmig = nmig[c * L + l];
      g->mig_cap = mig_cap ? mig_cap[c * L + l] : g->nmig;
      g->mig_edge = mig_edge + moff;
      g->mig_to = mig_to + moff;
      g->mig_time = mig_time + moff;
      off += ne;
      moff += g->mig_cap;
      g->u = u[c * L + l];
      g->kappa = kappa[c * L + l];
      g->st.n = n;
      g->st.S = S[l];
      g->st.frac = (double *)calloc((size_t)(n - 1) * S[l] * 4, sizeof(double));
      g->st.newfrac = (double *)calloc((size_t)(n - 1) * S[l] * 4, sizeof(double));
      g->st.scale = (double *)calloc((size_t)(n - 1) * S[l], sizeof(double));
      g->st.oldscale = (double *)calloc((size_t)(n - 1) * S[l], sizeof(double));
      g->pdg = hky(&pb, l, g, -1);
      hky_settle(g, 1);
      g->rec = (double *)malloc(sizeof(double) * R);
      weights(&pb, l, g, h->tv, g->rec);
      for (int i = 0; i < R; i++) h->all[i] += g->rec[i];
    }
    size_t nt = 0, nm = 0;
    for (int l = 0; l < L; l++) {
      nt += 2 * (size_t)h->g[l].n - 1;
      nm += (size_t)h->g[l].mig_cap; /* NW may add migrations up to the room */
    }
    h->tback = (double *)malloc(sizeof(double) * (nt + 1));
    h->mback = (double *)malloc(sizeof(double) * (nm + 1));
    h->probg = prior_terms(&pb, h->all, NULL, NULL, NULL, h->qc, h->qf, h->qt);
  }
  rng_t r;
  uint64_t sm = seed;
  for (int i = 0; i < 4; i++) r.s[i] = splitmix(&sm);
  long acc[3] = {0, 0, 0}, props[3] = {0, 0, 0}, sweeps = 0;
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  double el = 0.0;
  do {
    for (int c = 0; c < nchains; c++) {
      if (pb.nw == 2) { /* the NW update alone */
        nw(&pb, &ch[c], &r, &acc[1]);
        props[1]++;
        continue;
      }
      for (int l = 0; l < L; l++) {
        if (pb.topo && nmigp > 0)
          topo_move_mig(&pb, &ch[c], l, &r, &acc[0]);
        else if (pb.topo)
          topo_move(&pb, &ch[c], l, &r, &acc[0]);
        else
          slide(&pb, &ch[c], l, &r, &acc[0]);
        props[0]++;
      }
      if (sweeps & 1) {
        /* qupdate's time update: RY1 or NW with equal probability when
         * the model has migration (ima_main_mpi.cpp:1867-1911) */
        if (pb.nw == 1 && unif(&r) < 0.5)
          nw(&pb, &ch[c], &r, &acc[1]);
        else if (pb.ry1)
          ry1(&pb, &ch[c], &r, &acc[1]);
        else
          rescale(&pb, &ch[c], &r, &acc[1]);
        props[1]++;
      }
    }
    /* swap (swapchains.cpp:527-653) on S = sum pdg + probg */
    const int swaptries = nchains / 10 > 1 ? nchains / 10 : 1;
    for (int i = 0; i < swaptries && nchains > 1; i++) {
      const int ci = (int)(unif(&r) * nchains);
      int cj, cjmin = 0, cjrange = nchains;
      if (nchains >= 17) {
        cjmin = ci - 7 > 0 ? ci - 7 : 0;
        cjrange = (ci + 7 < nchains ? ci + 7 : nchains) - cjmin;
      }
      do cj = cjmin + (int)(unif(&r) * cjrange);
      while (cj == ci);
      double Si = ch[ci].probg, Sj = ch[cj].probg;
      for (int l = 0; l < L; l++) {
        Si += ch[ci].g[l].pdg;
        Sj += ch[cj].g[l].pdg;
      }
      const double w = exp((ch[ci].beta - ch[cj].beta) * (Sj - Si));
      if (w >= 1.0 || w > unif(&r)) {
        const double b = ch[ci].beta;
        ch[ci].beta = ch[cj].beta;
        ch[cj].beta = b;
      }
    }
    if (urate_update && sweeps % 5 == 4)
      for (int c = 0; c < nchains; c++) change_rates(&pb, &ch[c], &r, &props[2], &acc[2]);
    sweeps++;
    clock_gettime(CLOCK_MONOTONIC, &t1);
    el = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  } while (max_sweeps > 0 ? sweeps < max_sweeps : el < seconds);
  if (stats) {
    stats[0] = props[0];
    stats[1] = acc[0];
    stats[2] = props[1];
    stats[3] = acc[1];
    stats[4] = props[2];
    stats[5] = acc[2];
  }
  for (int c = 0; c < nchains; c++) {
    for (int k = 0; k < P - 1; k++) tvs[c * (P - 1) + k] = ch[c].tv[k];
    for (int l = 0; l < L; l++) {
      gene_t *g = &ch[c].g[l];
      if (roots_out) roots_out[c * L + l] = g->root;
      nmig[c * L + l] = g->nmig;
      u[c * L + l] = g->u;
      kappa[c * L + l] = g->kappa;
      free(g->st.frac);
      free(g->st.newfrac);
      free(g->st.scale);
      free(g->st.oldscale);
      free(g->rec);
    }
    free(ch[c].g);
    free(ch[c].tv);
    free(ch[c].all);
    free(ch[c].hold);
    free(ch[c].nrec);
    free(ch[c].tback);
    free(ch[c].mback);
    free(ch[c].qc);
    free(ch[c].nc);
    free(ch[c].qf);
    free(ch[c].qt);
    free(ch[c].nf);
    free(ch[c].nt);
  }
  free(ch);
  return sweeps;
}
