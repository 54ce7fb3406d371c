/*
 * ora_mig.c -- updategenealogy's proposal for a model with migration.
 *
 * TEST INFRASTRUCTURE ONLY (see ima_oracle.h): the oracle's sweep calls it
 * and the tests use it as the checker of the engine's device move; nothing
 * in the product calls it.
 *
 * The move (update_gtree.cpp:723-826): an edge other than the root's is
 * picked uniformly; joinsisdown (:375-454) frees its parent node by
 * joining the sister and parent edges (migrations concatenated); the edge
 * loses its migrations and its lower end slides a normal distance
 * (standard deviation min(20, root time / 3)) along the genealogy as
 * slider does (:256-372: up to the edge's own daughters' node, reflecting
 * there, down toward the root, a random branch at every node passed);
 * splitsisdown (:457-528) puts the freed node back where the slide stops,
 * the new sister's migrations split at that time and its population there
 * labelling the node; addmigration (:589-665) draws the edge's migrations
 * again -- and the new sister's too when the edge now ends at the root --
 * period by period with Poisson counts whose rates come from the old
 * edges' counts and times (mwork_single_edge, mwork_two_edges,
 * update_gtree_common.cpp:1354-1583; simmpath :330-441; fillmiginfo and
 * fillmiginfoperiods :1101-1249); getmprob (:851-1099) gives the forward
 * and reverse path probabilities; copynewmig_to_gtree (:1251-1292) writes
 * the paths and the root's population.  The Hastings term is migweight
 * (reverse minus forward) plus the root move's slideweight.
 *
 * Randomness comes from the caller's uniform source, in the reference's
 * order of draws (normdev is the polar method without its cached second
 * deviate; a random bit is U < 1/2).  The one value changed: when an edge
 * ending at the root leaves its final population open, the reference
 * stores the root edge's index in fpop and copynewmig_to_gtree overwrites
 * the root's population afterwards; here it is the root population, which
 * is equally unused.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ima_oracle.h"

#define MG_TIMEMAX 1000000.0
#define MG_MIGCLOSEFRAC 0.9
#define MG_LOG2 0.69314718055994530941723212146
#define MG_ADDMIGMAX 1000
#define MG_MAXP 10
#define MG_MAX_TRIES (1 << 20)

typedef struct {
  double (*u)(void *);
  void *s;
} usrc_t;

static double U_(usrc_t *r) { return r->u(r->s); }
static int randposint_(usrc_t *r, int n) { return (int)floor(U_(r) * n); }

static double normdev_(usrc_t *r, double mean, double sd) {
  double v1, v2, rsq;
  do {
    v1 = 2.0 * U_(r) - 1.0;
    v2 = 2.0 * U_(r) - 1.0;
    rsq = v1 * v1 + v2 * v2;
  } while (rsq >= 1.0 || rsq == 0.0);
  return v2 * sqrt(-2.0 * log(rsq) / rsq) * sd + mean;
}

/* poisson (utilities.cpp:517-596): condition -1 any, 0 even, 1 odd,
 * 2 not 0, 3 not 1 */
static int poisson_(usrc_t *r, double param, int cond) {
  if (param < 0.25 && cond == 1) {
    const double x = U_(r);
    if (x < param / sinh(param)) return 1;
    if (x < param * (6 + param * param) / (6 * sinh(param))) return 3;
    return 5;
  }
  if (param < 0.25 && cond == 2) {
    double raised = exp(-param);
    double rcheck = raised = param * raised / (1 - raised);
    const double x = U_(r);
    int i = 1;
    while (x > rcheck) {
      raised *= param / (i + 1);
      rcheck += raised;
      i++;
    }
    return i;
  }
  int i, stop, tries = 0;
  do {
    if (param >= 100.0) {
      const long v = (long)(normdev_(r, param, param) + 0.5);
      i = v > 0 ? (int)(v > 2000000000L ? 2000000000L : v) : 0;
    } else {
      const double raised = exp(-param);
      double x = U_(r);
      for (i = 0; x > raised && i < MG_MAX_TRIES; i++) x *= U_(r);
    }
    stop = cond == -1 ? 1 : cond == 0 ? !(i & 1) : cond == 1 ? (i & 1) : cond == 2 ? i != 0 : i != 1;
    if (!stop && ++tries >= MG_MAX_TRIES) {
      i = cond == 1 || cond == 2 ? 1 : 0;
      stop = 1;
    }
  } while (!stop);
  if (i > MG_ADDMIGMAX) i = cond == 1 ? MG_ADDMIGMAX - 1 : MG_ADDMIGMAX;
  return i;
}

static double calcmrate_(int mc, double mt) {
  if (mt <= 0.0) return 1.0;
  if (mc == 0) return mt < 1 ? 0.1 : 0.1 / mt;
  return mt < 1 ? (double)mc : (double)mc / mt;
}

static double logcosh__(double x) { return x < 100 ? log(cosh(x)) : x - MG_LOG2; }
static double logsinh__(double x) { return x < 100 ? log(sinh(x)) : x - MG_LOG2; }

/* a growable list of migration events (time, destination) */
typedef struct {
  int n, cap;
  double *t;
  int *p;
} mlist_t;

static void ml_push(mlist_t *l, double t, int p) {
  if (l->n == l->cap) {
    l->cap = l->cap ? 2 * l->cap : 8;
    l->t = (double *)realloc(l->t, sizeof(double) * l->cap);
    l->p = (int *)realloc(l->p, sizeof(int) * l->cap);
  }
  l->t[l->n] = t;
  l->p[l->n] = p;
  l->n++;
}

/* the model as the move sees it */
typedef struct {
  int P, lastp, rootpop;
  const int *plist, *npresent, *popdown, *pope;
  double tv[MG_MAXP]; /* split times, tv[P - 1] = TIMEMAX */
} mmodel_t;

static int findperiod_(const mmodel_t *md, double t) {
  int k = 0;
  while (k < md->lastp && md->tv[k] <= t) k++;
  return k;
}

static int picktopop_(usrc_t *r, const mmodel_t *md, int period, int nowpop, int notother) {
  const int *pl = md->plist + period * md->P, np = md->npresent[period];
  int q, tries = 0;
  do q = randposint_(r, np);
  while ((pl[q] == nowpop || pl[q] == notother) && ++tries < MG_MAX_TRIES);
  return pl[q];
}

/* edgemiginfo (update_gtree_common.cpp) */
typedef struct {
  int edgeid;
  double upt, dnt;
  int pop, temppop, fpop;
  int b, e;
  double mtimeavail[MG_MAXP];
  double mtall;
  int mp[MG_MAXP];
  int mpall;
  mlist_t mig;
} emi_t;

static void emi_reset(emi_t *m) {
  m->edgeid = -1;
  m->b = m->e = -1;
  m->pop = m->temppop = m->fpop = -1;
  m->upt = m->dnt = -1.0;
  memset(m->mtimeavail, 0, sizeof m->mtimeavail);
  m->mtall = 0.0;
  memset(m->mp, 0, sizeof m->mp);
  m->mpall = 0;
  m->mig.n = 0;
}

/* fillmiginfoperiods (:1101-1162) */
static void fill_periods(const mmodel_t *md, emi_t *m) {
  m->b = 0;
  while (m->upt > md->tv[m->b]) m->b++;
  m->e = m->b;
  while (m->dnt > md->tv[m->e]) m->e++;
  if (m->e == m->b) {
    m->mtimeavail[m->b] = m->b == md->lastp ? 0.0 : m->dnt - m->upt;
  } else {
    m->mtimeavail[m->b] = md->tv[m->b] - m->upt;
    m->mtimeavail[m->e] = m->e == md->lastp ? 0.0 : m->dnt - md->tv[m->e - 1];
    for (int i = m->b + 1; i < m->e; i++) m->mtimeavail[i] = md->tv[i] - md->tv[i - 1];
  }
  m->mtall = m->b < md->lastp ? fmin(md->tv[md->lastp - 1], m->dnt) - m->upt : 0.0;
}

typedef struct {
  int n, ne, root;
  int *up0, *up1, *down, *pop;
  double *time;
  mlist_t *L; /* per edge */
} mgene_t;

static double uptime_(const mgene_t *g, int e) { return e < g->n ? 0.0 : g->time[g->up0[e]]; }

/* fillmiginfo (:1170-1249) for one edge */
static void fill_old(const mmodel_t *md, const mgene_t *g, int edge, emi_t *m) {
  emi_reset(m);
  m->edgeid = edge;
  m->pop = g->pop[edge];
  m->dnt = g->time[edge];
  m->upt = uptime_(g, edge);
  m->fpop = g->pop[g->down[edge]];
  fill_periods(md, m);
  int j = m->b;
  for (int i = 0; i < g->L[edge].n; i++) {
    while (g->L[edge].t[i] > md->tv[j]) j++;
    m->mp[j]++;
    m->mpall++;
    ml_push(&m->mig, g->L[edge].t[i], g->L[edge].p[i]);
  }
}

/* simmpath (:330-441): numm events on (upt, upt + timein), sorted, the
 * destinations from the period's populations (the last two fixed when the
 * path must end in constrainpop); returns the population reached */
static int simmpath_(usrc_t *r, const mmodel_t *md, emi_t *m, int period, int numm,
                     double timein, double upt, int pop, int constrainpop) {
  const int start = m->mig.n;
  for (int i = 0; i < numm; i++) ml_push(&m->mig, 0.0, -1);
  double *t = m->mig.t + start;
  int dup, tries = 0;
  do {
    for (int i = 0; i < numm; i++) t[i] = upt + U_(r) * timein;
    for (int a = 1; a < numm; a++) {
      const double x = t[a];
      int b = a - 1;
      while (b >= 0 && t[b] > x) {
        t[b + 1] = t[b];
        b--;
      }
      t[b + 1] = x;
    }
    dup = 0;
    for (int i = 0; i + 1 < numm; i++)
      if (t[i] == t[i + 1]) dup = 1;
  } while (dup && ++tries < MG_MAX_TRIES);
  int last = pop;
  for (int i = 0; i < numm; i++) {
    int q;
    if (constrainpop >= 0 && i >= numm - 2)
      q = i == numm - 2 ? picktopop_(r, md, period, last, constrainpop) : constrainpop;
    else
      q = picktopop_(r, md, period, last, -1);
    m->mig.p[start + i] = q;
    last = q;
  }
  return last;
}

static void settle_(const mmodel_t *md, int *pop, int period) {
  while (md->pope[*pop] <= period && md->pope[*pop] != -1) *pop = md->popdown[*pop];
}

/* mwork_single_edge (:1354-1429) */
static int mwork_single(usrc_t *r, const mmodel_t *md, emi_t *m, const emi_t *old, int lastmig) {
  if (lastmig < m->b) return 0;
  int mpall = 0, k;
  double tstart = m->upt;
  for (k = m->b; k <= lastmig; k++) {
    settle_(md, &m->temppop, k);
    const double rate = calcmrate_(old->mp[k], old->mtimeavail[k]) * m->mtimeavail[k];
    int constrain = -1;
    if (k < m->e) {
      m->mp[k] = poisson_(r, rate, -1);
    } else if (md->P - k == 2) {
      m->mp[k] = poisson_(r, rate, m->temppop == m->fpop ? 0 : 1);
    } else {
      m->mp[k] = poisson_(r, rate, m->temppop == m->fpop ? 3 : 2);
      constrain = m->fpop;
    }
    if (m->mp[k] > 0) {
      m->temppop = simmpath_(r, md, m, k, m->mp[k], m->mtimeavail[k], tstart, m->temppop, constrain);
      mpall += m->mp[k];
    }
    tstart = md->tv[k];
  }
  if (m->mtimeavail[k] > 0 && md->pope[m->temppop] == k) m->temppop = md->popdown[m->temppop];
  return mpall;
}

/* mwork_two_edges (:1431-1583) */
static void mwork_two(usrc_t *r, const mmodel_t *md, emi_t *em, emi_t *sm, const emi_t *oe,
                      const emi_t *os, int lastmig) {
  emi_t *mm[2] = {em, sm};
  const emi_t *om[2] = {oe, os};
  double tstart[2] = {em->upt, sm->upt};
  const int b = em->b < sm->b ? em->b : sm->b;
  const int lastk = em->e == md->lastp ? lastmig : lastmig - 1;
  int k;
  for (k = b; k <= lastk; k++)
    for (int ii = 0; ii < 2; ii++) {
      emi_t *m = mm[ii];
      if (m->b > k) continue;
      settle_(md, &m->temppop, k);
      const double rate = calcmrate_(om[ii]->mp[k], om[ii]->mtimeavail[k]) * m->mtimeavail[k];
      m->mp[k] = poisson_(r, rate, -1);
      if (m->mp[k] > 0) {
        m->temppop = simmpath_(r, md, m, k, m->mp[k], m->mtimeavail[k], tstart[ii], m->temppop, -1);
        m->mpall += m->mp[k];
      }
      tstart[ii] = md->tv[k];
    }
  if (k == md->lastp) {
    em->fpop = sm->fpop = md->rootpop;
    return;
  }
  /* the period the two edges end in: their joint final population */
  for (int ii = 0; ii < 2; ii++)
    if (md->pope[mm[ii]->temppop] == k) mm[ii]->temppop = md->popdown[mm[ii]->temppop];
  int f;
  if (em->temppop == sm->temppop) {
    f = U_(r) < MG_MIGCLOSEFRAC ? em->temppop : picktopop_(r, md, k, em->temppop, -1);
  } else if (md->P - k == 2) {
    f = U_(r) < 0.5 ? em->temppop : sm->temppop;
  } else if (U_(r) < MG_MIGCLOSEFRAC) {
    f = U_(r) < 0.5 ? em->temppop : sm->temppop;
  } else {
    f = picktopop_(r, md, k, em->temppop, sm->temppop);
  }
  em->fpop = sm->fpop = f;
  for (int ii = 0; ii < 2; ii++) {
    emi_t *m = mm[ii];
    const double rate = calcmrate_(om[ii]->mp[k], om[ii]->mtimeavail[k]) * m->mtimeavail[k];
    int constrain = -1;
    if (md->P - k == 2) {
      m->mp[k] = poisson_(r, rate, m->temppop == m->fpop ? 0 : 1);
    } else {
      m->mp[k] = poisson_(r, rate, m->temppop == m->fpop ? 3 : 2);
      constrain = m->fpop;
    }
    if (m->mp[k] > 0) {
      m->temppop = simmpath_(r, md, m, k, m->mp[k], m->mtimeavail[k], tstart[ii], m->temppop, constrain);
      m->mpall += m->mp[k];
    }
    if (m->mtimeavail[k] > 0 && md->pope[m->temppop] == k) m->temppop = md->popdown[m->temppop];
  }
}

/* one edge's part of getmprob: the free periods, then the last period */
static double mprob_edge(const mmodel_t *md, const emi_t *m, const emi_t *old, int lastmig) {
  double tempp = 0.0;
  int cm = 0;
  for (int k = m->b; k <= m->e; k++)
    if (k < lastmig || (k == lastmig && m->e == md->lastp)) {
      const double rate = calcmrate_(old->mp[k], old->mtimeavail[k]) * m->mtimeavail[k];
      tempp += m->mp[k] * log(rate / (m->mtimeavail[k] * (md->P - (k + 1)))) - rate;
      cm += m->mp[k];
    }
  if (m->e < md->lastp) {
    const int e = m->e;
    const double rate = calcmrate_(old->mp[e], old->mtimeavail[e]) * m->mtimeavail[e];
    if (e == md->lastp - 1) {
      tempp += (m->mp[e] & 1) ? m->mp[e] * log(rate / m->mtimeavail[e]) - logsinh__(rate)
                              : m->mp[e] * log(rate / m->mtimeavail[e]) - logcosh__(rate);
    } else {
      int p0 = cm == 0 ? m->pop : m->mig.p[cm - 1];
      while (md->pope[p0] <= e && md->pope[p0] != -1) p0 = md->popdown[p0];
      const int popc = md->P - e - 1;
      const double d = p0 == m->fpop ? log(1 - rate * exp(-rate)) : log(1 - exp(-rate));
      double nn;
      if (m->mp[e] == 0) {
        nn = -rate;
      } else if (m->mp[e] == 1) {
        nn = log(rate / m->mtimeavail[e]) - rate;
      } else {
        const int l2 = m->mp[e] == 2 ? p0 : m->mig.p[m->mpall - 3];
        const double pathc = l2 == m->fpop ? -log((double)popc) : -log((double)popc - 1);
        nn = m->mp[e] * log(rate / m->mtimeavail[e]) - rate + (2 - m->mp[e]) * log((double)popc) +
             pathc;
      }
      tempp += nn - d;
    }
  }
  return tempp;
}

/* getmprob (:851-1099) */
static double getmprob_(const mmodel_t *md, const emi_t *em, const emi_t *sm, const emi_t *oe,
                        const emi_t *os) {
  const int lastmig = em->e < md->lastp - 1 ? em->e : md->lastp - 1;
  if (sm->mtall <= 0) return mprob_edge(md, em, oe, lastmig);
  double tempp = 0.0;
  if (em->mtall > 0 && em->e < md->lastp) {
    /* the probability of the two edges' joint final population */
    int pp[2];
    const emi_t *mm[2] = {em, sm};
    for (int ii = 0; ii < 2; ii++) {
      const emi_t *m = mm[ii];
      pp[ii] = m->pop;
      if (m->mpall > 0 && m->e > 0) {
        const double t = md->tv[m->e - 1];
        int j = -1;
        while (j + 1 < m->mig.n && m->mig.t[j + 1] < t) j++;
        if (j >= 0) pp[ii] = m->mig.p[j];
      }
      if (m->e > 0)
        while (md->pope[pp[ii]] <= m->e && md->pope[pp[ii]] != -1) pp[ii] = md->popdown[pp[ii]];
    }
    if (pp[0] == pp[1])
      tempp = pp[0] == em->fpop ? log(MG_MIGCLOSEFRAC)
                                : log((1.0 - MG_MIGCLOSEFRAC) / (double)(md->P - em->e - 1));
    else if (md->P - em->e == 2)
      tempp = log(0.5);
    else if (em->fpop == pp[0] || em->fpop == pp[1])
      tempp = log(0.5 * MG_MIGCLOSEFRAC);
    else
      tempp = log((1.0 - MG_MIGCLOSEFRAC) / (double)(md->P - em->e - 2));
  }
  if (em->mtall > 0) tempp += mprob_edge(md, em, oe, lastmig);
  if (sm->mtall > 0) tempp += mprob_edge(md, sm, os, lastmig);
  return tempp;
}

static double lognormprob_(double sd, double x) {
  return -0.91893853320467274178 - log(sd) - 0.5 * (x / sd) * (x / sd);
}

int ora_gene_move_mig(int n, int *up0, int *up1, int *down, double *time, int *root, int *pop,
                      int *nmig, int cap, int *mig_edge, double *mig_time, int *mig_to, int P,
                      const int *plist, const int *npresent, const int *popdown,
                      const int *pope, const double *tv, double (*unif)(void *), void *ustate,
                      double *hastings, int *rootmove_out) {
  if (P < 2 || P > MG_MAXP || n < 2) return -1;
  usrc_t rs = {unif, ustate};
  usrc_t *r = &rs;
  mmodel_t md;
  md.P = P;
  md.lastp = P - 1;
  md.rootpop = 2 * P - 2;
  md.plist = plist;
  md.npresent = npresent;
  md.popdown = popdown;
  md.pope = pope;
  for (int k = 0; k < P - 1; k++) md.tv[k] = tv[k];
  md.tv[P - 1] = MG_TIMEMAX;
  const int ne = 2 * n - 1;
  mgene_t g = {n, ne, *root, up0, up1, down, pop, time, (mlist_t *)calloc(ne, sizeof(mlist_t))};
  for (int j = 0; j < *nmig; j++) ml_push(&g.L[mig_edge[j]], mig_time[j], mig_to[j]);
  emi_t oe, os, nw, ns;
  memset(&oe, 0, sizeof oe);
  memset(&os, 0, sizeof os);
  memset(&nw, 0, sizeof nw);
  memset(&ns, 0, sizeof ns);
  /* the edge, its parent (freed) and its sister */
  int edge;
  do edge = randposint_(r, ne);
  while (edge < 0 || edge >= ne || g.down[edge] == -1);
  const int freed = g.down[edge];
  const int oldsis = g.up0[freed] == edge ? g.up1[freed] : g.up0[freed];
  fill_old(&md, &g, edge, &oe);
  emi_reset(&os);
  if (freed == g.root) fill_old(&md, &g, oldsis, &os);
  g.L[edge].n = 0;
  double roottime = g.time[g.up0[g.root]];
  const double sd0 = fmin(20.0, roottime / 3.0);
  const double hold = normdev_(r, 0.0, sd0);
  double dist = hold;
  int rootmove = 0;
  /* joinsisdown */
  for (int i = 0; i < g.L[freed].n; i++) ml_push(&g.L[oldsis], g.L[freed].t[i], g.L[freed].p[i]);
  g.time[oldsis] = g.time[freed];
  g.down[oldsis] = g.down[freed];
  const int dd0 = g.down[oldsis];
  if (dd0 != -1) {
    if (g.up0[dd0] == freed) g.up0[dd0] = oldsis;
    else g.up1[dd0] = oldsis;
  } else {
    rootmove = 1;
    g.root = oldsis;
    g.time[oldsis] = MG_TIMEMAX;
    g.L[oldsis].n = 0;
    roottime = uptime_(&g, oldsis);
  }
  g.L[freed].n = 0;
  /* slider: the lower end of edge moves from its time; sis follows it */
  int sis = oldsis;
  double tp = g.time[edge];
  const double uplimit = uptime_(&g, edge);
  for (int step = 0; step < (1 << 20) && dist != 0.0; step++) {
    if (dist < 0) { /* up */
      const double d = -dist;
      const int su = g.up0[sis];
      if (su == -1 || uplimit >= g.time[su]) {
        if (d < tp - uplimit) {
          tp -= d;
          dist = 0.0;
        } else {
          dist = d - (tp - uplimit); /* reflect at the edge's own top: go down */
          tp = uplimit;
        }
      } else if (d < tp - g.time[su]) {
        tp -= d;
        dist = 0.0;
      } else {
        const double left = d - (tp - g.time[su]);
        tp = g.time[su];
        sis = U_(r) < 0.5 ? g.up0[sis] : g.up1[sis];
        dist = -left;
      }
    } else { /* down */
      if (g.down[sis] == -1 || tp + dist < g.time[sis]) {
        tp += dist;
        if (tp >= MG_TIMEMAX) tp = MG_TIMEMAX;
        dist = 0.0;
      } else {
        const double left = dist - (g.time[sis] - tp);
        tp = g.time[sis];
        if (U_(r) < 0.5) {
          sis = g.down[sis];
          dist = left;
        } else {
          const int dn = g.down[sis];
          sis = g.up0[dn] == sis ? g.up1[dn] : g.up0[dn];
          dist = -left;
        }
      }
    }
  }
  /* splitsisdown */
  g.time[edge] = tp;
  g.time[freed] = g.time[sis];
  g.time[sis] = tp;
  const int dd = g.down[sis];
  if (dd != -1) {
    if (g.up0[dd] == sis) g.up0[dd] = freed;
    else g.up1[dd] = freed;
  } else {
    g.root = freed;
    roottime = tp;
    rootmove = 1;
  }
  g.down[freed] = dd;
  {
    int i = 0;
    while (i < g.L[sis].n && g.L[sis].t[i] < tp) i++;
    int nowpop = i > 0 ? g.L[sis].p[i - 1] : g.pop[sis];
    settle_(&md, &nowpop, findperiod_(&md, tp));
    g.pop[freed] = nowpop;
    g.L[freed].n = 0;
    if (dd != -1)
      for (int j = i; j < g.L[sis].n; j++) ml_push(&g.L[freed], g.L[sis].t[j], g.L[sis].p[j]);
    g.L[sis].n = i;
  }
  g.down[sis] = g.down[edge] = freed;
  g.up0[freed] = sis;
  g.up1[freed] = edge;
  const double slideweight =
      rootmove ? lognormprob_(fmin(20.0, roottime / 3.0), hold) - lognormprob_(sd0, hold) : 0.0;
  /* addmigration */
  emi_reset(&nw);
  emi_reset(&ns);
  nw.edgeid = edge;
  nw.upt = uptime_(&g, edge);
  nw.pop = nw.temppop = g.pop[edge];
  nw.dnt = g.time[edge];
  fill_periods(&md, &nw);
  const int lastmig = nw.e < md.lastp - 1 ? nw.e : md.lastp - 1;
  int newsis = -1;
  if (g.down[edge] == g.root) {
    nw.fpop = -1;
    newsis = g.up0[g.root] == edge ? g.up1[g.root] : g.up0[g.root];
    ns.edgeid = newsis;
    ns.upt = uptime_(&g, newsis);
    ns.fpop = -1;
    ns.pop = ns.temppop = g.pop[newsis];
    ns.dnt = g.time[newsis];
    fill_periods(&md, &ns);
  } else {
    nw.fpop = g.pop[g.down[edge]];
  }
  /* getm (update_gtree.cpp:536-569) */
  if (ns.edgeid == -1)
    nw.mpall = mwork_single(r, &md, &nw, &oe, lastmig);
  else if (nw.mtall <= 0)
    ns.mpall = mwork_single(r, &md, &ns, &os, lastmig);
  else
    mwork_two(r, &md, &nw, &ns, &oe, &os, lastmig);
  if (nw.fpop == -1) nw.fpop = ns.fpop = md.rootpop;
  if (g.down[edge] == g.root) g.pop[g.root] = nw.fpop;
  const double denom = getmprob_(&md, &nw, &ns, &oe, &os);
  const double num = getmprob_(&md, &oe, &os, &nw, &ns);
  /* copynewmig_to_gtree */
  g.L[edge].n = 0;
  for (int i = 0; i < nw.mig.n; i++) ml_push(&g.L[edge], nw.mig.t[i], nw.mig.p[i]);
  if (ns.edgeid >= 0) {
    g.L[newsis].n = 0;
    for (int i = 0; i < ns.mig.n; i++) ml_push(&g.L[newsis], ns.mig.t[i], ns.mig.p[i]);
    int p = ns.mig.n > 0 ? ns.mig.p[ns.mig.n - 1] : g.pop[newsis];
    settle_(&md, &p, findperiod_(&md, g.time[newsis]));
    g.pop[g.down[newsis]] = p;
  }
  /* pack the events in order of edge, then time */
  int tot = 0, rc = 0;
  for (int e = 0; e < ne; e++) tot += g.L[e].n;
  if (tot > cap) {
    rc = 1;
  } else {
    int w = 0;
    for (int e = 0; e < ne; e++)
      for (int i = 0; i < g.L[e].n; i++, w++) {
        mig_edge[w] = e;
        mig_time[w] = g.L[e].t[i];
        mig_to[w] = g.L[e].p[i];
      }
    *nmig = tot;
  }
  *root = g.root;
  *hastings = (num - denom) + slideweight;
  if (rootmove_out) *rootmove_out = rootmove;
  for (int e = 0; e < ne; e++) {
    free(g.L[e].t);
    free(g.L[e].p);
  }
  free(g.L);
  emi_t *all[4] = {&oe, &os, &nw, &ns};
  for (int i = 0; i < 4; i++) {
    free(all[i]->mig.t);
    free(all[i]->mig.p);
  }
  return rc;
}
