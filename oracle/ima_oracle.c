/*
 * ima_oracle.c -- CPU restatement of the IMa2p hot path (TEST INFRASTRUCTURE).
 *
 * This file is the oracle: the HIP engine is checked against it.  It is
 * built by oracle/Makefile into oracle/_build/libima_oracle.so and loaded
 * only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * Nothing in the product path calls it.
 *
 * Each function follows the cited reference function step by step, in code
 * of our own, on plain arrays instead of the reference's global structs.
 */
#include "ima_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORA_LOG2 0.69314718055994530941723212146   /* imamp.hpp:188 */
#define ORA_LOG10 2.3025850929940456840            /* imamp.hpp:187 */
#define ORA_MYDBL_MAX (DBL_MAX / 1e10)             /* imamp.hpp:171 */
#define ORA_BESSI_MIN (-1e+100)                    /* imamp.hpp:173 */

static double branch_len(const ora_tree *t, int e) {
  int u = t->up0[e];
  return (u == -1) ? t->time[e] : t->time[e] - t->time[u];
}

/* =====================================================================
 * HKY -- calc_prob_data.cpp:118-471 (makefrac), :474-493 (getstandfactor),
 * :584-607 (likelihoodHKY); commit/rollback update_gtree_common.cpp:2140-2181
 * ===================================================================== */

double ora_getstandfactor(const double pi[4], double kappa) {
  /* calc_prob_data.cpp:474-493: sum over i != j of pi_i pi_j, times kappa
   * for transitions (purine<->purine A/G = 0/2, pyrimidine C/T = 1/3). */
  double p = 0;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      if (i == j) continue;
      if (i + j == 2 || i + j == 4)
        p += pi[i] * pi[j] * kappa;
      else
        p += pi[i] * pi[j];
    }
  return p;
}

/* HKY transition matrix P[from][to] for mutation-scaled branch length
 * (-mutrate * t) = mt, in the closed form used by makefrac
 * (calc_prob_data.cpp:197-208, 250-293): the class sum PIj and the
 * eigenvalue A = 1 + PIj (kappa - 1) are those of the destination base. */
static void hky_pmatrix(const double pi[4], double kappa, double mt,
                        double P[4][4]) {
  double kpa = kappa - 1.0;
  double expp = exp(mt);
  for (int to = 0; to < 4; to++) {
    double PIj = (to == 0 || to == 2) ? pi[0] + pi[2] : pi[1] + pi[3];
    double A = 1.0 + PIj * kpa;
    double divm = 1.0 / PIj - 1.0;
    double eA = exp(mt * A);
    for (int from = 0; from < 4; from++) {
      double v;
      if (from == to)
        v = pi[to] + pi[to] * expp * divm + eA * ((PIj - pi[to]) / PIj);
      else if (from + to == 2 || from + to == 4)
        v = pi[to] + pi[to] * expp * divm - (pi[to] / PIj) * eA;
      else
        v = pi[to] * (1.0 - expp);
      P[from][to] = v;
    }
  }
}

typedef struct {
  const ora_tree *t;
  int S;
  const int *seq;
  const double *pi;
  double mutrate, kappa;
  int e1, e2, e3, e4;
  ora_hky_state *st;
} hky_ctx;

static double *nf_row(ora_hky_state *st, int n, int edge) {
  return st->newfrac + (size_t)(edge - n) * st->S * 4;
}
static double *f_row(ora_hky_state *st, int n, int edge) {
  return st->frac + (size_t)(edge - n) * st->S * 4;
}
static double *sc_row(ora_hky_state *st, int n, int edge) {
  return st->scale + (size_t)(edge - n) * st->S;
}

/* makefrac (calc_prob_data.cpp:118-471): returns 1 when the partial at the
 * top of `node` (or below a tip) must be recomputed, i.e. node is one of
 * e1..e4 (or e1 == -1) or any descendant was recomputed. */
static int makefrac(hky_ctx *c, int node) {
  const ora_tree *t = c->t;
  int n = t->n;
  int ret = (node == c->e1 || node == c->e2 || node == c->e3 ||
             node == c->e4 || c->e1 == -1);
  int up1 = t->up0[node];
  if (up1 == -1) return ret;
  int up2 = t->up1[node];
  const double *fp1 = NULL, *fp2 = NULL;
  int r1 = makefrac(c, up1);
  if (up1 >= n) fp1 = r1 ? nf_row(c->st, n, up1) : f_row(c->st, n, up1);
  if (r1) ret = 1;
  int r2 = makefrac(c, up2);
  if (up2 >= n) fp2 = r2 ? nf_row(c->st, n, up2) : f_row(c->st, n, up2);
  if (r2) ret = 1;
  if (!ret) return 0;

  double P1[4][4], P2[4][4];
  hky_pmatrix(c->pi, c->kappa, -c->mutrate * branch_len(t, up1), P1);
  hky_pmatrix(c->pi, c->kappa, -c->mutrate * branch_len(t, up2), P2);
  double *out = nf_row(c->st, n, node);
  double *osc = sc_row(c->st, n, node);
  const double *s1 = (up1 >= n) ? sc_row(c->st, n, up1) : NULL;
  const double *s2 = (up2 >= n) ? sc_row(c->st, n, up2) : NULL;
  int S = c->S;
  for (int i = 0; i < S; i++) {
    double v1[4], v2[4], mx = 0.0;
    for (int j = 0; j < 4; j++) {
      if (fp1 == NULL) {
        v1[j] = P1[j][c->seq[(size_t)up1 * S + i]];
      } else {
        const double *f = fp1 + (size_t)i * 4;
        v1[j] = P1[j][0] * f[0] + P1[j][1] * f[1] + P1[j][2] * f[2] + P1[j][3] * f[3];
      }
      if (fp2 == NULL) {
        v2[j] = P2[j][c->seq[(size_t)up2 * S + i]];
      } else {
        const double *f = fp2 + (size_t)i * 4;
        v2[j] = P2[j][0] * f[0] + P2[j][1] * f[1] + P2[j][2] * f[2] + P2[j][3] * f[3];
      }
    }
    for (int j = 0; j < 4; j++) {
      out[i * 4 + j] = v1[j] * v2[j];
      if (out[i * 4 + j] > mx) mx = out[i * 4 + j];
    }
    for (int j = 0; j < 4; j++) out[i * 4 + j] = out[i * 4 + j] / mx;
    osc[i] = (s1 ? s1[i] : 0.0) + (s2 ? s2[i] : 0.0) + log(mx);
  }
  return 1;
}

double ora_likelihoodHKY(const ora_tree *t, int S, const int *seq,
                         const int *mult, int totsites, const double pi[4],
                         double u, double kappa, int e1, int e2, int e3, int e4,
                         ora_hky_state *st) {
  int n = t->n;
  /* calc_prob_data.cpp:593: per-site rate normalised by totsites and the
   * expected substitution rate of the unnormalised HKY matrix. */
  double mutrate = u / (totsites * ora_getstandfactor(pi, kappa));
  for (int i = n; i < 2 * n - 1; i++) nf_row(st, n, i)[0] = -1; /* :594-595 */
  hky_ctx c = {t, S, seq, pi, mutrate, kappa, e1, e2, e3, e4, st};
  makefrac(&c, t->root);
  const double *r = nf_row(st, n, t->root);
  const double *rs = sc_row(st, n, t->root);
  double p = 0;
  for (int i = 0; i < S; i++) { /* :597-605 */
    double fracp = 0;
    for (int j = 0; j < 4; j++) fracp += pi[j] * r[i * 4 + j];
    p += mult[i] * (log(fracp) + rs[i]);
  }
  return p;
}

void ora_copyfraclike(const ora_tree *t, ora_hky_state *st) {
  int n = t->n;
  for (int i = n; i < 2 * n - 1; i++) {
    double *nf = nf_row(st, n, i);
    if (nf[0] != -1) memcpy(f_row(st, n, i), nf, sizeof(double) * st->S * 4);
  }
}
void ora_storescalefactors(const ora_tree *t, ora_hky_state *st) {
  memcpy(st->oldscale, st->scale, sizeof(double) * (size_t)(t->n - 1) * st->S);
}
void ora_restorescalefactors(const ora_tree *t, ora_hky_state *st) {
  memcpy(st->scale, st->oldscale, sizeof(double) * (size_t)(t->n - 1) * st->S);
}

/* =====================================================================
 * IS -- calc_prob_data.cpp:537-581 (labelgtree), :610-719 (calc_sumlogk),
 * :732-836 (likelihoodIS)
 * ===================================================================== */

void ora_labelgtree(const ora_tree *t, int *mut, int edge) {
  /* Push the label of `edge` down toward the root, combining with the
   * sister's label (2 = conflict, -1 = not yet labelled). */
  for (;;) {
    int dow1 = t->down[edge];
    if (dow1 == -1) break;
    int sis = t->up0[dow1];
    if (sis == edge) sis = t->up1[dow1];
    int changed = 0;
    if (mut[dow1] != mut[edge]) {
      if (mut[edge] == 2) {
        mut[dow1] = (mut[sis] != -1) ? mut[sis] : 2;
      } else if (mut[sis] != -1 && mut[sis] != 2 && mut[sis] != mut[edge]) {
        mut[dow1] = 2;
      } else {
        mut[dow1] = mut[edge];
      }
      changed = 1;
    } else if (mut[dow1] == 2) {
      if (mut[sis] != -1 && mut[sis] != 2) {
        mut[dow1] = mut[sis];
        changed = 1;
      }
    }
    if (!changed) break;
    edge = dow1;
  }
}

static void label_site(const ora_tree *t, const int *seq, int S, int site,
                       int *mut) {
  int n = t->n;
  for (int j = 0; j < n; j++) mut[j] = seq[(size_t)j * S + site];
  for (int j = n; j < 2 * n - 1; j++) mut[j] = -1;
  for (int j = 0; j < n; j++) ora_labelgtree(t, mut, j);
}

static int is01(const int *mut, int a, int b) {
  return (mut[a] == 0 && mut[b] == 1) || (mut[a] == 1 && mut[b] == 0);
}

int ora_calc_sumlogk(const ora_tree *t, int S, const int *seq,
                     double *sumlogk) {
  int n = t->n;
  int *mut = (int *)malloc(sizeof(int) * (2 * n - 1));
  int *mutcount = (int *)calloc((size_t)n, sizeof(int));
  for (int site = 0; site < S; site++) {
    label_site(t, seq, S, site, mut);
    int ret = -1;
    for (int node = n; node < 2 * n - 1; node++) {
      if (is01(mut, t->up0[node], t->up1[node])) {
        if (node != ret && ret != -1) { /* :654-658 leaves the sum untouched */
          free(mut);
          free(mutcount);
          return -2;
        }
        ret = node;
        mutcount[node - n]++;
      }
    }
    if (ret == -1) {
      free(mut);
      free(mutcount);
      return -1;
    }
  }
  double sum = 0.0;
  for (int i = 0; i < n - 1; i++) { /* :710-716 log(k!) as a product */
    double fact = 1.0;
    for (int j = 1; j <= mutcount[i]; j++) fact *= (double)j;
    sum += log(fact);
  }
  *sumlogk = sum;
  free(mut);
  free(mutcount);
  return 0;
}

double ora_tree_length(const ora_tree *t) {
  double L = 0;
  for (int i = 0; i < 2 * t->n - 1; i++)
    if (t->down[i] != -1) L += branch_len(t, i);
  return L;
}

double ora_likelihoodIS(const ora_tree *t, int S, const int *seq,
                        double length, double u, double sumlogk, int *status,
                        int *site_node) {
  int n = t->n;
  int *mut = (int *)malloc(sizeof(int) * (2 * n - 1));
  double p = -length * u; /* :766 */
  *status = 0;
  for (int site = 0; site < S; site++) {
    label_site(t, seq, S, site, mut);
    int ret = -1;
    double ptime = 0;
    for (int node = n; node < 2 * n - 1; node++) {
      int up1 = t->up0[node], up2 = t->up1[node];
      if (!is01(mut, up1, up2)) continue;
      if (node != ret && ret != -1) {
        free(mut);
        return ORA_REJECT_IS; /* :784-787 */
      }
      if (t->down[node] == -1) { /* root: the two daughter branches */
        ptime = branch_len(t, up1) + branch_len(t, up2);
      } else if (mut[t->down[node]] == mut[up1]) {
        ptime = branch_len(t, up2);
      } else {
        ptime = branch_len(t, up1);
      }
      ret = node;
    }
    if (site_node) site_node[site] = ret;
    if (ret == -1) { /* reference: IM_err(IMERR_INFINITESITESFAIL) */
      *status = -1;
      free(mut);
      return ORA_REJECT_IS;
    }
    p = p + log(ptime * u);
  }
  p -= sumlogk;
  free(mut);
  return p;
}

/* =====================================================================
 * SW -- utilities.cpp:55-124 (bessi0/bessi1, Abramowitz & Stegun 9.8.1-4
 * polynomial approximations), :1453-1491 (bessi, Miller downward
 * recurrence), calc_prob_data.cpp:842-909 (likelihoodSW)
 * ===================================================================== */

double ora_bessi0(double x) {
  double ax = fabs(x), y;
  if (ax < 3.75) {
    y = x / 3.75;
    y *= y;
    return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 +
           y * (0.2659732 + y * (0.360768e-1 + y * 0.45813e-2)))));
  }
  y = 3.75 / ax;
  return (exp(ax) / sqrt(ax)) *
         (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 +
          y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 +
          y * (0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2))))))));
}

double ora_bessi1(double x) {
  double ax = fabs(x), y, ans;
  if (ax < 3.75) {
    y = x / 3.75;
    y *= y;
    ans = ax * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 +
          y * (0.2658733e-1 + y * (0.301532e-2 + y * 0.32411e-3))))));
  } else {
    y = 3.75 / ax;
    ans = 0.2282967e-1 + y * (-0.2895312e-1 + y * (0.1787654e-1 - y * 0.420059e-2));
    ans = 0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 +
          y * (0.163801e-2 + y * (-0.1031555e-1 + y * ans))));
    ans *= (exp(ax) / sqrt(ax));
  }
  return x < 0.0 ? -ans : ans;
}

double ora_bessi(int n, double x) {
  const double ACC = 40.0, BIGNO = 1.0e10, BIGNI = 1.0e-10;
  n = abs(n);
  if (x > 700) return ORA_MYDBL_MAX;
  if (n == 0) return ora_bessi0(x);
  if (n == 1) return ora_bessi1(x);
  if (x == 0.0) return 0.0;
  double tox = 2.0 / fabs(x), bip = 0.0, ans = 0.0, bi = 1.0, bim;
  for (int j = 2 * (n + (int)sqrt(ACC * n)); j > 0; j--) {
    bim = bip + j * tox * bi;
    bip = bi;
    bi = bim;
    if (fabs(bi) > BIGNO) {
      ans *= BIGNI;
      bi *= BIGNI;
      bip *= BIGNI;
    }
    if (j == n) ans = bip;
  }
  ans *= ora_bessi0(x) / bi;
  return (x < 0.0 && (n & 1)) ? -ans : ans;
}

double ora_likelihoodSW(const ora_tree *t, const int *A, double u, double tr,
                        double *dlikeA) {
  double like = 0.0;
  int iszero = 0;
  for (int i = 0; i < 2 * t->n - 1; i++) {
    if (t->down[i] == -1) {
      if (dlikeA) dlikeA[i] = 0.0;
      continue;
    }
    int d = A[i] - A[t->down[i]];
    double tt = (t->up0[i] == -1) ? t->time[i] * tr
                                  : tr * (t->time[i] - t->time[t->up0[i]]);
    double b = ora_bessi(d, tt * u);
    double v;
    if (!(b > 0.0)) {
      iszero = 1;
      v = ORA_BESSI_MIN;
    } else {
      v = -(tt * u) + log(b);
    }
    if (dlikeA) dlikeA[i] = v;
    like += v;
  }
  if (iszero) like = -DBL_MAX;
  return like;
}

/* =====================================================================
 * finishSWupdateA -- update_gtree_common.cpp:2219-2362
 * ===================================================================== */

static double ora_upper_time(const ora_tree *t, int e) {
  return e >= t->n ? t->time[t->up0[e]] : 0.0; /* :2247-2256 */
}

double ora_finishSWupdateA(const ora_tree *t, int *A, const ora_sw_move *m, int calc,
                           double *Aterm, int *newA_out, double *edge_dlikeA) {
  const int oldA = A[m->downedge];
  /* holdsisdlikeA (:2232-2235) and the old adjustment (:2240-2242) */
  const double holdsis = m->newsisedge != m->sisedge ? m->newsis_dlikeA : 0.0;
  const double oldlikeadj = m->copy_dlikeA[0] + m->copy_dlikeA[1] + m->copy_dlikeA[2] + holdsis;
  int e[3];
  double tt[3] = {0.0, 0.0, 0.0};
  e[0] = m->edge;
  e[1] = m->newsisedge;
  e[2] = t->down[e[0]];
  for (int i = 0; i < 3; i++)
    if (t->down[e[i]] != -1) tt[i] = t->time[e[i]] - ora_upper_time(t, e[i]);
  /* forward geometric parameter from the neighbours of the node (:2264-2278) */
  int wsum = 0, j = 0;
  for (int i = 0; i < 3; i++) {
    if (t->down[e[i]] == -1) continue;
    const int d = (i < 2 ? A[e[i]] : A[t->down[e[i]]]) - oldA;
    wsum += abs(d);
    j++;
  }
  double gnew = j / (double)(wsum + j);
  if (gnew > 0.95) gnew = 0.95;
  int dA = (int)ceil(log(m->uni) / log(1.0 - gnew)) - 1;
  if (m->flip) dA = -dA;
  int newA;
  if (dA >= 0)
    newA = oldA + dA < m->maxA ? oldA + dA : m->maxA;
  else
    newA = oldA + dA > m->minA ? oldA + dA : m->minA;
  A[m->downedge] = newA;
  dA = newA - oldA;
  /* the old sister now runs into the old downedge's parent (:2294-2313) */
  double sis_dl = 0.0;
  if (t->down[m->sisedge] != -1 && m->sisedge != m->newsisedge) {
    const double ts = t->time[m->sisedge] - ora_upper_time(t, m->sisedge);
    const int d = A[m->sisedge] - A[t->down[m->sisedge]];
    sis_dl = -ts * m->u + log(ora_bessi(d, ts * m->u));
  }
  double likeadj = sis_dl;
  double dl[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < 3; i++) {
    if (t->down[e[i]] == -1) continue;
    const int d = (i < 2 ? A[e[i]] : A[t->down[e[i]]]) - newA;
    dl[i] = -(tt[i] * m->u) + log(ora_bessi(d, tt[i] * m->u));
    likeadj += dl[i];
  }
  /* reverse geometric parameter from the node's neighbours before the move
   * (:2327-2343) */
  wsum = abs(m->copyA0 - newA) + abs(m->copyA1 - newA);
  j = 2;
  if (!m->down_was_root) {
    wsum += abs(m->hold_down_A - newA);
    j++;
  }
  double gold = j / (double)(wsum + j);
  if (gold > 0.95) gold = 0.95;
  if (edge_dlikeA) {
    edge_dlikeA[0] = m->sisedge == m->newsisedge ? dl[1] : sis_dl;
    edge_dlikeA[1] = dl[0];
    edge_dlikeA[2] = dl[1];
    edge_dlikeA[3] = dl[2];
  }
  if (newA_out) *newA_out = newA;
  if (!calc) {
    *Aterm = 0.0;
    return 0.0;
  }
  *Aterm = (abs(dA) * log(1 - gold) + log(gold)) - (abs(dA) * log(1 - gnew) + log(gnew));
  return likeadj - oldlikeadj;
}

/* =====================================================================
 * MC3 swap -- swapchains.cpp:13-34 (swapweight), :72-178 (setheat)
 * ===================================================================== */

double ora_swapweight(double beta_i, double beta_j, double sum_i,
                      double sum_j) {
  return exp((beta_i - beta_j) * (sum_j - sum_i));
}

double ora_setheat_geometric(int i, int nchains, double hval1, double hval2) {
  return 1 - (1 - hval2) * (i) * pow(hval1, (double)(nchains - 1 - i)) /
                 (double)(nchains - 1);
}

double ora_setheat_linear(int i, double hval1) {
  if (hval1 < 0.05) hval1 = 0.05;
  return 1.0 / (1.0 + hval1 * i);
}

/* =====================================================================
 * L mode -- utilities.cpp:1507-1539 (eexp), jointfind.cpp:885-1047 (jointp),
 * surface_call_functions.cpp:120-173 (margincalc)
 * ===================================================================== */

static const double eexp_us[10] = {
    1.0, 0.5, 0.16666666666666666666666666667, 0.04166666666666666666666666667,
    0.00833333333333333333333333333, 0.001388888888888888888888888889,
    0.000198412698412698412698412698, 0.000024801587301587301587301587301,
    2.75573192239858906525573192239859e-6, 2.75573192239858906525573192239e-7};

void ora_eexp(double x, double *m, int *z) {
  /* e^x = m * 10^z with a Taylor series for e^u, u = x - n ln2 in [0, ln2) */
  const double LOG_10_2 = 0.30102999566398119521;
  int n = (int)floor(x / ORA_LOG2);
  double zr = LOG_10_2 * (double)n;
  *z = (int)zr;
  zr -= (double)*z;
  double u = x - (((double)n) * ORA_LOG2);
  double temp = 1 + u * (eexp_us[0] + u * (eexp_us[1] + u * (eexp_us[2] +
                u * (eexp_us[3] + u * (eexp_us[4] + u * (eexp_us[5] +
                u * (eexp_us[6] + u * (eexp_us[7] + u * (eexp_us[8] +
                u * (eexp_us[9]))))))))));
  *m = temp * pow(10.0, zr);
  if (fabs(*m) > 10) {
    *m = *m / 10.0;
    *z = *z + 1;
  }
  if (fabs(*m) < 1) {
    *m = *m * 10.0;
    *z = *z - 1;
  }
}

/* Per-genealogy log term of jointp (jointfind.cpp:966-995). */
static double jointp_term(const float *g, int npsize, int nmig, int npops,
                          int modeltype, int r0, int r1, const double *x,
                          const double *log2diffx, const double *logx,
                          const double *divx) {
  int ccp = 0, fcp = ccp + npsize, hccp = fcp + npsize, mcp = hccp + npsize;
  int fmp = mcp + nmig, qip = fmp + nmig, mip = qip + npsize;
  int pdgp = mip + nmig, probgp = pdgp + 1;
  double p;
  if (npops == 2) {
    p = -g[probgp];
  } else {
    double s = 0;
    if (modeltype == 1)
      for (int i = 0; i < npsize; i++) s += g[qip + i];
    else
      for (int i = 0; i < nmig; i++) s += g[mip + i];
    p = -s;
  }
  for (int i = r0; i < r1; i++) {
    if (i < npsize) {
      double dg = 2.0 * g[fcp + i];
      p += g[ccp + i] * log2diffx[i] - g[hccp + i] - dg * divx[i];
    } else {
      int i1 = i - npsize;
      p += g[mcp + i1] * logx[i] - g[fmp + i1] * x[i];
    }
  }
  return p;
}

static int cmp_desc(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x < y) - (x > y);
}

double ora_jointp(const float *table, int G, int K, int npsize, int nmig,
                  int npops, int modeltype, int r0, int r1, const double *x,
                  int calc_ess, double *effective_n) {
  int np = npsize + nmig;
  double *log2diffx = (double *)calloc((size_t)np + 1, sizeof(double));
  double *logx = (double *)calloc((size_t)np + 1, sizeof(double));
  double *divx = (double *)calloc((size_t)np + 1, sizeof(double));
  if (modeltype == 0 || modeltype == 1)
    for (int i = 0; i < npsize; i++) log2diffx[i] = ORA_LOG2 - log(x[i]);
  for (int i = r0; i < r1; i++) {
    logx[i] = log(x[i]);
    divx[i] = 1.0 / x[i];
  }
  /* The reference keeps a sorted list of the values that lie within
   * PRANGELOG = 10 of the running maximum at the time they are seen
   * (jointfind.cpp:997-1009; the first value is always kept).  It then walks
   * down from the maximum, taking at most (list size - 1) values while they
   * stay within 10 of the maximum (:1010-1022), and sums them as eexp
   * mantissa/exponent pairs (:1016-1044). */
  double *kept = (double *)malloc(sizeof(double) * (size_t)G);
  int nkept = 0;
  double runmax = -INFINITY;
  for (int gi = 0; gi < G; gi++) {
    double p = jointp_term(table + (size_t)gi * K, npsize, nmig, npops,
                           modeltype, r0, r1, x, log2diffx, logx, divx);
    if (gi == 0 || runmax - p < 10) {
      kept[nkept++] = p;
      if (p > runmax) runmax = p;
    }
  }
  qsort(kept, (size_t)nkept, sizeof(double), cmp_desc);
  int iin = nkept - 1;
  int gin = 0;
  double top = kept[0];
  double *mm = (double *)malloc(sizeof(double) * (size_t)nkept);
  int *zz = (int *)malloc(sizeof(int) * (size_t)nkept);
  int maxz = -10000000;
  do {
    ora_eexp(kept[gin], &mm[gin], &zz[gin]);
    if (zz[gin] > maxz) maxz = zz[gin];
    gin++;
  } while (gin < iin && top - kept[gin] < 10);
  maxz -= 10;
  double acumm = 0, acumm_sqr = 0;
  for (int gi = 0; gi < gin; gi++) {
    int zadj = zz[gi] - maxz;
    double m = mm[gi];
    if (zadj > -308 && zadj < 308)
      m *= pow(10.0, (double)zadj);
    else
      m = (zadj <= -308) ? 0.0 : DBL_MAX;
    acumm += m;
    if (calc_ess) acumm_sqr += m * m;
  }
  if (calc_ess && effective_n) *effective_n = acumm * acumm / acumm_sqr;
  double sum = log(acumm) + maxz * ORA_LOG10;
  double q = log((double)G) - sum;
  free(kept);
  free(mm);
  free(zz);
  free(log2diffx);
  free(logx);
  free(divx);
  return q;
}

double ora_jointp_exact(const float *table, int G, int K, int npsize, int nmig,
                        int npops, int modeltype, int r0, int r1,
                        const double *x) {
  int np = npsize + nmig;
  double *log2diffx = (double *)calloc((size_t)np + 1, sizeof(double));
  double *logx = (double *)calloc((size_t)np + 1, sizeof(double));
  double *divx = (double *)calloc((size_t)np + 1, sizeof(double));
  for (int i = 0; i < npsize; i++) log2diffx[i] = ORA_LOG2 - log(x[i]);
  for (int i = r0; i < r1; i++) {
    logx[i] = log(x[i]);
    divx[i] = 1.0 / x[i];
  }
  double mx = -INFINITY, s = 0;
  for (int gi = 0; gi < G; gi++) {
    double p = jointp_term(table + (size_t)gi * K, npsize, nmig, npops,
                           modeltype, r0, r1, x, log2diffx, logx, divx);
    if (p > mx) {
      s = s * exp(mx - p) + 1.0;
      mx = p;
    } else {
      s += exp(p - mx);
    }
  }
  free(log2diffx);
  free(logx);
  free(divx);
  return log((double)G) - (mx + log(s));
}

static long int_round(double x) {
  return (x >= 0) ? (long)(x + 0.5) : (long)(x - 0.5); /* imamp.hpp:180 */
}

double ora_margincalc(const float *table, int G, int K, int npsize, int nmig,
                      double x, double yadjust, int pi, int logi,
                      int expo_prior, double prior_mean) {
  int ccp = 0, fcp = ccp + npsize, hccp = fcp + npsize, mcp = hccp + npsize;
  int fmp = mcp + nmig, qip = fmp + nmig, mip = qip + npsize;
  double sum = 0, meani = 0;
  (void)K;
  if (expo_prior && pi >= npsize && pi < npsize + nmig) meani = 1.0 / prior_mean;
  for (int ei = 0; ei < G; ei++) {
    const float *g = table + (size_t)ei * K;
    if (pi < npsize) {
      int p = pi;
      double hval = g[hccp + p];
      double temp = -g[qip + p] +
                    int_round(g[ccp + p]) * (ORA_LOG2 - log(x)) - hval -
                    2 * g[fcp + p] / x;
      sum += exp(temp);
    } else {
      int p = pi - npsize;
      if (expo_prior)
        sum += exp(-g[mip + p] + log(meani) - x * meani +
                   int_round(g[mcp + p]) * log(x) - g[fmp + p] * x);
      else
        sum += exp(-g[mip + p] + int_round(g[mcp + p]) * log(x) -
                   g[fmp + p] * x);
    }
  }
  sum /= G;
  if (logi) sum = (sum <= 0) ? -1e200 : log(sum);
  sum -= yadjust;
  return sum;
}

/* =====================================================================
 * swapchains (swapchains.cpp:527-653), serial within-process form, drawing
 * its uniforms from uni[] in order (the reference calls uniform()).
 * Returns the number of accepted swaps; *used = uniforms consumed;
 * pairs (optional, 3 per try) = (ci, cj, accepted).  -1 when uni runs out.
 * ===================================================================== */
int ora_swapchains(int nchains, const double *S, double *beta, int swaptries,
                   const double *uni, int nuni, int *used, int *pairs) {
  const int SWAPDIST = 7;
  int k = 0, acc = 0;
#define NEXTU() (k < nuni ? uni[k++] : -1.0)
  for (int i = 0; i < swaptries; i++) {
    int ci, cj, cjmin, cjrange;
    do {
      double u = NEXTU();
      if (u < 0) return -1;
      ci = (int)(u * nchains);
    } while (ci < 0 || ci >= nchains);
    if (nchains < 2 * SWAPDIST + 3) {
      cjmin = 0;
      cjrange = nchains;
    } else {
      cjmin = (ci - SWAPDIST > 0) ? ci - SWAPDIST : 0;
      cjrange = ((nchains < ci + SWAPDIST) ? nchains : ci + SWAPDIST) - cjmin;
    }
    do {
      double u = NEXTU();
      if (u < 0) return -1;
      cj = cjmin + (int)(u * cjrange);
    } while (cj == ci || cj < 0 || cj >= nchains);
    double w = ora_swapweight(beta[ci], beta[cj], S[ci], S[cj]);
    int ok = 0;
    if (w >= 1.0) {
      ok = 1;
    } else {
      double u = NEXTU();
      if (u < 0) return -1;
      ok = w > u;
    }
    if (ok) {
      double b = beta[ci];
      beta[ci] = beta[cj];
      beta[cj] = b;
      acc++;
    }
    if (pairs) {
      pairs[3 * i] = ci;
      pairs[3 * i + 1] = cj;
      pairs[3 * i + 2] = ok;
    }
  }
#undef NEXTU
  if (used) *used = k;
  return acc;
}

/* =====================================================================
 * Genealogy prior -- treeweight (update_gtree_common.cpp:1680-1931),
 * integrate_tree_prob (:1945-2053), integrate_coalescent_term (:109-203),
 * integrate_migration_term (:206-296), incomplete gamma functions
 * (utilities.cpp:812-1122: series and continued fraction with EPS = 3e-7,
 * ITMAX = 1000, FPMIN = 1e-30; exponential integral MAXIT = 100).
 * ===================================================================== */

#define PRI_EPS 3.0e-7
#define PRI_ITMAX 1000
#define PRI_FPMIN 1.0e-30

static double logfact_of(int k) { /* setlogfact: running sum of log(i) */
  double s = 0.0;
  for (int i = 1; i <= k; i++) s = s + log((double)i);
  return s;
}

double ora_logfact(int k) { return logfact_of(k); }

/* continued fraction for the upper incomplete gamma (log form) */
static double gcf_log(double a, double x, double *gln) {
  *gln = logfact_of((int)a - 1);
  double b = x + 1.0 - a, c = 1.0 / PRI_FPMIN, d = 1.0 / b, h = d;
  int i;
  for (i = 1; i <= PRI_ITMAX; i++) {
    double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < PRI_FPMIN) d = PRI_FPMIN;
    c = b + an / c;
    if (fabs(c) < PRI_FPMIN) c = PRI_FPMIN;
    d = 1.0 / d;
    double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < PRI_EPS) break;
  }
  return (-x + a * log(x) - (*gln)) + log(h);
}

static double gcf_lin(double a, double x, double *gln) {
  *gln = logfact_of((int)a - 1);
  double b = x + 1.0 - a, c = 1.0 / PRI_FPMIN, d = 1.0 / b, h = d;
  int i;
  for (i = 1; i <= PRI_ITMAX; i++) {
    double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < PRI_FPMIN) d = PRI_FPMIN;
    c = b + an / c;
    if (fabs(c) < PRI_FPMIN) c = PRI_FPMIN;
    d = 1.0 / d;
    double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < PRI_EPS) break;
  }
  return exp(-x + a * log(x) - (*gln)) * h;
}

/* series for the lower incomplete gamma (linear and log forms) */
static double gser_lin(int a, double x, double *gln) {
  *gln = logfact_of(a - 1);
  if (x <= 0.0) return 0.0;
  double ap = a, del = 1.0 / a, sum = del;
  for (int n = 1; n <= PRI_ITMAX; n++) {
    ++ap;
    del *= x / ap;
    sum += del;
    if (fabs(del) < fabs(sum) * PRI_EPS) return sum * exp(-x + a * log(x) - (*gln));
  }
  return sum * exp(-x + a * log(x) - (*gln));
}

static double gser_log(int a, double x, double *gln) {
  *gln = logfact_of(a - 1);
  if (x <= 0.0) return 0.0;
  double ap = a, del = 1.0 / a, sum = del;
  for (int n = 1; n <= PRI_ITMAX; n++) {
    ++ap;
    del *= x / ap;
    sum += del;
    if (fabs(del) < fabs(sum) * PRI_EPS) return log(sum) + (-x + a * log(x) - (*gln));
  }
  return log(sum) + (-x + a * log(x) - (*gln));
}

/* exponential integral E_n(x) (value, or its log when *islog = 1) */
static double expint_n(int n, double x, int *islog) {
  const double EULER = 0.5772156649;
  *islog = 0;
  int nm1 = n - 1;
  if (n == 0) return exp(-x) / x;
  if (x == 0.0) return 1.0 / nm1;
  if (x > 1.0) {
    double b = x + n, c = 1.0 / PRI_FPMIN, d = 1.0 / b, h = d;
    for (int i = 1; i <= 100; i++) {
      double a = -i * (nm1 + i);
      b += 2.0;
      d = 1.0 / (a * d + b);
      c = b + a / c;
      double del = c * d;
      h *= del;
      if (fabs(del - 1.0) < PRI_EPS) {
        *islog = 1;
        return log(h) - x;
      }
    }
    *islog = 1;
    return log(h) - x;
  }
  double ans = (nm1 != 0 ? 1.0 / nm1 : -log(x) - EULER), fact = 1.0;
  for (int i = 1; i <= 100; i++) {
    double del;
    fact *= -x / i;
    if (i != nm1) {
      del = -fact / (i - nm1);
    } else {
      double psi = -EULER;
      for (int ii = 1; ii <= nm1; ii++) psi += 1.0 / ii;
      del = fact * (-log(x) + psi);
    }
    ans += del;
    if (fabs(del) < fabs(ans) * PRI_EPS) return ans;
  }
  return ans;
}

double ora_uppergamma(int a, double x) {
  double gln, p;
  if (a == 0) {
    int islog;
    double t = expint_n(1, x, &islog);
    p = islog ? t : log(t);
  } else if (x < (a + 1.0)) {
    double g = gser_lin(a, x, &gln);
    p = gln + log(1.0 - g);
  } else {
    double g = gcf_log((double)a, x, &gln);
    p = gln + g;
  }
  if (p < -1e200) p = -1e200;
  return p;
}

double ora_lowergamma(int a, double x) {
  double gln, p;
  if (x < (a + 1.0)) {
    double g = gser_log(a, x, &gln);
    p = gln + g;
  } else {
    double g = gcf_lin((double)a, x, &gln);
    p = gln + log(1 - g);
  }
  if (p < -1e200) p = -1e200;
  return p;
}

static double log_diff(double a, double b) { /* LogDiff imamp.hpp:257-263 */
  if (a - b < 7.0978271289338397e+02) return b + log(exp(a - b) - 1.0);
  return a;
}

double ora_integrate_coalescent_term(int cc, double fc, double hcc, double max, double min) {
  double p;
  if (cc > 0) {
    if (min == 0) {
      double ug = ora_uppergamma(cc - 1, 2 * fc / max);
      if (cc > 1) {
        double fullg = logfact_of(cc - 2);
        if (fullg - ug < 1e-15 || fullg - ug > 7.0978271289338397e+02) {
          double lg = ora_lowergamma(cc - 1, 2 * fc / max);
          if (fullg > lg) {
            double ugalt = log_diff(fullg, lg);
            if (fabs(ugalt - ug) > 1e-10) ug = ugalt;
          }
        }
      }
      p = ug + ORA_LOG2 - hcc + (1 - cc) * log(fc);
    } else {
      double a = ora_uppergamma(cc - 1, 2 * fc / max);
      double b = ora_uppergamma(cc - 1, 2 * fc / min);
      p = log_diff(a, b) + (ORA_LOG2 - hcc + (1 - cc) * log(fc));
    }
  } else {
    if (2 * fc / max > 0) {
      if (min == 0) {
        double a = log(max) - 2.0 * fc / max;
        double b = ORA_LOG2 + log(fc) + ora_uppergamma(0, 2.0 * fc / max);
        p = log_diff(a, b);
      } else {
        double a = ora_uppergamma(0, 2 * fc / max), b = ora_uppergamma(0, 2 * fc / min);
        double c = log_diff(a, b) + ORA_LOG2 + log(fc);
        a = log(max) - 2.0 * fc / max;
        b = log(min) - 2.0 * fc / min;
        double d = log_diff(a, b);
        p = log_diff(d, c);
      }
    } else {
      p = log(max - min);
    }
  }
  return p;
}

double ora_integrate_migration_term(int cm, double fm, double max, double min) {
  double p;
  const double MPRIORMIN = 0.000001; /* imamp.hpp:141 */
  if (cm > 0) {
    if (min == 0) {
      double lg = ora_lowergamma(cm + 1, fm * max);
      double fullg = logfact_of(cm);
      if (fullg - lg < 1e-15 || fullg - lg > 7.0978271289338397e+02) {
        double ug = ora_uppergamma(cm + 1, fm * max);
        if (fullg > ug) {
          double lgalt = log_diff(fullg, ug);
          if (fabs(lgalt - lg) > 1e-12) lg = lgalt;
        }
      }
      p = (-1 - cm) * log(fm) + lg;
    } else {
      double a = ora_uppergamma(cm + 1, fm * min), b = ora_uppergamma(cm + 1, fm * max);
      p = (-1 - cm) * log(fm) + log_diff(a, b);
    }
  } else {
    if (fm > MPRIORMIN) {
      if (min == 0) {
        if (max == MPRIORMIN)
          p = 0;
        else
          p = log_diff(0.0, -fm * max) - log(fm);
      } else {
        p = log_diff(-fm * min, -fm * max) - log(fm);
      }
    } else {
      p = log(max - min);
    }
  }
  return p;
}

/* treeweight over one genealogy.  The model is flattened: periods
 * k = 0..nper-1, plist[k*P + i] = i-th population present in period k
 * (npresent[k] of them), popdown[p] / pope[p] = parent population and the
 * period where p ends (-1 for the root population), join[k] = {a, b, new}
 * for the split closing period k.  mig_* hold the migration events of the
 * genealogy sorted by edge then time: edge, time, destination.  Outputs
 * cc[k*P+i] (int), fc[k*P+i], hcc, mc[(k*P+i)*P+j] (int), fm (same). */
typedef struct {
  double t;
  int type; /* 0 coalescence, 1 migration, -1 split */
  int pop, topop;
} pri_event;

static int cmp_event(const void *a, const void *b) {
  double x = ((const pri_event *)a)->t, y = ((const pri_event *)b)->t;
  return (x > y) - (x < y);
}

int ora_treeweight(const ora_tree *t, const int *pop, int nmig, const int *mig_edge,
                   const double *mig_time, const int *mig_to, int npops, int nper,
                   const int *plist, const int *npresent, const int *popdown,
                   const int *pope, const int *join, const double *tvals,
                   const int *samppop, double hval, int *cc, double *fc, double *hcc,
                   int *mc, double *fm, double *length) {
  const int P = npops;
  int ng = t->n, ne = 2 * ng - 1;
  double roottime = t->time[t->up0[t->root]];
  int kroot = 0;
  while (kroot < nper - 1 && tvals[kroot] <= roottime) kroot++;
  int nev = (ng - 1) + nmig + kroot;
  pri_event *ev = (pri_event *)malloc(sizeof(pri_event) * (size_t)(nev + 1));
  int ec = 0, mi = 0;
  for (int i = 0; i < ne; i++) {
    int nowpop = pop[i];
    if (i >= ng) {
      ev[ec].t = t->time[t->up0[i]];
      ev[ec].type = 0;
      ev[ec].pop = nowpop;
      ev[ec].topop = -1;
      ec++;
    }
    while (mi < nmig && mig_edge[mi] == i) {
      double mt = mig_time[mi];
      int period = 0;
      while (period < nper - 1 && tvals[period] <= mt) period++;
      while (pope[nowpop] != -1 && pope[nowpop] <= period) nowpop = popdown[nowpop];
      ev[ec].t = mt;
      ev[ec].type = 1;
      ev[ec].pop = nowpop;
      ev[ec].topop = mig_to[mi];
      nowpop = mig_to[mi];
      ec++;
      mi++;
    }
  }
  for (int k = 0; k < kroot; k++) {
    ev[ec].t = tvals[k];
    ev[ec].type = -1;
    ev[ec].pop = ev[ec].topop = -1;
    ec++;
  }
  if (ec != nev || mi != nmig) {
    free(ev);
    return -1;
  }
  qsort(ev, (size_t)nev, sizeof(pri_event), cmp_event);
  int n[2 * 10 - 1];
  for (int p = 0; p < 2 * P - 1; p++) n[p] = (p < P) ? samppop[p] : 0;
  memset(cc, 0, sizeof(int) * (size_t)nper * P);
  memset(fc, 0, sizeof(double) * (size_t)nper * P);
  memset(hcc, 0, sizeof(double) * (size_t)nper * P);
  memset(mc, 0, sizeof(int) * (size_t)nper * P * P);
  memset(fm, 0, sizeof(double) * (size_t)nper * P * P);
  int nsum = ng, k = 0;
  double last = 0.0, len = 0.0, h2term = 1 / (2 * hval);
  for (int j = 0; j < nev; j++) {
    double dt = ev[j].t - last;
    len += nsum * dt;
    last = ev[j].t;
    for (int ii = 0; ii < npresent[k]; ii++) {
      int ip = plist[k * P + ii];
      fc[k * P + ii] += ((double)n[ip] * ((double)n[ip] - 1)) * dt * h2term;
      if (k < nper - 1) {
        double fmtemp = n[ip] * dt;
        for (int jj = 0; jj < npresent[k]; jj++)
          if (jj != ii) fm[(k * P + ii) * P + jj] += fmtemp;
      }
    }
    if (ev[j].type == 0) {
      int ii = 0;
      while (ii < npresent[k] && plist[k * P + ii] != ev[j].pop) ii++;
      if (ii == npresent[k]) { /* population absent from the period */
        free(ev);
        return -2;
      }
      cc[k * P + ii]++;
      n[ev[j].pop]--;
      nsum--;
    } else if (ev[j].type == 1) {
      int ii = 0, jj = 0;
      while (ii < npresent[k] && plist[k * P + ii] != ev[j].pop) ii++;
      while (jj < npresent[k] && plist[k * P + jj] != ev[j].topop) jj++;
      if (ii == npresent[k] || jj == npresent[k]) {
        free(ev);
        return -2;
      }
      mc[(k * P + ii) * P + jj]++;
      n[ev[j].pop]--;
      n[ev[j].topop]++;
    } else {
      const int *jn = join + 3 * k;
      k++;
      n[jn[2]] = n[jn[0]] + n[jn[1]];
      n[jn[0]] = n[jn[1]] = 0;
    }
  }
  double hlog = log(hval);
  if (hlog != 0.0)
    for (int kk = 0; kk < nper; kk++)
      for (int i = 0; i < npresent[kk]; i++) hcc[kk * P + i] += hlog * cc[kk * P + i];
  *length = len;
  free(ev);
  return 0;
}

/* ---------------------------------------------------------------------------
 * 2NM density -- popmig.cpp:9-351 (calc_popmig, calc_pop_expomig,
 * marginpopmig, marginpop_expomig) and marginalopt_popmig (:361-429) with
 * goldenmod (surface_search_functions.cpp:134-184).
 */
typedef struct {
  int ccp, fcp, hcp, mcp, fmp, qip, mip;
} pm_cols;

static pm_cols pm_columns(int npsize, int nmig, int thetai, int mi) {
  int ccp = 0, fcp = ccp + npsize, hccp = fcp + npsize, mcp = hccp + npsize;
  int fmp = mcp + nmig, qip = fmp + nmig, mip = qip + npsize;
  pm_cols c = {ccp + thetai, fcp + thetai, hccp + thetai, mcp + mi, fmp + mi, qip + thetai,
               mip + mi};
  return c;
}

/* sum over rows [first, last) of the uniform-prior terms (popmig.cpp:27-85) */
static double pm_sum_uniform(const float *table, int K, pm_cols c, int first, int last,
                             double x, double qmax, double mmax) {
  double sum = 0;
  for (int ei = first; ei < last; ei++) {
    const float *g = table + (size_t)ei * K;
    int cc = (int)g[c.ccp];
    double fc = g[c.fcp], hc = g[c.hcp];
    int mc = (int)g[c.mcp];
    double fm = g[c.fmp], qintg = g[c.qip], mintg = g[c.mip], temp1, temp2;
    if (fc == 0 && cc == 0 && fm > 0) {
      temp1 = ORA_LOG2 - (mc * log(fm)) - hc - qintg - mintg;
      temp2 = log(exp(ora_uppergamma(mc, 2 * fm * x / qmax)) -
                  exp(ora_uppergamma(mc, mmax * fm)));
    } else if (fm == 0 && mc == 0 && fc > 0) {
      temp1 = ORA_LOG2 - (cc * log(fc)) - hc - qintg - mintg;
      temp2 = log(exp(ora_uppergamma(cc, 2 * fc / qmax)) -
                  exp(ora_uppergamma(cc, fc * mmax / x)));
    } else if (fc == 0 && cc == 0 && mc == 0 && fm == 0) {
      temp1 = log(2 * log(mmax * qmax / (2 * x))) - hc - qintg - mintg;
      temp2 = 0;
    } else {
      temp1 = ORA_LOG2 + (mc * log(x)) - ((cc + mc) * log(fc + fm * x)) - hc - qintg - mintg;
      double a = ora_uppergamma(cc + mc, 2 * (fc + fm * x) / qmax);
      double b = ora_uppergamma(cc + mc, mmax * (fm + fc / x));
      if (a == b) {
        b = ora_lowergamma(cc + mc, 2 * (fc + fm * x) / qmax);
        a = ora_lowergamma(cc + mc, mmax * (fm + fc / x));
      }
      if (a > b)
        temp2 = log_diff(a, b);
      else
        temp1 = temp2 = 0.0;
    }
    if ((temp1 + temp2 < 700) && (temp1 + temp2 > -700)) sum += exp(temp1 + temp2);
  }
  return sum;
}

/* log-sum of the exponential-prior terms over rows [first, last) as the
 * reference sums them: eexp per term, scaled to the largest decade less
 * OCUTOFF = 10 (popmig.cpp:105-164) */
static double pm_logsum_expo(const float *table, int K, pm_cols c, int first, int last,
                             double x, double qmax, double mmean) {
  int n = last - first, maxz = -1000000000;
  double *m = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  int *z = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int ei = first; ei < last; ei++) {
    const float *g = table + (size_t)ei * K;
    int cc = (int)g[c.ccp];
    double fc = g[c.fcp], hc = g[c.hcp];
    int mc = (int)g[c.mcp];
    double fm = g[c.fmp], qintg = g[c.qip], mintg = g[c.mip];
    double temp1 = x + fc * mmean + fm * mmean * x;
    double temp2 = ORA_LOG2 - hc - log(mmean) - qintg - mintg;
    double temp3 = 2 * temp1 / (mmean * qmax);
    if (mc == 0 && cc == 0) {
      temp2 += ora_uppergamma(0, temp3);
    } else if (cc == 0) {
      temp2 += ora_uppergamma(mc, temp3) - mc * log(fm + 1 / mmean);
    } else if (mc == 0) {
      temp2 += ora_uppergamma(cc, temp3) + -cc * log(fc + x * (fm + 1 / mmean));
    } else {
      temp2 += ora_uppergamma(mc + cc, temp3) - cc * log(x) +
               (cc + mc) * log(mmean * x / temp1);
    }
    ora_eexp(temp2, &m[ei - first], &z[ei - first]);
    if (z[ei - first] > maxz) maxz = z[ei - first];
  }
  double acumm = 0;
  for (int i = 0; i < n; i++) {
    int zadj = z[i] - (maxz - 10);
    acumm += m[i] * pow(10.0, (double)zadj);
  }
  free(m);
  free(z);
  return log(acumm) + (maxz - 10) * ORA_LOG10;
}

double ora_popmig(const float *table, int G, int K, int npsize, int nmig, int thetai, int mi,
                  double qmax, double mprior, int expo, int prob_or_like, double x) {
  pm_cols c = pm_columns(npsize, nmig, thetai, mi);
  double sum, temp;
  if (!expo) {
    sum = pm_sum_uniform(table, K, c, 0, G, x, qmax, mprior) / G;
    if (prob_or_like) {
      temp = 2 * (log(qmax) + log(mprior) - log(2 * x)) / (qmax * mprior);
      sum /= temp;
    }
  } else {
    sum = exp(pm_logsum_expo(table, K, c, 0, G, x, qmax, mprior) - log((double)G));
    if (prob_or_like) {
      temp = 2 * exp(ora_uppergamma(0, 2 * x / (mprior * qmax))) / (qmax * mprior);
      sum /= temp;
    }
  }
  return sum;
}

double ora_marginpopmig(const float *table, int K, int npsize, int nmig, int thetai, int mi,
                        double qmax, double mprior, int expo, int first, int last, double x) {
  pm_cols c = pm_columns(npsize, nmig, thetai, mi);
  double max = expo ? 20.0 * mprior : qmax * mprior / 2.0; /* EXPOMIGPLOTSCALE */
  if (x < 0 || x > max) return 1.0;                        /* OFFSCALEVAL */
  double div = (double)(last - first + (first == 0));
  if (!expo) return -pm_sum_uniform(table, K, c, first, last, x, qmax, mprior) / div;
  return -exp(pm_logsum_expo(table, K, c, first, last, x, qmax, mprior) - log(div));
}

double ora_popmig_peak(const float *table, int K, int npsize, int nmig, int thetai, int mi,
                       double qmax, double mprior, int expo, int first, int last, double *peak) {
  const int bins = 100;       /* POPMIG_MARGINBINS */
  const double minp = 1e-7;   /* MINPARAMVAL */
  const double r = 0.61803399, cg = 1.0 - r, ftol = 1e-7;
  double upper = expo ? 20.0 * mprior : qmax * mprior / 2.0;
  double maxf = 1e100, ax, bx, cx;
  int maxj = -1;
#define PMF(xx) ora_marginpopmig(table, K, npsize, nmig, thetai, mi, qmax, mprior, expo, first, \
                                 last, (xx))
  for (int j = 0; j < bins; j++) {
    double a = minp + j * (upper / (double)bins), fa = PMF(a);
    if (fa < maxf) {
      maxf = fa;
      maxj = j;
    }
  }
  if (maxj == 0) {
    ax = minp;
    cx = minp + upper / (double)bins;
    bx = (ax + cx) / 2.0;
  } else if (maxj == bins - 1) {
    ax = minp + ((double)bins - 1) * (upper / (double)bins);
    bx = minp + ((double)bins - 2) * (upper / (double)bins);
    cx = upper;
  } else {
    bx = minp + ((double)maxj - 1) * (upper / (double)bins);
    cx = minp + ((double)maxj + 1) * (upper / (double)bins);
    ax = minp + ((double)maxj) * (upper / (double)bins);
  }
  double x0 = ax, x3 = cx, x1, x2;
  if (fabs(cx - bx) > fabs(bx - ax)) {
    x1 = bx;
    x2 = bx + cg * (cx - bx);
  } else {
    x2 = bx;
    x1 = bx - cg * (bx - ax);
  }
  double f1 = PMF(x1), f2 = PMF(x2);
  while (fabs(x3 - x0) > ftol * (fabs(x1) + fabs(x2))) {
    if (f2 < f1) {
      x0 = x1;
      x1 = x2;
      x2 = r * x1 + cg * x3;
      f1 = f2;
      f2 = PMF(x2);
      continue;
    }
    x3 = x2;
    x2 = x1;
    x1 = r * x2 + cg * x0;
    f2 = f1;
    f1 = PMF(x1);
  }
#undef PMF
  *peak = f1 < f2 ? x1 : x2;
  return -(f1 < f2 ? f1 : f2);
}

/* ---------------------------------------------------------------------------
 * Probability that one parameter exceeds another -- gtint.cpp:1-335
 * (mgt_wj_gt_0, pgt_fcj_gt_0, trapzd, qtrap, gtmig, gtpops) and the
 * genealogy subsampling of print_greater_than_tests (:335-349).
 */
#define GT_SWITCH_CRIT 1e-15 /* SWITCH_TO_LOWERGAMMA_CRIT */
#define GT_MINPARAM 1e-7     /* MINPARAMVAL */

typedef struct {
  int cci, ccj, wi, wj;
  double fci, fcj, hval, denom, qmax, fmi, fmj, mmax;
} gt_state;

static double gt_mgt(const gt_state *s, double mi) { /* mgt_wj_gt_0 */
  double temp1, temp2, a, b;
  if (mi < GT_MINPARAM) return 0.0;
  a = ora_logfact(s->wj);
  b = ora_uppergamma(s->wj + 1, s->fmj * mi);
  if (a <= b) return 0.0;
  if ((a - b) < GT_SWITCH_CRIT)
    temp1 = ora_lowergamma(s->wj + 1, s->fmj * mi);
  else
    temp1 = log_diff(a, b);
  temp2 = s->wi * log(mi) - s->fmi * mi - (s->wj + 1) * log(s->fmj) + temp1;
  temp2 -= s->denom;
  return exp(temp2);
}

static double gt_pgt(const gt_state *s, double qi) { /* pgt_fcj_gt_0 */
  double fcj2, fci2, temp1, temp2, temp3, a, b;
  if (qi < GT_MINPARAM) return 0.0;
  fcj2 = 2 * s->fcj;
  fci2 = 2 * s->fci;
  if (s->ccj == 0) {
    a = log(qi) - fcj2 / qi;
    b = log(fcj2) + ora_uppergamma(0, fcj2 / qi);
    if (a > b) {
      temp1 = log_diff(a, b);
      temp2 = -fci2 / qi + s->cci * log(2 / qi);
      temp3 = temp1 + temp2 - s->hval - s->denom;
      return exp(temp3);
    }
    return 0.0;
  }
  temp1 = ora_uppergamma(s->ccj - 1, fcj2 / qi);
  temp2 = ORA_LOG2 + s->cci * log(2 / qi) + (1 - s->ccj) * log(s->fcj) - fci2 / qi;
  temp3 = temp2 + temp1 - s->hval - s->denom;
  return exp(temp3);
}

/* trapzd / qtrap (Numerical Recipes, as gtint.cpp:75-122): EPS 1e-4,
 * JMAX 20, convergence tested only after the fifth refinement */
static double gt_qtrap(const gt_state *s, int mig, double a, double b) {
  double st = 0, olds = -1.0e100;
  for (int j = 1; j <= 20; j++) {
    if (j == 1) {
      double fa = mig ? gt_mgt(s, a) : gt_pgt(s, a);
      double fb = mig ? gt_mgt(s, b) : gt_pgt(s, b);
      st = 0.5 * (b - a) * (fa + fb);
    } else {
      int it = 1;
      for (int k = 1; k < j - 1; k++) it <<= 1;
      double tnm = it, del = (b - a) / tnm, x = a + 0.5 * del, sum = 0.0;
      for (int k = 1; k <= it; k++, x += del) sum += mig ? gt_mgt(s, x) : gt_pgt(s, x);
      st = 0.5 * (st + (b - a) * sum / tnm);
    }
    if (j > 5) {
      if (fabs(st - olds) < 1.0e-4 * fabs(olds) || (st == 0.0 && olds == 0.0)) return st;
    }
    olds = st;
  }
  return st;
}

/* the term of one genealogy of gtmig(mi, mj) (gtint.cpp:127-248) */
double ora_gtmig_one(const float *g, int npsize, int nmig, int mi, int mj, double mmax) {
  int fmp = 3 * npsize + nmig, mcp = 3 * npsize, mip = 4 * npsize + 2 * nmig;
  gt_state s;
  double temp, temp1, temp2, temp3, temp4, a, b, c;
  memset(&s, 0, sizeof s);
  s.mmax = mmax;
  s.fmi = g[fmp + mi];
  s.fmj = g[fmp + mj];
  s.wi = (int)g[mcp + mi];
  s.wj = (int)g[mcp + mj];
  s.denom = g[mip + mi] + g[mip + mj];
  if (s.wj == 0) {
    if (s.fmj > 0.0) {
      if (s.wi > 0) {
        a = ora_logfact(s.wi);
        b = ora_uppergamma(s.wi + 1, (s.fmi + s.fmj) * mmax);
        c = ora_uppergamma(s.wi + 1, s.fmi * mmax);
        if (a <= b || a <= c) {
          temp = 0.0;
        } else {
          if ((a - b) < GT_SWITCH_CRIT)
            temp1 = ora_lowergamma(s.wi + 1, (s.fmi + s.fmj) * mmax);
          else
            temp1 = log_diff(a, b);
          temp1 += -(s.wi + 1) * log(s.fmi + s.fmj);
          if ((a - c) < GT_SWITCH_CRIT)
            temp2 = ora_lowergamma(s.wi + 1, s.fmi * mmax);
          else
            temp2 = log_diff(a, c);
          temp2 += -(s.wi + 1) * log(s.fmi);
          if (temp2 <= temp1) {
            temp = 0.0;
          } else {
            temp3 = log_diff(temp2, temp1);
            temp4 = temp3 - log(s.fmj) - s.denom;
            temp = exp(temp4);
          }
        }
      } else if (s.fmi > 0.0) {
        temp1 = s.fmi * (exp(-(s.fmi + s.fmj) * mmax) - exp(-s.fmi * mmax));
        temp2 = s.fmj * (1 - exp(-s.fmi * mmax));
        temp3 = (temp1 + temp2) / (s.fmi * s.fmj * (s.fmi + s.fmj));
        temp4 = log(temp3) - s.denom;
        temp = exp(temp4);
      } else {
        temp1 = (s.fmj * mmax - 1.0 + exp(-s.fmj * mmax)) / (s.fmj * s.fmj);
        temp4 = log(temp1) - s.denom;
        temp = exp(temp4);
      }
    } else {
      if (s.wi > 0) {
        a = ora_logfact(s.wi + 1);
        b = ora_uppergamma(s.wi + 2, s.fmi * mmax);
        if (a <= b) {
          temp = 0.0;
        } else {
          if ((a - b) < GT_SWITCH_CRIT)
            temp1 = ora_lowergamma(s.wi + 2, s.fmi * mmax);
          else
            temp1 = log_diff(a, b);
          temp2 = -(s.wi + 2) * log(s.fmi);
          temp3 = temp2 + temp1 - s.denom;
          temp = exp(temp3);
        }
      } else if (s.fmi > 0.0) {
        temp1 = (1.0 - exp(-s.fmi * mmax) * (s.fmi * mmax + 1.0)) / (s.fmi * s.fmi);
        temp2 = log(temp1) - s.denom;
        temp = exp(temp2);
      } else {
        temp1 = log(mmax * mmax / 2.0) - s.denom;
        temp = exp(temp1);
      }
    }
  } else {
    temp = gt_qtrap(&s, 1, GT_MINPARAM, mmax);
  }
  return temp < 1.0 ? temp : 1.0;
}

/* the term of one genealogy of gtpops(pi, pj) (gtint.cpp:251-331) */
double ora_gtpops_one(const float *g, int npsize, int nmig, int pi, int pj, double qmax) {
  int ccp = 0, fcp = npsize, hccp = 2 * npsize, qip = 3 * npsize + 2 * nmig;
  gt_state s;
  double temp, temp1, temp2, temp3;
  memset(&s, 0, sizeof s);
  s.qmax = qmax;
  s.cci = (int)g[ccp + pi];
  s.ccj = (int)g[ccp + pj];
  s.fci = g[fcp + pi];
  s.fcj = g[fcp + pj];
  s.hval = g[hccp + pi] + g[hccp + pj];
  s.denom = g[qip + pi] + g[qip + pj];
  if (s.ccj == 0 && s.fcj == 0) {
    if (s.fci == 0) {
      temp = exp(2.0 * log(qmax) - ORA_LOG2 - s.hval - s.denom);
    } else if (s.cci >= 2) {
      temp1 = 2 * ORA_LOG2 + (2 - s.cci) * log(s.fci);
      temp2 = ora_uppergamma(s.cci - 2, 2 * s.fci / qmax);
      temp3 = temp1 + temp2 - s.hval - s.denom;
      temp = exp(temp3);
    } else if (s.cci == 1) {
      temp1 = 4 * s.fci * exp(ora_uppergamma(0, 2 * s.fci / qmax));
      temp2 = 2 * qmax * exp(-2 * s.fci / qmax) - temp1;
      temp3 = log(temp2) - s.hval - s.denom;
      temp = exp(temp3);
    } else {
      temp1 = exp(ora_uppergamma(0, 2 * s.fci / qmax));
      temp2 = (qmax / 2) * (qmax - 2 * s.fci) * exp(-2 * s.fci / qmax) + 2 * s.fci * s.fci * temp1;
      temp3 = log(temp2) - s.hval - s.denom;
      temp = exp(temp3);
    }
  } else {
    temp = gt_qtrap(&s, 0, GT_MINPARAM, qmax);
  }
  return temp < 1.0 ? temp : 1.0;
}

/* gtpops / gtmig over the table with print_greater_than_tests's
 * subsampling: at most 20000 genealogies (USETREESMAX), every
 * (G / 20000)-th one when there are more */
double ora_greater_than(const float *table, int G, int K, int npsize, int nmig, int kind, int i,
                        int j, double prior_max) {
  int treeinc = 1, numtrees = G, hitree = G;
  if (G > 20000) {
    treeinc = G / 20000;
    numtrees = 20000;
    hitree = treeinc * 20000;
  }
  double sum = 0;
  for (int ei = 0; ei < hitree; ei += treeinc) {
    const float *g = table + (size_t)ei * K;
    sum += kind ? ora_gtmig_one(g, npsize, nmig, i, j, prior_max)
                : ora_gtpops_one(g, npsize, nmig, i, j, prior_max);
  }
  return sum / numtrees;
}
