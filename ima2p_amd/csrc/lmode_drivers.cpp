// lmode_drivers.cpp -- IMa2p's L-mode optimisers on top of the batched
// jointp launch (host code over the engine's own entry points).
//
// Joint peaks by differential evolution (jointfind.cpp:596-812): a
// population of (varying parameters x 100) vectors drawn uniformly within
// the bounds; each generation makes one trial per member -- three members
// A, B, C at random, and per coordinate with probability 0.9
// C + 0.8 (A - B), moved only part of the way toward a bound it would
// cross -- evaluates every trial, and keeps a trial that lowers the
// member's -log P.  Generations repeat until the population's values span
// at most 1e-7; the search restarts from a new population until the same
// best value (within 1e-6) has been found twice, or ten times.  Nested
// models map their varying values onto the full parameter vector with
// mapvals (:563-573), their bounds with reversemapvals (:547-560).
//
// The reference evaluates the trials of a generation one jointp call at a
// time, but every trial is built from the population as it was before the
// generation's selections, so evaluating the generation's trials in one
// batched launch gives the same values.  The random numbers come from a
// Mersenne twister (MT19937, as the reference's uniform()) seeded by the
// caller; the reference's exact stream is not reproduced.
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "../../include/ima_gpu.h"

namespace {

constexpr double kSpreadTol = 1.0e-7;  // SPREADTOL
constexpr double kLoopTol = 1.0e-6;    // PLOOPTOL
constexpr int kLoopMatch = 2;          // LOOPMATCHCRITERIA
constexpr int kMaxRestart = 10;        // MAXRESTART
constexpr double kRecRate = 0.9;       // DEFAULTRECRATE
constexpr double kFWeight = 0.8;       // DEFAULTFWEIGHT
constexpr int kMaxGenerations = 20000; // a generation cap the reference does not have

struct DE {
  ima_ctx *ctx = nullptr;
  int tid = 0, np = 0, r0 = 0, r1 = 0, jmax = 0, mode = 0, modeltype = 0, popn = 0;
  const double *xmap = nullptr;
  std::vector<double> lo, hi;     // mapped bounds (reversemapvals), by vector position
  std::vector<double> pop, trial; // popn vectors of np values + the value at np
  std::vector<double> full, val;
  std::mt19937 rng;
  std::uniform_real_distribution<double> U{0.0, 1.0};
  int64_t evals = 0;

  double uniform() { return U(rng); }
  int randposint(int n) {
    int k;
    do k = (int)(uniform() * n);
    while (k < 0 || k >= n);
    return k;
  }
  double *row(std::vector<double> &v, int i) { return &v[(size_t)i * (np + 1)]; }
  // mapvals (jointfind.cpp:563-573); parameters outside the model's range
  // are not read by jointp and get 1
  void mapvals(const double *from, double *to) const {
    for (int i = 0; i < np; ++i) {
      if (i < r0 || i >= r1) to[i] = 1.0;
      else if (!xmap) to[i] = from[i];
      else if (xmap[i] < 0) to[i] = -xmap[i];
      else to[i] = from[(int)xmap[i]];
    }
  }
  // -log P of every vector of v (popn of them) into v[.][np], one launch
  int eval(std::vector<double> &v) {
    full.assign((size_t)popn * np, 1.0);
    for (int i = 0; i < popn; ++i) mapvals(row(v, i), &full[(size_t)i * np]);
    val.resize(popn);
    const int rc = ima_lmode_jointp(ctx, tid, full.data(), popn, modeltype, r0, r1, mode,
                                    val.data(), nullptr);
    if (rc != IMA_OK) return rc;
    for (int i = 0; i < popn; ++i) row(v, i)[np] = val[i];
    evals += popn;
    return IMA_OK;
  }
  // startpop (:659-699)
  int startpop() {
    for (int i = 0; i < popn; ++i) {
      double *p = row(pop, i);
      for (int j = 0; j < np; ++j)
        p[j] = (j >= r0 && j < jmax) ? lo[j] + uniform() * (hi[j] - lo[j]) : -1.0;
    }
    return eval(pop);
  }
  // nextgen (:599-656)
  int nextgen(double *lowpd, double *hipd) {
    for (int i = 0; i < popn; ++i) {
      const int A = randposint(popn), B = randposint(popn), Cv = randposint(popn);
      const double *pa = row(pop, A), *pb = row(pop, B), *pc = row(pop, Cv), *pi = row(pop, i);
      double *t = row(trial, i);
      for (int j = 0; j < np; ++j) t[j] = pi[j];
      for (int j = r0; j < jmax; ++j) {
        if (uniform() < kRecRate) {
          double x = pc[j] + kFWeight * (pa[j] - pb[j]);
          if (x < lo[j]) x = pc[j] - uniform() * (pc[j] - lo[j]);
          if (x > hi[j]) x = pc[j] + uniform() * (hi[j] - pc[j]);
          t[j] = x;
        } else {
          t[j] = pi[j];
        }
      }
    }
    const int rc = eval(trial);
    if (rc != IMA_OK) return rc;
    *lowpd = 1e200;
    *hipd = -1e200;
    for (int i = 0; i < popn; ++i) {
      double *p = row(pop, i);
      const double *t = row(trial, i);
      if (t[np] < p[np])
        for (int j = 0; j <= np; ++j) p[j] = t[j];
      if (p[np] < *lowpd) *lowpd = p[np];
      if (p[np] > *hipd) *hipd = p[np];
    }
    return IMA_OK;
  }
  int best() const {
    int k = 0;
    for (int i = 1; i < popn; ++i)
      if (pop[(size_t)i * (np + 1) + np] < pop[(size_t)k * (np + 1) + np]) k = i;
    return k;
  }
};

}  // namespace

extern "C" {

int ima_lmode_joint_peak(ima_ctx *ctx, int table_id, int modeltype, const double *xmap,
                         const double *lower, const double *upper, int popsize_multiplier,
                         uint64_t seed, int mode, double *peak, double *logp, double *ess,
                         int64_t *evaluations) {
  if (!ctx || !lower || !upper || !peak || modeltype < 0 || modeltype > 2 ||
      popsize_multiplier < 4)
    return IMA_ERR_ARG;
  int G = 0, npsize = 0, nmig = 0, npops = 0, rc;
  if ((rc = ima_lmode_table_info(ctx, table_id, &G, nullptr, &npsize, &nmig, &npops)) != IMA_OK)
    return rc;
  DE d;
  d.ctx = ctx;
  d.tid = table_id;
  d.mode = mode;
  d.modeltype = modeltype;
  d.xmap = xmap;
  d.np = npsize + nmig;
  d.rng.seed((uint32_t)seed);
  if (modeltype == 0) {
    d.r0 = 0;
    d.r1 = d.np;
  } else if (modeltype == 1) {
    d.r0 = 0;
    d.r1 = npsize;
  } else {
    d.r0 = npsize;
    d.r1 = d.np;
  }
  if (d.r1 <= d.r0) return IMA_ERR_ARG;
  for (int i = 0; i < d.np; ++i)
    if (!(lower[i] > 0) || !(upper[i] > lower[i])) return IMA_ERR_ARG;
  // varying values of the (nested) model: nparams_nested, the distinct
  // positions xmap refers to inside the model's range
  int modelparams = d.r1 - d.r0;
  if (xmap) {
    int mx = -1;
    for (int i = 0; i < d.np; ++i) {
      if (xmap[i] < 0) {
        if (i < d.r0 || i >= d.r1) return IMA_ERR_ARG;  // constants inside the range only
        continue;
      }
      const int k = (int)xmap[i];
      if ((double)k != xmap[i] || k >= d.np) return IMA_ERR_ARG;
      if ((i < d.r0 || i >= d.r1) && k != i) return IMA_ERR_ARG;
      if (i >= d.r0 && i < d.r1 && (k < d.r0 || k >= d.r1)) return IMA_ERR_ARG;
      if (i >= d.r0 && i < d.r1) mx = mx > k ? mx : k;
    }
    modelparams = mx < d.r0 ? 0 : mx - d.r0 + 1;
    if (modelparams < 1) return IMA_ERR_ARG;
  }
  d.jmax = d.r0 + modelparams;
  // reversemapvals (:547-560) of the bounds
  d.lo.assign(d.np, 0.0);
  d.hi.assign(d.np, 0.0);
  if (!xmap) {
    d.lo.assign(lower, lower + d.np);
    d.hi.assign(upper, upper + d.np);
  } else {
    for (int i = 0, toi = 0; i < d.np; ++i)
      if (toi == (int)xmap[i]) {
        d.lo[toi] = lower[i];
        d.hi[toi] = upper[i];
        ++toi;
      }
  }
  d.popn = modelparams * popsize_multiplier;
  d.pop.assign((size_t)d.popn * (d.np + 1), 0.0);
  d.trial.assign((size_t)d.popn * (d.np + 1), 0.0);
  // modelloop (:744-812)
  std::vector<double> bestvals(d.np + 1, 0.0);
  if ((rc = d.startpop()) != IMA_OK) return rc;
  int k = d.best();
  for (int j = 0; j <= d.np; ++j) bestvals[j] = d.row(d.pop, k)[j];
  double global_pd = bestvals[d.np];
  int newstart = 0, countloop = 0;
  do {
    if (newstart > 0 && (rc = d.startpop()) != IMA_OK) return rc;
    double lowpd = 0.0, hipd = 0.0;
    int gen = 0;
    do {
      if ((rc = d.nextgen(&lowpd, &hipd)) != IMA_OK) return rc;
    } while (hipd - lowpd > kSpreadTol && ++gen < kMaxGenerations);
    const double local_pd = lowpd;
    if (std::fabs(local_pd - global_pd) < kLoopTol)
      countloop++;
    else if (local_pd < global_pd)
      countloop = 0;
    if (local_pd < global_pd) {
      k = d.best();
      for (int j = 0; j <= d.np; ++j) bestvals[j] = d.row(d.pop, k)[j];
      global_pd = local_pd;
    }
    newstart++;
  } while (countloop < kLoopMatch && newstart < kMaxRestart);
  // the peak on the full parameter vector, with its effective sample number
  d.mapvals(bestvals.data(), peak);
  double v = 0.0, e = 0.0;
  if ((rc = ima_lmode_jointp(ctx, table_id, peak, 1, modeltype, d.r0, d.r1, mode, &v, &e)) !=
      IMA_OK)
    return rc;
  if (logp) *logp = -v;
  if (ess) *ess = e;
  if (evaluations) *evaluations = d.evals;
  return IMA_OK;
}

}  // extern "C"

// ---- marginal peaks (surface_call_functions.cpp:175-315, 330-470) -----
//
// The marginal posterior density of one parameter, margincalc's mean over
// genealogies (:120-173), is maximised over [lower, upper]: the reference
// brackets it (mnbrakmod) and refines by golden section to a relative
// tolerance of 1e-7 (goldenmod).  Here the bracket comes from one batched
// launch over a grid of kGrid points, the golden section then refines
// within the grid cells around the best point.  The 95% limits are where
// the density falls e^1.92 below the peak's, found by bisection on each side
// (margin95 / marginbis, :265-290); for a migration rate the likelihood-
// ratio statistic 2 log(p(peak) / p(MINPARAMVAL)) is the reference's test
// of no migration (:404-409).

namespace {

constexpr int kGrid = 1024;
constexpr double kMinParam = 1e-7;  // MINPARAMVAL
constexpr double kDown95 = 1.92;    // DOWN95

struct Marg {
  ima_ctx *ctx;
  int tid, param, expo;
  double mean;
  int64_t evals = 0;
  // log marginal density at each x (margincalc with logi = 1)
  int eval(const double *x, int n, double *out) {
    evals += n;
    return ima_lmode_margincalc(ctx, tid, param, x, n, 1, 0.0, expo, mean, out);
  }
  int eval1(double x, double *out) { return eval(&x, 1, out); }
};

}  // namespace

extern "C" int ima_lmode_marginal_peak(ima_ctx *ctx, int table_id, int param, double lower,
                                       double upper, int expo, double prior_mean, double *out) {
  if (!ctx || !out || !(lower > 0) || !(upper > lower)) return IMA_ERR_ARG;
  int npsize = 0, nmig = 0, rc;
  if ((rc = ima_lmode_table_info(ctx, table_id, nullptr, nullptr, &npsize, &nmig, nullptr)) !=
      IMA_OK)
    return rc;
  if (param < 0 || param >= npsize + nmig) return IMA_ERR_ARG;
  Marg f{ctx, table_id, param, expo, prior_mean};
  // bracket: the best point of a geometric-then-linear grid over the range
  std::vector<double> x(kGrid), y(kGrid);
  for (int i = 0; i < kGrid; ++i) x[i] = lower + (upper - lower) * (double)i / (kGrid - 1);
  if ((rc = f.eval(x.data(), kGrid, y.data())) != IMA_OK) return rc;
  int b = 0;
  for (int i = 1; i < kGrid; ++i)
    if (y[i] > y[b]) b = i;
  // golden section on [x[b-1], x[b+1]] (goldenmod's tolerance)
  double a = x[b > 0 ? b - 1 : 0], c = x[b < kGrid - 1 ? b + 1 : kGrid - 1];
  const double gr = 0.61803398874989484820;
  double x1 = c - gr * (c - a), x2 = a + gr * (c - a), f1, f2;
  if ((rc = f.eval1(x1, &f1)) != IMA_OK) return rc;
  if ((rc = f.eval1(x2, &f2)) != IMA_OK) return rc;
  while (std::fabs(c - a) > 1e-7 * (std::fabs(x1) + std::fabs(x2)) && f.evals < kGrid + 400) {
    if (f1 > f2) {
      c = x2;
      x2 = x1;
      f2 = f1;
      x1 = c - gr * (c - a);
      if ((rc = f.eval1(x1, &f1)) != IMA_OK) return rc;
    } else {
      a = x1;
      x1 = x2;
      f1 = f2;
      x2 = a + gr * (c - a);
      if ((rc = f.eval1(x2, &f2)) != IMA_OK) return rc;
    }
  }
  double peak = f1 > f2 ? x1 : x2, fp = f1 > f2 ? f1 : f2;
  if (y[b] > fp) {  // the range's edge is the maximum
    peak = x[b];
    fp = y[b];
  }
  out[0] = peak;
  out[1] = fp;
  // 95% limits: where the log density is fp - 1.92, by bisection between
  // the peak and the range's ends (-1 where the density never falls so low)
  const double target = fp - kDown95;
  for (int side = 0; side < 2; ++side) {
    double lo = side == 0 ? lower : peak, hi = side == 0 ? peak : upper, fl, fh;
    if ((rc = f.eval1(lo, &fl)) != IMA_OK) return rc;
    if ((rc = f.eval1(hi, &fh)) != IMA_OK) return rc;
    if ((fl - target) * (fh - target) > 0) {
      out[2 + side] = -1.0;
      continue;
    }
    for (int it = 0; it < 100 && hi - lo > 1e-9 * (hi + lo); ++it) {
      const double mid = 0.5 * (lo + hi);
      double fm;
      if ((rc = f.eval1(mid, &fm)) != IMA_OK) return rc;
      if ((fm - target) * (fl - target) <= 0) {
        hi = mid;
        fh = fm;
      } else {
        lo = mid;
        fl = fm;
      }
    }
    out[2 + side] = 0.5 * (lo + hi);
  }
  // no-migration test (migration rates only)
  out[4] = 0.0;
  if (param >= npsize) {
    double f0;
    if ((rc = f.eval1(kMinParam, &f0)) != IMA_OK) return rc;
    out[4] = 2.0 * (fp - f0);
  }
  return IMA_OK;
}

// ---- peak of the 2NM density (popmig.cpp:355-429) ----------------------
//
// marginalopt_popmig for one (population size, migration rate) pair: the
// curve often has several peaks, so the reference scans 100 bins for the
// lowest -density (the first of equal values), brackets around it, and
// refines by golden section (goldenmod, surface_search_functions.cpp:
// 134-184, r = 0.61803399, tolerance 1e-7).  The 100 bins are one launch;
// each golden-section step evaluates one point.

extern "C" int ima_lmode_popmig_peak(ima_ctx *ctx, int table_id, int theta_i, int mig_i,
                                     double qmax, double mprior, int expo, int first, int last,
                                     double *peak, double *mlval, int64_t *evaluations) {
  if (!ctx || !peak) return IMA_ERR_ARG;
  constexpr int kBins = 100;          // POPMIG_MARGINBINS
  constexpr double kMinParam = 1e-7;  // MINPARAMVAL
  constexpr double r = 0.61803399, cg = 1.0 - r;
  const double upper = expo ? 20.0 * mprior : qmax * mprior / 2.0;  // EXPOMIGPLOTSCALE
  int64_t evals = 0;
  auto f = [&](const double *x, int n, double *out) {
    evals += n;
    return ima_lmode_popmig(ctx, table_id, theta_i, mig_i, qmax, mprior, expo, IMA_POPMIG_MARGIN,
                            0, first, last, x, n, out);
  };
  std::vector<double> bx(kBins), fb(kBins);
  for (int j = 0; j < kBins; ++j) bx[j] = kMinParam + j * (upper / (double)kBins);
  int rc = f(bx.data(), kBins, fb.data());
  if (rc != IMA_OK) return rc;
  double maxf = 1e100;
  int maxj = -1;
  for (int j = 0; j < kBins; ++j)
    if (fb[j] < maxf) {
      maxf = fb[j];
      maxj = j;
    }
  double ax, bxm, cx;
  const double w = upper / (double)kBins;
  if (maxj == 0) {
    ax = kMinParam;
    cx = kMinParam + w;
    bxm = (ax + cx) / 2.0;
  } else if (maxj == kBins - 1) {
    ax = kMinParam + ((double)kBins - 1) * w;
    bxm = kMinParam + ((double)kBins - 2) * w;
    cx = upper;
  } else {
    bxm = kMinParam + ((double)maxj - 1) * w;
    cx = kMinParam + ((double)maxj + 1) * w;
    ax = kMinParam + ((double)maxj) * w;
  }
  // goldenmod
  double x0 = ax, x3 = cx, x1, x2, f1, f2;
  if (std::fabs(cx - bxm) > std::fabs(bxm - ax)) {
    x1 = bxm;
    x2 = bxm + cg * (cx - bxm);
  } else {
    x2 = bxm;
    x1 = bxm - cg * (bxm - ax);
  }
  if ((rc = f(&x1, 1, &f1)) != IMA_OK) return rc;
  if ((rc = f(&x2, 1, &f2)) != IMA_OK) return rc;
  while (std::fabs(x3 - x0) > 1e-7 * (std::fabs(x1) + std::fabs(x2)) && evals < kBins + 2000) {
    if (f2 < f1) {
      x0 = x1;
      x1 = x2;
      x2 = r * x1 + cg * x3;
      f1 = f2;
      if ((rc = f(&x2, 1, &f2)) != IMA_OK) return rc;
      continue;
    }
    x3 = x2;
    x2 = x1;
    x1 = r * x2 + cg * x0;
    f2 = f1;
    if ((rc = f(&x1, 1, &f1)) != IMA_OK) return rc;
  }
  const double res = f1 < f2 ? f1 : f2;
  *peak = f1 < f2 ? x1 : x2;
  if (mlval) *mlval = -res;
  if (evaluations) *evaluations = evals;
  return IMA_OK;
}
