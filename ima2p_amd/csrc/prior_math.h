// prior_math.h -- log incomplete gamma functions and the analytic
// integrals of the genealogy prior over uniform priors on population size
// and migration rate (integrate_coalescent_term / integrate_migration_term,
// update_gtree_common.cpp:109-296).
//
// The incomplete gammas follow the published Numerical Recipes methods the
// reference uses (utilities.cpp:812-1122): series for x < a + 1, Lentz
// continued fraction otherwise, relative tolerance 3e-7, at most 1000
// terms, FPMIN 1e-30; E_1 by its series (x <= 1) or continued fraction
// (at most 100 terms).  The termination rules decide the value to ~3e-7,
// so they are kept as they are.  The functions take lga = log Gamma(a) =
// log((a-1)!) for their order a; callers read it from `lf`, a table built
// on the host as the running sum of log(i), as the reference's setlogfact
// builds it.  Everything is FP64.  The batch prior kernels compile these
// without floating-point contraction, so sums round as written.
#pragma once

#include <hip/hip_runtime.h>

#include "ima_internal.h"

namespace ima {

constexpr double kPriEps = 3.0e-7;
constexpr int kPriItmax = 1000;
constexpr double kPriFpmin = 1.0e-30;
constexpr double kLog2 = 0.69314718055994530941723212146;

__device__ inline double pri_gcf(double a, double x, double lga, double *gln,
                                 double *h_out) {
  *gln = lga;
  double b = x + 1.0 - a, c = 1.0 / kPriFpmin, d = 1.0 / b, h = d;
  for (int i = 1; i <= kPriItmax; i++) {
    const double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < kPriFpmin) d = kPriFpmin;
    c = b + an / c;
    if (fabs(c) < kPriFpmin) c = kPriFpmin;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < kPriEps) break;
  }
  *h_out = h;
  return -x + a * log(x) - (*gln);
}

// series: returns the sum and the log prefactor -x + a log x - log Gamma(a)
__device__ inline double pri_gser(int a, double x, double lga, double *gln,
                                  double *pre) {
  *gln = lga;
  *pre = 0.0;
  if (x <= 0.0) return 0.0;
  double ap = a, del = 1.0 / a, sum = del;
  for (int n = 1; n <= kPriItmax; n++) {
    ++ap;
    del *= x / ap;
    sum += del;
    if (fabs(del) < fabs(sum) * kPriEps) break;
  }
  *pre = -x + a * log(x) - (*gln);
  return sum;
}

// E_1(x) or its log (*islog)
__device__ inline double pri_expint1(double x, int *islog) {
  const double euler = 0.5772156649;
  *islog = 0;
  if (x > 1.0) {
    double b = x + 1.0, c = 1.0 / kPriFpmin, d = 1.0 / b, h = d;
    for (int i = 1; i <= 100; i++) {
      const double a = -i * (0 + i);
      b += 2.0;
      d = 1.0 / (a * d + b);
      c = b + a / c;
      const double del = c * d;
      h *= del;
      if (fabs(del - 1.0) < kPriEps) break;
    }
    *islog = 1;
    return log(h) - x;
  }
  double ans = -log(x) - euler, fact = 1.0;
  for (int i = 1; i <= 100; i++) {
    fact *= -x / i;
    const double del = -fact / (i - 0);
    ans += del;
    if (fabs(del) < fabs(ans) * kPriEps) break;
  }
  return ans;
}

__device__ inline double pri_uppergamma(int a, double x, double lga) {
  double gln, p;
  if (a == 0) {
    int islog;
    const double t = pri_expint1(x, &islog);
    p = islog ? t : log(t);
  } else if (x < (a + 1.0)) {
    double pre;
    const double s = pri_gser(a, x, lga, &gln, &pre);
    const double g = (x <= 0.0) ? 0.0 : s * exp(pre);
    p = gln + log(1.0 - g);
  } else {
    double h;
    const double pre = pri_gcf((double)a, x, lga, &gln, &h);
    p = gln + (pre + log(h));
  }
  if (p < -1e200) p = -1e200;
  return p;
}

__device__ inline double pri_lowergamma(int a, double x, double lga) {
  double gln, p;
  if (x < (a + 1.0)) {
    double pre;
    const double s = pri_gser(a, x, lga, &gln, &pre);
    const double g = (x <= 0.0) ? 0.0 : log(s) + pre;
    p = gln + g;
  } else {
    double h;
    const double pre = pri_gcf((double)a, x, lga, &gln, &h);
    p = gln + log(1 - exp(pre) * h);
  }
  if (p < -1e200) p = -1e200;
  return p;
}

// Both log incomplete gammas of one (a, x) share their series or continued
// fraction: run it once, then pri_upper_of / pri_lower_of give exactly the
// values pri_uppergamma / pri_lowergamma would (a >= 1).
struct GammaPrep {
  bool series;
  double sum_or_h, pre, gln, x;
};
__device__ inline GammaPrep pri_gamma_prep(int a, double x, double lga) {
  GammaPrep g;
  g.x = x;
  g.series = x < (a + 1.0);
  if (g.series)
    g.sum_or_h = pri_gser(a, x, lga, &g.gln, &g.pre);
  else
    g.pre = pri_gcf((double)a, x, lga, &g.gln, &g.sum_or_h);
  return g;
}
__device__ inline double pri_upper_of(const GammaPrep &g) {
  double p;
  if (g.series) {
    const double gg = (g.x <= 0.0) ? 0.0 : g.sum_or_h * exp(g.pre);
    p = g.gln + log(1.0 - gg);
  } else {
    p = g.gln + (g.pre + log(g.sum_or_h));
  }
  return p < -1e200 ? -1e200 : p;
}
__device__ inline double pri_lower_of(const GammaPrep &g) {
  double p;
  if (g.series) {
    const double gg = (g.x <= 0.0) ? 0.0 : log(g.sum_or_h) + g.pre;
    p = g.gln + gg;
  } else {
    p = g.gln + log(1 - exp(g.pre) * g.sum_or_h);
  }
  return p < -1e200 ? -1e200 : p;
}

// log(e^a - e^b) for a > b (LogDiff, imamp.hpp:257-263)
__device__ inline double pri_logdiff(double a, double b) {
  if (a - b < 7.0978271289338397e+02) return b + log(exp(a - b) - 1.0);
  return a;
}

// log int_min^max (2/t)^cc e^(-2 fc/t - hcc) dt; lga = log((cc-2)!)
__device__ inline double pri_coalescent_term(int cc, double fc, double hcc, double max,
                                             double min, double lga) {
  double p;
  if (cc > 0) {
    if (min == 0) {
      // cc = 1: order 0, the exponential integral; else one series serves
      // both gammas
      GammaPrep gp{};
      if (cc > 1) gp = pri_gamma_prep(cc - 1, 2 * fc / max, lga);
      double ug = cc > 1 ? pri_upper_of(gp) : pri_uppergamma(0, 2 * fc / max, lga);
      if (cc > 1) {
        const double fullg = lga;
        if (fullg - ug < 1e-15 || fullg - ug > 7.0978271289338397e+02) {
          const double lg = pri_lower_of(gp);
          if (fullg > lg) {
            const double ugalt = pri_logdiff(fullg, lg);
            if (fabs(ugalt - ug) > 1e-10) ug = ugalt;
          }
        }
      }
      p = ug + kLog2 - hcc + (1 - cc) * log(fc);
    } else {
      const double a = pri_uppergamma(cc - 1, 2 * fc / max, lga);
      const double b = pri_uppergamma(cc - 1, 2 * fc / min, lga);
      p = pri_logdiff(a, b) + (kLog2 - hcc + (1 - cc) * log(fc));
    }
  } else if (2 * fc / max > 0) {
    if (min == 0) {
      const double a = log(max) - 2.0 * fc / max;
      const double b = kLog2 + log(fc) + pri_uppergamma(0, 2.0 * fc / max, 0.0);
      p = pri_logdiff(a, b);
    } else {
      double a = pri_uppergamma(0, 2 * fc / max, 0.0), b = pri_uppergamma(0, 2 * fc / min, 0.0);
      const double c = pri_logdiff(a, b) + kLog2 + log(fc);
      a = log(max) - 2.0 * fc / max;
      b = log(min) - 2.0 * fc / min;
      p = pri_logdiff(pri_logdiff(a, b), c);
    }
  } else {
    p = log(max - min);
  }
  return p;
}

// log int_min^max m^cm e^(-fm m) dm; lga = log(cm!)
__device__ inline double pri_migration_term(int cm, double fm, double max, double min,
                                            double lga) {
  const double mpriormin = 0.000001;  // MPRIORMIN, imamp.hpp:141
  double p;
  if (cm > 0) {
    if (min == 0) {
      const GammaPrep gp = pri_gamma_prep(cm + 1, fm * max, lga);
      double lg = pri_lower_of(gp);
      const double fullg = lga;
      if (fullg - lg < 1e-15 || fullg - lg > 7.0978271289338397e+02) {
        const double ug = pri_upper_of(gp);
        if (fullg > ug) {
          const double lgalt = pri_logdiff(fullg, ug);
          if (fabs(lgalt - lg) > 1e-12) lg = lgalt;
        }
      }
      p = (-1 - cm) * log(fm) + lg;
    } else {
      const double a = pri_uppergamma(cm + 1, fm * min, lga);
      const double b = pri_uppergamma(cm + 1, fm * max, lga);
      p = (-1 - cm) * log(fm) + pri_logdiff(a, b);
    }
  } else if (fm > mpriormin) {
    if (min == 0)
      p = (max == mpriormin) ? 0.0 : pri_logdiff(0.0, -fm * max) - log(fm);
    else
      p = pri_logdiff(-fm * min, -fm * max) - log(fm);
  } else {
    p = log(max - min);
  }
  return p;
}

// Count, weight and hcc sums of parameter t (population sizes first, then
// migration rates) over its (period, position) entries of a weight record,
// in parameter-map order (integrate_tree_prob's loops).
// get(r) returns record entry r.
template <typename Get>
__device__ inline void prior_param_sums_of(const ModelView &md, int P, int npsize, int t,
                                           const Get &get, int *c, double *f, double *h) {
  int cc = 0;
  double ff = 0.0, hh = 0.0;
  if (t < npsize) {
    for (int w = md.theta_off[t]; w < md.theta_off[t + 1]; ++w) {
      const int kk = md.theta_wp[2 * w], ii = md.theta_wp[2 * w + 1];
      cc += (int)get(kk * P + ii);
      ff += get(P * P + kk * P + ii);
      hh += get(2 * P * P + kk * P + ii);
    }
  } else {
    const int m = t - npsize;
    for (int w = md.mig_off[m]; w < md.mig_off[m + 1]; ++w) {
      const int kk = md.mig_wp[3 * w], ii = md.mig_wp[3 * w + 1], jj = md.mig_wp[3 * w + 2];
      cc += (int)get(3 * P * P + (kk * P + ii) * P + jj);
      ff += get(3 * P * P + P * P * P + (kk * P + ii) * P + jj);
    }
  }
  *c = cc;
  *f = ff;
  *h = hh;
}

__device__ inline void prior_param_sums(const ModelView &md, int P, int npsize, int t,
                                        const double *rec, int *c, double *f, double *h) {
  prior_param_sums_of(md, P, npsize, t, [rec](int r) { return rec[r]; }, c, f, h);
}

// log Gamma of the incomplete gamma order parameter t's integral uses at
// count c: log((c-2)!) for a population size (no gamma below c = 2),
// log(c!) for a migration rate.
__device__ inline double prior_param_lga(int npsize, int t, int c, const double *lf) {
  return t < npsize ? (c >= 2 ? lf[c - 2] : 0.0) : lf[c];
}

// The analytic integral of parameter t at those sums (uniform prior on
// [0, qmax] / [0, mmax]), given prior_param_lga.
__device__ inline double prior_param_term_lga(const ModelView &md, int npsize, int t, int c,
                                              double f, double h, double lga) {
  return t < npsize ? pri_coalescent_term(c, f, h, md.qmax[t], 0.0, lga)
                    : pri_migration_term(c, f, md.mmax[t - npsize], 0.0, lga);
}

__device__ inline double prior_param_term(const ModelView &md, int npsize, int t, int c,
                                          double f, double h, const double *lf) {
  return prior_param_term_lga(md, npsize, t, c, f, h, prior_param_lga(npsize, t, c, lf));
}

// sum_{i < n} rec[i * stride + r] in order, with eight loads in flight
__device__ inline double ordered_rec_sum(const double *rec, int64_t stride, int n, int r) {
  double s = 0.0;
  int i = 0;
  for (; i + 8 <= n; i += 8) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = rec[(int64_t)(i + k) * stride + r];
#pragma unroll
    for (int k = 0; k < 8; ++k) s += v[k];
  }
  for (; i < n; ++i) s += rec[(int64_t)i * stride + r];
  return s;
}

}  // namespace ima
