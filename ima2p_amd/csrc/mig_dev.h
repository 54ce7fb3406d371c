// mig_dev.h -- device helpers shared by the moves that draw migration
// paths (changet_NW in kern_nw.hip, updategenealogy's move with migration
// in kern_topo.hip): the reference's conditioned Poisson draws, calcmrate,
// the polar normal deviate and log sinh / log cosh with their large-value
// forms (utilities.cpp:236-255, 471-596; update_gtree_common.cpp:463-482).
#pragma once

#include "philox.h"

namespace ima {

constexpr double kLog2 = 0.69314718055994530941723212146;
constexpr int kAddMigMax = 1000;     // ADDMIGMAX
constexpr int kMaxTries = 1 << 20;   // bound of every rejection loop
constexpr int kPoissonMax = 1 << 20;  // bound of one Poisson product loop

__device__ __forceinline__ int mig_randposint(PhiloxStream &r, int n) {
  return (int)floor(r.uniform() * n);
}

__device__ __forceinline__ double calcmrate(int mc, double mt) {
  if (mt <= 0.0) return 1.0;
  if (mc == 0) return mt < 1 ? 0.1 : 0.1 / mt;
  return mt < 1 ? (double)mc : (double)mc / mt;
}

__device__ __forceinline__ double mig_normdev(PhiloxStream &r, double mean, double stdev) {
#pragma clang fp contract(off)
  double v1, v2, rsq;
  do {
    v1 = 2.0 * r.uniform() - 1.0;
    v2 = 2.0 * r.uniform() - 1.0;
    rsq = v1 * v1 + v2 * v2;
  } while (rsq >= 1.0 || rsq == 0.0);
  return v2 * sqrt(-2.0 * log(rsq) / rsq) * stdev + mean;
}

// poisson() (utilities.cpp:517-596): condition -1 any, 0 even, 1 odd,
// 2 not 0, 3 not 1
__device__ __forceinline__ int poisson_cond(PhiloxStream &r, double param, int cond) {
#pragma clang fp contract(off)
  if (param < 0.25 && cond == 1) {
    const double U = r.uniform();
    if (U < param / sinh(param)) return 1;
    if (U < param * (6 + param * param) / (6 * sinh(param))) return 3;
    return 5;
  }
  if (param < 0.25 && cond == 2) {
    double raised = exp(-param);
    double rcheck = raised = param * raised / (1 - raised);
    const double U = r.uniform();
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
      const long v = (long)(mig_normdev(r, param, param) + 0.5);
      i = v > 0 ? (int)(v > 2000000000L ? 2000000000L : v) : 0;
    } else {
      const double raised = exp(-param);
      double U = r.uniform();
      for (i = 0; U > raised && i < kPoissonMax; i++) U *= r.uniform();
    }
    stop = cond == -1   ? 1
           : cond == 0 ? !(i & 1)
           : cond == 1 ? (i & 1)
           : cond == 2 ? i != 0
                       : i != 1;
    if (!stop && ++tries >= kMaxTries) {  // a bound no valid rate reaches
      i = cond == 1 || cond == 2 ? 1 : 0;
      stop = 1;
    }
  } while (!stop);
  if (i > kAddMigMax) i = cond == 1 ? kAddMigMax - 1 : kAddMigMax;
  return i;
}

__device__ inline double logcosh_(double x) { return x < 100 ? log(cosh(x)) : x - kLog2; }
__device__ inline double logsinh_(double x) { return x < 100 ? log(sinh(x)) : x - kLog2; }


}  // namespace ima
