// kern_prior.hip -- genealogy prior: per-genealogy weights (treeweight,
// update_gtree_common.cpp:1680-1931) and their integral over the uniform
// priors on population sizes and migration rates for a set of genealogies
// (integrate_tree_prob, :1945-2053).
//
// Weights.  One wavefront per genealogy (tw_dev.h).
//
// Integral.  One wavefront per group of genealogies (a chain's loci):
// lanes sum the records over the group in group order, then one lane per
// parameter adds its (period, position) entries in parameter-map order
// and evaluates the analytic integral; lane 0 adds the terms.
#pragma clang fp contract(off)

#include "ima_internal.h"
#include "prior_math.h"
#include "tw_dev.h"

namespace ima {

namespace {

__global__ void __launch_bounds__(kWave)
treeweight_kernel(DeviceView v, const int32_t *slots, int nslots, const double *tvals,
                  const int32_t *tv_row, double *rec_out, int32_t *status) {
  extern __shared__ double lds[];
  if ((int)blockIdx.x >= nslots) return;
  twdev::treeweight_one(v, slots, blockIdx.x, tvals, tv_row, rec_out, status, lds);
}

__global__ void __launch_bounds__(kWave)
prior_kernel(DeviceView v, int model, const double *recs, const int32_t *group_off,
             int ngroups, double *probg, double *qint, double *mint) {
  extern __shared__ double sums[];
  const int g = blockIdx.x;
  if (g >= ngroups) return;
  const int lane = threadIdx.x;
  const ModelMeta mm = v.models[model];
  const ModelView md = model_view(v, mm);
  const int P = mm.npops, R = prior_rec_len(P);
  const int s0 = group_off[g], s1 = group_off[g + 1];
  for (int r = lane; r < R; r += kWave) sums[r] = ordered_rec_sum(recs + (int64_t)s0 * R, R, s1 - s0, r);
  __syncthreads();
  double *terms = sums + R;  // [npsize + nmig]
  const int np = mm.npsize, nm = mm.nmig;
  for (int t = lane; t < np + nm; t += kWave) {
    int c;
    double f, h;
    prior_param_sums(md, P, np, t, sums, &c, &f, &h);
    terms[t] = prior_param_term(md, np, t, c, f, h, v.logfact);
  }
  __syncthreads();
  if (lane == 0) {
    double q = 0.0, m = 0.0;
    for (int t = 0; t < np; ++t) q += terms[t];
    for (int t = 0; t < nm; ++t) m += terms[np + t];
    probg[g] = q + m;
  }
  for (int t = lane; t < np; t += kWave) qint[(int64_t)g * np + t] = terms[t];
  for (int t = lane; t < nm; t += kWave) mint[(int64_t)g * nm + t] = terms[np + t];
}

}  // namespace

size_t treeweight_lds_bytes(int P, int emax) {
  // record, unsorted and sorted events, one chunk's terms
  return (size_t)prior_rec_len(P) * 8 + (size_t)emax * 24 +
         (size_t)(2 * (2 * P - 1) + 1) * kWave * 8;
}

hipError_t launch_treeweight(const DeviceView &v, const int32_t *slots, int nslots,
                             const double *tvals, const int32_t *tv_row, int P, int emax,
                             double *rec_out, int32_t *status, hipStream_t st) {
  if (nslots <= 0) return hipSuccess;
  hipLaunchKernelGGL(treeweight_kernel, dim3(nslots), dim3(kWave),
                     treeweight_lds_bytes(P, emax), st, v, slots, nslots, tvals, tv_row,
                     rec_out, status);
  return hipGetLastError();
}

hipError_t launch_prior(const DeviceView &v, int model, int P, int nparams, const double *recs,
                        const int32_t *group_off, int ngroups, double *probg, double *qint,
                        double *mint, hipStream_t st) {
  if (ngroups <= 0) return hipSuccess;
  const size_t lds = ((size_t)prior_rec_len(P) + nparams) * 8;
  hipLaunchKernelGGL(prior_kernel, dim3(ngroups), dim3(kWave), lds, st, v, model, recs,
                     group_off, ngroups, probg, qint, mint);
  return hipGetLastError();
}

}  // namespace ima
