// kern_swap.hip -- Metropolis-coupled chain swap (swapchains.cpp:527-662).
//
// All heated chains live on this device, so the swap is a single-lane
// kernel over (beta_c, S_c) pairs: no messages, no barrier.  Random numbers
// come from Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), keyed by
// (seed, counter), so a swap sequence is reproducible from its key.
#include "ima_internal.h"
#include "philox.h"

namespace ima {

// swapchains: pick ci uniformly, cj within SWAPDIST = 7 of ci unless there
// are fewer than 2*SWAPDIST+3 chains, w = exp((b_i - b_j)(S_j - S_i))
// (swapweight :13-34), accept when w >= 1 or w > U, then swap the betas.
__global__ void swap_kernel(int nchains, const double *S, double *beta, int swaptries,
                            uint64_t seed, uint64_t counter, int32_t *nacc,
                            int32_t *pairs) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  PhiloxStream rng(seed, counter);
  const int SWAPDIST = 7;
  int acc = 0;
  for (int i = 0; i < swaptries; ++i) {
    int ci, cj, cjmin, cjrange;
    do {
      ci = (int)(rng.uniform() * nchains);
    } while (ci < 0 || ci >= nchains);
    if (nchains < 2 * SWAPDIST + 3) {
      cjmin = 0;
      cjrange = nchains;
    } else {
      cjmin = max(0, ci - SWAPDIST);
      cjrange = min(nchains, ci + SWAPDIST) - cjmin;
    }
    do {
      cj = cjmin + (int)(rng.uniform() * cjrange);
    } while (cj == ci || cj < 0 || cj >= nchains);
    const double w = exp((beta[ci] - beta[cj]) * (S[cj] - S[ci]));
    int ok = 0;
    if (w >= 1.0 || w > rng.uniform()) {
      const double b = beta[ci];
      beta[ci] = beta[cj];
      beta[cj] = b;
      ok = 1;
      ++acc;
    }
    if (pairs) {
      pairs[3 * i] = ci;
      pairs[3 * i + 1] = cj;
      pairs[3 * i + 2] = ok;
    }
  }
  if (nacc) *nacc = acc;
}

hipError_t launch_swap(int nchains, const double *S, double *beta, int swaptries,
                       uint64_t seed, uint64_t counter, int32_t *nacc, int32_t *pairs,
                       hipStream_t st) {
  hipLaunchKernelGGL(swap_kernel, dim3(1), dim3(64), 0, st, nchains, S, beta, swaptries,
                     seed, counter, nacc, pairs);
  return hipGetLastError();
}

}  // namespace ima
