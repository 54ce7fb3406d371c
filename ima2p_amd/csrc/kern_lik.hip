// kern_lik.hip -- batched per-locus data likelihood kernels for gfx950.
//
//   IS   infinite-sites mutation placement (calc_prob_data.cpp:537-836)
//   SW   stepwise-mutation Bessel terms (calc_prob_data.cpp:842-909,
//        utilities.cpp:55-124, 1453-1491)
//
// Every kernel runs one 64-lane wavefront per scoring item (one genealogy
// of one chain at one locus).  Items of all chains and loci, with uneven
// sample sizes and pattern counts, go into a single launch.
#include <float.h>

#include "ima_internal.h"

namespace ima {

// ------------------------------------------------------------ helpers --

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ int wave_or(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off, kWave);
  return v;
}

// ================================================================= IS ==
//
// The reference labels every site bottom-up (labelgtree) and looks for the
// nodes whose daughters are labelled 0 and 1; one such node means the site
// fits the genealogy, two mean rejection.  Those labels are Fitch parsimony
// sets, so exactly one such node exists iff the tips carrying 1 (or those
// carrying 0) are exactly the clade below one edge x.  The 0/1 node is then
// down(x) and the mutation sits on x, or on both root daughters when
// down(x) is the root (calc_prob_data.cpp:790-818).  Each lane takes one
// site and compares its tip masks against every edge's clade mask.

template <int W>
__global__ __launch_bounds__(kWave) void is_kernel(
    DeviceView v, const ima_item *__restrict__ items,
    const int32_t *__restrict__ idx, double *__restrict__ out,
    int32_t *__restrict__ status, int32_t *__restrict__ site_node,
    int32_t *__restrict__ counts, const int32_t *__restrict__ list_cnt, int32_t *cnt_next) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x;
  if (list_cnt) {
    if (blockIdx.x == 0 && lane == 0) *cnt_next = 0;
    if ((int)blockIdx.x >= *list_cnt) return;
  }
  const int ii = idx[blockIdx.x];
  const ima_item item = items[ii];
  const SlotMeta sm = v.slots[item.slot];
  const LocusMeta lm = v.loci[sm.locus];
  const int n = sm.n, ne = 2 * n - 1, S = lm.S;
  unsigned long long *clade = reinterpret_cast<unsigned long long *>(smem);  // [ne][W]
  double *bl = reinterpret_cast<double *>(clade + (size_t)ne * W);           // [ne]
  int32_t *dn = reinterpret_cast<int32_t *>(bl + ne);                        // [ne]
  int32_t *u0 = dn + ne;
  int32_t *u1 = u0 + ne;
  int32_t *cnt = u1 + ne;                                                    // [ne]

  const int32_t *gup0 = v.up0 + sm.node_off, *gup1 = v.up1 + sm.node_off;
  const int32_t *gdn = v.down + sm.node_off;
  const double *tm = v.time + sm.node_off;
  double len = 0.0;
  for (int e = lane; e < ne; e += kWave) {
    const int a = gup0[e];
    dn[e] = gdn[e];
    u0[e] = a;
    u1[e] = gup1[e];
    cnt[e] = 0;
    const double b = (gdn[e] == -1) ? 0.0 : tm[e] - (a >= 0 ? tm[a] : 0.0);
    bl[e] = b;
    len += b;
#pragma unroll
    for (int w = 0; w < W; ++w) clade[(size_t)e * W + w] = 0ull;
  }
  len = wave_sum(len);
  __syncthreads();
  for (int t = lane; t < n; t += kWave) {
    const unsigned long long bit = 1ull << (t & 63);
    for (int e = t; e != -1; e = dn[e]) atomicOr(&clade[(size_t)e * W + (t >> 6)], bit);
  }
  __syncthreads();

  unsigned long long full[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const int lo = w * 64, hi = min(n, lo + 64);
    full[w] = (hi <= lo) ? 0ull : (hi - lo == 64 ? ~0ull : ((1ull << (hi - lo)) - 1ull));
  }
  const double u = item.u;
  double acc = 0.0;
  int bad = 0;  // bit0: reject, bit1: no 0/1 node
  const uint64_t *mask = v.mask + lm.mask_off;
  const int Wl = lm.W;  // this locus' mask stride (<= W)
  for (int s0 = 0; s0 < S; s0 += kWave) {
    const int s = s0 + lane;
    if (s < S) {
      unsigned long long m1[W], m0[W];
#pragma unroll
      for (int w = 0; w < W; ++w) {
        m1[w] = (w < Wl) ? mask[(size_t)s * Wl + w] : 0ull;
        m0[w] = full[w] & ~m1[w];
      }
      int x = -1, y = -1;
      for (int e = 0; e < ne; ++e) {
        if (e == sm.root) continue;
        bool eq1 = true, eq0 = true;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const unsigned long long c = clade[(size_t)e * W + w];
          eq1 = eq1 && (c == m1[w]);
          eq0 = eq0 && (c == m0[w]);
        }
        x = eq1 ? e : x;
        y = eq0 ? e : y;
      }
      const int c = (x >= 0) ? x : y;
      if (c < 0) {
        bool empty = true;
#pragma unroll
        for (int w = 0; w < W; ++w) empty = empty && (m1[w] == 0ull);
        bad |= empty ? 2 : 1;
        if (site_node) site_node[s] = -1;
      } else {
        const int N = dn[c];
        const int sis = (u0[N] == c) ? u1[N] : u0[N];
        const double ptime = (dn[N] == -1) ? bl[c] + bl[sis] : bl[c];
        acc += log(ptime * u);
        if (site_node) site_node[s] = N;
        if (counts) atomicAdd(&cnt[N], 1);
      }
    }
  }
  acc = wave_sum(acc);
  bad = wave_or(bad);
  __syncthreads();
  if (counts)
    for (int e = lane; e < ne; e += kWave) counts[e] = cnt[e];
  if (lane == 0) {
    if (bad & 2) {
      out[ii] = IMA_REJECT_IS;
      if (status) status[ii] = IMA_STATUS_IS_NOSITE;
    } else if (bad & 1) {
      out[ii] = IMA_REJECT_IS;
      if (status) status[ii] = IMA_STATUS_IS_REJECT;
    } else {
      out[ii] = -len * u + acc - lm.sumlogk;
      if (status) status[ii] = IMA_STATUS_OK;
    }
  }
}

hipError_t launch_is(const DeviceView &v, const ima_item *items,
                     const int32_t *item_idx, int nitems, int max_n, int max_w,
                     double *out, int32_t *status, int32_t *site_node,
                     int32_t *counts, hipStream_t st, const int32_t *cnt,
                     int32_t *cnt_next) {
  if (nitems == 0) return hipSuccess;
  const int ne = 2 * max_n - 1;
  const size_t lds = (size_t)ne * max_w * 8 + (size_t)ne * 8 + (size_t)ne * 4 * 4;
  switch (max_w) {
    case 1:
      hipLaunchKernelGGL(is_kernel<1>, dim3(nitems), dim3(kWave), lds, st, v, items,
                         item_idx, out, status, site_node, counts, cnt, cnt_next);
      break;
    case 2:
      hipLaunchKernelGGL(is_kernel<2>, dim3(nitems), dim3(kWave), lds, st, v, items,
                         item_idx, out, status, site_node, counts, cnt, cnt_next);
      break;
    default:
      hipLaunchKernelGGL(is_kernel<4>, dim3(nitems), dim3(kWave), lds, st, v, items,
                         item_idx, out, status, site_node, counts, cnt, cnt_next);
      break;
  }
  return hipGetLastError();
}

// ================================================================= SW ==
//
// Modified Bessel functions I_n in the reference's own method:
// Abramowitz & Stegun 9.8.1-9.8.4 polynomial approximations for I_0 and
// I_1 and Miller's downward recurrence (start 2(n + floor(sqrt(40 n))),
// renormalised at 1e10) for n >= 2, as in utilities.cpp:55-124, 1453-1491.

__device__ double d_bessi0(double x) {
  const double ax = fabs(x);
  if (ax < 3.75) {
    double y = x / 3.75;
    y *= y;
    return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 +
           y * (0.2659732 + y * (0.360768e-1 + y * 0.45813e-2)))));
  }
  const double y = 3.75 / ax;
  return (exp(ax) / sqrt(ax)) *
         (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 +
          y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 +
          y * (0.2635537e-1 + y * (-0.1647633e-1 + y * 0.392377e-2))))))));
}

__device__ double d_bessi1(double x) {
  const double ax = fabs(x);
  double ans;
  if (ax < 3.75) {
    double y = x / 3.75;
    y *= y;
    ans = ax * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 +
          y * (0.2658733e-1 + y * (0.301532e-2 + y * 0.32411e-3))))));
  } else {
    const double y = 3.75 / ax;
    ans = 0.2282967e-1 + y * (-0.2895312e-1 + y * (0.1787654e-1 - y * 0.420059e-2));
    ans = 0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 +
          y * (0.163801e-2 + y * (-0.1031555e-1 + y * ans))));
    ans *= (exp(ax) / sqrt(ax));
  }
  return x < 0.0 ? -ans : ans;
}

__device__ double d_bessi(int n, double x) {
  n = abs(n);
  if (x > 700) return DBL_MAX / 1e10;
  if (n == 0) return d_bessi0(x);
  if (n == 1) return d_bessi1(x);
  if (x == 0.0) return 0.0;
  const double tox = 2.0 / fabs(x);
  double bip = 0.0, ans = 0.0, bi = 1.0;
  for (int j = 2 * (n + (int)sqrt(40.0 * n)); j > 0; j--) {
    const double bim = bip + j * tox * bi;
    bip = bi;
    bi = bim;
    if (fabs(bi) > 1.0e10) {
      ans *= 1.0e-10;
      bi *= 1.0e-10;
      bip *= 1.0e-10;
    }
    if (j == n) ans = bip;
  }
  ans *= d_bessi0(x) / bi;
  return (x < 0.0 && (n & 1)) ? -ans : ans;
}

__global__ __launch_bounds__(kWave) void sw_kernel(
    DeviceView v, const ima_item *__restrict__ items,
    const int32_t *__restrict__ idx, double *__restrict__ out,
    int32_t *__restrict__ status, double *__restrict__ dlike, const int32_t *__restrict__ cnt,
    int32_t *cnt_next) {
  const int lane = threadIdx.x;
  if (cnt) {
    if (blockIdx.x == 0 && lane == 0) *cnt_next = 0;
    if ((int)blockIdx.x >= *cnt) return;
  }
  const int ii = idx[blockIdx.x];
  const ima_item item = items[ii];
  const SlotMeta sm = v.slots[item.slot];
  const int n = sm.n, ne = 2 * n - 1;
  const int32_t *up0 = v.up0 + sm.node_off, *dn = v.down + sm.node_off;
  const int32_t *A = v.A + sm.node_off;
  const double *tm = v.time + sm.node_off;
  const double u = item.u, tr = item.tr;
  double acc = 0.0;
  int zero = 0;
  for (int e = lane; e < ne; e += kWave) {
    double d_l = 0.0;
    if (dn[e] != -1) {
      const int d = A[e] - A[dn[e]];
      const double t = (up0[e] == -1) ? tm[e] * tr : tr * (tm[e] - tm[up0[e]]);
      const double b = d_bessi(d, t * u);
      if (!(b > 0.0)) {
        zero = 1;
        d_l = -1e+100;  // IM_BESSI_MIN imamp.hpp:173
      } else {
        d_l = -(t * u) + log(b);
      }
    }
    if (dlike) dlike[e] = d_l;
    acc += d_l;
  }
  acc = wave_sum(acc);
  zero = wave_or(zero);
  if (lane == 0) {
    out[ii] = zero ? -DBL_MAX : acc;
    if (status) status[ii] = zero ? IMA_STATUS_SW_ZERO : IMA_STATUS_OK;
  }
}

// finishSWupdateA (update_gtree_common.cpp:2219-2362), one lane per move,
// in the reference's order of operations.  Pass 0 checks every move's
// structure (status 1: edge and newsisedge are not under downedge) and
// writes nothing else; pass 1, run only when all moves passed, proposes,
// stores and scores.
__device__ inline double sw_upper(const int32_t *up0, const double *tm, int n, int e) {
  return e >= n ? tm[up0[e]] : 0.0;
}

__global__ __launch_bounds__(kWave) void sw_update_kernel(
    DeviceView v, int32_t *__restrict__ Aw, const ima_sw_move *__restrict__ moves, int nmoves,
    int pass, int calc, double *__restrict__ out, int32_t *__restrict__ iout) {
  const int i = blockIdx.x * kWave + threadIdx.x;
  if (i >= nmoves) return;
  const ima_sw_move m = moves[i];
  const SlotMeta sm = v.slots[m.slot];
  const int n = sm.n;
  const int32_t *up0 = v.up0 + sm.node_off, *dn = v.down + sm.node_off;
  const double *tm = v.time + sm.node_off;
  int32_t *A = Aw + sm.node_off;  // the view's allele array, writable
  if (pass == 0) {
    iout[2 * i + 1] = (dn[m.edge] == m.downedge && dn[m.newsisedge] == m.downedge) ? 0 : 1;
    return;
  }
  const int oldA = A[m.downedge];
  const double holdsis = m.newsisedge != m.sisedge ? m.newsis_dlikeA : 0.0;
  const double oldlikeadj = m.copy_dlikeA[0] + m.copy_dlikeA[1] + m.copy_dlikeA[2] + holdsis;
  const int e[3] = {m.edge, m.newsisedge, dn[m.edge]};
  double t[3] = {0.0, 0.0, 0.0};
  int wsum = 0, j = 0;
  for (int k = 0; k < 3; ++k) {
    if (dn[e[k]] == -1) continue;
    t[k] = tm[e[k]] - sw_upper(up0, tm, n, e[k]);
    const int d = (k < 2 ? A[e[k]] : A[dn[e[k]]]) - oldA;
    wsum += abs(d);
    ++j;
  }
  double gnew = j / (double)(wsum + j);
  if (gnew > 0.95) gnew = 0.95;
  int dA = (int)ceil(log(m.uni) / log(1.0 - gnew)) - 1;
  if (m.flip) dA = -dA;
  const int newA = dA >= 0 ? min(m.maxA, oldA + dA) : max(m.minA, oldA + dA);
  A[m.downedge] = newA;
  dA = newA - oldA;
  double sis_dl = 0.0;
  if (dn[m.sisedge] != -1 && m.sisedge != m.newsisedge) {
    const double ts = tm[m.sisedge] - sw_upper(up0, tm, n, m.sisedge);
    const int d = A[m.sisedge] - A[dn[m.sisedge]];
    sis_dl = -ts * m.u + log(d_bessi(d, ts * m.u));
  }
  double likeadj = sis_dl;
  double dl[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < 3; ++k) {
    if (dn[e[k]] == -1) continue;
    const int d = (k < 2 ? A[e[k]] : A[dn[e[k]]]) - newA;
    dl[k] = -(t[k] * m.u) + log(d_bessi(d, t[k] * m.u));
    likeadj += dl[k];
  }
  int rsum = abs(m.copyA0 - newA) + abs(m.copyA1 - newA), r = 2;
  if (!m.down_was_root) {
    rsum += abs(m.hold_down_A - newA);
    ++r;
  }
  double gold = r / (double)(rsum + r);
  if (gold > 0.95) gold = 0.95;
  double *o = out + 6 * i;
  o[0] = calc ? likeadj - oldlikeadj : 0.0;
  o[1] = calc ? (abs(dA) * log(1 - gold) + log(gold)) - (abs(dA) * log(1 - gnew) + log(gnew))
              : 0.0;
  o[2] = m.sisedge == m.newsisedge ? dl[1] : sis_dl;
  o[3] = dl[0];
  o[4] = dl[1];
  o[5] = dl[2];
  iout[2 * i] = newA;
}

hipError_t launch_sw_update(const DeviceView &v, int32_t *A, const ima_sw_move *moves,
                            int nmoves, int pass, int calc, double *out, int32_t *iout,
                            hipStream_t st) {
  if (nmoves == 0) return hipSuccess;
  hipLaunchKernelGGL(sw_update_kernel, dim3((nmoves + kWave - 1) / kWave), dim3(kWave), 0, st,
                     v, A, moves, nmoves, pass, calc, out, iout);
  return hipGetLastError();
}

hipError_t launch_sw(const DeviceView &v, const ima_item *items,
                     const int32_t *item_idx, int nitems, double *out,
                     int32_t *status, double *dlike, hipStream_t st, const int32_t *cnt,
                     int32_t *cnt_next) {
  if (nitems == 0) return hipSuccess;
  hipLaunchKernelGGL(sw_kernel, dim3(nitems), dim3(kWave), 0, st, v, items, item_idx,
                     out, status, dlike, cnt, cnt_next);
  return hipGetLastError();
}

}  // namespace ima
