// kern_lmode.hip -- L-mode surfaces over stored genealogy tables.
//
//   jointp      jointfind.cpp:885-1047   joint -log posterior at a point
//   margincalc  surface_call_functions.cpp:120-173   marginal for one param
//
// Tables hold one row of K floats per saved genealogy in the savegsampinf
// layout (ginfo.cpp:288-316): cc | fc | hcc (npsize each), mc | fm (nmig
// each), qintegrate (npsize), mintegrate (nmig), pdg, probg, split times.
//
// Mapping (jointp): one workgroup of 256 lanes per (tile of points, chunk
// of rows), each lane carrying one or two points.  The workgroup stages
// its rows in LDS as compact FP64 records that no longer depend on the
// point, and every lane walks them in order, so each row is read from HBM
// once per point tile and broadcast from LDS; a point's coefficients sit
// in registers.  Chunk partials are merged by a second kernel in chunk
// order.  The chunk size depends on the table only, never on how many
// points share a launch, so a point's result is the same alone and batched
// with others.  margincalc: one thread per value, 64-value tiles.
#include <float.h>

#include "ima_internal.h"

namespace ima {

constexpr int kLTile = 64;      // points per workgroup
constexpr int kLStage = 128;    // rows staged in LDS at a time

struct JointLayout {
  int ccp, fcp, hccp, mcp, fmp, qip, mip, pdgp, probgp;
};

__host__ __device__ inline JointLayout joint_layout(int npsize, int nmig) {
  JointLayout L;
  L.ccp = 0;
  L.fcp = L.ccp + npsize;
  L.hccp = L.fcp + npsize;
  L.mcp = L.hccp + npsize;
  L.fmp = L.mcp + nmig;
  L.qip = L.fmp + nmig;
  L.mip = L.qip + npsize;
  L.pdgp = L.mip + nmig;
  L.probgp = L.pdgp + 1;
  return L;
}

// eexp (utilities.cpp:1507-1539): e^x as m * 10^z, with a degree-10 Taylor
// series of e^u on u = x - n ln2 in [0, ln2).  Coefficients are 1/k!.
__device__ inline void d_eexp(double x, double *m, int *z) {
  const double LOG_10_2 = 0.30102999566398119521;
  const int nn = (int)floor(x / kLn2);
  double zr = LOG_10_2 * (double)nn;
  *z = (int)zr;
  zr -= (double)*z;
  const double u = x - (((double)nn) * kLn2);
  const double temp = 1 + u * (1.0 + u * (0.5 + u * (0.16666666666666666666666666667 +
      u * (0.04166666666666666666666666667 + u * (0.00833333333333333333333333333 +
      u * (0.001388888888888888888888888889 + u * (0.000198412698412698412698412698 +
      u * (0.000024801587301587301587301587301 + u * (2.75573192239858906525573192239859e-6 +
      u * (2.75573192239858906525573192239e-7))))))))));
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

struct LArgs {
  const float *rows;
  int G, K, npsize, nmig, npops, modeltype, r0, r1;
  const double *x;  // [npoints][np]
  int npoints, nchunks, chunk_rows, stage;
};

// --------------------------------------------------------------- EXACT --
// Per (point, chunk): running max M and s = sum exp(p - M), s2 = sum
// exp(2(p - M)) for the effective sample number.
// Staging folds each row into a compact FP64
// record that no longer depends on the point: c0 = (the row's base term)
// minus the hcc of the parameters in range, then (count, weight) of each
// parameter in [r0, r1).  A point's term is then
//   p = c0 + sum_k count_k * A_k + weight_k * F_k
// with A = log 2 - log x (theta, models 0/1; else 0) or log x (migration),
// F = -2/x (theta: 2 fc / x) or -x (migration); (2/x) fc equals (2 fc)/x
// exactly, the remaining differences from the reference's order are the
// fused multiply-adds and the summation order (about 1e-16 relative).
// 256 lanes per workgroup share one staged block; each lane carries PTS
// points, so every record read from LDS serves PTS independent exp chains.
constexpr int kXLanes = 256;

__host__ __device__ inline int xrec_len(int nr) { return 1 + 2 * nr; }

__device__ inline void stage_compact(const LArgs &a, double *sh, int g0, int cnt, int nr) {
  const JointLayout L = joint_layout(a.npsize, a.nmig);
  const int rl = xrec_len(nr);
  for (int r = threadIdx.x; r < cnt; r += blockDim.x) {
    const float *g = a.rows + (size_t)(g0 + r) * a.K;
    double base;
    if (a.npops == 2) {
      base = -(double)g[L.probgp];
    } else {
      double t = 0.0;
      if (a.modeltype == 1)
        for (int i = 0; i < a.npsize; ++i) t += (double)g[L.qip + i];
      else
        for (int i = 0; i < a.nmig; ++i) t += (double)g[L.mip + i];
      base = -t;
    }
    double *o = sh + (size_t)r * rl;
    for (int k = 0; k < nr; ++k) {
      const int i = a.r0 + k;
      if (i < a.npsize) {
        base -= (double)g[L.hccp + i];
        o[1 + 2 * k] = (double)g[L.ccp + i];
        o[2 + 2 * k] = (double)g[L.fcp + i];
      } else {
        o[1 + 2 * k] = (double)g[L.mcp + i - a.npsize];
        o[2 + 2 * k] = (double)g[L.fmp + i - a.npsize];
      }
    }
    o[0] = base;
  }
}

// Per-point coefficients of the compact records: A = log 2 - log x
// (theta, models 0/1; else 0) or log x (migration), F = -2/x or -x.
template <int NR>
__device__ inline void point_coef_compact(const LArgs &a, const double *xp, double (&A)[NR],
                                          double (&F)[NR]) {
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    const int i = a.r0 + k;
    const double x = xp[i];
    if (i < a.npsize) {
      A[k] = (a.modeltype == 0 || a.modeltype == 1) ? kLn2 - log(x) : 0.0;
      F[k] = -2.0 * (1.0 / x);
    } else {
      A[k] = log(x);
      F[k] = -x;
    }
  }
}

template <int NR>
__device__ inline double compact_term(const double *v, const double (&A)[NR],
                                      const double (&F)[NR]) {
  double p = v[0];
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    p = fma(v[1 + 2 * k], A[k], p);
    p = fma(v[2 + 2 * k], F[k], p);
  }
  return p;
}

template <int NR, int PTS>
__global__ __launch_bounds__(kXLanes) void jointp_exact_fast(LArgs a, double *work) {
  extern __shared__ double shrows[];
  const int np = a.npsize + a.nmig;
  double A[PTS][NR], F[PTS][NR], M[PTS], s[PTS], s2[PTS];
  int pt[PTS];
#pragma unroll
  for (int j = 0; j < PTS; ++j) {
    pt[j] = blockIdx.x * (kXLanes * PTS) + j * kXLanes + threadIdx.x;
    const double *xp = a.x + (size_t)(pt[j] < a.npoints ? pt[j] : 0) * np;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      const int i = a.r0 + k;
      const double x = xp[i];
      if (i < a.npsize) {
        A[j][k] = (a.modeltype == 0 || a.modeltype == 1) ? kLn2 - log(x) : 0.0;
        F[j][k] = -2.0 * (1.0 / x);
      } else {
        A[j][k] = log(x);
        F[j][k] = -x;
      }
    }
    M[j] = -DBL_MAX;
    s[j] = 0.0;
    s2[j] = 0.0;
  }
  const int chunk = blockIdx.y;
  const int gbeg = chunk * a.chunk_rows;
  const int gend = min(a.G, gbeg + a.chunk_rows);
  constexpr int RL = 1 + 2 * NR;
  for (int g0 = gbeg; g0 < gend; g0 += a.stage) {
    const int cnt = min(a.stage, gend - g0);
    __syncthreads();
    stage_compact(a, shrows, g0, cnt, NR);
    __syncthreads();
    for (int r = 0; r < cnt; ++r) {
      const double *o = shrows + (size_t)r * RL;
      double v[RL];
#pragma unroll
      for (int q = 0; q < RL; ++q) v[q] = o[q];
#pragma unroll
      for (int j = 0; j < PTS; ++j) {
        double p = v[0];
#pragma unroll
        for (int k = 0; k < NR; ++k) {
          p = fma(v[1 + 2 * k], A[j][k], p);
          p = fma(v[2 + 2 * k], F[j][k], p);
        }
        if (p > M[j]) {
          const double sc = exp(M[j] - p);
          s[j] = s[j] * sc + 1.0;
          s2[j] = s2[j] * sc * sc + 1.0;
          M[j] = p;
        } else {
          const double w = exp(p - M[j]);
          s[j] += w;
          s2[j] = fma(w, w, s2[j]);
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < PTS; ++j)
    if (pt[j] < a.npoints) {
      double *o = work + ((size_t)pt[j] * a.nchunks + chunk) * 3;
      o[0] = M[j];
      o[1] = s[j];
      o[2] = s2[j];
    }
}

template <int NR>
hipError_t launch_exact_fast(const LArgs &a, dim3 grid_y, size_t lds, hipStream_t st,
                             double *work) {
  constexpr int PTS = NR <= 8 ? 2 : 1;
  const dim3 grid((a.npoints + kXLanes * PTS - 1) / (kXLanes * PTS), grid_y.y);
  hipLaunchKernelGGL((jointp_exact_fast<NR, PTS>), grid, dim3(kXLanes), lds, st, a, work);
  return hipSuccess;
}

__global__ void jointp_exact_final(LArgs a, const double *work, double *out, double *ess) {
  const int pt = blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= a.npoints) return;
  const double *w = work + (size_t)pt * a.nchunks * 3;
  double M = -DBL_MAX;
  for (int c = 0; c < a.nchunks; ++c) M = fmax(M, w[c * 3]);
  double s = 0.0, s2 = 0.0;
  for (int c = 0; c < a.nchunks; ++c) {
    if (w[c * 3 + 1] == 0.0) continue;
    const double sc = exp(w[c * 3] - M);
    s += w[c * 3 + 1] * sc;
    s2 += w[c * 3 + 2] * sc * sc;
  }
  out[pt] = log((double)a.G) - (M + log(s));
  if (ess) ess[pt] = s * s / s2;
}

// ---------------------------------------------------------- REF_COMPAT --
// The reference keeps only values within PRANGELOG = 10 nats of the running
// maximum (the first value always), then sums at most (kept - 1) of them
// from the top while they stay within 10 nats of the final maximum, as
// eexp mantissa/exponent pairs scaled to 10^(z - (z_max - 10)).  Every
// value within the final window is always kept, so the sum covers the
// whole final window, minus its smallest member when nothing outside the
// window was kept and the window holds more than one value.
//
// pass 1: chunk maxima; pass 2: per chunk, the running maximum starts at
// the maximum of the earlier chunks and each lane walks its rows in order;
// pass 3: merge chunks in order and apply the list cap.

// Per (point, chunk) record of pass 2.
struct RefPartial {
  double sum, sum2;   // sum of scaled eexp terms (and squares) in the window
  double minp, minterm;
  int32_t cnt;        // values in the final window
  int32_t outside;    // a kept value lies outside the final window
};

// REF_COMPAT from the same compact records as EXACT, PTS points per lane.
// Pass 1: chunk maxima.
template <int NR, int PTS>
__global__ __launch_bounds__(kXLanes) void jointp_ref_fast1(LArgs a, double *cmax) {
  extern __shared__ double shrows[];
  const int np = a.npsize + a.nmig;
  double A[PTS][NR], F[PTS][NR], M[PTS];
  int pt[PTS];
#pragma unroll
  for (int j = 0; j < PTS; ++j) {
    pt[j] = blockIdx.x * (kXLanes * PTS) + j * kXLanes + threadIdx.x;
    point_coef_compact<NR>(a, a.x + (size_t)(pt[j] < a.npoints ? pt[j] : 0) * np, A[j], F[j]);
    M[j] = -DBL_MAX;
  }
  const int chunk = blockIdx.y;
  const int gbeg = chunk * a.chunk_rows;
  const int gend = min(a.G, gbeg + a.chunk_rows);
  constexpr int RL = 1 + 2 * NR;
  for (int g0 = gbeg; g0 < gend; g0 += a.stage) {
    const int cnt = min(a.stage, gend - g0);
    __syncthreads();
    stage_compact(a, shrows, g0, cnt, NR);
    __syncthreads();
    for (int r = 0; r < cnt; ++r) {
      const double *o = shrows + (size_t)r * RL;
      double v[RL];
#pragma unroll
      for (int q = 0; q < RL; ++q) v[q] = o[q];
#pragma unroll
      for (int j = 0; j < PTS; ++j) M[j] = fmax(M[j], compact_term<NR>(v, A[j], F[j]));
    }
  }
#pragma unroll
  for (int j = 0; j < PTS; ++j)
    if (pt[j] < a.npoints) cmax[(size_t)pt[j] * a.nchunks + chunk] = M[j];
}

// Pass 2: per (point, chunk), the window sums.  A kept value's term is
// e^p / 10^maxz (the reference's eexp mantissa m 10^(z - maxz)), taken here
// as exp(p - maxz ln 10): the window, the list cap and the ordering of the
// running maximum are the reference's, the term itself differs from the
// degree-10 Taylor series of eexp by at most its truncation, ~4e-10
// relative, which moves the result by less than 1e-12 relative.
template <int NR, int PTS>
__global__ __launch_bounds__(kXLanes) void jointp_ref_fast2(LArgs a, const double *cmax,
                                                            RefPartial *part) {
  extern __shared__ double shrows[];
  const int np = a.npsize + a.nmig;
  const int chunk = blockIdx.y;
  double A[PTS][NR], F[PTS][NR], Mfin[PTS], run[PTS], sc[PTS];
  RefPartial r[PTS];
  int pt[PTS];
#pragma unroll
  for (int j = 0; j < PTS; ++j) {
    pt[j] = blockIdx.x * (kXLanes * PTS) + j * kXLanes + threadIdx.x;
    const int ptc = pt[j] < a.npoints ? pt[j] : 0;
    point_coef_compact<NR>(a, a.x + (size_t)ptc * np, A[j], F[j]);
    const double *cm = cmax + (size_t)ptc * a.nchunks;
    double mf = -DBL_MAX, ru = -DBL_MAX;
    for (int c = 0; c < a.nchunks; ++c) {
      mf = fmax(mf, cm[c]);
      if (c < chunk) ru = fmax(ru, cm[c]);
    }
    Mfin[j] = mf;
    run[j] = ru;
    double mM;
    int zM;
    d_eexp(mf, &mM, &zM);
    sc[j] = (double)(zM - 10) * 2.3025850929940456840;  // maxz ln 10
    r[j] = RefPartial{0.0, 0.0, DBL_MAX, 0.0, 0, 0};
  }
  const int gbeg = chunk * a.chunk_rows;
  const int gend = min(a.G, gbeg + a.chunk_rows);
  constexpr int RL = 1 + 2 * NR;
  for (int g0 = gbeg; g0 < gend; g0 += a.stage) {
    const int cnt = min(a.stage, gend - g0);
    __syncthreads();
    stage_compact(a, shrows, g0, cnt, NR);
    __syncthreads();
    for (int rr = 0; rr < cnt; ++rr) {
      const double *o = shrows + (size_t)rr * RL;
      double v[RL];
#pragma unroll
      for (int q = 0; q < RL; ++q) v[q] = o[q];
      const int g = g0 + rr;
#pragma unroll
      for (int j = 0; j < PTS; ++j) {
        const double p = compact_term<NR>(v, A[j], F[j]);
        const bool kept = (g == 0) || (run[j] - p < 10);
        if (kept && p > run[j]) run[j] = p;
        if (Mfin[j] - p < 10) {
          const double term = exp(p - sc[j]);
          r[j].sum += term;
          r[j].sum2 += term * term;
          r[j].cnt += 1;
          if (p < r[j].minp) {
            r[j].minp = p;
            r[j].minterm = term;
          }
        } else if (kept) {
          r[j].outside = 1;
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < PTS; ++j)
    if (pt[j] < a.npoints) part[(size_t)pt[j] * a.nchunks + chunk] = r[j];
}

template <int NR>
hipError_t launch_ref_fast(const LArgs &a, dim3 grid_y, size_t lds, hipStream_t st,
                           double *cmax, RefPartial *part) {
  constexpr int PTS = NR <= 8 ? 2 : 1;
  const dim3 grid((a.npoints + kXLanes * PTS - 1) / (kXLanes * PTS), grid_y.y);
  hipLaunchKernelGGL((jointp_ref_fast1<NR, PTS>), grid, dim3(kXLanes), lds, st, a, cmax);
  hipLaunchKernelGGL((jointp_ref_fast2<NR, PTS>), grid, dim3(kXLanes), lds, st, a, cmax, part);
  return hipSuccess;
}

__global__ void jointp_ref_final(LArgs a, const double *cmax, const RefPartial *part,
                                 double *out, double *ess) {
  const int pt = blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= a.npoints) return;
  const RefPartial *P = part + (size_t)pt * a.nchunks;
  const double *cm = cmax + (size_t)pt * a.nchunks;
  double Mfin = -DBL_MAX;
  for (int c = 0; c < a.nchunks; ++c) Mfin = fmax(Mfin, cm[c]);
  double mM;
  int zM;
  d_eexp(Mfin, &mM, &zM);
  const int maxz = zM - 10;
  double sum = 0.0, sum2 = 0.0, minp = DBL_MAX, minterm = 0.0;
  int cnt = 0, outside = 0;
  for (int c = 0; c < a.nchunks; ++c) {
    sum += P[c].sum;
    sum2 += P[c].sum2;
    cnt += P[c].cnt;
    outside |= P[c].outside;
    if (P[c].cnt > 0 && P[c].minp < minp) {
      minp = P[c].minp;
      minterm = P[c].minterm;
    }
  }
  if (!outside && cnt > 1) {  // the list cap drops the smallest kept value
    sum -= minterm;
    sum2 -= minterm * minterm;
  }
  out[pt] = log((double)a.G) - (log(sum) + maxz * 2.3025850929940456840);
  if (ess) ess[pt] = sum * sum / sum2;
}

// Rows staged in LDS at a time (FP64 rows, at most 32 KB) and rows per
// chunk: a function of the table alone, aiming at about 256 chunks.
static int stage_of(const LTable &t) {
  int st = 32768 / (8 * t.K);
  st = st >= kLStage ? kLStage : (st / 8) * 8;
  return st < 8 ? 8 : st;
}
static void chunking(const LTable &t, int npoints, int *nchunks, int *chunk_rows) {
  (void)npoints;
  const int st = stage_of(t);
  int rows = (t.G + 255) / 256;
  rows = ((rows + st - 1) / st) * st;
  if (rows < st) rows = st;
  *chunk_rows = rows;
  *nchunks = (t.G + rows - 1) / rows;
}

size_t jointp_lds_bytes(const LTable &t) { return (size_t)stage_of(t) * t.K * 8; }

// Compact records of the fast EXACT kernel: stage_of(t) rows of at most
// 1 + 2 * 16 doubles.
static size_t exact_lds_bytes(const LTable &t, int nr) {
  return (size_t)stage_of(t) * xrec_len(nr) * 8;
}

size_t jointp_work_bytes(const LTable &t, int npoints) {
  int nc, cr;
  chunking(t, npoints, &nc, &cr);
  return (size_t)npoints * nc * (sizeof(double) + sizeof(RefPartial)) + 64;
}

hipError_t launch_jointp(const LTable &t, const double *x, int npoints, int modeltype,
                         int r0, int r1, int mode, double *work, size_t work_bytes,
                         double *out, double *ess, hipStream_t st) {
  if (npoints == 0) return hipSuccess;
  LArgs a{t.rows, t.G, t.K, t.npsize, t.nmig, t.npops, modeltype, r0, r1, x, npoints, 0, 0,
          stage_of(t)};
  chunking(t, npoints, &a.nchunks, &a.chunk_rows);
  if (jointp_work_bytes(t, npoints) > work_bytes) return hipErrorInvalidValue;
  const dim3 grid((npoints + kLTile - 1) / kLTile, a.nchunks);
  const size_t lds = jointp_lds_bytes(t);
  const int fin_blocks = (npoints + 255) / 256;
  if (mode == IMA_LMODE_EXACT) {
    const int nr = r1 - r0;
    const size_t xl = exact_lds_bytes(t, nr);
    switch (nr) {
#define IMA_XCASE(N) \
  case N:            \
    launch_exact_fast<N>(a, grid, xl, st, work); \
    break;
      IMA_XCASE(1) IMA_XCASE(2) IMA_XCASE(3) IMA_XCASE(4) IMA_XCASE(5) IMA_XCASE(6)
      IMA_XCASE(7) IMA_XCASE(8) IMA_XCASE(9) IMA_XCASE(10) IMA_XCASE(11) IMA_XCASE(12)
      IMA_XCASE(13) IMA_XCASE(14) IMA_XCASE(15) IMA_XCASE(16)
#undef IMA_XCASE
      default:
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(jointp_exact_final, dim3(fin_blocks), dim3(256), 0, st, a, work, out,
                       ess);
  } else {
    double *cmax = work;
    RefPartial *part = reinterpret_cast<RefPartial *>(cmax + (size_t)npoints * a.nchunks);
    const int nr = r1 - r0;
    const size_t xl = exact_lds_bytes(t, nr);
    switch (nr) {
#define IMA_RCASE(N) \
  case N:            \
    launch_ref_fast<N>(a, grid, xl, st, cmax, part); \
    break;
      IMA_RCASE(1) IMA_RCASE(2) IMA_RCASE(3) IMA_RCASE(4) IMA_RCASE(5) IMA_RCASE(6)
      IMA_RCASE(7) IMA_RCASE(8) IMA_RCASE(9) IMA_RCASE(10) IMA_RCASE(11) IMA_RCASE(12)
      IMA_RCASE(13) IMA_RCASE(14) IMA_RCASE(15) IMA_RCASE(16)
#undef IMA_RCASE
      default:
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(jointp_ref_final, dim3(fin_blocks), dim3(256), 0, st, a, cmax, part,
                       out, ess);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------- margincalc --

struct MArgs {
  const float *rows;
  int G, K, npsize, nmig, param, expo;
  double meani;
  const double *x;
  int nx, nchunks, chunk_rows, stage;
};

__global__ __launch_bounds__(kLTile) void margin_partial(MArgs a, double *work) {
  extern __shared__ double shrows_raw[];
  float *shrows = reinterpret_cast<float *>(shrows_raw);
  const int pt = blockIdx.x * kLTile + threadIdx.x;
  const int chunk = blockIdx.y;
  const bool active = pt < a.nx;
  const double x = a.x[active ? pt : 0];
  const JointLayout L = joint_layout(a.npsize, a.nmig);
  const double lx = log(x), l2 = kLn2 - log(x), lmean = log(a.meani);
  double s = 0.0;
  const int gbeg = chunk * a.chunk_rows;
  const int gend = min(a.G, gbeg + a.chunk_rows);
  for (int g0 = gbeg; g0 < gend; g0 += a.stage) {
    const int cnt = min(a.stage, gend - g0);
    __syncthreads();
    const int tot = cnt * a.K;
    const float *src = a.rows + (size_t)g0 * a.K;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) shrows[i] = src[i];
    __syncthreads();
    for (int r = 0; r < cnt; ++r) {
      const float *g = shrows + (size_t)r * a.K;
      if (a.param < a.npsize) {
        const int p = a.param;
        const float cc = g[L.ccp + p];
        const long ccr = cc >= 0 ? (long)(cc + 0.5) : (long)(cc - 0.5);  // INTEGERROUND
        const double temp = -g[L.qip + p] + ccr * l2 - (double)g[L.hccp + p] -
                            (double)(2 * g[L.fcp + p]) / x;
        s += exp(temp);
      } else {
        const int p = a.param - a.npsize;
        const float mc = g[L.mcp + p];
        const long mcr = mc >= 0 ? (long)(mc + 0.5) : (long)(mc - 0.5);
        if (a.expo)
          s += exp(-g[L.mip + p] + lmean - x * a.meani + mcr * lx - g[L.fmp + p] * x);
        else
          s += exp(-g[L.mip + p] + mcr * lx - g[L.fmp + p] * x);
      }
    }
  }
  if (active) work[(size_t)pt * a.nchunks + chunk] = s;
}

__global__ void margin_final(MArgs a, int logi, double yadjust, const double *work,
                             double *out) {
  const int pt = blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= a.nx) return;
  double s = 0.0;
  for (int c = 0; c < a.nchunks; ++c) s += work[(size_t)pt * a.nchunks + c];
  s /= a.G;
  if (logi) s = (s <= 0) ? -1e200 : log(s);
  out[pt] = s - yadjust;
}

size_t margincalc_work_bytes(const LTable &t, int nx) {
  int nc, cr;
  chunking(t, nx, &nc, &cr);
  return (size_t)nx * nc * sizeof(double) + 64;
}

hipError_t launch_margincalc(const LTable &t, int param, const double *x, int nx, int logi,
                             double yadjust, int expo, double prior_mean, double *work,
                             double *out, hipStream_t st) {
  if (nx == 0) return hipSuccess;
  MArgs a{t.rows, t.G, t.K, t.npsize, t.nmig, param, expo,
          expo ? 1.0 / prior_mean : 1.0, x, nx, 0, 0, stage_of(t)};
  chunking(t, nx, &a.nchunks, &a.chunk_rows);
  const dim3 grid((nx + kLTile - 1) / kLTile, a.nchunks);
  const size_t lds = (size_t)a.stage * t.K * sizeof(float);
  hipLaunchKernelGGL(margin_partial, grid, dim3(kLTile), lds, st, a, work);
  hipLaunchKernelGGL(margin_final, dim3((nx + 255) / 256), dim3(256), 0, st, a, logi,
                     yadjust, work, out);
  return hipGetLastError();
}

hipError_t upload_pow10_table(hipStream_t) { return hipSuccess; }

}  // namespace ima
