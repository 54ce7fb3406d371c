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
#include "prior_math.h"

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

// -------------------------------------------------------------- popmig --
//
// The density of 2NM = theta_i m_i / 2 at x, per genealogy the closed form
// of integrating the coalescent and migration likelihoods over the priors
// of theta_i and m_i along x = theta m / 2 (popmig.cpp:9-351):
//   uniform priors: 2 x^mc (fc + fm x)^-(cc+mc) [G(cc+mc, 2(fc+fm x)/qmax)
//     - G(cc+mc, mmax(fm + fc/x))] e^(-hc - qint - mint), with the
//     reference's special cases when the counts or sums vanish;
//   exponential migration prior (mean mmean): its own closed form with one
//     upper incomplete gamma.
// One thread per point, genealogy chunks across workgroups, rows staged in
// LDS as in margincalc.  Uniform terms are summed (terms whose log lies
// outside (-700, 700) are skipped, as the reference does); exponential-prior
// terms are summed relative to their running maximum (log-sum-exp), the
// reference's eexp sum to rounding.

struct PArgs {
  const float *rows;
  int K, g0;               // rows [g0, g0 + n)
  int n, cc, fc, hc, mc, fm, qi, mi;  // column offsets
  double qmax, mmax, mmean;
  int expo, kind, prob_or_like;
  double xmax;             // MARGIN: 1.0 outside [0, xmax]
  double denom;            // rows averaged over (MARGIN: the reference's divisor)
  const double *x;
  int nx, nchunks, chunk_rows, stage;
  const double *lf;        // log(k!) for k < nlf
  int nlf;
};

__device__ inline double pm_lga(const PArgs &a, int k) {  // log Gamma(k) = log((k-1)!)
  return (k >= 1 && k - 1 < a.nlf) ? a.lf[k - 1] : lgamma((double)k);
}
__device__ inline double pm_upper(const PArgs &a, int k, double x) {
  return pri_uppergamma(k, x, pm_lga(a, k));
}
__device__ inline double pm_lower(const PArgs &a, int k, double x) {
  return pri_lowergamma(k, x, pm_lga(a, k));
}

// log of the term of one genealogy, or flags a skipped term (uniform: the
// reference adds exp(t) only for -700 < t < 700).
__device__ inline double popmig_term(const PArgs &a, const float *g, double x, bool *use) {
  const int cc = (int)g[a.cc], mc = (int)g[a.mc];
  const double fc = g[a.fc], hc = g[a.hc], fm = g[a.fm], qintg = g[a.qi], mintg = g[a.mi];
  *use = true;
  if (!a.expo) {
    double t1, t2;
    if (fc == 0 && cc == 0 && fm > 0) {
      t1 = kLog2 - (mc * log(fm)) - hc - qintg - mintg;
      t2 = log(exp(pm_upper(a, mc, 2 * fm * x / a.qmax)) - exp(pm_upper(a, mc, a.mmax * fm)));
    } else if (fm == 0 && mc == 0 && fc > 0) {
      t1 = kLog2 - (cc * log(fc)) - hc - qintg - mintg;
      t2 = log(exp(pm_upper(a, cc, 2 * fc / a.qmax)) - exp(pm_upper(a, cc, fc * a.mmax / x)));
    } else if (fc == 0 && cc == 0 && mc == 0 && fm == 0) {
      t1 = log(2 * log(a.mmax * a.qmax / (2 * x))) - hc - qintg - mintg;
      t2 = 0;
    } else {
      t1 = kLog2 + (mc * log(x)) - ((cc + mc) * log(fc + fm * x)) - hc - qintg - mintg;
      double ua = pm_upper(a, cc + mc, 2 * (fc + fm * x) / a.qmax);
      double ub = pm_upper(a, cc + mc, a.mmax * (fm + fc / x));
      if (ua == ub) {  // both gammas saturate: the lower ones differ instead
        ub = pm_lower(a, cc + mc, 2 * (fc + fm * x) / a.qmax);
        ua = pm_lower(a, cc + mc, a.mmax * (fm + fc / x));
      }
      if (ua > ub) {
        t2 = pri_logdiff(ua, ub);
      } else {
        t1 = t2 = 0.0;  // the reference then adds exp(0)
      }
    }
    const double t = t1 + t2;
    *use = (t < 700) && (t > -700);
    return t;
  }
  const double mmean = a.mmean;
  const double u1 = x + fc * mmean + fm * mmean * x;
  double t2 = kLog2 - hc - log(mmean) - qintg - mintg;
  const double t3 = 2 * u1 / (mmean * a.qmax);
  if (mc == 0 && cc == 0) {
    t2 += pm_upper(a, 0, t3);
  } else if (cc == 0) {
    t2 += pm_upper(a, mc, t3) - mc * log(fm + 1 / mmean);
  } else if (mc == 0) {
    t2 += pm_upper(a, cc, t3) + -cc * log(fc + x * (fm + 1 / mmean));
  } else {
    t2 += pm_upper(a, mc + cc, t3) - cc * log(x) + (cc + mc) * log(mmean * x / u1);
  }
  return t2;
}

// work[pt][chunk] = (sum) for uniform priors, (max, sum relative to max)
// for the exponential prior
__global__ __launch_bounds__(kLTile) void popmig_partial(PArgs a, double *work) {
  extern __shared__ double shrows_raw[];
  float *shrows = reinterpret_cast<float *>(shrows_raw);
  const int pt = blockIdx.x * kLTile + threadIdx.x;
  const int chunk = blockIdx.y;
  const bool active = pt < a.nx;
  const double x = a.x[active ? pt : 0];
  const bool inside = !(a.kind == IMA_POPMIG_MARGIN && (x < 0 || x > a.xmax));
  double s = 0.0, mx = -INFINITY;
  const int gbeg = chunk * a.chunk_rows;
  const int gend = min(a.n, gbeg + a.chunk_rows);
  for (int r0 = gbeg; r0 < gend; r0 += a.stage) {
    const int cnt = min(a.stage, gend - r0);
    __syncthreads();
    const int tot = cnt * a.K;
    const float *src = a.rows + (size_t)(a.g0 + r0) * a.K;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) shrows[i] = src[i];
    __syncthreads();
    if (!active || !inside) continue;
    for (int r = 0; r < cnt; ++r) {
      bool use;
      const double t = popmig_term(a, shrows + (size_t)r * a.K, x, &use);
      if (!a.expo) {
        if (use) s += exp(t);
      } else if (t > mx) {
        s = s * exp(mx - t) + 1.0;
        mx = t;
      } else {
        s += exp(t - mx);
      }
    }
  }
  if (!active) return;
  if (!a.expo) {
    work[(size_t)pt * a.nchunks + chunk] = s;
  } else {
    work[2 * ((size_t)pt * a.nchunks + chunk)] = mx;
    work[2 * ((size_t)pt * a.nchunks + chunk) + 1] = s;
  }
}

__global__ void popmig_final(PArgs a, const double *work, double *out) {
  const int pt = blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= a.nx) return;
  const double x = a.x[pt];
  if (a.kind == IMA_POPMIG_MARGIN && (x < 0 || x > a.xmax)) {
    out[pt] = 1.0;  // OFFSCALEVAL
    return;
  }
  double v;
  if (!a.expo) {
    double s = 0.0;
    for (int c = 0; c < a.nchunks; ++c) s += work[(size_t)pt * a.nchunks + c];
    v = s / a.denom;
  } else {
    double mx = -INFINITY;
    for (int c = 0; c < a.nchunks; ++c) mx = fmax(mx, work[2 * ((size_t)pt * a.nchunks + c)]);
    double s = 0.0;
    for (int c = 0; c < a.nchunks; ++c) {
      const double m = work[2 * ((size_t)pt * a.nchunks + c)];
      if (m > -INFINITY) s += work[2 * ((size_t)pt * a.nchunks + c) + 1] * exp(m - mx);
    }
    v = exp(mx + log(s) - log(a.denom));
  }
  if (a.kind == IMA_POPMIG_MARGIN) {
    out[pt] = -v;
    return;
  }
  if (a.prob_or_like) {
    const double prior = a.expo ? 2 * exp(pri_uppergamma(0, 2 * x / (a.mmean * a.qmax), 0.0)) /
                                      (a.qmax * a.mmean)
                                : 2 * (log(a.qmax) + log(a.mmax) - log(2 * x)) / (a.qmax * a.mmax);
    v /= prior;
  }
  out[pt] = v;
}

size_t popmig_work_bytes(const LTable &t, int nx) {
  int nc, cr;
  chunking(t, nx, &nc, &cr);
  return (size_t)nx * nc * 2 * sizeof(double) + 64;
}

hipError_t launch_popmig(const LTable &t, int theta_i, int mig_i, double qmax, double mprior,
                         int expo, int kind, int prob_or_like, int first, int last,
                         const double *x, int nx, const double *lf, int nlf, double *work,
                         double *out, hipStream_t st) {
  if (nx == 0) return hipSuccess;
  const JointLayout L = joint_layout(t.npsize, t.nmig);
  PArgs a{};
  a.rows = t.rows;
  a.K = t.K;
  a.g0 = kind == IMA_POPMIG_MARGIN ? first : 0;
  a.n = kind == IMA_POPMIG_MARGIN ? last - first : t.G;
  a.cc = L.ccp + theta_i;
  a.fc = L.fcp + theta_i;
  a.hc = L.hccp + theta_i;
  a.mc = L.mcp + mig_i;
  a.fm = L.fmp + mig_i;
  a.qi = L.qip + theta_i;
  a.mi = L.mip + mig_i;
  a.qmax = qmax;
  a.mmax = expo ? 0.0 : mprior;
  a.mmean = expo ? mprior : 0.0;
  a.expo = expo;
  a.kind = kind;
  a.prob_or_like = prob_or_like;
  a.xmax = expo ? 20.0 * mprior : qmax * mprior / 2.0;  // EXPOMIGPLOTSCALE x mean
  a.denom = kind == IMA_POPMIG_MARGIN ? (double)(last - first + (first == 0)) : (double)t.G;
  a.x = x;
  a.nx = nx;
  a.stage = stage_of(t);
  a.lf = lf;
  a.nlf = nlf;
  LTable sub = t;
  sub.G = a.n;
  chunking(sub, nx, &a.nchunks, &a.chunk_rows);
  const dim3 grid((nx + kLTile - 1) / kLTile, a.nchunks);
  const size_t lds = (size_t)a.stage * t.K * sizeof(float);
  hipLaunchKernelGGL(popmig_partial, grid, dim3(kLTile), lds, st, a, work);
  hipLaunchKernelGGL(popmig_final, dim3((nx + 255) / 256), dim3(256), 0, st, a, work, out);
  return hipGetLastError();
}

// ------------------------------------------------------- greater-than --
//
// P(parameter i > parameter j) per genealogy (gtint.cpp:1-331): closed
// forms where the reference has them, else the reference's trapezoid
// refinement (qtrap: relative change below 1e-4, tested from the sixth
// refinement, at most 20) of the integrand over (MINPARAMVAL, prior max).
// One lane per genealogy: each runs the reference's sequence of points and
// sums them in the same order.  A second launch averages the terms of each
// ordered pair.

struct GArgs {
  const float *rows;
  int K, treeinc, ntrees, kind;
  int ccp, fcp, hccp, mcp, fmp, qip, mip;
  const int2 *pairs;  // [npairs] (i, j)
  const double *pmax; // [npairs] prior maximum of the pair
  const double *lf;
  int nlf;
};

struct GtOne {
  int cci, ccj, wi, wj;
  double fci, fcj, hval, denom, fmi, fmj;
};

__device__ inline double gt_lf(const GArgs &a, int k) {  // log(k!)
  return k < a.nlf ? a.lf[k] : lgamma((double)k + 1.0);
}
__device__ inline double gt_upper(const GArgs &a, int k, double x) {
  return pri_uppergamma(k, x, k >= 1 ? gt_lf(a, k - 1) : 0.0);
}
__device__ inline double gt_lower(const GArgs &a, int k, double x) {
  return pri_lowergamma(k, x, gt_lf(a, k - 1));
}

constexpr double kGtSwitch = 1e-15;  // SWITCH_TO_LOWERGAMMA_CRIT
constexpr double kGtMinParam = 1e-7; // MINPARAMVAL

__device__ inline double gt_mgt(const GArgs &a, const GtOne &s, double mi) {  // mgt_wj_gt_0
  if (mi < kGtMinParam) return 0.0;
  const double la = gt_lf(a, s.wj), lb = gt_upper(a, s.wj + 1, s.fmj * mi);
  if (la <= lb) return 0.0;
  const double t1 = (la - lb) < kGtSwitch ? gt_lower(a, s.wj + 1, s.fmj * mi) : pri_logdiff(la, lb);
  double t2 = s.wi * log(mi) - s.fmi * mi - (s.wj + 1) * log(s.fmj) + t1;
  t2 -= s.denom;
  return exp(t2);
}

__device__ inline double gt_pgt(const GArgs &a, const GtOne &s, double qi) {  // pgt_fcj_gt_0
  if (qi < kGtMinParam) return 0.0;
  const double fcj2 = 2 * s.fcj, fci2 = 2 * s.fci;
  if (s.ccj == 0) {
    const double la = log(qi) - fcj2 / qi;
    const double lb = log(fcj2) + gt_upper(a, 0, fcj2 / qi);
    if (la > lb) {
      const double t1 = pri_logdiff(la, lb);
      const double t2 = -fci2 / qi + s.cci * log(2 / qi);
      return exp(t1 + t2 - s.hval - s.denom);
    }
    return 0.0;
  }
  const double t1 = gt_upper(a, s.ccj - 1, fcj2 / qi);
  const double t2 = kLog2 + s.cci * log(2 / qi) + (1 - s.ccj) * log(s.fcj) - fci2 / qi;
  return exp(t2 + t1 - s.hval - s.denom);
}

__device__ inline double gt_f(const GArgs &a, const GtOne &s, bool mig, double x) {
  return mig ? gt_mgt(a, s, x) : gt_pgt(a, s, x);
}

__device__ double gt_qtrap(const GArgs &a, const GtOne &s, bool mig, double lo, double hi) {
  double st = 0.0, olds = -1.0e100;
  for (int j = 1; j <= 20; ++j) {
    if (j == 1) {
      st = 0.5 * (hi - lo) * (gt_f(a, s, mig, lo) + gt_f(a, s, mig, hi));
    } else {
      const int it = 1 << (j - 2);
      const double tnm = it, del = (hi - lo) / tnm;
      double x = lo + 0.5 * del, sum = 0.0;
      for (int k = 1; k <= it; ++k, x += del) sum += gt_f(a, s, mig, x);
      st = 0.5 * (st + (hi - lo) * sum / tnm);
    }
    if (j > 5 && (fabs(st - olds) < 1.0e-4 * fabs(olds) || (st == 0.0 && olds == 0.0)))
      return st;
    olds = st;
  }
  return st;
}

__device__ double gtmig_one(const GArgs &a, const float *g, int mi, int mj, double mmax) {
  GtOne s{};
  s.fmi = g[a.fmp + mi];
  s.fmj = g[a.fmp + mj];
  s.wi = (int)g[a.mcp + mi];
  s.wj = (int)g[a.mcp + mj];
  s.denom = (double)g[a.mip + mi] + (double)g[a.mip + mj];
  double temp;
  if (s.wj == 0) {
    if (s.fmj > 0.0) {
      if (s.wi > 0) {
        const double la = gt_lf(a, s.wi);
        const double lb = gt_upper(a, s.wi + 1, (s.fmi + s.fmj) * mmax);
        const double lc = gt_upper(a, s.wi + 1, s.fmi * mmax);
        if (la <= lb || la <= lc) {
          temp = 0.0;
        } else {
          double t1 = (la - lb) < kGtSwitch ? gt_lower(a, s.wi + 1, (s.fmi + s.fmj) * mmax)
                                             : pri_logdiff(la, lb);
          t1 += -(s.wi + 1) * log(s.fmi + s.fmj);
          double t2 = (la - lc) < kGtSwitch ? gt_lower(a, s.wi + 1, s.fmi * mmax)
                                             : pri_logdiff(la, lc);
          t2 += -(s.wi + 1) * log(s.fmi);
          temp = t2 <= t1 ? 0.0 : exp(pri_logdiff(t2, t1) - log(s.fmj) - s.denom);
        }
      } else if (s.fmi > 0.0) {
        const double t1 = s.fmi * (exp(-(s.fmi + s.fmj) * mmax) - exp(-s.fmi * mmax));
        const double t2 = s.fmj * (1 - exp(-s.fmi * mmax));
        const double t3 = (t1 + t2) / (s.fmi * s.fmj * (s.fmi + s.fmj));
        temp = exp(log(t3) - s.denom);
      } else {
        const double t1 = (s.fmj * mmax - 1.0 + exp(-s.fmj * mmax)) / (s.fmj * s.fmj);
        temp = exp(log(t1) - s.denom);
      }
    } else {
      if (s.wi > 0) {
        const double la = gt_lf(a, s.wi + 1), lb = gt_upper(a, s.wi + 2, s.fmi * mmax);
        if (la <= lb) {
          temp = 0.0;
        } else {
          const double t1 = (la - lb) < kGtSwitch ? gt_lower(a, s.wi + 2, s.fmi * mmax)
                                                   : pri_logdiff(la, lb);
          temp = exp(-(s.wi + 2) * log(s.fmi) + t1 - s.denom);
        }
      } else if (s.fmi > 0.0) {
        const double t1 = (1.0 - exp(-s.fmi * mmax) * (s.fmi * mmax + 1.0)) / (s.fmi * s.fmi);
        temp = exp(log(t1) - s.denom);
      } else {
        temp = exp(log(mmax * mmax / 2.0) - s.denom);
      }
    }
  } else {
    temp = gt_qtrap(a, s, true, kGtMinParam, mmax);
  }
  return fmin(1.0, temp);
}

__device__ double gtpops_one(const GArgs &a, const float *g, int pi, int pj, double qmax) {
  GtOne s{};
  s.cci = (int)g[a.ccp + pi];
  s.ccj = (int)g[a.ccp + pj];
  s.fci = g[a.fcp + pi];
  s.fcj = g[a.fcp + pj];
  s.hval = (double)g[a.hccp + pi] + (double)g[a.hccp + pj];
  s.denom = (double)g[a.qip + pi] + (double)g[a.qip + pj];
  double temp;
  if (s.ccj == 0 && s.fcj == 0) {
    if (s.fci == 0) {
      temp = exp(2.0 * log(qmax) - kLog2 - s.hval - s.denom);
    } else if (s.cci >= 2) {
      const double t1 = 2 * kLog2 + (2 - s.cci) * log(s.fci);
      const double t2 = gt_upper(a, s.cci - 2, 2 * s.fci / qmax);
      temp = exp(t1 + t2 - s.hval - s.denom);
    } else if (s.cci == 1) {
      const double t1 = 4 * s.fci * exp(gt_upper(a, 0, 2 * s.fci / qmax));
      const double t2 = 2 * qmax * exp(-2 * s.fci / qmax) - t1;
      temp = exp(log(t2) - s.hval - s.denom);
    } else {
      const double t1 = exp(gt_upper(a, 0, 2 * s.fci / qmax));
      const double t2 = (qmax / 2) * (qmax - 2 * s.fci) * exp(-2 * s.fci / qmax) +
                        2 * s.fci * s.fci * t1;
      temp = exp(log(t2) - s.hval - s.denom);
    }
  } else {
    temp = gt_qtrap(a, s, false, kGtMinParam, qmax);
  }
  return fmin(1.0, temp);
}

// per[pair][tree]
__global__ __launch_bounds__(64) void gt_terms(GArgs a, double *per) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (t >= a.ntrees) return;
  const float *g = a.rows + (size_t)t * a.treeinc * a.K;
  const int2 ij = a.pairs[p];
  per[(size_t)p * a.ntrees + t] =
      a.kind ? gtmig_one(a, g, ij.x, ij.y, a.pmax[p]) : gtpops_one(a, g, ij.x, ij.y, a.pmax[p]);
}

// one wavefront per pair: lane-strided partial sums in a fixed order
__global__ __launch_bounds__(64) void gt_mean(GArgs a, const double *per, double *out) {
  const int p = blockIdx.x, lane = threadIdx.x;
  double s = 0.0;
  for (int t = lane; t < a.ntrees; t += 64) s += per[(size_t)p * a.ntrees + t];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[p] = s / a.ntrees;
}

hipError_t launch_greater_than(const LTable &t, int kind, const int2 *pairs, const double *pmax,
                               int npairs, const double *lf, int nlf, double *per, double *out,
                               hipStream_t st) {
  if (npairs == 0) return hipSuccess;
  const JointLayout L = joint_layout(t.npsize, t.nmig);
  GArgs a{};
  a.rows = t.rows;
  a.K = t.K;
  // print_greater_than_tests (gtint.cpp:337-349): at most USETREESMAX = 20000
  if (t.G > 20000) {
    a.treeinc = t.G / 20000;
    a.ntrees = 20000;
  } else {
    a.treeinc = 1;
    a.ntrees = t.G;
  }
  a.kind = kind;
  a.ccp = L.ccp;
  a.fcp = L.fcp;
  a.hccp = L.hccp;
  a.mcp = L.mcp;
  a.fmp = L.fmp;
  a.qip = L.qip;
  a.mip = L.mip;
  a.pairs = pairs;
  a.pmax = pmax;
  a.lf = lf;
  a.nlf = nlf;
  hipLaunchKernelGGL(gt_terms, dim3((a.ntrees + 63) / 64, npairs), dim3(64), 0, st, a, per);
  hipLaunchKernelGGL(gt_mean, dim3(npairs), dim3(64), 0, st, a, per, out);
  return hipGetLastError();
}

hipError_t upload_pow10_table(hipStream_t) { return hipSuccess; }

}  // namespace ima
