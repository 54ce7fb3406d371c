// kern_hky.hip -- batched HKY Felsenstein pruning (calc_prob_data.cpp:118-607).
//
// One 64-lane wavefront per work unit = (scoring item, block of 64 * PPL
// site patterns), PPL patterns per lane (PPL = 1 or 2, the site-block size
// the tuner picks: 64 or 128 patterns).  A scoring item is a genealogy of
// one chain at one locus; items of all chains and loci, with uneven gene
// counts and pattern counts, share one launch, and every wavefront does the
// same amount of work.  Items spanning several blocks are summed in block
// order by a second kernel, so results do not depend on scheduling.  With
// PPL = 2 a lane carries two patterns through the same program: the
// program, coefficient and tip-row reads, the addresses and the branches
// are shared by both, only the arithmetic is doubled.
//
// The host turns each genealogy into a postorder stack program (ProgOp):
// the daughter with the larger need is evaluated first, so the stack never
// holds more than floor(log2 n) partials (Strahler number), at most 8 for
// n <= 256.  The two top partials live in registers, the levels below them
// in the wavefront's LDS, one 16-byte column per lane and pattern.  Nothing
// per node goes to HBM: every proposal is re-scored over the whole
// genealogy, ~(36n-58)S flops on chip, instead of reading and writing ~36 B
// per node-pattern of stored partials.
//
// Branch transition (HKY85 closed form, calc_prob_data.cpp:197-208):
//   P_jk(t) = pi_k (1-e) + [cls(j)=cls(k)] pi_k (e - eA_c)/PI_c + [j=k] eA_c
// with e = exp(-mu t), eA_c = exp(-mu t A_c), A_c = 1 + PI_c (kappa-1), c the
// purine {A,G} or pyrimidine {C,T} class of the destination base.  For a
// daughter with partial L:  (P L)_j = a W + b_c W_c + eA_c L_j with a = 1-e,
// b_c = (e-eA_c)/PI_c, W = sum pi_k L_k, W_c = sum_{k in c} pi_k L_k.  For a
// tip in state s the vector is column s of P.  Per branch the five
// coefficients, and per tip branch the four columns of P, are computed once
// per unit in FP64 and kept in LDS; a tip operand is one 16-byte LDS read.
// Partials are rescaled by an exact power of two when they fall below a
// floor; one integer exponent per lane and pattern carries the sum of the
// scale factors (the reference keeps log(max) per node, :211-217, 465).
#include "ima_internal.h"
#include "tw_dev.h"

namespace ima {
namespace {

template <typename real>
struct Q4 {
  real x, y, z, w;
};

template <typename real>
struct __align__(16) HkyCoef {
  real a, bR, bY, eAR, eAY, pad0, pad1, pad2;
};

template <typename real>
__device__ __forceinline__ Q4<real> int_vec(const HkyCoef<real> &c, const Q4<real> &L,
                                            const real pi[4]) {
  const real WR = pi[0] * L.x + pi[2] * L.z;
  const real WY = pi[1] * L.y + pi[3] * L.w;
  const real aW = c.a * (WR + WY);
  const real tR = aW + c.bR * WR;
  const real tY = aW + c.bY * WY;
  Q4<real> v;
  v.x = tR + c.eAR * L.x;
  v.z = tR + c.eAR * L.z;
  v.y = tY + c.eAY * L.y;
  v.w = tY + c.eAY * L.w;
  return v;
}

__device__ __forceinline__ float ldexp_r(float x, int e) { return ldexpf(x, e); }
__device__ __forceinline__ double ldexp_r(double x, int e) { return ldexp(x, e); }
__device__ __forceinline__ int frexp_e(float x) {
  int e;
  frexpf(x, &e);
  return e;
}
__device__ __forceinline__ int frexp_e(double x) {
  int e;
  frexp(x, &e);
  return e;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Three-stage operand pipeline.  Mid: a decoded program entry whose tip
// states are being read from LDS.  Nxt: an entry whose branch coefficients
// and tip columns of P are being read.  The compute step of an entry reads
// nothing from LDS.  Every read is made whatever the entry's type, so the
// staging has no branches: an operand that is not a tip reads tip row
// min(edge, n - 1) (any row: its value is not used), and an edge index is
// always below 2n - 1.
template <int PPL>
struct Mid {
  int type, a, b;        // the same in every lane
  int ra, rb;            // tip rows read for a and b
  int ta[PPL], tb[PPL];  // this lane's states at those rows
};

template <typename real, int PPL>
struct Nxt {
  int type;
  HkyCoef<real> ca, cb;
  Q4<real> pa[PPL], pb[PPL];  // columns of P for tip operands
};

// tips: [n][64 * PPL] states; pattern j * 64 + lane of the block at byte
// row * 64 * PPL + j * 64 + lane
template <int PPL>
__device__ __forceinline__ void to_mid(const ProgOp &raw, const uint8_t *tips, int lane, int nlast,
                                       Mid<PPL> &m) {
  // a and b stay in vector registers (every lane holds the same value) and
  // feed the LDS addresses directly; only the type is made scalar, for the
  // branch
  m.a = raw.a;
  m.b = raw.b;
  m.type = uni(raw.type);
  m.ra = min(m.a, nlast);
  m.rb = min(m.b, nlast);
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    m.ta[j] = tips[m.ra * (kWave * PPL) + j * kWave + lane];
    m.tb[j] = tips[m.rb * (kWave * PPL) + j * kWave + lane];
  }
}

template <typename real, int PPL>
__device__ __forceinline__ void to_nxt(const Mid<PPL> &m, const HkyCoef<real> *coef,
                                       const Q4<real> *colP, Nxt<real, PPL> &x) {
  x.type = m.type;
  x.ca = coef[m.a];
  x.cb = coef[m.b];
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    x.pa[j] = colP[m.ra * 4 + m.ta[j]];
    x.pb[j] = colP[m.rb * 4 + m.tb[j]];
  }
}

// Below this a node's partials are rescaled (a power of two: 2^-40 in
// FP32, 2^-400 in FP64), far above the smallest normal number.
__device__ __forceinline__ float rescale_floor(float) { return 0x1p-40f; }
__device__ __forceinline__ double rescale_floor(double) { return 0x1p-400; }

template <typename real>
__device__ __forceinline__ Q4<real> mul4(const Q4<real> &a, const Q4<real> &b) {
  Q4<real> L;
  L.x = a.x * b.x;
  L.y = a.y * b.y;
  L.z = a.z * b.z;
  L.w = a.w * b.w;
  return L;
}
template <typename real>
__device__ __forceinline__ real max4(const Q4<real> &L) {
  return fmax(fmax(L.x, L.y), fmax(L.z, L.w));
}

// A node's partials from the two operand vectors of each pattern.  When some
// lane's largest partial of some pattern falls below the floor, every
// pattern of every lane is rescaled by an exact power of two (its exponent
// carried in E), so the mantissas are those of rescaling at every node.
template <typename real, int PPL>
__device__ __forceinline__ void combine(const Q4<real> (&va)[PPL], const Q4<real> (&vb)[PPL],
                                        Q4<real> (&L)[PPL], int (&E)[PPL]) {
  real mx[PPL];
  bool small = false;
#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    L[j] = mul4(va[j], vb[j]);
    mx[j] = max4(L[j]);
    small = small || mx[j] < rescale_floor(mx[j]);
  }
  if (__ballot(small)) {
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
      const int ex = frexp_e(mx[j]);
      E[j] += ex;
      L[j].x = ldexp_r(L[j].x, -ex);
      L[j].y = ldexp_r(L[j].y, -ex);
      L[j].z = ldexp_r(L[j].z, -ex);
      L[j].w = ldexp_r(L[j].w, -ex);
    }
  }
}

// One program operation.  The two top partials of the stack live in
// registers (tos = stk[d-1], nos = stk[d-2]); the levels below them live in
// this lane's column of the wavefront's LDS stack (level l of pattern j at
// low[(l * PPL + j) * 64]).  TT pushes: nos spills to level d-2, tos becomes
// nos.  ST rewrites tos.  SS pops: the combination of nos (daughter a) and
// tos (daughter b) becomes tos, and level d-3 is read back into nos.  The
// type and d are wave-uniform, so the three cases are scalar branches.
// While the stack holds fewer than three partials, level 0 is free: a push
// spills its meaningless nos there and a pop reads it back unused, so the
// LDS accesses need no branches.
template <typename real, int PPL>
__device__ __forceinline__ void apply_op(const Nxt<real, PPL> &c, Q4<real> (&tos)[PPL],
                                         Q4<real> (&nos)[PPL], int &d, Q4<real> *low,
                                         const real pi[4], int (&E)[PPL]) {
  // three independent wave-uniform tests (not an else-chain), so each case
  // is one scalar compare and branch
  const int type = c.type;
  if (type == OP_TT) {
    const int lv = max(d - 2, 0) * PPL;
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
      low[(lv + j) * kWave] = nos[j];
      nos[j] = tos[j];
    }
    combine<real, PPL>(c.pa, c.pb, tos, E);
    ++d;
  }
  if (type == OP_ST) {
    Q4<real> va[PPL];
#pragma unroll
    for (int j = 0; j < PPL; ++j) va[j] = int_vec<real>(c.ca, tos[j], pi);
    combine<real, PPL>(va, c.pb, tos, E);
  }
  if (type == OP_SS) {
    Q4<real> va[PPL], vb[PPL];
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
      va[j] = int_vec<real>(c.ca, nos[j], pi);
      vb[j] = int_vec<real>(c.cb, tos[j], pi);
    }
    combine<real, PPL>(va, vb, tos, E);
    // read into nos itself, last: the wait for it falls at nos's next use,
    // after the next entry's staging, not here
    const int lv = max(d - 3, 0) * PPL;
#pragma unroll
    for (int j = 0; j < PPL; ++j) nos[j] = low[(lv + j) * kWave];
    --d;
  }
}

}  // namespace

// Stack levels kept in LDS for programs of up to `max_slots` partials (the
// launch's depth class 4, 6 or 8, less the two held in registers).
static int hky_low_levels(int max_slots) { return max_slots <= 4 ? 2 : max_slots <= 6 ? 4 : 6; }

size_t hky_lds_bytes(int precision, int max_n, int max_slots, int ppl) {
  const size_t ne = 2 * (size_t)max_n - 1;
  const size_t r = precision == IMA_PREC_FP32 ? 4 : 8;
  // lower stack levels, coefficients per edge, P columns per tip edge,
  // program, tip states
  return (size_t)hky_low_levels(max_slots) * ppl * kWave * (4 * r) + ne * (8 * r) +
         (size_t)max_n * (16 * r) + ((size_t)max_n - 1) * sizeof(ProgOp) +
         (size_t)max_n * kWave * ppl;
}

void hky_plan_add(std::vector<int4> &units, std::vector<int4> &red, int ii, int S,
                  int *nblocks_total, int ppl) {
  const int bs = kWave * ppl;
  const int nb = (S + bs - 1) / bs;
  const int pbase = *nblocks_total;
  for (int b = 0; b < nb; ++b) units.push_back(make_int4(ii, b, pbase, nb));
  if (nb > 1) red.push_back(make_int4(ii, pbase, nb, 0));
  *nblocks_total += nb;
}

// Work unit `unit` (item, pattern block, first partial, blocks) of one
// 64-lane workgroup, with smem its dynamic LDS of hky_lds_bytes.
template <typename real, int D, int PPL>
__device__ __forceinline__ void hky_unit(const DeviceView &v, const ima_item *__restrict__ items,
                                         const int4 unit, double *__restrict__ partial,
                                         double *__restrict__ out,
                                         int32_t *__restrict__ status, unsigned char *smem) {
  // the unit is one wavefront's own work: its LDS slice is its own and it
  // synchronises with wavefront-scope fences only
  const int lane = threadIdx.x & (kWave - 1);
  const int ii = unit.x, blk = unit.y;
  const ima_item item = items[ii];
  const SlotMeta sm = v.slots[item.slot];
  const LocusMeta lm = v.loci[sm.locus];
  const int n = sm.n, ne = 2 * n - 1, S = lm.S, nops = n - 1;

  constexpr int kLow = D - 2;  // the launch's lower stack levels (D is 4, 6 or 8)
  Q4<real> *low = reinterpret_cast<Q4<real> *>(smem) + lane;     // [kLow][PPL][64]
  HkyCoef<real> *coef = reinterpret_cast<HkyCoef<real> *>(
      reinterpret_cast<Q4<real> *>(smem) + kLow * PPL * kWave);  // [ne]
  Q4<real> *colP = reinterpret_cast<Q4<real> *>(coef + ne);       // [n][4] tips
  ProgOp *ops = reinterpret_cast<ProgOp *>(colP + (size_t)n * 4); // [nops]
  uint8_t *tips = reinterpret_cast<uint8_t *>(ops + nops);        // [n][64 PPL]

  const double *pid = lm.pi;
  const double kappa = item.kappa;
  // getstandfactor (calc_prob_data.cpp:474-493) and the per-site rate
  // u / (totsites * standfactor) (:593).
  double sf = 0.0;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      if (i != j) sf += (i + j == 2 || i + j == 4) ? pid[i] * pid[j] * kappa : pid[i] * pid[j];
  const double mutrate = item.u / (lm.totsites * sf);
  const double PR = pid[0] + pid[2], PY = pid[1] + pid[3];
  const double AR = 1.0 + PR * (kappa - 1.0), AY = 1.0 + PY * (kappa - 1.0);

  const int32_t *up0 = v.up0 + sm.node_off;
  const double *tm = v.time + sm.node_off;
  for (int e = lane; e < ne; e += kWave) {
    if (e == sm.root) continue;
    const int u0 = up0[e];
    const double bl = tm[e] - (u0 >= 0 ? tm[u0] : 0.0);
    const double mt = mutrate * bl;
    const double e1 = exp(-mt), eR = exp(-mt * AR), eY = exp(-mt * AY);
    const double bR = (e1 - eR) / PR, bY = (e1 - eY) / PY, a = 1.0 - e1;
    HkyCoef<real> c;
    c.a = real(a);
    c.bR = real(bR);
    c.bY = real(bY);
    c.eAR = real(eR);
    c.eAY = real(eY);
    c.pad0 = c.pad1 = c.pad2 = real(0);
    coef[e] = c;
    if (e >= n) continue;  // P columns are read for tip operands only
    // column s of P: P_js = pi_s a + [cls j = cls s] pi_s b_c + [j = s] eA_c
    for (int s = 0; s < 4; ++s) {
      const bool Rs = (s & 1) == 0;
      const double ps = pid[s], base = ps * a, same = base + ps * (Rs ? bR : bY);
      const double es = Rs ? eR : eY;
      Q4<real> col;
      col.x = real((Rs ? same : base) + (s == 0 ? es : 0.0));
      col.y = real((Rs ? base : same) + (s == 1 ? es : 0.0));
      col.z = real((Rs ? same : base) + (s == 2 ? es : 0.0));
      col.w = real((Rs ? base : same) + (s == 3 ? es : 0.0));
      colP[e * 4 + s] = col;
    }
  }
  const ProgOp *prog = v.prog + sm.prog_off;
  for (int k = lane; k < nops; k += kWave) ops[k] = prog[k];
  const real pi[4] = {real(pid[0]), real(pid[1]), real(pid[2]), real(pid[3])};
  const uint32_t *seq32 = reinterpret_cast<const uint32_t *>(v.seq + lm.seq_off);
  const int row32 = lm.Spad / 4;  // Spad is a multiple of 128: no block reads past its row
  const int32_t *mult = v.mult + lm.mult_off;
  uint32_t *tips32 = reinterpret_cast<uint32_t *>(tips);
  constexpr int kRowWords = kWave / 4 * PPL;  // dwords of one row of a block
  const int words = n * kRowWords;

  double acc = 0.0;
  {
    wave_lds_fence();
    // tip states of this block: n rows of kRowWords aligned dwords, all loads
    // issued before the LDS writes
    for (int w0 = 0; w0 < words; w0 += 16 * kWave) {
      uint32_t r[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int w = min(w0 + j * kWave + lane, words - 1);
        r[j] = seq32[(size_t)(w / kRowWords) * row32 + blk * kRowWords + (w % kRowWords)];
      }
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int w = w0 + j * kWave + lane;
        if (w < words) tips32[w] = r[j];
      }
    }
    wave_lds_fence();

    Q4<real> tos[PPL], nos[PPL];
    int E[PPL];
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
      tos[j].x = tos[j].y = tos[j].z = tos[j].w = real(0);
      nos[j] = tos[j];
      E[j] = 0;
    }
    int d = 0;
    // op k computes while the coefficients and P columns of op k+1, the tip
    // states of op k+2 and the program entry of op k+3 are read from LDS;
    // unrolled by two so no staged operand is ever copied
    const int last = nops - 1;
    Mid<PPL> M0, M1;
    Nxt<real, PPL> N0, N1;
    to_mid<PPL>(ops[0], tips, lane, n - 1, M0);
    to_mid<PPL>(ops[min(1, last)], tips, lane, n - 1, M1);
    to_nxt<real, PPL>(M0, coef, colP, N0);
    ProgOp raw = ops[min(2, last)];
    for (int k = 0; k < nops; k += 2) {
      // even: compute N0 (op k), N1 <- M1 (op k+1), M0 <- raw (op k+2)
      to_nxt<real, PPL>(M1, coef, colP, N1);
      to_mid<PPL>(raw, tips, lane, n - 1, M0);
      raw = ops[min(k + 3, last)];
      apply_op<real, PPL>(N0, tos, nos, d, low, pi, E);
      if (k + 1 >= nops) break;
      // odd: compute N1 (op k+1), N0 <- M0 (op k+2), M1 <- raw (op k+3)
      to_nxt<real, PPL>(M0, coef, colP, N0);
      to_mid<PPL>(raw, tips, lane, n - 1, M1);
      raw = ops[min(k + 4, last)];
      apply_op<real, PPL>(N1, tos, nos, d, low, pi, E);
    }
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
      const int p = blk * kWave * PPL + j * kWave + lane;
      const real fr =
          pi[0] * tos[j].x + pi[1] * tos[j].y + pi[2] * tos[j].z + pi[3] * tos[j].w;
      if (p < S) acc += (double)mult[p] * (log((double)fr) + (double)E[j] * kLn2);
    }
  }
  acc = wave_sum_d(acc);
  if (lane == 0) {
    if (unit.w == 1) {
      out[ii] = acc;
      if (status) status[ii] = IMA_STATUS_OK;
    } else {
      partial[unit.z + blk] = acc;
    }
  }
}

template <typename real, int D, int PPL>
__global__ __launch_bounds__(kWave) void hky_kernel(
    DeviceView v, const ima_item *__restrict__ items, const int4 *__restrict__ units,
    double *__restrict__ partial, double *__restrict__ out, int32_t *__restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem[];
  hky_unit<real, D, PPL>(v, items, units[blockIdx.x], partial, out, status, smem);
}

// The mutation-rate round's scoring: the pattern-block units of one
// level's proposals, listed by the previous launch (*cnt of them).  Block 0
// clears the other list's counter for the next level.
template <typename real, int D, int PPL>
__global__ __launch_bounds__(kWave) void hky_list_kernel(
    DeviceView v, const ima_item *__restrict__ items, const int4 *__restrict__ list,
    const int32_t *__restrict__ cnt, int32_t *__restrict__ cnt_next,
    double *__restrict__ partial, double *__restrict__ out, int32_t *__restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (blockIdx.x == 0 && threadIdx.x == 0) *cnt_next = 0;
  const int n = *cnt;
  for (int k = blockIdx.x; k < n; k += gridDim.x) {
    wave_lds_fence();
    hky_unit<real, D, PPL>(v, items, list[k], partial, out, status, smem);
  }
}

// The sweep's scoring in one launch: workgroups [0, tw.n) weigh the
// proposals' genealogies (treeweight), the rest run the likelihood units.
// One launch instead of two side by side on two streams saves the
// cross-stream fork and join between the accept steps.
template <typename real, int D, int PPL>
__global__ __launch_bounds__(kWave) void hky_tw_kernel(
    DeviceView v, const ima_item *__restrict__ items, const int4 *__restrict__ units,
    double *__restrict__ partial, double *__restrict__ out, int32_t *__restrict__ status,
    TwLaunch tw) {
  extern __shared__ __align__(16) unsigned char smem[];
  if ((int)blockIdx.x < tw.n) {
    twdev::treeweight_one(v, tw.slots, blockIdx.x, tw.tvals, tw.tv_row, tw.rec_out, tw.status,
                          reinterpret_cast<double *>(smem));
    return;
  }
  hky_unit<real, D, PPL>(v, items, units[blockIdx.x - tw.n], partial, out, status, smem);
}

__global__ void hky_reduce(const int4 *__restrict__ red, int nred,
                           const double *__restrict__ partial, double *__restrict__ out,
                           int32_t *__restrict__ status) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nred) return;
  const int4 r = red[k];
  double s = 0.0;
  for (int b = 0; b < r.z; ++b) s += partial[r.y + b];
  out[r.x] = s;
  if (status) status[r.x] = IMA_STATUS_OK;
}

// Calls f.template operator()<D, PPL>() for the launch's depth class
// (max_slots up to 4, 6 or 8) and patterns per lane (1 or 2).
template <typename F>
static void by_shape(int max_slots, int ppl, const F &f) {
  if (ppl == 2) {
    if (max_slots <= 4) f.template operator()<4, 2>();
    else if (max_slots <= 6) f.template operator()<6, 2>();
    else f.template operator()<8, 2>();
  } else {
    if (max_slots <= 4) f.template operator()<4, 1>();
    else if (max_slots <= 6) f.template operator()<6, 1>();
    else f.template operator()<8, 1>();
  }
}

// The three launchers' kernels for one (depth class, patterns per lane).
struct ListLaunch {
  dim3 g;
  size_t lds;
  hipStream_t st;
  int precision;
  const DeviceView *v;
  const ima_item *items;
  const int4 *list;
  const int32_t *cnt;
  int32_t *cnt_next;
  double *partial, *out;
  int32_t *status;
  template <int D, int PPL>
  void operator()() const {
    if (precision == IMA_PREC_FP32)
      hipLaunchKernelGGL((hky_list_kernel<float, D, PPL>), g, dim3(kWave), lds, st, *v, items,
                         list, cnt, cnt_next, partial, out, status);
    else
      hipLaunchKernelGGL((hky_list_kernel<double, D, PPL>), g, dim3(kWave), lds, st, *v, items,
                         list, cnt, cnt_next, partial, out, status);
  }
};

struct TwLaunchFn {
  dim3 g;
  size_t lds;
  hipStream_t st;
  int precision;
  const DeviceView *v;
  const ima_item *items;
  const HkyPlan *plan;
  double *out;
  int32_t *status;
  const TwLaunch *tw;
  template <int D, int PPL>
  void operator()() const {
    if (precision == IMA_PREC_FP32)
      hipLaunchKernelGGL((hky_tw_kernel<float, D, PPL>), g, dim3(kWave), lds, st, *v, items,
                         plan->units, plan->partial, out, status, *tw);
    else
      hipLaunchKernelGGL((hky_tw_kernel<double, D, PPL>), g, dim3(kWave), lds, st, *v, items,
                         plan->units, plan->partial, out, status, *tw);
  }
};

struct PlainLaunch {
  dim3 g;
  size_t lds;
  hipStream_t st;
  int precision;
  const DeviceView *v;
  const ima_item *items;
  const HkyPlan *plan;
  double *out;
  int32_t *status;
  template <int D, int PPL>
  void operator()() const {
    if (precision == IMA_PREC_FP32)
      hipLaunchKernelGGL((hky_kernel<float, D, PPL>), g, dim3(kWave), lds, st, *v, items,
                         plan->units, plan->partial, out, status);
    else
      hipLaunchKernelGGL((hky_kernel<double, D, PPL>), g, dim3(kWave), lds, st, *v, items,
                         plan->units, plan->partial, out, status);
  }
};

hipError_t launch_hky_list(const DeviceView &v, const ima_item *items, const int4 *list,
                           const int32_t *cnt, int32_t *cnt_next, int grid, int precision,
                           int max_n, int max_slots, int ppl, double *partial, double *out,
                           int32_t *status, hipStream_t st) {
  if (grid <= 0 || max_slots > 8 || (ppl != 1 && ppl != 2)) return hipErrorInvalidValue;
  const size_t lds = hky_lds_bytes(precision, max_n, max_slots, ppl);
  by_shape(max_slots, ppl,
           ListLaunch{dim3(grid), lds, st, precision, &v, items, list, cnt, cnt_next, partial,
                      out, status});
  return hipGetLastError();
}

size_t hky_tw_lds_bytes(int precision, int max_n, int max_slots, int ppl, int P, int emax) {
  return std::max(hky_lds_bytes(precision, max_n, max_slots, ppl),
                  treeweight_lds_bytes(P, emax));
}

hipError_t launch_hky_tw(const DeviceView &v, const ima_item *items, const HkyPlan &plan,
                         int precision, int max_n, int max_slots, double *out, int32_t *status,
                         const TwLaunch &tw, int P, hipStream_t st) {
  if (plan.nunits == 0 || tw.n == 0) return hipErrorInvalidValue;
  if (max_slots > 8 || plan.nred > 0 || (plan.ppl != 1 && plan.ppl != 2))
    return hipErrorInvalidValue;
  const size_t lds = hky_tw_lds_bytes(precision, max_n, max_slots, plan.ppl, P, tw.emax);
  if (lds > kTwLdsBudget) return hipErrorInvalidValue;
  by_shape(max_slots, plan.ppl,
           TwLaunchFn{dim3(plan.nunits + tw.n), lds, st, precision, &v, items, &plan, out,
                      status, &tw});
  return hipGetLastError();
}

hipError_t launch_hky(const DeviceView &v, const ima_item *items, const HkyPlan &plan,
                      int precision, int max_n, int max_slots, double *out, int32_t *status,
                      hipStream_t st) {
  if (plan.nunits == 0) return hipSuccess;
  if (max_slots > 8 || (plan.ppl != 1 && plan.ppl != 2)) return hipErrorInvalidValue;
  const size_t lds = hky_lds_bytes(precision, max_n, max_slots, plan.ppl);
  by_shape(max_slots, plan.ppl,
           PlainLaunch{dim3(plan.nunits), lds, st, precision, &v, items, &plan, out, status});
  if (plan.nred > 0)
    hipLaunchKernelGGL(hky_reduce, dim3((plan.nred + 255) / 256), dim3(256), 0, st, plan.red,
                       plan.nred, plan.partial, out, status);
  return hipGetLastError();
}

}  // namespace ima
