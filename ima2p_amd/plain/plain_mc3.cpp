// plain_mc3.cpp -- the plain per-locus FP64 path: the MC3 genealogy sweep
// driven from the host one proposal at a time through libima_gpu.so, the
// way IMa2p's own updategenealogy (update_gtree.cpp:723-966) calls its
// likelihood and prior functions once per (chain, locus).
//
// Per proposal: the host moves one coalescence time of its own copy of the
// genealogy, uploads that one slot (ima_slots_write), scores it alone in
// FP64 (ima_score, one item), weighs it alone (ima_treeweight, one slot),
// updates the chain's all-loci sums on the host (sum_subtract_treeinfo,
// ginfo.cpp:249-285), integrates them (ima_integrate_tree_prob, one group)
// and decides with U < min(1, exp(beta (dprobg + dpdg)))
// (update_gtree.cpp:919-927).  After every sweep one ima_swap.  This is the
// baseline the batched device sweep is measured against: same move, same
// scoring, same decisions, no batching.
//
// Built with the host compiler into libima_plain.so, linked to
// libima_gpu.so.  Not part of the engine's ABI.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/ima_gpu.h"

namespace {

struct Gen {
  int n, slot, locus_kind;
  std::vector<int32_t> up0, up1, down, pop;
  std::vector<double> time;
  int32_t root;
  std::vector<int32_t> me, mto;  // migrations sorted by edge, then time
  std::vector<double> mt;
  std::vector<int32_t> mfirst, mcount;  // per edge
};

// sum_subtract_treeinfo's rules for one entry of the all-loci record:
// counts add the difference, fc / fm subtract then add and stay >= 0.
double wsum_update(int r, int P, double all, double old, double nw) {
  const int P2 = P * P, P3 = P2 * P;
  if (r < P2) return all + (nw - old);
  if (r < 2 * P2) return std::fmax(0.0, (all - old) + nw);
  if (r < 3 * P2) return (all - old) + nw;
  if (r < 3 * P2 + P3) return all + (nw - old);
  if (r < 3 * P2 + 2 * P3) return std::fmax(0.0, (all - old) + nw);
  return (all - old) + nw;
}

}  // namespace

extern "C" {

// Runs nsweeps plain sweeps over nchains x nloci genealogies already
// written to slots first_slot + c*nloci + l (with populations and
// migrations when model >= 0), whose host copies are passed packed slot
// after slot as for ima_slots_write / ima_slots_write_pops.  model < 0:
// no genealogy prior.  Outputs: wall seconds of the sweeps, final pdg,
// probg, beta, accepted proposals, and (time_out, may be NULL) the final
// node times packed like time.
int ima_plain_mc3_run(ima_ctx *ctx, int nchains, int nloci, int first_slot, int model, int P,
                      const int32_t *up0, const int32_t *up1, const int32_t *down,
                      const double *time, const int32_t *root, const int32_t *pop,
                      const int32_t *nmig, const int32_t *mig_edge, const double *mig_time,
                      const int32_t *mig_to, const double *u, const double *kappa,
                      const double *beta_in, const double *tvals, uint64_t seed, int nsweeps,
                      int swaptries, double *seconds_out, double *pdg_out, double *probg_out,
                      double *beta_out, uint64_t *accepts_out, double *time_out) {
  if (!ctx || nchains < 1 || nloci < 1 || !up0 || !up1 || !down || !time || !root || !u ||
      !beta_in || nsweeps < 0)
    return IMA_ERR_ARG;
  const bool prior = model >= 0;
  if (prior && (!pop || !nmig || (P > 1 && !tvals))) return IMA_ERR_ARG;
  const int N = nchains * nloci, TW = P > 1 ? P - 1 : 1;
  const int R = 3 * P * P + 2 * P * P * P + 1;
  std::vector<Gen> g(N);
  size_t off = 0, moff = 0;
  for (int i = 0; i < N; ++i) {
    Gen &x = g[i];
    int locus = 0, n = 0;
    int rc = ima_slot_info(ctx, first_slot + i, &locus, &n);
    if (rc != IMA_OK) return rc;
    const int ne = 2 * n - 1;
    x.n = n;
    x.slot = first_slot + i;
    x.up0.assign(up0 + off, up0 + off + ne);
    x.up1.assign(up1 + off, up1 + off + ne);
    x.down.assign(down + off, down + off + ne);
    x.time.assign(time + off, time + off + ne);
    x.root = root[i];
    x.mfirst.assign(ne, 0);
    x.mcount.assign(ne, 0);
    if (prior) {
      x.pop.assign(pop + off, pop + off + ne);
      const int k = nmig[i];
      x.me.assign(mig_edge + moff, mig_edge + moff + k);
      x.mt.assign(mig_time + moff, mig_time + moff + k);
      x.mto.assign(mig_to + moff, mig_to + moff + k);
      for (int j = k - 1; j >= 0; --j) {
        x.mfirst[x.me[j]] = j;
        x.mcount[x.me[j]] += 1;
      }
      moff += (size_t)k;
    }
    off += (size_t)ne;
  }
  std::vector<double> tv((size_t)nchains * TW, 0.0);
  if (prior && P > 1) std::copy(tvals, tvals + (size_t)nchains * TW, tv.begin());
  std::vector<double> beta(beta_in, beta_in + nchains);
  // initial state: every locus scored, weighed and summed
  std::vector<double> pdg(N), probg(nchains, 0.0), w((size_t)N * R, 0.0), all((size_t)nchains * R, 0.0);
  std::vector<ima_item> items(N);
  for (int i = 0; i < N; ++i)
    items[i] = ima_item{first_slot + i, 0, u[i], kappa ? kappa[i] : 1.0, 1.0};
  int rc = ima_score(ctx, items.data(), N, IMA_PREC_FP64, pdg.data(), nullptr);
  if (rc != IMA_OK) return rc;
  std::vector<int32_t> cc(P * P), mc(P * P * P);
  std::vector<double> fc(P * P), hcc(P * P), fm(P * P * P);
  auto weigh = [&](int i, double *rec) -> int {
    double len = 0.0;
    const int32_t sl = first_slot + i;
    const int r = ima_treeweight(ctx, &sl, 1, tv.data() + (size_t)(i / nloci) * TW, cc.data(),
                                 fc.data(), hcc.data(), mc.data(), fm.data(), &len);
    if (r != IMA_OK) return r;
    const int P2 = P * P, P3 = P2 * P;
    for (int k = 0; k < P2; ++k) {
      rec[k] = cc[k];
      rec[P2 + k] = fc[k];
      rec[2 * P2 + k] = hcc[k];
    }
    for (int k = 0; k < P3; ++k) {
      rec[3 * P2 + k] = mc[k];
      rec[3 * P2 + P3 + k] = fm[k];
    }
    rec[R - 1] = len;
    return IMA_OK;
  };
  if (prior) {
    for (int i = 0; i < N; ++i)
      if ((rc = weigh(i, w.data() + (size_t)i * R)) != IMA_OK) return rc;
    for (int c = 0; c < nchains; ++c) {
      double *a = all.data() + (size_t)c * R;
      for (int l = 0; l < nloci; ++l)
        for (int r = 0; r < R; ++r) a[r] += w[((size_t)c * nloci + l) * R + r];
      if ((rc = ima_integrate_tree_prob(ctx, model, a, 1, &probg[c], nullptr, nullptr)) != IMA_OK)
        return rc;
    }
  }
  std::mt19937 mt((uint32_t)seed);
  std::uniform_real_distribution<double> unif(0.0, 1.0);
  std::vector<double> wn(R), an(R);
  uint64_t accepts = 0;
  const auto t0 = std::chrono::steady_clock::now();
  for (int sweep = 0; sweep < nsweeps; ++sweep) {
    for (int c = 0; c < nchains; ++c) {
      for (int l = 0; l < nloci; ++l) {
        const int i = c * nloci + l;
        Gen &x = g[i];
        const int n = x.n;
        // the time-slide move of the device sweep: one coalescence time
        // within its daughters, parent, period and neighbouring migrations
        int e = n + (int)(unif(mt) * (n - 1));
        if (e > 2 * n - 2) e = 2 * n - 2;
        const int a = x.up0[e], b = x.up1[e];
        const double T = x.time[a];
        double lo = std::fmax(a >= n ? x.time[x.up0[a]] : 0.0, b >= n ? x.time[x.up0[b]] : 0.0);
        double hi = x.down[e] == -1 ? HUGE_VAL : x.time[e];
        if (prior) {
          const double *t = tv.data() + (size_t)c * TW;
          int k = 0;
          while (k < P - 1 && t[k] <= T) k++;
          if (k > 0) lo = std::fmax(lo, t[k - 1]);
          if (k < P - 1) hi = std::fmin(hi, t[k]);
          if (x.mcount[a]) lo = std::fmax(lo, x.mt[x.mfirst[a] + x.mcount[a] - 1]);
          if (x.mcount[b]) lo = std::fmax(lo, x.mt[x.mfirst[b] + x.mcount[b] - 1]);
          if (x.mcount[e]) hi = std::fmin(hi, x.mt[x.mfirst[e]]);
        }
        const bool open = std::isinf(hi);
        const double sigma = open ? 0.5 * std::fmax(lo, 0.05) : 0.5 * (hi - lo);
        const double u1 = 1.0 - unif(mt), u2 = unif(mt);
        double Tn = T + sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        if (open) {
          Tn = lo + std::fabs(Tn - lo);
        } else {
          const double wd = hi - lo;
          double y = std::fmod(Tn - lo, 2.0 * wd);
          if (y < 0) y += 2.0 * wd;
          if (y > wd) y = 2.0 * wd - y;
          Tn = lo + y;
        }
        if (!(Tn > lo) || !(Tn < hi)) Tn = T;
        x.time[a] = Tn;
        x.time[b] = Tn;
        // upload, score and weigh this one genealogy
        if ((rc = ima_slots_write(ctx, x.slot, 1, x.up0.data(), x.up1.data(), x.down.data(),
                                  x.time.data(), &x.root, nullptr)) != IMA_OK)
          return rc;
        double npd = 0.0;
        int32_t st = 0;
        if ((rc = ima_score(ctx, &items[i], 1, IMA_PREC_FP64, &npd, &st)) != IMA_OK) return rc;
        double pn = 0.0;
        double *wo = w.data() + (size_t)i * R, *a0 = all.data() + (size_t)c * R;
        if (prior) {
          if ((rc = weigh(i, wn.data())) != IMA_OK) return rc;
          for (int r = 0; r < R; ++r) an[r] = wsum_update(r, P, a0[r], wo[r], wn[r]);
          if ((rc = ima_integrate_tree_prob(ctx, model, an.data(), 1, &pn, nullptr, nullptr)) !=
              IMA_OK)
            return rc;
        }
        const double U = unif(mt);
        const double xmh = beta[c] * ((prior ? pn - probg[c] : 0.0) + (npd - pdg[i]));
        const bool acc = st == IMA_STATUS_OK && U < std::fmin(1.0, std::exp(xmh));
        if (acc) {
          pdg[i] = npd;
          if (prior) {
            std::copy(wn.begin(), wn.end(), wo);
            std::copy(an.begin(), an.end(), a0);
            probg[c] = pn;
          }
          ++accepts;
        } else {
          x.time[a] = T;
          x.time[b] = T;
        }
      }
    }
    // the heated-chain swap on the chains' sums
    std::vector<double> S(nchains);
    for (int c = 0; c < nchains; ++c) {
      double s = probg[c];
      for (int l = 0; l < nloci; ++l) s += pdg[c * nloci + l];
      S[c] = s;
    }
    if ((rc = ima_swap(ctx, nchains, S.data(), beta.data(), swaptries, seed,
                       (uint64_t)sweep << 16, nullptr, nullptr)) != IMA_OK)
      return rc;
  }
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (seconds_out) *seconds_out = secs;
  if (pdg_out) std::copy(pdg.begin(), pdg.end(), pdg_out);
  if (probg_out) std::copy(probg.begin(), probg.end(), probg_out);
  if (beta_out) std::copy(beta.begin(), beta.end(), beta_out);
  if (accepts_out) *accepts_out = accepts;
  if (time_out) {
    size_t o = 0;
    for (const Gen &x : g) {
      std::copy(x.time.begin(), x.time.end(), time_out + o);
      o += x.time.size();
    }
  }
  return IMA_OK;
}

}  // extern "C"
