// mcf_io.cpp -- the MCMC state file (.mcf) of IMa2p's checkpoint / resume,
// host side, no device needed.
//
// Format (writemcf / readmcf, mcmcfile.cpp:204-442; one line per item as
// awrite prints it, :33-76): "<name> <type> <count> " then the count values
// each followed by a space -- integers "%d ", doubles "%.10lg " -- and a
// newline.  Per chain: one "tvalue" per split time; per locus: "uvalue",
// "kappavalue" (HKY only), "pi[4]"; per edge of the genealogy (2n - 1,
// tips first): "up[2]", "down", "mut", "pop", for stepwise loci "A[]" and
// "dlikeA[]", "time", the migration events of the edge in time order as
// "mig[].mt" + "mig[].mp", closed by a "mig[].mt" of -1; for the internal
// edges of HKY loci the conditional likelihoods of every site pattern,
// "C[ci]->G[li].gtree[i].hkyi.frac[j]" and "...newfrac[j]", four doubles
// each.  The engine keeps no stored partials, so it writes those as zeros:
// readmcf reads them and init_p (:131-193) recomputes every likelihood
// from the genealogies, so they do not reach the state.
//
// Reading checks every item's name and type as aread does (:79-128); when
// the file ends before every chain is loaded it is read again from the
// start for the remaining chains, and a chain whose split times exceed the
// prior's upper bounds is read over by the next one (readmcf, :310-442).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ima_gpu.h"

namespace {

const char *kFrac = "C[ci]->G[li].gtree[i].hkyi.frac[j]";
const char *kNewFrac = "C[ci]->G[li].gtree[i].hkyi.newfrac[j]";

void put_i(FILE *f, const char *name, const int32_t *v, int n) {
  fprintf(f, "%s %d %d ", name, 0, n);
  for (int i = 0; i < n; ++i) fprintf(f, "%d ", v[i]);
  fprintf(f, "\n");
}
void put_d(FILE *f, const char *name, const double *v, int n) {
  fprintf(f, "%s %d %d ", name, 3, n);
  for (int i = 0; i < n; ++i) fprintf(f, "%.10lg ", v[i]);
  fprintf(f, "\n");
}

struct Reader {
  FILE *f = nullptr;
  const char *path;
  bool bad = false;
  // aread: the item's name and type must match; returns its count
  int head(const char *name, int type) {
    char got[100];
    int t = -1, n = -1;
    if (fscanf(f, "%99s %d %d ", got, &t, &n) != 3 || strcmp(got, name) != 0 || t != type ||
        n < 0) {
      bad = true;
      return -1;
    }
    return n;
  }
  bool get_i(const char *name, int32_t *v, int want) {
    const int n = head(name, 0);
    if (n != want) return !(bad = true);
    for (int i = 0; i < n; ++i)
      if (fscanf(f, "%d ", &v[i]) != 1) return !(bad = true);
    return true;
  }
  bool get_d(const char *name, double *v, int want) {
    const int n = head(name, 3);
    if (n != want) return !(bad = true);
    for (int i = 0; i < n; ++i)
      if (fscanf(f, "%lg ", &v[i]) != 1) return !(bad = true);
    return true;
  }
};

int64_t edges_before(const ima_mcf_locus *loci, int nloci, int c, int l) {
  int64_t per_chain = 0, k = 0;
  for (int j = 0; j < nloci; ++j) per_chain += 2 * (int64_t)loci[j].n - 1;
  for (int j = 0; j < l; ++j) k += 2 * (int64_t)loci[j].n - 1;
  return (int64_t)c * per_chain + k;
}

bool layout_ok(int nchains, int nsplit, int nloci, const ima_mcf_locus *loci) {
  if (nchains < 1 || nsplit < 0 || nloci < 1 || !loci) return false;
  for (int l = 0; l < nloci; ++l)
    if (loci[l].n < 2 || (loci[l].model != IMA_MODEL_HKY && loci[l].model != IMA_MODEL_IS &&
                          loci[l].model != IMA_MODEL_SW) ||
        (loci[l].model == IMA_MODEL_HKY && loci[l].nsites < 1))
      return false;
  return true;
}

}  // namespace

extern "C" int ima_mcf_write(const char *path, int nchains, int nsplit, int nloci,
                             const ima_mcf_locus *loci, const ima_mcf_state *st) {
  if (!path || !st || !layout_ok(nchains, nsplit, nloci, loci)) return IMA_ERR_ARG;
  if ((nsplit && !st->tvals) || !st->u || !st->pi || !st->up0 || !st->up1 || !st->down ||
      !st->pop || !st->time || !st->mig_count)
    return IMA_ERR_ARG;
  FILE *f = fopen(path, "w");
  if (!f) return IMA_ERR_ARG;
  const int32_t zero_i = 0;
  const double minus1 = -1.0;
  const double zero4[4] = {0, 0, 0, 0};
  int64_t m = 0;  // migration event cursor
  for (int c = 0; c < nchains; ++c) {
    for (int k = 0; k < nsplit; ++k) put_d(f, "tvalue", &st->tvals[(size_t)c * nsplit + k], 1);
    for (int l = 0; l < nloci; ++l) {
      const ima_mcf_locus &L = loci[l];
      const size_t cl = (size_t)c * nloci + l;
      put_d(f, "uvalue", &st->u[cl], 1);  // one linked locus: nlinked = 1
      if (L.model == IMA_MODEL_HKY) put_d(f, "kappavalue", st->kappa ? &st->kappa[cl] : zero4, 1);
      put_d(f, "pi[4]", &st->pi[cl * 4], 4);
      const int64_t e0 = edges_before(loci, nloci, c, l);
      for (int i = 0; i < 2 * L.n - 1; ++i) {
        const int64_t e = e0 + i;
        const int32_t up[2] = {st->up0[e], st->up1[e]};
        put_i(f, "up[2]", up, 2);
        put_i(f, "down", &st->down[e], 1);
        put_i(f, "mut", st->mut ? &st->mut[e] : &zero_i, 1);
        put_i(f, "pop", &st->pop[e], 1);
        if (L.model == IMA_MODEL_SW) {
          put_i(f, "A[]", st->A ? &st->A[e] : &zero_i, 1);
          put_d(f, "dlikeA[]", st->dlikeA ? &st->dlikeA[e] : zero4, 1);
        }
        put_d(f, "time", &st->time[e], 1);
        for (int j = 0; j < st->mig_count[e]; ++j, ++m) {
          put_d(f, "mig[].mt", &st->mig_time[m], 1);
          put_i(f, "mig[].mp", &st->mig_to[m], 1);
        }
        put_d(f, "mig[].mt", &minus1, 1);
        if (L.model == IMA_MODEL_HKY && i >= L.n)
          for (int j = 0; j < L.nsites; ++j) {
            put_d(f, kFrac, zero4, 4);
            put_d(f, kNewFrac, zero4, 4);
          }
      }
    }
  }
  const bool ok = !ferror(f);
  return (fclose(f) == 0 && ok) ? IMA_OK : IMA_ERR_ARG;
}

extern "C" int ima_mcf_read(const char *path, int nchains, int nsplit, int nloci,
                            const ima_mcf_locus *loci, const double *tmax, ima_mcf_state *st,
                            int64_t *nmig) {
  if (!path || !st || !layout_ok(nchains, nsplit, nloci, loci)) return IMA_ERR_ARG;
  if ((nsplit && !st->tvals) || !st->u || !st->pi || !st->up0 || !st->up1 || !st->down ||
      !st->pop || !st->time || !st->mig_count || (st->mig_capacity > 0 && (!st->mig_time ||
                                                                          !st->mig_to)))
    return IMA_ERR_ARG;
  Reader r;
  r.path = path;
  if (!(r.f = fopen(path, "r"))) return IMA_ERR_ARG;
  int lastc = -1;
  std::vector<double> tv(nsplit > 0 ? nsplit : 1);
  std::vector<int64_t> mig_start(nchains + 1, 0);  // event cursor at each chain's start
  int64_t m = 0;
  double scratch[4];
  for (int c = 0; c < nchains; ++c) {
    m = mig_start[c];
    bool large = false;
    for (int k = 0; k < nsplit; ++k) {
      if (!r.get_d("tvalue", &st->tvals[(size_t)c * nsplit + k], 1)) break;
      if (tmax && st->tvals[(size_t)c * nsplit + k] > tmax[k]) large = true;
    }
    for (int l = 0; l < nloci && !r.bad; ++l) {
      const ima_mcf_locus &L = loci[l];
      const size_t cl = (size_t)c * nloci + l;
      if (!r.get_d("uvalue", &st->u[cl], 1)) break;
      if (L.model == IMA_MODEL_HKY &&
          !r.get_d("kappavalue", st->kappa ? &st->kappa[cl] : scratch, 1))
        break;
      if (!r.get_d("pi[4]", &st->pi[cl * 4], 4)) break;
      const int64_t e0 = edges_before(loci, nloci, c, l);
      for (int i = 0; i < 2 * L.n - 1 && !r.bad; ++i) {
        const int64_t e = e0 + i;
        int32_t up[2], mut;
        if (!r.get_i("up[2]", up, 2)) break;
        st->up0[e] = up[0];
        st->up1[e] = up[1];
        if (!r.get_i("down", &st->down[e], 1) || !r.get_i("mut", &mut, 1)) break;
        if (st->mut) st->mut[e] = mut;
        if (!r.get_i("pop", &st->pop[e], 1)) break;
        if (L.model == IMA_MODEL_SW) {
          int32_t a;
          if (!r.get_i("A[]", &a, 1) || !r.get_d("dlikeA[]", scratch, 1)) break;
          if (st->A) st->A[e] = a;
          if (st->dlikeA) st->dlikeA[e] = scratch[0];
        }
        if (!r.get_d("time", &st->time[e], 1)) break;
        int cnt = 0;
        for (;;) {
          double mt;
          if (!r.get_d("mig[].mt", &mt, 1)) break;
          if (!(mt > 0)) break;
          int32_t mp;
          if (!r.get_i("mig[].mp", &mp, 1)) break;
          if (m >= st->mig_capacity) {
            fclose(r.f);
            return IMA_ERR_SIZE;
          }
          st->mig_time[m] = mt;
          st->mig_to[m] = mp;
          ++m;
          ++cnt;
        }
        st->mig_count[e] = cnt;
        if (L.model == IMA_MODEL_HKY && i >= L.n)
          for (int j = 0; j < L.nsites && !r.bad; ++j)
            if (!r.get_d(kFrac, scratch, 4) || !r.get_d(kNewFrac, scratch, 4)) break;
      }
    }
    if (r.bad) {
      fclose(r.f);
      return IMA_ERR_ARG;
    }
    // end of file before the last chain: read the file again from the start
    // (the same chain index twice means every chain in it is out of range)
    const int ch = fgetc(r.f);
    if (ch == EOF) {
      if (c == lastc) {
        fclose(r.f);
        return IMA_ERR_ARG;
      }
      if (c < nchains - 1) {
        fclose(r.f);
        if (!(r.f = fopen(path, "r"))) return IMA_ERR_ARG;
        lastc = c;
      }
    } else {
      ungetc(ch, r.f);
    }
    if (large) {
      --c;  // read over by the next chain in the file
    } else {
      mig_start[c + 1] = m;
    }
  }
  fclose(r.f);
  if (nmig) *nmig = m;
  return IMA_OK;
}
