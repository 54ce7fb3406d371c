"""Device MC3 sweep with the genealogy prior (update_gtree.cpp:827-961 with
integrate_tree_prob over the chain's all-loci weights).

* State consistency: after genealogy proposals, rescalings (node,
  migration and split times) and swaps, every stored log-likelihood equals
  the oracle's on the genealogy now in its slot, and every chain's probg
  equals the oracle's integrate_tree_prob over its current genealogies and
  split times (1e-9: the sweep updates the all-loci sums incrementally, the
  oracle sums afresh).  Genealogies stay valid for the model.
* The same with changet_NW drawn as the other time update: migrations
  read back as NW rewrote them.
* Graph replay equals eager launches bit for bit.
* Stationary law, flat likelihood (u -> 0), one population: the sampled
  time to the most recent common ancestor of two gene copies follows the
  prior integrated over theta ~ U(0, q), density (2/q) E1(2t/q), CDF
  x E1(x) - e^-x + 1 with x = 2t/q; for three gene copies the ratio of the
  first coalescence interval to the tree height follows 3r / (1 + 2r)
  whatever theta is.  Checked at five quantiles of each law, with 64
  independent chains giving the standard error.
"""
import math

import numpy as np
import pytest
from scipy.special import exp1
from scipy.stats import kstest

import oracle
from ima2p_amd import _lib as L
from ima2p_amd import immodel, sim
from ima2p_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _three_pop_chains(C, Lc, seed, prec, site_block=0):
    rng = np.random.default_rng(seed)
    pt, theta, ts, mig = sim.im_model_three_pop(rng)
    model = immodel.build(pt, qmax=20.0, mmax=5.0)
    samp = [5, 4, 6]
    e = Engine(0)
    e.set_site_block(site_block)
    mid = e.model_add(model)
    locs, lids = [], []
    for _ in range(Lc):
        g0 = sim.simulate_genealogy(rng, samp, theta, ts, mig, pt)
        loc = sim.encode_hky(sim.simulate_hky_sequences(rng, g0, 300, rng.dirichlet([6] * 4),
                                                        4.0, 3.0))
        locs.append(loc)
        lid = e.add_locus_hky(loc)
        e.locus_set_model(lid, mid, samp)
        lids.append(lid)
    gens = [sim.relabel_internal(sim.simulate_genealogy(rng, samp, theta, ts, mig, pt), rng)
            for _ in range(C * Lc)]
    first = e.alloc_slots([lids[l] for c in range(C) for l in range(Lc)])
    e.write_slots(first, gens)
    e.write_pops(first, gens)
    u = rng.uniform(1.0, 3.0, Lc)
    betas = np.array([0.97 ** c for c in range(C)])
    e.mc3_setup(C, Lc, first, np.tile(u, C), 4.0, betas, seed=seed, precision=prec,
                scale_width=0.3)
    e.mc3_set_prior(np.tile(ts, (C, 1)), 10.0, model.npops)
    return e, model, samp, locs, gens, first, u, np.asarray(ts, float)


def _current(e, g, slot, times):
    """Genealogy g with the node and migration times the device now holds."""
    h = g.copy()
    h.time = times.copy()
    nm = sum(len(x) for x in g.mig)
    mt = e.read_mig_times(slot, nm)
    k = 0
    for edge in range(2 * g.n - 1):
        lst = sorted(g.mig[edge])
        h.mig[edge] = [(float(mt[k + j]), q) for j, (_, q) in enumerate(lst)]
        k += len(lst)
    return h


@pytest.mark.parametrize("prec,tol,ry1", [(L.PREC_FP64, 1e-9, False), (L.PREC_FP32, 1e-6, False),
                                          (L.PREC_FP64, 1e-9, True)])
def test_state_matches_oracle_after_sweeps(prec, tol, ry1):
    """ry1: the time update is changet_RY1's rubber band on one split time."""
    C, Lc = 6, 5
    e, model, samp, locs, gens, first, u, ts = _three_pop_chains(C, Lc, 31, prec)
    p0, _ = e.mc3_read_prior()
    e.mc3_run(40, L.MC3_GRAPH | L.MC3_SCALE | (L.MC3_RY1 if ry1 else 0))
    pdg, beta, stats, sweeps = e.mc3_read()
    probg, tv = e.mc3_read_prior()
    assert sweeps == 40 and stats[1] > 0 and stats[3] > 0
    assert np.any(np.abs(tv - ts[None, :]) > 1e-12)  # some rescaling accepted
    up0, up1, dn, tm, rt, _ = e.read_slots(first, C * Lc)
    off = 0
    for c in range(C):
        ws = []
        for l in range(Lc):
            i = c * Lc + l
            g = gens[i]
            ne = 2 * g.n - 1
            assert (up0[off:off + ne] == g.up0).all() and (dn[off:off + ne] == g.down).all()
            h = _current(e, g, first + i, tm[off:off + ne])
            off += ne
            ref, _ = oracle.likelihood_hky(h, locs[l], u[l], 4.0)
            assert abs(pdg[c, l] - ref) <= tol * abs(ref), (c, l, pdg[c, l], ref)
            ws.append(oracle.treeweight(h, model, tv[c], samp))  # raises if labels broke
        rp, _, _ = oracle.integrate_tree_prob(ws, model)
        assert abs(probg[c] - rp) <= 1e-9 * abs(rp), (c, probg[c], rp)
    assert np.any(probg != p0)
    e.close()


def _current_full(e, g, slot, times):
    """Genealogy g with the node times, populations and migrations the
    device now holds (the NW update changes the migrations)."""
    h = g.copy()
    h.time = times.copy()
    h.pop = e.read_pops(slot, 1)
    me, mt, mo = e.read_mig_events(slot)
    h.mig = [[] for _ in range(2 * g.n - 1)]
    for x in range(len(me)):
        h.mig[int(me[x])].append((float(mt[x]), int(mo[x])))
    return h


@pytest.mark.parametrize("prec,tol", [(L.PREC_FP64, 1e-9), (L.PREC_FP32, 1e-6)])
def test_state_matches_oracle_after_nw_sweeps(prec, tol):
    """Time updates drawn between changet_RY1 and changet_NW (IMA_MC3_NW):
    after 60 sweeps every stored likelihood equals the oracle's on the
    genealogy in its slot, every chain's probg equals integrate_tree_prob
    over its genealogies -- populations and migrations read back, as NW
    rewrote them -- at its split times, and the migration counts moved."""
    C, Lc = 6, 5
    e, model, samp, locs, gens, first, u, ts = _three_pop_chains(C, Lc, 37, prec)
    nm0 = [len(e.read_mig_events(first + i)[0]) for i in range(C * Lc)]
    e.mc3_run(60, L.MC3_GRAPH | L.MC3_SCALE | L.MC3_RY1 | L.MC3_NW)
    pdg, beta, stats, sweeps = e.mc3_read()
    probg, tv = e.mc3_read_prior()
    assert sweeps == 60 and stats[1] > 0 and stats[3] > 0
    up0, up1, dn, tm, rt, _ = e.read_slots(first, C * Lc)
    off, nm1 = 0, []
    for c in range(C):
        ws = []
        for l in range(Lc):
            i = c * Lc + l
            g = gens[i]
            ne = 2 * g.n - 1
            h = _current_full(e, g, first + i, tm[off:off + ne])
            off += ne
            nm1.append(sum(len(x) for x in h.mig))
            for lst in h.mig:
                assert [t for t, _ in lst] == sorted(t for t, _ in lst)
            ref, _ = oracle.likelihood_hky(h, locs[l], u[l], 4.0)
            assert abs(pdg[c, l] - ref) <= tol * abs(ref), (c, l, pdg[c, l], ref)
            ws.append(oracle.treeweight(h, model, tv[c], samp))  # raises if labels broke
        rp, _, _ = oracle.integrate_tree_prob(ws, model)
        assert abs(probg[c] - rp) <= 1e-9 * abs(rp), (c, probg[c], rp)
    assert nm1 != nm0
    e.close()


@pytest.mark.parametrize("ry1", [0, L.MC3_RY1, L.MC3_RY1 | L.MC3_NW],
                         ids=["rescale", "rubber_band", "rubber_band_nw"])
def test_graph_equals_eager_with_prior(ry1):
    out = []
    for flags in (L.MC3_GRAPH | L.MC3_SCALE | ry1, L.MC3_SCALE | ry1):
        e, *_ = _three_pop_chains(4, 3, 44, L.PREC_FP64)
        e.mc3_run(25, flags)
        pdg, beta, stats, _ = e.mc3_read()
        probg, tv = e.mc3_read_prior()
        out.append((pdg, beta, probg, tv, stats))
        e.close()
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


def _one_pop(n, C, seed, qmax):
    rng = np.random.default_rng(seed)
    pt = sim.PopTree(1, [])
    model = immodel.build(pt, qmax=qmax, mmax=5.0)
    e = Engine(0)
    mid = e.model_add(model)
    g0 = sim.simulate_genealogy(rng, [n], [5.0], [], np.zeros((1, 1)), pt)
    seq = np.zeros((n, 1), np.int8)
    loc = sim.HKYLocus(n, seq, np.ones(1, np.int32), 1, np.array([0.25, 0.25, 0.25, 0.25]))
    lid = e.add_locus_hky(loc)
    e.locus_set_model(lid, mid, [n])
    gens = [sim.simulate_genealogy(rng, [n], [5.0], [], np.zeros((1, 1)), pt) for _ in range(C)]
    first = e.alloc_slots([lid] * C)
    e.write_slots(first, gens)
    e.write_pops(first, gens)
    # u -> 0: the likelihood is flat to ~1e-9, the chains sample the prior
    e.mc3_setup(C, 1, first, 1e-9, 1.0, np.ones(C), seed=seed, precision=L.PREC_FP64,
                scale_width=1.5)
    e.mc3_set_prior(np.zeros((C, 0)), 10.0, 1)
    return e, first, gens


def _samples(e, first, C, n, nsamp, thin):
    out = []
    for _ in range(nsamp):
        e.mc3_run(thin, L.MC3_GRAPH | L.MC3_SCALE)
        _, up1, dn, tm, rt, _ = e.read_slots(first, C)
        ne = 2 * n - 1
        for c in range(C):
            t = tm[c * ne:(c + 1) * ne]
            d = dn[c * ne:(c + 1) * ne]
            out.append(np.sort(t[d != -1]))  # node times at the edges' upper ends
    return out


def _check_law(values_per_chain, cdf, quantiles):
    """Fraction of samples below each quantile of the target law against the
    law, with the spread of the independent chains' fractions as the error
    (the samples within a chain are autocorrelated, the chains are not)."""
    from scipy.optimize import brentq
    v = np.asarray(values_per_chain)  # [chains, samples]
    C = v.shape[0]
    worst = 0.0
    for qn in quantiles:
        tau = brentq(lambda t: cdf(t) - qn, 1e-12, 1e6)
        frac = (v <= tau).mean(axis=1)
        se = frac.std(ddof=1) / math.sqrt(C)
        z = abs(frac.mean() - qn) / max(se, 1e-3)
        print(f"  q={qn:.2f} tau={tau:.4f} frac={frac.mean():.4f} se={se:.4f} z={z:.2f}")
        worst = max(worst, z)
    return worst


def test_two_gene_copies_sample_the_integrated_prior():
    q = 8.0
    C = 64
    e, first, gens = _one_pop(2, C, 5, q)
    e.mc3_run(400)
    s = _samples(e, first, C, 2, 200, 40)
    tmrca = np.array([x.max() for x in s]).reshape(-1, C).T  # [chain, sample]

    def cdf(t):
        x = 2.0 * np.asarray(t) / q
        return float(x * exp1(x) - np.exp(-x) + 1.0)

    print("two copies: KS D", kstest(tmrca.ravel(), np.vectorize(cdf)).statistic)
    assert _check_law(tmrca, cdf, (0.1, 0.25, 0.5, 0.75, 0.9)) < 4.5
    e.close()


def test_three_gene_copies_interval_ratio():
    C = 64
    e, first, gens = _one_pop(3, C, 6, 8.0)
    e.mc3_run(400)
    s = _samples(e, first, C, 3, 200, 40)
    r = np.array([x.min() / x.max() for x in s]).reshape(-1, C).T

    def cdf(x):
        return 3 * x / (1 + 2 * x)

    print("three copies: KS D", kstest(r.ravel(), cdf).statistic)
    assert _check_law(r, cdf, (0.1, 0.25, 0.5, 0.75, 0.9)) < 4.5
    e.close()


def test_decisions_equal_for_every_group_size():
    """Deciding the loci in speculative groups of 1..6 gives the sweep of
    deciding them one at a time, bit for bit."""
    out = []
    for D in (1, 2, 3, 6):
        e, *_ = _three_pop_chains(5, 7, 52, L.PREC_FP64)
        assert e.mc3_prior_group(D) == D
        e.mc3_run(30, L.MC3_GRAPH | L.MC3_SCALE)
        pdg, beta, stats, _ = e.mc3_read()
        probg, tv = e.mc3_read_prior()
        out.append((pdg, beta, probg, tv, stats))
        e.close()
    for other in out[1:]:
        for a, b in zip(out[0], other):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("flags", [L.MC3_GRAPH | L.MC3_SCALE,
                                   L.MC3_GRAPH | L.MC3_SCALE | L.MC3_RY1 | L.MC3_UPDATE_U,
                                   L.MC3_GRAPH | L.MC3_SCALE | L.MC3_RY1 | L.MC3_UPDATE_U
                                   | L.MC3_NW], ids=["rescale", "rubber_band_rates", "with_nw"])
def test_tuner_leaves_the_chains_where_they_were(flags):
    """ima_mc3_tune times every group size, on the caller's sweeps (with
    the rubber band and the mutation-rate rounds in the second case, and
    NW drawn against RY1 in the third), from a saved state and restores it
    -- rates and migrations included: the sweeps that follow equal those of
    an untuned context.  The site block is fixed at 64 patterns: a tuned
    block of 128 would sum the FP32 pattern terms in another order."""
    out = []
    for tune in (False, True):
        e, model, samp, locs, gens, first, u, ts = _three_pop_chains(6, 9, 61, L.PREC_FP32,
                                                                     site_block=64)
        e.mc3_run(5, flags)
        if tune:
            best, ms = e.mc3_tune(8, flags)
            assert 1 <= best <= 6 and ms[best - 1] > 0 and ms[best - 1] == ms[ms > 0].min()
            assert e.mc3_prior_group() == best
        e.mc3_run(20, flags)
        pdg, beta, stats, sweeps = e.mc3_read()
        probg, tv = e.mc3_read_prior()
        u, kappa = e.mc3_read_rates()
        migs = np.concatenate([np.append(e.read_mig_events(first + i)[1], -1.0)
                               for i in range(6 * 9)])
        out.append((pdg, beta, probg, tv, stats, sweeps, u, kappa, migs))
        e.close()
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


def test_swaps_permute_the_heats_beyond_the_lds_staged_chains():
    """With more chains than the swap stages in LDS (256) the last chain
    swaps on the arrays in HBM: after sweeps with swaps every beta is still
    one of the initial ones, each used once, the swap counters advance, and
    graph replay equals eager launches."""
    out = []
    for flags in (L.MC3_GRAPH | L.MC3_SCALE | L.MC3_RY1, L.MC3_SCALE | L.MC3_RY1):
        e, *_ = _three_pop_chains(300, 2, 63, L.PREC_FP32)
        pdg0, beta0, _, _ = e.mc3_read()
        e.mc3_run(12, flags)
        pdg, beta, stats, sweeps = e.mc3_read()
        assert sweeps == 12 and stats[4] == 12 * 30 and stats[5] > 0
        np.testing.assert_array_equal(np.sort(beta), np.sort(beta0))
        probg, tv = e.mc3_read_prior()
        out.append((pdg, beta, probg, tv, stats))
        e.close()
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)
