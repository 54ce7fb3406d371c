"""The CPU baseline's C sweep (oracle/ora_mc3.c), pinned by the laws the
device sweep is held to: with a flat likelihood and one population its
chains sample the genealogy prior integrated over theta ~ U(0, q) -- the
two-copy coalescence time has CDF x E1(x) - e^-x + 1 (x = 2t/q), the
three-copy interval ratio 3r / (1 + 2r) -- and on the three-population
workload every genealogy stays valid for its chain's split times."""
import math

import numpy as np
import pytest
from scipy.optimize import brentq
from scipy.special import exp1

import oracle
from ima2p_amd import immodel, sim, workloads


def _flat_one_pop(n, C, seed):
    rng = np.random.default_rng(seed)
    pt = sim.PopTree(1, [])
    model = immodel.build(pt, qmax=8.0, mmax=5.0)
    loc = sim.HKYLocus(n, np.zeros((n, 1), np.int8), np.ones(1, np.int32), 1,
                       np.array([0.25, 0.25, 0.25, 0.25]))
    chains = [[sim.simulate_genealogy(rng, [n], [5.0], [], np.zeros((1, 1)), pt)]
              for _ in range(C)]
    return model, loc, chains


def _law(values, cdf, quantiles):
    v = np.asarray(values)  # [chains, samples]
    worst = 0.0
    for qn in quantiles:
        tau = brentq(lambda t: cdf(t) - qn, 1e-12, 1e6)
        frac = (v <= tau).mean(axis=1)
        se = frac.std(ddof=1) / math.sqrt(v.shape[0])
        worst = max(worst, abs(frac.mean() - qn) / max(se, 1e-3))
    return worst


@pytest.mark.parametrize("n", [2, 3])
def test_flat_likelihood_samples_the_prior(n):
    C = 32
    model, loc, chains = _flat_one_pop(n, C, 11 + n)
    tvs = np.zeros((C, 0))
    samples = []
    for k in range(160):
        oracle.mc3_bench([loc], [1e-9], 1.0, [[n]], model, chains, tvs, np.ones(C), 10.0,
                         seed=1000 * n + k, max_sweeps=30, scale_width=1.5)
        if k >= 10:
            row = []
            for c in range(C):
                g = chains[c][0]
                t = np.sort(g.time[g.down != -1])
                row.append(t.max() if n == 2 else t.min() / t.max())
            samples.append(row)
    v = np.array(samples).T
    if n == 2:
        def cdf(t):
            x = 2.0 * t / 8.0
            return x * exp1(x) - math.exp(-x) + 1.0
    else:
        def cdf(r):
            return 3 * r / (1 + 2 * r)
    assert _law(v, cdf, (0.1, 0.25, 0.5, 0.75, 0.9)) < 4.5


@pytest.mark.parametrize("n", [2, 3, 4])
def test_topology_move_flat_likelihood_samples_the_prior(n):
    """The oracle's restatement of the reference's own move (slider_nomigration,
    joinsisdown, splitsisdown, root-move Hastings term) with the time
    update switched off (scale width 0): n = 2, every move slides the root,
    the coalescence time follows x E1(x) - e^-x + 1; n = 3, each gene copy
    is the root's lone daughter with probability 1/3 and the interval ratio
    follows 3r / (1 + 2r); n = 4, the balanced topology has probability 1/3.
    Dropping the root move's Hastings term fails the n = 2 law."""
    C = 32
    model, loc, chains = _flat_one_pop(n, C, 21 + n)
    tvs = np.zeros((C, 0))
    stat, ratio = [], []
    for k in range(170):
        oracle.mc3_bench([loc], [1e-9], 1.0, [[n]], model, chains, tvs, np.ones(C), 10.0,
                         seed=5000 * n + k, max_sweeps=30, scale_width=0.0, topo=True)
        if k >= 20:
            row, rrow = [], []
            for c in range(C):
                g = chains[c][0]
                t = np.sort(g.time[g.down != -1])
                a, b = g.up0[g.root], g.up1[g.root]
                if n == 2:
                    row.append(t.max())
                elif n == 3:
                    row.append(a if a < 3 else b)
                    rrow.append(t.min() / t.max())
                else:
                    row.append(int(a >= 4 and b >= 4))
            stat.append(row)
            ratio.append(rrow)
    v = np.array(stat).T  # [chain, sample]
    if n == 2:
        def cdf(t):
            x = 2.0 * t / 8.0
            return x * exp1(x) - math.exp(-x) + 1.0
        assert _law(v, cdf, (0.1, 0.25, 0.5, 0.75, 0.9)) < 4.5
        return
    cats = range(3) if n == 3 else [1]
    for kcat in cats:
        f = (v == kcat).mean(axis=1)
        se = f.std(ddof=1) / math.sqrt(C)
        assert abs(f.mean() - 1 / 3) < 4.5 * max(se, 1e-3), (kcat, f.mean(), se)
    if n == 3:
        assert _law(np.array(ratio).T, lambda r: 3 * r / (1 + 2 * r),
                    (0.1, 0.25, 0.5, 0.75, 0.9)) < 4.5


def test_topology_move_keeps_two_population_labels_valid():
    """Two populations without migration: after topology moves every
    genealogy still fits the model (treeweight accepts its labels) and the
    stored likelihoods equal fresh evaluations."""
    rng = np.random.default_rng(31)
    pt, theta, ts, mig = sim.im_model_two_pop(t=0.6, m12=0.0, m21=0.0)
    model = immodel.build(pt, qmax=15.0, migration=False)
    samp = [4, 5]
    g0 = sim.simulate_genealogy(rng, samp, theta, ts, np.zeros((2, 2)), pt)
    loc = sim.encode_hky(sim.simulate_hky_sequences(rng, g0, 120, [0.3, 0.2, 0.2, 0.3], 4.0, 1.0))
    C = 6
    chains = [[g0.copy()] for _ in range(C)]
    tvs = np.tile(np.asarray(ts, float), (C, 1))
    oracle.mc3_bench([loc], [1.0], 4.0, [samp], model, chains, tvs, np.ones(C), 5.0, seed=9,
                     max_sweeps=200, scale_width=0.3, ry1=True, topo=True)
    changed = 0
    for c in range(C):
        g = chains[c][0]
        changed += int(not np.array_equal(g.up0, g0.up0))
        oracle.treeweight(g, model, tvs[c], samp)  # raises on labels that do not fit
    assert changed > 0


def test_three_population_sweep_keeps_genealogies_valid():
    wl = workloads.three_pop_hky(nchains=3, nloci=4, nsites=300, per_pop=5, pool=3)
    model = immodel.build(wl.poptree)
    chains = [[g.copy() for g in wl.chains[c]] for c in range(3)]
    tvs = np.tile(wl.tsplit, (3, 1))
    sw, st = oracle.mc3_bench(wl.loci, wl.u, wl.kappa, wl.samppop, model, chains, tvs,
                              workloads.geometric_betas(3), wl.tmax, max_sweeps=60)
    assert sw == 60 and st[0] == 60 * 3 * 4 and 0 < st[1] < st[0] and st[2] == 30 * 3
    for c in range(3):
        assert np.all(np.diff(tvs[c]) > 0) and tvs[c].max() <= wl.tmax
        for g in chains[c]:
            g.check()
            oracle.treeweight(g, model, tvs[c], wl.samppop)  # raises on a broken label
            for e in range(2 * g.n - 1):
                lo = g.time[g.up0[e]] if g.up0[e] >= 0 else 0.0
                for t, _ in g.mig[e]:
                    assert lo < t < g.time[e]


def _two_pop_posterior(seed):
    rng = np.random.default_rng(seed)
    pt, theta, ts, mig = sim.im_model_two_pop()
    model = immodel.build(pt, qmax=15.0, mmax=3.0)
    samp = [3, 3]
    loci, base = [], []
    for _ in range(2):
        g = sim.simulate_genealogy(rng, samp, theta, ts, mig, pt)
        loci.append(sim.encode_hky(sim.simulate_hky_sequences(
            rng, g, 150, np.array([0.3, 0.2, 0.2, 0.3]), 4.0, 2.0)))
        base.append(g)
    return model, samp, loci, base, np.asarray(ts, float)


def _sweep_samples(problem, C, burn, nsamp, thin, ry1, seed0):
    model, samp, loci, base, ts = problem
    chains = [[g.copy() for g in base] for _ in range(C)]
    tvs = np.ascontiguousarray(np.tile(ts, (C, 1)))
    k = 0

    def run(n):
        nonlocal k
        oracle.mc3_bench(loci, [2.0, 2.0], 4.0, samp, model, chains, tvs, np.ones(C), 5.0,
                         seed=seed0 + k, max_sweeps=n, scale_width=0.3, ry1=ry1)
        k += 1

    run(burn)
    out = []
    for _ in range(nsamp):
        run(thin)
        out.append([[float(g.time[g.up0[g.root]]) for g in chains[c]] + [float(tvs[c][0])]
                    for c in range(C)])
    return np.array(out).transpose(1, 0, 2)  # [chain, sample, stat]


def test_rubber_band_and_rescale_sample_the_same_posterior():
    """changet_RY1's rubber band (one split time, times between the
    neighbouring splits stretched, Hastings term from the counts of moved
    events) and the global rescale are different moves with the same
    target: two populations, two HKY loci with data, the genealogy prior
    and the split time, every chain at beta = 1.  The split time's and the
    root times' marginals agree at five quantiles (|z| < 4.5 with the
    chain-level spread); a wrong Hastings term shifts the split time."""
    problem = _two_pop_posterior(8)
    C, burn, nsamp, thin = 32, 300, 120, 20
    a = _sweep_samples(problem, C, burn, nsamp, thin, True, 50000)
    b = _sweep_samples(problem, C, burn, nsamp, thin, False, 60000)
    worst = 0.0
    for s in range(a.shape[2]):
        pooled = np.concatenate([a[:, :, s].ravel(), b[:, :, s].ravel()])
        for qn in (0.1, 0.25, 0.5, 0.75, 0.9):
            tau = np.quantile(pooled, qn)
            fa = (a[:, :, s] <= tau).mean(axis=1)
            fb = (b[:, :, s] <= tau).mean(axis=1)
            se = math.sqrt(fa.var(ddof=1) / len(fa) + fb.var(ddof=1) / len(fb))
            z = abs(fa.mean() - fb.mean()) / max(se, 1e-3)
            print(f"stat {s} q={qn:.2f} ry1={fa.mean():.4f} rescale={fb.mean():.4f} z={z:.2f}")
            worst = max(worst, z)
    assert worst < 4.5


def _fracs(values, taus):
    """Worst |mean fraction below tau - q| / chain standard error over the
    (q, tau) pairs; values is [chains, samples]."""
    v = np.asarray(values)
    worst = 0.0
    for qn, tau in taus:
        frac = (v <= tau).mean(axis=1)
        se = frac.std(ddof=1) / math.sqrt(v.shape[0])
        worst = max(worst, abs(frac.mean() - qn) / max(se, 1e-3))
    return worst


QS = (0.1, 0.25, 0.5, 0.75, 0.9)


def test_flat_likelihood_rates_walk_their_reflecting_bounds():
    """changeu and its kappa moves under a flat likelihood (u ~ 1e-15): every
    pair move is accepted, so log(u0/u1) is a walk reflected at
    +-3 log(10000) and each kappa one reflected at 0 and 100; both are
    uniform on their ranges at stationarity (the chains start there), and
    u0 u1 never changes."""
    C, n = 32, 2
    rng = np.random.default_rng(4)
    model, loc, chains = _flat_one_pop(n, C, 5)
    chains = [row + other for row, other in zip(chains, _flat_one_pop(n, C, 6)[2])]
    tvs = np.zeros((C, 0))
    R = 3 * math.log(10000.0)
    r0 = rng.uniform(-R, R, C)
    u = np.ascontiguousarray(np.stack([1e-15 * np.exp(r0 / 2), 1e-15 * np.exp(-r0 / 2)], 1))
    kappa = np.ascontiguousarray(rng.uniform(0.0, 100.0, (C, 2)))
    lr, ks = [], []
    for k in range(60):
        _, st = oracle.mc3_bench([loc, loc], None, None, [[n]], model, chains, tvs, np.ones(C),
                                 10.0, seed=500 + k, max_sweeps=500, scale_width=1.5,
                                 urate=True, rates=(u, kappa))
        assert st[4] == 100 * C and st[5] == st[4]  # one step a round, all accepted
        np.testing.assert_allclose(np.log(u).sum(axis=1), 2 * math.log(1e-15), rtol=0,
                                   atol=1e-9)
        lr.append(np.log(u[:, 0] / u[:, 1]))
        ks.append(kappa[:, 0].copy())
    lr, ks = np.array(lr).T, np.array(ks).T
    assert np.abs(lr).max() <= R and ks.min() >= 0 and ks.max() <= 100
    assert _fracs(lr, [(qn, -R + 2 * R * qn) for qn in QS]) < 4.5
    assert _fracs(ks, [(qn, 100.0 * qn) for qn in QS]) < 4.5


def test_changekappa_samples_the_kappa_posterior():
    """A lone HKY locus of two gene copies (changekappa, update_mc_params.cpp
    :387-429): the sampled kappa follows its marginal posterior
    p(kappa) ~ int L(t, kappa) (2/q) E1(2t/q) dt on [0, 100], the
    integral taken on a grid with the oracle's likelihood.  A flipped
    acceptance ratio or a missing beta moves the quantiles."""
    rng = np.random.default_rng(21)
    pt = sim.PopTree(1, [])
    q = 8.0
    model = immodel.build(pt, qmax=q, mmax=5.0)
    truth = sim.simulate_genealogy(rng, [2], [2.0], [], np.zeros((1, 1)), pt)
    truth.time[truth.down == truth.root] = 1.0  # the coalescence: top of both tip edges
    pi = np.array([0.3, 0.2, 0.2, 0.3])
    u = 60.0
    loc = sim.encode_hky(sim.simulate_hky_sequences(rng, truth, 300, pi, 4.0, u))
    # marginal posterior of kappa on a grid
    tg = np.geomspace(0.02, 20.0, 400)
    kg = np.linspace(0.05, 30.0, 240)
    g = truth.copy()
    ll = np.empty((len(kg), len(tg)))
    for i, kap in enumerate(kg):
        for j, t in enumerate(tg):
            g.time[g.down == g.root] = t
            ll[i, j] = oracle.likelihood_hky(g, loc, u, kap)[0]
    prior = (2 / q) * exp1(2 * tg / q)
    w = np.exp(ll - ll.max()) * prior
    mk = np.trapezoid(w, tg, axis=1)
    assert mk[-1] < 1e-6 * mk.max() and w[:, [0, -1]].max() < 1e-4 * w.max()
    cdfk = np.concatenate([[0.0], np.cumsum(0.5 * (mk[1:] + mk[:-1]) * np.diff(kg))])
    cdfk /= cdfk[-1]

    C = 32
    chains = [[truth.copy()] for _ in range(C)]
    rates = (np.full((C, 1), u), np.full((C, 1), 4.0))
    tvs = np.zeros((C, 0))
    samples = []
    for k in range(150):
        _, st = oracle.mc3_bench([loc], None, None, [[2]], model, chains, tvs, np.ones(C), 10.0,
                                 seed=90 + k, max_sweeps=50, scale_width=1.5, urate=True,
                                 rates=rates)
        assert st[4] == 10 * C and 0 < st[5] < st[4]
        if k >= 20:
            samples.append(rates[1][:, 0].copy())
    v = np.array(samples).T
    assert _fracs(v, [(qn, float(np.interp(qn, cdfk, kg))) for qn in QS]) < 4.5
