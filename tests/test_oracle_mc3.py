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
