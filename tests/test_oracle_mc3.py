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
