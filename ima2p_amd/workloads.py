"""Synthetic workloads at the benchmark configurations.

``three_pop_hky`` is the headline configuration: three-population IM model
with tree ((0,1),2), theta ~ U(1,10) for the five populations, split times
U(0.5,3), migration U(0,1) between coexisting populations (seed 2); fifty
HKY loci of 1000 bp with kappa = 4, base frequencies ~ Dirichlet(10,10,10,10),
15 gene copies per population, per-locus mutation scalars ~ LogNormal(0,0.5);
64 heated chains.

``two_pop_is`` is the reference's CPU-runnable case: two populations,
theta 5/5/5, t = 1, m = 0.5 both ways, five IS loci with 10+10 gene copies
(seed 1), 4 chains with geometric heating (-ha 0.96 -hb 0.9).

Chains start from genealogies simulated under the true model (a pool per
locus, relabelled per chain), since the reference's starting-tree builder
is not part of this engine.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import sim


@dataclass
class Workload:
    name: str
    nchains: int
    model: str                     # "hky" or "is"
    loci: list                     # HKYLocus / ISLocus
    u: np.ndarray                  # per-locus mutation scalar
    kappa: float
    chains: list = field(default_factory=list)  # chains[c][l] -> Genealogy
    beta: np.ndarray = None
    seed: int = 0
    poptree: sim.PopTree = None    # population tree of the IM model
    tsplit: np.ndarray = None      # split times (chains start at the truth)
    samppop: list = None           # gene copies per sampled population
    tmax: float = 5.0              # split-time prior upper bound (-t)

    @property
    def nloci(self):
        return len(self.loci)

    def patterns(self):
        return np.array([loc.seq.shape[1] for loc in self.loci])


def geometric_betas(nchains, ha=0.96, hb=0.9):
    """setheat HGEOMETRIC (swapchains.cpp:152-153)."""
    if nchains == 1:
        return np.ones(1)
    i = np.arange(nchains, dtype=float)
    return 1 - (1 - hb) * i * np.power(ha, nchains - 1 - i) / (nchains - 1)


def _chain_pool(rng, simulate, nloci, nchains, pool):
    trees = []
    for _ in range(nloci):
        base = [simulate() for _ in range(min(pool, nchains))]
        trees.append([sim.relabel_internal(base[c % len(base)], rng) for c in range(nchains)])
    return [[trees[l][c] for l in range(nloci)] for c in range(nchains)]


def three_pop_hky(nchains=64, nloci=50, nsites=1000, per_pop=15, seed=2, pool=16):
    rng = np.random.default_rng(seed)
    pt, theta, ts, mig = sim.im_model_three_pop(rng)
    u = rng.lognormal(0.0, 0.5, nloci)
    loci = []
    for l in range(nloci):
        g = sim.simulate_genealogy(rng, [per_pop] * 3, theta, ts, mig, pt)
        pi = rng.dirichlet([10, 10, 10, 10])
        raw = sim.simulate_hky_sequences(rng, g, nsites, pi, 4.0, u[l])
        loci.append(sim.encode_hky(raw))
    chains = _chain_pool(rng, lambda: sim.simulate_genealogy(rng, [per_pop] * 3, theta, ts,
                                                             mig, pt), nloci, nchains, pool)
    return Workload("3pop-50xHKY1000bp-n45-64chains", nchains, "hky", loci, u, 4.0, chains,
                    geometric_betas(nchains), seed, pt, np.asarray(ts, float), [per_pop] * 3)


def two_pop_is(nchains=4, nloci=5, per_pop=10, seed=1, pool=4):
    rng = np.random.default_rng(seed)
    pt, theta, ts, mig = sim.im_model_two_pop()
    loci, truth = [], []
    for _ in range(nloci):
        g = sim.simulate_genealogy(rng, [per_pop] * 2, theta, ts, mig, pt)
        loci.append(sim.simulate_is(rng, g, 1.0))
        truth.append(g)
    # IS chains must start on genealogies compatible with the data: use the
    # generating genealogy (times perturbed per chain keep the topology)
    chains = []
    for c in range(nchains):
        row = []
        for l in range(nloci):
            g = sim.relabel_internal(truth[l], rng)
            row.append(g)
        chains.append(row)
    return Workload("2pop-5xIS-n20-4chains", nchains, "is", loci, np.ones(nloci), 1.0, chains,
                    geometric_betas(nchains), seed, pt, np.asarray(ts, float), [per_pop] * 2)


# ----------------------------------------------------------------- serving --

@dataclass
class LocusPool:
    """Loci simulated under one IM model, drawn from by serving analyses.
    entries[i] = (kind, data, genealogy, allele states or None, u, samppop)."""
    npops: int
    poptree: sim.PopTree
    tsplit: np.ndarray
    model: object
    entries: list


def serving_pool(rng, P, size=48, mix=(0.5, 1 / 3, 1 / 6), copies=(2, 10),
                 hky_sites=(200, 1500)):
    """Mixed HKY / IS / SW loci (the 150:100:50 proportions of the uneven
    batch) with copies[0]..copies[1] gene copies per sampled population."""
    from .immodel import build
    pt, theta, ts, mig = sim.im_model_random(rng, P)
    model = build(pt, qmax=20.0, mmax=5.0)
    kinds = rng.choice(["hky", "is", "sw"], size, p=list(mix))
    entries = []
    for kind in kinds:
        samp = [int(x) for x in rng.integers(copies[0], copies[1] + 1, P)]
        while True:
            g = sim.simulate_genealogy(rng, samp, theta, ts, mig, pt)
            if kind == "hky":
                u = float(rng.lognormal(0.0, 0.5))
                nsites = int(rng.integers(hky_sites[0], hky_sites[1] + 1))
                data = sim.encode_hky(sim.simulate_hky_sequences(
                    rng, g, nsites, rng.dirichlet([10] * 4), 4.0, u))
                A = None
            elif kind == "is":
                u = float(rng.uniform(1.0, 3.0))
                data = sim.simulate_is(rng, g, u)
                A = None
                if data.seq.shape[1] == 0:
                    continue
            else:
                u = float(rng.uniform(1.0, 3.0))
                A = sim.simulate_sw(rng, g, u)
                data = g.n
            break
        entries.append((str(kind), data, g, A, u, samp))
    return LocusPool(P, pt, np.asarray(ts, float), model, entries)


def serving_analysis(rng, pool: LocusPool, nloci, nchains, nsweeps=100, seed=1,
                     precision=None):
    """An IM analysis of nloci loci drawn from the pool.  Every chain starts
    on the genealogies that generated the data (compatible with the IS
    sites), at the generating split times."""
    from . import _lib as L
    from .serve import AnalysisSpec
    idx = rng.integers(0, len(pool.entries), nloci)
    ent = [pool.entries[i] for i in idx]
    loci = [(k, d) for k, d, *_ in ent]
    base = [g for _, _, g, *_ in ent]
    As = [A for _, _, _, A, *_ in ent]
    u = np.array([e[4] for e in ent])
    kappa = np.array([4.0 if e[0] == "hky" else 1.0 for e in ent])
    P = pool.npops
    return AnalysisSpec(
        loci=loci, model=pool.model, samppop=[e[5] for e in ent],
        chains=[base] * nchains, u=np.tile(u, (nchains, 1)),
        kappa=np.tile(kappa, (nchains, 1)), beta=geometric_betas(nchains),
        tvals=np.tile(pool.tsplit, (nchains, 1)) if P > 1 else np.zeros((nchains, 0)),
        tmax=5.0, seed=seed, precision=L.PREC_FP32 if precision is None else precision,
        nsweeps=nsweeps, A=[As] * nchains)


def uneven_mixed(nchains=256, seed=3, hky=150, is_=100, sw=50, copies=(4, 60),
                 hky_sites=(200, 3000), precision=None):
    """The uneven batch: a two-population IM model (theta 5/5/5, t = 1,
    m = 0.5 both ways) with 300 loci, 150 HKY of U{200..3000} bp, 100
    infinite-sites and 50 stepwise (starting allele 20, walk simulated on
    every branch), U{4..60} gene copies per locus split across the two
    populations, seed 3; 256 chains starting on the generating genealogies
    at the true split time."""
    from . import _lib as L
    from .immodel import build
    from .serve import AnalysisSpec
    rng = np.random.default_rng(seed)
    pt, theta, ts, mig = sim.im_model_two_pop()
    model = build(pt, qmax=20.0, mmax=5.0)
    kinds = ["hky"] * hky + ["is"] * is_ + ["sw"] * sw
    loci, base, As, u, samp = [], [], [], [], []
    for kind in kinds:
        n = int(rng.integers(copies[0], copies[1] + 1))
        a = int(rng.integers(1, n))
        sp = [a, n - a]
        while True:
            g = sim.simulate_genealogy(rng, sp, theta, ts, mig, pt)
            A = None
            if kind == "hky":
                uu = float(rng.lognormal(0.0, 0.5))
                nsites = int(rng.integers(hky_sites[0], hky_sites[1] + 1))
                data = sim.encode_hky(sim.simulate_hky_sequences(
                    rng, g, nsites, rng.dirichlet([10] * 4), 4.0, uu))
            elif kind == "is":
                uu = float(rng.uniform(0.5, 2.0))
                data = sim.simulate_is(rng, g, uu)
                if data.seq.shape[1] == 0:
                    continue
            else:
                uu = float(rng.uniform(0.5, 2.0))
                A = sim.simulate_sw(rng, g, uu, start=20)
                data = g.n
            break
        loci.append((kind, data))
        base.append(g)
        As.append(A)
        u.append(uu)
        samp.append(sp)
    kappa = np.array([4.0 if k == "hky" else 1.0 for k in kinds])
    return AnalysisSpec(loci=loci, model=model, samppop=samp, chains=[base] * nchains,
                        u=np.tile(u, (nchains, 1)), kappa=np.tile(kappa, (nchains, 1)),
                        beta=geometric_betas(nchains), tvals=np.tile(np.asarray(ts, float),
                                                                     (nchains, 1)),
                        tmax=5.0, seed=seed,
                        precision=L.PREC_FP32 if precision is None else precision,
                        nsweeps=0, A=[As] * nchains)


def serving_requests(rng, duration_s=20.0, rate=20.0, nsweeps=100,
                     tab_sizes=(10000, 100000, 1000000), pool_size=48):
    """The serving stream's requests, generated before it starts: Poisson
    arrivals at `rate`/s for `duration_s` s; 70% L-mode query batches of
    1..4096 points (log-uniform) on resident tables of `tab_sizes`
    genealogies, 30% new IM analyses with 2..4 populations, 5..300 mixed
    HKY/IS/SW loci and 16..128 chains running `nsweeps` sweeps.

    Returns (tables, pools, requests); requests[i] = (arrival time, "q",
    (table index, points)) or (arrival time, "a", AnalysisSpec)."""
    pools = {P: serving_pool(rng, P, size=pool_size) for P in (2, 3, 4)}
    tabs = [sim.lmode_table(rng, G) for G in tab_sizes]
    lo, hi = np.array([0.5, 0.5, 0.5, 0.01, 0.01]), np.array([10, 10, 10, 2, 2])
    t, reqs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration_s:
            break
        if rng.random() < 0.7:
            k = int(rng.integers(0, len(tabs)))
            npts = int(np.exp(rng.uniform(0.0, np.log(4096.0))))
            reqs.append((t, "q", (k, rng.uniform(lo, hi, (npts, 5)))))
        else:
            P = int(rng.integers(2, 5))
            spec = serving_analysis(rng, pools[P], int(rng.integers(5, 301)),
                                    int(rng.integers(16, 129)), nsweeps=nsweeps,
                                    seed=int(rng.integers(1, 1 << 30)))
            reqs.append((t, "a", spec))
    return tabs, pools, reqs
