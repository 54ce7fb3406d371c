"""The device sweep and the oracle's C sweep sample the same posterior.

Two populations (theta 5/5/5, t = 1, m = 0.5 both ways), two HKY loci of 6
gene copies and 150 bp with real (simulated) data, the genealogy prior and
the split time; every chain at beta = 1, so each chain samples the full
posterior.  Both sweeps start from the same genealogies and run the same
moves with independent random numbers.  At five quantiles of the pooled
samples of each statistic -- the root time of each locus and the split
time -- the fraction of each sampler's draws below the quantile must agree:
|difference| / (standard error from the independent chains of both) < 4.5.
This is the two-sample comparison of marginal posteriors, with the
autocorrelation inside a chain accounted for by the chain-level spread.
Both time updates are compared: the global rescale, and changet_RY1's
rubber band on one split time (IMA_MC3_RY1), alone and drawn against
changet_NW (IMA_MC3_NW) as qupdate does, with the migration counts
compared as well; with the mutation-rate rounds
(IMA_MC3_UPDATE_U, changeu every fifth sweep) the rate ratio log(u0/u1)
and both loci's kappa are compared too.
"""
import math

import numpy as np
import pytest

import oracle
from ima2p_amd import _lib as L
from ima2p_amd import immodel, sim
from ima2p_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _setup(seed):
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


def _stats(gens, tv):
    return [float(g.time[g.up0[g.root]]) for g in gens] + [float(tv[0])]


def _gpu_samples(model, samp, loci, base, ts, C, burn, nsamp, thin, ry1, urate, nw=False):
    e = Engine(0)
    mid = e.model_add(model)
    lids = []
    for loc in loci:
        lid = e.add_locus_hky(loc)
        e.locus_set_model(lid, mid, samp)
        lids.append(lid)
    Lc = len(loci)
    gens = [base[l].copy() for c in range(C) for l in range(Lc)]
    first = e.alloc_slots([lids[l] for c in range(C) for l in range(Lc)])
    e.write_slots(first, gens)
    e.write_pops(first, gens)
    e.mc3_setup(C, Lc, first, 2.0, 4.0, np.ones(C), seed=77, precision=L.PREC_FP64,
                scale_width=0.3)
    e.mc3_set_prior(np.tile(ts, (C, 1)), 5.0, 2)
    flags = (L.MC3_GRAPH | L.MC3_SCALE | (L.MC3_RY1 if ry1 else 0)
             | (L.MC3_UPDATE_U if urate else 0) | (L.MC3_NW if nw else 0))
    e.mc3_run(burn, flags)
    out = []
    for _ in range(nsamp):
        e.mc3_run(thin, flags)
        _, _, dn, tm, rt, _ = e.read_slots(first, C * Lc)
        _, tv = e.mc3_read_prior()
        un, kn = e.mc3_read_rates()
        row, off = [], 0
        for c in range(C):
            vals = []
            for l in range(Lc):
                g = gens[c * Lc + l]
                ne = 2 * g.n - 1
                t = tm[off:off + ne]
                vals.append(float(t[g.up0[g.root]]))
                off += ne
            vals.append(float(tv[c][0]))
            if urate:
                vals += [float(np.log(un[c, 0] / un[c, 1])), float(kn[c, 0]), float(kn[c, 1])]
            if nw:
                vals.append(float(sum(len(e.read_mig_events(first + c * Lc + l)[0])
                                      for l in range(Lc))))
            row.append(vals)
        out.append(row)
    e.close()
    return np.array(out).transpose(1, 0, 2)  # [chain, sample, stat]


def _cpu_samples(model, samp, loci, base, ts, C, burn, nsamp, thin, ry1, urate, nw=False):
    chains = [[g.copy() for g in base] for _ in range(C)]
    tvs = np.ascontiguousarray(np.tile(ts, (C, 1)))
    rates = (np.full((C, 2), 2.0), np.full((C, 2), 4.0))
    k = 0

    def run(n):
        nonlocal k
        oracle.mc3_bench(loci, None, None, samp, model, chains, tvs, np.ones(C), 5.0,
                         seed=90000 + k, max_sweeps=n, scale_width=0.3, ry1=ry1, urate=urate,
                         rates=rates, nw=1 if nw else 0)
        k += 1

    def stats(c):
        v = _stats(chains[c], tvs[c])
        if urate:
            u, kap = rates
            v += [float(np.log(u[c, 0] / u[c, 1])), float(kap[c, 0]), float(kap[c, 1])]
        if nw:
            v.append(float(sum(len(m) for g in chains[c] for m in g.mig)))
        return v

    run(burn)
    out = []
    for _ in range(nsamp):
        run(thin)
        out.append([stats(c) for c in range(C)])
    return np.array(out).transpose(1, 0, 2)


@pytest.mark.parametrize("ry1,urate,nw", [(False, False, False), (True, False, False),
                                           (True, True, False), (True, False, True)],
                         ids=["rescale", "rubber_band", "rubber_band_rates", "rubber_band_nw"])
def test_device_and_oracle_sample_the_same_posterior(ry1, urate, nw):
    """With nw, each time update is changet_NW with probability 1/2 (device:
    IMA_MC3_NW; oracle: nw=1) and the chains' migration counts are compared
    too."""
    model, samp, loci, base, ts = _setup(8)
    C, burn, nsamp, thin = 48, 400, 150, 20
    a = _gpu_samples(model, samp, loci, base, ts, C, burn, nsamp, thin, ry1, urate, nw)
    b = _cpu_samples(model, samp, loci, base, ts, C, burn, nsamp, thin, ry1, urate, nw)
    worst = 0.0
    for s in range(a.shape[2]):
        pooled = np.concatenate([a[:, :, s].ravel(), b[:, :, s].ravel()])
        for qn in (0.1, 0.25, 0.5, 0.75, 0.9):
            tau = np.quantile(pooled, qn)
            fa = (a[:, :, s] <= tau).mean(axis=1)
            fb = (b[:, :, s] <= tau).mean(axis=1)
            se = math.sqrt(fa.var(ddof=1) / len(fa) + fb.var(ddof=1) / len(fb))
            z = abs(fa.mean() - fb.mean()) / max(se, 1e-3)
            print(f"stat {s} q={qn:.2f} gpu={fa.mean():.4f} cpu={fb.mean():.4f} z={z:.2f}")
            worst = max(worst, z)
    assert worst < 4.5


def _topo_setup(seed, mig):
    rng = np.random.default_rng(seed)
    if mig:
        pt, theta, ts, m = sim.im_model_two_pop(t=0.8, m12=0.5, m21=0.5)
        model = immodel.build(pt, qmax=15.0, mmax=3.0)
    else:
        pt, theta, ts, m = sim.im_model_two_pop(t=0.8, m12=0.0, m21=0.0)
        model = immodel.build(pt, qmax=15.0, migration=False)
    samp = [3, 3]
    loci, base = [], []
    for _ in range(2):
        g = sim.simulate_genealogy(rng, samp, theta, ts, m, pt)
        loci.append(sim.encode_hky(sim.simulate_hky_sequences(
            rng, g, 150, np.array([0.3, 0.2, 0.2, 0.3]), 4.0, 2.0)))
        base.append(g)
    return model, samp, loci, base, np.asarray(ts, float)


def _topo_stats(up0, up1, dn, tm, roots, n, Lc):
    """Per locus: the root time and whether gene copies 0 and 1 form a
    cherry; every array packed over the chain's loci."""
    out, off = [], 0
    ne = 2 * n - 1
    for l in range(Lc):
        u0, u1, t, rt = up0[off:off + ne], up1[off:off + ne], tm[off:off + ne], roots[l]
        off += ne
        cherry = any((u0[e] == 0 and u1[e] == 1) or (u0[e] == 1 and u1[e] == 0)
                     for e in range(n, ne))
        out += [float(t[u0[rt]]), float(cherry)]
    return out


@pytest.mark.parametrize("mig", [False, True], ids=["no_migration", "migration_nw"])
def test_device_and_oracle_sample_the_same_posterior_with_topology_moves(mig):
    """The reference's own genealogy move (IMA_MC3_TOPOLOGY) on the device
    and in the oracle's sweep (two populations, two HKY loci, the
    rubber-band split-time update on alternate sweeps): root times, the
    split time and a topology statistic (gene copies 0 and 1 a cherry) have
    the same marginal posteriors (|z| < 4.5).  With migration the move draws
    migration paths, each time update is changet_NW with probability 1/2,
    and the chains' migration counts are compared too."""
    model, samp, loci, base, ts = _topo_setup(19, mig)
    C, burn, nsamp, thin = 48, 400, 150, 20
    n, Lc = base[0].n, len(loci)
    # device
    e = Engine(0)
    mid = e.model_add(model)
    lids = []
    for loc in loci:
        lid = e.add_locus_hky(loc)
        e.locus_set_model(lid, mid, samp)
        lids.append(lid)
    gens = [base[l].copy() for c in range(C) for l in range(Lc)]
    first = e.alloc_slots([lids[l] for c in range(C) for l in range(Lc)])
    e.write_slots(first, gens)
    e.write_pops(first, gens)
    e.mc3_setup(C, Lc, first, 2.0, 4.0, np.ones(C), seed=78, precision=L.PREC_FP64,
                scale_width=0.3)
    e.mc3_set_prior(np.tile(ts, (C, 1)), 5.0, 2)
    flags = L.MC3_GRAPH | L.MC3_SCALE | L.MC3_RY1 | L.MC3_TOPOLOGY | (L.MC3_NW if mig else 0)
    e.mc3_run(burn, flags)
    a = []
    for _ in range(nsamp):
        e.mc3_run(thin, flags)
        up0, up1, dn, tm, rt, _ = e.read_slots(first, C * Lc)
        _, tv = e.mc3_read_prior()
        ne = (2 * n - 1) * Lc
        a.append([_topo_stats(up0[c * ne:(c + 1) * ne], up1[c * ne:(c + 1) * ne],
                              dn[c * ne:(c + 1) * ne], tm[c * ne:(c + 1) * ne],
                              rt[c * Lc:(c + 1) * Lc], n, Lc) + [float(tv[c][0])]
                  + ([float(sum(len(e.read_mig_events(first + c * Lc + l)[0])
                                for l in range(Lc)))] if mig else [])
                  for c in range(C)])
    e.close()
    a = np.array(a).transpose(1, 0, 2)
    # oracle
    chains = [[g.copy() for g in base] for _ in range(C)]
    tvs = np.ascontiguousarray(np.tile(ts, (C, 1)))
    b = []
    for k in range(burn // thin + nsamp):
        oracle.mc3_bench(loci, [2.0, 2.0], 4.0, samp, model, chains, tvs, np.ones(C), 5.0,
                         seed=70000 + k, max_sweeps=thin, scale_width=0.3, ry1=True, topo=True,
                         nw=1 if mig else 0)
        if k >= burn // thin:
            row = []
            for c in range(C):
                gs = chains[c]
                row.append(_topo_stats(np.concatenate([g.up0 for g in gs]),
                                       np.concatenate([g.up1 for g in gs]),
                                       np.concatenate([g.down for g in gs]),
                                       np.concatenate([g.time for g in gs]),
                                       [g.root for g in gs], n, Lc) + [float(tvs[c][0])]
                           + ([float(sum(len(m) for g in gs for m in g.mig))] if mig else []))
            b.append(row)
    b = np.array(b).transpose(1, 0, 2)
    worst = 0.0
    for s in range(a.shape[2]):
        binary = set(np.unique(a[:, :, s])) <= {0.0, 1.0}
        pooled = np.concatenate([a[:, :, s].ravel(), b[:, :, s].ravel()])
        for qn in ((0.5,) if binary else (0.1, 0.25, 0.5, 0.75, 0.9)):
            tau = 0.5 if binary else np.quantile(pooled, qn)
            fa = (a[:, :, s] <= tau).mean(axis=1)
            fb = (b[:, :, s] <= tau).mean(axis=1)
            se = math.sqrt(fa.var(ddof=1) / len(fa) + fb.var(ddof=1) / len(fb))
            z = abs(fa.mean() - fb.mean()) / max(se, 1e-3)
            print(f"stat {s} q={qn:.2f} gpu={fa.mean():.4f} cpu={fb.mean():.4f} z={z:.2f}")
            worst = max(worst, z)
    assert worst < 4.5
