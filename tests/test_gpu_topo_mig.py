"""The reference's genealogy move for models with migration on the device
(IMA_MC3_TOPOLOGY with migration parameters; slider and addmigration,
update_gtree.cpp:256-372, 589-665, with mwork_single_edge, mwork_two_edges,
simmpath and getmprob, update_gtree_common.cpp:330-1583).

* Stationary law: chains started from exact draws of the joint prior of
  genealogy and split times (parameters from their uniform priors, then the
  structured coalescent), flat likelihood, split times held: after the
  device's moves the migration counts, root times and a topology statistic
  (gene copies 0 and 1 a cherry) still follow that prior (KS / binomial
  against fresh draws), for two and three populations.  The oracle's
  restatement passes the same test (test_oracle_nw.py), and a dropped
  Hastings term fails it there.
* State after sweeps with HKY data and the prior (moves, RY1 or NW time
  updates, swaps): every stored likelihood equals the oracle's on the
  genealogy read back (topology, times, populations, migrations), and every
  chain's probg equals integrate_tree_prob over its genealogies.
"""
import math

import numpy as np
import pytest
from scipy import stats

import oracle
from ima2p_amd import _lib as L
from ima2p_amd import immodel, sim
from ima2p_amd.engine import Engine
from tests.test_oracle_nw import _cherry, _prior_draw

pytestmark = pytest.mark.gpu


def _read_gene(e, slot, g, up0, up1, dn, tm, rt, off):
    ne = 2 * g.n - 1
    h = g.copy()
    h.up0 = up0[off:off + ne].copy()
    h.up1 = up1[off:off + ne].copy()
    h.down = dn[off:off + ne].copy()
    h.time = tm[off:off + ne].copy()
    h.root = int(rt)
    h.pop = e.read_pops(slot, 1)
    me, mt, mo = e.read_mig_events(slot)
    h.mig = [[] for _ in range(ne)]
    for x in range(len(me)):
        h.mig[int(me[x])].append((float(mt[x]), int(mo[x])))
    return h


@pytest.mark.parametrize("P", [2, 3])
def test_device_move_keeps_the_joint_prior(P):
    Q, TM = 4.0, 3.0
    M = 1.5 if P == 2 else 3.0
    pt = sim.PopTree.two_pop() if P == 2 else sim.PopTree.three_pop_01_2()
    model = immodel.build(pt, qmax=Q, mmax=M)
    samp = [2, 2] if P == 2 else [3, 3, 2]
    n = sum(samp)
    loc = sim.HKYLocus(n, np.zeros((n, 1), np.int8), np.ones(1, np.int32), 1,
                       np.array([0.25] * 4))
    rng = np.random.default_rng(80 + P)
    C = 3000
    gens, tvs = [], []
    for _ in range(C):
        g, t = _prior_draw(rng, pt, model, samp, Q, M, TM)
        gens.append(g)
        tvs.append(t)
    e = Engine(0)
    mid = e.model_add(model)
    lid = e.add_locus_hky(loc)
    e.locus_set_model(lid, mid, samp)
    first = e.alloc_slots([lid] * C)
    e.write_slots(first, gens)
    e.write_pops(first, gens)
    e.mc3_setup(C, 1, first, 1e-9, 1.0, np.ones(C), seed=90 + P, precision=L.PREC_FP64)
    e.mc3_set_prior(np.array(tvs), TM, P)
    e.mc3_run(150, L.MC3_GRAPH | L.MC3_TOPOLOGY)
    _, _, stats_, sweeps = e.mc3_read()
    assert sweeps == 150 and 0.2 * stats_[0] < stats_[1] < 0.9 * stats_[0]
    up0, up1, dn, tm, rt, _ = e.read_slots(first, C)
    ne = 2 * n - 1
    got = [_read_gene(e, first + i, gens[i], up0, up1, dn, tm, rt[i], i * ne) for i in range(C)]
    e.close()
    fresh = [_prior_draw(rng, pt, model, samp, Q, M, TM)[0] for _ in range(C)]
    nm = lambda xs: [sum(len(m) for m in g.mig) for g in xs]
    rtm = lambda xs: [g.time[g.up0[g.root]] for g in xs]
    assert stats.ks_2samp(nm(got), nm(fresh)).pvalue > 1e-3
    assert stats.ks_2samp(rtm(got), rtm(fresh)).pvalue > 1e-3
    a, b = np.mean([_cherry(g) for g in got]), np.mean([_cherry(g) for g in fresh])
    assert abs(a - b) < 4.5 * math.sqrt(a * (1 - a) / C + b * (1 - b) / C)
    for i in range(300):
        oracle.treeweight(got[i], model, tvs[i], samp)  # labels fit the split times


@pytest.mark.parametrize("nw", [False, True], ids=["ry1", "ry1_nw"])
def test_state_matches_oracle_after_moves_with_migration(nw):
    rng = np.random.default_rng(23)
    pt, theta, ts, mig = sim.im_model_three_pop(rng)
    model = immodel.build(pt, qmax=20.0, mmax=5.0)
    samp = [5, 4, 6]
    C, Lc = 6, 4
    e = Engine(0)
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
    e.mc3_setup(C, Lc, first, np.tile(u, C), 4.0, np.array([0.97 ** c for c in range(C)]),
                seed=29, precision=L.PREC_FP64, scale_width=0.3)
    e.mc3_set_prior(np.tile(ts, (C, 1)), 10.0, model.npops)
    flags = (L.MC3_GRAPH | L.MC3_SCALE | L.MC3_RY1 | L.MC3_TOPOLOGY
             | (L.MC3_NW if nw else 0))
    e.mc3_run(60, flags)
    pdg, beta, st, sweeps = e.mc3_read()
    probg, tv = e.mc3_read_prior()
    assert sweeps == 60 and st[1] > 0
    up0, up1, dn, tm, rt, _ = e.read_slots(first, C * Lc)
    off, changed = 0, 0
    for c in range(C):
        ws = []
        for l in range(Lc):
            i = c * Lc + l
            g = gens[i]
            h = _read_gene(e, first + i, g, up0, up1, dn, tm, rt[i], off)
            off += 2 * g.n - 1
            changed += int(not np.array_equal(h.up0, g.up0))
            h.check()
            for lst in h.mig:
                assert [t for t, _ in lst] == sorted(t for t, _ in lst)
            ref, _ = oracle.likelihood_hky(h, locs[l], u[l], 4.0)
            assert abs(pdg[c, l] - ref) <= 1e-9 * abs(ref), (c, l, pdg[c, l], ref)
            ws.append(oracle.treeweight(h, model, tv[c], samp))
        rp, _, _ = oracle.integrate_tree_prob(ws, model)
        assert abs(probg[c] - rp) <= 1e-9 * abs(rp), (c, probg[c], rp)
    assert changed > 0
    e.close()
