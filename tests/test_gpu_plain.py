"""The plain per-locus FP64 path (libima_plain.so: the host drives one
proposal at a time through the ABI, as IMa2p's updategenealogy calls its
likelihood and prior per (chain, locus)).  After its sweeps, every pdg and
probg it holds equals the oracle's on the genealogies it ends with (1e-10 /
1e-9), so the baseline the batched sweep is measured against does the same
work correctly."""
import numpy as np
import pytest

import oracle
from ima2p_amd import _lib as L
from ima2p_amd import immodel, plain, sim
from ima2p_amd.engine import Engine

pytestmark = pytest.mark.gpu


def test_plain_path_state_matches_oracle():
    rng = np.random.default_rng(17)
    pt, theta, ts, mig = sim.im_model_three_pop(rng)
    model = immodel.build(pt, qmax=20.0, mmax=5.0)
    samp = [4, 3, 5]
    C, Lc = 4, 3
    e = Engine(0)
    mid = e.model_add(model)
    locs, lids = [], []
    for _ in range(Lc):
        g0 = sim.simulate_genealogy(rng, samp, theta, ts, mig, pt)
        loc = sim.encode_hky(sim.simulate_hky_sequences(rng, g0, 250, rng.dirichlet([6] * 4),
                                                        4.0, 2.0))
        locs.append(loc)
        lid = e.add_locus_hky(loc)
        e.locus_set_model(lid, mid, samp)
        lids.append(lid)
    gens = [sim.simulate_genealogy(rng, samp, theta, ts, mig, pt) for _ in range(C * Lc)]
    first = e.alloc_slots([lids[l] for c in range(C) for l in range(Lc)])
    e.write_slots(first, gens)
    e.write_pops(first, gens)
    u = np.tile(rng.uniform(1.0, 3.0, Lc), C)
    betas = np.array([0.9 ** c for c in range(C)])
    r = plain.run(e, first, mid, model.npops, gens, C, Lc, u, 4.0, betas,
                  np.tile(ts, (C, 1)), nsweeps=6)
    assert r["accepts"] > 0
    off = 0
    for c in range(C):
        ws = []
        for l in range(Lc):
            g = gens[c * Lc + l].copy()
            ne = 2 * g.n - 1
            g.time = r["times"][off:off + ne].copy()
            off += ne
            ref, _ = oracle.likelihood_hky(g, locs[l], u[c * Lc + l], 4.0)
            assert abs(r["pdg"][c, l] - ref) <= 1e-10 * abs(ref)
            ws.append(oracle.treeweight(g, model, np.asarray(ts, float), samp))
        rp, _, _ = oracle.integrate_tree_prob(ws, model)
        assert abs(r["probg"][c] - rp) <= 1e-9 * abs(rp)
    assert sorted(r["beta"]) == sorted(betas)
    e.close()


def test_integrate_tree_prob_from_summed_weights():
    """ima_integrate_tree_prob on a chain's summed weight records equals
    ima_prior over the chain's slots and the oracle (1e-10)."""
    rng = np.random.default_rng(5)
    pt, theta, ts, mig = sim.im_model_three_pop(rng)
    model = immodel.build(pt, qmax=20.0, mmax=5.0)
    samp = [4, 3, 5]
    e = Engine(0)
    mid = e.model_add(model)
    lid = e.add_locus_hky(sim.encode_hky(rng.integers(0, 4, (12, 30)).astype(np.int8)))
    e.locus_set_model(lid, mid, samp)
    gens = [sim.simulate_genealogy(rng, samp, theta, ts, mig, pt) for _ in range(5)]
    first = e.alloc_slots([lid] * 5)
    e.write_slots(first, gens)
    e.write_pops(first, gens)
    tw = e.treeweight(first + np.arange(5), np.tile(ts, (5, 1)), model.npops)
    P = model.npops
    rec = np.concatenate([tw["cc"].sum(0).ravel(), tw["fc"].sum(0).ravel(),
                          tw["hcc"].sum(0).ravel(), tw["mc"].sum(0).ravel(),
                          tw["fm"].sum(0).ravel(), [tw["length"].sum()]]).astype(np.float64)
    assert rec.size == 3 * P * P + 2 * P ** 3 + 1
    import ctypes as C_
    pg = C_.c_double()
    lib = L.load()
    rc = lib.ima_integrate_tree_prob(e.h, mid, rec.ctypes.data_as(L.dp), 1, C_.byref(pg),
                                     None, None)
    assert rc == 0
    probg, _, _ = e.prior([list(first + np.arange(5))], np.asarray(ts, float).reshape(1, -1), model)
    ws = [oracle.treeweight(g, model, np.asarray(ts, float), samp) for g in gens]
    ref, _, _ = oracle.integrate_tree_prob(ws, model)
    assert abs(pg.value - ref) <= 1e-10 * abs(ref)
    assert abs(pg.value - probg[0]) <= 1e-12 * abs(ref)
    e.close()
