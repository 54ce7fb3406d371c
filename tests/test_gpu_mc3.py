"""Device-resident MC3 sweep: after accepts, rejects, rescaling and swaps the
stored log-likelihood of every (chain, locus) equals the oracle's value for
the genealogy now held in its slot (1e-10 FP64, 1e-6 FP32), genealogies stay
valid with unchanged topology, and graph replay equals eager launches."""
import numpy as np
import pytest

import oracle
from ima2p_amd import _lib as L
from ima2p_amd import sim
from ima2p_amd.engine import Engine
from tests import datasets as D

pytestmark = pytest.mark.gpu


def build(engine, C, prec, seed=3):
    rng = np.random.default_rng(seed)
    pt, theta, ts, mig = sim.im_model_two_pop()
    loci, models, gens, As = [], [], [], []
    hky = []
    for l in range(3):
        g = sim.simulate_genealogy(rng, [6 + l, 5], theta, ts, mig, pt)
        loc = sim.encode_hky(sim.simulate_hky_sequences(rng, g, 300, rng.dirichlet([5] * 4), 3.0, 2.0))
        hky.append(loc)
        loci.append(engine.add_locus_hky(loc))
        models.append(("hky", loc, g))
    g_is = sim.simulate_genealogy(rng, [7, 7], theta, ts, mig, pt)
    loc_is = sim.simulate_is(rng, g_is, 3.0)
    loci.append(engine.add_locus_is(loc_is))
    models.append(("is", loc_is, g_is))
    g_sw = sim.simulate_genealogy(rng, [6, 6], theta, ts, mig, pt)
    A_sw = sim.simulate_sw(rng, g_sw, 2.0)
    loci.append(engine.add_locus_sw(g_sw.n))
    models.append(("sw", A_sw, g_sw))
    Lc = len(loci)
    first = engine.alloc_slots([loci[l] for c in range(C) for l in range(Lc)])
    for c in range(C):
        for l, (kind, loc, g) in enumerate(models):
            gg = sim.relabel_internal(g, rng) if kind != "sw" else g
            A = A_sw if kind == "sw" else np.zeros(2 * g.n - 1, np.int32)
            engine.write_slots(first + c * Lc + l, [gg], A=[A])
    rc, sl = oracle.calc_sumlogk(g_is, loc_is)
    engine.set_sumlogk(loci[3], sl)
    u = np.tile([1.5, 2.0, 1.0, 2.5, 1.2], C)
    kappa = np.tile([3.0, 2.0, 5.0, 1.0, 1.0], C)
    beta = 1.0 / (1.0 + 0.2 * np.arange(C))
    engine.mc3_setup(C, Lc, first, u, kappa, beta, seed=99, precision=prec, scale_width=0.3)
    return first, Lc, models, u, kappa, sl


def check(engine, first, C, Lc, models, u, kappa, sl, tol):
    pdg, beta, stats, sweeps = engine.mc3_read()
    up0, up1, dn, tm, rt, A = engine.read_slots(first, C * Lc)
    off = 0
    worst = 0.0
    for c in range(C):
        for l, (kind, loc, g0) in enumerate(models):
            n = g0.n
            ne = 2 * n - 1
            g = sim.Genealogy(n, up0[off:off + ne].copy(), up1[off:off + ne].copy(),
                              dn[off:off + ne].copy(), tm[off:off + ne].copy(), int(rt[c * Lc + l]),
                              np.zeros(ne, np.int32))
            g.check()
            it = c * Lc + l
            if kind == "hky":
                ref, _ = oracle.likelihood_hky(g, loc, u[it], kappa[it])
            elif kind == "is":
                ref, _, _ = oracle.likelihood_is(g, loc, u[it], sl)
            else:
                ref, _ = oracle.likelihood_sw(g, A[off:off + ne], u[it], 1.0)
            worst = max(worst, abs(pdg[c, l] - ref) / abs(ref))
            off += ne
    assert worst < tol, worst
    return stats, sweeps


@pytest.mark.parametrize("prec,tol", [(L.PREC_FP64, 1e-10), (L.PREC_FP32, 1e-6)])
def test_sweep_state_consistent(prec, tol):
    e = Engine(0)
    C = 6
    first, Lc, models, u, kappa, sl = build(e, C, prec)
    e.mc3_run(5, L.MC3_SCALE)                      # eager
    e.mc3_run(6, L.MC3_SCALE | L.MC3_GRAPH)        # graph replay
    e.mc3_run(3, L.MC3_SCALE | L.MC3_TIMED)        # timed eager
    stats, sweeps = check(e, first, C, Lc, models, u, kappa, sl, tol)
    assert sweeps == 14
    assert stats[0] == 14 * C * Lc
    assert 0 < stats[1] < stats[0]
    assert stats[2] == 7 * C
    e.close()


def test_graph_equals_eager():
    res = []
    for flags in (L.MC3_SCALE, L.MC3_SCALE | L.MC3_GRAPH):
        e = Engine(0)
        first, Lc, models, u, kappa, sl = build(e, 4, L.PREC_FP64, seed=5)
        e.mc3_run(8, flags)
        pdg, beta, stats, _ = e.mc3_read()
        res.append((pdg.copy(), beta.copy(), stats.copy()))
        e.close()
    assert (res[0][0] == res[1][0]).all()
    assert (res[0][1] == res[1][1]).all()
    assert (res[0][2] == res[1][2]).all()
