"""Device mutation-rate rounds (IMA_MC3_UPDATE_U): changeu
(update_mc_params.cpp:24-382) for every chain after every fifth sweep, and
changekappa (:387-429) for a lone HKY locus.

* The grouped schedule (steps touching disjoint loci scored together)
  equals the sequential loop bit for bit: IMA_MC3_U_SERIAL launches one
  step at a time in the loop's order.  Graph replay equals eager launches.
  Both on HKY, IS and SW loci together, and with the genealogy prior.
* After rounds every stored log-likelihood equals the oracle's at the
  chain's current (u, kappa) for the genealogy in its slot (1e-10 FP64,
  1e-6 FP32); each chain's product of rates is unchanged (u_j d, u_k / d);
  kappa moves only for HKY loci and stays in [0, 100].
* A lone HKY locus moves kappa only.
"""
import numpy as np
import pytest

import oracle
from ima2p_amd import _lib as L
from ima2p_amd import sim
from ima2p_amd.engine import Engine
from tests.test_gpu_mc3 import build, check
from tests.test_gpu_mc3_prior import _three_pop_chains

pytestmark = pytest.mark.gpu

F = L.MC3_GRAPH | L.MC3_SCALE | L.MC3_RY1 | L.MC3_UPDATE_U


def _state(e, first, n):
    pdg, beta, stats, sweeps = e.mc3_read()
    u, k = e.mc3_read_rates()
    tm = e.read_slots(first, n)[3]
    return pdg, beta, stats, u, k, tm


def test_grouped_equals_serial_and_graph_equals_eager_mixed_loci():
    C, runs = 5, []
    for flags in (F, F | L.MC3_U_SERIAL, F & ~L.MC3_GRAPH):
        e = Engine(0)
        first, Lc, models, u, kappa, sl = build(e, C, L.PREC_FP64)
        e.mc3_run(10, flags)  # a round after the last sweep of each run
        e.mc3_run(10, flags)
        runs.append(_state(e, first, C * Lc))
        e.close()
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            np.testing.assert_array_equal(a, b)
    pdg, beta, stats, un, kn, _ = runs[0]
    # sweeps 4, 9, 14, 19 end with a round of 5 steps per chain
    assert stats[6] == 4 * C * 5 and 0 < stats[7] < stats[6]
    assert np.any(un != u.reshape(C, Lc))


def test_grouped_equals_serial_with_the_prior():
    runs = []
    for flags in (F, F | L.MC3_U_SERIAL):
        e, *_ = _three_pop_chains(8, 12, 17, L.PREC_FP64)
        e.mc3_run(30, flags)
        pdg, beta, stats, _ = e.mc3_read()
        probg, tv = e.mc3_read_prior()
        u, k = e.mc3_read_rates()
        runs.append((pdg, beta, stats, probg, tv, u, k))
        e.close()
    for a, b in zip(*runs):
        np.testing.assert_array_equal(a, b)
    assert runs[0][2][6] == 6 * 8 * 12


@pytest.mark.parametrize("prec,tol", [(L.PREC_FP64, 1e-10), (L.PREC_FP32, 1e-6)])
def test_state_consistent_after_rate_rounds(prec, tol):
    C = 6
    e = Engine(0)
    first, Lc, models, u, kappa, sl = build(e, C, prec)
    e.mc3_run(40, F)
    un, kn = e.mc3_read_rates()
    stats, _ = check(e, first, C, Lc, models, un.ravel(), kn.ravel(), sl, tol)
    assert stats[7] > 0
    np.testing.assert_allclose(np.log(un).sum(axis=1), np.log(u.reshape(C, Lc)).sum(axis=1),
                               rtol=0, atol=1e-12)
    k0 = kappa.reshape(C, Lc)
    np.testing.assert_array_equal(kn[:, 3:], k0[:, 3:])  # IS and SW loci have no kappa
    assert np.any(kn[:, :3] != k0[:, :3]) and kn.min() >= 0 and kn.max() <= 100
    e.close()


def test_lone_hky_locus_moves_kappa_only():
    rng = np.random.default_rng(12)
    pt, theta, ts, mig = sim.im_model_two_pop()
    g = sim.simulate_genealogy(rng, [5, 4], theta, ts, mig, pt)
    loc = sim.encode_hky(sim.simulate_hky_sequences(rng, g, 400, rng.dirichlet([5] * 4), 6.0, 3.0))
    C = 4
    e = Engine(0)
    lid = e.add_locus_hky(loc)
    first = e.alloc_slots([lid] * C)
    e.write_slots(first, [sim.relabel_internal(g, rng) for _ in range(C)])
    e.mc3_setup(C, 1, first, np.full(C, 3.0), np.full(C, 2.0), np.ones(C), seed=8,
                precision=L.PREC_FP64, scale_width=0.3)
    e.mc3_run(50, L.MC3_GRAPH | L.MC3_SCALE | L.MC3_UPDATE_U)
    pdg, _, stats, _ = e.mc3_read()
    un, kn = e.mc3_read_rates()
    assert stats[6] == 10 * C and 0 < stats[7] <= stats[6]
    np.testing.assert_array_equal(un, 3.0)
    assert np.all(kn != 2.0)
    up0, up1, dn, tm, rt, _ = e.read_slots(first, C)
    ne = 2 * g.n - 1
    for c in range(C):
        h = sim.Genealogy(g.n, up0[c * ne:(c + 1) * ne].copy(), up1[c * ne:(c + 1) * ne].copy(),
                          dn[c * ne:(c + 1) * ne].copy(), tm[c * ne:(c + 1) * ne].copy(),
                          int(rt[c]), np.zeros(ne, np.int32))
        ref, _ = oracle.likelihood_hky(h, loc, 3.0, kn[c, 0])
        assert abs(pdg[c, 0] - ref) <= 1e-10 * abs(ref)
    e.close()
