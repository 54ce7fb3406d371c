"""GPU parity of the L-mode kernels and the MC3 swap against the oracle.

jointp (REF_COMPAT and EXACT) and margincalc within 1e-10 relative of the
oracle; swap decisions and betas identical to the oracle's swapchains fed
with the same Philox uniforms.
"""
import numpy as np
import pytest

import oracle
from ima2p_amd import _lib as L
from ima2p_amd import sim
from tests.philox_ref import uniforms

pytestmark = pytest.mark.gpu


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("G", [1, 2, 37, 1000, 20000])
def test_jointp_two_pop(engine, G):
    rng = np.random.default_rng(G)
    tab = sim.lmode_table(rng, G)
    tid = engine.lmode_table(tab, 3, 2, 2)
    pts = np.column_stack([rng.uniform(1, 12, (40, 3)), rng.uniform(0.05, 2.0, (40, 2))])
    ex, ess_ex = engine.jointp(tid, pts, mode=L.LMODE_EXACT)
    rc, ess_rc = engine.jointp(tid, pts, mode=L.LMODE_REF_COMPAT)
    for i, x in enumerate(pts):
        r_rc, r_ess = oracle.jointp(tab, 3, 2, 2, x, ess=True)
        r_ex = oracle.jointp_exact(tab, 3, 2, 2, x)
        assert rel(rc[i], r_rc) < 1e-10, (i, rc[i], r_rc)
        assert rel(ess_rc[i], r_ess) < 1e-9
        assert rel(ex[i], r_ex) < 1e-10, (i, ex[i], r_ex)
    engine.lmode_free(tid)


def test_jointp_three_pop_models(engine):
    rng = np.random.default_rng(77)
    tab = sim.lmode_table(rng, 5000, npsize=5, nmig=6, nsplit=2)
    tid = engine.lmode_table(tab, 5, 6, 3)
    x = np.column_stack([rng.uniform(1, 12, (16, 5)), rng.uniform(0.05, 2.0, (16, 6))])
    for modeltype, pr in ((1, (0, 5)), (2, (5, 11))):
        out, _ = engine.jointp(tid, x, mode=L.LMODE_REF_COMPAT, modeltype=modeltype, prange=pr)
        for i in range(len(x)):
            ref = oracle.jointp(tab, 5, 6, 3, x[i], modeltype=modeltype, prange=pr)
            assert rel(out[i], ref) < 1e-10


def test_margincalc(engine):
    rng = np.random.default_rng(3)
    tab = sim.lmode_table(rng, 3000)
    tid = engine.lmode_table(tab, 3, 2, 2)
    xs = np.linspace(0.1, 15, 200)
    for param in range(5):
        for expo in (0, 1):
            got = engine.margincalc(tid, param, xs, logi=1, yadjust=0.25, expo=expo, mean=0.7)
            for j in range(0, 200, 17):
                ref = oracle.margincalc(tab, 3, 2, xs[j], param, 1, 0.25, bool(expo), 0.7)
                assert rel(got[j], ref) < 1e-10


@pytest.mark.parametrize("nchains,tries", [(4, 1), (16, 3), (64, 6), (256, 25)])
def test_swap_matches_oracle(engine, nchains, tries):
    rng = np.random.default_rng(nchains)
    S = rng.normal(-1000, 3, nchains)
    beta = np.array([oracle.setheat_geometric(i, nchains, 0.96, 0.9) for i in range(nchains)])
    for counter in range(5):
        seed = 1234 + nchains
        got_b, nacc, pairs = engine.swap(S, beta, tries, seed, counter * 1000)
        uni = uniforms(seed, counter * 1000, 64 * tries + 64)
        ref_b, racc, rpairs, _ = oracle.swapchains(S, beta, tries, uni)
        assert nacc == racc
        assert (pairs == rpairs).all()
        assert (got_b == ref_b).all()
        beta = got_b
