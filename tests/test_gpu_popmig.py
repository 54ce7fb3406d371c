"""ima_lmode_popmig / ima_lmode_popmig_peak against the oracle's restatement
of popmig.cpp (calc_popmig, calc_pop_expomig, marginpopmig,
marginpop_expomig, marginalopt_popmig): every point of both priors, both
kinds, with and without the prior division, on synthetic tables and on
rows built to hit each closed-form branch; relative 1e-10 for the uniform
prior (FP64 sums in another order) and 1e-9 for the exponential prior,
whose reference sum goes through eexp: its degree-10 series of e^u on
[0, ln 2) is accurate to (ln 2)^11 / 11! = 4.4e-10 relative, while the
device sums the terms exactly (log-sum-exp).  The peak: location within 1e-6 relative of the oracle's (the
golden section may stop on either side of a near tie), density within
1e-9."""
import numpy as np
import pytest

import oracle
from ima2p_amd import _lib as L
from ima2p_amd import sim
from ima2p_amd.engine import Engine, EngineError

pytestmark = pytest.mark.gpu

NP, NM = 3, 2


def _branch_rows(rng, G):
    """sim rows with some columns zeroed to reach every branch."""
    tab = sim.lmode_table(rng, G, npsize=NP, nmig=NM)
    for i in range(G):
        k = i % 5
        if k == 1:   # no coalescence in population 0
            tab[i, 0] = 0
            tab[i, NP] = 0
        elif k == 2:  # no migration time for rate 1
            tab[i, 3 * NP + 1] = 0
            tab[i, 3 * NP + NM + 1] = 0
        elif k == 3:  # nothing
            tab[i, 0] = tab[i, NP] = 0
            tab[i, 3 * NP + 1] = tab[i, 3 * NP + NM + 1] = 0
    # log integrals of moderate size, so that the densities are O(1)
    tab[:, 3 * NP + 2 * NM:4 * NP + 3 * NM] *= 0.05
    return tab


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("expo", [False, True])
def test_popmig_calc_and_margin_equal_the_oracle(eng, expo):
    rng = np.random.default_rng(11 + expo)
    tab = _branch_rows(rng, 700)
    tid = eng.lmode_table(tab, NP, NM, 3)
    qmax, mpr = 15.0, (0.8 if expo else 3.0)
    upper = 20 * mpr if expo else qmax * mpr / 2
    x = np.concatenate([np.geomspace(1e-4, upper * 0.999, 61), [upper * 0.5]])
    for ti, mi in ((0, 1), (2, 0)):
        for pol in (0, 1):
            got = eng.popmig(tid, ti, mi, x, qmax, mpr, expo=expo, prob_or_like=pol)
            ref = np.array([oracle.popmig(tab, NP, NM, ti, mi, v, qmax, mpr, expo, pol)
                            for v in x])
            np.testing.assert_allclose(got, ref, rtol=1e-9 if expo else 1e-10, atol=0)
        for first, last in ((0, 700), (100, 333)):
            xm = np.concatenate([x, [-1.0, upper * 1.01]])
            got = eng.popmig(tid, ti, mi, xm, qmax, mpr, expo=expo, kind=L.POPMIG_MARGIN,
                             first=first, last=last)
            ref = np.array([oracle.marginpopmig(tab, NP, NM, ti, mi, v, qmax, mpr, expo, first,
                                                last) for v in xm])
            np.testing.assert_allclose(got, ref, rtol=1e-9 if expo else 1e-10, atol=0)
            assert got[-1] == 1.0 and got[-2] == 1.0
    eng.lmode_table_free(tid)


@pytest.mark.parametrize("expo", [False, True])
def test_popmig_peak_equals_the_oracle(eng, expo):
    rng = np.random.default_rng(21 + expo)
    tab = _branch_rows(rng, 400)
    tid = eng.lmode_table(tab, NP, NM, 3)
    qmax, mpr = 15.0, (0.8 if expo else 3.0)
    for ti, mi, first, last in ((0, 1, 0, 400), (1, 0, 50, 300)):
        pk, ml, ev = eng.popmig_peak(tid, ti, mi, qmax, mpr, expo=expo, first=first, last=last)
        rpk, rml = oracle.popmig_peak(tab, NP, NM, ti, mi, qmax, mpr, expo, first, last)
        assert pk == pytest.approx(rpk, rel=1e-6), (ti, mi)
        assert ml == pytest.approx(rml, rel=1e-9), (ti, mi)
        assert ev >= 100
    eng.lmode_table_free(tid)


def test_popmig_errors(eng):
    rng = np.random.default_rng(3)
    tab = sim.lmode_table(rng, 50, npsize=NP, nmig=NM)
    tid = eng.lmode_table(tab, NP, NM, 3)
    with pytest.raises(EngineError):
        eng.popmig(tid, NP, 0, [1.0], 10.0, 2.0)        # no such population size
    with pytest.raises(EngineError):
        eng.popmig(tid, 0, NM, [1.0], 10.0, 2.0)        # no such migration rate
    with pytest.raises(EngineError):
        eng.popmig(tid, 0, 0, [0.0], 10.0, 2.0)         # 2NM must be > 0
    with pytest.raises(EngineError):
        eng.popmig(tid, 0, 0, [1.0], 10.0, 2.0, kind=L.POPMIG_MARGIN, first=10, last=60)
    eng.lmode_table_free(tid)
