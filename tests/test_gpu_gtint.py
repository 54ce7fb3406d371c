"""ima_lmode_greater_than against the oracle's restatement of gtint.cpp
(gtpops, gtmig, qtrap, print_greater_than_tests's subsampling).  Per
genealogy both run the same closed forms or the same trapezoid points in
the same order; libm's last bit may still move qtrap's stopping level for
a genealogy, which changes its term by less than the 1e-4 relative
tolerance qtrap stops at.  Averaged over 600 genealogies that is below
2e-7 relative per flip, so the bar is 1e-5 relative.  Pairs with unequal
prior bounds, the diagonal and migration bounds <= MINPARAMVAL are -1, and
P(i > j) + P(j > i) is checked near 1 on consistent rows."""
import numpy as np
import pytest

import oracle
from ima2p_amd import sim
from ima2p_amd.engine import Engine, EngineError

pytestmark = pytest.mark.gpu

NP, NM = 3, 4


def _table(rng, G):
    tab = sim.lmode_table(rng, G, npsize=NP, nmig=NM, nloci=2)
    # log integrals of moderate size and some rows without coalescences or
    # migration events in a parameter, to reach the closed forms
    tab[:, 3 * NP + 2 * NM:4 * NP + 3 * NM] *= 0.1
    for g in range(0, G, 4):
        tab[g, 1] = tab[g, NP + 1] = 0           # population 1: nothing
        tab[g, 3 * NP + 2] = 0                   # migration rate 2: no event
    for g in range(1, G, 7):
        tab[g, 3 * NP + NM + 2] = 0              # ... and no time
    return tab


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("kind", [0, 1])
def test_greater_than_matrix_equals_the_oracle(eng, kind):
    rng = np.random.default_rng(31 + kind)
    G = 600
    tab = _table(rng, G)
    tid = eng.lmode_table(tab, NP, NM, 3)
    n = NM if kind else NP
    pmax = np.full(n, 8.0 if kind == 0 else 2.5)
    pmax[-1] = 5.0  # a parameter with another prior: no comparisons
    got = eng.greater_than(tid, kind, pmax)
    for i in range(n):
        for j in range(n):
            if i == j or pmax[i] != pmax[j]:
                assert got[i, j] == -1.0
                continue
            ref = oracle.greater_than(tab, NP, NM, kind, i, j, pmax[i])
            assert got[i, j] == pytest.approx(ref, rel=1e-5, abs=1e-12), (kind, i, j)
    eng.lmode_table_free(tid)


def test_subsampling_and_errors(eng):
    rng = np.random.default_rng(4)
    G = 41000  # more than 20000: every second genealogy of the first 40000
    tab = np.zeros((G, 4 * NP + 3 * NM + 3), np.float32)
    tab[:, 0] = rng.integers(2, 9, G)                 # closed form: population 1 empty
    tab[:, NP] = rng.gamma(3.0, 2.0, G)
    tab[:, 3 * NP + 2 * NM] = rng.normal(0.5, 0.3, G)
    tid = eng.lmode_table(tab, NP, NM, 3)
    got = eng.greater_than(tid, 0, np.full(NP, 10.0))
    ref = oracle.greater_than(tab, NP, NM, 0, 0, 1, 10.0)
    assert got[0, 1] == pytest.approx(ref, rel=1e-9)
    with pytest.raises(EngineError):
        eng.greater_than(tid, 0, np.full(NP + 1, 10.0))   # wrong parameter count
    m = eng.greater_than(tid, 1, np.array([1e-8, 1e-8, 2.0, 3.0]))
    assert (m[:2, :2] == -1).all()                        # bounds <= MINPARAMVAL
    eng.lmode_table_free(tid)
