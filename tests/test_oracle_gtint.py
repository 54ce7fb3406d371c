"""The oracle's P(parameter i > parameter j) (gtint.cpp:1-349) pinned by
two-dimensional quadrature of its definition.  For one genealogy with
population-size likelihoods L(t) = (2/t)^cc e^(-2 fc/t - hcc) under a
uniform prior on (0, qmax), and qintegrate = log of L's integral, the
probability is the double integral of L_i(t_i) L_j(t_j) over t_i > t_j
divided by both integrals; migration rates likewise with L(m) = m^w
e^(-fm m) on (0, mmax).  The closed-form branches (no coalescence and no
time in population j, no migration event of rate j) must match to 1e-6 (the incomplete gammas stop at 3e-7 relative);
the branches the reference integrates by trapezoid refinement (qtrap,
relative tolerance 1e-4, convergence tested from the sixth refinement) to
2e-3, and P(i > j) + P(j > i) to 1 within 2e-3.  The subsampling of
tables larger than 20000 genealogies is checked on closed-form rows."""
import numpy as np
import pytest
from scipy import integrate

import oracle

NP, NM = 2, 2


def _row(cc=(0, 0), fc=(0.0, 0.0), hcc=(0.0, 0.0), qint=(0.0, 0.0), mc=(0, 0), fm=(0.0, 0.0),
         mint=(0.0, 0.0)):
    K = 4 * NP + 3 * NM + 2 + 1
    r = np.zeros(K, np.float32)
    r[0:NP] = cc
    r[NP:2 * NP] = fc
    r[2 * NP:3 * NP] = hcc
    r[3 * NP:3 * NP + NM] = mc
    r[3 * NP + NM:3 * NP + 2 * NM] = fm
    r[3 * NP + 2 * NM:4 * NP + 2 * NM] = qint
    r[4 * NP + 2 * NM:4 * NP + 3 * NM] = mint
    return r


def _Lq(cc, fc, hcc):
    return lambda t: np.exp(cc * np.log(2 / t) - 2 * fc / t - hcc) if t > 0 else 0.0


def _Lm(w, fm):
    return lambda m: np.exp(w * np.log(m) - fm * m) if m > 0 else (1.0 if w == 0 else 0.0)


def _p_greater(Li, Lj, hi):
    Zi = integrate.quad(Li, 0, hi, limit=200, epsrel=1e-12)[0]
    Zj = integrate.quad(Lj, 0, hi, limit=200, epsrel=1e-12)[0]
    num = integrate.dblquad(lambda ti, tj: Li(ti) * Lj(tj), 0, hi, lambda tj: tj,
                            lambda tj: hi, epsrel=1e-10)[0]
    return num / (Zi * Zj), np.log(Zi), np.log(Zj)


POP_CASES = [  # (cc_i, fc_i, cc_j, fc_j)
    (6, 8.0, 4, 3.0),     # qtrap
    (3, 2.5, 0, 1.5),     # ccj = 0, fcj > 0: qtrap
    (5, 9.0, 0, 0.0),     # closed form, cci >= 2
    (1, 2.0, 0, 0.0),     # closed form, cci = 1
    (0, 1.0, 0, 0.0),     # closed form, cci = 0
    (0, 0.0, 0, 0.0),     # closed form, nothing
]


@pytest.mark.parametrize("case", POP_CASES)
def test_population_sizes_equal_the_double_integral(case):
    ci, fi, cj, fj = case
    qmax, hi_, hj_ = 12.0, 0.2, 0.1
    f32 = lambda v: float(np.float32(v))  # the table holds floats
    fi, fj, hi_, hj_ = f32(fi), f32(fj), f32(hi_), f32(hj_)
    p, zi, zj = _p_greater(_Lq(ci, fi, hi_), _Lq(cj, fj, hj_), qmax)
    row = _row(cc=(ci, cj), fc=(fi, fj), hcc=(hi_, hj_), qint=(zi, zj))
    # the probability with the float-rounded log integrals the row holds
    p *= np.exp(zi + zj - float(row[3 * NP + 2 * NM]) - float(row[3 * NP + 2 * NM + 1]))
    got = oracle.greater_than_one(row, NP, NM, 0, 0, 1, qmax)
    closed = cj == 0 and fj == 0
    assert got == pytest.approx(min(p, 1.0), rel=1e-6 if closed else 2e-3, abs=1e-12), case
    # and the reverse comparison sums to one (the reference's own check)
    p2, _, _ = _p_greater(_Lq(cj, fj, hj_), _Lq(ci, fi, hi_), qmax)
    p2 *= np.exp(zi + zj - float(row[3 * NP + 2 * NM]) - float(row[3 * NP + 2 * NM + 1]))
    rev = oracle.greater_than_one(row, NP, NM, 0, 1, 0, qmax)
    assert rev == pytest.approx(min(p2, 1.0), rel=2e-3, abs=1e-12)
    assert got + rev == pytest.approx(1.0, abs=2e-3)


MIG_CASES = [  # (w_i, fm_i, w_j, fm_j)
    (3, 4.0, 2, 1.5),     # qtrap
    (4, 3.0, 0, 2.0),     # wj = 0, fmj > 0, wi > 0
    (0, 1.2, 0, 2.5),     # wj = 0, fmj > 0, wi = 0, fmi > 0
    (0, 0.0, 0, 2.5),     # ... fmi = 0
    (2, 1.0, 0, 0.0),     # wj = 0, fmj = 0, wi > 0
    (0, 0.7, 0, 0.0),     # ... wi = 0, fmi > 0
    (0, 0.0, 0, 0.0),     # ... nothing
]


@pytest.mark.parametrize("case", MIG_CASES)
def test_migration_rates_equal_the_double_integral(case):
    wi, fi, wj, fj = case
    mmax = 3.0
    fi, fj = float(np.float32(fi)), float(np.float32(fj))
    p, zi, zj = _p_greater(_Lm(wi, fi), _Lm(wj, fj), mmax)
    row = _row(mc=(wi, wj), fm=(fi, fj), mint=(zi, zj))
    p *= np.exp(zi + zj - float(row[4 * NP + 2 * NM]) - float(row[4 * NP + 2 * NM + 1]))
    got = oracle.greater_than_one(row, NP, NM, 1, 0, 1, mmax)
    closed = wj == 0
    assert got == pytest.approx(min(p, 1.0), rel=1e-6 if closed else 2e-3, abs=1e-12), case


def test_subsampling_of_large_tables():
    """More than 20000 genealogies: every (G // 20000)-th of the first
    (G // 20000) * 20000, divided by 20000."""
    rng = np.random.default_rng(2)
    G = 45001
    rows = np.zeros((G, 4 * NP + 3 * NM + 3), np.float32)
    rows[:, 0] = rng.integers(2, 9, G)               # cc_i (closed form: ccj = fcj = 0)
    rows[:, NP] = rng.gamma(3.0, 2.0, G)             # fc_i
    rows[:, 3 * NP + 2 * NM] = rng.normal(0.5, 0.3, G)  # qint_i
    qmax = 10.0
    per = np.array([oracle.greater_than_one(rows[g], NP, NM, 0, 0, 1, qmax)
                    for g in range(0, 40000, 2)])
    assert oracle.greater_than(rows, NP, NM, 0, 0, 1, qmax) == \
        pytest.approx(per.sum() / 20000, rel=1e-13)
    small = rows[:1000]
    per = np.array([oracle.greater_than_one(r, NP, NM, 0, 0, 1, qmax) for r in small])
    assert oracle.greater_than(small, NP, NM, 0, 0, 1, qmax) == pytest.approx(per.mean(),
                                                                            rel=1e-13)
