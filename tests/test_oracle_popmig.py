"""The oracle's 2NM density (popmig.cpp:9-429) pinned by the integral it
closes: for one genealogy, the density of x = theta m / 2 is

  e^(-hc - qint - mint) * integral over theta of (2/theta)^cc e^(-2 fc/theta)
      * prior_m(2x/theta) * (2x/theta)^mc e^(-fm 2x/theta) * (2/theta) dtheta

with theta in (2x/mmax, qmax) for the uniform migration prior (density 1 on
(0, mmax), as the reference leaves the 1/mmax out) and theta in (0, qmax)
with prior_m(m) = e^(-m/mean)/mean for the exponential one.  scipy's quad of
that integral checks every branch of the closed forms (both counts
positive, either count zero, both zero); the prior densities of 2NM that
prob_or_like divides by are checked to integrate to 1; marginpopmig's
divisor, sign and off-scale value, and marginalopt_popmig's bin scan and
golden section, against their definitions."""
import numpy as np
import pytest
from scipy import integrate

import oracle

NP, NM = 2, 2


def _row(cc, fc, hc, mc, fm, qint, mint, thetai=0, mi=1):
    K = 4 * NP + 3 * NM + 2 + 1
    r = np.zeros(K, np.float32)
    r[thetai] = cc
    r[NP + thetai] = fc
    r[2 * NP + thetai] = hc
    r[3 * NP + mi] = mc
    r[3 * NP + NM + mi] = fm
    r[3 * NP + 2 * NM + thetai] = qint
    r[4 * NP + 2 * NM + mi] = mint
    return r


def _quad_uniform(row, x, qmax, mmax, thetai=0, mi=1):
    r = row.astype(np.float64)
    cc, fc, hc = r[thetai], r[NP + thetai], r[2 * NP + thetai]
    mc, fm = r[3 * NP + mi], r[3 * NP + NM + mi]
    qint, mint = r[3 * NP + 2 * NM + thetai], r[4 * NP + 2 * NM + mi]

    def f(t):
        m = 2 * x / t
        return np.exp(cc * np.log(2 / t) - 2 * fc / t + mc * np.log(m) - fm * m) * (2 / t)
    lo = 2 * x / mmax
    pts = np.geomspace(lo, qmax, 12)[1:-1]
    v, _ = integrate.quad(f, lo, qmax, points=pts, limit=400, epsabs=0, epsrel=1e-11)
    return v * np.exp(-hc - qint - mint)


def _quad_expo(row, x, qmax, mean, thetai=0, mi=1):
    r = row.astype(np.float64)
    cc, fc, hc = r[thetai], r[NP + thetai], r[2 * NP + thetai]
    mc, fm = r[3 * NP + mi], r[3 * NP + NM + mi]
    qint, mint = r[3 * NP + 2 * NM + thetai], r[4 * NP + 2 * NM + mi]

    def f(t):
        m = 2 * x / t
        return (np.exp(cc * np.log(2 / t) - 2 * fc / t + mc * np.log(m) - fm * m - m / mean)
                / mean * (2 / t))
    pts = np.geomspace(qmax * 1e-4, qmax, 12)[1:-1]
    v, _ = integrate.quad(f, 0, qmax, points=pts, limit=400, epsabs=0, epsrel=1e-11)
    return v * np.exp(-hc - qint - mint)


CASES = [  # (cc, fc, hc, mc, fm): every branch of the closed forms
    (7, 9.5, 0.3, 3, 4.25),    # both counts positive
    (0, 0.0, 0.0, 4, 3.5),     # no coalescence in theta_i's population
    (0, 0.0, 0.1, 0, 2.0),     # ... and no migration event, fm > 0
    (5, 6.0, 0.0, 0, 0.0),     # no migration time at all
    (0, 0.0, 0.0, 0, 0.0),     # nothing: 2 log(qmax mmax / 2x)
    (3, 2.0, 0.0, 0, 1.5),     # cc > 0, mc = 0, fm > 0 (general branch)
]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("x", [0.05, 0.7, 3.0])
def test_uniform_closed_form_equals_quadrature(case, x):
    cc, fc, hc, mc, fm = case
    qmax, mmax = 10.0, 2.0
    if 2 * x / mmax >= qmax:
        pytest.skip("x beyond the prior box")
    row = _row(cc, fc, hc, mc, fm, -1.5, -0.5)
    tab = np.stack([row])
    got = oracle.popmig(tab, NP, NM, 0, 1, x, qmax, mmax)
    ref = _quad_uniform(row, x, qmax, mmax)
    assert got == pytest.approx(ref, rel=2e-6), (case, x)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("x", [0.05, 0.7, 3.0])
def test_exponential_closed_form_equals_quadrature(case, x):
    cc, fc, hc, mc, fm = case
    if cc == 0 and mc == 0 and fc == 0 and fm == 0:
        pass  # E1 branch of the exponential form
    qmax, mean = 10.0, 0.8
    row = _row(cc, fc, hc, mc, fm, -1.5, -0.5)
    tab = np.stack([row])
    got = oracle.popmig(tab, NP, NM, 0, 1, x, qmax, mean, expo=True)
    ref = _quad_expo(row, x, qmax, mean)
    assert got == pytest.approx(ref, rel=2e-6), (case, x)


def test_prior_densities_of_2NM_integrate_to_one():
    """prob_or_like divides by the prior density of 2NM: uniform theta and
    m give 2 log(qmax mmax / 2x) / (qmax mmax) on (0, qmax mmax / 2), an
    exponential m gives 2 E1(2x / (mean qmax)) / (qmax mean)."""
    qmax, mmax, mean = 10.0, 2.0, 0.8
    tab = np.stack([_row(0, 0.0, 0.0, 0, 0.0, 0.0, 0.0)])
    # with a flat genealogy (all sums 0, log integrals 0) the uniform
    # density divided by the prior density is qmax mmax at every x
    for x in (0.01, 1.0, 9.0):
        d = oracle.popmig(tab, NP, NM, 0, 1, x, qmax, mmax)
        r = oracle.popmig(tab, NP, NM, 0, 1, x, qmax, mmax, prob_or_like=1)
        prior = d / r
        assert prior == pytest.approx(2 * np.log(qmax * mmax / (2 * x)) / (qmax * mmax),
                                      rel=1e-12)
    f = lambda x: 2 * np.log(qmax * mmax / (2 * x)) / (qmax * mmax)
    assert integrate.quad(f, 0, qmax * mmax / 2)[0] == pytest.approx(1.0, rel=1e-9)
    # exponential prior: recover the divisor the same way and integrate it
    def prior_e(x):
        d = oracle.popmig(tab, NP, NM, 0, 1, x, qmax, mean, expo=True)
        r = oracle.popmig(tab, NP, NM, 0, 1, x, qmax, mean, expo=True, prob_or_like=1)
        return d / r
    from scipy.special import exp1
    for x in (0.01, 0.5, 4.0):
        assert prior_e(x) == pytest.approx(2 * exp1(2 * x / (mean * qmax)) / (qmax * mean),
                                           rel=1e-6)
    g = lambda x: 2 * exp1(2 * x / (mean * qmax)) / (qmax * mean)
    assert integrate.quad(g, 0, np.inf, limit=200)[0] == pytest.approx(1.0, rel=1e-9)


def _table(rng, G):
    rows = []
    for _ in range(G):
        cc, mc = rng.poisson(6), rng.poisson(2)
        fc = rng.gamma(cc + 1, 2.0) if cc or rng.random() < 0.7 else 0.0
        fm = rng.gamma(mc + 1, 1.5) if mc or rng.random() < 0.7 else 0.0
        rows.append(_row(cc, fc, rng.uniform(0, 0.5), mc, fm, rng.normal(-3, 0.5),
                         rng.normal(-1, 0.3)))
    return np.stack(rows)


def test_mean_over_genealogies_and_marginpopmig_divisor():
    rng = np.random.default_rng(3)
    tab = _table(rng, 40)
    qmax, mmax = 10.0, 2.0
    for x in (0.1, 1.3, 4.0):
        per = [oracle.popmig(tab[i:i + 1], NP, NM, 0, 1, x, qmax, mmax) for i in range(40)]
        whole = oracle.popmig(tab, NP, NM, 0, 1, x, qmax, mmax)
        assert whole == pytest.approx(np.mean(per), rel=1e-12)
        # marginpopmig: minus the sum over [first, last) / (last - first + (first == 0))
        m0 = oracle.marginpopmig(tab, NP, NM, 0, 1, x, qmax, mmax, first=0, last=40)
        assert m0 == pytest.approx(-np.sum(per) / 41, rel=1e-12)
        m1 = oracle.marginpopmig(tab, NP, NM, 0, 1, x, qmax, mmax, first=10, last=30)
        assert m1 == pytest.approx(-np.sum(per[10:30]) / 20, rel=1e-12)
    # off the search interval: OFFSCALEVAL
    assert oracle.marginpopmig(tab, NP, NM, 0, 1, qmax * mmax / 2 + 1e-6, qmax, mmax) == 1.0
    assert oracle.marginpopmig(tab, NP, NM, 0, 1, 20 * 0.8 + 1e-6, qmax, 0.8, expo=True) == 1.0
    # exponential prior: the eexp-scaled sum is the plain mean
    for x in (0.1, 1.3):
        per = [oracle.popmig(tab[i:i + 1], NP, NM, 0, 1, x, qmax, 0.8, expo=True)
               for i in range(40)]
        assert oracle.popmig(tab, NP, NM, 0, 1, x, qmax, 0.8, expo=True) == \
            pytest.approx(np.mean(per), rel=1e-12)


def test_peak_is_the_grid_maximum():
    rng = np.random.default_rng(5)
    tab = _table(rng, 30)
    qmax, mmax = 10.0, 2.0
    pk, ml = oracle.popmig_peak(tab, NP, NM, 0, 1, qmax, mmax)
    xs = np.linspace(1e-7, qmax * mmax / 2 * (1 - 1e-6), 4001)
    ys = np.array([-oracle.marginpopmig(tab, NP, NM, 0, 1, x, qmax, mmax) for x in xs])
    assert ml >= ys.max() * (1 - 1e-9)
    assert abs(pk - xs[np.argmax(ys)]) <= 2 * (xs[1] - xs[0])
    assert ml == pytest.approx(-oracle.marginpopmig(tab, NP, NM, 0, 1, pk, qmax, mmax),
                               rel=1e-15)


def test_general_branch_at_the_upper_edge_adds_one():
    """At x = qmax mmax / 2 both incomplete gammas of the general branch have
    the same argument, a == b, and the reference adds exp(0) = 1 for the
    genealogy (popmig.cpp:74-78); the special branches give log(0) there
    and are skipped."""
    rng = np.random.default_rng(7)
    tab = _table(rng, 25)
    qmax, mmax = 10.0, 2.0
    r = tab.astype(np.float64)
    cc, fc, mc, fm = r[:, 0], r[:, NP], r[:, 3 * NP + 1], r[:, 3 * NP + NM + 1]
    special = ((fc == 0) & (cc == 0) & (fm > 0)) | ((fm == 0) & (mc == 0) & (fc > 0)) | \
        ((fc == 0) & (cc == 0) & (mc == 0) & (fm == 0))
    x = qmax * mmax / 2
    assert oracle.popmig(tab, NP, NM, 0, 1, x, qmax, mmax) == \
        pytest.approx((~special).sum() / len(tab), rel=1e-15)
