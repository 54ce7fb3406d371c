"""L-mode joint peaks by differential evolution on the batched jointp
(ima_lmode_joint_peak; findjointpeaks / modelloop, jointfind.cpp:596-812,
1089-1183), against the oracle's jointp optimised independently:

* full two-population model: the DE peak's log P is at least as high as
  every bounded quasi-Newton search (scipy L-BFGS-B on the oracle's
  jointp, several starts, among them the DE peak itself) within 1e-6 --
  the DE found the maximum;
* nested models through mapvals (equal population sizes; both migration
  rates fixed at MINPARAMVAL): the same against searches over the nested
  model's own values, and 2LLR = 2 (log P full - log P nested) >= 0;
* three populations: the population-size and migration surfaces
  (modeltype 1 and 2) likewise.
"""
import numpy as np
import pytest
from scipy.optimize import minimize

import oracle
from ima2p_amd import _lib as L
from ima2p_amd import sim
from ima2p_amd.engine import Engine

pytestmark = pytest.mark.gpu

MINPARAMVAL = 1e-7


def _mapvals(f, xmap):
    return np.array([-m if m < 0 else f[int(m)] for m in xmap])


def _scipy_best(fun, lo, hi, starts):
    best = np.inf
    for s in starts:
        r = minimize(fun, np.clip(s, lo, hi), method="L-BFGS-B", bounds=list(zip(lo, hi)),
                     options={"ftol": 1e-13, "gtol": 1e-10, "maxiter": 2000})
        best = min(best, r.fun, fun(np.clip(s, lo, hi)))
    return best


def _check(e, tid, tab, npsize, nmig, npops, modeltype, xmap, lo, hi, seed):
    np_ = npsize + nmig
    r0, r1 = {0: (0, np_), 1: (0, npsize), 2: (npsize, np_)}[modeltype]
    peak, logp, ess, ev = e.joint_peak(tid, lo, hi, modeltype=modeltype, xmap=xmap, seed=seed,
                                       mode=L.LMODE_EXACT)
    assert ev > 0 and ess > 0
    assert np.all(peak >= np.asarray(lo) - 1e-12) and np.all(peak <= np.asarray(hi) + 1e-12)
    ref = oracle.jointp_exact(tab, npsize, nmig, npops, peak, modeltype=modeltype,
                              prange=(r0, r1))
    assert abs(-logp - ref) <= 1e-10 * abs(ref)
    # independent search over the model's own varying values
    if xmap is None:
        idx = list(range(r0, r1))
        def full_of(f):
            x = peak.copy()
            x[r0:r1] = f
            return x
    else:
        nv = int(max(m for i, m in enumerate(xmap) if r0 <= i < r1 and m >= 0)) + 1
        idx = list(range(r0, nv))
        def full_of(f):
            v = peak.copy()
            v[r0:nv] = f
            return _mapvals(v, xmap)
    vlo = np.array([lo[i] for i in idx])
    vhi = np.array([hi[i] for i in idx])
    fun = lambda f: oracle.jointp_exact(tab, npsize, nmig, npops, full_of(f), modeltype=modeltype,
                                        prange=(r0, r1))
    rng = np.random.default_rng(seed)
    starts = [peak[idx] if xmap is None else None] + \
        [rng.uniform(vlo, vhi) for _ in range(4)]
    if xmap is not None:
        # the DE peak's varying values: the first full position mapping to each
        starts[0] = np.array([peak[[i for i, m in enumerate(xmap) if m == k][0]] for k in idx])
    best = _scipy_best(fun, vlo, vhi, starts)
    assert -logp >= best - 1e-9 and -logp <= best + 1e-6, (logp, -best)
    return logp


def test_two_population_full_and_nested():
    rng = np.random.default_rng(12)
    tab = sim.lmode_table(rng, 3000)
    e = Engine(0)
    tid = e.lmode_table(tab, 3, 2, 2)
    lo = [MINPARAMVAL] * 5
    hi = [20.0, 20.0, 20.0, 5.0, 5.0]
    full = _check(e, tid, tab, 3, 2, 2, 0, None, lo, hi, 1)
    for xmap in ([0, 0, 0, 1, 2], [0, 1, 2, -MINPARAMVAL, -MINPARAMVAL]):
        nested = _check(e, tid, tab, 3, 2, 2, 0, xmap, lo, hi, 2)
        assert 2 * (full - nested) >= -1e-6
    e.close()


def test_three_population_surfaces():
    rng = np.random.default_rng(13)
    tab = sim.lmode_table(rng, 2000, npsize=5, nmig=6, nsplit=2)
    e = Engine(0)
    tid = e.lmode_table(tab, 5, 6, 3)
    lo = [MINPARAMVAL] * 11
    hi = [20.0] * 5 + [5.0] * 6
    p_full = _check(e, tid, tab, 5, 6, 3, 1, None, lo, hi, 3)
    p_eq = _check(e, tid, tab, 5, 6, 3, 1, [0, 0, 1, 1, 2, 5, 6, 7, 8, 9, 10], lo, hi, 4)
    assert 2 * (p_full - p_eq) >= -1e-6
    _check(e, tid, tab, 5, 6, 3, 2, None, lo, hi, 5)
    e.close()


def test_marginal_peaks_and_intervals():
    """Marginal peaks (ima_lmode_marginal_peak; marginalopt / margin95,
    surface_call_functions.cpp:175-290) against scipy on the oracle's
    margincalc: the peak's log density within 1e-9 of the best bounded
    search, the 95% limits where the oracle's density is 1.92 below the
    peak (brentq, 1e-6 relative), and the no-migration statistic."""
    from scipy.optimize import brentq, minimize_scalar
    rng = np.random.default_rng(21)
    tab = sim.lmode_table(rng, 4000)
    e = Engine(0)
    tid = e.lmode_table(tab, 3, 2, 2)
    for param, hi in ((0, 20.0), (2, 20.0), (3, 5.0), (4, 5.0)):
        peak, logd, lo95, hi95, llr = e.marginal_peak(tid, param, MINPARAMVAL, hi)
        f = lambda x: oracle.margincalc(tab, 3, 2, x, param)
        assert abs(logd - f(peak)) <= 1e-10 * max(1.0, abs(logd))
        xs = np.linspace(MINPARAMVAL, hi, 4001)
        k = int(np.argmax([f(x) for x in xs]))
        r = minimize_scalar(lambda x: -f(x), bounds=(xs[max(k - 1, 0)], xs[min(k + 1, 4000)]),
                            method="bounded", options={"xatol": 1e-10})
        best = max(-r.fun, f(xs[k]))
        assert logd >= best - 1e-9, (param, logd, best)
        target = logd - 1.92
        for lim, a, b in ((lo95, MINPARAMVAL, peak), (hi95, peak, hi)):
            if (f(a) - target) * (f(b) - target) > 0:
                assert lim == -1.0
            else:
                ref = brentq(lambda x: f(x) - target, a, b, xtol=1e-12)
                assert abs(lim - ref) <= 1e-6 * max(1.0, ref), (param, lim, ref)
        if param >= 3:
            assert abs(llr - 2 * (logd - f(MINPARAMVAL))) <= 1e-9 * max(1.0, abs(llr))
        else:
            assert llr == 0.0
    e.close()
