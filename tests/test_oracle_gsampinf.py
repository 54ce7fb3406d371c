"""Pins the oracle's savegsampinf row (ginfo.cpp:319-377), the checker of
ima_mc3_record:

* the layout is ginfo.cpp:288-316's -- cc | fc | hcc (per population-size
  parameter) | mc | fm (per migration parameter) | qintegrate | mintegrate
  | pdg | probg | t -- with the columns L mode reads (sim.lmode_table's
  layout, which the L-mode kernels and the oracle's jointp share);
* counts are the integer sums of the parameter's cc / mc entries (a
  three-population ancestral population sums two periods);
* fc, hcc and fm are accumulated in float, entry by entry: for entries
  1e8 and then 1.0 four times the float sum stays 1e8, where summing in
  double and rounding once gives 100000004;
* a row of sweep weights computed by treeweight / integrate_tree_prob
  gives jointp at any point the same value as the formula of jointfind.cpp
  971-994 evaluated on the row.
"""
import math

import numpy as np

import oracle
from ima2p_amd import immodel, sim


def _three_pop():
    pt = sim.PopTree.three_pop_01_2()
    return immodel.build(pt, qmax=20.0, mmax=5.0)


def test_layout_and_counts():
    model = _three_pop()
    P = model.npops
    R = 3 * P * P + 2 * P ** 3 + 1
    rng = np.random.default_rng(3)
    w = rng.uniform(0.1, 9.0, R)
    w[:P * P] = rng.integers(0, 20, P * P)
    w[3 * P * P:3 * P * P + P ** 3] = rng.integers(0, 7, P ** 3)
    nps, nm = model.npsize, model.nmig
    integ = rng.normal(-10, 2, nps + nm)
    row = oracle.gsampinf_row(w, integ, [-3.5, -4.25], -7.0, [0.7, 1.9], model)
    assert row.dtype == np.float32 and len(row) == 4 * nps + 3 * nm + 2 + (P - 1)
    cc = w[:P * P].reshape(P, P)
    for i, wp in enumerate(model.theta_wp):
        assert row[i] == sum(int(cc[k, r]) for k, r in wp)
    mc = w[3 * P * P:3 * P * P + P ** 3].reshape(P, P, P)
    for i, wp in enumerate(model.mig_wp):
        assert row[3 * nps + i] == sum(int(mc[k, a, b]) for k, a, b in wp)
    assert any(len(wp) > 1 for wp in model.theta_wp)  # a two-period parameter
    np.testing.assert_array_equal(row[3 * nps + 2 * nm:4 * nps + 2 * nm], np.float32(integ[:nps]))
    np.testing.assert_array_equal(row[4 * nps + 2 * nm:4 * nps + 3 * nm], np.float32(integ[nps:]))
    assert row[4 * nps + 3 * nm] == np.float32(-7.75)
    assert row[4 * nps + 3 * nm + 1] == np.float32(-7.0)
    np.testing.assert_array_equal(row[-2:], np.float32([0.7, 1.9]))


def test_float_accumulation():
    """Population 2 of ((0,1),2) lives in two periods: its parameter sums two
    fc entries in float, each rounded to float first.  For the entries
    8.574042765875694 and 0.33585575305464355 that gives 8.909898, where
    summing in double and rounding once gives 8.909899."""
    model = _three_pop()
    P = model.npops
    R = 3 * P * P + 2 * P ** 3 + 1
    two = [j for j, wp in enumerate(model.theta_wp) if len(wp) >= 2]
    assert two
    j = two[0]
    (k0, r0), (k1, r1) = model.theta_wp[j][:2]
    w = np.zeros(R)
    a, b = 8.574042765875694, 0.33585575305464355
    w[P * P + k0 * P + r0] = a
    w[P * P + k1 * P + r1] = b
    w[2 * P * P + k0 * P + r0] = b
    w[2 * P * P + k1 * P + r1] = a
    row = oracle.gsampinf_row(w, np.zeros(model.npsize + model.nmig), [0.0], 0.0, [1.0, 2.0], model)
    assert row[model.npsize + j] == np.float32(8.909898)
    assert row[2 * model.npsize + j] == np.float32(8.909898)
    assert np.float32(a + b) == np.float32(8.909899)  # the double-then-round sum differs


def test_row_of_real_weights_feeds_jointp():
    rng = np.random.default_rng(8)
    pt, theta, ts, mig = sim.im_model_two_pop(t=0.9, m12=0.4, m21=0.7)
    model = immodel.build(pt, qmax=20.0, mmax=5.0)
    samp = [5, 4]
    rows = []
    for _ in range(40):
        gs = [sim.simulate_genealogy(rng, samp, theta, ts, mig, pt) for _ in range(3)]
        ws = [oracle.treeweight(g, model, np.asarray(ts, float), samp) for g in gs]
        probg, q, m = oracle.integrate_tree_prob(ws, model)
        P = model.npops
        wall = np.concatenate([sum(w["cc"] for w in ws).ravel(), sum(w["fc"] for w in ws).ravel(),
                               sum(w["hcc"] for w in ws).ravel(), sum(w["mc"] for w in ws).ravel(),
                               sum(w["fm"] for w in ws).ravel(), [sum(w["length"] for w in ws)]])
        rows.append(oracle.gsampinf_row(wall, list(q) + list(m), [-10.0, -11.0, -12.0], probg,
                                        np.asarray(ts, float), model))
    tab = np.array(rows, np.float32)
    x = np.array([4.0, 6.0, 3.0, 0.5, 0.8])
    L = tab.astype(np.float64)
    # jointfind.cpp:971-994 on the rows: -probg + sum_theta [cc (ln2 - ln x) - hcc - 2 fc / x]
    # + sum_m [mc ln m - fm m]
    p = -L[:, 19]  # probg (ginfo.cpp:288-316 with 3 + 2 parameters)
    for i in range(3):
        p += np.floor(L[:, i] + 0.5) * (math.log(2) - math.log(x[i])) - L[:, 6 + i] - \
            2 * L[:, 3 + i] / x[i]
    for i in range(2):
        p += np.floor(L[:, 9 + i] + 0.5) * math.log(x[3 + i]) - L[:, 11 + i] * x[3 + i]
    mx = p.max()
    ref = math.log(len(p)) - (mx + math.log(np.exp(p - mx).sum()))
    assert abs(oracle.jointp_exact(tab, 3, 2, 2, x) - ref) <= 1e-10 * abs(ref)
