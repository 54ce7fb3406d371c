"""Oracle finishSWupdateA (update_gtree_common.cpp:2219-2362), pinned by
what the mathematics fixes:

* the returned log-likelihood change equals likelihoodSW of the whole
  genealogy after the move (with the proposed allele) minus that before
  it, and the four returned edge terms equal likelihoodSW's per-edge
  terms of the new genealogy: the incremental bookkeeping drops no edge;
* the proposed allele is the geometric step of the forward parameter,
  scipy's geometric quantile at the same uniform, clamped to the range;
* Aterm is the ratio of the reverse and forward geometric probabilities
  (scipy's logpmf);
* the reverse move (edge back onto its old sister at its old time) that
  draws the reverse step returns the negated likelihood change and Aterm.
"""
import numpy as np
import pytest
from scipy.stats import geom

import oracle
from ima2p_amd import sim


def _tree(rng, n, u=2.0):
    pt = sim.PopTree(1, [])
    g = sim.simulate_genealogy(rng, [n], [4.0], [], np.zeros((1, 1)), pt)
    A = sim.simulate_sw(rng, g, u)
    return g, A


def _neighbour_g(vals, ref):
    j = len(vals)
    g = j / float(sum(abs(v - ref) for v in vals) + j)
    return min(g, 0.95)


@pytest.mark.parametrize("seed", range(12))
def test_update_matches_full_likelihood_difference(seed):
    rng = np.random.default_rng(seed)
    u = float(rng.uniform(0.5, 3.0))
    for _ in range(40):
        n = int(rng.integers(3, 14))
        g, A = _tree(rng, n, u)
        minA, maxA = int(A.min()) - 3, int(A.max()) + 3
        old, dl_old = oracle.likelihood_sw(g, A, u)
        h, mv = sim.random_sw_move(rng, g, A, dl_old, u, minA, maxA)
        dlike, aterm, newA, ed, A2 = oracle.finish_sw_update_A(h, A, mv)
        assert A2[mv.downedge] == newA and minA <= newA <= maxA
        new, dl_new = oracle.likelihood_sw(h, A2, u)
        assert abs(dlike - (new - old)) < 1e-9 * max(1.0, abs(old)), (dlike, new - old)
        for k, e in enumerate((mv.sisedge, mv.edge, mv.newsisedge, mv.downedge)):
            assert abs(ed[k] - dl_new[e]) < 1e-12 * max(1.0, abs(dl_new[e])), (k, e)
        # geometric step by inversion, then the clamp
        oldA = int(A[mv.downedge])
        nb = [A2[e] for e in (mv.edge, mv.newsisedge)]
        if h.down[mv.downedge] != -1:
            nb.append(A2[h.down[mv.downedge]])
        gnew = _neighbour_g(nb, oldA)
        k = int(geom.ppf(1.0 - mv.uni, gnew)) - 1
        step = -k if mv.flip else k
        assert newA == min(max(oldA + step, minA), maxA)
        # reverse over forward proposal probability
        back = [mv.copyA0, mv.copyA1] + ([] if mv.down_was_root else [mv.hold_down_A])
        gold = _neighbour_g(back, newA)
        d = abs(newA - oldA)
        want = geom.logpmf(d + 1, gold) - geom.logpmf(d + 1, gnew)
        assert abs(aterm - want) < 1e-12 * max(1.0, abs(want))


def _reverse(rng, g, A, h, A2, mv, u):
    """The move that puts edge back on its old sister at its old time, with
    the uniform that draws the step back to the old allele (bisection over
    the monotone quantile), or None if no uniform does."""
    _, dl_h = oracle.likelihood_sw(h, A2, u)
    par = int(h.down[mv.downedge])
    oldA, newA = int(A[mv.downedge]), int(A2[mv.downedge])
    rv = sim.SWMove(mv.edge, mv.downedge, mv.newsisedge, mv.sisedge, int(A2[mv.edge]),
                    int(A2[mv.newsisedge]), par == -1, int(A2[par]) if par != -1 else -1,
                    mv.minA, mv.maxA, int(oldA < newA),
                    (float(dl_h[mv.edge]), float(dl_h[mv.newsisedge]),
                     float(dl_h[mv.downedge]) if par != -1 else 0.0),
                    float(dl_h[mv.sisedge]), u, 0.5)
    g2 = sim.move_edge(h, mv.edge, mv.sisedge, float(g.time[mv.edge]))
    lo, hi = 1e-300, 1.0
    for _ in range(200):
        rv.uni = 0.5 * (lo + hi)
        _, _, a, _, _ = oracle.finish_sw_update_A(g2, A2, rv)
        if a == oldA:
            return g2, rv
        # a larger uniform gives a smaller step
        if abs(a - newA) > abs(oldA - newA):
            lo = rv.uni
        else:
            hi = rv.uni
    return None


@pytest.mark.parametrize("seed", range(6))
def test_reverse_move_negates(seed):
    rng = np.random.default_rng(100 + seed)
    u = 1.5
    done = 0
    while done < 25:
        g, A = _tree(rng, int(rng.integers(3, 10)), u)
        minA, maxA = int(A.min()) - 50, int(A.max()) + 50
        _, dl_old = oracle.likelihood_sw(g, A, u)
        h, mv = sim.random_sw_move(rng, g, A, dl_old, u, minA, maxA)
        dlike, aterm, newA, _, A2 = oracle.finish_sw_update_A(h, A, mv)
        r = _reverse(rng, g, A, h, A2, mv, u)
        if r is None:
            continue
        g2, rv = r
        np.testing.assert_allclose(np.sort(g2.time), np.sort(g.time), rtol=0, atol=0)
        dl2, at2, a2, _, _ = oracle.finish_sw_update_A(g2, A2, rv)
        assert a2 == A[mv.downedge]
        assert abs(dl2 + dlike) < 1e-9 * max(1.0, abs(dlike))
        assert abs(at2 + aterm) < 1e-9 * max(1.0, abs(aterm))
        done += 1


def test_no_likelihood_option():
    rng = np.random.default_rng(7)
    g, A = _tree(rng, 6)
    _, dl = oracle.likelihood_sw(g, A, 2.0)
    h, mv = sim.random_sw_move(rng, g, A, dl, 2.0, -100, 200)
    v, at, newA, _, A2 = oracle.finish_sw_update_A(h, A, mv, calc=False)
    assert v == 0.0 and at == 0.0 and A2[mv.downedge] == newA
