"""Pins the oracle's infinite-sites likelihood (labelgtree, calc_sumlogk,
likelihoodIS: calc_prob_data.cpp:537-836) against independent statements:

* an independent Fitch parsimony count: the number of "0/1 nodes" the
  reference's labelling finds equals the Fitch score, so a site is accepted
  iff it is compatible with the genealogy;
* an independent clade construction of the mutation branch: the tips
  carrying the derived state are exactly the clade below one edge; the 0/1
  node is that edge's parent and the branch length is that edge's, or the
  sum of both root branches when the parent is the root;
* the closed form -length*u + sum log(ptime*u) - sum log(k!);
* sumlogk is the same on every genealogy compatible with the data that
  puts each site on the same node.
"""
import math

import numpy as np
import pytest

import oracle
from ima2p_amd import sim
from tests import datasets as D


def fitch_score(g, col):
    sets = {}

    def go(e):
        if g.up0[e] < 0:
            sets[e] = {int(col[e])}
            return 0
        c = go(int(g.up0[e])) + go(int(g.up1[e]))
        a, b = sets[int(g.up0[e])], sets[int(g.up1[e])]
        if a & b:
            sets[e] = a & b
            return c
        sets[e] = a | b
        return c + 1
    return go(g.root)


def clade_placement(g, col):
    """(0/1 node, branch length) from clades, or None when incompatible."""
    cl = g.clades()
    bl = g.branch_lengths()
    full = (1 << g.n) - 1
    m1 = sum(1 << t for t in range(g.n) if col[t] == 1)
    m0 = full & ~m1
    x = [e for e in range(2 * g.n - 1) if e != g.root and cl[e] == m1]
    y = [e for e in range(2 * g.n - 1) if e != g.root and cl[e] == m0]
    c = x[0] if x else (y[0] if y else None)
    if c is None:
        return None
    N = int(g.down[c])
    sis = int(g.up1[N]) if g.up0[N] == c else int(g.up0[N])
    return N, (bl[c] + bl[sis]) if g.down[N] == -1 else bl[c]


@pytest.mark.parametrize("seed", range(8))
def test_labelling_counts_fitch_changes(seed):
    rng = np.random.default_rng(seed)
    g = D.random_topology(rng, int(rng.integers(3, 25)))
    for _ in range(60):
        col = rng.integers(0, 2, g.n)
        col[0] = 0
        if col.sum() == 0:
            continue
        mut = oracle.labelgtree_site(g, col)
        n01 = sum(1 for e in range(g.n, 2 * g.n - 1)
                  if {mut[g.up0[e]], mut[g.up1[e]]} == {0, 1})
        assert n01 == fitch_score(g, col)


@pytest.mark.parametrize("seed,samppop", [(1, (10, 10)), (2, (3, 40)), (3, (25, 25)),
                                          (4, (15, 15, 15))])
def test_placement_and_loglik_match_clade_construction(seed, samppop):
    model = "3pop" if len(samppop) == 3 else "2pop"
    g, loc = D.is_case(seed, samppop, u=5.0, model=model)
    S = loc.seq.shape[1]
    assert S > 0
    rc, sl = oracle.calc_sumlogk(g, loc)
    assert rc == 0
    u = 1.7
    v, st, nodes = oracle.likelihood_is(g, loc, u, sl)
    assert st == 0
    counts = {}
    acc = -g.length() * u
    for s in range(S):
        N, pt = clade_placement(g, loc.seq[:, s])
        assert nodes[s] == N
        counts[N] = counts.get(N, 0) + 1
        acc += math.log(pt * u)
    sl_ref = sum(math.lgamma(k + 1) for k in counts.values())
    assert abs(sl - sl_ref) <= 1e-12 * max(1.0, sl_ref)
    acc -= sl_ref
    assert abs(v - acc) <= 1e-11 * abs(acc)


def test_incompatible_genealogy_is_rejected():
    rng = np.random.default_rng(12)
    g, loc = D.is_case(12, (10, 10), u=8.0)
    nrej = 0
    for _ in range(30):
        other = D.random_topology(rng, g.n)
        v, st, _ = oracle.likelihood_is(other, loc, 1.0, 0.0)
        compatible = all(clade_placement(other, loc.seq[:, s]) is not None
                         for s in range(loc.seq.shape[1]))
        assert (v != oracle.REJECT_IS) == compatible
        nrej += v == oracle.REJECT_IS
    assert nrej > 0


def test_sumlogk_from_incompatible_genealogy_is_left_zero():
    """calc_prob_data.cpp:654-658 returns before setting the constant."""
    rng = np.random.default_rng(3)
    g, loc = D.is_case(3, (10, 10), u=8.0)
    for _ in range(50):
        other = D.random_topology(rng, g.n)
        rc, sl = oracle.calc_sumlogk(other, loc)
        if rc == -2:
            assert sl == 0.0
            return
    pytest.skip("no incompatible genealogy drawn")


def test_tree_length_is_sum_of_branches():
    g, _ = D.is_case(5, (7, 8))
    assert abs(oracle.tree_length(g) - g.branch_lengths().sum()) < 1e-12
