"""Pins the oracle's likelihoodHKY (calc_prob_data.cpp:118-607) without the
reference binary, against mathematics that fixes the result independently:

* pruning with transition matrices from scipy's matrix exponential of the
  HKY rate matrix, normalised by its own stationary substitution rate
  (pins the closed-form P(t), getstandfactor and the per-site rate
  u / (totsites * standfactor) of :593);
* brute-force summation over every internal-state assignment (n <= 5);
* the two-tip closed form pi_a P_ab(t1 + t2) (reversibility);
* root-position (pulley) and time/rate scaling invariances;
* compressed patterns with counts vs one column per site (sortseq);
* incremental evaluation (e1..e4) against a full one.
A dropped term, a wrong class sum or eigenvalue, or a transposed P fails at
least one of these.
"""
import itertools

import numpy as np
import pytest
from scipy.linalg import expm

import oracle
from ima2p_amd import sim
from tests import datasets as D


def hky_Q(pi, kappa):
    Q = np.zeros((4, 4))
    for i in range(4):
        for j in range(4):
            if i != j:
                Q[i, j] = pi[j] * (kappa if (i + j) in (2, 4) else 1.0)
        Q[i, i] = -Q[i].sum()
    return Q


def expm_loglik(g, seq, mult, pi, kappa, u, totsites):
    Q = hky_Q(pi, kappa)
    rate = -(pi * np.diag(Q)).sum()
    mr = u / (totsites * rate)
    bl = g.branch_lengths()
    S = seq.shape[1]

    def part(e):
        if g.up0[e] < 0:
            L = np.zeros((S, 4))
            L[np.arange(S), seq[e]] = 1.0
            return L
        a, b = g.up0[e], g.up1[e]
        return (part(a) @ expm(Q * mr * bl[a]).T) * (part(b) @ expm(Q * mr * bl[b]).T)

    return float((mult * np.log(part(g.root) @ pi)).sum())


@pytest.mark.parametrize("seed,samppop,nsites", [(1, (5, 6), 200), (2, (10, 10), 600),
                                                 (3, (2, 30), 900), (4, (15, 15, 15), 1000)])
def test_matches_expm_pruning(seed, samppop, nsites):
    model = "3pop" if len(samppop) == 3 else "2pop"
    g, loc = D.hky_case(seed, samppop, nsites, model=model)
    for u, kappa in ((0.3, 1.0), (2.0, 4.0), (7.5, 25.0)):
        ref = expm_loglik(g, loc.seq, loc.mult, loc.pi, kappa, u, loc.totsites)
        got, _ = oracle.likelihood_hky(g, loc, u, kappa)
        assert abs(got - ref) <= 1e-10 * abs(ref), (u, kappa, got, ref)


def test_standfactor_is_stationary_rate():
    rng = np.random.default_rng(0)
    for _ in range(20):
        pi = rng.dirichlet([3, 3, 3, 3])
        kappa = rng.uniform(0.5, 30)
        Q = hky_Q(pi, kappa)
        assert abs(oracle.getstandfactor(pi, kappa) + (pi * np.diag(Q)).sum()) < 1e-14


def test_bruteforce_internal_states():
    rng = np.random.default_rng(7)
    for trial in range(6):
        n = int(rng.integers(2, 6))
        g = D.random_topology(rng, n)
        pi = rng.dirichlet([2, 2, 2, 2])
        kappa, u = rng.uniform(1, 10), rng.uniform(0.5, 3)
        S = 7
        seq = rng.integers(0, 4, (n, S)).astype(np.int8)
        loc = sim.HKYLocus(n, seq, np.ones(S, np.int32), S, pi)
        Q = hky_Q(pi, kappa)
        mr = u / (S * oracle.getstandfactor(pi, kappa))
        bl = g.branch_lengths()
        P = {e: expm(Q * mr * bl[e]) for e in range(2 * n - 1) if g.down[e] != -1}
        internal = [e for e in range(n, 2 * n - 1)]
        total = 0.0
        for s in range(S):
            lik = 0.0
            for states in itertools.product(range(4), repeat=len(internal)):
                st = dict(zip(internal, states))
                for t in range(n):
                    st[t] = seq[t, s]
                # the state of edge e sits at its top; the root edge's top is the root
                p = pi[st[g.root]]
                for e in range(2 * n - 1):
                    if g.down[e] != -1:
                        p *= P[e][st[g.down[e]], st[e]]
                lik += p
            total += np.log(lik)
        got, _ = oracle.likelihood_hky(g, loc, u, kappa)
        assert abs(got - total) <= 1e-11 * abs(total)


def test_two_tips_closed_form():
    rng = np.random.default_rng(3)
    for _ in range(10):
        t1 = t2 = rng.uniform(0.01, 3.0)
        up0 = np.array([-1, -1, 0]); up1 = np.array([-1, -1, 1])
        down = np.array([2, 2, -1]); time = np.array([t1, t2, sim.TIMEMAX])
        g = sim.Genealogy(2, up0, up1, down, time, 2, np.zeros(3, np.int32))
        pi = rng.dirichlet([4, 4, 4, 4])
        kappa, u = rng.uniform(1, 8), rng.uniform(0.2, 4)
        seq = np.array([[0, 1, 2, 3, 0, 2], [0, 3, 2, 1, 2, 0]], np.int8)
        loc = sim.HKYLocus(2, seq, np.array([3, 1, 2, 1, 1, 5], np.int32), 100, pi)
        mr = u / (100 * oracle.getstandfactor(pi, kappa))
        P = expm(hky_Q(pi, kappa) * mr * (t1 + t2))
        ref = sum(m * np.log(pi[a] * P[a, b]) for (a, b), m in zip(seq.T, loc.mult))
        got, _ = oracle.likelihood_hky(g, loc, u, kappa)
        assert abs(got - ref) <= 1e-11 * abs(ref)


def test_root_position_and_scaling_invariance():
    g, loc = D.hky_case(11, (8, 9), 500)
    base, _ = oracle.likelihood_hky(g, loc, 1.5, 3.0)
    # pulley: move the root's coalescence so one root branch gains what the
    # other loses is not expressible with a shared node time, so check the
    # equivalent statement that scaling every time by c and u by 1/c keeps
    # the likelihood, and that relabelling internal edges keeps it
    g2 = g.copy()
    mask = g2.down != -1
    g2.time[mask] = g2.time[mask] * 2.5
    scaled, _ = oracle.likelihood_hky(g2, loc, 1.5 / 2.5, 3.0)
    assert abs(scaled - base) <= 1e-11 * abs(base)
    rl = sim.relabel_internal(g, np.random.default_rng(1))
    relab, _ = oracle.likelihood_hky(rl, loc, 1.5, 3.0)
    assert abs(relab - base) <= 1e-12 * abs(base)


def test_compressed_patterns_equal_per_site_columns():
    rng = np.random.default_rng(5)
    g = D.genealogy(rng, (9, 9))
    raw = sim.simulate_hky_sequences(rng, g, 400, [0.1, 0.2, 0.3, 0.4], 5.0, 3.0)
    loc = sim.encode_hky(raw)
    assert loc.seq.shape[1] < 400 and loc.mult.sum() == 400
    per_site = sim.HKYLocus(g.n, raw.astype(np.int8), np.ones(400, np.int32), 400, loc.pi)
    a, _ = oracle.likelihood_hky(g, loc, 2.0, 5.0)
    b, _ = oracle.likelihood_hky(g, per_site, 2.0, 5.0)
    assert abs(a - b) <= 1e-11 * abs(b)


def test_empirical_pi_counts_dropped_gap_columns():
    """readata.cpp:296-316: pi counts every read base, including columns
    later removed for gaps; totsites counts the remaining columns."""
    raw = np.array([[0, 1, -1, 3], [0, 2, 2, 3], [1, 1, 2, 3]], np.int16)
    loc = sim.encode_hky(raw)
    counts = np.array([2, 3, 3, 3], float)  # A C G T over every non-gap entry
    np.testing.assert_allclose(loc.pi, counts / counts.sum())
    assert loc.totsites == 3 and loc.seq.shape[1] == 3


def test_incremental_equals_full():
    rng = np.random.default_rng(21)
    g, loc = D.hky_case(21, (12, 12), 600)
    _, st = oracle.likelihood_hky(g, loc, 2.0, 4.0)
    for _ in range(20):
        g2 = g.copy()
        v = int(rng.integers(g.n, 2 * g.n - 1))
        a, b = int(g2.up0[v]), int(g2.up1[v])
        lo = max(g2.time[g2.up0[a]] if a >= g.n else 0.0, g2.time[g2.up0[b]] if b >= g.n else 0.0)
        hi = g2.time[v] if g2.down[v] != -1 else g2.time[a] * 1.5
        g2.time[a] = g2.time[b] = lo + (hi - lo) * rng.uniform(0.1, 0.9)
        inc, _ = oracle.likelihood_hky(g2, loc, 2.0, 4.0, e=(a, b, v, -1), state=st, commit=False)
        full, _ = oracle.likelihood_hky(g2, loc, 2.0, 4.0)
        assert abs(inc - full) <= 1e-13 * abs(full)
        accept = rng.random() < 0.5
        oracle.hky_commit(g2, st, accept)
        if accept:
            g = g2
    final, _ = oracle.likelihood_hky(g, loc, 2.0, 4.0)
    again, _ = oracle.likelihood_hky(g, loc, 2.0, 4.0, e=(g.n, -1, -1, -1), state=st, commit=False)
    assert abs(final - again) <= 1e-13 * abs(final)
