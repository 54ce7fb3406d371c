"""GPU parity of the batched data likelihoods against the oracle.

Bars: FP64 HKY, IS and SW within 1e-10 relative of the oracle; IS 0/1 node
of every site bit-exact; FP32-scaled HKY within 1e-6 relative (the stated
FP32 tolerance, measured error is printed).
"""
import numpy as np
import pytest

import oracle
from ima2p_amd import _lib as L
from ima2p_amd import sim
from tests import datasets as D

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-10
FP32_RTOL = 1e-6


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_hky_batch_fp64_fp32_mixed_sizes(engine):
    cases = []
    rng = np.random.default_rng(11)
    for k in range(12):
        n1, n2 = int(rng.integers(2, 30)), int(rng.integers(2, 30))
        nsites = int(rng.integers(50, 1500))
        cases.append(D.hky_case(100 + k, (n1, n2), nsites))
    loci = [engine.add_locus_hky(loc) for _, loc in cases]
    first = engine.alloc_slots(loci)
    engine.write_slots(first, [g for g, _ in cases])
    us = rng.uniform(0.5, 3.0, len(cases))
    ks = rng.uniform(1.0, 8.0, len(cases))
    items = engine.items(np.arange(first, first + len(cases)), us, ks)
    out64, st = engine.score(items, L.PREC_FP64)
    out32, _ = engine.score(items, L.PREC_FP32)
    assert (st == 0).all()
    worst32 = 0.0
    for i, (g, loc) in enumerate(cases):
        ref, _ = oracle.likelihood_hky(g, loc, us[i], ks[i])
        assert rel(out64[i], ref) < FP64_RTOL, (i, out64[i], ref)
        worst32 = max(worst32, rel(out32[i], ref))
    print("HKY FP32 worst relative error", worst32)
    assert worst32 < FP32_RTOL


def test_hky_alone_equals_in_mix(engine):
    cases = [D.hky_case(300 + k, (8 + k, 9), 400) for k in range(6)]
    loci = [engine.add_locus_hky(loc) for _, loc in cases]
    first = engine.alloc_slots(loci)
    engine.write_slots(first, [g for g, _ in cases])
    items = engine.items(np.arange(first, first + 6), 1.7, 3.0)
    mix, _ = engine.score(items, L.PREC_FP32)
    for i in range(6):
        alone, _ = engine.score(engine.items([first + i], 1.7, 3.0), L.PREC_FP32)
        assert alone[0] == mix[i]


def test_hky_three_pop_45_genes(engine):
    g, loc = D.hky_case(7, (15, 15, 15), 1000, model="3pop")
    li = engine.add_locus_hky(loc)
    s = engine.alloc_slots([li])
    engine.write_slots(s, [g])
    out, _ = engine.score(engine.items([s], 1.3, 4.0), L.PREC_FP64)
    ref, _ = oracle.likelihood_hky(g, loc, 1.3, 4.0)
    assert rel(out[0], ref) < FP64_RTOL


def test_hky_single_call_matches_incremental_oracle(engine):
    """The reference-shaped call with e1..e4 equals the oracle's incremental
    evaluation after a genealogy change (same committed state)."""
    rng = np.random.default_rng(5)
    g, loc = D.hky_case(21, (12, 12), 600)
    li = engine.add_locus_hky(loc)
    _, st = oracle.likelihood_hky(g, loc, 2.0, 4.0)
    # perturb one internal node time within its bounds
    g2 = g.copy()
    v = int(rng.integers(g.n, 2 * g.n - 1))
    while g2.down[v] == -1:
        v = int(rng.integers(g.n, 2 * g.n - 1))
    a, b = g2.up0[v], g2.up1[v]
    lo = max(g2.time[g2.up0[a]] if g2.up0[a] >= 0 else 0, g2.time[g2.up0[b]] if g2.up0[b] >= 0 else 0)
    hi = g2.time[v]
    t = lo + (hi - lo) * 0.37
    g2.time[a] = g2.time[b] = t
    inc, _ = oracle.likelihood_hky(g2, loc, 2.0, 4.0, e=(a, b, v, -1), state=st, commit=False)
    full, _ = oracle.likelihood_hky(g2, loc, 2.0, 4.0)
    assert rel(inc, full) < 1e-13
    got = engine.likelihood_hky(li, g2, 2.0, 4.0, e=(a, b, v, -1))
    assert rel(got, inc) < FP64_RTOL


@pytest.mark.parametrize("seed,samppop", [(1, (10, 10)), (2, (4, 60)), (3, (30, 30)),
                                          (4, (2, 2)), (5, (40, 45))])
def test_is_matches_oracle_bit_exact_nodes(engine, seed, samppop):
    g, loc = D.is_case(seed, samppop, u=4.0)
    if loc.seq.shape[1] == 0:
        pytest.skip("no segregating sites")
    li = engine.add_locus_is(loc)
    s = engine.alloc_slots([li])
    engine.write_slots(s, [g])
    sl_gpu = engine.sumlogk_from_slot(li, s)
    rc, sl_ref = oracle.calc_sumlogk(g, loc)
    assert rc == 0 and sl_gpu == sl_ref
    out, st = engine.score(engine.items([s], 1.7))
    ref, rst, nodes = oracle.likelihood_is(g, loc, 1.7, sl_ref)
    assert st[0] == 0 and rst == 0
    assert rel(out[0], ref) < FP64_RTOL
    gn = engine.is_site_nodes(s, loc.seq.shape[1])
    assert (gn == nodes).all()


def test_is_reject_and_other_genealogy(engine):
    rng = np.random.default_rng(9)
    g, loc = D.is_case(31, (10, 10), u=6.0)
    li = engine.add_locus_is(loc)
    others = [D.random_topology(rng, 20) for _ in range(20)]
    first = engine.alloc_slots([li] * (len(others) + 1))
    engine.write_slots(first, [g] + others)
    rc, sl = oracle.calc_sumlogk(g, loc)
    engine.set_sumlogk(li, sl)
    out, st = engine.score(engine.items(np.arange(first, first + 21), 2.0))
    nrej = 0
    for i, gg in enumerate([g] + others):
        ref, rst, _ = oracle.likelihood_is(gg, loc, 2.0, sl)
        if ref == oracle.REJECT_IS:
            assert st[i] == L.STATUS_IS_REJECT and out[i] == L.REJECT_IS
            nrej += 1
        else:
            assert st[i] == 0 and rel(out[i], ref) < FP64_RTOL
    assert nrej > 0


@pytest.mark.parametrize("seed,u", [(1, 0.5), (2, 3.0), (3, 12.0)])
def test_sw_matches_oracle(engine, seed, u):
    g, A = D.sw_case(seed, (20, 25), u=u)
    li = engine.add_locus_sw(g.n)
    s = engine.alloc_slots([li])
    engine.write_slots(s, [g], A=[A])
    for uu, tr in ((u, 1.0), (0.3 * u, 1.7)):
        out, st = engine.score(engine.items([s], uu, tr=tr))
        ref, dl = oracle.likelihood_sw(g, A, uu, tr)
        if ref == -np.finfo(float).max:
            assert st[0] == L.STATUS_SW_ZERO and out[0] == ref
        else:
            assert st[0] == 0 and rel(out[0], ref) < FP64_RTOL
            gd = engine.sw_dlike(s, uu, tr)
            np.testing.assert_allclose(gd, dl, rtol=1e-12, atol=1e-300)


def test_sw_zero_bessel(engine):
    g, A = D.sw_case(4, (5, 5), u=1.0)
    A = A.copy()
    A[0] += 120  # I_120(x) underflows to 0 for tiny x (calc_prob_data.cpp:870-880)
    li = engine.add_locus_sw(g.n)
    s = engine.alloc_slots([li])
    engine.write_slots(s, [g], A=[A])
    out, st = engine.score(engine.items([s], 1e-6))
    ref, _ = oracle.likelihood_sw(g, A, 1e-6, 1.0)
    assert ref == -np.finfo(float).max
    assert st[0] == L.STATUS_SW_ZERO and out[0] == ref


def test_mixed_models_one_call(engine):
    gh, lh = D.hky_case(41, (10, 12), 500)
    gi, li_ = D.is_case(42, (10, 12), u=3.0)
    gs, A = D.sw_case(43, (10, 12))
    a = engine.add_locus_hky(lh)
    b = engine.add_locus_is(li_)
    c = engine.add_locus_sw(gs.n)
    first = engine.alloc_slots([a, b, c])
    engine.write_slots(first, [gh, gi, gs], A=[np.zeros(2 * gh.n - 1, np.int32),
                                              np.zeros(2 * gi.n - 1, np.int32), A])
    rc, sl = oracle.calc_sumlogk(gi, li_)
    engine.set_sumlogk(b, sl)
    items = engine.items([first, first + 1, first + 2], [1.5, 2.5, 3.0], [4.0, 1.0, 1.0],
                         [1.0, 1.0, 1.2])
    out, st = engine.score(items)
    assert rel(out[0], oracle.likelihood_hky(gh, lh, 1.5, 4.0)[0]) < FP64_RTOL
    assert rel(out[1], oracle.likelihood_is(gi, li_, 2.5, sl)[0]) < FP64_RTOL
    assert rel(out[2], oracle.likelihood_sw(gs, A, 3.0, 1.2)[0]) < FP64_RTOL
    # repeated passes (the kernel-timing entry point) give the same values
    rep, st2 = engine.score_repeat(items, 3)
    np.testing.assert_array_equal(rep, out)
    np.testing.assert_array_equal(st2, st)
