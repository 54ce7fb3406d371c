"""Genealogy prior on the GPU (treeweight, update_gtree_common.cpp:1680-1931;
integrate_tree_prob, :1945-2053) against the oracle on the same
genealogies: weights bit for bit (same event order, same arithmetic, no
contraction), the integrated prior within 1e-10 relative (the incomplete
gamma series stop at a 3e-7 relative term, so an ulp of difference in the
device's exp/log can only move the value by far less than that; 1e-10 is
the path's FP64 tolerance)."""
import numpy as np
import pytest

import oracle
from ima2p_amd import immodel, sim
from ima2p_amd.engine import EngineError
from tests import datasets as D

pytestmark = pytest.mark.gpu

PRIOR_RTOL = 1e-10


def _case(seed, three, nloci, samp, hval=1.0):
    rng = np.random.default_rng(seed)
    if three:
        pt, theta, ts, mig = sim.im_model_three_pop(rng)
    else:
        pt, theta, ts, mig = sim.im_model_two_pop(t=0.9, m12=0.7, m21=0.4)
    model = immodel.build(pt)
    gens = [sim.relabel_internal(sim.simulate_genealogy(rng, samp, theta, ts, mig, pt), rng)
            for _ in range(nloci)]
    return model, np.asarray(ts, float), gens


def _load(engine, model, gens, samp, hval=1.0):
    mid = engine.model_add(model)
    loci = []
    for g in gens:
        loc = engine.add_locus_sw(g.n)
        engine.locus_set_model(loc, mid, samp, hval)
        loci.append(loc)
    first = engine.alloc_slots(loci)
    engine.write_slots(first, gens, A=[np.zeros(2 * g.n - 1, np.int32) for g in gens])
    engine.write_pops(first, gens)
    return first


@pytest.mark.parametrize("seed,three,hval", [(1, False, 1.0), (2, False, 0.25),
                                             (3, True, 1.0), (4, True, 1.5)])
def test_treeweight_bit_exact(engine, seed, three, hval):
    samp = [6, 5, 7] if three else [9, 8]
    model, ts, gens = _case(seed, three, 6, samp)
    first = _load(engine, model, gens, samp, hval)
    got = engine.treeweight(np.arange(first, first + len(gens)), ts, model.npops)
    assert sum(len(m) for g in gens for m in g.mig) > 0
    for i, g in enumerate(gens):
        ref = oracle.treeweight(g, model, ts, samp, hval)
        np.testing.assert_array_equal(got["cc"][i], ref["cc"])
        np.testing.assert_array_equal(got["mc"][i], ref["mc"])
        np.testing.assert_array_equal(got["fc"][i], ref["fc"])
        np.testing.assert_array_equal(got["fm"][i], ref["fm"])
        np.testing.assert_array_equal(got["hcc"][i], ref["hcc"])
        assert got["length"][i] == ref["length"]


@pytest.mark.parametrize("seed,three", [(5, False), (6, True), (7, True)])
def test_prior_matches_integrate_tree_prob(engine, seed, three):
    samp = [5, 6, 4] if three else [10, 10]
    rng = np.random.default_rng(seed)
    if three:
        pt, theta, ts, mig = sim.im_model_three_pop(rng)
    else:
        pt, theta, ts, mig = sim.im_model_two_pop(t=0.9, m12=0.7, m21=0.4)
    model = immodel.build(pt)
    # three "chains" of four loci, each chain with its own split times
    tv = np.stack([np.asarray(ts, float) * (1.0 + 0.15 * c) for c in range(3)])
    gens = [sim.simulate_genealogy(rng, samp, theta, tv[c], mig, pt)
            for c in range(3) for _ in range(4)]
    first = _load(engine, model, gens, samp)
    groups = [list(range(first + 4 * c, first + 4 * c + 4)) for c in range(3)]
    probg, q, m = engine.prior(groups, tv, model)
    for c in range(3):
        w = [oracle.treeweight(gens[4 * c + j], model, tv[c], samp) for j in range(4)]
        rp, rq, rm = oracle.integrate_tree_prob(w, model)
        assert abs(probg[c] - rp) <= PRIOR_RTOL * abs(rp), (c, probg[c], rp)
        np.testing.assert_allclose(q[c], rq, rtol=PRIOR_RTOL, atol=1e-12)
        np.testing.assert_allclose(m[c], rm, rtol=PRIOR_RTOL, atol=1e-12)


def test_prior_many_loci_one_group(engine):
    """A chain's all-loci sum (allgweight) over 40 loci: large counts take
    the continued-fraction branch of the incomplete gamma."""
    samp = [12, 12]
    model, ts, gens = _case(8, False, 40, samp)
    first = _load(engine, model, gens, samp)
    probg, q, m = engine.prior([list(range(first, first + 40))], ts[None, :], model)
    w = [oracle.treeweight(g, model, ts, samp) for g in gens]
    rp, rq, rm = oracle.integrate_tree_prob(w, model)
    assert abs(probg[0] - rp) <= PRIOR_RTOL * abs(rp)


def test_population_label_outside_its_period_is_an_error(engine):
    samp = [6, 6]
    model, ts, gens = _case(9, False, 1, samp)
    g = gens[0].copy()
    # the root's coalescence happens after the split, where population 0 no
    # longer exists
    if g.time[g.up0[g.root]] <= ts[0]:
        pytest.skip("root below the split time in this draw")
    g.pop[g.root] = 0
    first = _load(engine, model, [g], samp)
    with pytest.raises(EngineError):
        engine.treeweight([first], ts, model.npops)


def test_slot_without_model_is_an_error(engine):
    g, loc = D.hky_case(3, (4, 4), 100)
    a = engine.add_locus_hky(loc)
    first = engine.alloc_slots([a])
    engine.write_slots(first, [g])
    with pytest.raises(EngineError):
        engine.treeweight([first], [1.0], 2)
