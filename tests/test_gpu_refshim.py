"""RefShim: the reference's per-(chain, locus) call shapes
(likelihoodHKY(ci, li, mutrate, kappa, e1..e4), likelihoodIS(ci, li,
mutrate), likelihoodSW(ci, li, ai, u, tr); calc_prob_data.cpp:584, 732,
842) over the engine -- each call uploads one genealogy and scores it
alone, in FP64 -- equal the oracle's functions on the same genealogies
(1e-10), including the infinite-sites constant taken from the first
genealogy a locus sees (calc_prob_data.cpp:752-765) and the reject
sentinel of an incompatible genealogy."""
import numpy as np
import pytest

import oracle
from ima2p_amd import sim
from ima2p_amd.engine import Engine, RefShim
from tests import datasets as D

pytestmark = pytest.mark.gpu


def test_reference_shaped_calls_equal_the_oracle():
    rng = np.random.default_rng(61)
    e = Engine(0)
    g0 = D.genealogy(rng, (5, 6))
    hky = D.hky_locus(rng, g0, nsites=250)
    isl = sim.simulate_is(rng, g0, 2.0)
    A = sim.simulate_sw(rng, g0, 1.5)
    lids = [e.add_locus_hky(hky), e.add_locus_is(isl), e.add_locus_sw(g0.n)]
    shim = RefShim(e, lids, nchains=2)
    _, sl = oracle.calc_sumlogk(g0, isl)
    for ci in range(2):
        g = g0 if ci == 0 else sim.relabel_internal(g0, rng)
        got = shim.likelihoodHKY(ci, 0, g, 1.7, 3.0, 4, 5, -1, -1)
        ref, _ = oracle.likelihood_hky(g, hky, 1.7, 3.0)
        assert abs(got - ref) <= 1e-10 * abs(ref)
        got = shim.likelihoodIS(ci, 1, g, 2.0)
        ref, st, _ = oracle.likelihood_is(g, isl, 2.0, sl)
        assert st == 0 and abs(got - ref) <= 1e-10 * abs(ref)
        got = shim.likelihoodSW(ci, 2, g, A, 1.5)
        ref, _ = oracle.likelihood_sw(g, A, 1.5)
        assert abs(got - ref) <= 1e-10 * abs(ref)
    # a genealogy the infinite-sites data reject
    other = D.genealogy(np.random.default_rng(5), (5, 6))
    _, st, _ = oracle.likelihood_is(other, isl, 2.0, sl)
    if st != 0:
        assert shim.likelihoodIS(1, 1, other, 2.0) == oracle.REJECT_IS
    e.close()
