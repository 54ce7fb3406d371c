"""Pins the oracle's restatement of updategenealogy's accept step
(update_gtree.cpp:900-961) for a chain's loci in order, used as the
checker of ima_mc3_decide:

* proposals equal to the committed genealogies: every decision is exactly
  U < min(1, exp(h)) and the chain's probg is unchanged;
* a rejected likelihood (incompatible infinite-sites genealogy) is never
  accepted, whatever the Hastings term;
* U = 0 accepts every proposal: the final probg is integrate_tree_prob over
  all the proposals' weights (whatever the order), and the accepted tpw
  terms telescope to it; U = 1 rejects all and changes nothing;
* gbeta = 0 removes the likelihood from the decision.
"""
import math

import numpy as np

import oracle
from ima2p_amd import immodel, sim


def _chain(seed, Lc=4):
    rng = np.random.default_rng(seed)
    pt, theta, ts, mig = sim.im_model_two_pop(t=0.8, m12=0.4, m21=0.6)
    model = immodel.build(pt, qmax=15.0, mmax=4.0)
    samp = [4, 5]
    gs = [sim.simulate_genealogy(rng, samp, theta, ts, mig, pt) for _ in range(Lc)]
    ps = [sim.simulate_genealogy(rng, samp, theta, ts, mig, pt) for _ in range(Lc)]
    w = [oracle.treeweight(g, model, np.asarray(ts, float), samp) for g in gs]
    wn = [oracle.treeweight(g, model, np.asarray(ts, float), samp) for g in ps]
    return rng, model, w, wn


def test_identical_proposals_decide_on_the_hastings_term():
    rng, model, w, _ = _chain(1)
    pdg = rng.normal(-100, 5, len(w))
    h = np.array([-0.5, 0.3, -2.0, -0.01])
    U = np.array([0.7, 0.99, 0.1, 0.995])
    dec, pdg2, _, probg = oracle.updategenealogy_decisions(pdg, w, pdg, w, [True] * 4, h, U,
                                                           0.8, model)
    assert dec == [0.7 < math.exp(-0.5), True, 0.1 < math.exp(-2.0), 0.995 < math.exp(-0.01)]
    assert probg == oracle.integrate_tree_prob(w, model)[0]
    assert list(pdg2) == list(pdg)


def test_rejected_likelihood_is_never_accepted():
    rng, model, w, wn = _chain(2)
    pdg = rng.normal(-100, 5, len(w))
    dec, pdg2, w2, probg = oracle.updategenealogy_decisions(
        pdg, w, pdg + 50.0, wn, [False] * 4, [100.0] * 4, [0.0] * 4, 1.0, model)
    assert dec == [False] * 4 and list(pdg2) == list(pdg)
    assert probg == oracle.integrate_tree_prob(w, model)[0]


def test_all_accepted_telescopes_and_all_rejected_is_inert():
    rng, model, w, wn = _chain(3)
    pdg = rng.normal(-100, 5, len(w))
    npd = pdg + rng.normal(0, 3, len(w))
    p0 = oracle.integrate_tree_prob(w, model)[0]
    dec, pdg2, w2, probg = oracle.updategenealogy_decisions(
        pdg, w, npd, wn, [True] * 4, [0.0] * 4, [0.0] * 4, 0.6, model)
    assert all(dec) and list(pdg2) == list(npd)
    pall = oracle.integrate_tree_prob(wn, model)[0]
    assert abs(probg - pall) <= 1e-12 * abs(pall)
    assert probg != p0
    dec, pdg2, w2, probg = oracle.updategenealogy_decisions(
        pdg, w, npd, wn, [True] * 4, [0.0] * 4, [1.0] * 4, 0.6, model)
    assert not any(dec) and list(pdg2) == list(pdg) and probg == p0


def test_gbeta_zero_removes_the_likelihood():
    rng, model, w, _ = _chain(4)
    pdg = rng.normal(-100, 5, len(w))
    # identical genealogies, likelihood 1000 nats worse: with gbeta = 0 the
    # decision is the Hastings term's alone
    dec, *_ = oracle.updategenealogy_decisions(pdg, w, pdg - 1000.0, w, [True] * 4,
                                               [0.0] * 4, [0.5] * 4, 1.0, model, gbeta=0.0)
    assert dec == [True] * 4
    dec, *_ = oracle.updategenealogy_decisions(pdg, w, pdg - 1000.0, w, [True] * 4,
                                               [0.0] * 4, [0.5] * 4, 1.0, model, gbeta=1.0)
    assert dec == [False] * 4
