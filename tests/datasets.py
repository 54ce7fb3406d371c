"""Seeded synthetic cases shared by the tests."""
from __future__ import annotations

import numpy as np

from ima2p_amd import sim


def genealogy(rng, samppop=(10, 10), model="2pop", relabel=True):
    if model == "2pop":
        pt, theta, ts, mig = sim.im_model_two_pop()
    else:
        pt, theta, ts, mig = sim.im_model_three_pop(rng)
    g = sim.simulate_genealogy(rng, samppop, theta, ts, mig, pt)
    if relabel:
        g = sim.relabel_internal(g, rng)
    return g


def hky_locus(rng, g, nsites=300, pi=None, kappa=4.0, u=2.0, gaps=0):
    if pi is None:
        pi = rng.dirichlet([10, 10, 10, 10])
    raw = sim.simulate_hky_sequences(rng, g, nsites, pi, kappa, u).astype(np.int16)
    for _ in range(gaps):
        raw[rng.integers(g.n), rng.integers(nsites)] = -1
    return sim.encode_hky(raw)


def hky_case(seed, samppop=(10, 10), nsites=300, model="2pop"):
    rng = np.random.default_rng(seed)
    g_true = genealogy(rng, samppop, model)
    loc = hky_locus(rng, g_true, nsites)
    g = genealogy(rng, samppop, model)  # a genealogy that did not make the data
    return g, loc


def is_case(seed, samppop=(10, 10), u=3.0, model="2pop"):
    rng = np.random.default_rng(seed)
    g = genealogy(rng, samppop, model)
    loc = sim.simulate_is(rng, g, u)
    return g, loc


def sw_case(seed, samppop=(10, 10), u=2.0, model="2pop"):
    rng = np.random.default_rng(seed)
    g = genealogy(rng, samppop, model)
    A = sim.simulate_sw(rng, g, u)
    return g, A


def random_topology(rng, n):
    """A random genealogy (Kingman coalescent, single population)."""
    return sim.relabel_internal(
        sim.simulate_genealogy(rng, [n], np.array([4.0]), np.array([]), np.zeros((1, 1)),
                               sim.PopTree(1, [])), rng)
