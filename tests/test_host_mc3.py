"""The host-driven sweep (ima2p_amd/hostmc3.py, the plain per-locus path's
control loop) with the rubber-band time update (changet_RY1,
update_t_RY.cpp:223-517), run on the oracle's likelihood and prior: after
sweeps that accept some split-time moves, every stored log-likelihood and
every chain's probg equal a fresh evaluation of the genealogies and split
times now held, genealogies stay valid for their split times (treeweight
raises otherwise), and the moved split time stays inside its neighbours
and the prior bound."""
import numpy as np

import oracle
from ima2p_amd import immodel, sim
from ima2p_amd.hostmc3 import HostMC3, HostPrior


class OraclePrior(HostPrior):
    def __init__(self, model, samp):
        self.model, self.samp = model, samp
        self.gens, self.tv = {}, {}

    def _eval(self, gens, tv):
        ws = [oracle.treeweight(g, self.model, tv, self.samp) for g in gens]
        return oracle.integrate_tree_prob(ws, self.model)[0]

    def init(self, c, gens, tv):
        self.gens[c] = gens
        self.tv[c] = np.array(tv, float)
        return self._eval(gens, tv)

    def try_locus(self, c, l, g):
        return self._eval(self.gens[c], self.tv[c])

    def settle_locus(self, c, l, accept, old=None):
        pass

    def try_chain(self, c, gens, tv):
        return self._eval(gens, tv)

    def settle_chain(self, c, accept, old=None, tv=None):
        if accept:
            self.tv[c] = np.array(tv, float)


def test_rubber_band_sweeps_keep_the_state_consistent():
    rng = np.random.default_rng(21)
    pt, theta, ts, mig = sim.im_model_three_pop(rng)
    model = immodel.build(pt, qmax=20.0, mmax=5.0)
    samp = [3, 2, 3]
    loci, base = [], []
    for _ in range(2):
        g = sim.simulate_genealogy(rng, samp, theta, ts, mig, pt)
        loci.append(sim.encode_hky(sim.simulate_hky_sequences(
            rng, g, 80, np.array([0.3, 0.2, 0.2, 0.3]), 4.0, 2.0)))
        base.append(g)
    C, tmax = 3, 10.0
    chains = [[g.copy() for g in base] for _ in range(C)]
    u = [2.0, 1.5]

    def lik(c, l, g, dirty, commit):
        return oracle.likelihood_hky(g, loci[l], u[l], 4.0)[0]

    h = HostMC3(chains, np.ones(C), lik, lambda c, l, a: None, seed=5,
                prior=OraclePrior(model, samp), tv=np.asarray(ts, float), tmax=tmax, ry1=True)
    for k in range(16):
        h.sweep(with_scale=bool(k & 1))
    assert h.stats[2] == 8 * C and h.stats[3] > 0
    assert any(np.any(h.tv[c] != np.asarray(ts)) for c in range(C))
    for c in range(C):
        tv = h.tv[c]
        assert np.all(np.diff(tv) > 0) and tv[0] > 0 and tv[-1] < tmax
        for l in range(2):
            ref = oracle.likelihood_hky(chains[c][l], loci[l], u[l], 4.0)[0]
            assert abs(h.pdg[c, l] - ref) <= 1e-9 * abs(ref)
        ws = [oracle.treeweight(g, model, tv, samp) for g in chains[c]]
        assert abs(h.probg[c] - oracle.integrate_tree_prob(ws, model)[0]) <= 1e-9 * abs(h.probg[c])
