"""The reference's genealogy move on the device (IMA_MC3_TOPOLOGY:
slider_nomigration, joinsisdown, splitsisdown, the root move's Hastings
term and finishSWupdateA's allele proposal; update_gtree.cpp:79-528,
723-826, update_gtree_common.cpp:2219-2362).

Flat likelihood (u -> 0), one population, uniform prior on theta: the
chains must sample the coalescent prior integrated over theta.
* n = 2: every move slides the root, so the Hastings term of the root move
  is in every decision; the time to the common ancestor follows
  CDF x E1(x) - e^-x + 1, x = 2t/q.
* n = 3: each of the three rooted topologies has probability 1/3, and the
  ratio of the first coalescence time to the root time follows 3r/(1+2r).
* n = 4: the balanced topology has probability 1/3.
Each law is checked with the independent chains' spread as the error
(|z| < 4.5), the same batch-means test as the time-slide move's.

With data and a two-population model without migration (HKY, infinite
sites and stepwise loci): after sweeps that change topologies, every
genealogy is valid for the model, and every stored log-likelihood and
probg equal the oracle's on the genealogies, labels and alleles the device
holds (1e-10 / 1e-9).  Graph replay equals eager launches bit for bit, and
a model with migration parameters is refused.
"""
import math

import numpy as np
import pytest
from scipy.special import exp1

import oracle
from ima2p_amd import _lib as L
from ima2p_amd import immodel, sim
from ima2p_amd.engine import Engine, EngineError
from tests.test_gpu_mc3_prior import _check_law, _one_pop

pytestmark = pytest.mark.gpu

TOPO = L.MC3_GRAPH | L.MC3_TOPOLOGY


def _read(e, first, C, n):
    up0, up1, dn, tm, rt, A = e.read_slots(first, C)
    ne = 2 * n - 1
    return [(up0[c * ne:(c + 1) * ne], up1[c * ne:(c + 1) * ne], dn[c * ne:(c + 1) * ne],
             tm[c * ne:(c + 1) * ne], int(rt[c])) for c in range(C)]


def _tips_below(u0, u1, n, e):
    if e < n:
        return {e}
    return _tips_below(u0, u1, n, u0[e]) | _tips_below(u0, u1, n, u1[e])


def test_two_copies_root_moves_sample_the_prior():
    q, C = 8.0, 64
    e, first, _ = _one_pop(2, C, 15, q)
    e.mc3_run(300, TOPO)
    tm = []
    for _ in range(200):
        e.mc3_run(30, TOPO)
        for u0, u1, dn, t, rt in _read(e, first, C, 2):
            tm.append(t[u0[rt]])
    v = np.array(tm).reshape(-1, C).T

    def cdf(t):
        x = 2.0 * np.asarray(t) / q
        return float(x * exp1(x) - np.exp(-x) + 1.0)

    assert _check_law(v, cdf, (0.1, 0.25, 0.5, 0.75, 0.9)) < 4.5
    e.close()


def _fractions(vals, C, cats):
    v = np.asarray(vals).reshape(-1, C).T  # [chain, sample]
    out = []
    for k in cats:
        f = (v == k).mean(axis=1)
        out.append((f.mean(), f.std(ddof=1) / math.sqrt(C)))
    return out


def test_three_copies_topologies_and_interval_ratio():
    C = 64
    e, first, _ = _one_pop(3, C, 16, 8.0)
    e.mc3_run(300, TOPO)
    lone, ratio = [], []
    for _ in range(200):
        e.mc3_run(30, TOPO)
        for u0, u1, dn, t, rt in _read(e, first, C, 3):
            # the tip that is a daughter of the root
            a, b = u0[rt], u1[rt]
            lone.append(a if a < 3 else b)
            times = np.sort(t[dn != -1])
            ratio.append(times.min() / times.max())
    for k, (f, se) in zip(range(3), _fractions(lone, C, range(3))):
        print(f"  tip {k} joins last: {f:.4f} +- {se:.4f}")
        assert abs(f - 1 / 3) < 4.5 * max(se, 1e-3)
    r = np.array(ratio).reshape(-1, C).T
    assert _check_law(r, lambda x: 3 * x / (1 + 2 * x), (0.1, 0.25, 0.5, 0.75, 0.9)) < 4.5
    e.close()


def test_four_copies_balanced_topology_one_third():
    C = 64
    e, first, _ = _one_pop(4, C, 17, 8.0)
    e.mc3_run(300, TOPO)
    bal = []
    for _ in range(200):
        e.mc3_run(30, TOPO)
        for u0, u1, dn, t, rt in _read(e, first, C, 4):
            bal.append(int(u0[rt] >= 4 and u1[rt] >= 4))
    (f, se), = _fractions(bal, C, [1])
    print(f"  balanced: {f:.4f} +- {se:.4f}")
    assert abs(f - 1 / 3) < 4.5 * max(se, 1e-3)
    e.close()


def _two_pop_data(seed, C=6, with_mig=False):
    rng = np.random.default_rng(seed)
    if with_mig:
        pt, theta, ts, mrates = sim.im_model_two_pop(t=0.8, m12=0.5, m21=0.5)
        model = immodel.build(pt, qmax=15.0, mmax=3.0)
    else:
        pt, theta, ts, mrates = sim.im_model_two_pop(t=0.8, m12=0.0, m21=0.0)
        model = immodel.build(pt, qmax=15.0, migration=False)
    samp = [4, 4]
    e = Engine(0)
    mid = e.model_add(model)
    kinds = ["hky", "is", "sw", "hky"]
    loci, lids, base, As, u = [], [], [], [], []
    for kind in kinds:
        while True:
            g0 = sim.simulate_genealogy(rng, samp, theta, ts, mrates, pt)
            A = None
            if kind == "hky":
                data = sim.encode_hky(sim.simulate_hky_sequences(rng, g0, 200,
                                                                 rng.dirichlet([8] * 4), 4.0, 1.0))
                lid, uu = e.add_locus_hky(data), 1.0
            elif kind == "is":
                data = sim.simulate_is(rng, g0, 1.0)
                if data.seq.shape[1] == 0:
                    continue
                lid, uu = e.add_locus_is(data), 1.0
            else:
                A = sim.simulate_sw(rng, g0, 1.0)
                data, lid, uu = g0.n, e.add_locus_sw(g0.n), 1.0
            break
        e.locus_set_model(lid, mid, samp)
        loci.append((kind, data))
        lids.append(lid)
        base.append(g0)
        As.append(A)
        u.append(uu)
    Lc = len(kinds)
    first = e.alloc_slots([lids[l] for _ in range(C) for l in range(Lc)])
    gens = [base[l] for _ in range(C) for l in range(Lc)]
    e.write_slots(first, gens, A=[As[l] if As[l] is not None else np.zeros(2 * base[l].n - 1,
                                                                             np.int32)
                                  for _ in range(C) for l in range(Lc)])
    e.write_pops(first, gens)
    sl = {l: e.sumlogk_from_slot(lids[l], first + l) for l, k in enumerate(kinds) if k == "is"}
    kappa = np.array([4.0 if k == "hky" else 1.0 for k in kinds])
    e.mc3_setup(C, Lc, first, np.tile(u, C), np.tile(kappa, C), 0.8 ** np.arange(C), seed=seed,
                precision=L.PREC_FP64, scale_width=0.3)
    e.mc3_set_prior(np.tile(ts, (C, 1)), 5.0, 2)
    return dict(e=e, first=first, model=model, samp=samp, loci=loci, u=np.array(u), kappa=kappa,
                sumlogk=sl, base=base, C=C, Lc=Lc, with_mig=with_mig)


def _state_matches_oracle(S, u, kappa):
    """Every chain's genealogies (topology, times, labels, alleles as the
    device holds them) against the oracle; u, kappa [chains, loci].
    Returns how many genealogies differ in topology from the start."""
    e, C, Lc, first = S["e"], S["C"], S["Lc"], S["first"]
    pdg, _, _, _ = e.mc3_read()
    probg, tv = e.mc3_read_prior()
    up0, up1, dn, tm, rt, A = e.read_slots(first, C * Lc)
    pops = e.read_pops(first, C * Lc)
    off, changed = 0, 0
    for c in range(C):
        ws = []
        for l in range(Lc):
            g0 = S["base"][l]
            ne = 2 * g0.n - 1
            sl = slice(off, off + ne)
            g = sim.Genealogy(g0.n, up0[sl].copy(), up1[sl].copy(), dn[sl].copy(), tm[sl].copy(),
                              int(rt[c * Lc + l]), pops[sl].copy(), [[] for _ in range(ne)])
            if S.get("with_mig"):
                me, mt, mo = e.read_mig_events(first + c * Lc + l)
                for x in range(len(me)):
                    g.mig[int(me[x])].append((float(mt[x]), int(mo[x])))
            a = A[sl].copy()
            off += ne
            changed += int(not np.array_equal(g.up0, g0.up0) or g.root != g0.root)
            assert (g.pop[:g.n] == g0.pop[:g.n]).all()  # gene copies keep their populations
            kind, data = S["loci"][l]
            if kind == "hky":
                ref = oracle.likelihood_hky(g, data, u[c, l], kappa[c, l])[0]
            elif kind == "is":
                ref, st, _ = oracle.likelihood_is(g, data, u[c, l], S["sumlogk"][l])
                assert st == 0
            else:
                assert (a[:g.n] == S["tipA"][l]).all()  # the data never move
                ref = oracle.likelihood_sw(g, a, u[c, l])[0]
            assert abs(pdg[c, l] - ref) <= 1e-10 * abs(ref), (c, l, kind, pdg[c, l], ref)
            ws.append(oracle.treeweight(g, S["model"], tv[c], S["samp"]))  # raises on bad labels
        rp, _, _ = oracle.integrate_tree_prob(ws, S["model"])
        assert abs(probg[c] - rp) <= 1e-9 * abs(rp), (c, probg[c], rp)
    return changed


@pytest.mark.parametrize("with_mig", [False, True], ids=["no_migration", "migration_nw"])
def test_topology_sweeps_keep_the_state_consistent(with_mig):
    """HKY, infinite-sites and stepwise loci; with migration the moves draw
    migration paths (and move stepwise alleles) and NW is drawn against
    RY1."""
    S = _two_pop_data(41, with_mig=with_mig)
    nw = L.MC3_NW if with_mig else 0
    e, C = S["e"], S["C"]
    _, _, _, _, _, A0 = e.read_slots(S["first"], S["Lc"])
    off, S["tipA"] = 0, {}
    for l in range(S["Lc"]):
        n = S["base"][l].n
        S["tipA"][l] = A0[off:off + n].copy()
        off += 2 * n - 1
    e.mc3_run(20, TOPO | L.MC3_SCALE | L.MC3_RY1 | nw)
    changed = _state_matches_oracle(S, np.tile(S["u"], (C, 1)), np.tile(S["kappa"], (C, 1)))
    e.mc3_run(20, L.MC3_TOPOLOGY | L.MC3_SCALE | L.MC3_RY1 | L.MC3_UPDATE_U | nw)
    u, kappa = e.mc3_read_rates()
    changed += _state_matches_oracle(S, u, kappa)
    _, _, stats, _ = e.mc3_read()
    assert stats[1] > 0 and changed > 0
    print("genealogies with a new topology:", changed)
    e.close()


def test_graph_equals_eager_and_migration_models_run():
    out = []
    for flags in (TOPO | L.MC3_SCALE | L.MC3_RY1, L.MC3_TOPOLOGY | L.MC3_SCALE | L.MC3_RY1):
        S = _two_pop_data(42, C=4)
        S["e"].mc3_run(17, flags)
        pdg, beta, stats, _ = S["e"].mc3_read()
        probg, tv = S["e"].mc3_read_prior()
        up0, _, _, tm, rt, A = S["e"].read_slots(S["first"], S["C"] * S["Lc"])
        out.append((pdg, beta, stats, probg, tv, up0, tm, rt, A))
        S["e"].close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
    # a model with migration parameters: the move draws migration paths
    # (tested in test_gpu_topo_mig.py); a single chain runs and keeps a
    # genealogy whose labels fit its split time
    rng = np.random.default_rng(3)
    pt, theta, ts, mig = sim.im_model_two_pop()
    model = immodel.build(pt)
    e = Engine(0)
    mid = e.model_add(model)
    g = sim.simulate_genealogy(rng, [3, 3], theta, ts, mig, pt)
    lid = e.add_locus_hky(sim.encode_hky(sim.simulate_hky_sequences(rng, g, 50, [0.25] * 4, 2.0,
                                                                    1.0)))
    e.locus_set_model(lid, mid, [3, 3])
    first = e.alloc_slots([lid])
    e.write_slots(first, [g])
    e.write_pops(first, [g])
    e.mc3_setup(1, 1, first, 1.0, 2.0, [1.0])
    e.mc3_set_prior(np.asarray(ts, float).reshape(1, -1), 5.0, 2)
    e.mc3_run(20, TOPO)
    assert e.mc3_read()[3] == 20
    e.close()
