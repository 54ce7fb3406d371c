"""ima_mc3_decide: decisions for proposals written by the host, as IMa2p's
own updategenealogy makes them (update_gtree.cpp:723-966), scored in one
batched launch and decided per chain in locus order with the speculative
groups of the device sweep.

* With the host's uniforms and nonzero Hastings terms, every decision
  equals the oracle's accept step (update_gtree.cpp:900-961, the chain's
  loci in order with probg from integrate_tree_prob over the all-loci
  weights) bit for bit, over two rounds of mixed proposals: unchanged
  genealogies, new topologies with new migrations (HKY, and infinite-sites
  ones that the data mostly reject), and stepwise loci.
* Afterwards each chain's slots hold exactly the genealogies the decisions
  left; every stored log-likelihood and probg equal the oracle's on them
  (1e-10 / 1e-9), also after device sweeps on the new topologies.
* Without the prior the loci decide independently:
  U < min(1, exp(beta gbeta dpdg + h)).
"""
import numpy as np
import pytest

import oracle
from ima2p_amd import _lib as L
from ima2p_amd import immodel, sim
from ima2p_amd.engine import Engine

pytestmark = pytest.mark.gpu

KINDS = ["hky", "is", "hky", "is", "sw", "hky"]


def _setup(seed, C=8, prior=True):
    rng = np.random.default_rng(seed)
    pt, theta, ts, mig = sim.im_model_two_pop(t=0.8, m12=0.4, m21=0.6)
    model = immodel.build(pt, qmax=15.0, mmax=4.0)
    samp = [4, 5]
    e = Engine(0)
    mid = e.model_add(model) if prior else None
    loci, lids, base, As, u = [], [], [], [], []
    for kind in KINDS:
        while True:
            g0 = sim.simulate_genealogy(rng, samp, theta, ts, mig, pt)
            A = None
            if kind == "hky":
                data = sim.encode_hky(sim.simulate_hky_sequences(rng, g0, 150,
                                                                 rng.dirichlet([8] * 4), 4.0, 1.0))
                lid, uu = e.add_locus_hky(data), 1.0
            elif kind == "is":
                data = sim.simulate_is(rng, g0, 1.5)
                if data.seq.shape[1] == 0:
                    continue
                lid, uu = e.add_locus_is(data), 1.5
            else:
                A = sim.simulate_sw(rng, g0, 1.5)
                data, lid, uu = g0.n, e.add_locus_sw(g0.n), 1.5
            break
        if prior:
            e.locus_set_model(lid, mid, samp)
        loci.append((kind, data))
        lids.append(lid)
        base.append(g0)
        As.append(A)
        u.append(uu)
    Lc = len(KINDS)
    first = e.alloc_slots([lids[l] for _ in range(C) for l in range(Lc)])
    prop = e.alloc_slots([lids[l] for _ in range(C) for l in range(Lc)])
    cur = [[base[l].copy() for l in range(Lc)] for _ in range(C)]
    e.write_slots(first, [g for row in cur for g in row], A=[a if a is not None else
                                                             np.zeros(2 * base[l].n - 1, np.int32)
                                                             for _ in range(C)
                                                             for l, a in enumerate(As)])
    if prior:
        e.write_pops(first, [g for row in cur for g in row])
    sl = {}
    for l, (kind, data) in enumerate(loci):
        if kind == "is":
            sl[l] = e.sumlogk_from_slot(lids[l], first + l)
    kappa = np.array([4.0 if k == "hky" else 1.0 for k in KINDS])
    betas = np.array([0.85 ** c for c in range(C)])
    e.mc3_setup(C, Lc, first, np.tile(u, C), np.tile(kappa, C), betas, seed=seed,
                precision=L.PREC_FP64, scale_width=0.3)
    if prior:
        e.mc3_set_prior(np.tile(ts, (C, 1)), 5.0, 2)
    return dict(e=e, rng=rng, model=model, samp=samp, loci=loci, As=As, u=np.array(u),
                kappa=kappa, first=first, prop=prop, cur=cur, ts=np.asarray(ts, float),
                sumlogk=sl, betas=betas, pt=pt, theta=theta, mig=mig, C=C, Lc=Lc)


def _loglik(S, l, g):
    kind, data = S["loci"][l]
    if kind == "hky":
        return oracle.likelihood_hky(g, data, S["u"][l], S["kappa"][l])[0], True
    if kind == "is":
        ll, st, _ = oracle.likelihood_is(g, data, S["u"][l], S["sumlogk"][l])
        return ll, st == 0
    return oracle.likelihood_sw(g, S["As"][l], S["u"][l])[0], True


def _proposals(S):
    rng, out = S["rng"], []
    for c in range(S["C"]):
        for l, kind in enumerate(KINDS):
            g = S["cur"][c][l]
            if kind == "sw" or rng.random() < 0.35:
                out.append(g.copy())  # unchanged: decided on h alone
            else:
                out.append(sim.simulate_genealogy(rng, S["samp"], S["theta"], S["ts"], S["mig"],
                                                  S["pt"]))
    return out


def _write(S, props):
    e = S["e"]
    A = [S["As"][l] if S["As"][l] is not None else np.zeros(2 * props[i].n - 1, np.int32)
         for i, l in enumerate([l for _ in range(S["C"]) for l in range(S["Lc"])])]
    e.write_slots(S["prop"], props, A=A)
    e.write_pops(S["prop"], props)


def _check_state(S, tol=1e-10):
    """Slots hold S['cur'] (topology exact; times from the device), pdg and
    probg equal the oracle's on them."""
    e, C, Lc = S["e"], S["C"], S["Lc"]
    pdg, _, _, _ = e.mc3_read()
    probg, tv = e.mc3_read_prior()
    up0, up1, dn, tm, rt, _ = e.read_slots(S["first"], C * Lc)
    off = 0
    for c in range(C):
        ws = []
        for l in range(Lc):
            g = S["cur"][c][l]
            ne = 2 * g.n - 1
            assert (up0[off:off + ne] == g.up0).all() and (up1[off:off + ne] == g.up1).all()
            assert (dn[off:off + ne] == g.down).all() and rt[c * Lc + l] == g.root
            h = g.copy()
            h.time = tm[off:off + ne].copy()
            nm = sum(len(x) for x in g.mig)
            mt = e.read_mig_times(S["first"] + c * Lc + l, nm)
            k = 0
            for edge in range(ne):
                lst = sorted(g.mig[edge])
                h.mig[edge] = [(float(mt[k + j]), q) for j, (_, q) in enumerate(lst)]
                k += len(lst)
            off += ne
            ref, ok = _loglik(S, l, h)
            assert ok and abs(pdg[c, l] - ref) <= tol * abs(ref), (c, l, pdg[c, l], ref)
            ws.append(oracle.treeweight(h, S["model"], tv[c], S["samp"]))
        rp, _, _ = oracle.integrate_tree_prob(ws, S["model"])
        assert abs(probg[c] - rp) <= 1e-9 * abs(rp), (c, probg[c], rp)


def test_decisions_equal_the_oracle_bit_for_bit():
    S = _setup(101)
    e, C, Lc = S["e"], S["C"], S["Lc"]
    total = np.zeros(2, int)
    for rnd in range(2):
        props = _proposals(S)
        _write(S, props)
        h = S["rng"].normal(0.0, 3.0, C * Lc)
        U = S["rng"].uniform(0.0, 1.0, C * Lc)
        gb = 1.0 if rnd == 0 else 0.7
        dec = e.mc3_decide(S["prop"], h, U, gbeta=gb)
        print(f"round {rnd}: {int(dec.sum())} of {C * Lc} accepted", flush=True)
        for c in range(C):
            pdg = [_loglik(S, l, S["cur"][c][l])[0] for l in range(Lc)]
            w = [oracle.treeweight(S["cur"][c][l], S["model"], S["ts"], S["samp"])
                 for l in range(Lc)]
            pr = [props[c * Lc + l] for l in range(Lc)]
            nl = [_loglik(S, l, pr[l]) for l in range(Lc)]
            nw = [oracle.treeweight(pr[l], S["model"], S["ts"], S["samp"]) for l in range(Lc)]
            ref, *_ = oracle.updategenealogy_decisions(
                pdg, w, [x[0] for x in nl], nw, [x[1] for x in nl], h[c * Lc:(c + 1) * Lc],
                U[c * Lc:(c + 1) * Lc], S["betas"][c], S["model"], gbeta=gb)
            assert list(dec[c].astype(bool)) == ref, (rnd, c, dec[c], ref)
            for l in range(Lc):
                if dec[c, l]:
                    S["cur"][c][l] = pr[l]
        total += [int(dec.sum()), C * Lc - int(dec.sum())]
        _check_state(S)
    assert total[0] > 0 and total[1] > 0
    topo = sum(1 for c in range(C) for l in range(Lc)
               if S["cur"][c][l] is not None and KINDS[l] == "hky")
    assert topo > 0
    # device sweeps continue from the accepted topologies
    e.mc3_run(12, L.MC3_GRAPH | L.MC3_SCALE | L.MC3_RY1)
    _check_state(S, tol=1e-9)
    e.close()


def test_engine_stream_and_no_hastings_keep_the_state_consistent():
    S = _setup(202)
    e = S["e"]
    props = _proposals(S)
    _write(S, props)
    dec = e.mc3_decide(S["prop"])
    for c in range(S["C"]):
        for l in range(S["Lc"]):
            if dec[c, l]:
                S["cur"][c][l] = props[c * S["Lc"] + l]
    _check_state(S)
    e.close()


def test_without_prior_loci_decide_independently():
    S = _setup(303, prior=False)
    e, C, Lc = S["e"], S["C"], S["Lc"]
    props = _proposals(S)
    e.write_slots(S["prop"], props, A=[S["As"][l] if S["As"][l] is not None else
                                       np.zeros(2 * props[i].n - 1, np.int32)
                                       for i, l in enumerate([l for _ in range(C)
                                                              for l in range(Lc)])])
    h = S["rng"].normal(0.0, 2.0, C * Lc)
    U = S["rng"].uniform(0.0, 1.0, C * Lc)
    pdg0, *_ = e.mc3_read()
    dec = e.mc3_decide(S["prop"], h, U, gbeta=0.9)
    for c in range(C):
        for l in range(Lc):
            i = c * Lc + l
            npd, ok = _loglik(S, l, props[i])
            x = S["betas"][c] * 0.9 * (npd - pdg0[c, l]) + h[i]
            assert bool(dec[c, l]) == (ok and U[i] < min(1.0, np.exp(x))), (c, l)
    pdg, *_ = e.mc3_read()
    for c in range(C):
        for l in range(Lc):
            g = props[c * Lc + l] if dec[c, l] else S["cur"][c][l]
            ref, _ = _loglik(S, l, g)
            assert abs(pdg[c, l] - ref) <= 1e-10 * abs(ref)
    e.close()
