"""HKY site blocks of 64 and 128 patterns (one or two patterns per lane,
ima_set_site_block / ima_mc3_site_block).  Both block sizes give the
oracle's likelihoods (FP64 1e-10, FP32 1e-6) on loci whose pattern counts
fall on and around the block edges (1, 63, 64, 65, 127, 128, 129, ...),
and agree with each other to 1e-12 in FP64 (the per-pattern terms are the
same; only the order of the FP64 sums differs).  The chains' sweeps with
128-pattern blocks (scoring, rubber-band re-scoring, mutation-rate rounds)
keep every locus' likelihood equal to the oracle's on the chain's current
genealogy and rates, and the tuner's choice is one of the two sizes."""
import numpy as np
import pytest

import oracle
from ima2p_amd import _lib as L
from ima2p_amd import sim
from tests import datasets as D
from tests.test_gpu_mc3_prior import _three_pop_chains

pytestmark = pytest.mark.gpu


def _edge_loci(rng):
    """Loci whose pattern counts hit the block edges."""
    out = []
    for k, nsites in enumerate((3, 40, 90, 150, 240, 400, 700, 1500, 3000)):
        g, loc = D.hky_case(500 + k, (int(rng.integers(3, 20)), int(rng.integers(3, 20))),
                            nsites)
        out.append((g, loc))
    return out


def test_both_block_sizes_equal_the_oracle(engine):
    rng = np.random.default_rng(13)
    cases = _edge_loci(rng)
    S = [loc.seq.shape[1] for _, loc in cases]
    print("pattern counts", S)
    loci = [engine.add_locus_hky(loc) for _, loc in cases]
    first = engine.alloc_slots(loci)
    engine.write_slots(first, [g for g, _ in cases])
    us = rng.uniform(0.5, 3.0, len(cases))
    ks = rng.uniform(1.0, 8.0, len(cases))
    items = engine.items(np.arange(first, first + len(cases)), us, ks)
    res = {}
    for bs in (64, 128):
        engine.set_site_block(bs)
        res[bs, 64], st = engine.score(items, L.PREC_FP64)
        assert (st == 0).all()
        res[bs, 32], _ = engine.score(items, L.PREC_FP32)
        for i, (g, loc) in enumerate(cases):
            ref, _ = oracle.likelihood_hky(g, loc, us[i], ks[i])
            assert abs(res[bs, 64][i] - ref) <= 1e-10 * abs(ref), (bs, i, S[i])
            assert abs(res[bs, 32][i] - ref) <= 1e-6 * abs(ref), (bs, i, S[i])
    np.testing.assert_allclose(res[128, 64], res[64, 64], rtol=1e-12)
    engine.set_site_block(0)


@pytest.mark.parametrize("prec,tol", [(L.PREC_FP64, 1e-9), (L.PREC_FP32, 1e-6)])
def test_sweeps_with_128_pattern_blocks_keep_the_state(prec, tol):
    C, Lc = 4, 3
    # 300-site loci: about 100-200 patterns, two or three 64-pattern blocks
    e, model, samp, locs, gens, first, u, ts = _three_pop_chains(C, Lc, 91, prec,
                                                                 site_block=128)
    assert e.mc3_site_block() == 128
    e.mc3_run(20, L.MC3_GRAPH | L.MC3_SCALE | L.MC3_RY1 | L.MC3_UPDATE_U)
    pdg, _, _, _ = e.mc3_read()
    urate, kap = e.mc3_read_rates()
    up0, up1, dn, tm, rt, _ = e.read_slots(first, C * Lc)
    off = 0
    for c in range(C):
        for l in range(Lc):
            g = gens[c * Lc + l].copy()
            ne = 2 * g.n - 1
            g.time = tm[off:off + ne].copy()
            g.up0, g.up1, g.down = up0[off:off + ne], up1[off:off + ne], dn[off:off + ne]
            g.root = int(rt[c * Lc + l])
            off += ne
            ref, _ = oracle.likelihood_hky(g, locs[l], urate[c, l], kap[c, l])
            assert abs(pdg[c, l] - ref) <= tol * abs(ref), (c, l)
    e.close()


def test_tuner_picks_a_block_size():
    C, Lc = 8, 3
    e, *_ = _three_pop_chains(C, Lc, 92, L.PREC_FP32)
    best, ms = e.mc3_tune(4, L.MC3_GRAPH | L.MC3_SCALE)
    assert e.mc3_site_block() in (64, 128)
    e.mc3_run(4, L.MC3_GRAPH | L.MC3_SCALE)
    e.close()
