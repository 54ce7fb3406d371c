"""Host-side logic of the serving path and the workload builders (no GPU):
the flattened analysis has the array lengths the C ABI reads, the
population models of 2..4 populations are well formed, the uneven batch has
the stated composition, and the genealogy move used by the stepwise allele
update leaves a valid genealogy with the clades it should have."""
import numpy as np
import pytest

from ima2p_amd import immodel, sim, workloads


@pytest.mark.parametrize("P", [2, 3, 4])
def test_random_models_and_pools(P):
    rng = np.random.default_rng(P)
    pool = workloads.serving_pool(rng, P, size=12)
    m = pool.model
    assert m.npops == P and len(pool.tsplit) == P - 1
    assert (np.diff(pool.tsplit) >= 0).all()
    # every population of the tree is in exactly the periods it spans
    for q in range(2 * P - 1):
        periods = [k for k in range(P) if q in m.plist[k][: m.npresent[k]]]
        assert periods == list(range(periods[0], periods[-1] + 1))
    for kind, data, g, A, u, samp in pool.entries:
        g.check()
        assert sum(samp) == g.n and len(samp) == P
        if kind == "sw":
            assert A is not None and len(A) == 2 * g.n - 1
        if kind == "is":
            assert data.seq.shape == (g.n, data.seq.shape[1]) and data.seq.shape[1] > 0


def test_flatten_lengths():
    rng = np.random.default_rng(9)
    pool = workloads.serving_pool(rng, 3, size=10)
    spec = workloads.serving_analysis(rng, pool, 7, 5, nsweeps=3, seed=4)
    s, keep = spec.flatten()
    C, Lc = spec.nchains, spec.nloci
    n = [spec.chains[0][l].n for l in range(Lc)]
    nodes = C * sum(2 * x - 1 for x in n)

    def arr(ptr, count):
        return np.ctypeslib.as_array(ptr, shape=(count,))

    assert s.nloci == Lc and s.nchains == C and s.npops == 3
    assert len(arr(s.up0, nodes)) == nodes and len(arr(s.time, nodes)) == nodes
    nm = arr(s.nmigev, C * Lc)
    # the migrations of every slot are sorted by edge, then time
    me, mt = arr(s.mig_edge, int(nm.sum())), arr(s.mig_time, int(nm.sum()))
    off = 0
    for k in nm:
        e, t = me[off:off + k], mt[off:off + k]
        assert all((e[i], t[i]) <= (e[i + 1], t[i + 1]) for i in range(k - 1))
        off += k
    toff = arr(s.theta_off, spec.model.npsize + 1)
    assert toff[-1] == sum(len(w) for w in spec.model.theta_wp)
    assert arr(s.tvals, C * 2).reshape(C, 2).tolist() == spec.tvals.tolist()


def test_uneven_composition():
    spec = workloads.uneven_mixed(nchains=2)
    kinds = [k for k, _ in spec.loci]
    assert kinds.count("hky") == 150 and kinds.count("is") == 100 and kinds.count("sw") == 50
    ns = [spec.chains[0][l].n for l in range(spec.nloci)]
    assert min(ns) >= 4 and max(ns) <= 60
    for (k, d), samp in zip(spec.loci, spec.samppop):
        assert len(samp) == 2 and min(samp) >= 1
        if k == "hky":
            assert 200 <= d.totsites <= 3000


@pytest.mark.parametrize("seed", range(5))
def test_move_edge_keeps_a_valid_genealogy(seed):
    rng = np.random.default_rng(seed)
    pt = sim.PopTree(1, [])
    for _ in range(30):
        g = sim.simulate_genealogy(rng, [int(rng.integers(3, 15))], [4.0], [], np.zeros((1, 1)),
                                   pt)
        A = sim.simulate_sw(rng, g, 1.0)
        h, mv = sim.random_sw_move(rng, g, A, np.zeros(2 * g.n - 1), 1.0, -50, 100)
        h.check()
        assert h.down[mv.edge] == mv.downedge and h.down[mv.newsisedge] == mv.downedge
        # the moved edge keeps its subtree; every tip is still below the root
        clades = h.clades()
        assert clades[mv.edge] == g.clades()[mv.edge]
        assert clades[h.root] == (1 << g.n) - 1


def test_ladder_model_matches_three_population_tree():
    pt, theta, ts, mig = sim.im_model_random(np.random.default_rng(1), 3)
    assert pt.joins == sim.PopTree.three_pop_01_2().joins
    m = immodel.build(pt)
    assert m.npsize == 5 and m.nmig > 0
