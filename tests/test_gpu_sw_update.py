"""Device finishSWupdateA (update_gtree_common.cpp:2219-2362) against the
oracle: a batch of random genealogy moves at stepwise loci of 3..40 gene
copies, every move in one launch.  The proposed allele is bit-exact (and
stored in the slot), the log-likelihood change, Hastings term and the four
new edge terms within 1e-12 relative.  A move whose slot does not hold its
edges under the freed node is refused and nothing changes."""
import numpy as np
import pytest

import oracle
from ima2p_amd import sim
from ima2p_amd.engine import Engine, EngineError

pytestmark = pytest.mark.gpu


def _batch(rng, count):
    cases = []
    for _ in range(count):
        n = int(rng.integers(3, 41))
        u = float(rng.uniform(0.3, 4.0))
        pt = sim.PopTree(1, [])
        g = sim.simulate_genealogy(rng, [n], [4.0], [], np.zeros((1, 1)), pt)
        A = sim.simulate_sw(rng, g, u)
        _, dl = oracle.likelihood_sw(g, A, u)
        h, mv = sim.random_sw_move(rng, g, A, dl, u, int(A.min()) - 3, int(A.max()) + 3)
        cases.append((h, A, mv))
    return cases


def _load(e, cases):
    lids = [e.add_locus_sw(h.n) for h, _, _ in cases]
    first = e.alloc_slots(lids)
    for i, (h, A, mv) in enumerate(cases):
        e.write_slots(first + i, [h], A=[A])
        mv.slot = first + i
    return first


@pytest.mark.parametrize("calc", [True, False])
def test_sw_update_matches_oracle(calc):
    rng = np.random.default_rng(11 if calc else 12)
    cases = _batch(rng, 300)
    e = Engine(0)
    first = _load(e, cases)
    dl, at, na, ed = e.sw_update_A([mv for _, _, mv in cases], calc=calc)
    _, _, _, _, _, Adev = e.read_slots(first, len(cases))
    off = 0
    for i, (h, A, mv) in enumerate(cases):
        rdl, rat, rna, red, A2 = oracle.finish_sw_update_A(h, A, mv, calc=calc)
        assert na[i] == rna, i
        ne = 2 * h.n - 1
        assert (Adev[off:off + ne] == A2).all(), i
        off += ne
        for got, want in [(dl[i], rdl), (at[i], rat)] + list(zip(ed[i], red)):
            assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (i, got, want)
    e.close()


def test_bad_move_is_refused_and_changes_nothing():
    rng = np.random.default_rng(13)
    cases = _batch(rng, 4)
    e = Engine(0)
    first = _load(e, cases)
    _, _, _, _, _, A0 = e.read_slots(first, len(cases))
    moves = [mv for _, _, mv in cases]
    bad = moves[2]
    h = cases[2][0]
    # a downedge that is not edge's parent in the slot
    wrong = next(x for x in range(2 * h.n - 1)
                 if x not in (bad.edge, bad.downedge, bad.sisedge, bad.newsisedge)
                 and h.down[bad.edge] != x)
    bad.downedge = wrong
    with pytest.raises(EngineError):
        e.sw_update_A(moves)
    _, _, _, _, _, A1 = e.read_slots(first, len(cases))
    assert (A0 == A1).all()
    e.close()
