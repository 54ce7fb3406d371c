"""Uncoupled analyses in one batched sweep (ima_mc3_set_analyses): swaps
stay inside each analysis, every analysis keeps its own heating ladder, and
one analysis of all chains is the plain sweep bit for bit."""
import numpy as np
import pytest

from ima2p_amd import _lib as L
from ima2p_amd import immodel, workloads
from ima2p_amd.engine import Engine, EngineError
from ima2p_amd.serve import AnalysisSpec

pytestmark = pytest.mark.gpu

FLAGS = L.MC3_GRAPH | L.MC3_TOPOLOGY | L.MC3_SCALE | L.MC3_RY1 | L.MC3_UPDATE_U | L.MC3_NW


def _engine(analyses, ladder):
    wl = workloads.two_pop_is()
    model = immodel.build(wl.poptree, qmax=20.0, mmax=5.0)
    C = wl.nchains * analyses
    spec = AnalysisSpec(loci=[("is", d) for d in wl.loci], model=model,
                        samppop=[wl.samppop] * wl.nloci, chains=list(wl.chains) * analyses,
                        u=np.tile(wl.u, (C, 1)), kappa=np.ones((C, wl.nloci)),
                        beta=np.tile(ladder, analyses), tvals=np.tile(wl.tsplit, (C, 1)),
                        tmax=wl.tmax, seed=5, precision=L.PREC_FP64)
    e = Engine(0)
    spec.load(e)
    return e, wl


def test_one_analysis_of_all_chains_is_the_plain_sweep():
    out = []
    for split in (False, True):
        e, wl = _engine(2, workloads.two_pop_is().beta)
        if split:
            e.mc3_set_analyses(2 * wl.nchains)
        e.mc3_run(50, FLAGS)
        pdg, beta, stats, _ = e.mc3_read()
        out.append((pdg, beta, stats))
        e.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


def test_swaps_stay_inside_each_analysis():
    ladder = np.array([1.0, 0.9, 0.8, 0.7])
    A, sweeps = 3, 300
    e, wl = _engine(A, ladder)
    e.mc3_set_analyses(wl.nchains)
    e.mc3_run(sweeps, FLAGS)
    _, beta, stats, _ = e.mc3_read()
    e.close()
    for k in range(A):
        np.testing.assert_array_equal(np.sort(beta[k * 4:(k + 1) * 4]), np.sort(ladder))
    assert stats[4] == sweeps * A  # one try per analysis per sweep
    assert 0 < stats[5] < stats[4]
    # the analyses drew their own swaps: their ladders are not all in step
    orders = {tuple(np.argsort(beta[k * 4:(k + 1) * 4])) for k in range(A)}
    print("final orders", orders)


def test_chain_count_must_split():
    e, wl = _engine(3, workloads.two_pop_is().beta)
    with pytest.raises(EngineError) as ei:
        e.mc3_set_analyses(5)
    assert ei.value.code == L.ERR_ARG
    e.close()
