"""Request loop (ima_server_*): L-mode queries and IM analyses arriving
over time share one device with continuous batching.

* Every L-mode query's values equal ima_lmode_jointp's on the same table
  and points, bit for bit, alone and when its points share launches with
  other queries' (large queries split across launches): the table's fixed
  chunking fixes each point's summation order.
* Every analysis' chains end in the state a direct ima_mc3_run of the
  same set-up reaches (log-likelihoods, probg, heats, counters), bit for
  bit, alone and in a mix of analyses and L-mode traffic, run in quanta of
  sweeps: the sweep's random numbers depend on the sweep index only.
* Admission by footprint: under a budget that holds one analysis the
  others wait and run in turn; a request larger than the budget is
  refused with IMA_ERR_SIZE.
"""
import numpy as np
import pytest

from ima2p_amd import _lib as L
from ima2p_amd import sim, workloads
from ima2p_amd.engine import Engine, EngineError
from ima2p_amd.serve import REQ_ANALYSIS, REQ_DONE, REQ_LMODE, REQ_RUNNING, REQ_WAITING, Server

pytestmark = pytest.mark.gpu

# the sweep a served analysis runs (the reference's step)
SERVED = (L.MC3_GRAPH | L.MC3_TOPOLOGY | L.MC3_SCALE | L.MC3_RY1 | L.MC3_NW
          | L.MC3_UPDATE_U)


def _direct(spec):
    """The analysis run straight through the engine, all sweeps at once."""
    e = Engine(0)
    spec.load(e)
    e.mc3_run(spec.nsweeps, SERVED)
    pdg, beta, stats, _ = e.mc3_read()
    probg, _ = e.mc3_read_prior()
    e.close()
    return pdg, probg, beta, stats


def _analyses(seed=21):
    rng = np.random.default_rng(seed)
    p2 = workloads.serving_pool(rng, 2, size=10)
    p3 = workloads.serving_pool(rng, 3, size=10)
    return [workloads.serving_analysis(rng, p3, 6, 8, nsweeps=30, seed=7),
            workloads.serving_analysis(rng, p2, 9, 12, nsweeps=23, seed=8,
                                       precision=L.PREC_FP64)]


def _queries(seed=22):
    rng = np.random.default_rng(seed)
    tabs = [(sim.lmode_table(rng, 3000), 3, 2, 2), (sim.lmode_table(rng, 20000), 3, 2, 2)]
    lo, hi = [0.5, 0.5, 0.5, 0.01, 0.01], [10, 10, 10, 2, 2]
    q = []
    for t, npts, mode, mt, pr in [(0, 5, L.LMODE_EXACT, 0, None), (1, 10000, L.LMODE_EXACT, 0, None),
                                  (0, 300, L.LMODE_REF_COMPAT, 0, None), (1, 1, L.LMODE_EXACT, 0, None),
                                  (1, 700, L.LMODE_EXACT, 0, (0, 3)), (0, 4096, L.LMODE_EXACT, 0, None)]:
        q.append((t, rng.uniform(lo, hi, (npts, 5)), mode, mt, pr))
    return tabs, q


def _direct_lmode(tabs, queries):
    e = Engine(0)
    tids = [e.lmode_table(*t) for t in tabs]
    out = [e.jointp(tids[t], x, mode=mode, modeltype=mt, prange=pr) for t, x, mode, mt, pr in queries]
    e.close()
    return out


def _serve(analyses, tabs, queries, budget=0.0, quantum=7, order=None):
    srv = Server(0, budget, quantum)
    tids = [srv.table(*t) for t in tabs]
    items = [("a", a) for a in analyses] + [("q", q) for q in queries]
    if order is not None:
        items = [items[i] for i in order]
    ids = []
    for kind, it in items:
        if kind == "a":
            ids.append(("a", it, srv.submit_analysis(it)))
        else:
            t, x, mode, mt, pr = it
            ids.append(("q", it, srv.submit_lmode(tids[t], x, mode, mt, pr)))
        srv.step()  # arrivals interleave with progress
    srv.run(timeout_s=300)
    res = {}
    for kind, it, rid in ids:
        st, lat, _ = srv.status(rid)
        assert st == REQ_DONE and lat >= 0
        res[id(it)] = srv.lmode_result(rid) if kind == "q" else srv.analysis_result(rid)
    lat = srv.latency()
    assert lat["count"] == len(ids)
    srv.close()
    return res


def test_lmode_alone_and_mixed_equal_direct():
    tabs, queries = _queries()
    ref = _direct_lmode(tabs, queries)
    for subset in ([0], [1], [5], list(range(len(queries)))):
        qs = [queries[i] for i in subset]
        got = _serve([], tabs, qs)
        for i, q in zip(subset, qs):
            np.testing.assert_array_equal(got[id(q)][0], ref[i][0])
            np.testing.assert_array_equal(got[id(q)][1], ref[i][1])


def test_analyses_alone_and_mixed_equal_direct():
    specs = _analyses()
    tabs, queries = _queries()
    ref = [_direct(s) for s in specs]
    assert ref[0][3][1] > 0  # genealogy updates were accepted
    alone = [_serve([s], [], [])[id(s)] for s in specs]
    mixed = _serve(specs, tabs, queries, order=[2, 0, 3, 4, 1, 5, 6, 7])
    for s, r, a in zip(specs, ref, alone):
        m = mixed[id(s)]
        for x, y, z in zip(r, a, m):
            np.testing.assert_array_equal(np.ravel(x), np.ravel(y))
            np.testing.assert_array_equal(np.ravel(x), np.ravel(z))


def test_admission_by_footprint():
    rng = np.random.default_rng(31)
    pool = workloads.serving_pool(rng, 2, size=8)
    specs = [workloads.serving_analysis(rng, pool, 5, 8, nsweeps=12, seed=40 + i) for i in range(3)]
    probe = Server(0, 0.0, 4)
    rid = probe.submit_analysis(specs[0])
    _, _, est = probe.status(rid)
    probe.close()
    assert est > 0
    srv = Server(0, 1.5 * est, 4)
    ids = [srv.submit_analysis(s) for s in specs]
    assert srv.usage()["waiting"] == 3
    srv.step()
    u = srv.usage()
    assert u["running"] == 1 and u["waiting"] == 2, u
    assert srv.status(ids[0])[0] == REQ_RUNNING and srv.status(ids[2])[0] == REQ_WAITING
    assert u["used_bytes"] <= u["budget_bytes"]
    srv.run(timeout_s=300)
    for s, rid in zip(specs, ids):
        assert srv.status(rid)[0] == REQ_DONE
        got = srv.analysis_result(rid)
        ref = _direct(s)
        for x, y in zip(ref, got):
            np.testing.assert_array_equal(np.ravel(x), np.ravel(y))
    assert srv.usage()["used_bytes"] == 0
    lat = srv.latency(REQ_ANALYSIS)
    assert lat["count"] == 3 and lat["p99_ms"] >= lat["p50_ms"] > 0
    # larger than the whole budget: refused at submission
    small = Server(0, 0.5 * est, 4)
    with pytest.raises(EngineError) as ei:
        small.submit_analysis(specs[0])
    assert ei.value.code == L.ERR_SIZE
    tab = sim.lmode_table(rng, 1000)
    tid = small.table(tab, 3, 2, 2)
    qid = small.submit_lmode(tid, rng.uniform(0.5, 5, (10, 5)))
    small.run(60)
    assert small.status(qid)[0] == REQ_DONE and small.latency(REQ_LMODE)["count"] == 1
    small.close()
    srv.close()


def test_contexts_return_their_memory():
    """Building and destroying analyses' contexts (what the server does per
    request) leaves the device's free memory where it was."""
    import ctypes

    lib = L.load()

    def free():
        f, t = ctypes.c_double(), ctypes.c_double()
        assert lib.ima_device_memory(0, ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value

    specs = _analyses()

    def cycle():
        for s in specs:
            e = Engine(0)
            s.load(e)
            e.mc3_run(3, SERVED)
            e.close()

    cycle()  # first-use allocations of the runtime
    free0 = free()
    for _ in range(10):
        cycle()
    free1 = free()
    assert free0 - free1 < (8 << 20), (free0 - free1)
