"""The serving stream at its own sizes: Poisson arrivals at 20 requests/s
for 5 s (seed 5); 70% L-mode query batches of 1..4096 points on resident
tables of 1e4, 1e5 and 1e6 genealogies, 30% IM analyses with 2..4
populations, 5..300 mixed HKY/IS/SW loci and 16..128 chains.  Requests
are submitted at their arrival times while the device works on earlier
ones, so launches of different requests overlap.

* Every L-mode result equals a direct ima_lmode_jointp of the same table
  and points, bit for bit; a few points also equal the oracle's jointp
  (1e-10), including the 1e6-row table.
* Sampled analyses (the largest, the smallest, one of four populations)
  end in the state a direct ima_mc3_run of the same set-up reaches, bit
  for bit; the largest one's direct run then matches the oracle on two
  chains (every locus' log-likelihood, probg).
"""
import time

import numpy as np
import pytest

import oracle
from ima2p_amd import _lib as L
from ima2p_amd import workloads
from ima2p_amd.engine import Engine
from ima2p_amd.serve import REQ_DONE, Server
from tests.test_gpu_configs import _check_chains, rel

pytestmark = pytest.mark.gpu

FLAGS = (L.MC3_GRAPH | L.MC3_TOPOLOGY | L.MC3_SCALE | L.MC3_RY1 | L.MC3_NW
         | L.MC3_UPDATE_U)  # the step a served analysis runs


def _stream(srv, tids, reqs, limit_s=240.0):
    ids = []
    t0 = time.perf_counter()
    i, pending = 0, 0
    while i < len(reqs) or pending:
        now = time.perf_counter() - t0
        while i < len(reqs) and reqs[i][0] <= now:
            _, kind, d = reqs[i]
            if kind == "q":
                ids.append(srv.submit_lmode(tids[d[0]], d[1]))
            else:
                ids.append(srv.submit_analysis(d))
            i += 1
        pending = srv.step()
        assert now < limit_s, "stream did not drain"
        if not pending and i < len(reqs):
            time.sleep(max(0.0, reqs[i][0] - (time.perf_counter() - t0)))
    return ids


def test_cfg5_stream_results_equal_direct_runs():
    rng = np.random.default_rng(5)
    tabs, _, reqs = workloads.serving_requests(rng, duration_s=5.0, rate=20.0, nsweeps=20)
    kinds = [k for _, k, _ in reqs]
    assert kinds.count("q") > 0 and kinds.count("a") > 0
    print(f"{len(reqs)} requests: {kinds.count('q')} L-mode, {kinds.count('a')} analyses", flush=True)
    srv = Server(0, 0.0, 25)
    tids = [srv.table(t, 3, 2, 2) for t in tabs]
    ids = _stream(srv, tids, reqs)
    results = []
    for rid, (_, kind, d) in zip(ids, reqs):
        st, lat, _ = srv.status(rid)
        assert st == REQ_DONE and lat >= 0
        results.append(srv.lmode_result(rid) if kind == "q" else srv.analysis_result(rid))
    print("served; L-mode", srv.latency(0), flush=True)
    srv.close()

    # every L-mode query against a direct launch on the same table
    e = Engine(0)
    dt = [e.lmode_table(t, 3, 2, 2) for t in tabs]
    qn = 0
    for (_, kind, d), res in zip(reqs, results):
        if kind != "q":
            continue
        k, x = d
        out, ess = e.jointp(dt[k], x, mode=L.LMODE_EXACT)
        np.testing.assert_array_equal(res[0], out)
        np.testing.assert_array_equal(res[1], ess)
        if qn < 6:  # a few points against the oracle, every table size
            j = int(rng.integers(0, len(x)))
            assert rel(out[j], oracle.jointp_exact(tabs[k], 3, 2, 2, x[j])) < 1e-10
        qn += 1
    e.close()

    # sampled analyses against direct runs
    an = [(i, d) for i, (_, kind, d) in enumerate(reqs) if kind == "a"]
    size = [d.nchains * d.nloci for _, d in an]
    pick = {an[int(np.argmax(size))][0], an[int(np.argmin(size))][0]}
    four = [i for i, d in an if d.model.npops == 4]
    if four:
        pick.add(four[0])
    largest = an[int(np.argmax(size))][0]
    for i in sorted(pick):
        spec = reqs[i][2]
        e = Engine(0)
        first, _ = spec.load(e)
        e.mc3_run(spec.nsweeps, FLAGS)
        pdg, beta, stats, _ = e.mc3_read()
        probg, _ = e.mc3_read_prior()
        got = results[i]
        for x, y in zip((pdg, probg, beta, stats), got):
            np.testing.assert_array_equal(np.ravel(x), np.ravel(y))
        if i == largest:
            tol = 1e-9 if spec.precision == L.PREC_FP64 else 1e-6
            _check_chains(e, spec, first, [0, spec.nchains - 1], tol, moved=True)
        print(f"analysis {i}: {spec.model.npops} pops, {spec.nloci} loci x {spec.nchains} chains "
              "equal direct", flush=True)
        e.close()
