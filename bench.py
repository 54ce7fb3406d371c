#!/usr/bin/env python3
"""Benchmark of the batched MC3 hot path on one MI355X.

Workload (default): three-population IM model, 50 HKY loci x 1000 bp,
45 gene copies per locus, 64 heated chains (geometric heating), synthetic
data from the package's simulator.  One step = one MC3 sweep over all
chains: per (chain, locus) one genealogy proposal, updategenealogy's own
move (an edge cut and slid to a new sister, its migration path -- and the
new sister's when it ends at the root -- redrawn period by period, with
the move's Hastings term), scored in a single batched FP32-scaled
likelihood launch plus one launch of genealogy weights, then accepted or
rejected locus by locus with beta (dprobg + dpdg) + Hastings term, probg
the chain's genealogy prior integrated over the population-size and
migration priors (update_gtree.cpp:827-961); on alternate sweeps each
chain's time update, drawn as qupdate draws it (ima_main_mpi.cpp:
1867-1911): changet_RY1's rubber band on one split time, re-scoring all
50 loci and the prior, or changet_NW, which redraws the migrations around
the moved split time; every fifth sweep the mutation-rate round (changeu);
then the chain swap.  value = chain-steps/s = chains x sweeps / wall
time, with all state resident in HBM.

Also reported: the HKY kernel's roofline (HIP events on the context's
stream around the sweep's scoring launch repeated back to back, over the
chains' current genealogies), the same sweep in FP64, the plain
per-locus FP64 GPU path (one launch per likelihood call, host control
loop, one prior call per proposal), L-mode jointp points/s, and the CPU
baseline: the oracle's sweep in C (its restatement of the reference's
likelihoodHKY, incremental as the reference evaluates it, treeweight and
the prior integrals, with the same proposals and swap) on the host cores.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ima2p_amd import workloads  # noqa: E402

FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def hky_flops(n, S):
    """Algorithmic FP32 flops of one full pruning pass: (36n - 58) S."""
    return (36 * n - 58) * S


def hky_traffic():
    """HBM bytes per HKY launch of this workload, from the committed
    rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE passes, each scaled by
    the same counter's reading for a plain copy of known size)."""
    path = os.path.join(ROOT, "profiles", "hky_kernel_3pop_50x1000bp_64chains_pmc.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


# ------------------------------------------------------------- CPU baseline --

_WL = None


def _cpu_worker(args):
    """The oracle's C sweep (likelihoodHKY incremental as the reference
    evaluates it, treeweight, the prior integrals, the same proposals and
    swap) on a chain subset for `seconds`; returns (chain_steps, elapsed)."""
    wid, chain_ids, seconds = args
    import oracle
    from ima2p_amd import immodel
    wl = _WL
    chains = [[g.copy() for g in wl.chains[c]] for c in chain_ids]
    tvs = np.tile(wl.tsplit, (len(chain_ids), 1))
    t0 = time.perf_counter()
    sweeps, _ = oracle.mc3_bench(wl.loci, wl.u, wl.kappa, wl.samppop, immodel.build(wl.poptree),
                                 chains, tvs, workloads.geometric_betas(len(chain_ids)),
                                 wl.tmax, seed=100 + wid, seconds=seconds, ry1=True,
                                 urate=True, nw=1, topo=True)
    return len(chain_ids) * sweeps, time.perf_counter() - t0


def sweep_flags(L, nw=True):
    """The device sweep of the reference's step: graph replay, the
    reference's genealogy move, the time update on alternate sweeps (RY1,
    or NW with probability 1/2 for models with migration), the
    mutation-rate round every fifth sweep."""
    return (L.MC3_GRAPH | L.MC3_TOPOLOGY | L.MC3_SCALE | L.MC3_RY1 | L.MC3_UPDATE_U
            | (L.MC3_NW if nw else 0))


def cpu_baseline(wl, seconds, cores):
    global _WL
    _WL = wl
    per = max(1, 4)  # 4 chains per process, like -hn4 on each MPI rank
    jobs = [(w, [(w * per + i) % wl.nchains for i in range(per)], seconds) for w in range(cores)]
    ctx = mp.get_context("fork")
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_worker, jobs)
    steps = sum(r[0] for r in res)
    wall = max(r[1] for r in res)
    return steps / wall, steps, wall


# ----------------------------------------------------------------- GPU legs --

def load_engine(wl, prec, prior=True):
    from ima2p_amd import _lib as L
    from ima2p_amd import immodel
    from ima2p_amd.engine import Engine
    e = Engine(0)
    model = immodel.build(wl.poptree)
    mid = e.model_add(model)
    loci = [e.add_locus_hky(loc) for loc in wl.loci]
    for lid in loci:
        e.locus_set_model(lid, mid, wl.samppop)
    C, Lc = wl.nchains, wl.nloci
    first = e.alloc_slots([loci[l] for c in range(C) for l in range(Lc)])
    gens = [wl.chains[c][l] for c in range(C) for l in range(Lc)]
    e.write_slots(first, gens)
    e.write_pops(first, gens)
    u = np.array([wl.u[l] for c in range(C) for l in range(Lc)])
    e.mc3_setup(C, Lc, first, u, wl.kappa, wl.beta, seed=12345, precision=prec)
    if prior:
        e.mc3_set_prior(np.tile(wl.tsplit, (C, 1)), wl.tmax, model.npops)
    return e, loci, first, model


def plain_leg(wl, sweeps=2):
    """The plain per-locus FP64 GPU path (libima_plain.so, C++ over the
    ABI): every chain, every locus, one proposal at a time -- one slot
    upload, one FP64 likelihood launch, one treeweight launch, the chain's
    all-loci sums updated on the host, one prior-integral launch and the
    decision, then one swap per sweep.  The batched comparator runs the
    same step (genealogy update of every locus with the prior, then the
    swap; no time or rate updates) from captured graphs."""
    from ima2p_amd import _lib as L
    from ima2p_amd import immodel, plain
    from ima2p_amd.engine import Engine
    C, Lc = wl.nchains, wl.nloci
    e = Engine(0)
    model = immodel.build(wl.poptree)
    mid = e.model_add(model)
    loci = [e.add_locus_hky(loc) for loc in wl.loci]
    for lid in loci:
        e.locus_set_model(lid, mid, wl.samppop)
    first = e.alloc_slots([loci[l] for c in range(C) for l in range(Lc)])
    gens = [wl.chains[c][l] for c in range(C) for l in range(Lc)]
    e.write_slots(first, gens)
    e.write_pops(first, gens)
    u = np.array([wl.u[l] for c in range(C) for l in range(Lc)])
    plain.run(e, first, mid, model.npops, gens, C, Lc, u, wl.kappa, wl.beta,
              np.tile(wl.tsplit, (C, 1)), nsweeps=0)  # first use of every launch
    r = plain.run(e, first, mid, model.npops, gens, C, Lc, u, wl.kappa, wl.beta,
                  np.tile(wl.tsplit, (C, 1)), nsweeps=sweeps)
    e.close()
    out = {"chain_steps_per_s": C * sweeps / r["seconds"], "sweeps": sweeps,
           "us_per_proposal": r["seconds"] / (sweeps * C * Lc) * 1e6,
           "accept": r["accepts"] / (sweeps * C * Lc)}
    same = {}
    for name, prec in (("fp32", L.PREC_FP32), ("fp64", L.PREC_FP64)):
        eb, _, _, _ = load_engine(wl, prec)
        eb.mc3_tune(6, L.MC3_GRAPH)
        eb.mc3_run(16, L.MC3_GRAPH)
        t0 = time.perf_counter()
        eb.mc3_run(96, L.MC3_GRAPH)
        same[name] = C * 96 / (time.perf_counter() - t0)
        eb.close()
    out["batched_same_step_fp32_chain_steps_per_s"] = same["fp32"]
    out["batched_same_step_fp64_chain_steps_per_s"] = same["fp64"]
    out["speedup_batched_fp32_vs_plain_fp64"] = same["fp32"] / out["chain_steps_per_s"]
    out["speedup_batched_fp64_vs_plain_fp64"] = same["fp64"] / out["chain_steps_per_s"]
    return out


def lmode_leg(G=100000, npoints=4096, seed=4):
    from ima2p_amd import _lib as L
    from ima2p_amd import sim
    from ima2p_amd.engine import Engine
    rng = np.random.default_rng(seed)
    tab = sim.lmode_table(rng, G)
    e = Engine(0)
    tid = e.lmode_table(tab, 3, 2, 2)
    grid = sim.lmode_grid([1.0, 1.0, 1.0, 0.05, 0.05], [10.0, 10.0, 10.0, 2.0, 2.0], 6)
    x = grid[rng.permutation(len(grid))[:npoints]] if len(grid) >= npoints else \
        np.resize(grid, (npoints, 5))
    out = {}
    for name, mode in (("exact", L.LMODE_EXACT), ("ref_compat", L.LMODE_REF_COMPAT)):
        e.jointp(tid, x[:64], mode=mode)
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            e.jointp(tid, x, mode=mode)
        dt = (time.perf_counter() - t0) / reps
        out[name] = {"points_per_s": npoints / dt, "pairs_per_s": npoints * G / dt,
                     "ms_per_query": dt * 1e3}
    out["G"] = G
    out["points"] = npoints
    # per (genealogy, point) pair, two populations: F = 2 K_f + 4 + 16 with
    # K_f = 3 P_theta + 2 P_m + 1 = 14 table terms, one exp counted as 16
    out["flops_per_pair"] = 48
    out["fp64_roofline_pairs_per_s"] = FP64_PEAK_TFLOPS * 1e12 / 48.0
    out["exact_frac_of_fp64_peak"] = out["exact"]["pairs_per_s"] / out["fp64_roofline_pairs_per_s"]
    # the L-mode drivers on the batched launch: the differential-evolution
    # joint peak of the full model and one marginal peak with its 95% limits
    lo = [1e-7] * 5
    hi = [20.0, 20.0, 20.0, 5.0, 5.0]
    t0 = time.perf_counter()
    _, logp, _, evals = e.joint_peak(tid, lo, hi, seed=3, mode=L.LMODE_REF_COMPAT)
    dt = time.perf_counter() - t0
    out["joint_peak"] = {"seconds": dt, "jointp_points": evals, "points_per_s": evals / dt,
                         "log_p": logp}
    t0 = time.perf_counter()
    e.marginal_peak(tid, 0, 1e-7, 20.0)
    out["marginal_peak_ms"] = (time.perf_counter() - t0) * 1e3
    e.close()
    return out


def uneven_leg(sweeps=40):
    """The uneven batch: 256 chains x 300 loci (150 HKY of 200..3000 bp, 100
    infinite-sites, 50 stepwise; 4..60 gene copies), all scored in one
    launch per model each sweep, with the prior; group size autotuned."""
    from ima2p_amd import _lib as L
    from ima2p_amd.engine import Engine
    spec = workloads.uneven_mixed()
    e = Engine(0)
    spec.load(e)
    flags = sweep_flags(L)
    group, group_ms = e.mc3_tune(4, flags)
    block = e.mc3_site_block()
    e.mc3_run(4, flags)
    t0 = time.perf_counter()
    e.mc3_run(sweeps, flags)
    dt = time.perf_counter() - t0
    # the other site-block size, for the record
    other = 64 if block == 128 else 128
    e.close()
    e = Engine(0)
    e.set_site_block(other)
    spec.load(e)
    if group:
        e.mc3_prior_group(group)
    e.mc3_run(4, flags)
    t1 = time.perf_counter()
    e.mc3_run(sweeps, flags)
    dt_other = time.perf_counter() - t1
    e.close()
    kinds = [k for k, _ in spec.loci]
    return {"chain_steps_per_s": spec.nchains * sweeps / dt, "ms_per_sweep": dt / sweeps * 1e3,
            "chains": spec.nchains, "loci": spec.nloci,
            "loci_by_model": {k: kinds.count(k) for k in ("hky", "is", "sw")},
            "items_per_sweep": spec.nchains * spec.nloci, "tuned_group_loci": group,
            "tuned_ms_per_sweep_by_group": [round(float(x), 4) for x in group_ms],
            "tuned_site_block_patterns": block,
            "ms_per_sweep_other_site_block": {str(other): dt_other / sweeps * 1e3}}


def _is_engine(seed=1, analyses=1):
    """The CPU-runnable case (two populations, 5 infinite-sites loci of
    10+10 gene copies, 4 chains) on a context of its own, FP64, with the
    prior; with analyses > 1, that many uncoupled copies of it (each its
    own 4 chains and heating ladder, swaps inside each) in one sweep."""
    from ima2p_amd import _lib as L
    from ima2p_amd import immodel
    from ima2p_amd.engine import Engine
    from ima2p_amd.serve import AnalysisSpec
    wl = workloads.two_pop_is()
    model = immodel.build(wl.poptree, qmax=20.0, mmax=5.0)
    C = wl.nchains * analyses
    spec = AnalysisSpec(loci=[("is", d) for d in wl.loci], model=model,
                        samppop=[wl.samppop] * wl.nloci, chains=list(wl.chains) * analyses,
                        u=np.tile(wl.u, (C, 1)), kappa=np.ones((C, wl.nloci)),
                        beta=np.tile(wl.beta, analyses),
                        tvals=np.tile(wl.tsplit, (C, 1)), tmax=wl.tmax, seed=seed,
                        precision=L.PREC_FP64)
    e = Engine(0)
    spec.load(e, swaptries=max(1, wl.nchains // 10))  # per analysis, ima_main_mpi.cpp:1376-1378
    if analyses > 1:
        e.mc3_set_analyses(wl.nchains)
    return e, wl


def is_leg(sweeps=2000, analyses=1):
    """The CPU-runnable case on the device: one analysis of 4 chains, or
    `analyses` uncoupled ones in the same sweep (the shape of the
    reference's aggregate over 8 concurrent processes)."""
    from ima2p_amd import _lib as L
    e, wl = _is_engine(analyses=analyses)
    e.mc3_run(20, sweep_flags(L))
    t0 = time.perf_counter()
    e.mc3_run(sweeps, sweep_flags(L))
    dt = time.perf_counter() - t0
    e.close()
    return {"chain_steps_per_s": analyses * wl.nchains * sweeps / dt,
            "ms_per_sweep": dt / sweeps * 1e3, "analyses": analyses, "workload": wl.name}


def serving_leg(duration_s=20.0, rate=20.0, seed=5, nsweeps=100):
    """Poisson request stream at `rate`/s for `duration_s` s: 70% L-mode
    query batches of 1..4096 points (log-uniform) on resident tables of
    1e4, 1e5 and 1e6 genealogies, 30% new IM analyses with 2..4 populations,
    5..300 mixed HKY/IS/SW loci and 16..128 chains running `nsweeps` sweeps.
    Requests are generated (and analyses flattened) before the stream
    starts; latency runs from a request's arrival time to the step that
    sees it finish, so time a request waits behind the host loop counts."""
    from ima2p_amd import _lib as L
    from ima2p_amd import sim
    from ima2p_amd.serve import REQ_DONE, Server
    rng = np.random.default_rng(seed)
    tab_sizes = (10000, 100000, 1000000)
    tabs, pools, stream = workloads.serving_requests(rng, duration_s, rate, nsweeps, tab_sizes)
    lo = np.array([0.5, 0.5, 0.5, 0.01, 0.01])
    reqs = [(t, k, (d.nchains, d.flatten()) if k == "a" else d) for t, k, d in stream]
    srv = Server(0, 0.0, 25)
    tids = [srv.table(tb, 3, 2, 2) for tb in tabs]
    # one warm query per table and one small analysis before the stream
    for tid in tids:
        srv.submit_lmode(tid, lo[None, :] + 1.0)
    warm = workloads.serving_analysis(rng, pools[2], 5, 16, nsweeps=2)
    srv.submit_analysis(warm)
    srv.run(120)
    rec = []  # (kind, arrival, submit offset, id, size)
    steps = []  # host time of each step call
    t0 = time.perf_counter()
    i = 0
    pending = 0
    while i < len(reqs) or pending:
        now = time.perf_counter() - t0
        while i < len(reqs) and reqs[i][0] <= now:
            at, kind, data = reqs[i]
            sub = time.perf_counter() - t0
            if kind == "q":
                k, x = data
                rid = srv.submit_lmode(tids[k], x)
                rec.append(("q", at, sub, rid, len(x)))
            else:
                nch, flat = data
                rid = srv.submit_analysis(flat)
                rec.append(("a", at, sub, rid, nch))
            i += 1
        ts = time.perf_counter()
        pending = srv.step()
        steps.append(time.perf_counter() - ts)
        if now > duration_s + 180.0:
            raise TimeoutError("serving stream did not drain within 180 s of its end")
        if not pending and i < len(reqs):
            time.sleep(max(0.0, reqs[i][0] - (time.perf_counter() - t0)))
    wall = time.perf_counter() - t0
    out = {}
    for kind, name in (("q", "lmode"), ("a", "analysis")):
        lat, work = [], 0
        for k, at, sub, rid, size in rec:
            if k != kind:
                continue
            st, l, _ = srv.status(rid)
            assert st == REQ_DONE, (kind, rid, st)
            lat.append((sub - at) + l)
            work += size * (1 if kind == "q" else nsweeps)
        lat = np.sort(np.asarray(lat)) * 1e3
        out[name] = {"count": len(lat), "p50_ms": float(np.percentile(lat, 50)),
                     "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat[-1]),
                     ("points_per_s" if kind == "q" else "chain_steps_per_s"): work / wall}
    srv.close()
    out.update({"requests": len(rec), "arrival_rate_per_s": rate, "stream_s": duration_s,
                "wall_s": wall, "tables": list(tab_sizes), "sweeps_per_analysis": nsweeps,
                "step_calls": len(steps), "step_max_ms": max(steps) * 1e3,
                "step_p99_ms": float(np.percentile(steps, 99)) * 1e3})
    return out


def accept_model(avg_us, nloci, group, ms_by_group):
    """The accept step (mc3_accept, one workgroup per chain) is bound by
    latency, not throughput: each chain's loci are decided in order, in
    speculative groups of `group` loci, and every group waits for its
    scenario integrals (incomplete-gamma chains side by side) before its
    in-order decisions.  The tuner's sweep over group sizes D = 1..6 gives
    the cost of one group: the sweep time is linear in the number of groups
    ceil(L / D) (least squares; the slope is the latency of one group, the
    intercept the rest of the sweep).  The latency bound of the accept step
    is groups x that slope; frac = bound / measured accept-step time."""
    groups = -(-nloci // max(1, group))
    out = {"avg_us": avg_us, "groups_per_chain": groups, "us_per_group": avg_us / groups,
           "bound": "latency"}
    if ms_by_group is not None and len(ms_by_group) >= 2:
        g = np.array([-(-nloci // d) for d in range(1, len(ms_by_group) + 1)], float)
        t = np.asarray(ms_by_group, float) * 1e3
        slope, icpt = np.polyfit(g, t, 1)
        out["group_latency_us"] = float(slope)
        out["sweep_rest_us"] = float(icpt)
        out["latency_bound_us"] = float(groups * slope)
        out["frac"] = float(groups * slope / avg_us)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-cores", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip FP64/plain/L-mode legs")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="length of the serving stream (0 skips it)")
    ap.add_argument("--serve-only", action="store_true", help="run the serving stream only")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run 5 plain device copies of MIB MiB (fp32 copy_), the "
                         "known-size kernel that scales PMC FETCH_SIZE/WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs all heated chains on one GPU; "
                                   "multi-GPU runs are not part of this build (use --gpus 1)"}))
        sys.exit(2)
    if args.serve_only:
        print(json.dumps({"serving": serving_leg(duration_s=args.serve_seconds or 20.0)}))
        return

    if args.calibrate_copy > 0:
        import torch
        nel = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(nel, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    t_setup = time.perf_counter()
    wl = workloads.three_pop_hky()
    S = wl.patterns()
    n = wl.chains[0][0].n
    C, Lc = wl.nchains, wl.nloci

    cpu = None
    if not args.no_cpu:
        cores = args.cpu_cores or min(16, os.cpu_count() or 1)
        rate, steps, wall = cpu_baseline(wl, args.cpu_seconds, cores)
        cpu = {"value": rate, "unit": "chain-steps/s", "cores": cores, "kind": "port",
               "sample": f"the oracle's C sweep (likelihoodHKY incremental as the reference, "
                         f"treeweight, prior integrals, same proposals, changeu and swap): {cores} "
                         f"processes x 4 chains x 50 loci, {steps} chain-steps in {wall:.1f} s"}

    from ima2p_amd import _lib as L
    e, loci, first, model = load_engine(wl, L.PREC_FP32)
    setup_s = time.perf_counter() - t_setup
    # autotune the accept step's group size (the chains are left as they were)
    group, group_ms = e.mc3_tune(10, sweep_flags(L))
    e.mc3_run(args.warmup, sweep_flags(L))
    t0 = time.perf_counter()
    e.mc3_run(args.steps, sweep_flags(L))
    dt = time.perf_counter() - t0
    value = C * args.steps / dt
    pdg, beta, stats, sweeps = e.mc3_read()

    # eager sweeps with HIP events around each sweep
    e.timing(True)
    e.timing_reset()
    ksweeps = max(10, min(args.steps, 100))
    e.mc3_run(ksweeps, (sweep_flags(L) & ~L.MC3_GRAPH) | L.MC3_TIMED)
    sweep_n, sweep_ms = e.timing_get(L.K_MC3)
    acc_n, acc_ms = e.timing_get(L.K_ACCEPT)
    # HKY kernel: the sweep's scoring launch (every chain x locus of the
    # chains' current genealogies) repeated back to back between one pair
    # of HIP events on the context's stream
    e.timing_reset()
    items = e.items(first + np.arange(C * Lc), np.tile(wl.u, C), wl.kappa)
    reps = 200
    e.score_repeat(items, reps, L.PREC_FP32)
    launches, ms = e.timing_get(L.K_HKY32)
    e.timing(False)
    avg_s = ms / launches / 1e3
    flops = float(sum(hky_flops(n, S[l]) for l in range(Lc))) * C
    achieved = flops / avg_s / 1e12

    extra = {}
    if not args.no_extra:
        e64, loci64, _, model64 = load_engine(wl, L.PREC_FP64)
        e64.mc3_run(5, sweep_flags(L))
        t1 = time.perf_counter()
        k64 = max(10, args.steps // 4)
        e64.mc3_run(k64, sweep_flags(L))
        extra["batched_fp64_chain_steps_per_s"] = C * k64 / (time.perf_counter() - t1)
        e64.close()
        # the same FP32 sweep with every time update RY1 (no NW)
        e_r, _, _, _ = load_engine(wl, L.PREC_FP32)
        e_r.mc3_prior_group(group)
        e_r.mc3_run(args.warmup, sweep_flags(L, nw=False))
        t1 = time.perf_counter()
        e_r.mc3_run(args.steps, sweep_flags(L, nw=False))
        extra["ry1_only_chain_steps_per_s"] = C * args.steps / (time.perf_counter() - t1)
        e_r.close()
        pl = plain_leg(wl)
        extra["plain_per_locus_fp64"] = pl
        extra["plain_per_locus_fp64_chain_steps_per_s"] = pl["chain_steps_per_s"]
        # the north star's ratio: batched FP32-scaled over plain per-locus
        # FP64, both on the same step (genealogy updates with the prior + swap)
        extra["speedup_vs_plain_fp64"] = pl["speedup_batched_fp32_vs_plain_fp64"]
        extra["batched_fp32_vs_batched_fp64"] = value / extra["batched_fp64_chain_steps_per_s"]
        if cpu:
            extra["batched_fp32_vs_cpu_baseline"] = value / cpu["value"]
        extra["lmode"] = lmode_leg()
        extra["uneven_batch"] = uneven_leg()
        extra["is_case"] = is_leg()
        extra["is_case_8_uncoupled"] = is_leg(analyses=8)
    if args.serve_seconds > 0:
        extra["serving"] = serving_leg(duration_s=args.serve_seconds)

    line = {
        "metric": "chain-steps/sec (all loci)",
        "value": value,
        "unit": "chain-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": wl.name, "chains": C, "loci": Lc, "sites_per_locus": 1000,
                   "genes_per_locus": n, "patterns_mean": float(S.mean()),
                   "patterns_min": int(S.min()), "patterns_max": int(S.max()),
                   "sweep": "1 genealogy proposal per (chain, locus), updategenealogy's move "
                            "(slider + addmigration), scored with data likelihood "
                            "and integrated genealogy prior, loci in order + on alternate sweeps "
                            "one split-time update per chain, changet_RY1 (rubber band, every "
                            "locus re-scored) or changet_NW (migrations redrawn around the "
                            "moved split time) with probability 1/2 each + swap; every fifth "
                            "sweep also changeu: per chain 50 pair moves of (u, kappa), each "
                            "re-scoring two loci",
                   "parallelism": "all chains on one GPU"},
        "roofline": {"bound": "mfma", "pipe": "VALU FP32 (pruning is not a dense contraction)",
                     "kernel": "hky_kernel<float>", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": hky_traffic(),
                     "flops_per_launch": flops, "avg_launch_us": avg_s * 1e6,
                     "launches_timed": launches},
        "cpu_baseline": cpu,
        "accept": {"genealogy": float(stats[1]) / max(1, stats[0]),
                   "time_update": float(stats[3]) / max(1, stats[2]),
                   "swap": float(stats[5]) / max(1, stats[4]),
                   "rates": float(stats[7]) / max(1, stats[6])},
        "eager_sweep_ms": sweep_ms / max(1, sweep_n),
        "accept_step": accept_model(acc_ms / max(1, acc_n) * 1e3, Lc, group, group_ms),
        "tuned": {"accept_group_loci": group,
                  "ms_per_sweep_by_group": [round(float(x), 4) for x in group_ms]},
        "setup_s": setup_s,
    }
    line.update(extra)
    e.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
