"""Python front end of the engine's request loop (ima_server_* in the C ABI).

``Server`` wraps a native server: resident L-mode tables, L-mode queries
and IM analyses submitted as they arrive, ``step()`` to make progress, and
per-request results and latencies.  ``AnalysisSpec`` flattens an analysis
(loci, population model, genealogies of every chain) into the ABI's
``ima_analysis`` description.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .engine import EngineError
from .immodel import IMModel, mig_arrays

REQ_WAITING, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3
REQ_LMODE, REQ_ANALYSIS = 0, 1

_ip = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)


class _Analysis(C.Structure):
    _fields_ = [("nloci", C.c_int), ("nchains", C.c_int), ("npops", C.c_int),
                ("model", _ip), ("n", _ip), ("S", _ip), ("totsites", _ip),
                ("seq", C.POINTER(C.c_int8)), ("mult", _ip), ("pi", _dp), ("samppop", _ip),
                ("plist", _ip), ("npresent", _ip), ("popdown", _ip), ("pope", _ip),
                ("join", _ip), ("npsize", C.c_int), ("theta_off", _ip), ("theta_wp", _ip),
                ("nmig", C.c_int), ("mig_off", _ip), ("mig_wp", _ip), ("qmax", _dp),
                ("mmax", _dp), ("up0", _ip), ("up1", _ip), ("down", _ip), ("root", _ip),
                ("A", _ip), ("pop", _ip), ("nmigev", _ip), ("mig_edge", _ip),
                ("mig_to", _ip), ("time", _dp), ("mig_time", _dp), ("u", _dp),
                ("kappa", _dp), ("beta", _dp), ("tvals", _dp), ("tmax", C.c_double),
                ("seed", C.c_uint64), ("precision", C.c_int), ("nsweeps", C.c_int)]


@dataclass
class AnalysisSpec:
    """One IM analysis.  loci[l] is ("hky", HKYLocus), ("is", ISLocus) or
    ("sw", n); chains[c][l] a Genealogy (population labels, migrations);
    A[c][l] the allele states of a stepwise locus (else None)."""
    loci: list
    model: IMModel
    samppop: list                 # per locus, gene copies per sampled population
    chains: list
    u: np.ndarray                 # [C, L]
    kappa: np.ndarray             # [C, L]
    beta: np.ndarray              # [C]
    tvals: np.ndarray             # [C, P-1]
    tmax: float = 5.0
    seed: int = 1
    precision: int = L.PREC_FP32
    nsweeps: int = 100
    A: list = field(default_factory=list)

    @property
    def nchains(self):
        return len(self.chains)

    @property
    def nloci(self):
        return len(self.loci)

    def load(self, engine, scale_width=0.2, swaptries=None, prior=True):
        """Set the analysis up on an Engine as the server does (same loci
        order, slots chain-major, IS sumlogk from chain 0's genealogy, MC3
        with the prior).  Returns (first slot, locus ids)."""
        m = self.model
        mid = engine.model_add(m) if prior else None
        lids = []
        for l, (kind, data) in enumerate(self.loci):
            lid = (engine.add_locus_hky(data) if kind == "hky" else
                   engine.add_locus_is(data) if kind == "is" else engine.add_locus_sw(data))
            if prior:
                engine.locus_set_model(lid, mid, self.samppop[l])
            lids.append(lid)
        Cc, Lc = self.nchains, self.nloci
        first = engine.alloc_slots([lids[i % Lc] for i in range(Cc * Lc)])
        gens = [g for row in self.chains for g in row]
        A = [a if a is not None else np.zeros(2 * g.n - 1, np.int32)
             for row, grow in zip(self.A or [[None] * Lc] * Cc, self.chains)
             for a, g in zip(row, grow)]
        engine.write_slots(first, gens, A=A)
        if prior:
            engine.write_pops(first, gens)
        for l, (kind, _) in enumerate(self.loci):
            if kind == "is":
                engine.sumlogk_from_slot(lids[l], first + l)
        engine.mc3_setup(Cc, Lc, first, np.ravel(self.u), np.ravel(self.kappa), self.beta,
                         seed=self.seed, precision=self.precision, swaptries=swaptries,
                         scale_width=scale_width)
        if prior:
            engine.mc3_set_prior(self.tvals, self.tmax, m.npops)
        return first, lids

    def flatten(self):
        """(ctypes struct, arrays kept alive for the duration of a call)."""
        Lc, Cc, m = self.nloci, self.nchains, self.model
        P = m.npops
        keep = []

        def arr(a, dt):
            x = np.ascontiguousarray(a, dt)
            keep.append(x)
            return x

        def ptr(a, dt, ct):
            x = arr(a, dt)
            return x.ctypes.data_as(C.POINTER(ct))

        kinds = {"hky": L.MODEL_HKY, "is": L.MODEL_IS, "sw": L.MODEL_SW}
        model = [kinds[k] for k, _ in self.loci]
        n = [self.chains[0][l].n for l in range(Lc)]
        S, tot, seq, mult, pi = [], [], [], [], []
        for k, loc in self.loci:
            if k == "hky":
                S.append(loc.seq.shape[1])
                tot.append(loc.totsites)
                seq.append(np.asarray(loc.seq, np.int8).ravel())
                mult.append(np.asarray(loc.mult, np.int32))
                pi.append(np.asarray(loc.pi, float))
            elif k == "is":
                S.append(loc.seq.shape[1])
                tot.append(loc.seq.shape[1])
                seq.append(np.asarray(loc.seq, np.int8).ravel())
                pi.append(np.full(4, 0.25))
            else:
                S.append(0)
                tot.append(0)
                pi.append(np.full(4, 0.25))
        gens = [g for row in self.chains for g in row]
        Aall = []
        for c in range(Cc):
            for l in range(Lc):
                a = self.A[c][l] if self.A and self.A[c][l] is not None else None
                Aall.append(np.zeros(2 * n[l] - 1, np.int32) if a is None else np.asarray(a))
        cache = {}  # chains often share genealogy objects

        def migs_of(g):
            k = id(g)
            if k not in cache:
                cache[k] = mig_arrays(g)
            return cache[k]
        migs = [migs_of(g) for g in gens]
        toff = np.cumsum([0] + [len(w) for w in m.theta_wp]).astype(np.int32)
        twp = [x for w in m.theta_wp for kv in w for x in kv] or [0]
        moff = np.cumsum([0] + [len(w) for w in m.mig_wp]).astype(np.int32)
        mwp = [x for w in m.mig_wp for kv in w for x in kv] or [0]
        s = _Analysis()
        s.nloci, s.nchains, s.npops = Lc, Cc, P
        s.model = ptr(model, np.int32, C.c_int32)
        s.n = ptr(n, np.int32, C.c_int32)
        s.S = ptr(S, np.int32, C.c_int32)
        s.totsites = ptr(tot, np.int32, C.c_int32)
        s.seq = ptr(np.concatenate(seq) if seq else np.zeros(1), np.int8, C.c_int8)
        s.mult = ptr(np.concatenate(mult) if mult else np.zeros(1), np.int32, C.c_int32)
        s.pi = ptr(np.concatenate(pi), np.float64, C.c_double)
        s.samppop = ptr(np.asarray(self.samppop).reshape(Lc, P), np.int32, C.c_int32)
        s.plist = ptr(m.plist, np.int32, C.c_int32)
        s.npresent = ptr(m.npresent, np.int32, C.c_int32)
        s.popdown = ptr(m.popdown, np.int32, C.c_int32)
        s.pope = ptr(m.pope, np.int32, C.c_int32)
        s.join = ptr(m.join, np.int32, C.c_int32)
        s.npsize = m.npsize
        s.theta_off = ptr(toff, np.int32, C.c_int32)
        s.theta_wp = ptr(twp, np.int32, C.c_int32)
        s.nmig = m.nmig
        s.mig_off = ptr(moff, np.int32, C.c_int32)
        s.mig_wp = ptr(mwp, np.int32, C.c_int32)
        s.qmax = ptr(m.qmax if m.npsize else [1.0], np.float64, C.c_double)
        s.mmax = ptr(m.mmax if m.nmig else [1.0], np.float64, C.c_double)
        s.up0 = ptr(np.concatenate([g.up0 for g in gens]), np.int32, C.c_int32)
        s.up1 = ptr(np.concatenate([g.up1 for g in gens]), np.int32, C.c_int32)
        s.down = ptr(np.concatenate([g.down for g in gens]), np.int32, C.c_int32)
        s.root = ptr([g.root for g in gens], np.int32, C.c_int32)
        s.A = ptr(np.concatenate(Aall), np.int32, C.c_int32)
        s.pop = ptr(np.concatenate([g.pop for g in gens]), np.int32, C.c_int32)
        s.nmigev = ptr([len(x[0]) for x in migs], np.int32, C.c_int32)
        s.mig_edge = ptr(np.concatenate([x[0] for x in migs] + [np.zeros(1, np.int32)]),
                         np.int32, C.c_int32)
        s.mig_to = ptr(np.concatenate([x[2] for x in migs] + [np.zeros(1, np.int32)]),
                       np.int32, C.c_int32)
        s.time = ptr(np.concatenate([g.time for g in gens]), np.float64, C.c_double)
        s.mig_time = ptr(np.concatenate([x[1] for x in migs] + [np.zeros(1)]), np.float64,
                         C.c_double)
        s.u = ptr(np.asarray(self.u, float).reshape(Cc * Lc), np.float64, C.c_double)
        s.kappa = ptr(np.asarray(self.kappa, float).reshape(Cc * Lc), np.float64, C.c_double)
        s.beta = ptr(self.beta, np.float64, C.c_double)
        s.tvals = ptr(np.asarray(self.tvals, float).reshape(-1) if P > 1 else np.zeros(1),
                      np.float64, C.c_double)
        s.tmax = float(self.tmax)
        s.seed = int(self.seed)
        s.precision = int(self.precision)
        s.nsweeps = int(self.nsweeps)
        return s, keep


class Server:
    def __init__(self, device=0, budget_bytes=0.0, quantum=0):
        self.lib = L.load()
        h = C.c_void_p()
        rc = self.lib.ima_server_create(device, float(budget_bytes), int(quantum), C.byref(h))
        if rc != 0:
            raise EngineError(f"ima_server_create failed: code {rc}")
        self.h = h
        self.kinds = {}
        self.shapes = {}

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.lib.ima_server_last_error(self.h).decode(errors="replace")
            err = EngineError(f"{what}: code {rc}: {msg}")
            err.code = rc
            raise err

    def close(self):
        if self.h:
            self.lib.ima_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def table(self, tab, npsize, nmig, npops):
        tab = np.ascontiguousarray(tab, np.float32)
        out = C.c_int(-1)
        self._chk(self.lib.ima_server_table(self.h, tab.ctypes.data_as(L.fp), tab.shape[0],
                                            tab.shape[1], npsize, nmig, npops, C.byref(out)),
                  "table")
        return out.value

    def submit_lmode(self, table, x, mode=L.LMODE_EXACT, modeltype=0, prange=None):
        x = np.ascontiguousarray(x, np.float64)
        npts, np_ = x.shape
        r0, r1 = prange if prange else (0, np_)
        rid = C.c_int64()
        self._chk(self.lib.ima_server_submit_lmode(self.h, table, x.ctypes.data_as(L.dp), npts,
                                                   modeltype, r0, r1, mode, C.byref(rid)),
                  "submit_lmode")
        self.kinds[rid.value] = REQ_LMODE
        self.shapes[rid.value] = npts
        return rid.value

    def submit_analysis(self, spec):
        """spec: an AnalysisSpec, or the (struct, arrays) pair of its
        flatten() made ahead of time."""
        s, keep = spec.flatten() if isinstance(spec, AnalysisSpec) else spec
        nch, nl = s.nchains, s.nloci
        rid = C.c_int64()
        self._chk(self.lib.ima_server_submit_analysis(self.h, C.byref(s), C.byref(rid)),
                  "submit_analysis")
        del keep
        self.kinds[rid.value] = REQ_ANALYSIS
        self.shapes[rid.value] = (nch, nl)
        return rid.value

    def step(self):
        pend = C.c_int()
        self._chk(self.lib.ima_server_step(self.h, C.byref(pend)), "step")
        return pend.value

    def run(self, timeout_s=600.0):
        import time
        t0 = time.perf_counter()
        while self.step() > 0:
            if time.perf_counter() - t0 > timeout_s:
                raise TimeoutError("server did not drain")

    def status(self, rid):
        """(state, latency in s or -1, footprint in bytes)."""
        st, lat, b = C.c_int(), C.c_double(), C.c_double()
        self._chk(self.lib.ima_server_status(self.h, C.c_int64(rid), C.byref(st), C.byref(lat),
                                             C.byref(b)), "status")
        return st.value, lat.value, b.value

    def lmode_result(self, rid):
        n = self.shapes[rid]
        out, ess = np.zeros(n), np.zeros(n)
        self._chk(self.lib.ima_server_lmode_result(self.h, C.c_int64(rid), out.ctypes.data_as(L.dp),
                                                   ess.ctypes.data_as(L.dp)), "lmode_result")
        return out, ess

    def analysis_result(self, rid):
        Cc, Lc = self.shapes[rid]
        pdg, probg, beta = np.zeros(Cc * Lc), np.zeros(Cc), np.zeros(Cc)
        stats = np.zeros(8, np.uint64)
        self._chk(self.lib.ima_server_analysis_result(
            self.h, C.c_int64(rid), pdg.ctypes.data_as(L.dp), probg.ctypes.data_as(L.dp),
            beta.ctypes.data_as(L.dp), stats.ctypes.data_as(C.POINTER(C.c_uint64))),
            "analysis_result")
        return pdg.reshape(Cc, Lc), probg, beta, stats

    def release(self, rid):
        self._chk(self.lib.ima_server_release(self.h, C.c_int64(rid)), "release")

    def latency(self, kind=-1):
        cnt, p50, p99, mean = C.c_int(), C.c_double(), C.c_double(), C.c_double()
        self._chk(self.lib.ima_server_latency(self.h, kind, C.byref(cnt), C.byref(p50),
                                              C.byref(p99), C.byref(mean)), "latency")
        return {"count": cnt.value, "p50_ms": p50.value * 1e3, "p99_ms": p99.value * 1e3,
                "mean_ms": mean.value * 1e3}

    def usage(self):
        u, b, w, r = C.c_double(), C.c_double(), C.c_int(), C.c_int()
        self._chk(self.lib.ima_server_usage(self.h, C.byref(u), C.byref(b), C.byref(w),
                                            C.byref(r)), "usage")
        return {"used_bytes": u.value, "budget_bytes": b.value, "waiting": w.value,
                "running": r.value}
