"""ctypes front end of libima_plain.so: the plain per-locus FP64 path, the
MC3 genealogy sweep driven from the host one proposal at a time through the
engine's C ABI (one slot upload, one FP64 likelihood launch, one treeweight
launch and one prior integral per proposal, as IMa2p's updategenealogy
calls its likelihood and prior functions).  A measurement baseline for the
batched device sweep, not part of the engine's ABI."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L
from .immodel import mig_arrays

HERE = os.path.dirname(os.path.abspath(__file__))
PLAIN_PATH = os.path.join(HERE, "libima_plain.so")

_plain = None


def load():
    global _plain
    if _plain is None:
        L.load()  # libima_gpu.so first (the plain library links against it)
        if not os.path.exists(PLAIN_PATH):
            raise RuntimeError(f"{PLAIN_PATH} is missing: build with `python -m ima2p_amd.build`")
        lib = C.CDLL(PLAIN_PATH)
        f = lib.ima_plain_mc3_run
        f.restype = C.c_int
        i32p, dp = L.i32p, L.dp
        f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                      i32p, i32p, i32p, dp, i32p, i32p, i32p, i32p, dp, i32p,
                      dp, dp, dp, dp, C.c_uint64, C.c_int, C.c_int,
                      dp, dp, dp, dp, C.POINTER(C.c_uint64), dp]
        _plain = lib
    return _plain


def run(engine, first, model_id, P, gens, nchains, nloci, u, kappa, beta, tvals, seed=7,
        nsweeps=2, swaptries=None):
    """Runs nsweeps plain sweeps over gens (chain-major Genealogy list already
    written to slots first..); model_id < 0 without the prior.  Returns a
    dict: seconds, pdg [C, L], probg [C], beta [C], accepts, times (packed)."""
    lib = load()
    N = nchains * nloci

    def i32(a):
        a = np.ascontiguousarray(a, np.int32)
        return a, a.ctypes.data_as(L.i32p)

    def f64(a):
        a = np.ascontiguousarray(a, np.float64)
        return a, a.ctypes.data_as(L.dp)

    keep = []

    def ai(a):
        x, p = i32(a)
        keep.append(x)
        return p

    def ad(a):
        x, p = f64(a)
        keep.append(x)
        return p

    me, mt, mo, nm = [], [], [], []
    for g in gens:
        e, t, to = mig_arrays(g)
        me.append(e)
        mt.append(t)
        mo.append(to)
        nm.append(len(e))
    if swaptries is None:
        swaptries = max(1, nchains // 10)
    secs = C.c_double()
    pdg = np.zeros(N)
    probg = np.zeros(nchains)
    bout = np.zeros(nchains)
    acc = C.c_uint64()
    times = np.zeros(sum(2 * g.n - 1 for g in gens))
    tv = np.asarray(tvals, float).reshape(nchains, -1) if P > 1 else np.zeros((nchains, 1))
    rc = lib.ima_plain_mc3_run(
        engine.h, nchains, nloci, int(first), int(model_id), int(P),
        ai(np.concatenate([g.up0 for g in gens])), ai(np.concatenate([g.up1 for g in gens])),
        ai(np.concatenate([g.down for g in gens])), ad(np.concatenate([g.time for g in gens])),
        ai([g.root for g in gens]), ai(np.concatenate([g.pop for g in gens])), ai(nm),
        ai(np.concatenate(me) if sum(nm) else [0]), ad(np.concatenate(mt) if sum(nm) else [0.0]),
        ai(np.concatenate(mo) if sum(nm) else [0]),
        ad(np.broadcast_to(np.asarray(u, float), (N,))),
        ad(np.broadcast_to(np.asarray(kappa, float), (N,))), ad(beta), ad(tv),
        int(seed), int(nsweeps), int(swaptries), C.byref(secs), pdg.ctypes.data_as(L.dp),
        probg.ctypes.data_as(L.dp), bout.ctypes.data_as(L.dp), C.byref(acc),
        times.ctypes.data_as(L.dp))
    if rc != 0:
        msg = engine.lib.ima_ctx_last_error(engine.h).decode(errors="replace")
        raise RuntimeError(f"ima_plain_mc3_run failed ({rc}): {msg}")
    return {"seconds": secs.value, "pdg": pdg.reshape(nchains, nloci), "probg": probg,
            "beta": bout, "accepts": acc.value, "times": times}
