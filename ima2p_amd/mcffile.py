"""MCMC state files (.mcf) through the engine's host-only entry points
ima_mcf_write / ima_mcf_read (the reference's writemcf / readmcf,
mcmcfile.cpp:204-442).  No device is needed.

A state is a dict of numpy arrays, packed chain after chain and locus
after locus within a chain (edges tips first, 2n - 1 per genealogy):
tvals [C, nsplit], u / kappa [C, L], pi [C, L, 4], up0 / up1 / down / pop /
mut / A [edges] int32, time / dlikeA [edges] float64, mig_count [edges] and
mig_time / mig_to [events] in edge order, each edge's events in time order.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

_INT = ("up0", "up1", "down", "mut", "pop", "A", "mig_count", "mig_to")
_DBL = ("tvals", "u", "kappa", "pi", "time", "dlikeA", "mig_time")


def _loci(loci):
    arr = (L.McfLocus * len(loci))()
    for i, (model, n, nsites) in enumerate(loci):
        arr[i].model, arr[i].n, arr[i].nsites = int(model), int(n), int(nsites)
    return arr


def _edges(nchains, loci):
    return nchains * sum(2 * n - 1 for _, n, _ in loci)


def write(path, nchains, nsplit, loci, state):
    """loci: [(model, n, nsites)], model IMA_MODEL_HKY / IS / SW."""
    lib = L.load()
    keep = {}
    st = L.McfState()
    for k in _INT:
        if state.get(k) is not None:
            keep[k] = np.ascontiguousarray(state[k], np.int32)
            setattr(st, k, keep[k].ctypes.data_as(C.POINTER(C.c_int32)))
    for k in _DBL:
        if state.get(k) is not None:
            keep[k] = np.ascontiguousarray(state[k], np.float64)
            setattr(st, k, keep[k].ctypes.data_as(C.POINTER(C.c_double)))
    rc = lib.ima_mcf_write(str(path).encode(), int(nchains), int(nsplit), len(loci),
                           C.cast(_loci(loci), C.c_void_p), C.cast(C.pointer(st), C.c_void_p))
    if rc != 0:
        raise OSError(f"ima_mcf_write {path}: code {rc}")


def read(path, nchains, nsplit, loci, tmax=None, mig_capacity=1 << 20):
    """The state of nchains chains (see the module's docstring)."""
    lib = L.load()
    E = _edges(nchains, loci)
    nl = len(loci)
    out = {"tvals": np.zeros(nchains * nsplit), "u": np.zeros(nchains * nl),
           "kappa": np.zeros(nchains * nl), "pi": np.zeros(nchains * nl * 4),
           "time": np.zeros(E), "dlikeA": np.zeros(E), "mig_time": np.zeros(mig_capacity)}
    for k in _INT:
        out[k] = np.zeros(mig_capacity if k == "mig_to" else E, np.int32)
    st = L.McfState()
    for k in _INT:
        setattr(st, k, out[k].ctypes.data_as(C.POINTER(C.c_int32)))
    for k in _DBL:
        setattr(st, k, out[k].ctypes.data_as(C.POINTER(C.c_double)))
    st.mig_capacity = mig_capacity
    tm = None if tmax is None else np.ascontiguousarray(tmax, np.float64)
    nm = C.c_int64()
    rc = lib.ima_mcf_read(str(path).encode(), int(nchains), int(nsplit), nl,
                          C.cast(_loci(loci), C.c_void_p), None if tm is None else
                          tm.ctypes.data_as(L.dp), C.cast(C.pointer(st), C.c_void_p),
                          C.byref(nm))
    if rc == L.ERR_SIZE:
        raise ValueError(f"ima_mcf_read {path}: more than {mig_capacity} migration events")
    if rc != 0:
        raise ValueError(f"ima_mcf_read {path}: code {rc} (an item's name or type does not "
                         f"match the layout, or the file is short)")
    out["mig_time"] = out["mig_time"][:nm.value]
    out["mig_to"] = out["mig_to"][:nm.value]
    out["tvals"] = out["tvals"].reshape(nchains, nsplit)
    out["u"] = out["u"].reshape(nchains, nl)
    out["kappa"] = out["kappa"].reshape(nchains, nl)
    out["pi"] = out["pi"].reshape(nchains, nl, 4)
    return out
