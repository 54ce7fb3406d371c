"""Genealogy-table files (.ti) through the engine's host-only entry points
ima_ti_write / ima_ti_count / ima_ti_load (the reference's
savegenealogyfile, output.cpp:662-685, and loadgenealogyvalues,
ima_main_mpi.cpp:3216-3440).  No device is needed."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def write(path, rows, header="", append=False):
    lib = L.load()
    rows = np.ascontiguousarray(rows, np.float32)
    G, K = rows.shape
    rc = lib.ima_ti_write(str(path).encode(), header.encode(), rows.ctypes.data_as(L.fp), G, K,
                          int(append))
    if rc != 0:
        raise OSError(f"ima_ti_write {path}: code {rc}")


def count(path):
    lib = L.load()
    g = C.c_int()
    rc = lib.ima_ti_count(str(path).encode(), C.byref(g))
    if rc != 0:
        raise OSError(f"ima_ti_count {path}: code {rc}")
    return g.value


def load(paths, K, max_rows=300000):
    """Rows of the files, in order, subsampled as the reference does when
    more than max_rows (MAXGENEALOGIESTOSAVE = 300000 by default) are
    present.  Returns (rows [loaded, K] float32, total genealogies)."""
    lib = L.load()
    if isinstance(paths, (str, bytes)) or hasattr(paths, "__fspath__"):
        paths = [paths]
    arr = (C.c_char_p * len(paths))(*[str(p).encode() for p in paths])
    total = sum(count(p) for p in paths)
    cap = max(1, min(total, max_rows))
    out = np.zeros((cap, K), np.float32)
    loaded, tot = C.c_int(), C.c_int()
    rc = lib.ima_ti_load(arr, len(paths), K, cap, out.ctypes.data_as(L.fp), C.byref(loaded),
                         C.byref(tot))
    if rc != 0:
        raise ValueError(f"ima_ti_load: code {rc} (missing VALUESSTART, or a genealogy with "
                         f"other than {K} values)")
    return out[:loaded.value], tot.value
