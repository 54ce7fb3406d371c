"""ima_lmode_jointp_noalloc: jointp over device-resident points with the
library's static per-device work area (no device allocation in the call),
checked against ima_lmode_jointp on the same table, bit for bit (same
kernels, same chunking), on two calls in a row -- the second with fewer
points -- in both modes, with and without the effective sample numbers;
more points than the work area holds are refused with IMA_ERR_SIZE.
The test is the caller: it allocates its device buffers itself through the
HIP runtime the library is linked against, and synchronises the device
before reading the results (the entry point returns without waiting)."""
import ctypes as C

import numpy as np
import pytest

from ima2p_amd import _lib as L
from ima2p_amd import sim
from ima2p_amd.engine import Engine

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class DevBuf:
    """A caller-owned device buffer of float64 values."""

    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(max(n, 1) * 8)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, np.float64)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes),
                                  H2D) == 0

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p,
                                  C.c_size_t(out.nbytes), D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(lib, hip, e, tid, x, mode, with_ess=True):
    xd, out, ess = DevBuf(hip, x.size), DevBuf(hip, len(x)), DevBuf(hip, len(x))
    xd.put(x)
    out.put(np.full(len(x), np.nan))
    ess.put(np.full(len(x), np.nan))
    assert hip.hipDeviceSynchronize() == 0
    rc = lib.ima_lmode_jointp_noalloc(e.h, tid, xd.p, len(x), 0, 0, x.shape[1], mode, out.p,
                                      ess.p if with_ess else None)
    assert hip.hipDeviceSynchronize() == 0  # waits for the context's stream too
    res = (rc, out.get(), ess.get())
    for b in (xd, out, ess):
        b.free()
    return res


def test_noalloc_jointp_equals_jointp_twice():
    lib = L.load()
    hip = C.CDLL("libamdhip64.so.7")
    rng = np.random.default_rng(9)
    e = Engine(0)
    tab = sim.lmode_table(rng, 100000)
    tid = e.lmode_table(tab, 3, 2, 2)
    lo, hi = [0.5, 0.5, 0.5, 0.01, 0.01], [10, 10, 10, 2, 2]
    for mode in (L.LMODE_EXACT, L.LMODE_REF_COMPAT):
        for npts in (600, 37):  # two calls in a row, the second smaller
            x = rng.uniform(lo, hi, (npts, 5))
            rc, out, ess = _noalloc(lib, hip, e, tid, x, mode)
            assert rc == 0, e.lib.ima_ctx_last_error(e.h)
            ref, ref_ess = e.jointp(tid, x, mode=mode)
            np.testing.assert_array_equal(out, ref)
            np.testing.assert_array_equal(ess, ref_ess)
        rc, out, ess = _noalloc(lib, hip, e, tid, x, mode, with_ess=False)
        assert rc == 0
        np.testing.assert_array_equal(out, ref)
        assert np.isnan(ess).all()  # not written
    # more points than the 16 MiB work area holds at 1e5 genealogies
    x = rng.uniform(lo, hi, (4096, 5))
    rc, out, _ = _noalloc(lib, hip, e, tid, x, L.LMODE_EXACT)
    assert rc == L.ERR_SIZE and np.isnan(out).all()
    assert b"work area" in e.lib.ima_ctx_last_error(e.h)
    e.close()
