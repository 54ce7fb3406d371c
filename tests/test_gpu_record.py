"""The genealogy table built on the device (ima_mc3_record, savegsampinf
ginfo.cpp:319-377): every recording interval the cold chain's row is
appended to a resident table, which L mode then reads directly.

* Each row equals the oracle's savegsampinf on the chain state it was
  recorded from (the device's all-loci weights, integrals, pdg, probg and
  split times of the chain whose beta is 1), bit for bit as float32; the
  cold chain moves between chains as swaps are accepted.
* The weights the rows come from equal the oracle's treeweight sums over
  the genealogies the chain holds (1e-9).
* jointp and margincalc on the resident table equal the oracle's on the
  rows read back (1e-10): M mode feeds L mode without leaving the device.
* A full table refuses further rows with IMA_ERR_SIZE.
"""
import numpy as np
import pytest

import oracle
from ima2p_amd import _lib as L
from ima2p_amd.engine import EngineError
from tests.test_gpu_mc3_prior import _current, _three_pop_chains

pytestmark = pytest.mark.gpu


def test_rows_equal_savegsampinf_and_feed_lmode():
    C, Lc = 6, 4
    e, model, samp, locs, gens, first, u, ts = _three_pop_chains(C, Lc, 71, L.PREC_FP64)
    P = model.npops
    R = 3 * P * P + 2 * P ** 3 + 1
    npar = model.npsize + model.nmig
    tid = e.lmode_table_reserve(12, model.npsize, model.nmig, P)
    colds = set()
    for r in range(12):
        e.mc3_run(3, L.MC3_GRAPH | L.MC3_SCALE | L.MC3_RY1)
        e.mc3_record(tid)
        pdg, beta, _, _ = e.mc3_read()
        probg, tv = e.mc3_read_prior()
        wall, qint = e.mc3_read_weights(R, npar)
        z = int(np.nonzero(beta == 1.0)[0][-1])
        colds.add(z)
        ref = oracle.gsampinf_row(wall[z], qint[z], pdg[z], probg[z], tv[z], model)
        got = e.lmode_table_read(tid, r, 1)[0]
        assert got.tobytes() == ref.tobytes(), (r, got, ref)
    assert e.lmode_table_rows(tid)[0] == 12
    # the last row's weights are those of the cold chain's genealogies
    _, _, _, tm, _, _ = e.read_slots(first + z * Lc, Lc)
    off, ws = 0, []
    for l in range(Lc):
        g = gens[z * Lc + l]
        ne = 2 * g.n - 1
        h = _current(e, g, first + z * Lc + l, tm[off:off + ne])
        off += ne
        ws.append(oracle.treeweight(h, model, tv[z], samp))
    fresh = np.concatenate([sum(w["cc"] for w in ws).ravel(), sum(w["fc"] for w in ws).ravel(),
                            sum(w["hcc"] for w in ws).ravel(), sum(w["mc"] for w in ws).ravel(),
                            sum(w["fm"] for w in ws).ravel(), [sum(w["length"] for w in ws)]])
    np.testing.assert_allclose(wall[z], fresh, rtol=1e-9, atol=1e-9)
    # L mode on the resident table
    tab = e.lmode_table_read(tid)
    rng = np.random.default_rng(3)
    x = np.column_stack([rng.uniform(1, 12, (20, model.npsize)),
                         rng.uniform(0.05, 2, (20, model.nmig))])
    for mt, pr in ((1, (0, model.npsize)), (2, (model.npsize, npar))):
        out, _ = e.jointp(tid, x, mode=L.LMODE_EXACT, modeltype=mt, prange=pr)
        rc, _ = e.jointp(tid, x, mode=L.LMODE_REF_COMPAT, modeltype=mt, prange=pr)
        for i in range(len(x)):
            ref = oracle.jointp_exact(tab, model.npsize, model.nmig, P, x[i], modeltype=mt, prange=pr)
            assert abs(out[i] - ref) <= 1e-10 * abs(ref)
            ref = oracle.jointp(tab, model.npsize, model.nmig, P, x[i], modeltype=mt, prange=pr)
            assert abs(rc[i] - ref) <= 1e-10 * abs(ref)
    xs = np.linspace(0.5, 10, 15)
    got = e.margincalc(tid, 0, xs)
    for j in range(len(xs)):
        ref = oracle.margincalc(tab, model.npsize, model.nmig, xs[j], 0)
        assert abs(got[j] - ref) <= 1e-10 * max(1.0, abs(ref))
    with pytest.raises(EngineError) as ei:
        e.mc3_record(tid)
    assert ei.value.code == L.ERR_SIZE
    e.close()
