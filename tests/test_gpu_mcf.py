"""Checkpoint and resume through .mcf files (writemcf / readmcf /
init_p, mcmcfile.cpp:131-442): chains with migration events run sweeps
(time slides, rubber-band split-time updates, rescaling of migration
times), their state is written with Engine.mcf_save, a fresh engine reads
it with Engine.mcf_load and sets the chains up again.  The genealogies come
back with the same topology, populations and migration events and with
times to the file's ten significant digits, and the fresh engine's
likelihoods and genealogy prior equal the saved chains' to 1e-8."""
import numpy as np

from ima2p_amd import _lib as L
from ima2p_amd.engine import Engine
from tests.test_gpu_mc3_prior import _three_pop_chains

import pytest

pytestmark = pytest.mark.gpu


def test_save_and_resume_reproduces_the_state(tmp_path):
    C, Lc = 4, 3
    e, model, samp, locs, gens, first, u, ts = _three_pop_chains(C, Lc, 83, L.PREC_FP64)
    e.mc3_run(30, L.MC3_GRAPH | L.MC3_SCALE | L.MC3_RY1)
    pdg0, beta0, _, _ = e.mc3_read()
    probg0, tv0 = e.mc3_read_prior()
    layout = [(L.MODEL_HKY, loc.n, loc.seq.shape[1]) for loc in locs]
    pi = np.array([loc.pi for loc in locs])
    path = tmp_path / "chains.mcf"
    e.mcf_save(path, first, C, Lc, layout, pi)
    g0 = e.read_slots(first, C * Lc)
    pop0 = e.read_pops(first, C * Lc)
    mig0 = [e.read_mig_events(first + i) for i in range(C * Lc)]
    urate, kap = e.mc3_read_rates()
    e.close()

    f = Engine(0)
    mid = f.model_add(model)
    lids = []
    for loc in locs:
        lid = f.add_locus_hky(loc)
        f.locus_set_model(lid, mid, samp)
        lids.append(lid)
    first2 = f.alloc_slots([lids[l] for c in range(C) for l in range(Lc)])
    st = f.mcf_load(path, first2, C, Lc, layout, tmax=np.full(model.npops - 1, 10.0))
    g1 = f.read_slots(first2, C * Lc)
    for k in (0, 1, 2, 4):  # up0, up1, down, root
        np.testing.assert_array_equal(g1[k], g0[k])
    np.testing.assert_allclose(g1[3], g0[3], rtol=6e-10)
    np.testing.assert_array_equal(f.read_pops(first2, C * Lc), pop0)
    for i in range(C * Lc):
        e1, t1, o1 = f.read_mig_events(first2 + i)
        np.testing.assert_array_equal(e1, mig0[i][0])
        np.testing.assert_array_equal(o1, mig0[i][2])
        np.testing.assert_allclose(t1, mig0[i][1], rtol=6e-10)
    np.testing.assert_allclose(st["u"], urate, rtol=6e-10)
    np.testing.assert_allclose(st["kappa"], kap, rtol=6e-10)
    np.testing.assert_allclose(st["tvals"], tv0, rtol=6e-10)
    f.mc3_setup(C, Lc, first2, st["u"].ravel(), st["kappa"].ravel(), beta0, seed=5,
                precision=L.PREC_FP64)
    f.mc3_set_prior(st["tvals"], 10.0, model.npops)
    pdg1, _, _, _ = f.mc3_read()
    probg1, _ = f.mc3_read_prior()
    np.testing.assert_allclose(pdg1, pdg0, rtol=1e-8)
    np.testing.assert_allclose(probg1, probg0, rtol=1e-8)
    f.close()
