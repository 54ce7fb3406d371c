"""Host-only planning of the sweep's per-chain prior kernels (no device).

ima_mc3_prior_plan returns the accept group size ima_mc3_set_prior picks
for a population model and the dynamic LDS of the per-chain kernels at that
size.  Every launch stays within 64 KiB of dynamic LDS, whatever the device
allows per workgroup: models of four and five populations, whose weight
records are 177 and 326 doubles long, get smaller groups instead of larger
launches.  The sizes are the reference's default parameter maps
(setup_iparams, initialize.cpp:202-760) for ladder trees of 2..5
populations.
"""
import ctypes as C

import pytest

from ima2p_amd import _lib as L
from ima2p_amd import immodel
from ima2p_amd.sim import PopTree


def _ladder(P):
    joins, cur = [], 0
    for k in range(P - 1):
        joins.append((cur, k + 1))
        cur = P + k
    return PopTree(P, joins)


def _plan(model):
    lib = L.load()
    loci, lds = C.c_int(), C.c_int64()
    rc = lib.ima_mc3_prior_plan(model.npops, model.npsize, sum(len(w) for w in model.theta_wp),
                                model.nmig, sum(len(w) for w in model.mig_wp),
                                C.byref(loci), C.byref(lds))
    return rc, loci.value, lds.value


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
def test_group_size_fits_64k(P):
    model = immodel.build(_ladder(P) if P > 1 else PopTree(1, []))
    rc, loci, lds = _plan(model)
    assert rc == L.IMA_OK
    assert 1 <= loci <= 6
    assert 0 < lds <= 65536
    if P <= 3:
        assert loci == 6  # the headline models keep the largest group


def test_larger_models_get_smaller_groups():
    sizes = [_plan(immodel.build(_ladder(P)))[1] for P in (3, 4, 5)]
    assert sizes[0] >= sizes[1] >= sizes[2]
    assert sizes[2] < 6


def test_group_choice_is_maximal():
    """One more locus per group would exceed 64 KiB (or the group cap of 6):
    the planner shrinks groups only as far as the limit needs."""
    for P in (4, 5):
        model = immodel.build(_ladder(P))
        rc, loci, lds = _plan(model)
        assert rc == L.IMA_OK
        if loci < 6:
            # LDS grows with the group: 2^D scenario slots per parameter
            # plus D rows of records, so D + 1 must not fit
            per_locus_rows = 6 * (3 * P * P + 2 * P ** 3 + 1) * 8
            assert lds + per_locus_rows > 65536 or lds * 2 > 65536


def test_plan_rejects_bad_sizes():
    lib = L.load()
    loci, lds = C.c_int(), C.c_int64()
    assert lib.ima_mc3_prior_plan(0, 1, 1, 0, 0, C.byref(loci), C.byref(lds)) == L.ERR_ARG
    assert lib.ima_mc3_prior_plan(11, 1, 1, 0, 0, C.byref(loci), C.byref(lds)) == L.ERR_ARG
