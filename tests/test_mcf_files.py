"""MCMC state files (.mcf): ima_mcf_write / ima_mcf_read against the
reference's layout (writemcf / readmcf, mcmcfile.cpp:33-442).  The golden
text below is written out by hand from awrite's format rules (one item a
line: "<name> <type> <count> ", each value "%d " or "%.10lg ", newline; the
items and their order from writemcf); the round trips, the reread from the
file's start when it holds fewer chains than asked for, the skip of chains
whose split times exceed the prior, and the name/type checks follow
readmcf.  Host only, no device."""
import numpy as np
import pytest

from ima2p_amd import _lib as L
from ima2p_amd import mcffile

HKY, IS, SW = L.MODEL_HKY, L.MODEL_IS, L.MODEL_SW

GOLDEN = """tvalue 3 1 1.25 
uvalue 3 1 0.5 
kappavalue 3 1 4 
pi[4] 3 4 0.1 0.2 0.3 0.4 
up[2] 0 2 -1 -1 
down 0 1 2 
mut 0 1 0 
pop 0 1 0 
time 3 1 0.7 
mig[].mt 3 1 0.3 
mig[].mp 0 1 1 
mig[].mt 3 1 -1 
up[2] 0 2 -1 -1 
down 0 1 2 
mut 0 1 0 
pop 0 1 1 
time 3 1 0.7 
mig[].mt 3 1 -1 
up[2] 0 2 0 1 
down 0 1 -1 
mut 0 1 0 
pop 0 1 2 
time 3 1 1e+100 
mig[].mt 3 1 -1 
C[ci]->G[li].gtree[i].hkyi.frac[j] 3 4 0 0 0 0 
C[ci]->G[li].gtree[i].hkyi.newfrac[j] 3 4 0 0 0 0 
C[ci]->G[li].gtree[i].hkyi.frac[j] 3 4 0 0 0 0 
C[ci]->G[li].gtree[i].hkyi.newfrac[j] 3 4 0 0 0 0 
uvalue 3 1 1.123456789 
pi[4] 3 4 0.25 0.25 0.25 0.25 
up[2] 0 2 -1 -1 
down 0 1 2 
mut 0 1 0 
pop 0 1 0 
A[] 0 1 12 
dlikeA[] 3 1 -2.5 
time 3 1 0.2 
mig[].mt 3 1 -1 
up[2] 0 2 -1 -1 
down 0 1 2 
mut 0 1 0 
pop 0 1 0 
A[] 0 1 14 
dlikeA[] 3 1 -1.75 
time 3 1 0.2 
mig[].mt 3 1 -1 
up[2] 0 2 0 1 
down 0 1 -1 
mut 0 1 0 
pop 0 1 0 
A[] 0 1 13 
dlikeA[] 3 1 0 
time 3 1 1e+100 
mig[].mt 3 1 -1 
"""

LOCI = [(HKY, 2, 2), (SW, 2, 0)]


def _golden_state():
    return {
        "tvals": [1.25], "u": [0.5, 1.1234567891234], "kappa": [4.0, 0.0],
        "pi": [0.1, 0.2, 0.3, 0.4, 0.25, 0.25, 0.25, 0.25],
        "up0": [-1, -1, 0, -1, -1, 0], "up1": [-1, -1, 1, -1, -1, 1],
        "down": [2, 2, -1, 2, 2, -1], "pop": [0, 1, 2, 0, 0, 0],
        "time": [0.7, 0.7, 1e100, 0.2, 0.2, 1e100],
        "A": [0, 0, 0, 12, 14, 13], "dlikeA": [0, 0, 0, -2.5, -1.75, 0.0],
        "mig_count": [1, 0, 0, 0, 0, 0], "mig_time": [0.3], "mig_to": [1],
    }


def test_writer_matches_the_reference_layout(tmp_path):
    p = tmp_path / "a.mcf"
    mcffile.write(p, 1, 1, LOCI, _golden_state())
    assert p.read_text() == GOLDEN


def test_reader_parses_the_golden_text(tmp_path):
    p = tmp_path / "g.mcf"
    p.write_text(GOLDEN)
    st = mcffile.read(p, 1, 1, LOCI)
    g = _golden_state()
    assert st["tvals"].tolist() == [[1.25]]
    assert st["u"][0, 1] == 1.123456789  # 10 significant digits survive
    for k in ("up0", "up1", "down", "pop", "A", "mig_count"):
        assert st[k].tolist() == g[k], k
    np.testing.assert_array_equal(st["time"], g["time"])
    np.testing.assert_array_equal(st["dlikeA"], g["dlikeA"])
    assert st["mig_time"].tolist() == [0.3] and st["mig_to"].tolist() == [1]


def _random_state(rng, C, nsplit, loci):
    E = C * sum(2 * n - 1 for _, n, _ in loci)
    mig_count = rng.poisson(0.7, E).astype(np.int32)
    M = int(mig_count.sum())
    return {
        "tvals": rng.uniform(0.1, 5, (C, nsplit)), "u": rng.lognormal(0, 0.5, (C, len(loci))),
        "kappa": rng.uniform(1, 9, (C, len(loci))), "pi": rng.dirichlet([5] * 4, (C, len(loci))),
        "up0": rng.integers(-1, 40, E), "up1": rng.integers(-1, 40, E),
        "down": rng.integers(-1, 40, E), "pop": rng.integers(0, 5, E),
        "mut": rng.integers(0, 4, E), "A": rng.integers(1, 60, E),
        "time": rng.uniform(0, 10, E), "dlikeA": rng.normal(-3, 1, E),
        "mig_count": mig_count, "mig_time": rng.uniform(0, 10, M),
        "mig_to": rng.integers(0, 5, M),
    }


def test_round_trip_keeps_ten_significant_digits(tmp_path):
    rng = np.random.default_rng(1)
    loci = [(HKY, 7, 5), (IS, 4, 0), (SW, 5, 0), (HKY, 3, 1)]
    C, ns = 3, 2
    st = _random_state(rng, C, ns, loci)
    p = tmp_path / "r.mcf"
    mcffile.write(p, C, ns, loci, st)
    back = mcffile.read(p, C, ns, loci)
    for k in ("up0", "up1", "down", "pop", "mut", "mig_count", "mig_to"):
        np.testing.assert_array_equal(back[k], np.asarray(st[k]).ravel(), err_msg=k)
    sw = np.concatenate([[m == SW] * (2 * n - 1) for m, n, _ in loci] * C)
    np.testing.assert_array_equal(back["A"][sw], np.asarray(st["A"])[sw])
    for k in ("tvals", "u", "pi", "time", "mig_time"):
        np.testing.assert_allclose(back[k].ravel(), np.asarray(st[k]).ravel(), rtol=6e-10,
                                   err_msg=k)
    hk = np.array([[m == HKY for m, _, _ in loci]] * C)
    np.testing.assert_allclose(back["kappa"][hk], st["kappa"][hk], rtol=6e-10)
    np.testing.assert_allclose(back["dlikeA"][sw], np.asarray(st["dlikeA"])[sw], rtol=6e-10)
    # the frac lines of every internal HKY edge are present
    txt = p.read_text()
    assert txt.count("hkyi.frac[j]") == C * (6 * 5 + 2 * 1)


def test_fewer_chains_in_the_file_are_read_again(tmp_path):
    rng = np.random.default_rng(2)
    loci = [(IS, 3, 0)]
    st = _random_state(rng, 2, 1, loci)
    p = tmp_path / "w.mcf"
    mcffile.write(p, 2, 1, loci, st)
    back = mcffile.read(p, 5, 1, loci)
    np.testing.assert_allclose(back["tvals"][:, 0], st["tvals"][[0, 1, 0, 1, 0], 0], rtol=6e-10)
    E = 5
    np.testing.assert_array_equal(back["down"].reshape(5, E),
                                  np.asarray(st["down"]).reshape(2, E)[[0, 1, 0, 1, 0]])


def test_chains_beyond_the_split_time_prior_are_skipped(tmp_path):
    rng = np.random.default_rng(3)
    loci = [(IS, 3, 0)]
    st = _random_state(rng, 4, 1, loci)
    st["tvals"][:, 0] = [1.0, 9.0, 2.0, 3.0]  # chain 1 exceeds tmax = 5
    st["mig_count"][:] = 0
    st["mig_time"] = st["mig_time"][:0]
    st["mig_to"] = st["mig_to"][:0]
    p = tmp_path / "t.mcf"
    mcffile.write(p, 4, 1, loci, st)
    back = mcffile.read(p, 3, 1, loci, tmax=[5.0])
    assert back["tvals"][:, 0].tolist() == [1.0, 2.0, 3.0]
    # every chain out of range: the second pass over the file fails
    st["tvals"][:, 0] = 9.0
    mcffile.write(p, 4, 1, loci, st)
    with pytest.raises(ValueError):
        mcffile.read(p, 2, 1, loci, tmax=[5.0])


def test_layout_mismatch_is_refused(tmp_path):
    p = tmp_path / "g.mcf"
    p.write_text(GOLDEN)
    with pytest.raises(ValueError):
        mcffile.read(p, 1, 1, [(IS, 2, 0), (SW, 2, 0)])   # no kappavalue expected
    with pytest.raises(ValueError):
        mcffile.read(p, 1, 2, LOCI)                       # two split times
    with pytest.raises(ValueError):
        mcffile.read(p, 1, 1, LOCI, mig_capacity=0)       # one migration event
