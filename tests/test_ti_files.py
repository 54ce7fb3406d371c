"""The .ti genealogy-table files (host only, no device):

* the writer's layout is the reference's: header block, a VALUESSTART
  line, then one line per genealogy of "%.6f\\t" values
  (ima_main_mpi.cpp:2121-2142, output.cpp:662-685);
* loading gives each value as the float nearest its 6-decimal text;
* appending adds rows; a row with too few or too many values is an error,
  and so is a file without VALUESSTART;
* with more genealogies than wanted, the loader keeps a spread subsequence
  in file order across files, the first always, exactly `wanted` rows.
"""
import numpy as np
import pytest

from ima2p_amd import sim, tifile


def test_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    tab = sim.lmode_table(rng, 50)
    p = tmp_path / "run.ti"
    tifile.write(p, tab, header="model: 2 populations")
    text = p.read_text()
    head, values = text.split("VALUESSTART\n")
    assert "model: 2 populations" in head and head.startswith("-----")
    lines = values.split("\n")
    assert lines[-1] == "" and len(lines) == 51
    first = lines[0].split("\t")
    assert first[-1] == "" and len(first) == tab.shape[1] + 1
    assert first[0] == "%.6f" % float(tab[0, 0])
    rows, total = tifile.load(p, tab.shape[1])
    assert total == 50 and rows.shape == tab.shape
    ref = np.array([[np.float32(float("%.6f" % float(v))) for v in r] for r in tab], np.float32)
    np.testing.assert_array_equal(rows, ref)
    tifile.write(p, tab[:5], append=True)
    assert tifile.count(p) == 55
    rows, _ = tifile.load(p, tab.shape[1])
    np.testing.assert_array_equal(rows[50:], ref[:5])


def test_malformed_files(tmp_path):
    tab = np.ones((3, 4), np.float32)
    p = tmp_path / "a.ti"
    tifile.write(p, tab)
    with pytest.raises(ValueError):
        tifile.load(p, 5)  # too few values per genealogy
    with pytest.raises(ValueError):
        tifile.load(p, 3)  # too many
    q = tmp_path / "b.ti"
    q.write_text("no marker here\n1.0\t2.0\n")
    with pytest.raises(OSError):
        tifile.count(q)


def test_subsampling_across_files(tmp_path):
    sizes = [10, 7, 13]
    paths, allrows, k = [], [], 0
    for i, n in enumerate(sizes):
        tab = np.arange(k, k + n, dtype=np.float32)[:, None] * np.ones((1, 3), np.float32)
        k += n
        p = tmp_path / f"r{i}.ti"
        tifile.write(p, tab)
        paths.append(p)
        allrows.append(tab)
    allrows = np.concatenate(allrows)
    rows, total = tifile.load(paths, 3, max_rows=12)
    assert total == 30 and len(rows) == 12
    idx = rows[:, 0].astype(int)
    assert idx[0] == 0 and np.all(np.diff(idx) > 0)
    assert np.diff(idx).max() <= 4  # spread over all three files
    assert idx[-1] >= 25
    np.testing.assert_array_equal(rows, allrows[idx])
