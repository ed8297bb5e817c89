"""Readers of the reference's on-disk formats (orac_hip.formats): the Fortran
list-directed reading rules they rest on, and the restart record layout, on
small synthetic files.  The readers on the reference's own files are covered
by tests/test_oracle_fkbp12.py (the test1 system parsed from 1j4i.tpgprm and
md.rst reproduces the reference's forces bit for bit)."""
import numpy as np

from orac_hip import formats


def _reader(tmp_path, text):
    p = tmp_path / "f.txt"
    p.write_text(text)
    return formats._ListDirected(str(p))


def test_each_read_starts_a_new_line_and_drops_the_rest(tmp_path):
    r = _reader(tmp_path, "1 2 3 4\n5 6\n7 8 9\n10\n")
    assert list(r.ints(2)) == [1, 2]          # 3 4 are dropped
    assert list(r.ints(3)) == [5, 6, 7]       # spans two lines, 8 9 dropped
    assert list(r.ints(1)) == [10]


def test_repeat_counts_and_d_exponents(tmp_path):
    r = _reader(tmp_path, "3*2 5\n1.5D+01 2*-0.25d0 4.0E-1\n")
    assert list(r.ints(4)) == [2, 2, 2, 5]
    assert np.allclose(r.reals(4), [15.0, -0.25, -0.25, 0.4])


def test_a_read_of_nothing_consumes_a_line(tmp_path):
    r = _reader(tmp_path, "skipped\n42\n")
    assert r.values(0) == []
    assert list(r.ints(1)) == [42]


def test_counted_lists_cross_lines(tmp_path):
    r = _reader(tmp_path, "2 10 11 0 3\n20 21\n22 1 30\nnext\n")
    assert r.counted_lists(4) == [[10, 11], [], [20, 21, 22], [30]]
    assert r.values(1) == ["next"]


def test_restart_layout(tmp_path):
    co = np.array([[10.0, 0.5, 0.25], [0.0, 11.0, 0.75], [0.0, 0.0, 12.0]])
    oc = np.linalg.inv(co)
    x = np.arange(12, dtype=float).reshape(3, 4) * 0.5
    v = -x / 10.0
    lines = [" ".join(f"{c:.17g}" for c in co.ravel()), " ".join(f"{c:.17g}" for c in oc.ravel()),
             "1500 300.0", "4 2 1",
             " ".join(f"{c:.17g}" for c in np.concatenate([x.ravel(), v.ravel(), np.zeros(3 * 2 + 6 * 1)]))]
    p = tmp_path / "r.rst"
    p.write_text("\n".join(lines) + "\n")
    rs = formats.read_restart(str(p))
    # the cell is written row by row: co(1,2) is the second value
    assert np.array_equal(rs["co"], co) and rs["co"][0, 1] == 0.5
    assert np.allclose(rs["oc"], oc, rtol=0, atol=1e-15)
    assert rs["nstep"] == 1500 and rs["temp"] == 300.0 and rs["ngrp"] == 2 and rs["nprot"] == 1
    assert np.array_equal(rs["x"], x) and np.array_equal(rs["v"], v)
