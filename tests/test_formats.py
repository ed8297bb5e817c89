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


def _state(rng, ntap=7, ngrp=3, nprot=2):
    return dict(co=rng.standard_normal((3, 3)), oc=rng.standard_normal((3, 3)), nstep=123456, temp=299.75,
                x=rng.standard_normal((3, ntap)) * 30, v=rng.standard_normal((3, ntap)) * 1e-3,
                xpg=rng.standard_normal((3, ngrp)), xcm=rng.standard_normal((3, nprot)),
                vcm=rng.standard_normal((3, nprot)) * 1e-4, v1=rng.standard_normal((3, ntap)) * 1e-3,
                avg1=rng.standard_normal(5), avg2=rng.standard_normal(4))


def _same_state(a, b):
    for k in ("co", "oc", "x", "v", "xpg", "xcm", "vcm", "v1", "avg1", "avg2"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    assert a["nstep"] == b["nstep"] and a["temp"] == b["temp"]
    for k in ("hoover", "cpress", "igr"):
        assert bool(a.get(k, False)) == bool(b.get(k, False)), k


def test_restart_writer_round_trip(tmp_path):
    """write_restart then read_restart_full gives back every record, with the
    thermostat, barostat and rdf records switched on and off."""
    rng = np.random.default_rng(3)
    for hoover, cpress, igr in ((False, False, False), (True, True, True), (True, False, True)):
        st = _state(rng)
        st.update(hoover=hoover, cpress=cpress, igr=igr)
        if hoover:
            st.update(gh=rng.standard_normal(3), vh=rng.standard_normal(3), vh1=rng.standard_normal(3))
        if cpress:
            st.update(zz=rng.standard_normal((3, 3)), zz1=rng.standard_normal((3, 3)), lzz=1.25, lzz1=-0.5)
        if igr:
            st.update(krdf=np.array([3, 1, 4, 1, 5]))
        p = tmp_path / "w.rst"
        formats.write_restart(str(p), st)
        r = formats.read_restart_full(str(p))
        assert r["complete"]
        _same_state(st, r)
        if hoover:
            assert r["nh"] == 3 and all(np.array_equal(st[k], r[k]) for k in ("gh", "vh", "vh1"))
        if cpress:
            assert np.array_equal(st["zz"], r["zz"]) and np.array_equal(st["zz1"], r["zz1"])
            assert r["lzz"] == 1.25 and r["lzz1"] == -0.5
        if igr:
            assert list(r["krdf"]) == [3, 1, 4, 1, 5]
        # the subset reader sees the same leading records
        q = formats.read_restart(str(p))
        assert np.array_equal(q["x"], st["x"]) and np.array_equal(q["co"], st["co"])


def test_restart_of_a_system_state(tmp_path):
    """restart_state: group centres are the mass-weighted centres, molecule
    centres and their velocities the mass-weighted means (water box)."""
    from orac_hip import synth
    s = synth.config("water216")
    x = np.stack(s.cartesian())
    v = np.random.default_rng(1).standard_normal(x.shape) * 1e-3
    st = formats.restart_state(s, x, v, 10, 300.0)
    o = x[:, 0::3]
    assert st["xpg"].shape == (3, 216) and st["xcm"].shape == (3, 216)
    assert np.allclose(st["xpg"], st["xcm"], atol=1e-12)  # one group per water: centre of mass either way
    m = s.mass
    com = (x[:, 0::3] * m[0::3] + x[:, 1::3] * m[1::3] + x[:, 2::3] * m[2::3]) / (m[0::3] + m[1::3] + m[2::3])
    assert np.allclose(st["xcm"], com, rtol=0, atol=1e-12) and np.abs(st["xcm"] - o).max() < 0.2
    p = tmp_path / "s.rst"
    formats.write_restart(str(p), st)
    r = formats.read_restart_full(str(p))
    _same_state(st, r)


def test_reference_restart_files_round_trip(tmp_path):
    """The reference's own restart files (test1/md.rst, test2's replicas) are
    read to their last record (thermostat and barostat records included) and
    written back value for value."""
    import glob
    import os
    import pytest
    files = sorted(glob.glob("/root/reference/test1/md.rst")) + sorted(
        glob.glob("/root/reference/test2/b-RSTRT/ij4i000[01].rst"))
    if not files or not os.path.exists(files[0]):
        pytest.skip("reference tree not present")
    for f in files:
        a = formats.read_restart_full(f)
        assert a["complete"] and a["hoover"] and a["cpress"] and a["x"].shape == (3, 14900)
        p = tmp_path / "copy.rst"
        formats.write_restart(str(p), a)
        b = formats.read_restart_full(str(p))
        _same_state(a, b)
        for k in ("gh", "vh", "vh1", "zz", "zz1"):
            assert np.array_equal(a[k], b[k]), k
        assert a["lzz"] == b["lzz"] and a["lzz1"] == b["lzz1"]
