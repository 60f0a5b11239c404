"""Plotfile reader / writer / comparator (Source/Plot.cpp:29-38 layout,
py_utils/compare_fields_from_twoplotfiles.py metric)."""
import numpy as np
import pytest

from vidyut3d_amd import plotfile


@pytest.mark.parametrize("dim", [2, 3])
def test_write_read_round_trip(tmp_path, dim):
    rng = np.random.default_rng(dim)
    shape = (3, 1, 6, 5) if dim == 2 else (3, 4, 6, 5)
    data = rng.normal(size=shape) * 10.0 ** rng.integers(-20, 20, size=shape)
    names = ["E", "ION+", "Potential"]
    lo = [0.0, -1.0, 2.0][:dim]
    hi = [0.01, 1.0, 3.5][:dim]
    p = str(tmp_path / "plt00007")
    plotfile.write(p, names, data, dim, 1.25e-9, 7, lo, hi)
    pf = plotfile.read(p)
    assert pf.names == names and pf.dim == dim and pf.step == 7
    assert pf.time == 1.25e-9
    assert pf.prob_lo == lo and pf.prob_hi == hi
    assert pf.lo == [0] * dim and pf.hi == [shape[3] - 1, shape[2] - 1, shape[1] - 1][:dim]
    np.testing.assert_allclose(pf.dx, [(hi[d] - lo[d]) / [5, 6, 4][d] for d in range(dim)], rtol=0, atol=0)
    assert np.array_equal(pf.data, data)
    assert np.array_equal(pf.field("ION+"), data[1])


def test_reads_big_endian_fab(tmp_path):
    data = np.arange(2 * 1 * 3 * 4, dtype=np.float64).reshape(2, 1, 3, 4) * 0.5
    p = str(tmp_path / "plt")
    plotfile.write(p, ["a", "b"], data, 2, 0.0, 0, [0, 0], [1, 1])
    d = tmp_path / "plt" / "Level_0" / "Cell_D_00000"
    raw = d.read_bytes()
    head, body = raw.split(b"\n", 1)
    head = head.replace(b"(8 7 6 5 4 3 2 1)", b"(1 2 3 4 5 6 7 8)")
    d.write_bytes(head + b"\n" + data.astype(">f8").tobytes())
    assert np.array_equal(plotfile.read(p).data, data)


def test_compare_metric_closed_form(tmp_path):
    """err = sqrt(mean((f1-f2)^2)), rel = err/mean|f1|, rel_l2 = |f1-f2|/|f1|."""
    f1 = np.full((1, 1, 4, 4), 2.0)
    f2 = f1.copy()
    f2[0, 0, 0, :] += 0.5  # 4 of 16 cells off by 0.5
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    plotfile.write(a, ["phi"], f1, 2, 0.0, 0, [0, 0], [1, 1])
    plotfile.write(b, ["phi"], f2, 2, 0.0, 0, [0, 0], [1, 1])
    e = plotfile.compare(a, b)["phi"]
    assert e["err"] == pytest.approx(np.sqrt(4 * 0.25 / 16))
    assert e["rel"] == pytest.approx(0.25 / 2.0)
    assert e["rel_l2"] == pytest.approx(np.sqrt(4 * 0.25) / np.sqrt(16 * 4.0))
    assert plotfile.compare(a, a)["phi"] == {"err": 0.0, "rel": 0.0, "rel_l2": 0.0}


def test_cli(tmp_path, capsys):
    f = np.ones((2, 1, 2, 2))
    a = str(tmp_path / "a")
    plotfile.write(a, ["x", "y"], f, 2, 0.0, 0, [0, 0], [1, 1])
    assert plotfile.main([a, a, "y"]) == 0
    out = capsys.readouterr().out
    assert out.startswith("y") and "rel 0.000000e+00" in out


@pytest.mark.parametrize("dim", [2, 3])
def test_checkpoint_round_trip(tmp_path, dim):
    """Vidyut's checkpoint layout (Source/Plot.cpp:40-119): Header lines
    (title, finest level, istep, dt, t_new, BoxArray) and Level_0/phi."""
    rng = np.random.default_rng(10 + dim)
    shape = (4, 1, 5, 6) if dim == 2 else (4, 3, 5, 6)
    data = rng.normal(size=shape) * 1e15
    p = str(tmp_path / "chk00012")
    plotfile.write_checkpoint(p, data, dim, 12, 1e-12, 1.2e-11)
    lines = (tmp_path / "chk00012" / "Header").read_text().splitlines()
    assert lines[0] == "Checkpoint file for Vidyut" and lines[1] == "0"
    assert lines[2].split() == ["12"] and float(lines[3]) == 1e-12 and float(lines[4]) == 1.2e-11
    assert lines[5] == "(1 0" and lines[-1] == ")"
    c = plotfile.read_checkpoint(p)
    assert (c.step, c.dt, c.time) == (12, 1e-12, 1.2e-11)
    assert c.hi == [shape[3] - 1, shape[2] - 1, shape[1] - 1][:dim]
    assert np.array_equal(c.data, data)
    assert (tmp_path / "chk00012" / "Level_0" / "phi_H").read_text().count("FabOnDisk: phi_D_00000 0") == 1
