"""AMReX plotfile reader, writer and field comparator.

The on-disk format is the one Vidyut writes through
amrex::WriteMultiLevelPlotfile (Source/Plot.cpp:29-38):

* ``Header``: ``HyperCLaw-V1.1``, the variable count and names, the
  dimension, time, finest level, prob_lo / prob_hi, the (empty) refinement
  ratios, the domain box, the level step, the cell sizes, the coordinate
  system, the boundary width, then per level ``level nboxes time``, the step,
  the physical extent of each box and the path ``Level_0/Cell``;
* ``Level_0/Cell_H``: the VisMF header (box array, the FAB file and offset
  of each box, per-component minima and maxima);
* ``Level_0/Cell_D_00000``: per box a text line
  ``FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))<box> <ncomp>``
  followed by the box data, IEEE doubles in the byte order the descriptor
  names, component after component, x fastest.

Checkpoints (``read_checkpoint`` / ``write_checkpoint``) follow
Vidyut::WriteCheckpointFile (Source/Plot.cpp:40-119): a ``Header`` with the
title ``Checkpoint file for Vidyut``, the finest level, istep, dt and t_new
per level and the BoxArray, and the state as the VisMF MultiFab
``Level_0/phi`` (``phi_H`` + ``phi_D_00000``, the same FAB layout).

Both readers also take a second level (finest level 1): ``levels`` holds,
per level, the list of (lo, hi, data) of its boxes (the ``Level_1/Cell``
or ``Level_1/phi`` MultiFab, several FABs in one data file at the offsets
the VisMF header gives); ``data`` stays the level-0 state.

``compare`` follows py_utils/compare_fields_from_twoplotfiles.py:39-46
(err = sqrt(mean((f1-f2)^2)), rel = err / mean(|f1|)) and also reports
the relative L2 norm ||f1-f2||_2 / ||f1||_2.
"""
import os
import re
import sys

import numpy as np

_BOX = re.compile(r"\(\(([-\d,\s]+)\)\s*\(([-\d,\s]+)\)\s*\(([-\d,\s]+)\)\)")


def _ints(s):
    return [int(x) for x in s.replace(",", " ").split()]


def _box(text):
    m = _BOX.search(text)
    if not m:
        raise ValueError("no box in %r" % text)
    return _ints(m.group(1)), _ints(m.group(2))


class Plotfile:
    """Contents of a plotfile: names, geometry, the level-0 data of shape
    (ncomp, nz, ny, nx) (nz = 1 in 2-D) and, per level, its boxes' data."""

    def __init__(self, names, data, dim, time, step, prob_lo, prob_hi, dx, lo, hi, levels=None,
                 ref_ratio=None):
        self.names, self.data, self.dim = list(names), data, dim
        self.time, self.step = time, step
        self.prob_lo, self.prob_hi, self.dx = list(prob_lo), list(prob_hi), list(dx)
        self.lo, self.hi = list(lo), list(hi)
        self.levels = levels if levels is not None else [[(list(lo), list(hi), data)]]
        self.ref_ratio = list(ref_ratio or [])

    @property
    def finest_level(self):
        return len(self.levels) - 1

    def field(self, name):
        return self.data[self.names.index(name)]


def read(path):
    with open(os.path.join(path, "Header")) as f:
        lines = [ln.rstrip("\n") for ln in f]
    if not lines or lines[0].strip() != "HyperCLaw-V1.1":
        raise ValueError("%s: not a HyperCLaw-V1.1 plotfile" % path)
    nvar = int(lines[1])
    names = [lines[2 + q].strip() for q in range(nvar)]
    p = 2 + nvar
    dim = int(lines[p])
    time = float(lines[p + 1])
    finest = int(lines[p + 2])
    prob_lo = [float(x) for x in lines[p + 3].split()]
    prob_hi = [float(x) for x in lines[p + 4].split()]
    ratios = _ints(lines[p + 5])  # refinement ratios (empty for one level)
    lo, hi = _box(lines[p + 6])  # the level-0 domain comes first
    step = int(lines[p + 7].split()[0])
    dx = [float(x) for x in lines[p + 8].split()]
    cells = [ln.strip() for ln in lines[p + 9 + finest:] if ln.strip().endswith("/Cell")]
    if len(cells) != finest + 1:
        raise ValueError("%s: %d level paths for %d levels" % (path, len(cells), finest + 1))
    levels = [_read_fabs(os.path.join(path, c), dim) for c in cells]
    if len(levels[0]) != 1:
        raise ValueError("%s: %d boxes on level 0; this reader expects one" % (path, len(levels[0])))
    return Plotfile(names, levels[0][0][2], dim, time, step, prob_lo, prob_hi, dx, lo, hi, levels, ratios)


def _read_vismf(prefix, dim):
    """The single FAB of the VisMF MultiFab <prefix>_H / its data file, as
    an array (ncomp, nz, ny, nx)."""
    fabs = _read_fabs(prefix, dim)
    if len(fabs) != 1:
        raise ValueError("%s: %d boxes on level 0; this reader expects one" % (prefix, len(fabs)))
    return fabs[0][2]


def _read_fabs(prefix, dim):
    """Every FAB of the VisMF MultiFab <prefix>, in header order, as
    (lo, hi, array (ncomp, nz, ny, nx))."""
    with open(prefix + "_H") as f:
        vh = f.read()
    out = []
    for fname, off in re.findall(r"FabOnDisk:\s+(\S+)\s+(\d+)", vh):
        out.append(_read_fab(os.path.join(os.path.dirname(prefix), fname), int(off), dim))
    return out


def _read_fab(fpath, off, dim):
    fname = os.path.basename(fpath)
    with open(fpath, "rb") as f:
        f.seek(off)
        hdr = b""
        while not hdr.endswith(b"\n"):
            c = f.read(1)
            if not c:
                raise ValueError("%s: truncated FAB header" % fname)
            hdr += c
        hdr = hdr.decode()
        m = re.match(r"FAB \(\(8, \(64 11 52 0 1 12 0 1023\)\),\(8, \(([\d ]+)\)\)\)(.*)\s(\d+)\s*$", hdr)
        if not m:
            raise ValueError("%s: unsupported FAB header %r" % (fname, hdr))
        order = _ints(m.group(1))
        if order == [8, 7, 6, 5, 4, 3, 2, 1]:
            dt = np.dtype("<f8")
        elif order == [1, 2, 3, 4, 5, 6, 7, 8]:
            dt = np.dtype(">f8")
        else:
            raise ValueError("%s: byte order %s" % (fname, order))
        blo, bhi = _box(m.group(2))
        ncomp = int(m.group(3))
        ext = [bhi[d] - blo[d] + 1 for d in range(dim)] + [1] * (3 - dim)
        n = ext[0] * ext[1] * ext[2] * ncomp
        raw = np.frombuffer(f.read(n * 8), dtype=dt, count=n)
    return blo, bhi, raw.astype(np.float64).reshape(ncomp, ext[2], ext[1], ext[0])


def _write_vismf(prefix, data, box_text):
    """data (ncomp, nz, ny, nx) as the single-FAB VisMF MultiFab <prefix>."""
    ncomp = data.shape[0]
    base = os.path.basename(prefix)
    hdr = "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))%s %d\n" % (box_text, ncomp)
    with open(prefix + "_D_00000", "wb") as f:
        f.write(hdr.encode())
        f.write(np.ascontiguousarray(data, dtype="<f8").tobytes())
    with open(prefix + "_H", "w") as f:
        f.write("1\n1\n%d\n0\n(1 0\n%s\n)\n1\nFabOnDisk: %s_D_00000 0\n\n" % (ncomp, box_text, base))
        flat = data.reshape(ncomp, -1)
        f.write("1,%d\n%s\n\n1,%d\n%s\n" % (ncomp, "".join("%.16e," % v for v in flat.min(1)), ncomp,
                                          "".join("%.16e," % v for v in flat.max(1))))


class Checkpoint:
    """Contents of a Vidyut checkpoint: step, dt, time, the level-0 state
    (ncomp, nz, ny, nx) and, per level, its boxes' data."""

    def __init__(self, step, dt, time, data, lo, hi, levels=None):
        self.step, self.dt, self.time, self.data = step, dt, time, data
        self.lo, self.hi = list(lo), list(hi)
        self.levels = levels if levels is not None else [[(list(lo), list(hi), data)]]


def read_checkpoint(path):
    """Reads the checkpoint Vidyut::WriteCheckpointFile writes
    (Source/Plot.cpp:40-119): the Header "Checkpoint file for Vidyut",
    finest level, istep / dt / t_new per level, the BoxArray, and the VisMF
    MultiFab Level_0/phi."""
    with open(os.path.join(path, "Header")) as f:
        lines = [ln.rstrip("\n") for ln in f]
    if not lines or lines[0].strip() != "Checkpoint file for Vidyut":
        raise ValueError("%s: not a Vidyut checkpoint" % path)
    finest = int(lines[1])
    step = int(lines[2].split()[0])
    dt = float(lines[3].split()[0])
    time = float(lines[4].split()[0])
    lo, hi = _box("\n".join(lines[5:]))
    levels = [_read_fabs(os.path.join(path, "Level_%d" % l, "phi"), len(lo)) for l in range(finest + 1)]
    if len(levels[0]) != 1:
        raise ValueError("%s: %d boxes on level 0; this reader expects one" % (path, len(levels[0])))
    return Checkpoint(step, dt, time, levels[0][0][2], lo, hi, levels)


def write_checkpoint(path, data, dim, step, dt, time, lo=None):
    """Writes data (ncomp, nz, ny, nx) as a single-level checkpoint in the
    layout above."""
    data = np.ascontiguousarray(data, dtype="<f8")
    ncomp, nz, ny, nx = data.shape
    ext = [nx, ny, nz][:dim]
    lo = [0] * dim if lo is None else list(lo)
    hi = [lo[d] + ext[d] - 1 for d in range(dim)]
    j = lambda v: ",".join(str(x) for x in v)
    box = "((%s) (%s) (%s))" % (j(lo), j(hi), j([0] * dim))
    os.makedirs(os.path.join(path, "Level_0"), exist_ok=True)
    with open(os.path.join(path, "Header"), "w") as f:
        f.write("Checkpoint file for Vidyut\n0\n%d \n%r \n%r \n(1 0\n%s\n)\n" % (step, float(dt), float(time), box))
    _write_vismf(os.path.join(path, "Level_0", "phi"), data, box)


def write(path, names, data, dim, time, step, prob_lo, prob_hi, lo=None):
    """Writes data (ncomp, nz, ny, nx) as a single-level, single-box
    plotfile in the layout above (little-endian doubles)."""
    data = np.ascontiguousarray(data, dtype="<f8")
    ncomp, nz, ny, nx = data.shape
    ext = [nx, ny, nz][:dim]
    lo = [0] * dim if lo is None else list(lo)
    hi = [lo[d] + ext[d] - 1 for d in range(dim)]
    dx = [(prob_hi[d] - prob_lo[d]) / ext[d] for d in range(dim)]

    def box(a, b):
        j = lambda v: ",".join(str(x) for x in v)
        return "((%s) (%s) (%s))" % (j(a), j(b), j([0] * dim))

    os.makedirs(os.path.join(path, "Level_0"), exist_ok=True)
    r = lambda v: " ".join(repr(float(x)) for x in v)
    with open(os.path.join(path, "Header"), "w") as f:
        f.write("HyperCLaw-V1.1\n%d\n" % ncomp)
        f.writelines(n + "\n" for n in names)
        f.write("%d\n%r\n0\n%s \n%s \n\n%s \n%d \n%s \n0\n0\n" % (
            dim, float(time), r(prob_lo[:dim]), r(prob_hi[:dim]), box(lo, hi), step, r(dx)))
        f.write("0 1 %r\n%d\n" % (float(time), step))
        for d in range(dim):
            f.write("%r %r\n" % (prob_lo[d] + dx[d] * lo[d], prob_lo[d] + dx[d] * (hi[d] + 1)))
        f.write("Level_0/Cell\n")
    _write_vismf(os.path.join(path, "Level_0", "Cell"), data, box(lo, hi))


def compare(a, b, fields=None):
    """Per-field errors of plotfile b against plotfile a (paths or
    Plotfile objects): {'name': {'err', 'rel', 'rel_l2'}}."""
    a = read(a) if isinstance(a, str) else a
    b = read(b) if isinstance(b, str) else b
    if a.data.shape[1:] != b.data.shape[1:]:
        raise ValueError("plotfiles cover different boxes: %s vs %s" % (a.data.shape, b.data.shape))
    out = {}
    for name in (fields or a.names):
        f1, f2 = a.field(name), b.field(name)
        d = f1 - f2
        err = float(np.sqrt(np.mean(d * d)))
        mabs = float(np.mean(np.abs(f1)))
        n1 = float(np.sqrt(np.sum(f1 * f1)))
        out[name] = {"err": err, "rel": err / mabs if mabs > 0 else (0.0 if err == 0 else float("inf")),
                     "rel_l2": float(np.sqrt(np.sum(d * d))) / n1 if n1 > 0 else
                     (0.0 if err == 0 else float("inf"))}
    return out


def main(argv=None):
    """python -m vidyut3d_amd.plotfile PLT1 PLT2 [field ...]"""
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) < 2:
        print("usage: python -m vidyut3d_amd.plotfile PLT1 PLT2 [field ...]", file=sys.stderr)
        return 2
    res = compare(argv[0], argv[1], argv[2:] or None)
    for name, e in res.items():
        print("%-24s err %.6e  rel %.6e  rel_l2 %.6e" % (name, e["err"], e["rel"], e["rel_l2"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
