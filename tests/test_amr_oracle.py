"""Pins of the oracle's level transfers (AMR): average_down, the flux
average_down_faces and the two-level fill with conservative linear
interpolation (Source/GridWork.cpp:94-164, Source/ScalarSolve.cpp:88-94).

The AMReX routines behind them are not in the reference tree; their
published algorithms are restated in the oracle and pinned here by closed
forms (means of exactly representable data, hand-computed limited slopes,
the 3-D limiter that lands the extreme child exactly on the neighbourhood
maximum) and invariants (conservation, constants and linear fields kept,
children bounded by the coarse neighbourhood, fine data and boundary
extrapolation copied exactly)."""
import numpy as np
import pytest

from oracle import oracle as O
from vidyut3d_amd import abi


def ext(lo, hi, dim, ng=0, face=None):
    return abi.grown_extent(abi.make_box(lo, hi), ng, dim, face)


def geoms(nc, per, r):
    dim = len(nc)
    cg = abi.make_geom(nc, [0.0] * dim, [1.0] * dim, per)
    fg = abi.make_geom([n * q for n, q in zip(nc, r)], [0.0] * dim, [1.0] * dim, per)
    return cg, fg


def full3(v, dim, fill=0):
    return list(v) + [fill] * (3 - len(v))


# ------------------------------------------------------------ average_down
@pytest.mark.parametrize("dim,ratio", [(2, [2, 2, 1]), (3, [2, 2, 2]), (3, [4, 2, 4])])
def test_average_down_is_the_mean_of_the_children(dim, ratio):
    rng = np.random.default_rng(3)
    flo, fhi = full3([4, 8, 0][:dim], dim), full3([11, 15, 7][:dim], dim)
    if dim == 3:
        flo = [r * (x // r) for r, x in zip(ratio, [4, 8, 0])]
        fhi = [flo[d] + ratio[d] * 2 - 1 for d in range(3)]
    fb, fe = ext(flo, fhi, dim)
    cb, ce = ext([0, 0, 0], [7, 7, 7 if dim == 3 else 0], dim)
    fbox = abi.make_box(flo, fhi)
    # dyadic data: every sum and the scaling are exact in any order
    fine = rng.integers(-64, 64, size=abi.shape_of(fb, fe, 3)).astype(np.float64) / 8.0
    crse = np.full(abi.shape_of(cb, ce, 3), 7.0)
    O.average_down(fine, fb, fe, fbox, crse, cb, ce, 1, 2, ratio, dim)
    r = ratio if dim == 3 else ratio[:2] + [1]
    sub = fine[1:3].reshape(2, (fe[2] - fb[2]) // r[2], r[2], (fe[1] - fb[1]) // r[1], r[1],
                             (fe[0] - fb[0]) // r[0], r[0])
    mean = sub.mean(axis=(2, 4, 6))
    clo = [flo[d] // r[d] for d in range(3)]
    sl = tuple(slice(clo[d] - cb[d], clo[d] - cb[d] + mean.shape[3 - d]) for d in (2, 1, 0))
    assert np.array_equal(crse[(slice(1, 3),) + sl], mean)
    # untouched: component 0 and the cells outside the coarsened box
    assert np.all(crse[0] == 7.0)
    mask = np.ones(crse.shape[1:], bool)
    mask[sl] = False
    assert np.all(crse[1:3][:, mask] == 7.0)


@pytest.mark.parametrize("dim", [2, 3])
def test_average_down_faces_is_the_mean_of_the_tiling_faces(dim):
    rng = np.random.default_rng(5)
    ratio = [2, 2, 2] if dim == 3 else [2, 2, 1]
    flo, fhi = full3([2, 4, 2][:dim], dim), full3([9, 7, 5][:dim], dim)
    fbox = abi.make_box(flo, fhi)
    cbox = abi.make_box([0, 0, 0], [5, 5, 5 if dim == 3 else 0])
    fext = [abi.grown_extent(fbox, 0, dim, d) for d in range(3)]
    cext = [abi.grown_extent(cbox, 0, dim, d) for d in range(3)]
    fine = [rng.integers(-32, 32, size=abi.shape_of(*e, 1)).astype(np.float64) / 4.0 for e in fext]
    crse = [np.full(abi.shape_of(*e, 1), -3.0) for e in cext]
    O.average_down_faces(fine, fext, fbox, crse, cext, ratio, dim)
    for d in range(dim):
        f = fine[d][0]
        # fine faces on coarse face positions along d, then tangential means
        idx = [slice(None)] * 3
        ax = 2 - d
        idx[ax] = slice(0, None, ratio[d])
        f = f[tuple(idx)]
        t = [e for e in range(dim) if e != d]
        for e in t:
            a = 2 - e
            shp = list(f.shape)
            shp[a:a + 1] = [shp[a] // ratio[e], ratio[e]]
            f = f.reshape(shp).mean(axis=a + 1)
        lo = [flo[e] // ratio[e] if e < dim else 0 for e in range(3)]
        sl = tuple(slice(lo[e] - cext[d][0][e], lo[e] - cext[d][0][e] + f.shape[2 - e]) for e in (2, 1, 0))
        assert np.array_equal(crse[d][0][sl], f), d


# ---------------------------------------------------- conservative interp
def interp_all(crse_valid, cg, fg, ratio, ng=0):
    """FillCoarsePatch of the whole fine domain (grown by ng)."""
    dim = cg.dim
    cb, ce = abi.grown_extent(abi.domain_box(cg), 0, dim)
    fb, fe = abi.grown_extent(abi.domain_box(fg), ng, dim)
    dst = np.zeros(abi.shape_of(fb, fe, crse_valid.shape[0]))
    O.fill_patch_two_levels(crse_valid, (cb, ce), cg, [], [], [], dst, (fb, fe), fg, ratio)
    return dst


def children(fine, r, dim):
    """(ncomp, nzc, nyc, nxc, r^dim) view of the children of each coarse cell."""
    n, nz, ny, nx = fine.shape
    rz = r[2] if dim == 3 else 1
    v = fine.reshape(n, nz // rz, rz, ny // r[1], r[1], nx // r[0], r[0])
    return v.transpose(0, 1, 3, 5, 2, 4, 6).reshape(n, nz // rz, ny // r[1], nx // r[0], -1)


def test_interp_1d_profiles_closed_form():
    """Data varying along x only: the y slope is 0 and the x slope is the MC
    limited central difference.  Coarse row (0, 1, 4, 4, 4.5, 6): hand values
    of the children of cells 1..4."""
    cg, fg = geoms([6, 3], [0, 1], [2, 2, 1])
    row = np.array([0.0, 1.0, 4.0, 4.0, 4.5, 6.0])
    crse = np.tile(row, (1, 1, 3, 1))
    fine = interp_all(crse, cg, fg, [2, 2, 1])
    got = fine[0, 0, 0]
    # cell 1: cen 2, forw 6, back 2 -> slope 2 -> 1 -/+ 0.5
    assert got[2] == 0.5 and got[3] == 1.5
    # cell 2: forw 0 -> slope 0
    assert got[4] == 4.0 and got[5] == 4.0
    # cell 3: back 0 -> slope 0
    assert got[6] == 4.0 and got[7] == 4.0
    # cell 4: cen 1, forw 3, back 1 -> slope 1 -> 4.5 -/+ 0.25
    assert got[8] == 4.25 and got[9] == 4.75
    # y is periodic and uniform: every fine row equal
    assert np.all(fine[0, 0] == got)


def test_interp_limiter_lands_the_extreme_child_on_the_neighbourhood_max():
    """3-D, ratio 4: a cell whose six face neighbours are +-1 (others 0) has
    unit slopes; its far corner child would reach 3 * 0.375 = 1.125 > 1, so
    alpha = 1 / 1.125 scales it onto the neighbourhood maximum 1."""
    cg, fg = geoms([5, 5, 5], [0, 0, 0], [4, 4, 4])
    c = np.zeros((1, 5, 5, 5))
    for d in range(3):
        p = [2, 2, 2]
        p[d] = 3
        c[(0,) + tuple(p[::-1])] = 1.0
        p[d] = 1
        c[(0,) + tuple(p[::-1])] = -1.0
    fine = interp_all(c, cg, fg, [4, 4, 4])
    ch = fine[0, 8:12, 8:12, 8:12]
    assert abs(ch.max() - 1.0) < 1e-15 and abs(ch.min() + 1.0) < 1e-15
    assert abs(ch[3, 3, 3] - 1.0) < 1e-15
    assert abs(ch.mean()) < 1e-16


@pytest.mark.parametrize("dim,ratio,per", [(2, [2, 2, 1], [0, 1]), (3, [2, 2, 2], [1, 0, 0]),
                                           (3, [4, 2, 2], [0, 0, 1])])
def test_interp_invariants(dim, ratio, per):
    nc = [6, 5, 4][:dim]
    cg, fg = geoms(nc, per, ratio)
    rng = np.random.default_rng(11)
    crse = rng.normal(size=(2,) + tuple(([1] if dim == 2 else []) + nc[::-1]))
    crse = crse.reshape((2,) + abi.shape_of(*abi.grown_extent(abi.domain_box(cg), 0, dim), 1)[1:])
    fine = interp_all(crse, cg, fg, ratio)
    ch = children(fine, ratio, dim)
    # conservation: the children average to the coarse value
    assert np.max(np.abs(ch.mean(axis=-1) - crse)) < 1e-14 * np.max(np.abs(crse))
    # children within the min/max of the (wrapped / clamped) neighbourhood
    pad = []
    for ax in range(3):
        d = 2 - ax
        if d >= dim:
            pad.append((0, 0))
        else:
            pad.append((1, 1))
    mode_wrap = np.pad(crse, [(0, 0)] + pad, mode="wrap")
    mode_edge = np.pad(crse, [(0, 0)] + pad, mode="edge")
    nb = mode_edge.copy()
    for ax in range(3):
        d = 2 - ax
        if d < dim and per[d]:
            sl = [slice(None)] * 4
            sl[ax + 1] = [0, -1]
            nb[tuple(sl)] = mode_wrap[tuple(sl)]
    # every 3^dim neighbourhood
    offs = [(a, b, c) for a in ((-1, 0, 1) if dim == 3 else (0,)) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    zs = 1 if dim == 3 else 0
    views = [nb[:, zs + a:nb.shape[1] - zs + a if nb.shape[1] - zs + a else None,
                1 + b:nb.shape[2] - 1 + b if nb.shape[2] - 1 + b else None,
                1 + c:nb.shape[3] - 1 + c if nb.shape[3] - 1 + c else None] for a, b, c in offs]
    lo, hi = np.min(views, axis=0), np.max(views, axis=0)
    tol = 1e-14 * np.max(np.abs(crse))
    assert np.all(ch.min(axis=-1) >= lo - tol) and np.all(ch.max(axis=-1) <= hi + tol)
    # constants stay constant
    const = interp_all(np.full_like(crse, 3.25), cg, fg, ratio)
    assert np.all(const == 3.25)


def test_interp_reproduces_linear_fields_in_the_interior():
    cg, fg = geoms([8, 8, 8], [0, 0, 0], [2, 2, 2])
    x = np.arange(8) + 0.5
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    crse = (1.0 + 0.5 * X - 0.25 * Y + 0.125 * Z)[None]
    fine = interp_all(crse, cg, fg, [2, 2, 2])
    xf = (np.arange(16) + 0.5) / 2
    Zf, Yf, Xf = np.meshgrid(xf, xf, xf, indexing="ij")
    lin = 1.0 + 0.5 * Xf - 0.25 * Yf + 0.125 * Zf
    inner = (slice(2, 14),) * 3
    assert np.max(np.abs(fine[0][inner] - lin[inner])) < 1e-14


# -------------------------------------------------- two-level ghost fill
def test_two_level_fill_copies_fine_data_and_extrapolates_at_walls():
    """Fine level of two boxes on a 2-D domain periodic in y: dst = box 0
    grown by 3.  Cells that box 0 or box 1 covers (periodic images
    included) carry their fine values exactly; cells beyond the x walls
    repeat the nearest cell inside; the rest are coarse interpolation."""
    ratio = [2, 2, 1]
    cg, fg = geoms([8, 8], [0, 1], ratio)
    rng = np.random.default_rng(2)
    cb, ce = abi.grown_extent(abi.domain_box(cg), 0, 2)
    crse = rng.normal(size=abi.shape_of(cb, ce, 2))
    boxes = [abi.make_box([0, 0, 0], [5, 7, 0]), abi.make_box([6, 12, 0], [11, 15, 0])]
    fexts = [abi.grown_extent(b, 0, 2) for b in boxes]
    fines = [rng.normal(size=abi.shape_of(*e, 2)) + 10.0 for e in fexts]
    db, de = abi.grown_extent(boxes[0], 3, 2)
    dst = np.full(abi.shape_of(db, de, 2), np.nan)
    O.fill_patch_two_levels(crse, (cb, ce), cg, fines, fexts, boxes, dst, (db, de), fg, ratio)
    interp = interp_all(crse, cg, fg, ratio)
    assert not np.isnan(dst).any()
    for j in range(db[1], de[1]):
        for i in range(db[0], de[0]):
            jw = j % 16
            ic = min(max(i, 0), 15)
            v = dst[:, 0, j - db[1], i - db[0]]
            src = None
            for b, (e, f) in enumerate(zip(fexts, fines)):
                bx = boxes[b]
                if bx.lo[0] <= ic <= bx.hi[0] and bx.lo[1] <= jw <= bx.hi[1]:
                    src = f[:, 0, jw - e[0][1], ic - e[0][0]]
            if src is None:
                src = interp[:, 0, jw, ic]
            assert np.array_equal(v, src), (i, j)
