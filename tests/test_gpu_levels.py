"""Levels held in several boxes (a BoxArray chopped by amr.max_grid_size,
as every input of the reference sets it): the multi-box FillPatch against
the padded domain (numpy.pad, wrap for periodic directions and edge for
the others, the closed form of FillBoundary + foextrap), the per-box
transport against the single-box call (bit for bit), the gather/scatter of
a level for the solvers, and the checks that the boxes tile the domain."""
import ctypes as C

import numpy as np
import pytest

from vidyut3d_amd import abi, cases, native

pytestmark = pytest.mark.gpu


def padded(a, ng, geom):
    """a (ncomp, nz, ny, nx) over the domain, grown by ng: periodic
    directions wrap, the others repeat the edge cell."""
    out = a
    for d in range(geom.dim):
        ax = 3 - d
        w = [(0, 0)] * 4
        w[ax] = (ng, ng)
        out = np.pad(out, w, mode="wrap" if geom.is_periodic[d] else "edge")
    return out


def part(full, box, ng, geom, full_ng):
    sl = tuple(slice(box.lo[d] - ng + full_ng, box.hi[d] + ng + 1 + full_ng) if d < geom.dim else slice(0, 1)
               for d in (2, 1, 0))
    return full[(slice(None),) + sl]


CASES = [([40, 24], [0, 1], 16, 3), ([40, 24], [1, 0], 8, 1), ([24, 20, 16], [1, 0, 0], 8, 3),
         ([24, 20, 16], [0, 1, 1], 16, 2)]


@pytest.mark.parametrize("n,per,mgs,ng", CASES)
@pytest.mark.parametrize("inplace", [False, True])
def test_fill_patch_boxes_is_the_padded_domain(dev, n, per, mgs, ng, inplace):
    g = abi.make_geom(n, [0.0] * len(n), [1.0] * len(n), per)
    boxes = native.chop(g, mgs)
    lev = native.Level(dev, g, boxes)
    rng = np.random.default_rng(len(boxes))
    shape = (4,) + tuple(reversed(n + [1] * (3 - len(n))))
    a = rng.normal(size=shape)
    want = padded(a, ng, g)
    if inplace:
        host = padded(np.zeros_like(a), ng, g)
        gz = ng if g.dim == 3 else 0
        host[:, gz:gz + shape[1], ng:ng + shape[2], ng:ng + shape[3]] = a  # valid cells only
        arrs = lev.arrays(4, ng, host=host, host_ng=ng)
        native.fill_patch_boxes(lev, arrs, arrs)
        dst = arrs
    else:
        src = lev.arrays(4, 0, host=a)
        dst = lev.arrays(4, ng)
        native.fill_patch_boxes(lev, src, dst)
    for b, box in enumerate(lev.boxes):
        assert np.array_equal(dst[b].get(), part(want, box, ng, g, ng)), b


def test_per_box_transport_equals_the_single_box_call(dev):
    """WENO-Z transport of the electrons and an ion on 8^3 boxes of a 24 x 16
    x 16 air case, each box after the multi-box FillPatch, against the
    single-box fill and call: bit for bit."""
    c = cases.air_3d(n=[24, 16, 16], nsteps=1)
    g, st, m = c.geom, c.state, c.mech
    whole = native.Grid(dev, g)
    mech = native.Mech(dev, m)
    rng = np.random.default_rng(4)
    rxn = rng.normal(size=(m.nspec + 1,) + st.shape[1:]) * 1e10
    bclo, bchi = list(c.params.eden_bc_lo), list(c.params.eden_bc_hi)
    sb1 = whole.array(st.shape[0], 3)
    native.fill_patch(whole, whole.array(st.shape[0], 0, host=st), sb1)
    lev = native.Level(dev, g, native.chop(g, 8))
    src = lev.arrays(st.shape[0], 0, host=st)
    sbs = lev.arrays(st.shape[0], 3)
    native.fill_patch_boxes(lev, src, sbs)
    rx = lev.arrays(m.nspec + 1, 0, host=rxn)
    for spec in (0, 1):
        out1 = whole.array(1)
        native.transport_rhs(mech, c.params, whole, sb1, whole.array(m.nspec + 1, host=rxn), out1, spec, bclo, bchi)
        ref = out1.get()
        outs = lev.arrays(1)
        for b in range(len(lev.boxes)):
            native.transport_rhs(mech, c.params, lev.box_grid(b), sbs[b], rx[b], outs[b], spec, bclo, bchi)
        got = whole.array(1)
        native.level_gather(lev, outs, got)
        assert np.array_equal(got.get(), ref), spec


def test_gather_and_scatter_round_trip(dev):
    g = abi.make_geom([20, 12, 8], [0.0] * 3, [1.0] * 3, [0, 0, 1])
    lev = native.Level(dev, g, native.chop(g, 8))
    a = np.random.default_rng(2).normal(size=(3, 8, 12, 20))
    parts = lev.arrays(3, 1, host=np.pad(a, [(0, 0), (1, 1), (1, 1), (1, 1)]), host_ng=1)
    whole = lev.whole.array(3)
    native.level_gather(lev, parts, whole)
    assert np.array_equal(whole.get(), a)
    back = lev.arrays(3, 0)
    native.level_scatter(lev, whole, back)
    for b, box in enumerate(lev.boxes):
        assert np.array_equal(back[b].get(), part(a, box, 0, g, 0)), b


@pytest.mark.parametrize("boxes,what", [
    ([((0, 0, 0), (9, 7, 0)), ((8, 0, 0), (15, 7, 0))], "overlap"),
    ([((0, 0, 0), (7, 7, 0)), ((9, 0, 0), (15, 7, 0))], "uncovered"),
    ([((0, 0, 0), (7, 7, 0)), ((8, 0, 0), (16, 7, 0))], "leaves the domain"),
])
def test_boxes_must_tile_the_domain(dev, boxes, what):
    g = abi.make_geom([16, 8], [0.0, 0.0], [1.0, 1.0], [0, 0])
    lev = native.Level(dev, g, boxes)
    src = [native.DeviceArray(dev, *abi.grown_extent(b, 0, 2), 1) for b in lev.boxes]
    st = native.lib().vd_fill_patch_boxes(dev.ptr, len(lev.boxes), native.Level.views(src), native.Level.views(src),
                                          lev.cboxes, C.byref(g), 0, 1)
    assert st == abi.VD_EINVAL
    assert what in native.lib().vd_last_error().decode()
