"""GPU level transfers (AMR) against the oracle, bit for bit: average_down,
the flux average_down_faces and the two-level fill (FillPatchTwoLevels /
FillCoarsePatch with conservative linear interpolation)."""
import numpy as np
import pytest

from oracle import oracle as O
from vidyut3d_amd import abi, native

pytestmark = pytest.mark.gpu


def geoms(nc, per, r):
    dim = len(nc)
    cg = abi.make_geom(nc, [0.0] * dim, [1.0] * dim, per)
    fg = abi.make_geom([n * q for n, q in zip(nc, r)], [0.0] * dim, [1.0] * dim, per)
    return cg, fg


@pytest.mark.parametrize("dim,ratio", [(2, [2, 2, 1]), (3, [2, 2, 2]), (3, [4, 2, 4])])
def test_average_down_bit_exact(dev, dim, ratio):
    rng = np.random.default_rng(7)
    r = ratio if dim == 3 else ratio[:2] + [1]
    flo = [r[0] * 2, r[1] * 1, r[2] * 1 if dim == 3 else 0]
    fhi = [flo[0] + 3 * r[0] - 1, flo[1] + 4 * r[1] - 1, (flo[2] + 2 * r[2] - 1) if dim == 3 else 0]
    fbox = abi.make_box(flo, fhi)
    fb, fe = abi.grown_extent(fbox, 1, dim)
    cbox = abi.make_box([0, 0, 0], [9, 9, 9 if dim == 3 else 0])
    cb, ce = abi.grown_extent(cbox, 0, dim)
    fine = rng.normal(size=abi.shape_of(fb, fe, 4))
    crse = rng.normal(size=abi.shape_of(cb, ce, 4))
    ref = O.average_down(fine, fb, fe, fbox, crse.copy(), cb, ce, 1, 3, ratio, dim)
    df = native.DeviceArray(dev, fb, fe, 4, fine)
    dc = native.DeviceArray(dev, cb, ce, 4, crse)
    native.average_down(dev, df, fbox, dc, 1, 3, ratio, dim)
    assert np.array_equal(dc.get(), ref)


@pytest.mark.parametrize("dim,ratio", [(2, [2, 2, 1]), (3, [2, 2, 2]), (3, [2, 4, 2])])
def test_average_down_faces_bit_exact(dev, dim, ratio):
    rng = np.random.default_rng(9)
    r = ratio if dim == 3 else ratio[:2] + [1]
    flo = [r[0] * 1, r[1] * 2, r[2] if dim == 3 else 0]
    fhi = [flo[0] + 4 * r[0] - 1, flo[1] + 2 * r[1] - 1, (flo[2] + 3 * r[2] - 1) if dim == 3 else 0]
    fbox = abi.make_box(flo, fhi)
    cbox = abi.make_box([0, 0, 0], [7, 7, 7 if dim == 3 else 0])
    fext = [abi.grown_extent(fbox, 0, dim, d) for d in range(3)]
    cext = [abi.grown_extent(cbox, 0, dim, d) for d in range(3)]
    fine = [rng.normal(size=abi.shape_of(*e, 1)) for e in fext]
    crse = [rng.normal(size=abi.shape_of(*e, 1)) for e in cext]
    ref = O.average_down_faces(fine, fext, fbox, [c.copy() for c in crse], cext, ratio, dim)
    df = [native.DeviceArray(dev, *e, 1, f) for e, f in zip(fext, fine)]
    dc = [native.DeviceArray(dev, *e, 1, c) for e, c in zip(cext, crse)]
    native.average_down_faces(dev, df, fbox, dc, ratio, dim)
    for d in range(dim):
        assert np.array_equal(dc[d].get(), ref[d]), d


CASES = [
    # dim, coarse cells, periodic, ratio, fine boxes, dst box index, ghosts
    (2, [8, 8], [0, 1], [2, 2, 1], [([0, 0], [5, 7]), ([6, 12], [11, 15]), ([12, 2], [15, 9])], 0, 3),
    (2, [6, 4], [1, 1], [4, 4, 1], [([4, 0], [11, 7]), ([12, 8], [19, 15])], 1, 2),
    (3, [6, 6, 4], [1, 0, 0], [2, 2, 2], [([0, 0, 0], [3, 5, 3]), ([4, 2, 2], [9, 7, 7]), ([10, 6, 0], [11, 11, 3])],
     1, 3),
    (3, [4, 4, 4], [0, 0, 1], [2, 4, 2], [([2, 4, 0], [5, 11, 5]), ([0, 12, 6], [7, 15, 7])], 0, 3),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_fill_patch_two_levels_bit_exact(dev, case):
    dim, nc, per, ratio, bl, di, ng = CASES[case]
    cg, fg = geoms(nc, per, ratio)
    rng = np.random.default_rng(20 + case)
    cb, ce = abi.grown_extent(abi.domain_box(cg), 0, dim)
    crse = rng.normal(size=abi.shape_of(cb, ce, 3))
    boxes = [abi.make_box(lo + [0] * (3 - dim), hi + [0] * (3 - dim)) for lo, hi in bl]
    fexts = [abi.grown_extent(b, 0, dim) for b in boxes]
    fines = [rng.normal(size=abi.shape_of(*e, 3)) + 5.0 for e in fexts]
    db, de = abi.grown_extent(boxes[di], ng, dim)
    dst0 = np.full(abi.shape_of(db, de, 3), -99.0)
    ref = O.fill_patch_two_levels(crse, (cb, ce), cg, fines, fexts, boxes, dst0.copy(), (db, de), fg, ratio,
                                  scomp=1, ncomp=2)
    dc = native.DeviceArray(dev, cb, ce, 3, crse)
    dfs = [native.DeviceArray(dev, *e, 3, f) for e, f in zip(fexts, fines)]
    dd = native.DeviceArray(dev, db, de, 3, dst0)
    native.fill_patch_two_levels(dev, dc, cg, dfs, boxes, dd, fg, ratio, scomp=1, ncomp=2)
    got = dd.get()
    assert np.array_equal(got, ref)
    assert np.all(got[0] == -99.0)
    # FillCoarsePatch: no fine boxes, the whole grown box from the coarse level
    ref2 = O.fill_patch_two_levels(crse, (cb, ce), cg, [], [], [], dst0.copy(), (db, de), fg, ratio)
    native.fill_patch_two_levels(dev, dc, cg, [], [], dd, fg, ratio)
    assert np.array_equal(dd.get(), ref2)


def test_fill_patch_two_levels_rejects_mismatched_domains(dev):
    cg, fg = geoms([4, 4], [0, 0], [2, 2, 1])
    bad = abi.make_geom([10, 8], [0.0, 0.0], [1.0, 1.0], [0, 0])
    cb, ce = abi.grown_extent(abi.domain_box(cg), 0, 2)
    dc = native.DeviceArray(dev, cb, ce, 1)
    db, de = abi.grown_extent(abi.domain_box(fg), 1, 2)
    dd = native.DeviceArray(dev, db, de, 1)
    with pytest.raises(native.VidyutError):
        native.fill_patch_two_levels(dev, dc, cg, [], [], dd, bad, [2, 2, 1])
    with pytest.raises(native.VidyutError):
        native.fill_patch_two_levels(dev, dc, cg, [], [], dd, fg, [3, 2, 1])


def test_level_transfers_reject_bad_boxes_and_components(dev):
    """Fine boxes not aligned to the refinement ratio, components beyond an
    array and arrays that do not cover the coarsened box are errors, not
    launches."""
    cbox = abi.make_box([0, 0, 0], [7, 7, 7])
    cb, ce = abi.grown_extent(cbox, 0, 3)
    dc = native.DeviceArray(dev, cb, ce, 2)
    fbox = abi.make_box([2, 2, 2], [9, 9, 9])
    fb, fe = abi.grown_extent(fbox, 0, 3)
    df = native.DeviceArray(dev, fb, fe, 2)
    native.average_down(dev, df, fbox, dc, 0, 2, [2, 2, 2], 3)  # aligned: fine
    with pytest.raises(native.VidyutError):  # lo not a multiple of 4
        native.average_down(dev, df, fbox, dc, 0, 2, [4, 4, 4], 3)
    with pytest.raises(native.VidyutError):  # components beyond the arrays
        native.average_down(dev, df, fbox, dc, 1, 2, [2, 2, 2], 3)
    odd = abi.make_box([3, 2, 2], [9, 9, 9])
    with pytest.raises(native.VidyutError):
        native.average_down(dev, df, odd, dc, 0, 1, [2, 2, 2], 3)
    small = native.DeviceArray(dev, [0, 0, 0], [3, 3, 3], 2)  # does not hold coarse cells 1..4
    with pytest.raises(native.VidyutError):
        native.average_down(dev, df, fbox, small, 0, 1, [2, 2, 2], 3)
    fext = [abi.grown_extent(fbox, 0, 3, d) for d in range(3)]
    ffs = [native.DeviceArray(dev, *e, 1) for e in fext]
    cfs = [native.DeviceArray(dev, *abi.grown_extent(cbox, 0, 3, d), 1) for d in range(3)]
    native.average_down_faces(dev, ffs, fbox, cfs, [2, 2, 2], 3)
    with pytest.raises(native.VidyutError):
        native.average_down_faces(dev, ffs, fbox, cfs, [3, 2, 2], 3)


def test_sim_regrid_cycle_make_remake_average_clear(dev):
    """The level-1 cycle of a simulation: regrid builds level 1 on the
    grids make_new_grids gives for the level-0 tags, filled by
    FillCoarsePatch (bit for bit the oracle's interpolation); AverageDown
    then returns level 0 to itself (conservative interpolation, to
    rounding); a second regrid on other tags remakes level 1 from its old
    data and the coarse level (oracle FillPatchTwoLevels with the old
    boxes); with no tags the level is cleared."""
    from vidyut3d_amd import cases
    c = cases.discharge_2d(n=32, seed=1, nsteps=1)
    sim = native.Sim(dev, c)
    E = c.mech.e_idx
    st = sim.get_state()
    ratio = [2, 2, 1]
    prm = native.regrid_params(ref_ratio=2, blocking_factor=4, max_grid_size=16, n_error_buf=1)
    thr = 0.3 * st[E].max()
    boxes = sim.regrid([E], [thr], [1e300], prm)
    assert boxes
    tags = sim.error_est([E], [thr], [1e300])
    expect = native.make_new_grids(tags, c.geom, prm)
    assert [(tuple(b.lo), tuple(b.hi)) for b in boxes] == [(tuple(b.lo), tuple(b.hi)) for b in expect]
    cg = c.geom
    fg = abi.make_geom([2 * (cg.dom_hi[d] + 1) for d in range(2)], [cg.prob_lo[0], cg.prob_lo[1]],
                       [cg.prob_hi[0], cg.prob_hi[1]], [cg.is_periodic[0], cg.is_periodic[1]])
    cext = abi.grown_extent(abi.domain_box(cg), 0, 2)
    for b, box in enumerate(boxes):
        ext = abi.grown_extent(box, 0, 2)
        ref = O.fill_patch_two_levels(st, cext, cg, [], [], [], np.zeros(abi.shape_of(*ext, st.shape[0])), ext, fg,
                                      ratio)
        assert np.array_equal(sim.get_fine(b, box), ref), b
    # AverageDown of interpolated data gives level 0 back
    sim.average_down()
    st2 = sim.get_state()
    scale = np.max(np.abs(st), axis=(1, 2, 3), keepdims=True) + 1e-300
    assert np.max(np.abs(st2 - st) / scale) < 1e-14
    # new fine data, other tags: RemakeLevel
    rng = np.random.default_rng(4)
    fines = []
    for b, box in enumerate(boxes):
        f = sim.get_fine(b, box) * (1.0 + 0.1 * rng.uniform(size=sim.fine_shape(box)))
        sim.set_fine(b, f)
        fines.append(f)
    thr2 = 0.1 * st2[E].max()
    boxes2 = sim.regrid([E], [thr2], [1e300], prm)
    assert len(boxes2) >= 1 and boxes2 != boxes
    st3 = sim.get_state()
    fexts = [abi.grown_extent(b, 0, 2) for b in boxes]
    for b, box in enumerate(boxes2):
        ext = abi.grown_extent(box, 0, 2)
        ref = O.fill_patch_two_levels(st3, cext, cg, fines, fexts, boxes, np.zeros(abi.shape_of(*ext, st.shape[0])),
                                      ext, fg, ratio)
        assert np.array_equal(sim.get_fine(b, box), ref), b
    # nothing tagged: level 1 cleared
    assert sim.regrid([E], [1e300], [1e300], prm) == []
    sim.close()


def test_two_level_plotfile_and_checkpoint_round_trip(dev, tmp_path):
    """With a level 1, WritePlotFile and WriteCheckpointFile write both
    levels (Level_1/Cell, Level_1/phi: every box of the level in one data
    file); the readers give back each box's data bit for bit, and a new
    simulation restarted from the checkpoint holds the same two levels."""
    from vidyut3d_amd import cases, plotfile
    c = cases.discharge_2d(n=32, seed=2, nsteps=1)
    sim = native.Sim(dev, c)
    E = c.mech.e_idx
    st = sim.get_state()
    prm = native.regrid_params(ref_ratio=2, blocking_factor=4, max_grid_size=8, n_error_buf=1)
    boxes = sim.regrid([E], [0.2 * st[E].max()], [1e300], prm)
    assert len(boxes) >= 2  # several boxes in one level file
    fines = [sim.get_fine(b, box) for b, box in enumerate(boxes)]
    plt = str(tmp_path / "plt")
    sim.write_plotfile(plt)
    pf = plotfile.read(plt)
    assert pf.finest_level == 1 and pf.ref_ratio == [2]
    assert np.array_equal(pf.data, st)
    got = {(tuple(lo), tuple(hi)): d for lo, hi, d in pf.levels[1]}
    for box, f in zip(boxes, fines):
        assert np.array_equal(got[(tuple(box.lo[:2]), tuple(box.hi[:2]))], f)
    chk = str(tmp_path / "chk")
    sim.write_checkpoint(chk)
    ck = plotfile.read_checkpoint(chk)
    assert len(ck.levels) == 2 and len(ck.levels[1]) == len(boxes)
    sim2 = native.Sim(dev, c)
    sim2.read_checkpoint(chk)
    assert np.array_equal(sim2.get_state(), st)
    b2 = sim2.level_boxes(1)
    assert [(tuple(b.lo), tuple(b.hi)) for b in b2] == [(tuple(b.lo), tuple(b.hi)) for b in boxes]
    for b, box in enumerate(b2):
        assert np.array_equal(sim2.get_fine(b, box), fines[b])
    sim2.close()
    sim.close()
