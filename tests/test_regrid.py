"""Regrid (AmrMesh::MakeNewGrids as Vidyut's regrid(0, time) runs it,
Source/Evolve.cpp:66-72) on the host: known answers for small tag sets and
the invariants every grid generation must keep (all buffered tags covered,
boxes disjoint, inside the domain, aligned to the blocking factor, no side
longer than max_grid_size, efficiency of the clusters).  AMReX's own
clustering is not in the reference tree, so against it the boxes are
parity unpinned; these tests pin the algorithm as restated."""
import numpy as np
import pytest

from vidyut3d_amd import abi, native


def boxes_as_tuples(bl):
    return sorted((tuple(b.lo), tuple(b.hi)) for b in bl)


def buffered_coarse_tags(tags, nbuf, per, dim):
    """the tags dilated by nbuf (Chebyshev), periodic wrap / drop"""
    out = np.zeros_like(tags)
    idx = np.argwhere(tags)
    n = tags.shape
    for z, y, x in idx:
        for dz in (range(-nbuf, nbuf + 1) if dim == 3 else (0,)):
            for dy in range(-nbuf, nbuf + 1):
                for dx in range(-nbuf, nbuf + 1):
                    c = [z + dz, y + dy, x + dx]
                    ok = True
                    for ax in range(3):
                        d = 2 - ax
                        if 0 <= c[ax] < n[ax]:
                            continue
                        if d < dim and per[d]:
                            c[ax] %= n[ax]
                        else:
                            ok = False
                    if ok:
                        out[tuple(c)] = 1
    return out


def check_invariants(tags, geom, prm, bl):
    dim = geom.dim
    r = list(prm.ref_ratio)
    bf = list(prm.blocking_factor)
    mgs = list(prm.max_grid_size)
    nf = [(geom.dom_hi[d] + 1) * r[d] if d < dim else 1 for d in range(3)]
    cover = np.zeros((nf[2], nf[1], nf[0]), np.int32)
    for b in bl:
        for d in range(dim):
            assert 0 <= b.lo[d] <= b.hi[d] < nf[d]
            assert b.lo[d] % bf[d] == 0 and (b.hi[d] + 1) % bf[d] == 0
            assert b.hi[d] - b.lo[d] + 1 <= mgs[d]
        sl = tuple(slice(b.lo[d], b.hi[d] + 1) if d < dim else slice(0, 1) for d in (2, 1, 0))
        cover[sl] += 1
    assert cover.max() <= 1, "boxes overlap"
    buf = buffered_coarse_tags(tags, prm.n_error_buf, [geom.is_periodic[d] for d in range(3)], dim)
    for z, y, x in np.argwhere(buf):
        c = [x * r[0], y * r[1], z * r[2] if dim == 3 else 0]
        assert cover[c[2], c[1], c[0]] == 1, ("uncovered tag", x, y, z)


def test_two_separate_blobs_give_their_bounding_boxes():
    g = abi.make_geom([32, 32], [0, 0], [1, 1], [0, 0])
    tags = np.zeros((1, 32, 32), np.int8)
    tags[0, 4:8, 4:8] = 1
    tags[0, 20:26, 18:30] = 1
    prm = native.regrid_params(ref_ratio=2, blocking_factor=2, max_grid_size=64, n_error_buf=0, grid_eff=0.7)
    bl = native.make_new_grids(tags, g, prm)
    assert boxes_as_tuples(bl) == [((8, 8, 0), (15, 15, 0)), ((36, 40, 0), (59, 51, 0))]
    check_invariants(tags, g, prm, bl)


def test_buffer_and_blocking_factor_grow_and_align_the_box():
    """One tagged cell (5, 9): buffered by 1 -> cells 4..6 x 8..10; with
    blocking factor 8 and ratio 2 the tags are clustered on a 4-cell
    coarsening, giving coarse cells 4..7 x 8..11 -> fine 8..15 x 16..23."""
    g = abi.make_geom([32, 32], [0, 0], [1, 1], [0, 0])
    tags = np.zeros((1, 32, 32), np.int8)
    tags[0, 9, 5] = 1
    prm = native.regrid_params(ref_ratio=2, blocking_factor=8, max_grid_size=64, n_error_buf=1)
    bl = native.make_new_grids(tags, g, prm)
    assert boxes_as_tuples(bl) == [((8, 16, 0), (15, 23, 0))]


def test_periodic_buffer_wraps_and_walls_drop():
    g = abi.make_geom([16, 16], [0, 0], [1, 1], [1, 0])
    tags = np.zeros((1, 16, 16), np.int8)
    tags[0, 0, 0] = 1  # corner: x wraps (periodic), y drops below 0
    prm = native.regrid_params(ref_ratio=2, blocking_factor=2, max_grid_size=64, n_error_buf=1, grid_eff=0.5)
    bl = native.make_new_grids(tags, g, prm)
    assert boxes_as_tuples(bl) == [((0, 0, 0), (3, 3, 0)), ((30, 0, 0), (31, 3, 0))]
    check_invariants(tags, g, prm, bl)


def test_l_shape_is_cut_at_the_inflection():
    """An L of tags fills half its bounding box (efficiency 0.5 < 0.7): the
    signature along y has one strong inflection at the corner of the L, so
    the cluster is cut there into two full boxes."""
    g = abi.make_geom([32, 32], [0, 0], [1, 1], [0, 0])
    tags = np.zeros((1, 32, 32), np.int8)
    tags[0, 8:24, 8:12] = 1   # vertical bar x 8..11, y 8..23
    tags[0, 20:24, 12:24] = 1  # foot x 12..23, y 20..23
    prm = native.regrid_params(ref_ratio=2, blocking_factor=2, max_grid_size=64, n_error_buf=0)
    bl = native.make_new_grids(tags, g, prm)
    check_invariants(tags, g, prm, bl)
    assert len(bl) == 2
    vol = sum((b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1) for b in bl)
    assert vol == 4 * int(tags.sum())  # no untagged cell refined


def test_max_grid_size_chops_into_aligned_pieces():
    g = abi.make_geom([32, 32, 32], [0, 0, 0], [1, 1, 1], [0, 0, 0])
    tags = np.ones((32, 32, 32), np.int8)
    prm = native.regrid_params(ref_ratio=2, blocking_factor=8, max_grid_size=24, n_error_buf=0)
    bl = native.make_new_grids(tags, g, prm)
    check_invariants(tags, g, prm, bl)
    # 64 cells per side, pieces of at most 24 in units of 8: 24 + 24 + 16
    xs = sorted({(b.lo[0], b.hi[0]) for b in bl})
    assert xs == [(0, 23), (24, 47), (48, 63)]
    assert len(bl) == 27


@pytest.mark.parametrize("dim,seed", [(2, 1), (2, 2), (3, 3), (3, 4)])
def test_random_blobs_keep_the_invariants(dim, seed):
    rng = np.random.default_rng(seed)
    n = [48, 40, 32][:dim]
    per = [1, 0, 1][:dim]
    g = abi.make_geom(n, [0] * dim, [1] * dim, per)
    shape = (n[2] if dim == 3 else 1, n[1], n[0])
    tags = np.zeros(shape, np.int8)
    zz, yy, xx = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    for _ in range(4):
        c = [rng.integers(0, s) for s in shape]
        rad = rng.uniform(2, 7)
        d2 = (xx - c[2]) ** 2 + (yy - c[1]) ** 2 + ((zz - c[0]) ** 2 if dim == 3 else 0)
        tags[d2 <= rad * rad] = 1
    prm = native.regrid_params(ref_ratio=2, blocking_factor=4, max_grid_size=16, n_error_buf=1, grid_eff=0.7)
    bl = native.make_new_grids(tags, g, prm)
    check_invariants(tags, g, prm, bl)


def test_no_tags_no_boxes_and_bad_parameters_are_rejected():
    g = abi.make_geom([16, 16], [0, 0], [1, 1], [0, 0])
    tags = np.zeros((1, 16, 16), np.int8)
    assert native.make_new_grids(tags, g, native.regrid_params()) == []
    with pytest.raises(native.VidyutError):
        native.make_new_grids(tags, g, native.regrid_params(ref_ratio=3))
    with pytest.raises(native.VidyutError):
        native.make_new_grids(tags, g, native.regrid_params(blocking_factor=6, ref_ratio=4))
    with pytest.raises(native.VidyutError):
        native.make_new_grids(tags, g, native.regrid_params(blocking_factor=4, max_grid_size=10))


def test_ratio_four_in_3d_keeps_the_invariants():
    """Ratio 4 with blocking factor 8: tags clustered on a 2-cell coarsening,
    boxes on the new level aligned to 8 and chopped to 16."""
    g = abi.make_geom([16, 16, 16], [0, 0, 0], [1, 1, 1], [0, 1, 0])
    tags = np.zeros((16, 16, 16), np.int8)
    tags[3:9, 0:2, 5:12] = 1   # touches y = 0 (periodic in y)
    tags[12:14, 10:15, 1:3] = 1
    prm = native.regrid_params(ref_ratio=4, blocking_factor=8, max_grid_size=16, n_error_buf=1, grid_eff=0.7)
    bl = native.make_new_grids(tags, g, prm)
    check_invariants(tags, g, prm, bl)
    # the buffered y-range of the first blob wraps to the top of the domain
    assert any(b.hi[1] == 63 for b in bl)
