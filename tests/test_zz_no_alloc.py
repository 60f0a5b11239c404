"""vd_transport_rhs_noalloc against vd_transport_rhs through the C ABI:
the same dsdt bit for bit (the same kernels, face fluxes in the fixed
scratch array instead of the context's buffers), on two calls in a row
(the second on a smaller box, so stale contents of the shared array from
the first call would show), in 2-D and 3-D WENO and in 3-D upwind (every
direction's face array stored); boxes whose fluxes exceed the array are
rejected before any kernel runs."""
import ctypes as C

import numpy as np
import pytest

from vidyut3d_amd import abi, cases, native

pytestmark = pytest.mark.gpu


def _call(fn, dev, m, c, g, sb, rxn, out, spec, bclo, bchi):
    return fn(dev.ptr, m.ptr, C.byref(sb.view()), C.byref(rxn.view()), C.byref(out.view()), C.byref(g.box),
              C.byref(g.geom), abi.ivec(bclo), abi.ivec(bchi), C.byref(c.params), C.c_double(0.0),
              C.c_double(1e-12), int(spec))


def _rhs_pair(dev, c, spec, seed):
    """dsdt of species `spec` from vd_transport_rhs and from the no-allocation
    variant, on a seeded state with the case's boundary conditions."""
    L = native.lib()
    g = native.Grid(dev, c.geom)
    m = native.Mech(dev, c.mech)
    rng = np.random.default_rng(seed)
    st = c.state.copy()
    cc = abi.comp(c.mech.nspec)
    for d in range(c.geom.dim):
        st[cc["EFX"] + d] = rng.normal(size=st.shape[1:]) * 1e5
    src = g.array(st.shape[0], 0, host=st)
    sb = g.array(st.shape[0], 3)
    native.fill_patch(g, src, sb)
    rxn = g.array(c.mech.nspec + 1, host=rng.normal(size=(c.mech.nspec + 1,) + st.shape[1:]) * 1e10)
    bclo, bchi = list(c.params.ion_bc_lo), list(c.params.ion_bc_hi)
    ref = g.array(1)
    native.check(_call(L.vd_transport_rhs, dev, m, c, g, sb, rxn, ref, spec, bclo, bchi), "vd_transport_rhs")
    got = g.array(1, host=np.full((1,) + st.shape[1:], np.nan))
    native.check(_call(L.vd_transport_rhs_noalloc, dev, m, c, g, sb, rxn, got, spec, bclo, bchi),
                 "vd_transport_rhs_noalloc")
    return ref.get(), got.get()


@pytest.mark.parametrize("dim,hyp", [(3, 5), (2, 5), (3, 1)])
def test_noalloc_transport_matches_allocating_entry_point_twice(dim, hyp):
    dev = native.Device(0)
    try:
        sizes = [40, 24] if dim == 3 else [96, 48]
        for q, n in enumerate(sizes):
            c = cases.air_3d(n=n, seed=1) if dim == 3 else cases.discharge_2d(n=n, seed=q)
            c.params.hyp_order = hyp
            print("no-alloc transport: dim %d hyp %d n %d (call %d)" % (dim, hyp, n, q + 1), flush=True)
            for spec in (0, 1):
                ref, got = _rhs_pair(dev, c, spec, 10 * q + spec)
                assert np.isfinite(ref).all()
                assert np.array_equal(ref, got), (n, spec, float(np.max(np.abs(ref - got))))
    finally:
        dev.close()


def test_noalloc_transport_rejects_boxes_beyond_its_scratch():
    dev = native.Device(0)
    try:
        # 3-D WENO, 176^3: x + y faces = 2 * 177 * 176^2 * 8 B > 64 MiB
        n = 176
        g3 = abi.make_geom([n] * 3, [0.0] * 3, [0.1] * 3, [0, 0, 0])
        c = cases.air_3d(n=8)
        c.params.hyp_order = 5
        m = native.Mech(dev, c.mech)
        box = abi.domain_box(g3)
        # views describe the box but are never dereferenced: the size check
        # rejects the call before any launch
        sb = abi.array4_view(0, *abi.grown_extent(box, 3, 3), c.state.shape[0])
        rxn = abi.array4_view(0, *abi.grown_extent(box, 0, 3), c.mech.nspec + 1)
        out = abi.array4_view(0, *abi.grown_extent(box, 0, 3), 1)
        L = native.lib()
        st = L.vd_transport_rhs_noalloc(dev.ptr, m.ptr, C.byref(sb), C.byref(rxn), C.byref(out), C.byref(box),
                                        C.byref(g3), abi.ivec([2, 2, 3]), abi.ivec([2, 2, 3]), C.byref(c.params),
                                        C.c_double(0.0), C.c_double(1e-12), 0)
        msg = L.vd_last_error().decode()
        assert st == abi.VD_EINVAL and "no-allocation scratch" in msg, (st, msg)
    finally:
        dev.close()


def test_noalloc_transport_with_per_cell_drift_velocities():
    """Species whose drift velocity is computed per cell (a prescribed
    velocity; a fitted mobility) also take the velocity array from the fixed
    scratch: bit-equal to vd_transport_rhs, two calls in a row."""
    import test_gpu_plugins as tp
    dev = native.Device(0)
    try:
        for q, c in enumerate([cases.advect_1d(64), cases.advect_1d(32)]):
            print("no-alloc transport, prescribed velocity: n %d" % c.geom.dom_hi[0], flush=True)
            ref, got = _rhs_pair(dev, c, 1, 40 + q)
            assert np.array_equal(ref, got)
        for q, n in enumerate([24, 16]):
            c = tp.fitted_air(n=n)
            print("no-alloc transport, fitted mobilities: n %d" % n, flush=True)
            for spec in (0, 1, 10):
                ref, got = _rhs_pair(dev, c, spec, 50 + q)
                assert np.isfinite(ref).all()
                assert np.array_equal(ref, got), (n, spec)
    finally:
        dev.close()
