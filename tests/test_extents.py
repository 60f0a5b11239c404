"""32-bit cell offsets: every entry point rejects a box, or an array with
its ghost cells, of 2^31 cells or more with VD_EINVAL before it uses the
context or launches anything (CPU only: the context and mechanism given
here are placeholders that are never dereferenced)."""
import ctypes as C

import pytest

from vidyut3d_amd import abi, cases, native

CTX = C.c_void_p(0x10)   # never used: the extent check returns first
MECH = C.c_void_p(0x20)
PTR = 0x1000


def view(begin, end, ncomp=1):
    return abi.array4_view(PTR, begin, end, ncomp)


BIG = ([0, 0, 0], [2048, 2048, 2048])       # 2^33 cells
SMALL = ([0, 0, 0], [8, 8, 8])


def box(lo, hi_excl):
    return abi.make_box(lo, [h - 1 for h in hi_excl])


def geom3():
    return abi.make_geom([2048] * 3, [0.0] * 3, [1.0] * 3, [0, 0, 0])


def err():
    return native.lib().vd_last_error().decode()


@pytest.mark.parametrize("big_box", [True, False])
def test_transport_rhs_rejects_oversized_extents(big_box):
    b = box(*BIG) if big_box else box(*SMALL)
    sb = view([-3] * 3, [2051] * 3, 20) if not big_box else view([-3] * 3, [11] * 3, 20)
    small = view(*SMALL, 11)
    g, p = geom3(), cases.default_params()
    st = native.lib().vd_transport_rhs(CTX, MECH, C.byref(sb), C.byref(small), C.byref(small), C.byref(b),
                                       C.byref(g), abi.ivec([0] * 3), abi.ivec([0] * 3), C.byref(p),
                                       C.c_double(0.0), C.c_double(1e-12), 0)
    assert st == abi.VD_EINVAL
    assert "2^31" in err() and ("box" if big_box else "array") in err()


def test_solve_potential_rejects_an_oversized_box():
    b, g, p = box(*BIG), geom3(), cases.default_params()
    a = view([-1] * 3, [2049] * 3, 20)
    ef = (abi.vd_array4 * 3)(*[view(*SMALL) for _ in range(3)])
    st = native.lib().vd_solve_potential(CTX, MECH, C.byref(a), C.byref(a), ef, C.byref(b), C.byref(g), C.byref(p),
                                         C.c_double(0.0), None)
    assert st == abi.VD_EINVAL and "solve_potential" in err()


def test_fill_patch_and_multifab_ops_reject_oversized_arrays():
    b = box(*SMALL)
    big = view(*BIG, 4)
    small = view(*SMALL, 4)
    g = geom3()
    assert native.lib().vd_fill_patch(CTX, C.byref(big), C.byref(small), C.byref(b), C.byref(g), 0, 1) == abi.VD_EINVAL
    assert "fill_patch" in err()
    assert native.lib().vd_setval(CTX, C.byref(big), C.byref(b), 0, 1, C.c_double(0.0)) == abi.VD_EINVAL
    assert native.lib().vd_copy(CTX, C.byref(small), C.byref(big), C.byref(b), 0, 0, 1) == abi.VD_EINVAL


def test_mlmg_solve_rejects_an_oversized_face_array():
    b, g = box(*SMALL), abi.make_geom([8] * 3, [0.0] * 3, [1.0] * 3, [0, 0, 0])
    prob = abi.vd_mlmg_problem()
    face = view([0, 0, 0], [2049, 2048, 2048])
    prob.bface[0] = C.pointer(face)
    a = view(*SMALL)
    st = native.lib().vd_mlmg_solve(CTX, C.byref(prob), C.byref(a), C.byref(a), C.byref(b), C.byref(g), None)
    assert st == abi.VD_EINVAL and "mlmg_solve" in err()
