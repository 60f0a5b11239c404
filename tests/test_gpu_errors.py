"""Error paths of the C ABI: where the reference aborts the process
(Source/compute_explicit_flux.H:100,315, Source/Vidyut.cpp:203-207, MLMG
non-convergence), the library returns a status and a message, and the
context stays usable."""
import ctypes as C

import numpy as np
import pytest

from vidyut3d_amd import abi, cases, native

pytestmark = pytest.mark.gpu


def _status(fn, *args):
    st = fn(*args)
    return st, native.lib().vd_last_error().decode(errors="replace")


def _transport_setup(dev, hyp, scheme):
    c = cases.discharge_2d(n=16)
    c.params.hyp_order, c.params.weno_scheme = hyp, scheme
    g = native.Grid(dev, c.geom)
    m = native.Mech(dev, c.mech)
    sb = g.array(c.state.shape[0], 3)
    rxn = g.array(c.mech.nspec + 1)
    out = g.array(1)
    return c, g, m, sb, rxn, out


@pytest.mark.parametrize("hyp,scheme,word", [(3, 3, "hyp_order"), (4, 1, "hyp_order"),
                                             (5, 0, "WENO scheme"), (5, 4, "WENO scheme")])
def test_unknown_hyp_order_or_weno_scheme_is_einval(dev, hyp, scheme, word):
    c, g, m, sb, rxn, out = _transport_setup(dev, hyp, scheme)
    L = native.lib()
    st, msg = _status(L.vd_transport_rhs, dev.ptr, m.ptr, C.byref(sb.view()), C.byref(rxn.view()),
                      C.byref(out.view()), C.byref(g.box), C.byref(g.geom), abi.ivec([3, 0, 0]),
                      abi.ivec([3, 0, 0]), C.byref(c.params), C.c_double(0.0), C.c_double(1e-11), 0)
    assert st == abi.VD_EINVAL and word in msg, (st, msg)
    fl = g.faces()
    st, msg = _status(L.vd_compute_scalar_transport_flux, dev.ptr, m.ptr, C.byref(sb.view()),
                      native.Grid.face_views(fl), C.byref(g.box), C.byref(g.geom), abi.ivec([3, 0, 0]),
                      abi.ivec([3, 0, 0]), C.byref(c.params), C.c_double(0.0), C.c_double(1e-11), 0)
    assert st == abi.VD_EINVAL and word in msg, (st, msg)
    # the context still works
    c.params.hyp_order, c.params.weno_scheme = 5, 3
    native.transport_rhs(m, c.params, g, sb, rxn, out, 0, [3, 0, 0], [3, 0, 0])
    dev.sync()


def test_unknown_tag_variable_is_enotfound(dev):
    c = cases.discharge_2d(n=16)
    g = native.Grid(dev, c.geom)
    st1 = g.array(c.state.shape[0], 1)
    tags = native.DeviceArray(dev, *g.extent(0), 1, dtype=np.int8)
    L = native.lib()
    tv = tags.tag_view()
    for bad in (c.state.shape[0], -1):
        st, msg = _status(L.vd_error_est, dev.ptr, C.byref(st1.view()), C.byref(tv), C.byref(g.box), 1,
                          (C.c_int * 1)(bad), (C.c_double * 1)(1.0), (C.c_double * 1)(1.0), C.c_char(1))
        assert st == abi.VD_ENOTFOUND and "not in the state" in msg, (st, msg)
    sim = native.Sim(dev, c)
    with pytest.raises(native.VidyutError) as e:
        sim.error_est([c.state.shape[0] + 2], [1.0], [1.0])
    assert e.value.status == abi.VD_ENOTFOUND
    sim.close()


def test_mech_find_id(dev):
    """plasmachem::find_id (Source/PlasmaChem.cpp:21-30): index of a named
    species, -1 when absent."""
    c = cases.air_3d(n=8)
    m = native.Mech(dev, c.mech)
    for s, name in enumerate(cases.AIR_SPECIES):
        assert m.find_id(name) == s
    assert m.find_id("AR") == -1
    assert m.find_id("") == -1
    assert m.find_id("e") == -1  # names are case sensitive, as std::string compare


def test_solver_at_maxiter_is_enoconv_with_its_residual(dev):
    """A solve that reaches maxiter above its target returns VD_ENOCONV with
    the last residual; the reference's MLMG aborts there."""
    g, faces, rhs = cases.dielectric_poisson(n=16)
    dg = native.Grid(dev, g)
    df = [dg.array(1, 0, d, host=np.ascontiguousarray(faces[d][None])) for d in range(3)]
    p = abi.vd_mlmg_problem()
    p.alpha_a, p.beta = 0.0, 1.0
    views = [f.view() for f in df]
    for d in range(3):
        p.bface[d] = C.pointer(views[d])
    p.bc_lo[:] = [0, 0, abi.DIRCBC]
    p.bc_hi[:] = [0, 0, abi.DIRCBC]
    p.reltol, p.abstol, p.maxiter = 1e-14, 0.0, 1
    p.pre_smooth = p.post_smooth = 2
    p.max_coarsening_level = 10
    info = abi.vd_solve_info()
    phi = dg.array(1)
    st, msg = _status(native.lib().vd_mlmg_solve, dev.ptr, C.byref(p), C.byref(phi.view()),
                      C.byref(dg.array(1, host=rhs[None]).view()), C.byref(dg.box), C.byref(dg.geom),
                      C.byref(info))
    assert st == abi.VD_ENOCONV, (st, msg)
    assert info.iterations == 1 and info.resnorm > info.target > 0
    assert info.resnorm < info.resnorm0
    # the same problem with room to converge succeeds on the same context
    p.maxiter = 100
    phi2 = dg.array(1)
    info2 = native.mlmg_solve(dg, p, phi2, dg.array(1, host=rhs[None]))
    assert info2.resnorm <= info2.target
