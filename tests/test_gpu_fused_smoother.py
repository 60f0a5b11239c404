"""The fused red-black smoother gives the values of the per-colour sweeps.

The fused kernel applies two colours of one or two sweeps (and, in the
smoother-only solves, the residual norm of the result) in one pass over
the level, on tiles of 64 x 16 cells and chunks of planes with halos.
Solves through it and through the per-colour launches agree to round-off
(the arithmetic is the same expression per cell), and both agree with the
oracle.  Grids here span several tiles in x and y and several plane
chunks in z, with Dirichlet, Neumann and periodic sides, so halo, wrap and
chunk seams are all exercised.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O
from vidyut3d_amd import abi, cases, compare, native

pytestmark = pytest.mark.gpu


def _with_fused(flag, fn, launches_out=None):
    """Runs fn(dev) in a fresh context whose multigrid workspaces are
    created with the fused smoother on (flag True; 16-plane chunks, which
    also enables it on grids smaller than it is chosen for by default) or
    off, and checks from the kernel-timing counts that it ran or not
    (appended to launches_out when given)."""
    keys = ("VD_FUSED_SMOOTHER", "VD_FUSED_ZC")
    old = {k: os.environ.get(k) for k in keys}
    os.environ["VD_FUSED_SMOOTHER"] = "1" if flag else "0"
    os.environ["VD_FUSED_ZC"] = "16"
    dev = native.Device(0)
    try:
        dev.set_kernel_timing(abi.KT_FUSED)
        out = fn(dev)
        launches = dev.kernel_timing()[1]
        assert (launches > 0) == flag, launches
        if launches_out is not None:
            launches_out.append(launches)
        return out
    finally:
        dev.set_kernel_timing(abi.KT_NONE)
        dev.close()
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _problem(views, bcl, bch, prec):
    p = abi.vd_mlmg_problem()
    p.alpha_a, p.beta = 0.0, 1.0
    for d in range(3):
        p.bface[d] = C.pointer(views[d])
    p.bc_lo[:] = bcl
    p.bc_hi[:] = bch
    p.reltol, p.abstol, p.maxiter = 1e-11, 0.0, 200
    p.precision = prec
    p.pre_smooth = p.post_smooth = 2
    p.max_coarsening_level = 10
    return p


# shapes that coarsen to a few hundred cells (the bottom solve is CG in one
# workgroup) and span 2-3 tiles in x and y and 2-3 plane chunks in z
CASES = [
    # shape (nx, ny, nz), periodic x/y, bc_lo, bc_hi, precision
    ((128, 48, 64), (0, 0), [abi.HNEUBC, abi.DIRCBC, abi.DIRCBC], [abi.DIRCBC, abi.HNEUBC, abi.DIRCBC], 0),
    ((192, 32, 96), (1, 1), [0, 0, abi.DIRCBC], [0, 0, abi.DIRCBC], 0),
    ((128, 48, 64), (1, 0), [0, abi.HNEUBC, abi.DIRCBC], [0, abi.DIRCBC, abi.HNEUBC], 1),
]


@pytest.mark.parametrize("shape,per,bcl,bch,prec", CASES)
def test_fused_smoother_matches_colour_sweeps_and_oracle(shape, per, bcl, bch, prec):
    rng = np.random.default_rng(sum(shape) + prec)
    g = abi.make_geom(list(shape), [0.0] * 3, [1.0, 0.7, 1.3], [per[0], per[1], 0])
    og = O.Grid(g)
    faces = []
    for d in range(3):
        fs = list(reversed(shape))
        fs[2 - d] += 1
        faces.append(np.exp(0.3 * rng.normal(size=[1] + fs)))
    rhs = rng.normal(size=(1,) + tuple(reversed(shape)))

    def run(dev):
        dg = native.Grid(dev, g)
        df = [dg.array(1, 0, d, host=faces[d]) for d in range(3)]
        dv = [f.view() for f in df]
        phi = dg.array(1)
        info = native.mlmg_solve(dg, _problem(dv, bcl, bch, prec), phi, dg.array(1, host=rhs))
        return phi.get(), info.iterations

    fused, it_f = _with_fused(True, run)
    plain, it_p = _with_fused(False, run)
    diff = compare.field_errors(plain, fused)["rel_l2"]
    print("fused vs colour sweeps: rel L2 %.3e, bit-exact %s, iterations %d / %d"
          % (diff, np.array_equal(plain, fused), it_f, it_p))
    assert it_f == it_p
    assert diff <= 1e-13
    hv = [og.view(faces[d], 0, d) for d in range(3)]
    ref, _ = O.mlmg_solve(og, _problem(hv, bcl, bch, 0), og.zeros(1), rhs)
    err = compare.field_errors(ref, fused)["rel_l2"]
    assert err <= (1e-10 if prec == 0 else 1e-8), err


def test_fused_smoother_only_step_multi_tile_matches_oracle():
    """Whole steps at 80^3 (2 x 5 tiles, 3 plane chunks): the smoother-only
    implicit solves (sweep + residual norm fused) and the potential
    V-cycles through the fused kernel, against the oracle and against the
    per-colour launches."""
    n, steps = 80, 2
    c = cases.air_3d(n=n, nsteps=steps)
    ref, _ = O.evolve(c, steps)

    def run(dev):
        sim = native.Sim(dev, cases.air_3d(n=n, nsteps=steps))
        sim.evolve(steps)
        st = sim.get_state()
        sim.close()
        return st

    fused = _with_fused(True, run)
    plain = _with_fused(False, run)
    errs = compare.state_errors(ref, fused, c.mech.nspec, c.geom.dim)
    bad = {q: e for q, e in errs.items() if not e <= 1e-10}
    assert not bad, bad
    errs = compare.state_errors(plain, fused, c.mech.nspec, c.geom.dim)
    assert max(errs.values()) <= 1e-12, errs


# implicit species solves on the cell coefficient (the harmonic face values
# formed per plane inside the fused kernel): periodic sides, Dirichlet and
# Neumann walls, tiles and plane chunks cut by the level's edges
CC_CASES = [
    # shape (nx, ny, nz), periodic x/y, bc_lo, bc_hi
    ((80, 48, 40), (1, 0), [abi.PERBC, abi.DIRCBC, abi.DIRCBC], [abi.PERBC, abi.HNEUBC, abi.DIRCBC]),
    ((96, 32, 36), (1, 1), [abi.PERBC, abi.PERBC, abi.HNEUBC], [abi.PERBC, abi.PERBC, abi.DIRCBC]),
    ((72, 40, 34), (0, 0), [abi.DIRCBC, abi.HNEUBC, abi.HNEUBC], [abi.HNEUBC, abi.DIRCBC, abi.IHNEUBC]),
]


@pytest.mark.parametrize("shape,per,bcl,bch", CC_CASES)
@pytest.mark.parametrize("which", ["electrons", "ion", "energy"])
def test_fused_cell_coefficient_solve_matches_colour_sweeps_and_oracle(shape, per, bcl, bch, which):
    """The smoother-only implicit solve (Jacobi bound <= 1e-2 at this dt)
    through one fused sweep-plus-residual launch per iteration, against the
    per-colour sweeps with the streamed residual (same iteration count,
    round-off agreement) and against the oracle."""
    c = cases.air_3d(n=list(shape), nsteps=1)
    g = abi.make_geom(list(shape), [0.0] * 3, [0.1] * 3, [per[0], per[1], 0])
    og = O.Grid(g)
    cc = abi.comp(c.mech.nspec)
    spec = {"electrons": 0, "ion": 1, "energy": cc["EEN"]}[which]
    rng = np.random.default_rng(sum(shape) + spec)
    st = c.state.copy()
    dims = tuple(reversed(shape))
    st[cc["ETEMP"]] = rng.uniform(5e3, 5e4, dims)
    st[cc["EEN"]] = 1.5 * st[0] * abi.K_B * st[cc["ETEMP"]]
    for q in ("EFX", "EFY", "EFZ"):
        st[cc[q]] = rng.normal(size=dims) * 1e5
    sb = O.fill_patch(og, st, 3)
    old = st * (1 + 0.01 * rng.normal(size=st.shape))
    ds = rng.normal(size=(1,) + dims) * (1e24 if spec != cc["EEN"] else 1e5)
    prm = c.params
    prm.linsolve_reltol, prm.linsolve_abstol, prm.linsolve_maxiter = 1e-13, 0.0, 200
    dt = 1e-9

    def run(dev):
        dg = native.Grid(dev, g)
        mech = native.Mech(dev, c.mech)
        dpn = dg.array(st.shape[0], host=st)
        info = native.implicit_solve(mech, prm, dg, spec, dg.array(st.shape[0], 3, host=sb),
                                     dg.array(st.shape[0], host=old), dg.array(1, host=ds), dpn, dg.faces(),
                                     bcl, bch, dt)
        return dpn.get()[spec], info.iterations

    print("%s %s per=%s: fused" % (which, shape, per), flush=True)
    nl = []
    (fused, it_f) = _with_fused(True, run, nl)
    plain, it_p = _with_fused(False, run)
    diff = float(np.max(np.abs(plain - fused)) / np.max(np.abs(plain)))
    print("%s %s: fused vs colour sweeps max rel %.3e (bit-exact %s), iterations %d / %d"
          % (which, shape, diff, np.array_equal(plain, fused), it_f, it_p), flush=True)
    assert it_f == it_p and it_f >= 2
    # sweeps only: one fused launch per iteration (a V-cycle solve would
    # launch it for its pre- and post-smoothing too)
    assert nl[0] == it_f, nl
    assert diff <= 1e-13
    ref, _, info = O.implicit_solve(c.mech, prm, og, spec, sb, 3, old, ds, st, bcl, bch, dt)
    assert info.iterations == it_f
    err = float(np.max(np.abs(ref[spec] - fused)) / np.max(np.abs(ref[spec])))
    assert err <= 1e-10, err
