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


def _with_fused(flag, fn):
    """Runs fn(dev) in a fresh context whose multigrid workspaces are
    created with the fused smoother on (flag True; 16-plane chunks, which
    also enables it on grids smaller than it is chosen for by default) or
    off, and checks from the kernel-timing counts that it ran or not."""
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
