"""The z-streamed multigrid kernels (colour sweeps, residual and norms,
residual-restriction, prolongation) give the values of the row-per-block
kernels bit for bit: the same terms in the same order per cell, only the
loads are organised differently.  Each case runs in a fresh context whose
workspaces are created with the streamed kernels on and then off
(VD_STREAM_SWEEPS), on grids with walls, a periodic direction and a Robin
side (non-uniform boundary c0), in FP64 and with FP32 V-cycles."""
import ctypes as C
import os

import numpy as np
import pytest

from vidyut3d_amd import abi, cases, native

pytestmark = pytest.mark.gpu


def _with_stream(flag, fn):
    old = os.environ.get("VD_STREAM_SWEEPS")
    os.environ["VD_STREAM_SWEEPS"] = "1" if flag else "0"
    dev = native.Device(0)
    try:
        return fn(dev)
    finally:
        dev.close()
        if old is None:
            os.environ.pop("VD_STREAM_SWEEPS", None)
        else:
            os.environ["VD_STREAM_SWEEPS"] = old


CASES = [
    # shape, periodic (x, y, z), bc_lo, bc_hi, precision
    ((64, 48, 40), (1, 0, 0), [0, abi.DIRCBC, abi.HNEUBC], [0, abi.HNEUBC, abi.DIRCBC], 0),
    ((48, 64, 32), (0, 0, 0), [abi.DIRCBC, abi.HNEUBC, abi.DIRCBC], [abi.HNEUBC, abi.DIRCBC, abi.ROBINBC], 0),
    ((64, 32, 48), (0, 1, 0), [abi.HNEUBC, 0, abi.DIRCBC], [abi.DIRCBC, 0, abi.DIRCBC], 1),
]


@pytest.mark.parametrize("shape,per,bcl,bch,prec", CASES)
def test_streamed_mg_kernels_match_row_kernels(shape, per, bcl, bch, prec):
    rng = np.random.default_rng(sum(shape) + 7 * prec)
    g = abi.make_geom(list(shape), [0.0] * 3, [1.0, 0.8, 1.2], list(per))
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
        p = abi.vd_mlmg_problem()
        p.alpha_a, p.beta = 0.0, 1.0
        for d in range(3):
            p.bface[d] = C.pointer(dv[d])
        p.bc_lo[:] = bcl
        p.bc_hi[:] = bch
        p.reltol, p.abstol, p.maxiter = 1e-11, 0.0, 200
        p.precision = prec
        p.pre_smooth = p.post_smooth = 2
        p.max_coarsening_level = 10
        phi = dg.array(1)
        info = native.mlmg_solve(dg, p, phi, dg.array(1, host=rhs))
        return phi.get(), info.iterations

    streamed, it_s = _with_stream(True, run)
    rows, it_r = _with_stream(False, run)
    assert it_s == it_r
    assert np.array_equal(streamed, rows)


def test_streamed_kernels_whole_steps_match_row_kernels():
    """Whole 3-D steps: potential V-cycles, the constant-stencil and the
    cell-coefficient implicit solves."""
    n, steps = 48, 3

    def run(dev):
        sim = native.Sim(dev, cases.air_3d(n=n, nsteps=steps))
        sim.evolve(steps)
        st = sim.get_state()
        sim.close()
        return st

    assert np.array_equal(_with_stream(True, run), _with_stream(False, run))

