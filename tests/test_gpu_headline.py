"""The headline configurations themselves against the oracle, with the
launch choices the library makes by default (no environment overrides).

* One whole step of the 3-D 256^3 air-like discharge, the configuration the
  throughput figure is quoted on: its potential V-cycles run the fused
  smoother in 32-plane chunks, which smaller grids never select.
* A dielectric Poisson solve at 192 x 192 x 128, the smallest shape class
  that also selects 32-plane chunks, in FP64 and with FP32 V-cycles.
* The predictor-corrector driver (num_timestep_correctors = 2,
  Source/Evolve.cpp:290-302).
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O
from vidyut3d_amd import abi, cases, compare, native

pytestmark = pytest.mark.gpu

NAMES_AIR = cases.AIR_SPECIES + abi.PLASMA_VAR_NAMES


@pytest.fixture
def default_launch(monkeypatch):
    """A fresh context with no launch-choice overrides in the environment."""
    for k in ("VD_FUSED_SMOOTHER", "VD_FUSED_ZC", "VD_STREAM_SWEEPS"):
        monkeypatch.delenv(k, raising=False)
    d = native.Device(0)
    yield d
    d.set_kernel_timing(abi.KT_NONE)
    d.close()


def _assert_fields(ref, got, ns, dim, tol, names):
    errs = compare.state_errors(ref, got, ns, dim)
    print({names[q]: "%.2e" % e for q, e in errs.items()}, flush=True)
    bad = {names[q]: e for q, e in errs.items() if not e <= tol}
    assert not bad, bad
    worst, q = compare.worst_rms_over_meanabs(ref, got, ns, dim)
    print("worst rms/mean|f| %.2e (%s)" % (worst, names[q] if q is not None else "-"), flush=True)
    assert worst <= tol


def test_headline_air_3d_256_step_matches_oracle(default_launch):
    """Config 1 at full size, one step, default tolerances (1e-10 relative
    and absolute, the reference defaults) and default launch shapes."""
    dev = default_launch
    c = cases.air_3d(n=256, nsteps=1)
    print("headline 256^3: oracle step", flush=True)
    ref, tref = O.evolve(c, 1)
    print("headline 256^3: GPU step", flush=True)
    dev.set_kernel_timing(abi.KT_FUSED)
    sim = native.Sim(dev, c)
    sim.evolve(1)
    fused = dev.kernel_timing()[1]
    got = sim.get_state()
    assert sim.time == pytest.approx(tref)
    sim.close()
    assert fused > 0, "the fused smoother did not run at 256^3"
    _assert_fields(ref, got, c.mech.nspec, 3, 1e-10, NAMES_AIR)


def _dielectric_problem(views, precision):
    p = abi.vd_mlmg_problem()
    p.alpha_a, p.beta = 0.0, 1.0
    for d in range(3):
        p.bface[d] = C.pointer(views[d])
    p.bc_lo[:] = [abi.PERBC, abi.PERBC, abi.DIRCBC]
    p.bc_hi[:] = [abi.PERBC, abi.PERBC, abi.DIRCBC]
    p.reltol, p.abstol, p.maxiter = 1e-10, 0.0, 100
    p.precision = precision
    p.pre_smooth = p.post_smooth = 2
    p.max_coarsening_level = 10
    return p


@pytest.fixture(scope="module")
def dielectric_192():
    g, faces, rhs = cases.dielectric_poisson(n=[192, 192, 128])
    og = O.Grid(g)
    hf = [np.ascontiguousarray(faces[d][None]) for d in range(3)]
    hv = [og.view(hf[d], 0, d) for d in range(3)]
    print("dielectric 192x192x128: oracle solve", flush=True)
    ref, info = O.mlmg_solve(og, _dielectric_problem(hv, 0), og.zeros(1), np.ascontiguousarray(rhs[None]))
    return g, hf, rhs, ref, info


@pytest.mark.parametrize("prec", [0, 1])
def test_dielectric_32_plane_chunks_match_oracle(default_launch, dielectric_192, prec):
    dev = default_launch
    g, hf, rhs, ref, info_ref = dielectric_192
    dg = native.Grid(dev, g)
    df = [dg.array(1, 0, d, host=hf[d]) for d in range(3)]
    dphi = dg.array(1)
    dev.set_kernel_timing(abi.KT_FUSED)
    print("dielectric 192x192x128: GPU solve, precision %d" % prec, flush=True)
    info = native.mlmg_solve(dg, _dielectric_problem([f.view() for f in df], prec), dphi,
                             dg.array(1, host=rhs[None]))
    assert dev.kernel_timing()[1] > 0, "the fused smoother did not run"
    err = compare.field_errors(ref, dphi.get())["rel_l2"]
    print("prec %d: %d iterations (oracle %d), rel L2 vs FP64 oracle %.2e"
          % (prec, info.iterations, info_ref.iterations, err), flush=True)
    assert info.resnorm <= info.target
    assert err <= (1e-10 if prec == 0 else 1e-8), err


def test_two_correctors_match_oracle(dev):
    """num_timestep_correctors = 2: predictor, then the corrector with the
    0.5/0.5 averaging of the two stages (Source/Evolve.cpp:290-302)."""
    c = cases.discharge_2d(n=32, nsteps=4)
    c.params.linsolve_reltol, c.params.linsolve_abstol = 1e-12, 0.0
    ref, tref = O.evolve(c, 4, ncorr=2)
    one, _ = O.evolve(c, 4, ncorr=1)
    sim = native.Sim(dev, c)
    sim.evolve(4, ncorr=2)
    got = sim.get_state()
    assert sim.time == pytest.approx(tref)
    sim.close()
    names = [c.mech.names[s].value.decode() for s in range(3)] + abi.PLASMA_VAR_NAMES
    _assert_fields(ref, got, 3, 2, 1e-10, names)
    # the corrector changes the answer well above the tolerance
    assert max(compare.state_errors(ref, one, 3, 2).values()) > 1e-6
