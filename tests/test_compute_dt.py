"""Vidyut::ComputeDt (Source/TimeStep.cpp:14-31): the adaptive time step of
the library (vd_compute_dt, host code) and of the oracle's step driver."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import oracle as O
from vidyut3d_amd import abi, cases, native


def rule(p, t, old, ts):
    """TimeStep.cpp:17-27, restated."""
    if not (p.adaptive_dt and t > p.adaptive_dt_delay):
        return old
    trans = sys.float_info.max
    if p.do_transport:
        trans = min(ts[0] * p.advective_cfl, ts[1] * p.diffusive_cfl)
    adp = min(trans, ts[2] * p.dielectric_cfl)
    adp = min(max(adp, p.dt_min), p.dt_max)
    return min(old * p.dt_stretch, adp)


def lib_dt(p, t, old, ts):
    out = C.c_double()
    st = native.lib().vd_compute_dt(C.byref(p), C.c_double(t), C.c_double(old), (C.c_double * 3)(*ts),
                                    C.byref(out))
    assert st == abi.VD_OK
    return out.value


def test_defaults_are_the_reference_keys():
    p = abi.vd_params()
    native.lib().vd_params_default(C.byref(p))
    assert p.adaptive_dt == 0 and p.advective_cfl == 0.7 and p.diffusive_cfl == 1.0
    assert p.dielectric_cfl == 1.0 and p.dt_min == 1e-14 and p.dt_max == 1e-9
    assert p.adaptive_dt_delay == 0.0 and p.dt_stretch == sys.float_info.max


@pytest.mark.parametrize("ts", [(3e-11, 8e-11, 5e-11), (9e-12, 1e-13, 2e-10), (1e-8, 1e-8, 4e-12),
                                (1e-16, 1e-16, 1e-16), (1.0, 1.0, 1.0)])
@pytest.mark.parametrize("transport", [1, 0])
def test_library_compute_dt_follows_timestep_cpp(ts, transport):
    p = cases.default_params()
    p.adaptive_dt = 1
    p.do_transport = transport
    for stretch in (sys.float_info.max, 1.05):
        p.dt_stretch = stretch
        for old in (1e-12, 1e-10):
            assert lib_dt(p, 1e-9, old, ts) == rule(p, 1e-9, old, ts)
    # the first step (t = 0, not after the delay) and a fixed step keep old
    assert lib_dt(p, 0.0, 2e-12, ts) == 2e-12
    p.adaptive_dt = 0
    assert lib_dt(p, 1e-9, 2e-12, ts) == 2e-12
    # clamps
    p.adaptive_dt = 1
    p.dt_stretch = sys.float_info.max
    assert lib_dt(p, 1e-9, 1.0, (1.0, 1.0, 1.0)) == p.dt_max
    assert lib_dt(p, 1e-9, 1.0, (1e-20, 1e-20, 1e-20)) == p.dt_min


def _adaptive(c, stretch=sys.float_info.max, delay=0.0):
    p = c.params
    p.adaptive_dt = 1
    p.dt_stretch = stretch
    p.adaptive_dt_delay = delay
    p.dt_max = 1e-6  # the time scales decide
    return c


def test_oracle_steps_take_the_computed_dt():
    """Each step's dt is ComputeDt of the previous dt and the time scales of
    the state the step starts from; the first step (t = 0) keeps vidyut.dt."""
    c = _adaptive(cases.discharge_2d(n=32))
    grid = O.Grid(c.geom)
    st, t, dt = O.evolve(c, 1, return_dt=True)
    assert dt == c.dt and t == pytest.approx(c.dt, rel=1e-15, abs=0)
    cur_state, cur_t, cur_dt = st, t, dt
    seen = set()
    for _ in range(3):
        ts = O.time_scales(c.mech, c.params, grid, cur_state)
        want = rule(c.params, cur_t, cur_dt, ts)
        c2 = cases.discharge_2d(n=32)
        c2.params = c.params
        c2.state = cur_state
        nxt, nt, ndt = O.evolve(c2, 1, time=cur_t, dt=cur_dt, return_dt=True)
        assert ndt == want and nt == pytest.approx(cur_t + want, rel=1e-15, abs=0)
        seen.add(want)
        cur_state, cur_t, cur_dt = nxt, nt, ndt
    assert c.dt not in seen and min(seen) > 10 * c.dt  # the dielectric time, not vidyut.dt


def test_oracle_stretch_limits_growth_and_delay_holds_the_step():
    c = _adaptive(cases.discharge_2d(n=32), stretch=1.1)
    c.dt = 1e-13  # far below the CFL limits: the stretch decides
    _, t, dt = O.evolve(c, 6, return_dt=True)
    assert dt == pytest.approx(1e-13 * 1.1 ** 5, rel=1e-12, abs=0)
    assert t == pytest.approx(1e-13 * sum(1.1 ** k for k in range(6)), rel=1e-12, abs=0)
    c = _adaptive(cases.discharge_2d(n=32), stretch=1.1, delay=3.5e-13)
    c.dt = 1e-13
    _, t, dt = O.evolve(c, 5, return_dt=True)
    # steps 1-4 start at 0, 1, 2, 3 e-13 <= delay (fixed); step 5 at 4e-13: 1.1x
    assert dt == pytest.approx(1.1e-13, rel=1e-12, abs=0)
    assert t == pytest.approx(5.1e-13, rel=1e-12, abs=0)
