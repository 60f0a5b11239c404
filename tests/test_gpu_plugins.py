"""The case plug-ins on the GPU against the oracle: fitted specMob/specDiff
(Source/PlasmaChem.H:186-206 with a case's Chemistry.H), the prescribed
drift velocity of user_transport::compute_vel, user Robin species data and
user reaction sources, per entry point and through whole steps (the
reference's driftDiffuse and Advect verification problems)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from vidyut3d_amd import abi, cases, compare, native

pytestmark = pytest.mark.gpu


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    s = np.max(np.abs(a))
    return float(np.max(np.abs(a - b)) / (s if s > 0 else 1.0))


def fitted_air(n=16, nsteps=4):
    """The 3-D air case with state-dependent transport: the electron takes
    the Streamer power laws in |E| (mobility and diffusion), the positive
    ions hard-sphere mobilities in ndens and Tg (Einstein diffusion at Tg),
    the energy 5/3 of the electron's coefficients."""
    c = cases.air_3d(n=n, nsteps=nsteps)
    m = c.mech
    abi.set_fit(m.mob_fit[0], -1.0, c0=math.log(2.3987), c1=-0.26)
    abi.set_fit(m.diff_fit[0], 1.0, c0=math.log(4.3628e-3), c1=0.22)
    for s in range(m.nspec):
        if m.charge[s] != 0 and s != m.e_idx:
            sgn = 1.0 if m.charge[s] > 0 else -1.0
            # mu = e / (m nu), nu = sqrt(8 k Tg / (pi m)) N sigma (Ar_DC_1d form)
            mass = m.molwt[s] * abi.M_AMU
            c0 = math.log(abi.ECHARGE / mass / (math.sqrt(8.0 * abi.K_B / math.pi / mass) * 120e-20))
            abi.set_fit(m.mob_fit[s], sgn, c0=c0, c3=-0.5, c4=-1.0)
    c.params.linsolve_reltol, c.params.linsolve_abstol = 1e-12, 0.0
    return c


def _random_fields(c, seed):
    rng = np.random.default_rng(seed)
    st = c.state.copy()
    cc = abi.comp(c.mech.nspec)
    for d in range(c.geom.dim):
        st[cc["EFX"] + d] = rng.normal(size=st.shape[1:]) * 1e5
    st[cc["ETEMP"]] = rng.uniform(5e3, 5e4, st.shape[1:])
    st[cc["EEN"]] = 1.5 * st[0] * abi.K_B * st[cc["ETEMP"]]
    return st


@pytest.mark.parametrize("spec", [0, 1, 4, 10])
def test_transport_rhs_with_fitted_mobilities(dev, spec):
    """e-, N2+, O2- and the electron energy with state-dependent mobilities:
    the GPU precomputes the drift velocity per cell, the oracle calls
    specMob per face."""
    c = fitted_air(n=16)
    st = _random_fields(c, 3 + spec)
    og, dg = O.Grid(c.geom), native.Grid(dev, c.geom)
    sb = O.fill_patch(og, st, 3)
    rng = np.random.default_rng(spec)
    rxn = rng.normal(size=(c.mech.nspec + 1,) + st.shape[1:]) * 1e10
    bclo, bchi = [abi.HNEUBC, abi.HNEUBC, abi.IHNEUBC], [abi.HNEUBC, abi.HNEUBC, abi.IHNEUBC]
    fl = O.transport_flux(c.mech, c.params, og, sb, 3, spec, bclo, bchi)
    ref = O.dsdt(og, fl, rxn, spec)
    m = native.Mech(dev, c.mech)
    out = dg.array(1)
    native.transport_rhs(m, c.params, dg, dg.array(st.shape[0], 3, host=sb), dg.array(rxn.shape[0], host=rxn),
                         out, spec, bclo, bchi)
    assert relmax(ref, out.get()) < 1e-12


def test_transport_with_prescribed_velocity(dev):
    c = cases.advect_1d(40)
    og, dg = O.Grid(c.geom), native.Grid(dev, c.geom)
    st = c.state.copy()
    cc = abi.comp(2)
    st[cc["EFX"]] = 7e4  # ignored: the velocity is prescribed
    sb = O.fill_patch(og, st, 3)
    rxn = np.zeros((3,) + st.shape[1:])
    ref = O.dsdt(og, O.transport_flux(c.mech, c.params, og, sb, 3, 1, [0] * 3, [0] * 3), rxn, 1)
    m = native.Mech(dev, c.mech)
    out = dg.array(1)
    native.transport_rhs(m, c.params, dg, dg.array(st.shape[0], 3, host=sb), dg.array(3, host=rxn), out, 1,
                         [0] * 3, [0] * 3)
    assert relmax(ref, out.get()) < 1e-12


@pytest.mark.parametrize("spec", [0, 10])
def test_implicit_solve_with_fitted_diffusion(dev, spec):
    c = fitted_air(n=16)
    st = _random_fields(c, 11)
    og, dg = O.Grid(c.geom), native.Grid(dev, c.geom)
    sb1 = O.fill_patch(og, st, 1)
    rng = np.random.default_rng(5)
    q = spec if spec < c.mech.nspec else abi.comp(c.mech.nspec)["EEN"]
    ds = rng.normal(size=(1,) + st.shape[1:]) * 1e-3 * np.abs(st[q]).mean() / c.dt
    bclo, bchi = list(c.params.eden_bc_lo), list(c.params.eden_bc_hi)
    pn_ref, gr_ref, info_ref = O.implicit_solve(c.mech, c.params, og, spec, sb1, 1, st, ds, st.copy(), bclo,
                                                bchi, c.dt)
    m = native.Mech(dev, c.mech)
    pn = dg.array(st.shape[0], host=st)
    gr = dg.faces()
    native.implicit_solve(m, c.params, dg, spec, dg.array(st.shape[0], 1, host=sb1), dg.array(st.shape[0], host=st),
                          dg.array(1, host=ds), pn, gr, bclo, bchi, c.dt)
    got = pn.get()
    cc = abi.comp(c.mech.nspec)
    for comp in ([spec] if spec < c.mech.nspec else [cc["EEN"], cc["ETEMP"]]):
        assert relmax(pn_ref[comp], got[comp]) < 1e-10, comp


def test_joule_and_time_scales_with_fitted_electron_coefficients(dev):
    c = fitted_air(n=16)
    st = _random_fields(c, 21)
    og, dg = O.Grid(c.geom), native.Grid(dev, c.geom)
    m = native.Mech(dev, c.mech)
    ref = O.time_scales(c.mech, c.params, og, st)
    got = native.time_scales(m, c.params, dg, dg.array(st.shape[0], host=st))
    np.testing.assert_allclose(got, ref, rtol=1e-13)
    rng = np.random.default_rng(2)
    sb1 = O.fill_patch(og, st, 1)
    rxn = rng.normal(size=(c.mech.nspec + 1,) + st.shape[1:])
    ef = [rng.normal(size=f.shape) * 1e5 for f in og.faces()]
    gn = [rng.normal(size=f.shape) * 1e18 for f in og.faces()]
    ds0 = rng.normal(size=(1,) + st.shape[1:])
    ds_ref, pn_ref = O.elec_energy_source(c.mech, c.params, og, sb1, 1, rxn, ef, gn, ds0, st)
    ds = dg.array(1, host=ds0)
    pn = dg.array(st.shape[0], host=st)
    native.elec_energy_source(m, c.params, dg, dg.array(st.shape[0], 1, host=sb1), dg.array(rxn.shape[0], host=rxn),
                              [dg.array(1, 0, d, host=ef[d]) for d in range(3)],
                              [dg.array(1, 0, d, host=gn[d]) for d in range(3)], ds, pn)
    assert relmax(ds_ref, ds.get()) < 1e-12
    cc = abi.comp(c.mech.nspec)
    assert relmax(pn_ref[cc["EJH"]], pn.get()[cc["EJH"]]) < 1e-12


def test_user_reaction_sources(dev):
    c = cases.drift_diffuse(16)
    c.mech.user_src[0] = -3.5
    st = c.state.copy()
    st[1:3] = 0.25
    og, dg = O.Grid(c.geom), native.Grid(dev, c.geom)
    ref = O.rxn_source(c.mech, c.params, og, O.fill_patch(og, st, 1), 1)
    m = native.Mech(dev, c.mech)
    out = dg.array(c.mech.nspec + 1)
    native.rxn_source(m, c.params, dg, dg.array(st.shape[0], 1, host=O.fill_patch(og, st, 1)), out)
    assert np.array_equal(ref, out.get())
    assert np.all(ref[2] == 1.0) and np.all(ref[0] == -3.5)


def _names(c):
    return [c.mech.names[s].value.decode() for s in range(c.mech.nspec)] + abi.PLASMA_VAR_NAMES


def test_fitted_air_steps_match_oracle(dev):
    c = fitted_air(n=16, nsteps=4)
    ref, _ = O.evolve(c, 4)
    sim = native.Sim(dev, c)
    sim.evolve(4)
    got = sim.get_state()
    sim.close()
    errs = compare.state_errors(ref, got, c.mech.nspec, 3)
    print({n: "%.2e" % errs[q] for q, n in enumerate(_names(c))}, flush=True)
    bad = {q: e for q, e in errs.items() if not e <= 1e-10}
    assert not bad, bad


def test_drift_diffuse_on_the_gpu_matches_oracle_and_closed_forms(dev):
    """driftDiffuse to its steady state: the GPU run equals the oracle's and
    is within first order (S1) / second order (S2) of verify_spec.py's
    closed forms at 64 cells."""
    c = cases.drift_diffuse(64)
    c.params.linsolve_reltol, c.params.linsolve_abstol = 1e-12, 0.0
    ref, _ = O.evolve(c, 150)
    sim = native.Sim(dev, c)
    sim.evolve(150)
    got = sim.get_state()
    sim.close()
    for q in (1, 2):
        assert compare.field_errors(ref[q], got[q])["rel_l2"] < 1e-10, q
    x = (np.arange(64) + 0.5) / 64
    s1, s2 = cases.drift_diffuse_exact(x)
    assert np.max(np.abs(got[1, 0, 0] - s1)) < 0.03
    assert np.max(np.abs(got[2, 0, 0] - s2)) < 4e-5


def test_advect_two_correctors_on_the_gpu(dev):
    c = cases.advect_1d(20, nsteps=2000, dt=1e-3)
    ref, t = O.evolve(c, 2000, ncorr=2)
    sim = native.Sim(dev, c)
    sim.evolve(2000, ncorr=2)
    got = sim.get_state()
    assert sim.time == pytest.approx(t, rel=1e-13, abs=0)
    sim.close()
    assert compare.field_errors(ref[1], got[1])["rel_l2"] < 1e-10
    x = -1.0 + (np.arange(20) + 0.5) * 0.1
    exact = np.sin(np.pi * (x - t) - np.sin(np.pi * (x - t)) / np.pi)
    assert np.sqrt(np.mean((got[1, 0, 1] - exact) ** 2)) < 3e-3
