"""GPU parity: every entry point of the HIP library against the CPU oracle
on the same seeded inputs (called through the C ABI).

Tolerances: integer/index/copy work bit-exact; FP64 kernels within a few
ULP-scale relative differences (FMA contraction, exp/log ulps); solves run
to a tight residual and agree to 1e-10 relative; the full step driver
matches per field at relative L2 <= 1e-10.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from vidyut3d_amd import abi, cases, compare, native, plotfile

pytestmark = pytest.mark.gpu


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    s = np.max(np.abs(a))
    return float(np.max(np.abs(a - b)) / (s if s > 0 else 1.0))


def rand_state_2d(n, seed, ns=3):
    rng = np.random.default_rng(seed)
    cc = abi.comp(ns)
    st = np.zeros((cc["NVAR"], 1, n, n))
    st[0] = 1e15 * np.exp(rng.normal(size=(n, n)))
    st[1] = 1e15 * np.exp(rng.normal(size=(n, n)))
    st[2] = 2.4e25
    st[cc["ETEMP"]] = rng.uniform(5e3, 5e4, (n, n))
    st[cc["EEN"]] = 1.5 * st[0] * abi.K_B * st[cc["ETEMP"]]
    st[cc["EFX"]] = rng.normal(size=(n, n)) * 1e5
    st[cc["EFY"]] = rng.normal(size=(n, n)) * 1e5
    st[cc["POT"]] = rng.normal(size=(n, n)) * 100
    return st


def rand_state_3d(n, seed):
    c = cases.air_3d(n=n, seed=seed)
    rng = np.random.default_rng(seed + 7)
    cc = abi.comp(c.mech.nspec)
    shp = c.state.shape[1:]
    c.state[cc["EFX"]] = rng.normal(size=shp) * 1e5
    c.state[cc["EFY"]] = rng.normal(size=shp) * 1e5
    c.state[cc["EFZ"]] = rng.normal(size=shp) * 1e5
    c.state[cc["ETEMP"]] = rng.uniform(5e3, 5e4, shp)
    return c


# ------------------------------------------------------------ ghost fill
@pytest.mark.parametrize("dim", [2, 3])
def test_fill_patch_bit_exact(dev, dim):
    n = [9, 6, 5][:dim]
    per = [1, 0, 1][:dim]
    g = abi.make_geom(n, [0] * dim, [1] * dim, per)
    og, dg = O.Grid(g), native.Grid(dev, g)
    rng = np.random.default_rng(1)
    src = rng.normal(size=abi.shape_of(*og.extent(0), 4))
    ref = O.fill_patch(og, src, 3)
    s = dg.array(4, 0, host=src)
    d = dg.array(4, 3)
    native.fill_patch(dg, s, d)
    assert np.array_equal(d.get(), ref)


# ---------------------------------------------------------------- flux
@pytest.mark.parametrize("hyp,scheme", [(1, 3), (2, 3), (5, 1), (5, 2), (5, 3)])
@pytest.mark.parametrize("spec", [0, 1, 3])
def test_transport_flux_and_dsdt_2d(dev, hyp, scheme, spec):
    c = cases.discharge_2d(n=12)
    c.params.hyp_order, c.params.weno_scheme = hyp, scheme
    g = abi.make_geom([12, 10], [0, 0], [0.01, 0.008], [0, 1])
    og, dg = O.Grid(g), native.Grid(dev, g)
    st = rand_state_2d(12, 2)[:, :, :10, :]
    st = np.ascontiguousarray(st)
    ng = 3
    sb = O.fill_patch(og, st, ng)
    bclo, bchi = [abi.IHNEUBC, 0, 0], [abi.HNEUBC, 0, 0]
    fl = O.transport_flux(c.mech, c.params, og, sb, ng, spec, bclo, bchi, dt=1e-11)
    rng = np.random.default_rng(3)
    rxn = rng.normal(size=(4, 1, 10, 12)) * 1e20
    ds = O.dsdt(og, fl, rxn, spec)
    mech = native.Mech(dev, c.mech)
    dsb = dg.array(st.shape[0], ng, host=sb)
    dfl = dg.faces()
    native.transport_flux(mech, c.params, dg, dsb, spec, bclo, bchi, dfl, dt=1e-11)
    for d in range(2):
        assert relmax(fl[d], dfl[d].get()) < 1e-13, d
    drx = dg.array(4, host=rxn)
    dds = dg.array(1)
    native.dsdt(dg, dfl, drx, dds, spec)
    assert relmax(ds, dds.get()) < 1e-12
    dds2 = dg.array(1)
    native.transport_rhs(mech, c.params, dg, dsb, drx, dds2, spec, bclo, bchi, dt=1e-11)
    assert np.array_equal(dds.get(), dds2.get())


def test_transport_rhs_3d_weno_z(dev):
    c = rand_state_3d(10, 4)
    og, dg = O.Grid(c.geom), native.Grid(dev, c.geom)
    ns = c.mech.nspec
    sb = O.fill_patch(og, c.state, 3)
    rng = np.random.default_rng(5)
    rxn = rng.normal(size=(ns + 1,) + c.state.shape[1:]) * 1e20
    mech = native.Mech(dev, c.mech)
    dsb = dg.array(c.state.shape[0], 3, host=sb)
    drx = dg.array(ns + 1, host=rxn)
    for spec in (0, 2, 4, ns):
        fl = O.transport_flux(c.mech, c.params, og, sb, 3, spec, list(c.params.ion_bc_lo),
                              list(c.params.ion_bc_hi), dt=c.dt)
        ds = O.dsdt(og, fl, rxn, min(spec, ns))
        dds = dg.array(1)
        native.transport_rhs(mech, c.params, dg, dsb, drx, dds, spec, list(c.params.ion_bc_lo),
                             list(c.params.ion_bc_hi), dt=c.dt)
        assert relmax(ds, dds.get()) < 1e-12, spec


@pytest.mark.parametrize("scale", [1e30, 1e-30])
def test_transport_rhs_3d_weno_z_extreme_magnitudes(dev, scale):
    """WENO-Z forms its three tau/(beta+eps) ratios and its two weight
    normalisations from one reciprocal each; where those products leave the
    normal range (densities ~1e45, fluxes ~1e51: smoothness indicators
    ~1e100 and weight sums ~1e240) it falls back to separate quotients.
    Both regimes against the oracle's per-face IEEE evaluation.  (Above
    ~1e53 the weight-times-value numerators overflow, where the
    reference's weights-first order does not.)"""
    c = rand_state_3d(10, 6)
    ns = c.mech.nspec
    c.state[:ns] *= scale
    og, dg = O.Grid(c.geom), native.Grid(dev, c.geom)
    sb = O.fill_patch(og, c.state, 3)
    rng = np.random.default_rng(8)
    rxn = rng.normal(size=(ns + 1,) + c.state.shape[1:]) * 1e20 * scale
    mech = native.Mech(dev, c.mech)
    dsb = dg.array(c.state.shape[0], 3, host=sb)
    drx = dg.array(ns + 1, host=rxn)
    for spec in (0, 2):
        fl = O.transport_flux(c.mech, c.params, og, sb, 3, spec, list(c.params.ion_bc_lo),
                              list(c.params.ion_bc_hi), dt=c.dt)
        ds = O.dsdt(og, fl, rxn, spec)
        dds = dg.array(1)
        native.transport_rhs(mech, c.params, dg, dsb, drx, dds, spec, list(c.params.ion_bc_lo),
                             list(c.params.ion_bc_hi), dt=c.dt)
        assert relmax(ds, dds.get()) < 1e-12, spec


# ------------------------------------------------------------ chemistry
def test_rxn_ckwc_2d(dev):
    c = cases.discharge_2d(n=16)
    og, dg = O.Grid(c.geom), native.Grid(dev, c.geom)
    st = rand_state_2d(16, 6)
    ref = O.rxn_source(c.mech, c.params, og, st, 0)
    mech = native.Mech(dev, c.mech)
    out = dg.array(4)
    native.rxn_source(mech, c.params, dg, dg.array(st.shape[0], host=st), out)
    got = out.get()
    for q in range(4):
        assert relmax(ref[q], got[q]) < 1e-12, q


@pytest.mark.parametrize("specialised", [True, False])
def test_rxn_table_40_species(dev, specialised):
    """Table chemistry, 40 species / 200 reactions, through the
    mechanism-specialised kernel and through the generic table kernel."""
    m = cases.co2_mechanism(seed=3)
    n = 8
    g = abi.make_geom([n] * 3, [0] * 3, [1] * 3, [0] * 3)
    og, dg = O.Grid(g), native.Grid(dev, g)
    st = cases.co2_state(n, seed=4)
    prm = cases.default_params()
    ref = O.rxn_source(m, prm, og, st, 0)
    mech = native.Mech(dev, m)
    mech.specialise(specialised)
    assert mech.specialised == specialised
    out = dg.array(m.nspec + 1)
    native.rxn_source(mech, prm, dg, dg.array(st.shape[0], host=st), out)
    got = out.get()
    for q in range(m.nspec + 1):
        scale = np.max(np.abs(ref[q]))
        assert np.max(np.abs(ref[q] - got[q])) <= 1e-11 * scale, q


def test_elec_energy_source(dev):
    c = cases.discharge_2d(n=12)
    og, dg = O.Grid(c.geom), native.Grid(dev, c.geom)
    st = rand_state_2d(12, 7)
    sb = O.fill_patch(og, st, 3)
    rng = np.random.default_rng(8)
    ef = [rng.normal(size=f.shape) * 1e5 for f in og.faces()]
    gn = [rng.normal(size=f.shape) * 1e18 for f in og.faces()]
    rxn = rng.normal(size=(4,) + st.shape[1:])
    ds0 = rng.normal(size=(1,) + st.shape[1:])
    for floor in (0, 1):
        c.params.floor_jh = floor
        ds_ref, pn_ref = O.elec_energy_source(c.mech, c.params, og, sb, 3, rxn, ef, gn, ds0, st)
        mech = native.Mech(dev, c.mech)
        dims = [0, 1, 0]
        def_faces = lambda arrs: [dg.array(1, 0, dims[q], host=a) for q, a in enumerate(arrs)]
        dds = dg.array(1, host=ds0)
        dpn = dg.array(st.shape[0], host=st)
        native.elec_energy_source(mech, c.params, dg, dg.array(st.shape[0], 3, host=sb),
                                  dg.array(4, host=rxn), def_faces(ef), def_faces(gn), dds, dpn)
        assert relmax(ds_ref, dds.get()) < 1e-12
        assert relmax(pn_ref, dpn.get()) < 1e-12


# ------------------------------------------------------------- solvers
def tight(prm, tol=1e-13):
    prm.linsolve_reltol, prm.linsolve_abstol = tol, 0.0
    prm.linsolve_maxiter = 200
    return prm


@pytest.mark.parametrize("n", [16, 32])
def test_solve_potential_2d(dev, n):
    c = cases.discharge_2d(n=n)
    tight(c.params)
    og, dg = O.Grid(c.geom), native.Grid(dev, c.geom)
    st = rand_state_2d(n, 9)
    sb = O.fill_patch(og, st, 3)
    pn_ref, ef_ref, info_ref = O.solve_potential(c.mech, c.params, og, sb, 3, st)
    mech = native.Mech(dev, c.mech)
    dpn = dg.array(st.shape[0], host=st)
    def_ = dg.faces()
    info = native.solve_potential(mech, c.params, dg, dg.array(st.shape[0], 3, host=sb), dpn, def_)
    got = dpn.get()
    cc = abi.comp(3)
    for comp in ("POT", "EFX", "EFY", "REF"):
        assert relmax(pn_ref[cc[comp]], got[cc[comp]]) < 1e-10, comp
    for d in range(2):
        assert relmax(ef_ref[d], def_[d].get()) < 1e-10
    assert abs(info.iterations - info_ref.iterations) <= 1
    assert info.resnorm <= info.target


def test_solve_potential_3d_neumann_dirichlet(dev):
    c = rand_state_3d(16, 10)
    tight(c.params)
    og, dg = O.Grid(c.geom), native.Grid(dev, c.geom)
    sb = O.fill_patch(og, c.state, 3)
    pn_ref, ef_ref, _ = O.solve_potential(c.mech, c.params, og, sb, 3, c.state)
    mech = native.Mech(dev, c.mech)
    dpn = dg.array(c.state.shape[0], host=c.state)
    def_ = dg.faces()
    native.solve_potential(mech, c.params, dg, dg.array(c.state.shape[0], 3, host=sb), dpn, def_)
    got = dpn.get()
    cc = abi.comp(c.mech.nspec)
    for comp in ("POT", "EFX", "EFY", "EFZ", "REF"):
        assert relmax(pn_ref[cc[comp]], got[cc[comp]]) < 1e-10, comp


@pytest.mark.parametrize("spec", [0, 1, 3])
def test_implicit_solve_2d(dev, spec):
    c = cases.discharge_2d(n=16)
    tight(c.params)
    og, dg = O.Grid(c.geom), native.Grid(dev, c.geom)
    st = rand_state_2d(16, 11)
    sb = O.fill_patch(og, st, 3)
    rng = np.random.default_rng(12)
    ds = rng.normal(size=(1,) + st.shape[1:]) * 1e24
    old = st * (1 + 0.01 * rng.normal(size=st.shape))
    bl = list(c.params.eden_bc_lo) if spec in (0, 3) else list(c.params.ion_bc_lo)
    bh = list(c.params.eden_bc_hi) if spec in (0, 3) else list(c.params.ion_bc_hi)
    pn_ref, gr_ref, _ = O.implicit_solve(c.mech, c.params, og, spec, sb, 3, old, ds, st, bl, bh, 1e-11)
    mech = native.Mech(dev, c.mech)
    dpn = dg.array(st.shape[0], host=st)
    dgr = dg.faces()
    native.implicit_solve(mech, c.params, dg, spec, dg.array(st.shape[0], 3, host=sb),
                          dg.array(st.shape[0], host=old), dg.array(1, host=ds), dpn, dgr, bl, bh, 1e-11)
    got = dpn.get()
    cc = abi.comp(3)
    assert relmax(pn_ref[spec], got[spec]) < 1e-10
    if spec == 3:
        assert relmax(pn_ref[cc["ETEMP"]], got[cc["ETEMP"]]) < 1e-10
    if spec == 0:
        for d in range(2):
            assert relmax(gr_ref[d], dgr[d].get()) < 1e-8


def _dielectric(n, prec):
    g, faces, rhs = cases.dielectric_poisson(n=n)
    return g, faces, rhs


@pytest.mark.parametrize("prec", [0, 1])
def test_mlmg_dielectric_fp64_and_mixed(dev, prec):
    n = 32
    g, faces, rhs = cases.dielectric_poisson(n=n)
    og, dg = O.Grid(g), native.Grid(dev, g)
    rhs = rhs[None]

    def problem(views, precision):
        p = abi.vd_mlmg_problem()
        p.alpha_a, p.beta = 0.0, 1.0
        for d in range(3):
            p.bface[d] = C.pointer(views[d])
        p.bc_lo[:] = [0, 0, abi.DIRCBC]
        p.bc_hi[:] = [0, 0, abi.DIRCBC]
        p.reltol, p.abstol, p.maxiter = 1e-12, 0.0, 200
        p.precision = precision
        p.pre_smooth = p.post_smooth = 2
        p.max_coarsening_level = 10
        return p

    hv = [og.view(np.ascontiguousarray(faces[d][None]), 0, d) for d in range(3)]
    hf = [np.ascontiguousarray(faces[d][None]) for d in range(3)]
    hv = [og.view(hf[d], 0, d) for d in range(3)]
    ref, info_ref = O.mlmg_solve(og, problem(hv, 0), og.zeros(1), np.ascontiguousarray(rhs))
    df = [dg.array(1, 0, d, host=hf[d]) for d in range(3)]
    dv = [f.view() for f in df]
    dphi = dg.array(1)
    info = native.mlmg_solve(dg, problem(dv, prec), dphi, dg.array(1, host=rhs))
    got = dphi.get()
    err = compare.field_errors(ref, got)["rel_l2"]
    assert err < (1e-10 if prec == 0 else 1e-8), err
    assert info.resnorm <= info.target


@pytest.mark.parametrize("shape", [(5, 40, 36), (12, 20, 44), (8, 8, 1000), (6, 1000)])
def test_mlmg_odd_and_thin_shapes(dev, shape):
    """Levels that stop coarsening early: a 7200-cell bottom level that is
    also the finest (the coarse V-cycle then keeps its levels in memory,
    not LDS), odd dimensions after two coarsenings, a thin column whose
    coarse tail starts two levels down, and a thin 2-D strip."""
    rng = np.random.default_rng(sum(shape))
    dim = len(shape)
    g = abi.make_geom(list(shape), [0.0] * dim, [0.01 * n for n in shape], [0] * dim)  # cubic cells
    og, dg = O.Grid(g), native.Grid(dev, g)
    dims = [0, 1, 2] if dim == 3 else [0, 1, 0]
    faces = []
    for d in range(3):
        fs = [1] + list(reversed(shape)) if dim == 3 else [1, 1] + list(reversed(shape))
        if d < dim:
            fs[len(fs) - 1 - d] += 1
        faces.append(np.exp(0.3 * rng.normal(size=fs)))
    rhs = rng.normal(size=[1] + ([1] if dim == 2 else []) + list(reversed(shape)))

    def problem(views):
        p = abi.vd_mlmg_problem()
        p.alpha_a, p.beta = 0.0, 1.0
        for d in range(3):
            p.bface[d] = C.pointer(views[dims[d]])
        p.bc_lo[:] = [abi.DIRCBC, abi.HNEUBC, abi.DIRCBC]
        p.bc_hi[:] = [abi.HNEUBC, abi.DIRCBC, abi.DIRCBC]
        p.reltol, p.abstol, p.maxiter = 1e-12, 0.0, 400
        p.precision = 0
        p.pre_smooth = p.post_smooth = 2
        p.max_coarsening_level = 10
        return p

    hv = [og.view(faces[d], 0, d) for d in range(dim)]
    ref, _ = O.mlmg_solve(og, problem(hv), og.zeros(1), rhs)
    df = [dg.array(1, 0, d, host=faces[d]) for d in range(dim)]
    dv = [f.view() for f in df]
    phi = dg.array(1)
    info = native.mlmg_solve(dg, problem(dv), phi, dg.array(1, host=rhs))
    err = compare.field_errors(ref, phi.get())["rel_l2"]
    assert err <= 1e-10, err
    assert info.resnorm <= info.target


# ---------------------------------------------------------- CFL and tags
def test_time_scales(dev):
    c = rand_state_3d(12, 13)
    og, dg = O.Grid(c.geom), native.Grid(dev, c.geom)
    ref = O.time_scales(c.mech, c.params, og, c.state)
    mech = native.Mech(dev, c.mech)
    got = native.time_scales(mech, c.params, dg, dg.array(c.state.shape[0], host=c.state))
    for a, b in zip(ref, got):
        assert a == pytest.approx(b, rel=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_tags_bit_exact(dev, dim):
    n = [13, 11, 9][:dim]
    g = abi.make_geom(n, [0] * dim, [1] * dim, [0, 1, 0][:dim])
    og, dg = O.Grid(g), native.Grid(dev, g)
    rng = np.random.default_rng(14)
    st = rng.normal(size=abi.shape_of(*og.extent(0), 6))
    s1 = O.fill_patch(og, st, 1)
    comps, thr, gthr = [0, 2, 5], [1.0, 2.0, 1.5], [1.7, 2.2, 0.9]
    ref = O.error_est(og, s1, comps, thr, gthr)
    b, e = og.extent(0)
    tags = native.DeviceArray(dev, b, e, 1, dtype=np.int8)
    native.error_est(dg, dg.array(6, 1, host=s1), tags, comps, thr, gthr)
    assert np.array_equal(tags.get()[0], ref)


# ---------------------------------------------------------- whole steps
@pytest.mark.parametrize("nsteps", [3, 20])
def test_evolve_discharge_2d_matches_oracle(dev, nsteps):
    c = cases.discharge_2d(n=32, nsteps=nsteps)
    tight(c.params, 1e-12)
    ref, tref = O.evolve(c, nsteps)
    sim = native.Sim(dev, c)
    sim.evolve(nsteps)
    got = sim.get_state()
    assert sim.time == pytest.approx(tref)
    errs = compare.state_errors(ref, got, c.mech.nspec, c.geom.dim)
    bad = {q: e for q, e in errs.items() if not e <= 1e-10}
    assert not bad, bad


@pytest.mark.parametrize("bg", [0, 1])
def test_evolve_split_calls_match_one_call_and_oracle(dev, bg):
    """Evolving in calls of 1, 3 and 2 steps runs a single-step call and the
    first, middle and last steps of the one-corrector driver, which moves
    the state between phi_new and two ghosted buffers; the state equals that
    of one 6-step call bit for bit, and the oracle's (bg = 1: background
    species advanced by their reaction source, the electron source added to
    the solved density)."""
    c = cases.discharge_2d(n=32, nsteps=6)
    c.params.do_bg_reactions = bg
    c.params.bound_specden = bg
    tight(c.params, 1e-12)
    ref, _ = O.evolve(c, 6)
    one = native.Sim(dev, c)
    one.evolve(6)
    s_one = one.get_state()
    one.close()
    split = native.Sim(dev, c)
    for n in (1, 3, 2):
        split.evolve(n)
    s_split = split.get_state()
    split.close()
    assert np.array_equal(s_one, s_split)
    errs = compare.state_errors(ref, s_split, c.mech.nspec, c.geom.dim)
    bad = {q: e for q, e in errs.items() if not e <= 1e-10}
    assert not bad, bad


def test_evolve_air_3d_matches_oracle(dev):
    c = cases.air_3d(n=16, nsteps=4)
    tight(c.params, 1e-12)
    ref, _ = O.evolve(c, 4)
    sim = native.Sim(dev, c)
    sim.evolve(4)
    got = sim.get_state()
    errs = compare.state_errors(ref, got, c.mech.nspec, c.geom.dim)
    bad = {q: e for q, e in errs.items() if not e <= 1e-10}
    assert not bad, bad


def test_sim_error_est_matches_oracle_tags(dev):
    c = cases.discharge_2d(n=32, nsteps=2)
    sim = native.Sim(dev, c)
    sim.evolve(2)
    st = sim.get_state()
    og = O.Grid(c.geom)
    cc = abi.comp(3)
    comps, thr, gthr = [cc["POT"], 0], [1e20, 5e15], [10.0, 1e14]
    ref = O.error_est(og, O.fill_patch(og, st, 1), comps, thr, gthr)
    got = sim.error_est(comps, thr, gthr)
    assert np.array_equal(got, ref)
    assert ref.any()


def test_discharge_2d_full_config_matches_oracle(dev):
    """The 2-D config at its full size: 128^2, 100 steps of 1e-11 s, the
    reference's default solver tolerances (1e-10 relative and absolute)."""
    c = cases.discharge_2d(n=128, nsteps=100)
    ref, _ = O.evolve(c, 100)
    sim = native.Sim(dev, c)
    sim.evolve(100)
    got = sim.get_state()
    errs = compare.state_errors(ref, got, c.mech.nspec, c.geom.dim)
    print({q: "%.2e" % e for q, e in errs.items()})
    bad = {q: e for q, e in errs.items() if not e <= 1e-10}
    assert not bad, bad


def test_mixed_precision_step_within_1e8_of_fp64_oracle(dev):
    c = cases.air_3d(n=24, nsteps=4)
    c.params.mg_precision = 1
    ref_case = cases.air_3d(n=24, nsteps=4)
    ref, _ = O.evolve(ref_case, 4)
    sim = native.Sim(dev, c)
    sim.evolve(4)
    got = sim.get_state()
    errs = compare.state_errors(ref, got, c.mech.nspec, c.geom.dim)
    print({q: "%.2e" % e for q, e in errs.items()})
    bad = {q: e for q, e in errs.items() if not e <= 1e-8}
    assert not bad, bad


# ---------------------------------------------------------- plotfile output
def test_plotfile_from_sim_reads_back_and_matches_oracle(dev, tmp_path):
    """WritePlotFile (Source/Plot.cpp:29-38) of the device state reads back
    bit-exact, and the reference comparator metric
    (py_utils/compare_fields_from_twoplotfiles.py:44-46) against the
    oracle's state written as a plotfile stays within 1e-10."""
    from vidyut3d_amd import plotfile
    c = cases.discharge_2d(n=32, nsteps=5)
    tight(c.params, 1e-12)
    ref, tref = O.evolve(c, 5)
    sim = native.Sim(dev, c)
    sim.evolve(5)
    p = str(tmp_path / "plt00005")
    sim.write_plotfile(p)
    pf = plotfile.read(p)
    assert np.array_equal(pf.data, sim.get_state())
    names = [c.mech.names[s].value.decode() for s in range(c.mech.nspec)] + abi.PLASMA_VAR_NAMES
    assert pf.names == names and pf.step == 5 and pf.dim == 2
    assert pf.time == pytest.approx(tref, rel=1e-15)
    assert pf.prob_lo == [c.geom.prob_lo[0], c.geom.prob_lo[1]]
    q = str(tmp_path / "oracle00005")
    plotfile.write(q, names, ref, 2, tref, 5, pf.prob_lo, pf.prob_hi)
    errs = plotfile.compare(q, p)
    vec = np.sqrt(np.mean(ref[abi.comp(3)["EFX"]] ** 2 + ref[abi.comp(3)["EFY"]] ** 2))
    for name, e in errs.items():
        if name in ("Efieldx", "Efieldy"):
            assert e["err"] <= 1e-10 * vec, (name, e)  # component of a vector: scale by |E|
        else:
            assert e["rel"] <= 1e-10, (name, e)


def test_checkpoint_restart_is_bit_exact(dev, tmp_path):
    """WriteCheckpointFile / ReadCheckpointFile (Source/Plot.cpp:40-209):
    the checkpoint holds the state, step, dt and time (read back by the
    Python reader of the same layout), and a run restarted from it takes
    the same steps as the run that wrote it, bit for bit."""
    c = cases.air_3d(n=24, nsteps=5)
    a = native.Sim(dev, c)
    a.evolve(3)
    chk = str(tmp_path / "chk00003")
    a.write_checkpoint(chk)
    s3, t3 = a.get_state(), a.time
    a.evolve(2)
    s_a = a.get_state()
    a.close()
    cp = plotfile.read_checkpoint(chk)
    assert (cp.step, cp.dt, cp.time) == (3, c.dt, t3)
    assert np.array_equal(cp.data, s3)
    b = native.Sim(dev, cases.air_3d(n=24, nsteps=5))
    b.read_checkpoint(chk)
    assert (b.step, b.time) == (3, t3)
    assert np.array_equal(b.get_state(), s3)
    b.evolve(2)
    s_b = b.get_state()
    b.close()
    assert np.array_equal(s_a, s_b)


@pytest.mark.gpu
def test_handles_outliving_their_device_are_released_with_it():
    """Closing a device releases the arrays, mechanisms and sims made on it;
    dropping them afterwards frees nothing through the destroyed context."""
    d = native.Device(0)
    c = cases.discharge_2d(n=16, seed=1, nsteps=1)
    a = native.Grid(d, c.geom).array(2, 1)
    m = native.Mech(d, c.mech)
    s = native.Sim(d, c)
    d.close()
    assert not a.ptr and not m.ptr and not s.ptr
    del a, m, s


def test_memset_zero_clears_exactly_the_range(dev):
    """vd_memset_zero on ranges of every alignment and length: the bytes in
    the range become 0, every byte around it keeps its value."""
    nb = (1 << 20) + 64
    base = dev.alloc(nb)
    try:
        rng = np.random.default_rng(7)
        for off, n in [(0, 0), (3, 1), (1, 15), (0, 16), (5, 17), (16, 31), (7, 1000), (13, (1 << 20) + 3),
                       (0, 1 << 20)]:
            init = rng.integers(1, 256, nb, dtype=np.uint8)
            dev.upload(base, init)
            native.check(native.lib().vd_memset_zero(dev.ptr, C.c_void_p(base + off), C.c_size_t(n)),
                         "vd_memset_zero")
            got = dev.download(base, np.empty(nb, np.uint8))
            want = init.copy()
            want[off:off + n] = 0
            assert np.array_equal(got, want), (off, n)
    finally:
        dev.free(base)


# ---------------------------------------------------- voltage profiles
@pytest.mark.parametrize("profile,frac", [(1, 0.25), (1, 0.6), (2, 0.0), (2, 0.3), (2, 0.7)])
def test_potential_solve_voltage_profiles_match_closed_form_and_oracle(dev, profile, frac):
    """Sine and triangle applied voltages (Source/HelperFuncs.H:75-92) on a
    vacuum gap: the potential is linear between the two plate voltages at
    time t, E = -(V2 - V1)/L, and equals the oracle's solve."""
    import test_oracle_closed_forms as cf
    c, p, g, st, cc = cf.laplace_case(profile)
    t = frac / p.voltage_freq if profile == 1 else p.voltage_center + frac * p.voltage_dur
    v1, v2 = O.applied_potential(t, -1, p), O.applied_potential(t, 1, p)
    og, dg = O.Grid(g), native.Grid(dev, g)
    ref_pn, _, _ = O.solve_potential(c.mech, p, og, O.fill_patch(og, st, 1), 1, st, time=t)
    m = native.Mech(dev, c.mech)
    src = dg.array(st.shape[0], 0, host=st)
    sb = dg.array(st.shape[0], 1)
    native.fill_patch(dg, src, sb)
    pn = dg.array(st.shape[0], 0, host=st)
    ef = dg.faces()
    native.solve_potential(m, p, dg, sb, pn, ef, time=t)
    got = pn.get()
    x = (np.arange(16) + 0.5) * g.dx[0]
    scale = max(abs(v1), abs(v2), 1.0)
    np.testing.assert_allclose(got[cc["POT"], 0], np.broadcast_to(v1 + (v2 - v1) * x / 0.01, (4, 16)),
                               rtol=0, atol=1e-10 * scale)
    np.testing.assert_allclose(got[cc["EFX"], 0], np.full((4, 16), -(v2 - v1) / 0.01), rtol=1e-10,
                               atol=1e-8 * scale)
    assert relmax(ref_pn[cc["POT"]], got[cc["POT"]]) < 1e-12


# ------------------------------------------------------ adaptive time step
def test_adaptive_dt_steps_match_oracle_across_calls(dev):
    """ComputeDt (Source/TimeStep.cpp:14-31) in the GPU driver: the same dt
    sequence as the oracle's (time after each call), the same states, and
    calls of 2 + 3 steps equal one call of 5 (the dt carries over)."""
    c = cases.discharge_2d(n=32, nsteps=5)
    tight(c.params, 1e-12)
    c.params.adaptive_dt = 1
    c.params.dt_max = 1e-6
    c.params.dt_stretch = 4.0
    ref, tref, dref = O.evolve(c, 5, return_dt=True)
    sim = native.Sim(dev, c)
    sim.evolve(2)
    sim.evolve(3)
    got = sim.get_state()
    assert sim.time == pytest.approx(tref, rel=1e-12, abs=0)
    assert sim.dt == pytest.approx(dref, rel=1e-12, abs=0)
    assert dref > 10 * c.dt
    sim.close()
    errs = compare.state_errors(ref, got, c.mech.nspec, c.geom.dim)
    bad = {q: e for q, e in errs.items() if not e <= 1e-10}
    assert not bad, bad
