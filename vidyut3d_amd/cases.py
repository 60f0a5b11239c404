"""Discharge cases: the per-case plug-ins (Chemistry.H, UserFunctions.H,
Prob.H, inputs) of the reference, restated as data for the C ABI.

* ``discharge_2d``   -- 2-D electron / ion / background-neutral discharge
  with the electron energy equation; physics template
  test/models/ionelec_DC_1d/{Chemistry.H,UserFunctions.H,inputs1d} and the
  Einstein relation of test/models/Axisymmetric_streamer_Ar/Chemistry.H:
  207-240.  Ionisation k = 2.5e-12*exp(-2.5e5/Te) m^3/s, mu_e = -0.03,
  mu_i = 2e-4 m^2/Vs, 128^2 cells over 1 cm^2, 0 V / 1 kV on the x faces,
  periodic in y, dt = 1e-11 s.
* ``air_3d``         -- 3-D discharge with a seeded 10-species /
  30-reaction air-like mechanism (table form of Source/reaction_source.H).
* ``co2_mechanism``  -- seeded 40-species / 200-reaction mechanism for the
  stiff-chemistry source evaluation.
* ``dielectric_poisson`` -- variable-permittivity Poisson problem.
"""
import math
import sys

import numpy as np

from . import abi

TE_1EV = abi.EV  # 1 eV in K (Source/UnivConstants.H: eV = ECHARGE/K_B)


def _mech(species, charges, molwt, mob, diff_kind, bg):
    m = abi.vd_mech_desc()
    ns = len(species)
    assert ns <= abi.VD_MAX_SPECIES
    m.nspec = ns
    for s, name in enumerate(species):
        m.names[s].value = name.encode()
        m.charge[s] = int(charges[s])
        m.molwt[s] = float(molwt[s])
        m.mob[s] = float(mob[s])
        m.diff_kind[s] = int(diff_kind[s])
        m.is_bg[s] = int(bg[s])
    m.e_idx = species.index("E")
    return m


def _set_reaction(m, r, A, beta, Ea, te, reactants, products, energy=0.0, rxntype=0,
                  eexci=0.0, elidx=0):
    """reactants/products: dict species -> stoichiometric coefficient."""
    R = m.reac[r]
    R.A, R.beta, R.Ea, R.temp_is_te = A, beta, Ea, int(te)
    rs = sorted(reactants.items())
    ps = sorted(products.items())
    assert len(rs) <= abi.VD_MAX_REACTANTS and len(ps) <= abi.VD_MAX_PRODUCTS
    R.nreact, R.nprod = len(rs), len(ps)
    for q, (sp, nu) in enumerate(rs):
        R.react_sp[q], R.react_nu[q] = sp, float(nu)
    for q, (sp, nu) in enumerate(ps):
        R.prod_sp[q], R.prod_nu[q] = sp, float(nu)
    R.energy, R.rxntype, R.eexci, R.elidx = energy, rxntype, eexci, elidx


def default_params():
    """Source/Vidyut.H:149-272 defaults."""
    p = abi.vd_params()
    p.gas_temperature, p.gas_pressure = 300.0, 101325.0
    p.hyp_order, p.weno_scheme = 1, 3
    p.elecenergy_solve, p.do_reactions, p.do_transport = 0, 1, 1
    p.do_spacechrg, p.do_bg_reactions, p.bound_specden = 1, 0, 0
    p.min_species_density, p.min_electron_density, p.min_electron_temp = 1e-20, 1e12, 300.0
    p.floor_jh, p.const_ele_trans = 0, 0
    p.ele_mob, p.ele_diff = -1.0e24, 1.0e24
    p.voltage_profile = 0
    p.voltage_amp_1 = p.voltage_amp_2 = 0.0
    p.voltage_freq, p.voltage_dur, p.voltage_center = 1.0, 0.0, 0.0
    p.linsolve_reltol = p.linsolve_abstol = 1e-10
    p.linsolve_maxiter, p.linsolve_max_coarsening_level = 100, 10
    p.mg_pre_smooth = p.mg_post_smooth = 2
    p.mg_precision = 0
    p.adaptive_dt = 0
    p.advective_cfl, p.diffusive_cfl, p.dielectric_cfl = 0.7, 1.0, 1.0
    p.dt_min, p.dt_max, p.adaptive_dt_delay = 1e-14, 1e-9, 0.0
    p.dt_stretch = sys.float_info.max
    for name in ("pot", "eden", "eenrg", "ion", "neutral"):
        getattr(p, name + "_bc_lo")[:] = [abi.ROBINBC] * 3
        getattr(p, name + "_bc_hi")[:] = [abi.ROBINBC] * 3
    return p


class Case:
    """A runnable discharge case: mechanism, vidyut.* params, geometry,
    initial state (NVAR, nz, ny, nx) and the fixed time step."""

    def __init__(self, name, mech, params, geom, state, dt, nsteps):
        self.name, self.mech, self.params, self.geom = name, mech, params, geom
        self.state, self.dt, self.nsteps = state, dt, nsteps

    @property
    def ncells(self):
        return int(np.prod(self.state.shape[1:]))

    @property
    def box(self):
        return abi.domain_box(self.geom)


# --------------------------------------------------------------- 2-D case
def discharge_2d_mech():
    """E, ION+, NEUTRAL (background); CKWC form with one electron-impact
    ionisation NEUTRAL + E => ION + 2E, energy loss 24 eV and the
    elastic loss term of test/models/ionelec_DC_1d/Chemistry.H:321-341."""
    m = _mech(["E", "ION+", "NEUTRAL"], [-1, 1, 0], [0.000549, 4.002602, 4.002602],
              [-0.03, 2e-4, 0.0],
              [abi.VD_DIFF_EINSTEIN_TE, abi.VD_DIFF_EINSTEIN_TG, abi.VD_DIFF_EINSTEIN_TG],
              [0, 0, 1])
    m.chem_form = abi.VD_CHEM_CKWC
    m.elastic_loss = 1
    m.elastic_partner_amu = 4.0
    m.nreac = 1
    _set_reaction(m, 0, 2.5e-12, 0.0, 2.5e5, True, {0: 1, 2: 1}, {0: 2, 1: 1},
                  rxntype=2, eexci=24.0, elidx=0)
    return m


def discharge_2d(n=128, seed=0, nsteps=100, dt=1e-11):
    """2-D 1 cm x 1 cm, FP64, Gaussian n_e = n_i (peak 1e16 m^-3, 1 mm)."""
    m = discharge_2d_mech()
    p = default_params()
    p.hyp_order, p.weno_scheme = 5, 3
    p.elecenergy_solve = 1
    p.bound_specden = 1
    p.min_species_density, p.min_electron_density = 1e10, 1e10
    p.gas_temperature = 300.0
    p.gas_pressure = 2.4e25 * abi.K_B * 300.0
    p.voltage_amp_1, p.voltage_amp_2 = 0.0, 1000.0
    p.pot_bc_lo[:] = [abi.DIRCBC, abi.PERBC, abi.PERBC]
    p.pot_bc_hi[:] = [abi.DIRCBC, abi.PERBC, abi.PERBC]
    for nm in ("eden", "eenrg", "ion"):
        getattr(p, nm + "_bc_lo")[:] = [abi.IHNEUBC, abi.PERBC, abi.PERBC]
        getattr(p, nm + "_bc_hi")[:] = [abi.IHNEUBC, abi.PERBC, abi.PERBC]
    p.neutral_bc_lo[:] = [abi.HNEUBC, abi.PERBC, abi.PERBC]
    p.neutral_bc_hi[:] = [abi.HNEUBC, abi.PERBC, abi.PERBC]
    L = 0.01
    g = abi.make_geom([n, n], [0.0, 0.0], [L, L], [0, 1])
    ns = 3
    c = abi.comp(ns)
    st = np.zeros((c["NVAR"], 1, n, n))
    x = (np.arange(n) + 0.5) * g.dx[0]
    y = (np.arange(n) + 0.5) * g.dx[1]
    X, Y = np.meshgrid(x, y, indexing="xy")
    r2 = (X - 0.5 * L) ** 2 + (Y - 0.5 * L) ** 2
    sig = 1e-3
    noise = np.random.default_rng(seed).random((n, n))
    ne = 1e16 * np.exp(-r2 / (2.0 * sig * sig)) * (1.0 + 0.01 * noise)
    st[0, 0] = ne
    st[1, 0] = ne
    st[2, 0] = 2.4e25
    st[c["ETEMP"], 0] = TE_1EV
    st[c["EEN"], 0] = 1.5 * ne * abi.K_B * TE_1EV
    return Case("discharge_2d_%d" % n, m, p, g, st, dt, nsteps)


# ---------------------------------------------------------- air-like 3-D
AIR_SPECIES = ["E", "N2+", "O2+", "O+", "O2-", "N2", "O2", "O", "N", "NO"]
AIR_CHARGE = [-1, 1, 1, 1, -1, 0, 0, 0, 0, 0]
AIR_MW = [0.000549, 28.014, 31.998, 15.999, 31.999, 28.014, 31.998, 15.999, 14.007, 30.006]


def _draw_arrhenius(rng, te):
    A = 10.0 ** rng.uniform(-20.0, -14.0)
    beta = rng.uniform(-1.0, 1.0)
    Ea = rng.uniform(0.0, 1e5)
    Tref = TE_1EV if te else 300.0
    # k = A*(T/Tref)^beta*exp(-Ea/T) folded into the A*T^beta*exp(-Ea/T) form
    return A * Tref ** (-beta), beta, Ea


def _rate(A, beta, Ea, T):
    return A * math.exp(beta * math.log(T) - Ea / T)


def synthetic_mechanism(species, charges, molwt, nreac, frac_electron, seed, ion_cap=None,
                        neutral_density=2.4e25):
    """Seeded table-form mechanism (Source/reaction_source.H semantics).

    Charge-conserving templates: electron-impact ionisation, attachment,
    excitation and dissociative recombination on Te; charge transfer,
    detachment, ion-ion recombination and neutral exchange on Tg.
    A ~ logU[1e-20, 1e-14] m^3/s, beta ~ U[-1, 1], Ea/k ~ U[0, 1e5] K.
    ion_cap = (dt, Te0): ionisation draws whose rate at Te0 would grow the
    electron density by more than 1e-3 per step of dt are drawn again, so
    explicit chemistry stays bounded over a long run."""
    rng = np.random.default_rng(seed)
    ns = len(species)
    pos = [s for s in range(ns) if charges[s] > 0]
    negi = [s for s in range(ns) if charges[s] < 0 and species[s] != "E"]
    neu = [s for s in range(ns) if charges[s] == 0]
    E = species.index("E")
    mob = [(-0.03 if s == E else (2e-4 * np.sign(charges[s]) if charges[s] else 0.0))
           for s in range(ns)]
    dk = [abi.VD_DIFF_EINSTEIN_TE if s == E else abi.VD_DIFF_EINSTEIN_TG for s in range(ns)]
    m = _mech(species, charges, molwt, mob, dk, [1 if charges[s] == 0 else 0 for s in range(ns)])
    m.chem_form = abi.VD_CHEM_TABLE
    m.nreac = nreac
    n_e = int(round(frac_electron * nreac))
    eV = abi.ECHARGE
    for r in range(nreac):
        if r < n_e:
            kind = rng.integers(0, 4)
            while True:
                A, beta, Ea = _draw_arrhenius(rng, True)
                if kind != 0 or ion_cap is None:
                    break
                dt, te0 = ion_cap
                if _rate(A, beta, Ea, te0) * neutral_density * dt <= 1e-3:
                    break
            if kind == 0:    # ionisation  e + M -> M+ + 2e
                M, P = rng.choice(neu), rng.choice(pos)
                _set_reaction(m, r, A, beta, Ea, True, {E: 1, M: 1}, {E: 2, P: 1},
                              energy=-rng.uniform(10.0, 20.0) * eV)
            elif kind == 1 and negi:  # attachment e + M -> X-
                M, X = rng.choice(neu), rng.choice(negi)
                _set_reaction(m, r, A, beta, Ea, True, {E: 1, M: 1}, {X: 1},
                              energy=-rng.uniform(0.0, 2.0) * eV)
            elif kind == 2:  # dissociative recombination e + P+ -> M + N
                P = rng.choice(pos)
                M, N = rng.choice(neu), rng.choice(neu)
                prods = {M: 1} if M == N else {M: 1, N: 1}
                if M == N:
                    prods = {M: 2}
                _set_reaction(m, r, A, beta, Ea, True, {E: 1, P: 1}, prods,
                              energy=-rng.uniform(0.0, 1.0) * eV)
            else:            # excitation e + M -> e + M*
                M = rng.choice(neu)
                _set_reaction(m, r, A, beta, Ea, True, {E: 1, M: 1}, {E: 1, M: 1},
                              energy=-rng.uniform(1.0, 12.0) * eV)
        else:
            A, beta, Ea = _draw_arrhenius(rng, False)
            kind = rng.integers(0, 4)
            if kind == 0 and len(pos) > 1:  # charge transfer P+ + M -> Q+ + N
                P, Q = rng.choice(pos, 2, replace=False)
                M, N = rng.choice(neu, 2, replace=False)
                _set_reaction(m, r, A, beta, Ea, False, {P: 1, M: 1}, {Q: 1, N: 1})
            elif kind == 1 and negi:        # detachment X- + M -> e + N + M
                X = rng.choice(negi)
                M, N = rng.choice(neu, 2, replace=False)
                _set_reaction(m, r, A, beta, Ea, False, {X: 1, M: 1}, {E: 1, N: 1, M: 1})
            elif kind == 2 and negi:        # ion-ion recombination X- + P+ -> M + N
                X, P = rng.choice(negi), rng.choice(pos)
                M, N = rng.choice(neu, 2, replace=False)
                _set_reaction(m, r, A, beta, Ea, False, {X: 1, P: 1}, {M: 1, N: 1})
            else:                           # neutral exchange M + N -> O + Q
                M, N, O, Q = rng.choice(neu, 4, replace=True)
                re = {M: 1} if M != N else {M: 2}
                if M != N:
                    re[N] = 1
                pr = {O: 1} if O != Q else {O: 2}
                if O != Q:
                    pr[Q] = 1
                _set_reaction(m, r, A, beta, Ea, False, re, pr)
    return m


def air_3d(n=256, seed=1, nsteps=200, dt=1e-12, L=0.1, voltage=1.0e4, state_seed=None):
    """3-D air-like discharge: 10 species (E, 3 positive ions, O2-, 5
    background neutrals at 2.4e25 m^-3), 30 seeded reactions, charged
    densities log-normal (median 1e15 m^-3, sigma_ln = 1), Te = 1 eV,
    Dirichlet 0 / 10 kV on the z faces, homogeneous Neumann on x and y,
    WENO-Z transport, L = 10 cm cube, dt = 1e-12 s (Source/Vidyut.H
    default vidyut.dt).  `seed` draws the mechanism and, unless
    `state_seed` is given, the initial state."""
    m = synthetic_mechanism(AIR_SPECIES, AIR_CHARGE, AIR_MW, 30, 0.5, seed,
                            ion_cap=(dt, TE_1EV))
    p = default_params()
    p.hyp_order, p.weno_scheme = 5, 3
    p.elecenergy_solve = 1
    p.bound_specden = 1
    p.min_species_density, p.min_electron_density = 1e10, 1e10
    p.gas_temperature = 300.0
    p.gas_pressure = 2.4e25 * abi.K_B * 300.0
    p.voltage_amp_1, p.voltage_amp_2 = 0.0, voltage
    p.pot_bc_lo[:] = [abi.HNEUBC, abi.HNEUBC, abi.DIRCBC]
    p.pot_bc_hi[:] = [abi.HNEUBC, abi.HNEUBC, abi.DIRCBC]
    for nm in ("eden", "eenrg", "ion"):
        getattr(p, nm + "_bc_lo")[:] = [abi.HNEUBC, abi.HNEUBC, abi.IHNEUBC]
        getattr(p, nm + "_bc_hi")[:] = [abi.HNEUBC, abi.HNEUBC, abi.IHNEUBC]
    p.neutral_bc_lo[:] = [abi.HNEUBC] * 3
    p.neutral_bc_hi[:] = [abi.HNEUBC] * 3
    nn = [n] * 3 if np.isscalar(n) else list(n)
    g = abi.make_geom(nn, [0.0] * 3, [L] * 3, [0, 0, 0])
    ns = len(AIR_SPECIES)
    c = abi.comp(ns)
    rng = np.random.default_rng((seed if state_seed is None else state_seed) + 1000)
    st = np.zeros((c["NVAR"], nn[2], nn[1], nn[0]))
    for s in range(ns):
        if AIR_CHARGE[s] != 0:
            st[s] = 1e15 * np.exp(rng.standard_normal((nn[2], nn[1], nn[0])))
        else:
            st[s] = 2.4e25
    st[c["ETEMP"]] = TE_1EV
    st[c["EEN"]] = 1.5 * st[0] * abi.K_B * TE_1EV
    return Case("air_3d_%s" % "x".join(map(str, nn)), m, p, g, st, dt, nsteps)


# ------------------------------------------------------- 40 / 200 chemistry
def co2_species():
    pos = ["CO2+", "CO+", "O2+", "O+", "C+", "C2O2+", "C2O3+", "C2O4+", "Ar+", "Ar2+", "H2+",
           "H+", "H3+"]
    neg = ["O-", "O2-", "CO3-"]
    neu = ["CO2", "CO", "O2", "O", "C", "O3", "C2O", "Ar", "Ar*", "H2", "H", "H2O", "OH",
           "HO2", "H2O2", "CH4", "CH3", "CH2O", "CHO", "C2H2", "CO2v1", "CO2v2", "COv1"]
    species = ["E"] + pos + neg + neu
    assert len(species) == 40
    charges = [-1] + [1] * len(pos) + [-1] * len(neg) + [0] * len(neu)
    molwt = [0.000549] + [30.0] * (len(species) - 1)
    return species, charges, molwt


def co2_mechanism(seed=3):
    """40 species / 200 reactions, 20 % electron-impact on Te."""
    sp, ch, mw = co2_species()
    return synthetic_mechanism(sp, ch, mw, 200, 0.2, seed)


def co2_state(n, seed=4):
    """Log-normal densities for every species (median 1e15 m^-3 for charged,
    1e22 for neutrals, sigma_ln = 1) and Te ~ U[1e4, 1e5] K per cell."""
    sp, ch, _ = co2_species()
    ns = len(sp)
    c = abi.comp(ns)
    nn = [n] * 3 if np.isscalar(n) else list(n)
    rng = np.random.default_rng(seed)
    st = np.zeros((c["NVAR"], nn[2], nn[1], nn[0]))
    for s in range(ns):
        med = 1e15 if ch[s] != 0 else 1e22
        st[s] = med * np.exp(rng.standard_normal((nn[2], nn[1], nn[0])))
    st[c["ETEMP"]] = rng.uniform(1e4, 1e5, (nn[2], nn[1], nn[0]))
    st[c["EEN"]] = 1.5 * st[0] * abi.K_B * st[c["ETEMP"]]
    return st


# ----------------------------------------------------- dielectric Poisson
def dielectric_poisson(n=512, seed=5, eps_slab=4.0):
    """-div(eps_r grad phi) = rho/eps0 on the unit cube, eps_r = 4 for
    z in [0.4, 0.6] else 1, harmonic face eps, Dirichlet 0 on z, periodic
    x and y, rho = e*1e15*(U[-1,1] - mean).  Returns (geom, face eps list
    of (nz,ny,nx)-shaped face arrays, rhs of the reference sign
    convention alpha=0, beta=1, b=-eps_r: rhs = -rho/eps0)."""
    nn = [n] * 3 if np.isscalar(n) else list(n)
    g = abi.make_geom(nn, [0.0] * 3, [1.0] * 3, [1, 1, 0])
    z = (np.arange(nn[2]) + 0.5) * g.dx[2]
    epsc = np.where((z >= 0.4) & (z <= 0.6), eps_slab, 1.0)
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1.0, 1.0, (nn[2], nn[1], nn[0]))
    rho = abi.ECHARGE * 1e15 * (u - u.mean())
    rhs = -rho / abi.EPS0
    faces = []
    for d in range(3):
        shp = [nn[2], nn[1], nn[0]]
        shp[2 - d] += 1
        f = np.ones(shp)
        if d == 2:
            ez = np.concatenate([[epsc[0]], epsc, [epsc[-1]]])
            a, b = ez[:-1], ez[1:]
            fz = 2.0 * a * b / (a + b)
            f = -np.broadcast_to(fz[:, None, None], shp).copy()
        else:
            f = -np.broadcast_to(epsc[:, None, None], shp).copy()
        faces.append(np.ascontiguousarray(f))
    return g, faces, rhs


# ------------------------------------------- verification cases (plug-ins)
def drift_diffuse(n=32, dim=2, dt=0.1, nsteps=40):
    """test/verification/driftDiffuse: species E, S1, S2 on x in [0, 1]
    (periodic across, 0.125 wide).  The potential is 1 V at x = 0 and 0 V at
    x = 1 (Robin a = 1, b = 0), so E_x = 1 V/m; S1 drifts with velocity E
    (the case's user_transport::compute_vel, here a fitted mobility of 1)
    and diffuses with D = 0.1, Dirichlet 0 / 1 at the ends; S2 diffuses with
    D = 1 under a unit source (user_sources::add_user_react_sources),
    Dirichlet 0 at both ends.  Steady states (verify_spec.py:28,42):
    S1 = (exp(10 x) - 1) / (exp(10) - 1), S2 = (x - x^2) / 2.  Upwind
    fluxes and dt = 0.1 as in inputs_x."""
    m = _mech(["E", "S1", "S2"], [-1, 0, 0], [0.000549, 1.0, 1.0], [0.0, 0.0, 0.0],
              [abi.VD_DIFF_CONST] * 3, [0, 0, 0])
    m.diff[0], m.diff[1], m.diff[2] = 1.0, 0.1, 1.0
    m.chem_form = abi.VD_CHEM_TABLE
    m.nreac = 0
    abi.set_fit(m.mob_fit[1], 1.0, c0=0.0)  # mu = 1: velocity = E
    for s, fhi in ((1, 1.0), (2, 0.0)):
        m.robin_set[s] = 1
        for q, v in enumerate((1.0, 0.0, 0.0)):
            m.robin[s][0][0][q] = v
        for q, v in enumerate((1.0, 0.0, fhi)):
            m.robin[s][0][1][q] = v
    m.user_src[2] = 1.0
    p = default_params()
    p.hyp_order = 1
    p.elecenergy_solve = 0
    p.voltage_amp_1, p.voltage_amp_2 = 1.0, 0.0
    per = [abi.PERBC] * 3
    p.pot_bc_lo[:] = p.pot_bc_hi[:] = [abi.DIRCBC] + per[1:]
    for nm in ("eden", "eenrg", "ion"):
        getattr(p, nm + "_bc_lo")[:] = [abi.HNEUBC] + per[1:]
        getattr(p, nm + "_bc_hi")[:] = [abi.HNEUBC] + per[1:]
    p.neutral_bc_lo[:] = p.neutral_bc_hi[:] = [abi.ROBINBC] + per[1:]
    nn = [n, 2, 2][:dim]
    g = abi.make_geom(nn, [0.0] * dim, [1.0, 0.125, 0.125][:dim], [0] + [1] * (dim - 1))
    c = abi.comp(3)
    shp = (c["NVAR"],) + ((nn[2],) if dim == 3 else (1,)) + (nn[1], nn[0])
    st = np.zeros(shp)
    st[c["ETEMP"]] = 0.5 * TE_1EV
    return Case("drift_diffuse_%d" % n, m, p, g, st, dt, nsteps)


def drift_diffuse_exact(x):
    """verify_spec.py:28,42: the S1 and S2 steady states."""
    return (np.exp(10.0 * x) - 1.0) / (np.exp(10.0) - 1.0), 0.5 * (x - x * x)


def advect_1d(n=20, dt=1e-4, nsteps=20000, weno_scheme=3):
    """test/verification/Advect (inputs1d, Prob.H one_dim): species E, AR on
    x in [-1, 1] x [0, 0.4], periodic; AR moves with the prescribed velocity
    (1, 0) (user_transport::compute_vel), no diffusion; WENO fluxes with two
    timestep correctors.  AR(x, 0) = sin(pi x - sin(pi x) / pi); after t = 2
    it returns to its start (get_error_1d.py:52)."""
    m = _mech(["E", "AR"], [-1, 0], [0.000549, 39.948], [-1.0, -1.0], [abi.VD_DIFF_CONST] * 2, [0, 0])
    m.diff[0] = m.diff[1] = 0.0
    m.chem_form = abi.VD_CHEM_TABLE
    m.nreac = 0
    m.vel_set[1] = 1
    m.vel[1][0], m.vel[1][1], m.vel[1][2] = 1.0, 0.0, 0.0
    p = default_params()
    p.hyp_order, p.weno_scheme = 5, weno_scheme
    p.elecenergy_solve = 0
    for nm in ("pot", "eden", "eenrg", "ion", "neutral"):
        getattr(p, nm + "_bc_lo")[:] = [abi.PERBC] * 3
        getattr(p, nm + "_bc_hi")[:] = [abi.PERBC] * 3
    ny = max(2, n // 5)
    g = abi.make_geom([n, ny], [-1.0, 0.0], [1.0, 2.0 * ny / n], [1, 1])  # square cells
    c = abi.comp(2)
    st = np.zeros((c["NVAR"], 1, ny, n))
    x = -1.0 + (np.arange(n) + 0.5) * g.dx[0]
    st[1] = np.sin(np.pi * x - np.sin(np.pi * x) / np.pi)[None, :]
    st[c["ETEMP"]] = 300.0
    return Case("advect_%d" % n, m, p, g, st, dt, nsteps)


# ------------------------------------- generated mechanisms (VD_CHEM_GEN)
def generated_mech(name, mob_e=-0.04, mob_ion=2e-4, bg_neutrals=True, keep=None):
    """A mechanism converted from a reference case's generated Chemistry.H
    (vidyut3d_amd/mechanisms/<name>.json) with constant transport: the
    electron mobility mob_e (Einstein diffusion at Te), ions +-mob_ion
    (Einstein at Tg), neutrals as background species (updated by their
    reactions only, vidyut.bg_species_ids).  `keep(reaction)` selects a
    subset of the reactions."""
    from . import mechgen
    mj = mechgen.load(mechgen.builtin(name))
    if keep is not None:
        mj["reactions"] = [r for r in mj["reactions"] if keep(r)]
    m = mechgen.to_desc(mj)
    for s in range(m.nspec):
        q = m.charge[s]
        if s == m.e_idx:
            m.mob[s], m.diff_kind[s] = mob_e, abi.VD_DIFF_EINSTEIN_TE
        else:
            m.mob[s] = mob_ion * np.sign(q) if q else 0.0
            m.diff_kind[s] = abi.VD_DIFF_EINSTEIN_TG
        m.is_bg[s] = int(bg_neutrals and q == 0)
    return m


def co2_h2_ar_3d(n=32, nsteps=20, dt=1e-12, L=1e-2, voltage=2000.0, seed=5):
    """3-D discharge in the Ar / H2 / CO2 mixture of
    test/Reactor/CO2_H2_Ar_mechanism_reduced (0.9 / 0.075 / 0.025 of
    2.446e25 m^-3, inputs1d and Prob.H) with that case's converted
    mechanism (44 species, 391 reactions): electrons log-normal (median
    1e16 m^-3, sigma_ln = 1) neutralised by Ar+, Te = 2 eV, Tg = 300 K,
    Dirichlet 0 / `voltage` on z, WENO-Z transport, neutrals background
    species with background reactions on (vidyut.do_bg_reactions = 1)."""
    m = generated_mech("co2_h2_ar_reduced")
    names = [m.names[s].value.decode() for s in range(m.nspec)]
    p = default_params()
    p.hyp_order, p.weno_scheme = 5, 3
    p.elecenergy_solve = 1
    p.do_bg_reactions = 1
    p.bound_specden = 1
    p.min_species_density, p.min_electron_density = 0.0, 1e9
    p.gas_temperature = 300.0
    p.gas_pressure = 101325.0
    p.voltage_amp_1, p.voltage_amp_2 = 0.0, voltage
    p.pot_bc_lo[:] = [abi.HNEUBC, abi.HNEUBC, abi.DIRCBC]
    p.pot_bc_hi[:] = [abi.HNEUBC, abi.HNEUBC, abi.DIRCBC]
    for nm in ("eden", "eenrg", "ion"):
        getattr(p, nm + "_bc_lo")[:] = [abi.HNEUBC, abi.HNEUBC, abi.IHNEUBC]
        getattr(p, nm + "_bc_hi")[:] = [abi.HNEUBC, abi.HNEUBC, abi.IHNEUBC]
    p.neutral_bc_lo[:] = [abi.HNEUBC] * 3
    p.neutral_bc_hi[:] = [abi.HNEUBC] * 3
    nn = [n] * 3 if np.isscalar(n) else list(n)
    g = abi.make_geom(nn, [0.0] * 3, [L] * 3, [0, 0, 0])
    c = abi.comp(m.nspec)
    rng = np.random.default_rng(seed)
    st = np.zeros((c["NVAR"], nn[2], nn[1], nn[0]))
    N = 2.446e25
    st[names.index("AR")] = 0.9 * N
    st[names.index("H2")] = 0.075 * N
    st[names.index("CO2")] = 0.025 * N
    ne = 1e16 * np.exp(rng.standard_normal((nn[2], nn[1], nn[0])))
    st[m.e_idx] = ne
    st[names.index("AR+")] = ne
    te = 2.0 * TE_1EV
    st[c["ETEMP"]] = te
    st[c["EEN"]] = 1.5 * ne * abi.K_B * te
    return Case("co2_h2_ar_3d_%s" % "x".join(map(str, nn)), m, p, g, st, dt, nsteps)
