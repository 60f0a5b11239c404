"""Generated mechanisms (VD_CHEM_GEN) on the GPU against the oracle: the
source terms of the converted test/Reactor/CO2_H2_Ar_mechanism_reduced
(44 species, 391 reactions: Janev and 1/Te fits, Te power laws, gas-T
Arrhenius rates, third bodies, reverse rates from g/RT) and
test/verification/He_RF_1d mechanisms, through the mechanism-specialised
kernel and the generic one, and whole steps of a 3-D discharge with the
44-species mechanism.

Tolerances: 1e-13 where every rate is well conditioned (He_RF_1d, and the
CO2 mechanism without its fitted rates); with the Janev fits, the rounding
bound of their coefficient sums (fit_conditioning, 4.5e-10 here): GPU and
CPU pow / exp differ in the last place, and those sums amplify it."""
import numpy as np
import pytest

from oracle import oracle as O
from vidyut3d_amd import abi, cases, compare, mechgen, native

pytestmark = pytest.mark.gpu


def _state(m, n, seed):
    """Log-normal densities (charged median 1e16, neutrals 1e21, the three
    bath gases of the CO2 case at their mixture), Te ~ U[TE_EV] eV."""
    rng = np.random.default_rng(seed)
    ns = m.nspec
    cc = abi.comp(ns)
    st = np.zeros((cc["NVAR"], n, n, n))
    for s in range(ns):
        med = 1e16 if m.charge[s] != 0 else 1e21
        st[s] = med * np.exp(rng.standard_normal((n, n, n)))
    names = [m.names[s].value.decode() for s in range(ns)]
    for nm, f in (("AR", 0.9), ("H2", 0.075), ("CO2", 0.025), ("HE", 1.0)):
        if nm in names:
            st[names.index(nm)] = f * 2.446e25
    st[cc["ETEMP"]] = rng.uniform(*TE_EV, (n, n, n)) * abi.EV
    st[cc["EEN"]] = 1.5 * st[m.e_idx] * abi.K_B * st[cc["ETEMP"]]
    return st


def _percomp(ref, got):
    """max over components of max|diff| / max|ref| of that component."""
    worst = 0.0
    for q in range(ref.shape[0]):
        s = np.max(np.abs(ref[q]))
        if s == 0.0:
            assert np.max(np.abs(got[q])) == 0.0, q
            continue
        worst = max(worst, float(np.max(np.abs(ref[q] - got[q])) / s))
    return worst


FITS = ("janev", "ffit", "janev_ffit")
TE_EV = (0.3, 8.0)  # the Te range of _state


def fit_conditioning(name):
    """Rounding bound of the fitted rates: a Janev rate is exp of
    sum_j c_j (ln TeV)^j whose coefficients reach 2e4 in magnitude, so two
    correct evaluations (another pow, exp or FMA use) may differ in k by
    eps * sum_j |c_j (ln TeV)^j| relative; its maximum over the Te range."""
    mj = mechgen.load(mechgen.builtin(name))
    te = np.linspace(*TE_EV, 4000)
    worst = 0.0
    for r in mj["reactions"]:
        if r["kind"] in ("janev", "janev_ffit"):
            lt = np.log(te)
            k = np.abs(np.array([r["fit"][j] * lt ** j for j in range(9)])).sum(0)
            if r["kind"] == "janev_ffit":
                k = np.where(te < r["fit_switch_ev"], k, 0.0)
            worst = max(worst, float(k.max()))
        elif r["kind"] in ("ffit",):
            it = 1.0 / te
            worst = max(worst, float(np.abs(np.array([r["fit"][j] * it ** (j + 1) for j in range(4)])).sum(0).max()))
    return np.finfo(float).eps * worst


# (mechanism, reactions kept, bound): without the fits every rate is a
# product of well-conditioned factors
CASES = [("co2_h2_ar_reduced", "all", None), ("co2_h2_ar_reduced", "no_fits", 1e-13), ("he_rf_1d", "all", 1e-13)]


@pytest.mark.parametrize("name,subset,bound", CASES)
@pytest.mark.parametrize("specialise", [1, 0])
@pytest.mark.parametrize("Tg", [300.0, 1200.0])
def test_generated_source_matches_oracle(dev, name, subset, bound, specialise, Tg):
    keep = (lambda r: r["kind"] not in FITS) if subset == "no_fits" else None
    m = cases.generated_mech(name, keep=keep)
    if bound is None:
        bound = fit_conditioning(name)
        assert 1e-11 < bound < 1e-9  # 4.5e-10 for the H2 Janev fits of this mechanism
    n = 16
    st = _state(m, n, 3)
    g = abi.make_geom([n] * 3, [0.0] * 3, [1e-2] * 3, [1, 1, 1])
    og, dg = O.Grid(g), native.Grid(dev, g)
    p = cases.default_params()
    p.gas_temperature = Tg
    sb = O.fill_patch(og, st, 1)
    ref = O.rxn_source(m, p, og, sb, 1)
    print("%s specialise=%d Tg=%g: generating the source kernel" % (name, specialise, Tg), flush=True)
    mech = native.Mech(dev, m)
    mech.specialise(bool(specialise))
    assert mech.specialised == bool(specialise)
    out = dg.array(m.nspec + 1)
    native.rxn_source(mech, p, dg, dg.array(st.shape[0], 1, host=sb), out)
    got = out.get()
    err = _percomp(ref, got)
    print("%s (%s) specialise=%d Tg=%g: worst per-component max error %.2e, bound %.1e"
          % (name, subset, specialise, Tg, err, bound), flush=True)
    assert err < bound


def test_generated_mechanism_steps_match_oracle(dev):
    c = cases.co2_h2_ar_3d(n=16, nsteps=4)
    ref, _ = O.evolve(c, 4)
    sim = native.Sim(dev, c)
    sim.evolve(4)
    got = sim.get_state()
    sim.close()
    errs = compare.state_errors(ref, got, c.mech.nspec, 3)
    names = [c.mech.names[s].value.decode() for s in range(c.mech.nspec)] + abi.PLASMA_VAR_NAMES
    worst = max(errs, key=errs.get)
    print("44-species steps: worst field %s %.2e" % (names[worst], errs[worst]), flush=True)
    bad = {names[q]: e for q, e in errs.items() if not e <= 1e-10}
    assert not bad, bad
