"""Same-shape batching: cases stepped together in one batched context
(vd_ctx_create_batched) get exactly the states they get alone, bit for
bit, also when a case joins the batch late and another leaves it early;
the batched states also match the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from vidyut3d_amd import compare, native, serve

pytestmark = pytest.mark.gpu


def _cases(dim, n, prec, nsteps):
    reqs = [serve.Request(0, 0.0, dim, n, 800.0, nsteps, 21), serve.Request(1, 0.0, dim, n, 5200.0, nsteps, 22),
            serve.Request(2, 0.0, dim, n, 9100.0, nsteps, 23)]
    return reqs, [r.make_case(prec) for r in reqs]


def _alone(case, nsteps):
    dev = native.Device(0)
    try:
        sim = native.Sim(dev, case)
        sim.evolve(nsteps)
        return sim.get_state()
    finally:
        dev.close()


@pytest.mark.parametrize("dim,n,prec", [(2, 32, 0), (2, 64, 1), (3, 32, 0), (3, 32, 1), (3, 64, 1)])
def test_batched_cases_equal_their_runs_alone(dim, n, prec):
    _batched_vs_alone(dim, n, prec)


@pytest.mark.parametrize("prec", [0, 1])
def test_batched_fused_sweeps_equal_their_runs_alone(prec, monkeypatch):
    """The fused finest-level passes at 64^3 (16-plane chunks, which the
    default chooses from 192^3 up): the species solves' first sweep launched
    with the rhs pass, and the sweeps' copy-back into the state, per case."""
    monkeypatch.setenv("VD_FUSED_ZC", "16")
    monkeypatch.setattr(serve, "_REGION_BYTES", {})  # the fused passes' second solution buffer
    _batched_vs_alone(3, 64, prec)


def _batched_vs_alone(dim, n, prec):
    reqs, cs = _cases(dim, n, prec, 8)
    print("batch: %dD %d^%d, mg_precision %d, 3 cases" % (dim, n, dim, prec), flush=True)
    nbytes = serve.case_region_bytes(0, cs[0])
    dev = native.BatchDevice(0, 3, nbytes)
    try:
        sim = native.Sim(dev, cs[0])
        for c, case in enumerate(cs):
            sim.set_state_case(c, case.state)
            dev.set_case_voltage(c, case.params.voltage_amp_1, case.params.voltage_amp_2)
        # case 2 joins after 3 steps, case 0 leaves after 6: 6, 8 and 5 steps
        dev.set_active([0, 1])
        sim.evolve(3)
        dev.set_active([0, 1, 2])
        sim.evolve(2)
        sim.evolve(1)
        dev.set_active([1, 2])
        sim.evolve(2)
        got = [sim.get_state_case(c) for c in range(3)]
        infos = [dev.case_solve_info(c).iterations for c in range(3)]
    finally:
        dev.close()
    steps = [6, 8, 5]
    for c, case in enumerate(cs):
        alone = _alone(case, steps[c])
        assert np.isfinite(got[c]).all()
        assert np.array_equal(alone, got[c]), "case %d differs from its run alone (max |d| %.3e)" % (
            c, np.max(np.abs(alone - got[c])))
    assert min(infos) >= 0
    # the FP64 batch against the oracle
    if prec == 0:
        ref, _ = O.evolve(cs[1], 8)
        errs = compare.state_errors(ref, got[1], cs[1].mech.nspec, cs[1].geom.dim)
        assert max(errs.values()) <= 1e-10, errs


def test_batched_context_rejects_what_it_does_not_batch():
    reqs, cs = _cases(2, 32, 0, 2)
    nbytes = serve.case_region_bytes(0, cs[0])
    dev = native.BatchDevice(0, 2, nbytes)
    try:
        with pytest.raises(native.VidyutError):
            dev.set_active([0, 0])
        with pytest.raises(native.VidyutError):
            dev.set_active([2])
        c = cs[0]
        c.params.adaptive_dt = 1
        with pytest.raises(native.VidyutError):
            native.Sim(dev, c)
        c.params.adaptive_dt = 0
        c.params.voltage_profile = 1
        with pytest.raises(native.VidyutError):
            native.Sim(dev, c)
    finally:
        dev.close()
