"""Benchmark: cell-steps/s of the Vidyut per-timestep hot path on one MI355X.

Headline workload (the `value` of the JSON line): 3-D 256^3 single-level
discharge, FP64, 10 species / 30 seeded reactions (air-like), WENO-Z
transport of electrons, electron energy and 4 ions with implicit
multigrid diffusion, chemistry, Joule heating, the potential multigrid
solve and the CFL reduction every step, for the configuration's 200 steps
(after 2 warm-up steps).  One "step" = the body of Vidyut::Evolve's step
loop (Source/Evolve.cpp:33-369).

The same run also measures (under "configs"): the 2-D 128^2 discharge for
100 steps against the CPU oracle, the headline case with FP32 V-cycles in
FP64 iterative refinement, the 512^3 dielectric Poisson solve FP64 vs
mixed precision, one 40-species / 200-reaction source evaluation at 256^3
(FP64 FLOP/s), the 44-species mechanism converted from a reference case's
generated Chemistry.H, and the serving loop (p50 / p99 time-to-solution).

Prints ONE JSON line.  --gpus above 1 is refused: this build runs on one
GPU.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Canonical algorithmic bytes per launch (FP64 = 8 B per value, N = cells
# of the level the launch runs on), accumulated by the library per launch:
#  fused     the fused red-black smoother (two sweeps, or one sweep plus
#            the residual norm, in one pass): it must read x (unless the
#            correction starts from zero), rhs and the coefficient data
#            once and write x once, 8*N*(3 + n_b) (8*N*(2 + n_b) from zero)
#  gsrb      one colour of a red-black sweep on the finest level: a full
#            sweep reads phi, rhs and the coefficient data and writes phi,
#            8*N*(3 + n_b); one colour is half (n_b = 0 for the constant-
#            coefficient potential solve, 1 for the cell coefficient of the
#            implicit species solves, 3 for face arrays)
#  transport flux + divergence of one species, what the launches must
#            move: 8*N*4*dim with the divergence fused into the last
#            direction (u and E_d read and a face flux written per other
#            direction; u, E_d, those faces and rxn read and dsdt written by
#            the last), 8*N*(4*dim + 2) otherwise
#  chem      one source evaluation: 8*N*(2*NS + 3)
#  resid     finest residual + norm: 8*N*(4 + n_b)
#  mg_coarse the V-cycle legs below the finest level, per level of N_l
#            cells: (nu1 + nu2)*(3 + n_b) + (2 + n_b) + 2 + 1/4 arrays of N_l
#  species_solve  one implicit species solve: 8*N*(4 + n_b + iterations *
#            (3 + n_b)) (+ 3 arrays for the Te update of the energy)
#  pot_post  the potential's E faces, cell E and E/N: 8*N*(NS + 3 + 2*dim)
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print("bench:", *a, flush=True)


def pmc_traffic(kernel, n, precision):
    """HBM bytes per launch of the roofline kernel from the committed
    rocprofv3 --pmc summary of the same workload (FETCH_SIZE and
    WRITE_SIZE passes, scaled by a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "%s_%d_fp%d_pmc_traffic.md" % (kernel, n, 64 if precision == 0 else 32))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def run_steps(native, abi, dev, case, steps, warmup, families=None, state=True):
    """Times `steps` steps after `warmup` untimed ones; with `families`,
    HIP-event timing of those kernel families during the timed steps
    (returns {family: (ms, launches, canonical bytes)})."""
    sim = native.Sim(dev, case)
    sim.evolve(warmup)
    dev.sync()
    if families:
        dev.set_kernel_timing_mask(families)
    it0 = sim.mg_iterations
    t0 = time.perf_counter()
    sim.evolve(steps)
    dev.sync()
    wall = time.perf_counter() - t0
    fam = None
    if families:
        fam = {f: dev.kernel_timing_family(f) for f in families}
        dev.set_kernel_timing(abi.KT_NONE)
    st = sim.get_state() if state else None
    iters = sim.mg_iterations - it0
    sim.close()
    return wall, st, iters, fam


def family_rooflines(abi, fam, steps):
    """Per-family HIP-event time and canonical bytes -> GB/s and fraction of
    HBM peak."""
    out = {}
    for f, (ms, n, by) in fam.items():
        if n == 0:
            continue
        avg = ms / 1e3 / n
        gbs = by / n / avg / 1e9
        out[abi.KT_NAMES[f]] = {"ms_per_step": round(ms / steps, 3), "launches_per_step": round(n / steps, 2),
                                "avg_launch_us": round(avg * 1e6, 2), "bytes_per_launch": round(by / n),
                                "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline_steps(case_fn, sample_n, sample_steps):
    """The oracle (C, OpenMP over the host's cores) on `sample_steps` steps of
    the headline case at sample_n^3; returns the baseline and its state."""
    from oracle import oracle as O
    c = case_fn(sample_n)
    t0 = time.perf_counter()
    ref, _ = O.evolve(c, sample_steps)
    dt = time.perf_counter() - t0
    return {"value": c.ncells * sample_steps / dt, "unit": "cell-steps/s",
            "cores": int(O.lib().or_num_threads()), "kind": "port", "cpu_model": cpu_model(),
            "sample": "oracle (C, OpenMP) %d step(s) of the same case at %d^3 cells, %.1f s"
                      % (sample_steps, sample_n, dt)}, ref


def cpu_single_thread(case_fn, sample_n, sample_steps):
    """The same oracle on one thread, in a child process (OMP_NUM_THREADS=1)."""
    import subprocess
    code = ("import sys, time; sys.path.insert(0, %r)\n"
            "from oracle import oracle as O\nfrom vidyut3d_amd import cases\n"
            "c = cases.air_3d(n=%d)\nt = time.perf_counter(); O.evolve(c, %d)\n"
            "print(c.ncells * %d / (time.perf_counter() - t))\n" % (ROOT, sample_n, sample_steps, sample_steps))
    env = dict(os.environ, OMP_NUM_THREADS="1")
    try:
        out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
        v = float(out.stdout.strip().splitlines()[-1])
    except (subprocess.SubprocessError, ValueError, IndexError):
        return None
    return {"value": v, "unit": "cell-steps/s", "cores": 1,
            "sample": "oracle, 1 thread, %d step(s) of the same case at %d^3 cells" % (sample_steps, sample_n)}


# ------------------------------------------------------------ sub-benchmarks
def bench_batching(native, dev_index=0, steps=10, warmup=2, shapes=None):
    """Same-shape batching per shape bucket: ms per step of one case on a
    plain context, and of B cases of that shape (different seeds and
    voltages) stepped together in a batched context, in FP32 V-cycles as
    the serving loop runs them; speedup = B x (ms alone) / (ms batched)."""
    import time
    from vidyut3d_amd import serve
    out = {}
    for dim, n in shapes or ((2, 64), (2, 128), (2, 256), (3, 48), (3, 64), (3, 96), (3, 128)):
        base = serve.Request(0, 0.0, dim, n, 5000.0, steps, 7).make_case(1)
        d = native.Device(dev_index)
        sim = native.Sim(d, base)
        sim.evolve(warmup)
        d.sync()
        t0 = time.perf_counter()
        sim.evolve(steps)
        d.sync()
        alone = (time.perf_counter() - t0) / steps
        sim.close()
        d.close()
        row = {"cells": n ** dim, "alone_ms_per_step": round(alone * 1e3, 3)}
        region = serve.case_region_bytes(dev_index, base)
        for B in (4, 16):
            d = native.BatchDevice(dev_index, B, region)
            sim = native.Sim(d, base)
            for c in range(B):
                cc = serve.Request(c, 0.0, dim, n, 600.0 + 550.0 * c, steps, 100 + c).make_case(1)
                sim.set_state_case(c, cc.state)
                d.set_case_voltage(c, cc.params.voltage_amp_1, cc.params.voltage_amp_2)
            sim.evolve(warmup)
            d.sync()
            t0 = time.perf_counter()
            sim.evolve(steps)
            d.sync()
            tb = (time.perf_counter() - t0) / steps
            sim.close()
            d.close()
            row["batch%d_ms_per_step" % B] = round(tb * 1e3, 3)
            row["batch%d_cell_steps_per_s" % B] = B * n ** dim / tb
            row["batch%d_speedup" % B] = round(B * alone / tb, 2)
        out["%dD %d" % (dim, n)] = row
        log("batching %dD %d: alone %.3f ms/step, 4 cases %.3f (%.2fx), 16 cases %.3f (%.2fx)"
            % (dim, n, row["alone_ms_per_step"], row["batch4_ms_per_step"], row["batch4_speedup"],
               row["batch16_ms_per_step"], row["batch16_speedup"]))
    return out


def bench_discharge_2d(native, abi, cases, compare, dev, with_cpu):
    """Config 0: 2-D 128^2, 100 steps of dt = 1e-11 s, FP64, against the
    oracle run on this host (the CPU reference path)."""
    c = cases.discharge_2d(n=128, nsteps=100)
    wall, st, iters, _ = run_steps(native, abi, dev, c, 100, 0)
    out = {"cells": c.ncells, "steps": 100, "gpu_s": round(wall, 4),
           "cell_steps_per_s": c.ncells * 100 / wall, "mg_iterations": iters}
    if with_cpu:
        from oracle import oracle as O
        t0 = time.perf_counter()
        ref, _ = O.evolve(c, 100)
        tcpu = time.perf_counter() - t0
        errs = compare.state_errors(ref, st, c.mech.nspec, c.geom.dim)
        names = [c.mech.names[s].value.decode() for s in range(c.mech.nspec)] + abi.PLASMA_VAR_NAMES
        worst = max(errs, key=errs.get)
        rms = {names[q]: compare.field_errors(ref[q], st[q])["rel_rms_over_meanabs"]
               for q in range(ref.shape[0]) if np.max(np.abs(ref[q])) > 0}
        out.update({"cpu_s": round(tcpu, 3), "cpu_cell_steps_per_s": c.ncells * 100 / tcpu,
                    "cpu_cores": int(O.lib().or_num_threads()),
                    "max_rel_l2_vs_oracle": errs[worst], "worst_field": names[worst],
                    "max_rms_over_meanabs_vs_oracle": max(rms.values())})
    return out


def bench_dielectric(native, abi, cases, compare, dev, n, with_oracle):
    """Config 2: n^3 variable-permittivity Poisson, FP64 multigrid vs FP32
    V-cycles in FP64 iterative refinement, both to 1e-10 relative residual
    from a zero initial guess; with_oracle, both against the oracle's
    all-FP64 multigrid solve of the same problem."""
    import ctypes as C
    g, faces, rhs = cases.dielectric_poisson(n=n)
    oracle_sol = None
    if with_oracle:
        from oracle import oracle as O
        og = O.Grid(g)
        hf = [np.ascontiguousarray(faces[d][None]) for d in range(3)]
        hv = [og.view(hf[d], 0, d) for d in range(3)]
        p = abi.vd_mlmg_problem()
        p.alpha_a, p.beta = 0.0, 1.0
        for d in range(3):
            p.bface[d] = C.pointer(hv[d])
        p.bc_lo[:] = [0, 0, abi.DIRCBC]
        p.bc_hi[:] = [0, 0, abi.DIRCBC]
        p.reltol, p.abstol, p.maxiter = 1e-10, 0.0, 100
        p.pre_smooth = p.post_smooth = 2
        p.max_coarsening_level = 10
        t0 = time.perf_counter()
        oracle_sol, _ = O.mlmg_solve(og, p, og.zeros(1), np.ascontiguousarray(rhs[None]))
        log("oracle FP64 solve %.1f s" % (time.perf_counter() - t0))
        del hf, hv
    grid = native.Grid(dev, g)
    df = [grid.array(1, 0, d, host=np.ascontiguousarray(faces[d][None])) for d in range(3)]
    drhs = grid.array(1, host=rhs[None])
    del faces
    views = [f.view() for f in df]
    res = {}
    sols = {}
    for prec in (0, 1):
        p = abi.vd_mlmg_problem()
        p.alpha_a, p.beta = 0.0, 1.0
        for d in range(3):
            p.bface[d] = C.pointer(views[d])
        p.bc_lo[:] = [0, 0, abi.DIRCBC]
        p.bc_hi[:] = [0, 0, abi.DIRCBC]
        p.reltol, p.abstol, p.maxiter = 1e-10, 0.0, 100
        p.precision = prec
        p.pre_smooth = p.post_smooth = 2
        p.max_coarsening_level = 10
        phi = grid.array(1)
        native.mlmg_solve(grid, p, phi, drhs)  # warm-up (workspace allocation)
        best = None
        for _ in range(2):
            phi = grid.array(1)
            dev.sync()
            t0 = time.perf_counter()
            info = native.mlmg_solve(grid, p, phi, drhs)
            dev.sync()
            t = time.perf_counter() - t0
            best = t if best is None else min(best, t)
        sols[prec] = phi.get()
        res["fp64" if prec == 0 else "mixed"] = {"solve_s": round(best, 4),
                                                  "iterations": info.iterations,
                                                  "final_rel_residual": info.resnorm / max(info.resnorm0, 1e-300)}
    err = compare.field_errors(sols[0], sols[1])
    res.update({"cells": n ** 3, "speedup_mixed_over_fp64": res["fp64"]["solve_s"] / res["mixed"]["solve_s"],
                "mixed_rel_l2_vs_fp64": err["rel_l2"],
                "mixed_rms_over_meanabs_vs_fp64": err["rel_rms_over_meanabs"]})
    if oracle_sol is not None:
        for prec, key in ((0, "fp64"), (1, "mixed")):
            res[key]["rel_l2_vs_fp64_oracle"] = compare.field_errors(oracle_sol, sols[prec])["rel_l2"]
    return res


def chem_flops_per_cell(m):
    """Canonical FP64 flop count of one cell's source evaluation: per
    reaction 24 (rate: exp = 20 + 4) + 2*order + 2*(reactants + products),
    +6 per electron-impact reaction, plus 2 log + 2 div per cell (20 each)."""
    f = 80
    for r in range(m.nreac):
        R = m.reac[r]
        order = int(sum(R.react_nu[q] for q in range(R.nreact)))
        f += 24 + 2 * order + 2 * (R.nreact + R.nprod) + (6 if R.temp_is_te else 0)
    return f


def bench_chemistry(native, abi, cases, dev, n):
    """Config 3: one fused 40-species / 200-reaction source evaluation
    (species rates + electron-energy source) over n^3 cells."""
    m = cases.co2_mechanism(seed=3)
    blk = min(n, 128)
    st = cases.co2_state(blk, seed=4)
    reps = n // blk
    st = np.ascontiguousarray(np.tile(st, (1, reps, reps, reps)))
    g = abi.make_geom([n] * 3, [0] * 3, [1] * 3, [0] * 3)
    grid = native.Grid(dev, g)
    mech = native.Mech(dev, m)
    dst = grid.array(st.shape[0], host=st)
    del st
    out = grid.array(m.nspec + 1)
    prm = cases.default_params()
    native.rxn_source(mech, prm, grid, dst, out)
    dev.sync()
    dev.set_kernel_timing(abi.KT_CHEM)
    for _ in range(5):
        native.rxn_source(mech, prm, grid, dst, out)
    ms, launches, kbytes = dev.kernel_timing()
    dev.set_kernel_timing(abi.KT_NONE)
    t = ms / 1e3 / launches
    fpc = chem_flops_per_cell(m)
    tfl = fpc * n ** 3 / t / 1e12
    return {"cells": n ** 3, "nspec": m.nspec, "nreac": m.nreac, "launch_ms": round(t * 1e3, 3),
            "cells_per_s": n ** 3 / t, "flop_per_cell": fpc,
            "roofline": {"bound": "fp64", "achieved": round(tfl, 2), "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(tfl / FP64_PEAK_TFLOPS, 4),
                         "hbm_GBs": round(kbytes / launches / t / 1e9, 1)}}


def gen_flops_per_cell(m, abi):
    """Canonical FP64 flop count of one cell of a generated mechanism
    (VD_CHEM_GEN): per cell 13 pow + log + tanh + the mixture sum (20 per
    transcendental); per reaction the rate (gas-T Arrhenius: 0, taken from
    the per-launch table; Te power law 22; Te Arrhenius 44; Janev 40;
    1/Te fit 30), the forward product (order + 1, + 1 with a third body),
    the reverse rate (terms + 22 + its product), 1 for qf - qr, 2 per wdot
    update and 6 for an energy exchange."""
    f = 15 * 20 + m.nspec
    rate = {abi.VD_RATE_ARRH: 0, abi.VD_RATE_TEPOW: 22, abi.VD_RATE_TEARRH: 44,
            abi.VD_RATE_JANEV: 40, abi.VD_RATE_FFIT: 30, abi.VD_RATE_JANEV_FFIT: 40}
    for r in range(m.nreac):
        R = m.reac[r]
        order = int(sum(R.react_nu[q] for q in range(R.nreact)))
        f += rate[R.rate_kind] + order + 1 + (1 if R.third_body else 0) + 1 + 2 * R.nwd
        if R.reversible:
            f += R.ng + 22 + int(sum(R.rev_nu[q] for q in range(R.nrev)))
        if R.rxntype:
            f += 6
    return f


def bench_generated_chemistry(native, abi, cases, dev, n, with_oracle):
    """The 44-species / 391-reaction mechanism converted from
    test/Reactor/CO2_H2_Ar_mechanism_reduced/Chemistry.H (VD_CHEM_GEN,
    mechanism-specialised kernel), one source evaluation over n^3 cells of
    a 64^3 seeded state tiled; the first tile against the oracle."""
    m = cases.generated_mech("co2_h2_ar_reduced")
    blk = min(n, 64)
    rng = np.random.default_rng(11)
    cc = abi.comp(m.nspec)
    st = np.zeros((cc["NVAR"], blk, blk, blk))
    for sp in range(m.nspec):
        st[sp] = (1e16 if m.charge[sp] else 1e21) * np.exp(rng.standard_normal((blk, blk, blk)))
    names = [m.names[q].value.decode() for q in range(m.nspec)]
    for nm, f in (("AR", 0.9), ("H2", 0.075), ("CO2", 0.025)):
        st[names.index(nm)] = f * 2.446e25
    st[cc["ETEMP"]] = rng.uniform(0.3, 8.0, (blk, blk, blk)) * abi.EV
    st[cc["EEN"]] = 1.5 * st[m.e_idx] * abi.K_B * st[cc["ETEMP"]]
    tile = st
    reps = n // blk
    st = np.ascontiguousarray(np.tile(st, (1, reps, reps, reps)))
    g = abi.make_geom([n] * 3, [0] * 3, [1] * 3, [0] * 3)
    grid = native.Grid(dev, g)
    mech = native.Mech(dev, m)
    dst = grid.array(st.shape[0], host=st)
    del st
    out = grid.array(m.nspec + 1)
    prm = cases.default_params()
    native.rxn_source(mech, prm, grid, dst, out)
    dev.sync()
    dev.set_kernel_timing(abi.KT_CHEM)
    for _ in range(5):
        native.rxn_source(mech, prm, grid, dst, out)
    ms, launches, kbytes = dev.kernel_timing()
    dev.set_kernel_timing(abi.KT_NONE)
    t = ms / 1e3 / launches
    fpc = gen_flops_per_cell(m, abi)
    tfl = fpc * n ** 3 / t / 1e12
    res = {"mechanism": "test/Reactor/CO2_H2_Ar_mechanism_reduced (converted)", "cells": n ** 3,
           "nspec": m.nspec, "nreac": m.nreac, "specialised": mech.specialised, "launch_ms": round(t * 1e3, 3),
           "cells_per_s": n ** 3 / t, "flop_per_cell": fpc,
           "roofline": {"bound": "fp64", "achieved": round(tfl, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(tfl / FP64_PEAK_TFLOPS, 4),
                        "hbm_GBs": round(kbytes / launches / t / 1e9, 1)}}
    if with_oracle:
        from oracle import oracle as O
        got = out.get()[:, :blk, :blk, :blk]
        og = O.Grid(abi.make_geom([blk] * 3, [0] * 3, [1] * 3, [0] * 3))
        t0 = time.perf_counter()
        ref = O.rxn_source(m, prm, og, O.fill_patch(og, tile, 1), 1)
        cpu = time.perf_counter() - t0
        worst = max(float(np.max(np.abs(ref[q] - got[q])) / np.max(np.abs(ref[q])))
                    for q in range(ref.shape[0]) if np.max(np.abs(ref[q])) > 0)
        res["vs_oracle"] = {"tile_cells": blk ** 3, "worst_per_component_max_rel": worst,
                            "oracle_s": round(cpu, 3), "oracle_cells_per_s": blk ** 3 / cpu,
                            "oracle_threads": O.lib().or_num_threads()}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed steps of the headline (the configuration's 200)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", type=int, default=0, help="headline: 0 FP64 MG, 1 FP32 V-cycles in FP64 IR")
    ap.add_argument("--kernel", default="fused", choices=["fused", "gsrb", "transport", "chem", "resid", "none"])
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--headline-only", action="store_true", help="skip the other configurations")
    ap.add_argument("--only", default="", help="comma-separated subset of: mixed,discharge,poisson,chemistry,generated,batching,serving")
    ap.add_argument("--poisson-n", type=int, default=512)
    ap.add_argument("--chem-n", type=int, default=256)
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=4.0,
                    help="Poisson arrival rate of the serving leg (1/s; the default keeps the device saturated)")
    ap.add_argument("--serve-max-batch", type=int, default=16, help="cases per same-shape batch (1: no batching)")
    ap.add_argument("--batch-shape", default="", help="batching leg on one shape only, as DIM:N (e.g. 3:64)")
    ap.add_argument("--serve-sequential", default="sampled", choices=["sampled", "run", "none"],
                    help="FP64 one-at-a-time baseline of the serving leg: per-shape samples, or a full run")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the headline, copy 4 components of an n^3 array once (a plain copy of "
                         "8*4*n^3 B each way that scales rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from vidyut3d_amd import abi, cases, compare, native

    def case_fn(n, precision=args.precision):
        c = cases.air_3d(n=n)
        c.params.mg_precision = precision
        return c

    dev = native.Device(0)
    fam = {"fused": abi.KT_FUSED, "gsrb": abi.KT_GSRB, "transport": abi.KT_TRANSPORT, "chem": abi.KT_CHEM,
           "resid": abi.KT_RESID, "none": None}[args.kernel]
    case = case_fn(args.n)
    log("headline: %d^3 air-like discharge, %d warmup + %d timed steps" % (args.n, args.warmup, args.steps))
    wall, st, iters, _ = run_steps(native, abi, dev, case, args.steps, args.warmup)
    finite = bool(np.isfinite(st).all())
    del st
    value = case.ncells * args.steps / wall
    log("%d steps in %.3f s (%.2f ms/step), %d multigrid iterations, finite %s"
        % (args.steps, wall, wall / args.steps * 1e3, iters, finite))
    # the same steps again with HIP-event timing of every kernel family (the
    # value above is measured without the events)
    families = [abi.KT_FUSED, abi.KT_TRANSPORT, abi.KT_MGCOARSE, abi.KT_SPECIES, abi.KT_CHEM, abi.KT_POTPOST]
    if fam is not None and fam not in families:
        families.append(fam)
    wall_t, _, _, kt = run_steps(native, abi, dev, case_fn(args.n), args.steps, args.warmup, families, state=False)
    fam_roofs = family_rooflines(abi, kt, args.steps)
    log("families: " + ", ".join("%s %.3f ms/step %.2f" % (k, v["ms_per_step"], v["frac"]) for k, v in fam_roofs.items()))
    roof = None
    if fam is not None and kt[fam][1] > 0:
        ms_total, launches, kbytes = kt[fam]
        avg_s = ms_total / 1e3 / launches
        byts = kbytes / launches
        ach = byts / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": pmc_traffic(args.kernel, args.n, args.precision), "kernel": args.kernel,
                "launches": launches, "avg_launch_us": round(avg_s * 1e6, 2),
                "algorithmic_bytes_per_launch": round(byts)}

    if args.copy_probe:
        g = native.Grid(dev, case.geom)
        src, dst = g.array(4), g.array(4)
        native.fill_patch(g, src, dst, 0, 4)  # no ghost cells: a plain copy
        dev.sync()
        del src, dst, g
        log("copy probe: 4 x %d^3 x 8 B = %d B each way" % (args.n, 32 * case.ncells))

    configs = {}
    only = set(x for x in args.only.split(",") if x)

    def want(name):
        return not args.headline_only and (not only or name in only)

    if want("mixed"):
        log("headline case with FP32 V-cycles in FP64 iterative refinement")
        cmp_case = case_fn(args.n, precision=1)
        wall_mp, st_mp, iters_mp, _ = run_steps(native, abi, dev, cmp_case, args.steps, args.warmup)
        del st_mp
        configs["air_3d_mixed_precision"] = {
            "cell_steps_per_s": case.ncells * args.steps / wall_mp,
            "ms_per_step": wall_mp / args.steps * 1e3, "mg_iterations": iters_mp,
            "speedup_over_fp64": wall / wall_mp}
    if want("discharge"):
        log("2-D 128^2 discharge, 100 steps (GPU and oracle)")
        configs["discharge_2d_128"] = bench_discharge_2d(native, abi, cases, compare, dev, not args.no_cpu)
    if want("poisson"):
        log("%d^3 dielectric Poisson FP64 vs mixed" % args.poisson_n)
        configs["dielectric_poisson"] = bench_dielectric(native, abi, cases, compare, dev, args.poisson_n,
                                                         not args.no_cpu)
    if want("chemistry"):
        log("%d^3 40-species chemistry" % args.chem_n)
        configs["chemistry_40sp_200rx"] = bench_chemistry(native, abi, cases, dev, args.chem_n)
    if want("generated"):
        log("%d^3 44-species / 391-reaction generated mechanism" % args.chem_n)
        configs["chemistry_co2_h2_ar_44sp"] = bench_generated_chemistry(native, abi, cases, dev, args.chem_n,
                                                                         not args.no_cpu)
        log("generated mechanism: %.3f ms per launch, %s" % (configs["chemistry_co2_h2_ar_44sp"]["launch_ms"],
                                                            configs["chemistry_co2_h2_ar_44sp"].get("vs_oracle")))
    if want("batching"):
        log("same-shape batching per shape bucket")
        shp = [tuple(int(v) for v in args.batch_shape.split(":"))] if args.batch_shape else None
        configs["batching"] = bench_batching(native, shapes=shp)
    served = []
    if want("serving") and args.serve_requests > 0:
        from vidyut3d_amd import serve
        log("serving loop: %d requests, Poisson arrivals at %.2f/s" % (args.serve_requests, args.serve_rate))
        configs["serving"], served = serve.bench(
            dev_index=0, requests=args.serve_requests, rate_hz=args.serve_rate, sequential=args.serve_sequential,
            max_batch=args.serve_max_batch,
            progress=lambda done, waiting, t: log("serving: %d of %d requests done, %d waiting, %.0f s"
                                                  % (done, args.serve_requests, waiting, t)))
        log("serving: p50 %.2f s, p99 %.2f s, aggregate %.3g cell-steps/s"
            % (configs["serving"]["p50_time_to_solution_s"], configs["serving"]["p99_time_to_solution_s"],
               configs["serving"]["aggregate_cell_steps_per_s"]))

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle) %d step(s) at %d^3" % (args.cpu_steps, args.cpu_n))
        cpu, ref = cpu_baseline_steps(lambda n: case_fn(n, 0), args.cpu_n, args.cpu_steps)
        # the headline legs against the all-FP64 oracle on the same sample
        errs = {}
        for prec in (0, 1):
            c = case_fn(args.cpu_n, precision=prec)
            sim = native.Sim(dev, c)
            sim.evolve(args.cpu_steps)
            got = sim.get_state()
            sim.close()
            e = compare.state_errors(ref, got, c.mech.nspec, c.geom.dim)
            names = [c.mech.names[s].value.decode() for s in range(c.mech.nspec)] + abi.PLASMA_VAR_NAMES
            q = max(e, key=e.get)
            errs["fp64" if prec == 0 else "mixed"] = {"worst_rel_l2": e[q], "field": names[q]}
            del got
        del ref
        cpu["rel_l2_vs_fp64_oracle"] = errs
        log("headline vs oracle after %d step(s): %s" % (args.cpu_steps, errs))
        cpu["single_thread"] = cpu_single_thread(lambda n: case_fn(n, 0), 64, 1)
        kept = [r for r in served if r.state is not None]
        if kept:
            # served (FP32 V-cycle) results, and the same requests run alone
            # in FP64 on the GPU, against the all-FP64 oracle
            from oracle import oracle as O
            rows = []
            for r in kept:
                c = r.make_case(0)
                ref, _ = O.evolve(c, r.nsteps)
                names = [c.mech.names[s].value.decode() for s in range(c.mech.nspec)] + abi.PLASMA_VAR_NAMES
                row = {"request": r.rid, "shape": "%dD %d" % (r.dim, r.n), "voltage": round(r.voltage, 1)}
                for key, st in (("served_mixed", r.state), ("alone_fp64", r.state_fp64)):
                    e = compare.state_errors(ref, st, c.mech.nspec, c.geom.dim)
                    q = max(e, key=e.get)
                    row[key] = {"worst": e[q], "field": names[q]}
                if max(row["served_mixed"]["worst"], row["alone_fp64"]["worst"]) > 1e-8:
                    # how far the oracle itself moves when its initial state
                    # moves by 1e-14 (relative): the case's own sensitivity
                    cp = r.make_case(0)
                    st0 = np.array(cp.state, copy=True)
                    st0 *= 1.0 + 1e-14 * np.random.default_rng(r.rid).standard_normal(st0.shape)
                    cp.state = st0
                    ref_p, _ = O.evolve(cp, r.nsteps)
                    e = compare.state_errors(ref, ref_p, c.mech.nspec, c.geom.dim)
                    q = max(e, key=e.get)
                    row["oracle_perturbed_1e-14"] = {"worst": e[q], "field": names[q]}
                    del ref_p
                rows.append(row)
                r.state = r.state_fp64 = None
                log("serving request %d (%s) vs oracle: served %.2e (%s), FP64 alone %.2e (%s)%s"
                    % (r.rid, row["shape"], row["served_mixed"]["worst"], row["served_mixed"]["field"],
                       row["alone_fp64"]["worst"], row["alone_fp64"]["field"],
                       "; oracle vs oracle from a 1e-14 perturbed start %.2e (%s)"
                       % (row["oracle_perturbed_1e-14"]["worst"], row["oracle_perturbed_1e-14"]["field"])
                       if "oracle_perturbed_1e-14" in row else ""))
            configs["serving"]["rel_l2_vs_fp64_oracle"] = {
                "worst_served_mixed": max(x["served_mixed"]["worst"] for x in rows),
                "worst_alone_fp64": max(x["alone_fp64"]["worst"] for x in rows),
                "requests_checked": len(rows), "of": len(served), "per_request": rows}

    if fam_roofs:
        configs["kernel_families"] = fam_roofs
    out = {
        "metric": "cell-steps/s FP64 vs mixed-precision+batched; p50/p99 time-to-solution",
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f64+f32mg", "data": "synthetic",
        "config": {"workload": "3-D %d^3 single-level discharge, 10 species / 30 seeded reactions, "
                               "WENO-Z + implicit MG diffusion + chemistry + Poisson, FP64" % args.n,
                   "cells": case.ncells, "nspec": case.mech.nspec, "nreac": case.mech.nreac,
                   "dt": case.dt, "mg_precision": args.precision,
                   "mg_iterations": iters, "state_finite": finite},
        "roofline": roof, "cpu_baseline": cpu, "configs": configs,
    }
    print(json.dumps(out))
    dev.close()


if __name__ == "__main__":
    main()
