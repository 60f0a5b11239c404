"""Benchmark: cell-steps/s of the Vidyut per-timestep hot path on one MI355X.

Workload (default): 3-D 256^3 single-level discharge, FP64, 10 species /
30 seeded reactions (air-like), WENO-Z transport of electrons, electron
energy and 4 ions with implicit multigrid diffusion, chemistry, Joule
heating, the potential multigrid solve and the CFL reduction every step.
One "step" = Vidyut::Evolve's step loop body (Source/Evolve.cpp:33-369).

Prints ONE JSON line (rank 0).  --gpus above 1 is refused: this build runs
on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Canonical algorithmic bytes per launch (FP64 = 8 B per value, N = cells
# of the level the launch runs on), accumulated by the library per launch:
#  gsrb      one colour of a red-black sweep on the finest level: a full
#            sweep reads phi, rhs and n_b face arrays and writes phi,
#            8*N*(3 + n_b); one colour is half (n_b = 0 for the constant-
#            coefficient potential solve, 3 for the implicit species solves)
#  transport flux + divergence of one species: 8*N*(NS + 6)
#  chem      one source evaluation: 8*N*(2*NS + 3)
#  resid     finest residual + norm: 8*N*(4 + n_b)


def cpu_baseline(case_fn, sample_n, sample_steps):
    """The CPU oracle (OpenMP on this host's cores) on a bounded sample of
    the same workload: `sample_steps` steps of the same physics at
    sample_n^3 cells."""
    from oracle import oracle as O
    c = case_fn(sample_n)
    t0 = time.perf_counter()
    O.evolve(c, sample_steps)
    dt = time.perf_counter() - t0
    cells = c.ncells
    return {"value": cells * sample_steps / dt, "unit": "cell-steps/s",
            "cores": int(O.lib().or_num_threads()), "kind": "port",
            "sample": "oracle (C, OpenMP) %d step(s) of the same case at %d^3 cells, %.1f s"
                      % (sample_steps, sample_n, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", type=int, default=0, help="0 FP64 multigrid, 1 FP32 V-cycles in FP64 IR")
    ap.add_argument("--kernel", default="gsrb", choices=["gsrb", "transport", "chem", "resid", "none"])
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from vidyut3d_amd import abi, cases, native

    def case_fn(n):
        c = cases.air_3d(n=n)
        c.params.mg_precision = args.precision
        return c

    t_setup = time.perf_counter()
    case = case_fn(args.n)
    dev = native.Device(0)
    sim = native.Sim(dev, case)
    print("bench: setup %.1f s, %d^3 cells, warmup %d steps" % (time.perf_counter() - t_setup,
                                                               args.n, args.warmup), flush=True)
    sim.evolve(args.warmup)
    dev.sync()
    fam = {"gsrb": abi.KT_GSRB, "transport": abi.KT_TRANSPORT, "chem": abi.KT_CHEM,
           "resid": abi.KT_RESID, "none": abi.KT_NONE}[args.kernel]
    dev.set_kernel_timing(fam)
    it0 = sim.mg_iterations
    dev.sync()
    t0 = time.perf_counter()
    sim.evolve(args.steps)
    dev.sync()
    t1 = time.perf_counter()
    ms_total, launches, kbytes = dev.kernel_timing()
    dev.set_kernel_timing(abi.KT_NONE)
    mg_iters = sim.mg_iterations - it0
    wall = t1 - t0
    ncells = case.ncells
    value = ncells * args.steps / wall
    st = sim.get_state()
    import numpy as np
    finite = bool(np.isfinite(st).all())
    print("bench: %d steps in %.3f s, %d multigrid iterations, state finite: %s"
          % (args.steps, wall, mg_iters, finite), flush=True)

    roof = None
    if fam != abi.KT_NONE and launches > 0:
        avg_s = ms_total / 1e3 / launches
        byts = kbytes / launches
        ach = byts / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(ach / 8000.0, 4), "traffic": None,
                "kernel": args.kernel, "launches": launches,
                "avg_launch_us": round(avg_s * 1e6, 2),
                "algorithmic_bytes_per_launch": round(byts)}

    cpu = None
    if not args.no_cpu:
        print("bench: cpu baseline (oracle) %d step(s) at %d^3 ..." % (args.cpu_steps, args.cpu_n),
              flush=True)
        cpu = cpu_baseline(case_fn, args.cpu_n, args.cpu_steps)

    out = {
        "metric": "cell-steps/s FP64 vs mixed-precision+batched; p50/p99 time-to-solution",
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f64+f32mg", "data": "synthetic",
        "config": {"workload": "3-D %d^3 single-level discharge, 10 species / 30 seeded reactions, "
                               "WENO-Z + implicit MG diffusion + chemistry + Poisson, FP64"
                               % args.n,
                   "cells": ncells, "nspec": case.mech.nspec, "nreac": case.mech.nreac,
                   "dt": case.dt, "mg_precision": args.precision,
                   "mg_iterations": mg_iters, "state_finite": finite},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    sim.close()
    dev.close()


if __name__ == "__main__":
    main()
