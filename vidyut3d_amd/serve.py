"""Device-resident serving loop for discharge cases (continuous batching).

Cases of the 2-D / 3-D discharge physics arrive as a Poisson process with
grid sizes log-uniform between 64^2 and 256^3 cells and applied voltages
U[0.5, 10] kV.  Each case keeps its whole state in device memory from
admission to completion (vd_sim).  Admission is by HBM footprint: a case
is started as soon as a worker slot is free and its footprint fits in the
memory budget next to the cases already running; cases that arrive while
others are on the GPU are admitted as slots and memory free up.  Every
slot owns a vd_ctx (its own HIP stream), so the kernels of concurrent
cases overlap on the device.  Time-to-solution = completion - arrival.
"""
import math
import queue
import threading
import time

import numpy as np

from . import abi, cases, native


class Request:
    def __init__(self, rid, arrival, dim, n, voltage, nsteps, seed):
        self.rid, self.arrival, self.dim, self.n = rid, arrival, dim, n
        self.voltage, self.nsteps, self.seed = voltage, nsteps, seed
        self.start = self.done = None
        self.state = None
        self.keep = False  # keep the final state (host copy) of this request
        self.state_fp64 = None
        self.error = None

    @property
    def cells(self):
        return self.n ** self.dim

    def make_case(self, mg_precision):
        # one physics for every request (the config-0 / config-1 models); the
        # request's seed draws its initial state
        if self.dim == 2:
            c = cases.discharge_2d(n=self.n, seed=self.seed, nsteps=self.nsteps)
        else:
            c = cases.air_3d(n=self.n, seed=1, state_seed=self.seed, nsteps=self.nsteps)
        c.params.voltage_amp_2 = self.voltage
        c.params.mg_precision = mg_precision
        return c


def footprint_bytes(dim, n, nspec):
    """Device bytes of one case: state and old state (2*NVAR), the two
    ghosted state buffers of the step (Sborder and its partner, 2*NVAR with
    ghost growth), reaction sources (NS+1), face arrays (12 + 3 flux
    scratch) and the multigrid workspaces (~16 arrays)."""
    cells = n ** dim
    nvar = nspec + abi.NUM_PLASMAVARS
    grow = ((n + 6) / n) ** dim
    arrays = 2 * nvar + 2 * nvar * grow + (nspec + 1) + 1 + 15 + 16
    return int(8 * cells * arrays * 1.1)


def draw_requests(rate_hz, duration_s=None, seed=0, nsteps=500, min_cells=64 ** 2, max_cells=256 ** 3,
                  count=None):
    """Poisson arrivals, log-uniform cell counts (2-D up to 256^2, 3-D
    above), sizes rounded to multiples of 16, voltages U[0.5, 10] kV; the
    arrivals within duration_s, or the first `count` of them."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if (duration_s is not None and t > duration_s) or (count is not None and len(out) >= count):
            break
        cells = math.exp(rng.uniform(math.log(min_cells), math.log(max_cells)))
        if cells <= 256 ** 2:
            dim, n = 2, int(round(math.sqrt(cells) / 16.0)) * 16
            n = min(max(n, 64), 256)
        else:
            dim, n = 3, int(round(cells ** (1.0 / 3.0) / 16.0)) * 16
            n = min(max(n, 48), 256)
        out.append(Request(len(out), t, dim, n, rng.uniform(500.0, 1.0e4), nsteps,
                           int(rng.integers(0, 2 ** 31))))
    return out


_REGION_BYTES = {}


def case_region_bytes(dev_index, case):
    """Device bytes one case of `case`'s shape and options takes in a
    batched context: measured once per (dim, n, nspec, precision) by a
    one-case batched context that runs one step (every buffer and solver
    workspace of a step is allocated by then), plus a margin."""
    g = case.geom
    key = (g.dim, tuple(g.dom_hi[d] - g.dom_lo[d] + 1 for d in range(g.dim)), case.mech.nspec,
           case.params.mg_precision, case.params.hyp_order)
    if key not in _REGION_BYTES:
        nspec = case.mech.nspec
        n = key[1][0]
        guess = 3 * footprint_bytes(g.dim, n, nspec) + (64 << 20)
        dev = native.BatchDevice(dev_index, 1, guess)
        try:
            sim = native.Sim(dev, case)
            sim.evolve(1)
            used = dev.arena_used()
            sim.close()
        finally:
            dev.close()
        _REGION_BYTES[key] = int(used * 1.02) + (4 << 20)
    return _REGION_BYTES[key]


class Server:
    """Worker slots with their own contexts; admission by footprint."""

    def __init__(self, dev_index=0, slots=4, mem_budget_bytes=96 << 30, mg_precision=1,
                 keep_state=False):
        self.slots = slots
        self.budget = mem_budget_bytes
        self.mg_precision = mg_precision
        self.keep_state = keep_state
        self.devs = [native.Device(dev_index) for _ in range(slots)]
        self.free_slots = list(range(slots))
        self.in_use = 0
        self.lock = threading.Condition()
        self.done_q = queue.Queue()

    def close(self):
        for d in self.devs:
            d.close()

    def _run(self, slot, req):
        try:
            case = req.make_case(self.mg_precision)
            sim = native.Sim(self.devs[slot], case)
            req.t_ready = time.perf_counter()
            sim.evolve(req.nsteps)
            if self.keep_state or req.keep:
                req.state = sim.get_state()
            sim.close()
            self.devs[slot].release_workspaces()
        except Exception as e:  # reported per request, the loop keeps serving
            req.error = repr(e)
        req.done = time.perf_counter()
        self.done_q.put((slot, req))

    def run(self, requests, t0=None):
        """Serves `requests` (arrival times relative to t0); returns them
        with start/done filled."""
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter() if t0 is None else t0
        waiting, threads, finished = [], [], []
        nsp = {2: 3, 3: len(cases.AIR_SPECIES)}
        while len(finished) < len(requests):
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.pop(0))
            # admit in arrival order while a slot is free and memory fits
            admitted = True
            while admitted and waiting and self.free_slots:
                admitted = False
                req = waiting[0]
                fb = footprint_bytes(req.dim, req.n, nsp[req.dim])
                if self.in_use + fb <= self.budget or self.in_use == 0:
                    waiting.pop(0)
                    slot = self.free_slots.pop(0)
                    self.in_use += fb
                    req.footprint = fb
                    req.start = time.perf_counter()
                    th = threading.Thread(target=self._run, args=(slot, req), daemon=True)
                    th.start()
                    threads.append(th)
                    admitted = True
            timeout = 0.002
            if pending and not waiting:
                timeout = max(0.0, min(0.05, pending[0].arrival - (time.perf_counter() - t0)))
            try:
                slot, req = self.done_q.get(timeout=timeout)
                self.free_slots.append(slot)
                self.in_use -= req.footprint
                finished.append(req)
            except queue.Empty:
                pass
        for th in threads:
            th.join()
        for r in requests:
            r.arrival_abs = t0 + r.arrival
        return requests


def summarize(reqs):
    tts = np.array([r.done - r.arrival_abs for r in reqs])
    wait = np.array([r.start - r.arrival_abs for r in reqs])
    cell_steps = sum(r.cells * r.nsteps for r in reqs)
    span = max(r.done for r in reqs) - min(r.arrival_abs for r in reqs)
    setup = [r.t_ready - r.start for r in reqs if getattr(r, "t_ready", None)]
    return {"requests": len(reqs), "errors": sum(1 for r in reqs if r.error),
            "mean_setup_s": float(np.mean(setup)) if setup else None,
            "p50_time_to_solution_s": float(np.percentile(tts, 50)),
            "p99_time_to_solution_s": float(np.percentile(tts, 99)),
            "p50_queue_wait_s": float(np.percentile(wait, 50)),
            "p99_queue_wait_s": float(np.percentile(wait, 99)),
            "aggregate_cell_steps_per_s": cell_steps / span, "makespan_s": span,
            "sizes": sorted("%dD %d" % (r.dim, r.n) for r in reqs)}


def step_times(dev_index, shapes, steps=10, warmup=2, mg_precision=0):
    """ms per step of one case of each (dim, n) shape alone on the device,
    from `steps` steps after `warmup` (the same physics as the requests)."""
    dev = native.Device(dev_index)
    out = {}
    try:
        for dim, n in sorted(set(shapes)):
            c = Request(0, 0.0, dim, n, 5000.0, steps, 1).make_case(mg_precision)
            sim = native.Sim(dev, c)
            sim.evolve(warmup)
            dev.sync()
            t0 = time.perf_counter()
            sim.evolve(steps)
            dev.sync()
            out[(dim, n)] = (time.perf_counter() - t0) / steps * 1e3
            sim.close()
            dev.release_workspaces()
    finally:
        dev.close()
    return out


def run_sequential(dev_index, reqs, mg_precision=0):
    """The requests one after the other on one context (no overlap, no
    arrival gaps); returns the total wall time."""
    srv = Server(dev_index, slots=1, mg_precision=mg_precision)
    seq = [Request(r.rid, 0.0, r.dim, r.n, r.voltage, r.nsteps, r.seed) for r in reqs]
    t0 = time.perf_counter()
    try:
        srv.run(seq)
    finally:
        srv.close()
    return time.perf_counter() - t0, seq


def bench(dev_index=0, requests=48, rate_hz=1.0, nsteps=500, slots=4, seed=7, sequential="sampled",
          keep_cells_steps=2.0e7):
    """Serves `requests` Poisson arrivals with FP32 V-cycles (mixed
    precision); the baseline is the same request list in FP64, one at a time
    (sequential="run": actually run; "sampled": the sum over requests of
    nsteps times each shape's measured FP64 ms/step).  Requests whose
    cells * steps are at most keep_cells_steps keep their final state, and
    are run again alone in FP64 (state_fp64), for a check of both against
    the FP64 oracle."""
    reqs = draw_requests(rate_hz, seed=seed, nsteps=nsteps, count=requests)
    for r in reqs:
        r.keep = r.cells * r.nsteps <= keep_cells_steps
    srv = Server(dev_index, slots=slots)
    try:
        srv.run(reqs)
    finally:
        srv.close()
    out = summarize(reqs)
    out.update({"arrival_rate_hz": rate_hz, "steps_per_case": nsteps, "slots": slots, "mg_precision": 1})
    # the kept requests once more, alone and all-FP64: the state the FP64
    # path gives for the same request, beside the served one
    kept = [Request(r.rid, 0.0, r.dim, r.n, r.voltage, r.nsteps, r.seed) for r in reqs if r.keep]
    if kept:
        solo = Server(dev_index, slots=1, mg_precision=0, keep_state=True)
        try:
            solo.run(kept)
        finally:
            solo.close()
        by_id = {r.rid: r for r in kept}
        for r in reqs:
            if r.keep:
                r.state_fp64 = by_id[r.rid].state
    cell_steps = sum(r.cells * r.nsteps for r in reqs)
    if sequential == "run":
        wall, _ = run_sequential(dev_index, reqs, 0)
        out["fp64_sequential"] = {"kind": "run", "wall_s": wall, "cell_steps_per_s": cell_steps / wall}
    elif sequential == "sampled":
        ms = step_times(dev_index, [(r.dim, r.n) for r in reqs])
        wall = sum(r.nsteps * ms[(r.dim, r.n)] for r in reqs) / 1e3
        out["fp64_sequential"] = {"kind": "sampled: sum of nsteps x the shape's FP64 ms/step (10 steps alone)",
                                  "wall_s": wall, "cell_steps_per_s": cell_steps / wall}
    if "fp64_sequential" in out:
        out["speedup_over_fp64_sequential"] = out["aggregate_cell_steps_per_s"] / out["fp64_sequential"]["cell_steps_per_s"]
    return out, reqs
