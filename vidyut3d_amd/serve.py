"""Device-resident serving loop for discharge cases (continuous batching).

Cases of the 2-D / 3-D discharge physics arrive as a Poisson process with
grid sizes log-uniform between 64^2 and 256^3 cells (snapped to a ladder of
shape buckets) and applied voltages U[0.5, 10] kV.

* Same-shape batching.  The cases of one shape bucket run in a *group*: a
  batched context (vd_ctx_create_batched) whose regions each hold one
  case's state and solver workspaces.  One launch steps every active case
  of the group (a z-block per case), and each case keeps its own solver
  convergence, so it gets exactly the bits it gets alone.
* Continuous batching.  A group advances its cases in chunks of steps;
  between chunks, finished cases leave (their state is read back if kept)
  and requests of the same shape that arrived meanwhile take the free
  regions.  A case's state stays on the GPU from admission to completion.
* Admission by HBM footprint.  A group reserves its regions when it is
  made (capacity x the shape's measured bytes per case); a request starts
  when a group of its shape has a free region, or a new group fits next to
  the others within the memory budget and the stream count.  Idle groups
  are kept for later requests of their shape until their memory is needed.
* Overlap.  Each running group owns a HIP stream (its context), so the
  groups' kernels overlap on the device.
Time to solution = completion - arrival.
"""
import math
import queue
import threading
import time

import numpy as np

from . import abi, cases, native

# shape buckets (cells per side): every request is snapped to the nearest in
# log(cells); about 25 % apart, so a stream of log-uniform sizes falls into
# few shapes that batch
BUCKETS_2D = (64, 80, 96, 128, 160, 192, 256)
BUCKETS_3D = (48, 64, 80, 96, 128, 160, 192, 256)


class Request:
    def __init__(self, rid, arrival, dim, n, voltage, nsteps, seed):
        self.rid, self.arrival, self.dim, self.n = rid, arrival, dim, n
        self.voltage, self.nsteps, self.seed = voltage, nsteps, seed
        self.start = self.done = None
        self.state = None
        self.keep = False  # keep the final state (host copy) of this request
        self.state_fp64 = None
        self.error = None
        self.remaining = nsteps
        self.batch_peak = 0  # the most cases its group stepped together while it ran

    @property
    def cells(self):
        return self.n ** self.dim

    @property
    def shape(self):
        return (self.dim, self.n)

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


def snap(cells):
    """(dim, n) of the bucket nearest to `cells` in log scale: 2-D up to
    256^2 cells, 3-D above."""
    if cells <= 256 ** 2:
        dim, ladder = 2, BUCKETS_2D
    else:
        dim, ladder = 3, BUCKETS_3D
    lc = math.log(cells)
    n = min(ladder, key=lambda m: abs(dim * math.log(m) - lc))
    return dim, n


def draw_requests(rate_hz, duration_s=None, seed=0, nsteps=500, min_cells=64 ** 2, max_cells=256 ** 3,
                  count=None):
    """Poisson arrivals, log-uniform cell counts snapped to the shape
    buckets, voltages U[0.5, 10] kV; the arrivals within duration_s, or the
    first `count` of them."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if (duration_s is not None and t > duration_s) or (count is not None and len(out) >= count):
            break
        cells = math.exp(rng.uniform(math.log(min_cells), math.log(max_cells)))
        dim, n = snap(cells)
        out.append(Request(len(out), t, dim, n, rng.uniform(500.0, 1.0e4), nsteps,
                           int(rng.integers(0, 2 ** 31))))
    return out


_REGION_BYTES = {}
_REGION_LOCK = threading.Lock()


def case_region_bytes(dev_index, case):
    """Device bytes one case of `case`'s shape and options takes in a
    batched context: measured once per (dim, n, nspec, precision) by a
    one-case batched context that runs one step (every buffer and solver
    workspace of a step is allocated by then), plus a margin."""
    g = case.geom
    key = (g.dim, tuple(g.dom_hi[d] - g.dom_lo[d] + 1 for d in range(g.dim)), case.mech.nspec,
           case.params.mg_precision, case.params.hyp_order)
    with _REGION_LOCK:
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


class Group:
    """The batched context of one shape: `cap` regions, and a worker
    thread that steps the occupied ones chunk by chunk."""

    def __init__(self, server, shape, cap, region):
        self.server, self.shape, self.cap, self.region = server, shape, cap, region
        self.bytes = cap * region
        self.dev = native.BatchDevice(server.dev_index, cap, region)
        self.sim = None
        # the request of each region; one placed by the scheduler and not yet
        # loaded has remaining None
        self.slots = [None] * cap
        self.running = False
        self.last_used = time.perf_counter()

    def room(self):
        return sum(1 for s in self.slots if s is None)

    def place(self, req):
        """reserves a free region for req (the scheduler, under the lock)"""
        c = self.slots.index(None)
        req.remaining = None
        self.slots[c] = req

    def close(self):
        if self.sim is not None:
            self.sim.close()
        self.dev.close()

    def _load(self):
        """the placed requests' initial states into their regions (the
        worker, between chunks; outside the scheduler's lock)"""
        srv = self.server
        for c in range(self.cap):
            req = self.slots[c]
            if req is None or req.remaining is not None:
                continue
            try:
                case = req.make_case(srv.mg_precision)
                if self.sim is None:
                    self.sim = native.Sim(self.dev, case)
                self.sim.set_state_case(c, case.state)
                self.dev.set_case_voltage(c, case.params.voltage_amp_1, case.params.voltage_amp_2)
                del case
            except Exception as e:  # reported per request, the group keeps serving
                req.error = repr(e)
                req.done = time.perf_counter()
                req.remaining = 0
                self.slots[c] = None
                srv.done_q.put((self, req))
                continue
            req.t_ready = time.perf_counter()
            req.remaining = req.nsteps

    def work(self):
        srv = self.server
        while True:
            self._load()
            with srv.lock:
                if all(s is None for s in self.slots):
                    self.running = False
                    self.last_used = time.perf_counter()
                    srv.done_q.put((self, None))  # the group went idle
                    return
            active = [c for c in range(self.cap) if self.slots[c] is not None and self.slots[c].remaining]
            if not active:
                continue
            try:
                self.dev.set_active(active)
                k = min(srv.chunk, min(self.slots[c].remaining for c in active))
                self.sim.evolve(k)
            except Exception as e:
                for c in active:
                    r = self.slots[c]
                    r.error = repr(e)
                    r.done = time.perf_counter()
                    self.slots[c] = None
                    srv.done_q.put((self, r))
                continue
            for c in active:
                r = self.slots[c]
                r.remaining -= k
                r.batch_peak = max(r.batch_peak, len(active))
                if r.remaining == 0:
                    if srv.keep_state or r.keep:
                        r.state = self.sim.get_state_case(c)
                    r.done = time.perf_counter()
                    self.slots[c] = None
                    srv.done_q.put((self, r))


class Server:
    """Continuous-batching server: shape groups on up to `slots` streams,
    admission by HBM footprint within mem_budget_bytes.  max_batch = 1
    gives one case per group (no batching)."""

    def __init__(self, dev_index=0, slots=4, mem_budget_bytes=160 << 30, mg_precision=1, keep_state=False,
                 max_batch=16, group_bytes=32 << 30, chunk=25):
        self.dev_index = dev_index
        self.slots = slots
        self.budget = mem_budget_bytes
        self.mg_precision = mg_precision
        self.keep_state = keep_state
        self.max_batch = max_batch
        self.group_bytes = group_bytes
        self.chunk = chunk
        self.groups = {}  # shape -> [Group]
        self.lock = threading.Lock()
        self.done_q = queue.Queue()
        self.threads = []

    def close(self):
        for th in self.threads:
            th.join()
        for gs in self.groups.values():
            for g in gs:
                g.close()
        self.groups = {}

    def region(self, shape):
        dim, n = shape
        return case_region_bytes(self.dev_index, Request(0, 0.0, dim, n, 1000.0, 1, 1).make_case(self.mg_precision))

    def warm(self, shapes):
        """measures the bytes per case of each shape (one step each) before
        requests are timed"""
        for s in sorted(set(shapes)):
            self.region(s)

    def _used(self):
        return sum(g.bytes for gs in self.groups.values() for g in gs)

    def _running(self):
        return sum(1 for gs in self.groups.values() for g in gs if g.running)

    def _start(self, g):
        g.running = True
        th = threading.Thread(target=g.work, daemon=True)
        th.start()
        self.threads.append(th)

    def _place(self, req):
        """the group that takes `req` now (started if idle), or None"""
        gs = self.groups.setdefault(req.shape, [])
        for g in gs:
            if g.running and g.room() > 0:
                g.place(req)
                return g
        if self._running() >= self.slots:
            return None
        for g in gs:
            if not g.running and g.room() > 0:
                g.place(req)
                self._start(g)
                return g
        region = self.region(req.shape)
        cap = max(1, min(self.max_batch, int(self.group_bytes // region)))
        need = cap * region
        # make room: close idle groups, the longest idle first
        idle = sorted((g for x in self.groups.values() for g in x if not g.running), key=lambda g: g.last_used)
        while self._used() + need > self.budget and idle:
            g = idle.pop(0)
            self.groups[g.shape].remove(g)
            g.close()
        if self._used() + need > self.budget and self._used() > 0:
            return None
        g = Group(self, req.shape, cap, region)
        gs.append(g)
        g.place(req)
        self._start(g)
        return g

    def run(self, requests, t0=None, progress=None, every_s=30.0):
        """Serves `requests` (arrival times relative to t0); returns them
        with start/done filled.  progress(finished, waiting, elapsed) is
        called about every every_s seconds."""
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter() if t0 is None else t0
        waiting, finished = [], []
        t_report = time.perf_counter()
        while len(finished) < len(requests):
            if progress is not None and time.perf_counter() - t_report >= every_s:
                t_report = time.perf_counter()
                progress(len(finished), len(waiting), t_report - t0)
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.pop(0))
            # admission in arrival order; a request that cannot start now
            # does not hold back requests of other shapes
            with self.lock:
                for req in list(waiting):
                    if self._place(req) is not None:
                        req.start = time.perf_counter()
                        waiting.remove(req)
            timeout = 0.002
            if pending and not waiting:
                timeout = max(0.0, min(0.05, pending[0].arrival - (time.perf_counter() - t0)))
            try:
                g, req = self.done_q.get(timeout=timeout)
                if req is not None:
                    finished.append(req)
            except queue.Empty:
                pass
        for r in requests:
            r.arrival_abs = t0 + r.arrival
        return requests


def summarize(reqs):
    tts = np.array([r.done - r.arrival_abs for r in reqs])
    wait = np.array([r.start - r.arrival_abs for r in reqs])
    cell_steps = sum(r.cells * r.nsteps for r in reqs)
    span = max(r.done for r in reqs) - min(r.arrival_abs for r in reqs)
    setup = [r.t_ready - r.start for r in reqs if getattr(r, "t_ready", None)]
    shapes = {}
    for r in reqs:
        shapes.setdefault("%dD %d" % (r.dim, r.n), []).append(r.batch_peak)
    return {"requests": len(reqs), "errors": sum(1 for r in reqs if r.error),
            "mean_setup_s": float(np.mean(setup)) if setup else None,
            "p50_time_to_solution_s": float(np.percentile(tts, 50)),
            "p99_time_to_solution_s": float(np.percentile(tts, 99)),
            "p50_queue_wait_s": float(np.percentile(wait, 50)),
            "p99_queue_wait_s": float(np.percentile(wait, 99)),
            "aggregate_cell_steps_per_s": cell_steps / span, "makespan_s": span,
            "shapes": {k: {"requests": len(v), "max_batch": max(v)} for k, v in sorted(shapes.items())}}


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


def run_alone(dev_index, req, mg_precision=0):
    """one request on a plain (unbatched) context; its final state"""
    dev = native.Device(dev_index)
    try:
        sim = native.Sim(dev, req.make_case(mg_precision))
        sim.evolve(req.nsteps)
        out = sim.get_state()
        sim.close()
    finally:
        dev.close()
    return out


def run_sequential(dev_index, reqs, mg_precision=0):
    """The requests one after the other on one plain context (no overlap,
    no batching, no arrival gaps); returns the total wall time."""
    dev = native.Device(dev_index)
    t0 = time.perf_counter()
    try:
        for r in reqs:
            sim = native.Sim(dev, r.make_case(mg_precision))
            sim.evolve(r.nsteps)
            dev.sync()
            sim.close()
            dev.release_workspaces()
    finally:
        dev.close()
    return time.perf_counter() - t0


def bench(dev_index=0, requests=48, rate_hz=1.0, nsteps=500, slots=4, seed=7, sequential="sampled",
          keep_cells_steps=2.0e7, max_batch=16, max_kept=24, progress=None):
    """Serves `requests` Poisson arrivals with FP32 V-cycles (mixed
    precision) and same-shape batching; the baseline is the same request
    list in FP64, one at a time without batching (sequential="run":
    actually run; "sampled": the sum over requests of nsteps times each
    shape's measured FP64 ms/step).  Requests whose cells * steps are at
    most keep_cells_steps keep their final state (the first max_kept of
    them in arrival order), and are run again alone in FP64 (state_fp64),
    for a check of both against the FP64 oracle."""
    reqs = draw_requests(rate_hz, seed=seed, nsteps=nsteps, count=requests)
    kept = 0
    for r in reqs:
        r.keep = r.cells * r.nsteps <= keep_cells_steps and kept < max_kept
        kept += r.keep
    srv = Server(dev_index, slots=slots, max_batch=max_batch)
    tw = time.perf_counter()
    srv.warm(r.shape for r in reqs)
    warm_s = time.perf_counter() - tw
    try:
        srv.run(reqs, progress=progress)
    finally:
        srv.close()
    out = summarize(reqs)
    out.update({"arrival_rate_hz": rate_hz, "steps_per_case": nsteps, "slots": slots, "mg_precision": 1,
                "max_batch": max_batch, "shape_buckets": {"2D": list(BUCKETS_2D), "3D": list(BUCKETS_3D)},
                "warmup_s": warm_s})
    # the kept requests once more, alone and all-FP64 on a plain context:
    # the state the FP64 path gives for the same request, beside the served one
    for r in reqs:
        if r.keep:
            r.state_fp64 = run_alone(dev_index, r, 0)
    cell_steps = sum(r.cells * r.nsteps for r in reqs)
    if sequential == "run":
        wall = run_sequential(dev_index, reqs, 0)
        out["fp64_sequential"] = {"kind": "run", "wall_s": wall, "cell_steps_per_s": cell_steps / wall}
    elif sequential == "sampled":
        ms = step_times(dev_index, [(r.dim, r.n) for r in reqs])
        wall = sum(r.nsteps * ms[(r.dim, r.n)] for r in reqs) / 1e3
        out["fp64_sequential"] = {"kind": "sampled: sum of nsteps x the shape's FP64 ms/step (10 steps alone)",
                                  "wall_s": wall, "cell_steps_per_s": cell_steps / wall}
    if "fp64_sequential" in out:
        out["speedup_over_fp64_sequential"] = out["aggregate_cell_steps_per_s"] / out["fp64_sequential"]["cell_steps_per_s"]
    return out, reqs
