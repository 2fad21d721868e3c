"""A run server: many independent PUFFY runs resident on one GPU.

Runs arrive over time (Poisson), are admitted first come first served while
the device memory of the resident runs stays within an HBM budget, and are
advanced in batches of CFL steps captured once per run into a HIP graph
(kz_sim_capture_steps) and replayed on the run's own stream, so that small
runs share the GPU instead of queueing behind each other.  A run keeps its
whole state resident until it has taken its steps, then frees it.

Timing: the latency of a batch is the time from its graph launch to the
poll that sees it finished, divided by its steps; the aggregate throughput
is the zone-cycles of all runs over the wall time from the first arrival to
the last completion.  Admission work (device allocation and table upload,
the torus init sequence on the GPU, graph capture) runs on a second thread
inside that wall time.  The host-only part of a run's set-up (metric tables
and prepinit, seconds per run at the larger grids) is computed before the
clock starts, by a pool of threads.

The sequential baseline runs the same runs one after another, each as a
plain eager step loop (kz_step_many) on one stream.
"""
import collections
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import abi_types as T
from . import hip, host

C4_SIZES = ((64, 64), (128, 96), (192, 192), (256, 256), (384, 360))


class RunSpec:
    def __init__(self, rid, nx, ny, spin, seed, nsteps, arrival):
        self.rid, self.nx, self.ny, self.spin = rid, nx, ny, spin
        self.seed, self.nsteps, self.arrival = seed, nsteps, arrival

    def grid(self):
        return T.grid_puffy(self.nx, self.ny, rmin=T.rmin_for_spin(self.spin))

    def phys(self):
        return T.physics_puffy(bh_spin=self.spin)


def poisson_specs(n_runs, rate, nsteps, seed, sizes=C4_SIZES, spin_max=0.95):
    """n_runs arrivals of a Poisson process of `rate` per second (the first at
    t=0), grid sizes uniform over `sizes`, spins uniform over [0, spin_max],
    per-run perturbation seeds derived from `seed`."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_runs))
    t -= t[0]
    pick = rng.integers(0, len(sizes), n_runs)
    spin = rng.uniform(0.0, spin_max, n_runs)
    return [RunSpec(i, *sizes[pick[i]], float(spin[i]), (int(seed) + 1 + i) & 0xFFFFFFFFFFFFFFFF, nsteps, float(t[i]))
            for i in range(n_runs)]


def _prep(spec):
    """Host-only set-up of one run: metric tables and prepinit primitives."""
    grid = spec.grid()
    tab = host.metric_tables(grid, dict(host.MP_PUFFY, a=spec.spin))
    p0 = host.puffy_prepinit(grid, tab, spec.phys())
    return spec.rid, tab, p0


def prepare(specs, workers, log=None):
    """{rid: (tables, prepinit p)} for every spec, computed by `workers`
    threads (the host library's calls release the interpreter lock), or in
    the calling thread when workers <= 1."""
    out = {}
    t0 = time.perf_counter()
    order = sorted(specs, key=lambda s: -s.nx * s.ny)  # largest first
    pool = ThreadPoolExecutor(workers) if workers > 1 else None
    try:
        it = pool.map(_prep, order) if pool else map(_prep, order)
        for i, (rid, tab, p0) in enumerate(it):
            out[rid] = (tab, p0)
            if log and (i + 1) % 8 == 0:
                log(f"  prepared {i + 1}/{len(specs)} runs ({time.perf_counter() - t0:.1f}s)")
    finally:
        if pool is not None:
            pool.shutdown()
    return out


def start_run(spec, prepared, capture_steps=0, capture_lock=None):
    """A resident run in its initial state: the torus init sequence on the
    GPU from the prepared tables and prepinit, the run's seeded perturbation,
    and (capture_steps > 0) its step graph, captured while holding
    `capture_lock` when one is given."""
    tab, p0 = prepared
    grid, phys = spec.grid(), spec.phys()
    sim = hip.Sim(grid, phys, tab)
    try:
        host.init_puffy(sim, grid, tab, phys, seed=spec.seed, p0=p0)
        if capture_steps:
            if capture_lock is None:
                sim.capture_steps(capture_steps)
            else:
                with capture_lock:
                    sim.capture_steps(capture_steps)
    except Exception:
        sim.close()
        raise
    return sim


def footprints(specs, prepared):
    """Device bytes of a run of each grid size among `specs`, measured by
    creating (and freeing) one sim of that size."""
    out = {}
    for s in specs:
        if (s.nx, s.ny) not in out:
            sim = hip.Sim(s.grid(), s.phys(), prepared[s.rid][0])
            out[(s.nx, s.ny)] = sim.device_bytes()
            sim.close()
    return out


class _Resident:
    __slots__ = ("spec", "sim", "zones", "bytes", "done", "launched")

    def __init__(self, spec, sim, nbytes):
        self.spec, self.sim = spec, sim
        self.zones = spec.nx * spec.ny
        self.bytes = nbytes
        self.done = 0
        self.launched = None


class RunServer:
    """Serve `specs` to completion (see the module docstring).  `hbm_budget`
    bounds the summed device bytes of the resident runs; a run that would
    exceed it waits, except when nothing else is resident.

    grouped=False: each run replays its own captured batch on its own
    stream, as soon as its previous batch has finished.
    grouped=True: the resident runs form one run group (hip.Group) and
    advance together, batch by batch, by one captured graph of the whole
    group; the implicit cells of all of them are solved by one set of
    implicit launches per implicit pass.  Runs join and leave at batch
    boundaries, where the group's graph is captured again."""

    def __init__(self, hbm_budget, batch_steps=10, log=None, grouped=False):
        self.budget = int(hbm_budget)
        self.batch = int(batch_steps)
        self.log = log or (lambda *a: None)
        self.grouped = bool(grouped)

    def serve(self, specs, prepared, on_finish=None):
        """Returns the wall time, zone-cycles, batch latencies per step (ms),
        peak resident runs and bytes.  on_finish(spec, sim) sees each run's
        final state before its memory is freed."""
        for s in specs:
            if s.nsteps % self.batch:
                raise ValueError(f"run {s.rid}: {s.nsteps} steps is not a multiple of the batch {self.batch}")
        if self.grouped:
            return self._serve_grouped(specs, prepared, on_finish)
        need = footprints(specs, prepared)
        arrivals = collections.deque(sorted(specs, key=lambda s: s.arrival))
        lock = threading.Lock()
        # a graph capture on one thread and a device-wide operation on the
        # other (freeing a finished run's memory) must not overlap
        hip_lock = threading.Lock()
        ready = collections.deque()  # admitted, initialised runs for the stepping loop
        state = {"used": 0, "live": 0, "error": None, "stop": False}
        freed = threading.Condition(lock)
        t0 = time.perf_counter()

        def admit():
            try:
                while arrivals and not state["stop"]:
                    s = arrivals[0]
                    delay = s.arrival - (time.perf_counter() - t0)
                    if delay > 0:
                        with freed:
                            if not state["stop"]:
                                freed.wait(delay)
                        continue
                    b = need[(s.nx, s.ny)]
                    with freed:
                        while state["live"] and state["used"] + b > self.budget and not state["stop"]:
                            freed.wait(0.05)
                        if state["stop"]:
                            return
                        state["used"] += b
                        state["live"] += 1
                    arrivals.popleft()
                    sim = start_run(s, prepared.pop(s.rid), self.batch, hip_lock)
                    r = _Resident(s, sim, b)
                    with lock:
                        ready.append(r)
            except BaseException as e:  # surfaced by the stepping loop
                with lock:
                    state["error"] = e

        th = threading.Thread(target=admit, daemon=True)
        th.start()
        resident = []
        lat = []
        zc = 0
        peak_n = peak_b = 0
        n_done = 0
        last_log = t0
        try:
            while True:
                with lock:
                    if state["error"] is not None:
                        break
                    while ready:
                        resident.append(ready.popleft())
                    peak_n = max(peak_n, len(resident))
                    peak_b = max(peak_b, state["used"])
                if not resident and not th.is_alive() and not ready:
                    break
                for r in resident:
                    if r.launched is None:
                        r.sim.replay()
                        r.launched = time.perf_counter()
                finished = []
                for r in resident:
                    if not r.sim.busy():
                        now = time.perf_counter()
                        lat.append((now - r.launched) * 1e3 / self.batch)
                        r.launched = None
                        r.done += self.batch
                        zc += r.zones * self.batch
                        if r.done >= r.spec.nsteps:
                            finished.append(r)
                for r in finished:
                    resident.remove(r)
                    with hip_lock:
                        r.sim.sync()  # raises on a device-side error of the run
                        if on_finish is not None:
                            on_finish(r.spec, r.sim)
                        r.sim.close()
                    n_done += 1
                    with freed:
                        state["used"] -= r.bytes
                        state["live"] -= 1
                        freed.notify_all()
                now = time.perf_counter()
                if now - last_log > 5.0:
                    last_log = now
                    self.log(f"  t={now - t0:.1f}s resident={len(resident)} done={n_done}/{len(specs)} "
                             f"zone-cycles={zc:.3e}")
                if not finished:
                    time.sleep(1e-4)
            wall = time.perf_counter() - t0
        finally:
            with freed:
                state["stop"] = True
                freed.notify_all()
            th.join()
            with lock:
                resident.extend(ready)
            with hip_lock:
                for r in resident:
                    r.sim.close()
        if state["error"] is not None:
            raise state["error"]
        return {"wall_s": wall, "zone_cycles": zc, "latency_ms": np.asarray(lat),
                "peak_resident": peak_n, "peak_bytes": peak_b, "runs": n_done}


    def _serve_grouped(self, specs, prepared, on_finish):
        need = footprints(specs, prepared)
        arrivals = collections.deque(sorted(specs, key=lambda s: s.arrival))
        lock = threading.Lock()
        # run creation and init (admission thread) never overlap a group
        # capture or a device-wide free (stepping thread)
        hip_lock = threading.Lock()
        ready = collections.deque()
        state = {"used": 0, "live": 0, "error": None, "stop": False}
        freed = threading.Condition(lock)
        t0 = time.perf_counter()

        def admit():
            try:
                while arrivals and not state["stop"]:
                    s = arrivals[0]
                    delay = s.arrival - (time.perf_counter() - t0)
                    if delay > 0:
                        with freed:
                            if not state["stop"]:
                                freed.wait(delay)
                        continue
                    b = need[(s.nx, s.ny)]
                    with freed:
                        while state["live"] and state["used"] + b > self.budget and not state["stop"]:
                            freed.wait(0.05)
                        if state["stop"]:
                            return
                        state["used"] += b
                        state["live"] += 1
                    arrivals.popleft()
                    with hip_lock:
                        sim = start_run(s, prepared.pop(s.rid))
                    with lock:
                        ready.append(_Resident(s, sim, b))
            except BaseException as e:  # surfaced by the stepping loop
                with lock:
                    state["error"] = e
                    freed.notify_all()

        th = threading.Thread(target=admit, daemon=True)
        th.start()
        grp = hip.Group()
        members = []
        launched = None
        lat = []
        zc = 0
        peak_n = peak_b = 0
        n_done = 0
        regroups = 0
        last_log = t0
        try:
            while True:
                with lock:
                    if state["error"] is not None:
                        break
                changed = False
                if launched is not None:
                    if grp.busy():
                        time.sleep(1e-4)
                        continue
                    now = time.perf_counter()
                    finished = []
                    for r in members:
                        lat.append((now - launched) * 1e3 / self.batch)
                        r.done += self.batch
                        zc += r.zones * self.batch
                        if r.done >= r.spec.nsteps:
                            finished.append(r)
                    launched = None
                    if finished:
                        members = [r for r in members if r not in finished]
                        with hip_lock:
                            grp.set_runs([])  # waits for the batch; the group no longer refers to them
                            for r in finished:
                                r.sim.sync()  # raises on a device-side error of the run
                                if on_finish is not None:
                                    on_finish(r.spec, r.sim)
                                r.sim.close()
                        n_done += len(finished)
                        with freed:
                            for r in finished:
                                state["used"] -= r.bytes
                                state["live"] -= 1
                            freed.notify_all()
                        changed = True
                with lock:
                    joining = list(ready)
                    ready.clear()
                if joining:
                    members.extend(joining)
                    changed = True
                peak_n = max(peak_n, len(members))
                with lock:
                    peak_b = max(peak_b, state["used"])
                if not members:
                    if not th.is_alive():
                        with lock:
                            if not ready:
                                break
                    time.sleep(1e-4)
                    continue
                if changed:
                    with hip_lock:
                        grp.set_runs([r.sim for r in members])
                        grp.capture_steps(self.batch)
                    regroups += 1
                grp.replay()
                launched = time.perf_counter()
                now = launched
                if now - last_log > 5.0:
                    last_log = now
                    self.log(f"  t={now - t0:.1f}s resident={len(members)} done={n_done}/{len(specs)} "
                             f"zone-cycles={zc:.3e} regroups={regroups}")
            wall = time.perf_counter() - t0
        finally:
            with freed:
                state["stop"] = True
                freed.notify_all()
            th.join()
            with lock:
                members.extend(ready)
            with hip_lock:
                grp.close()  # waits for its stream
                for r in members:
                    r.sim.close()
        if state["error"] is not None:
            raise state["error"]
        return {"wall_s": wall, "zone_cycles": zc, "latency_ms": np.asarray(lat),
                "peak_resident": peak_n, "peak_bytes": peak_b, "runs": n_done, "regroups": regroups}


def serve_sequential(specs, prepared, log=None, on_finish=None):
    """The same runs one after another, each a plain eager loop of CFL steps
    on its own sim; returns the summed stepping time (init excluded), the
    wall time including each run's init, and the zone-cycles."""
    step_s = 0.0
    zc = 0
    t0 = time.perf_counter()
    for i, s in enumerate(sorted(specs, key=lambda s: s.arrival)):
        sim = start_run(s, prepared.pop(s.rid))
        try:
            sim.sync()
            t1 = time.perf_counter()
            sim.step_many(s.nsteps)
            sim.sync()
            step_s += time.perf_counter() - t1
            if on_finish is not None:
                on_finish(s, sim)
        finally:
            sim.close()
        zc += s.nx * s.ny * s.nsteps
        if log and (i + 1) % 16 == 0:
            log(f"  sequential {i + 1}/{len(specs)} runs ({time.perf_counter() - t0:.1f}s)")
    return {"step_s": step_s, "wall_s": time.perf_counter() - t0, "zone_cycles": zc}
