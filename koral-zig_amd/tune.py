"""Occupancy autotuner for the step's heavy kernels.

Each heavy kernel carries an `amdgpu_waves_per_eu` bound (KZ_WPE_* in
csrc/kz_sim_dev.h): a larger bound caps the registers per lane so that more
waves share a SIMD and hide each other's latency, at the price of spills.
The best bound depends on the kernel, the grid and the compiler, so it is
measured; so are the x width of the face flux kernel's y/z face tiles
(KZ_FT_T) and the width of the persistent implicit launches (blocks of
one wave per CU, KZ_PERSIST_PER_CU), the latter on the independent-cell
batch.  This module builds one library per candidate (only one knob
differs from the defaults in each), times every kernel class of the PUFFY
step on the GPU with HIP events in a subprocess per library, requires every
candidate to leave the state bit for bit where the default build leaves it,
and reports the fastest value per knob: the default unless another is
faster by at least --min-gain (3 %; repeated A/B runs of one build differ by
up to 2-3 % per kernel class and by up to 8 % on the batch, so a smaller
margin picks noise).

    python koral-zig_amd/tune.py [--jobs N] [--nx 384 --ny 360] [--out DIR]

from the repository root, on a machine with the GPU.  The two halves can
also run apart: `--build-only` compiles the libraries into --out (on any
machine with hipcc), `--prebuilt` measures the libraries found there.  The
defaults in kz_sim_dev.h are the values this tuner selected.
"""
import argparse
import concurrent.futures as cf
import hashlib
import json
import os
import subprocess
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent

# macro -> the kernel class (hip.KERNELS name) it governs, its default and
# the candidates tried: the waves-per-EU bounds, and the x width of the face
# flux kernel's y/z face tiles (KZ_FT_T columns by 64 / KZ_FT_T cells)
KNOBS = {
    "KZ_WPE_FACE": ("face_flux", 2, (1, 2, 3)),
    "KZ_FT_T": ("face_flux", 8, (4, 8, 16)),
    "KZ_WPE_UPDATE": ("update", 2, (1, 2, 3)),
    "KZ_WPE_U2P": ("u2p", 2, (1, 2, 3)),
    "KZ_WPE_WS": ("wavespeeds", 1, (1, 2, 4)),
    "KZ_WPE_PREP": ("implicit", 2, (1, 2, 3)),
    "KZ_WPE_RUNG0": ("implicit", 1, (1, 2)),
}
CLASSES = ("face_flux", "implicit", "u2p", "update", "wavespeeds", "ct", "bc")
# the persistent implicit launches' width (blocks of one wave per CU): timed
# on the independent-cell batch, where those launches run
BATCH_KNOBS = {"KZ_PERSIST_PER_CU": (4, (2, 4, 8))}


def variants():
    """(name, {macro: value}) for the default build and each single change."""
    out = [("default", {})]
    for macro, (cls, default, cands) in KNOBS.items():
        for v in cands:
            if v != default:
                out.append((f"{macro}={v}", {macro: v}))
    for macro, (default, cands) in BATCH_KNOBS.items():
        for v in cands:
            if v != default:
                out.append((f"{macro}={v}", {macro: v}))
    return out


def build_variant(name, defs, outdir):
    sys.path.insert(0, str(HERE))
    import build as B
    Path(outdir).mkdir(parents=True, exist_ok=True)
    lib = Path(outdir) / ("libkz_hip_" + name.replace("=", "_") + ".so")
    return B.compile_hip(lib, defs)


def measure(lib, nx, ny, steps):
    """Per-class kernel ms per step, ms per step and a hash of the state
    after a fixed number of steps, with `lib` loaded in place of the built
    library."""
    import importlib.util
    name = "koral_zig_amd"
    spec = importlib.util.spec_from_file_location(name, HERE / "__init__.py", submodule_search_locations=[str(HERE)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    from koral_zig_amd import abi_types as T
    from koral_zig_amd import hip, host
    hip.LIB_PATH = Path(lib)
    grid = T.grid_puffy(nx, ny)
    phys = T.physics_puffy()
    tab = host.metric_tables(grid)
    sim = hip.Sim(grid, phys, tab)
    host.init_puffy(sim, grid, tab, phys, seed=0x4B5A4D493335355F)
    sim.step_many(3)
    sim.sync()
    per = {}
    for k in CLASSES:
        sim.time_kernel(k)
        sim.step_many(2)
        ms, _ = sim.kernel_time()
        per[k] = ms / 2
    sim.time_kernel(None)
    t0 = time.perf_counter()
    sim.step_many(steps)
    sim.sync()
    wall = time.perf_counter() - t0
    p, u, f = sim.get_state()
    h = hashlib.sha1(p.tobytes() + u.tobytes() + f.tobytes()).hexdigest()[:16]
    sim.close()
    return {"per": per, "ms_per_step": wall / steps * 1e3, "hash": h}


def measure_batch(lib, cells, solves):
    """ms per solve of the independent-cell implicit batch and a hash of its
    results, with `lib` loaded in place of the built library."""
    import importlib.util
    import numpy as np
    name = "koral_zig_amd"
    spec = importlib.util.spec_from_file_location(name, HERE / "__init__.py", submodule_search_locations=[str(HERE)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    from koral_zig_amd import abi_types as T
    from koral_zig_amd import hip, workloads
    hip.LIB_PATH = Path(lib)
    phys = T.physics_puffy()
    pp, geom, dt = workloads.implicit_batch_inputs(
        cells, 0x4B5A4D493335355F, phys, lambda p, g: hip.radstate(p, g, phys.gam, phys.opac, 0)[:, 9])
    b = hip.ImplicitBatch(cells, phys)
    b.set(pp, geom, dt)
    b.solve()
    ms = float(np.median([b.solve() for _ in range(solves)]))
    uu, pq, res, cnt = b.get()
    b.close()
    h = hashlib.sha1(uu.tobytes() + pq.tobytes() + res["iters"].tobytes()).hexdigest()[:16]
    return {"ms_per_solve": ms, "hash": h, "counts": list(cnt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--nx", type=int, default=384)
    ap.add_argument("--ny", type=int, default=360)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--out", default=str(HERE / "lib" / "tune"))
    ap.add_argument("--build-only", action="store_true", help="compile the variants into --out and stop")
    ap.add_argument("--prebuilt", action="store_true", help="measure the variants already compiled in --out")
    ap.add_argument("--cells", type=int, default=1 << 20, help="cells of the batch that times the persistent width")
    ap.add_argument("--min-gain", type=float, default=0.03, help="relative gain a candidate needs over the default")
    ap.add_argument("--measure", help=argparse.SUPPRESS)
    ap.add_argument("--measure-batch", help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.measure:
        print(json.dumps(measure(a.measure, a.nx, a.ny, a.steps)), flush=True)
        return
    if a.measure_batch:
        print(json.dumps(measure_batch(a.measure_batch, a.cells, 3)), flush=True)
        return
    Path(a.out).mkdir(parents=True, exist_ok=True)
    vs = variants()
    t0 = time.perf_counter()
    libs = {}
    if a.prebuilt:
        for n, _ in vs:
            libs[n] = Path(a.out) / ("libkz_hip_" + n.replace("=", "_") + ".so")
            if not libs[n].exists():
                raise RuntimeError(f"{libs[n]} missing: run with --build-only first")
    else:
        with cf.ThreadPoolExecutor(a.jobs) as ex:
            futs = {ex.submit(build_variant, n, d, a.out): n for n, d in vs}
            for fu in cf.as_completed(futs):
                libs[futs[fu]] = fu.result()
                print(f"  built {futs[fu]} ({time.perf_counter() - t0:.0f}s)", flush=True)
        libs = {n: libs[n] for n, _ in vs}
        print(f"built {len(libs)} variants in {time.perf_counter() - t0:.0f}s", flush=True)
        if a.build_only:
            return
    res = {}
    for name, lib in libs.items():
        r = subprocess.run([sys.executable, str(Path(__file__)), "--measure", str(lib), "--nx", str(a.nx), "--ny",
                            str(a.ny), "--steps", str(a.steps)], capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"measuring {name} failed:\n{r.stderr[-2000:]}")
        res[name] = json.loads(r.stdout.strip().splitlines()[-1])
        print(f"{name:18s} {res[name]['ms_per_step']:.4f} ms/step " +
              " ".join(f"{k} {v:.4f}" for k, v in res[name]["per"].items()) + f" {res[name]['hash']}", flush=True)
    bres = {}
    for name in ["default"] + [f"{m}={v}" for m, (d, cs) in BATCH_KNOBS.items() for v in cs if v != d]:
        r = subprocess.run([sys.executable, str(Path(__file__)), "--measure-batch", str(libs[name]), "--cells",
                            str(a.cells)], capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError(f"measuring the batch with {name} failed:\n{r.stderr[-2000:]}")
        bres[name] = json.loads(r.stdout.strip().splitlines()[-1])
        print(f"batch {name:20s} {bres[name]['ms_per_solve']:.3f} ms/solve {bres[name]['hash']}", flush=True)
    bad = [n for n, r in bres.items() if r["hash"] != bres["default"]["hash"]]
    if bad:
        raise RuntimeError(f"variants changed the batch results: {bad}")
    ref = res["default"]["hash"]
    bad = [n for n, r in res.items() if r["hash"] != ref]
    if bad:
        raise RuntimeError(f"variants changed the state bits: {bad}")
    print("every variant leaves the state bit for bit as the default build")
    pick = {}
    for macro, (col, default, cands) in KNOBS.items():
        times = {default: res["default"]["per"][col]}
        for v in cands:
            if v != default:
                times[v] = res[f"{macro}={v}"]["per"][col]
        best = min(times, key=times.get)
        if times[best] > (1.0 - a.min_gain) * times[default]:  # keep the default within the timing noise
            best = default
        pick[macro] = best
        print(f"{macro:14s} default {default}: " + ", ".join(f"{v} -> {t:.4f} ms" for v, t in sorted(times.items())) +
              f"  best {best}")
    for macro, (default, cands) in BATCH_KNOBS.items():
        times = {v: bres["default" if v == default else f"{macro}={v}"]["ms_per_solve"] for v in cands}
        best = min(times, key=times.get)
        if times[best] > (1.0 - a.min_gain) * times[default]:
            best = default
        pick[macro] = best
        print(f"{macro:14s} default {default}: " + ", ".join(f"{v} -> {t:.3f} ms" for v, t in sorted(times.items())) +
              f"  best {best}")
    print(json.dumps({"selected": pick}))


if __name__ == "__main__":
    main()
