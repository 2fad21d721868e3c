#!/usr/bin/env python3
"""zone-cycles/s of the full RMHD+M1 step (PPM+LAXF, flux-CT, u2p, M1
implicit four-force) of a 384x360 PUFFY torus on one MI355X.

Workload: the seeded-perturbed PUFFY torus (M=10 Msun, spin 0, MKS2, 384x360
+ 3 ghosts, radiative viscosity and dynamo off), W warmup CFL steps, then K
timed CFL steps enqueued back to back on the sim's stream with the timestep
kept on the device.  The state is resident in HBM before the clock starts.

Prints ONE JSON line (see the keys below).  Beside the throughput it reports
  roofline      the dominant kernel class, timed live with HIP events on the
                sim's stream over the timed region, against MI355X's peak;
  cpu_baseline  the oracle (the CPU restatement, one thread) timed on this
                host over a bounded sample of the same workload;
  p50/p99 step latency over a short run of individually synchronised steps.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

SEED = 0x4B5A4D493335355F  # default perturbation seed of the reference's puffy runs
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# Algorithmic (compulsory) HBM bytes per launch unit of each HBM-bound kernel
# class, 2D, FP64 (13 slots = 104 B per state):
#  face_flux (per dim, per face): read p of the two cells' new stencil cell
#    104 + face geometry g,G symmetric 20 + gdet/gttpert 2 -> 22*8 = 176
#    + two wavespeeds 16 + write 13 fluxes 104                 = 400 B
#  update (per cell): read u 104 + p 104 + centre geometry 176 + 40 unique
#    Christoffels 320 + write u 104 + face fluxes: each face is shared by
#    the two cells beside it, so a cell owns one face per dimension,
#    2*104 in 2D                                                = 1016 B
#  u2p (per cell): read u 104 + p 104 + geometry 176 + write p 104
#    + 4 flags 16                                               = 504 B
BYTES_PER_UNIT = {"face_flux": 400.0, "update": 1016.0, "u2p": 504.0}
# The whole explicit operator (wavespeeds, face fluxes, EMF, update) per
# active zone: p 104 + u read 104 + u write 104 + centre geometry 184
# (symmetric g and G, gdet, alpha, gttpert) + 40 Christoffels 320 + face
# geometry 184 per active dimension = 1184 B in 2D, 1368 B in 3D; halo
# re-reads and face-array round trips are not compulsory.
EXPLICIT_OP_BYTES = {2: 1184.0, 3: 1368.0}
# FP64 operations of one implicit cell-solve, counted by the oracle built
# with a counting double (oracle/flops: +1 per + - * / sqrt, +2 per fused
# multiply-add, transcendentals counted through their own operations).  `python oracle/flops/count.py c1`
# over the 136 704 implicit cells of the 384x360 step after its first step:
# 28 135.0 per solve on average; least-squares fit over the cells
# flops = 14 588.6 + 9 854.7 * Newton iterations (rms 1.8e3).  The bench
# multiplies the fit by the solves and iterations the device counts.
IMPLICIT_FLOPS_PER_SOLVE = 14588.6
IMPLICIT_FLOPS_PER_ITER = 9854.7
# The independent-cell batch spends most of its operations in the later
# rungs and in failing ladders, which the fit above does not describe:
# `python oracle/flops/count.py c2` over 200 000 cells of the same generator
# counts 134 811.4 operations per cell (first-rung solves 55 463.7, later
# rungs 686 479.6, failed cells 2 631 508.7).
BATCH_FLOPS_PER_CELL = 134811.4
# HBM bytes per launch of a kernel class at 384x360, from rocprofv3 runs of
# FETCH_SIZE and WRITE_SIZE on this benchmark, each scaled by what the same
# counter reads for a copy kernel of known size in the same run (k_soa2aos:
# fetch x1.132, write x1.000).  Only valid for the default 384x360 workload.
TRAFFIC_BYTES_PER_LAUNCH_C1 = {"implicit": 1.68e8, "update": 1.39e8, "face_flux": 1.33e8, "u2p": 5.78e7}


def _pkg():
    import importlib.util
    name = "koral_zig_amd"
    if name not in sys.modules:
        pkg = ROOT / "koral-zig_amd"
        spec = importlib.util.spec_from_file_location(name, pkg / "__init__.py", submodule_search_locations=[str(pkg)])
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    """Threads for the all-cores CPU baseline: OMP_NUM_THREADS when set (the
    host's CPU share), else the CPUs this process may run on."""
    n = os.environ.get("OMP_NUM_THREADS")
    if n and n.isdigit() and int(n) > 0:
        return int(n)
    return len(os.sched_getaffinity(0))


def cpu_baseline(grid, phys, tab, p, u, clock, steps, threads=1):
    """The oracle over `steps` CFL steps from the GPU's post-warmup state on
    `threads` host threads (bitwise the serial result): zone-cycles/s."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import kzo
    kzo.set_threads(threads)
    sim = kzo.Sim(grid, phys, tab)
    sim.set_state(p, u)
    sim.set_time(clock["t"], clock["tstepdenmax"], clock["tstepdenmin"])
    per = []
    for _ in range(steps):
        t0 = time.perf_counter()
        sim.step()
        per.append(time.perf_counter() - t0)
    kzo.set_threads(1)
    return grid.nx * grid.ny * grid.nz / float(np.median(per)), float(np.sum(per))


def bench_implicit_batch(args):
    """The independent-cell implicit-solve batch (2^22 cells by default):
    cell-solves/s of the device-resident compacted solve, inputs resident in
    HBM; the oracle times a bounded sample of the same cells."""
    from koral_zig_amd import abi_types as T
    from koral_zig_amd import hip, workloads

    phys = T.physics_puffy()
    n = args.cells
    t0 = time.perf_counter()
    pp, geom, dt = workloads.implicit_batch_inputs(
        n, args.seed, phys, lambda p, g: hip.radstate(p, g, phys.gam, phys.opac, 0)[:, 9])
    b = hip.ImplicitBatch(n, phys)
    b.set(pp, geom, dt)
    log(f"setup {time.perf_counter() - t0:.1f}s ({n} cells, seed {args.seed:#x})")
    for _ in range(max(args.warmup, 1)):
        b.solve()
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kms.append(b.solve())
    wall = time.perf_counter() - t0
    _, _, res, (nfail, iters, solves) = b.get()
    value = n * args.steps / wall
    avg_ms = float(np.mean(kms))
    flops = BATCH_FLOPS_PER_CELL * n
    achieved = flops / (avg_ms * 1e-3) / 1e12
    log(f"timed {args.steps} solves in {wall:.3f}s: {value:.4e} cell-solves/s; kernels {avg_ms:.3f} ms/solve; "
        f"ok {solves}/{n}, {iters / max(solves, 1):.2f} iterations per solve")
    mixed = None
    if args.mixed:
        # the same cells with the FP32-Jacobian Newton, against the FP64 result
        _, p64, r64, _ = b.get()
        b.close()
        mp = T.physics_puffy()
        mp.implicit.jacobian_fp32 = 1
        b = hip.ImplicitBatch(n, mp)
        b.set(pp, geom, dt)
        for _ in range(max(args.warmup, 1)):
            b.solve()
        kms32 = []
        t1 = time.perf_counter()
        for _ in range(args.steps):
            kms32.append(b.solve())
        wall32 = time.perf_counter() - t1
        _, p32, r32, (nfail32, iters32, solves32) = b.get()
        ok64, ok32 = r64["ok"] == 1, r32["ok"] == 1
        both = ok64 & ok32 & (r64["rung"] == r32["rung"])
        rel = np.abs(p64[both] - p32[both]) / np.maximum(np.maximum(np.abs(p64[both]), np.abs(p32[both])), 1e-300)
        cell_max = rel.max(1)
        v32 = n * args.steps / wall32
        mixed = {
            "value": v32, "unit": "cell-solves/s", "speedup": v32 / value, "ms_per_solve": float(np.mean(kms32)),
            "ok_identical": float((ok64 == ok32).mean()), "rung_identical": float((r64["rung"] == r32["rung"]).mean()),
            "iters_identical": float((r64["iters"] == r32["iters"]).mean()),
            "iterations_per_solve": iters32 / max(solves32, 1), "failures": nfail32,
            "field_rel_err_vs_fp64": {
                "cells": int(both.sum()),
                "max_per_field": [float(x) for x in rel.max(0)],
                "p99_per_field": [float(x) for x in np.percentile(rel, 99, axis=0)],
                "cell_max_median": float(np.median(cell_max)),
                "frac_cells_le_1e-10": float((cell_max <= 1e-10).mean()),
                "frac_cells_le_1e-6": float((cell_max <= 1e-6).mean())},
        }
        log(f"FP32-Jacobian: {v32:.4e} cell-solves/s ({v32 / value:.3f}x), ok identical {mixed['ok_identical']:.4f}, "
            f"rung identical {mixed['rung_identical']:.4f}, median per-cell max rel err "
            f"{mixed['field_rel_err_vs_fp64']['cell_max_median']:.1e}")
    b.close()
    cpu = None
    if args.cpu_cells > 0:
        sys.path.insert(0, str(ROOT / "oracle"))
        import kzo
        m = min(args.cpu_cells, n)
        t1 = time.perf_counter()
        kzo.solve_implicit(pp[:m], geom[:m], dt[:m], phys.gam, phys.rad, phys.opac, phys.implicit)
        secs = time.perf_counter() - t1
        cpu = {"value": m / secs, "unit": "cell-solves/s", "cores": 1, "kind": "port",
               "sample": f"the first {m} cells of the same batch ({secs:.1f}s, one thread)"}
        log(f"cpu baseline {m / secs:.4e} cell-solves/s over {secs:.1f}s")
    out = {
        "metric": "cell-solves/s implicit radiation-gas four-force batch on 1 GPU",
        "value": value, "unit": "cell-solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"independent-cell implicit batch, {n} cells (rho, T_gas, E_hat/u, |F|/E, tau, "
                               f"MKS2 a=0 r in [2.5, 50] spread as specified in koral_zig_amd.workloads)",
                   "cells": n, "seed": args.seed},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "kernel": "implicit",
                     "avg_launch_ms": avg_ms, "launches": len(kms)},
        "cpu_baseline": cpu,
        "mixed_precision": mixed,
        "solver": {"failures": nfail, "iterations": iters, "solves": solves,
                   "rungs": np.bincount(res["rung"][res["ok"] == 1], minlength=6).tolist()},
    }
    print(json.dumps(out), flush=True)


def bench_run_server(args):
    """Many PUFFY runs served concurrently on one GPU: Poisson arrivals,
    admission by device footprint against an HBM budget, HIP-graph batches of
    CFL steps on per-run streams (or, grouped, one graph of all resident runs
    with their implicit cells in shared launches).  value = zone-cycles of all runs
    over the wall time from the first arrival to the last completion; the
    same runs stepped one after another with the plain eager loop give the
    sequential figure."""
    from koral_zig_amd import abi_types as T
    from koral_zig_amd import server

    specs = server.poisson_specs(args.runs, args.rate, args.run_steps, args.seed)
    workers = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    log(f"run server: {len(specs)} runs, preparing host set-up in {workers} threads")
    prep = server.prepare(specs, workers, log=log)
    log(f"host set-up {time.perf_counter() - t0:.1f}s (metric tables + prepinit, before the clock)")
    seq_prep = {k: v for k, v in prep.items()}  # the sequential pass consumes its own copy
    grp_prep = {k: v for k, v in prep.items()} if args.server_mode == "both" else None
    budget = int(args.hbm_gib * (1 << 30))

    def serve(grouped, pr):
        srv = server.RunServer(budget, batch_steps=args.batch, log=log, grouped=grouped)
        r = srv.serve(specs, pr)
        v = r["zone_cycles"] / r["wall_s"]
        la = r["latency_ms"]
        log(f"served {r['runs']} runs ({'one run group' if grouped else 'per-run streams'}) in {r['wall_s']:.2f}s: "
            f"{v:.4e} zone-cycles/s, peak resident {r['peak_resident']} ({r['peak_bytes'] / 2**30:.2f} GiB); step "
            f"latency p50 {np.percentile(la, 50):.3f} ms p99 {np.percentile(la, 99):.3f} ms"
            + (f"; {r['regroups']} regroups" if grouped else ""))
        return r, v, la

    res, value, lat = serve(args.server_mode == "grouped", prep)
    grouped = None
    if grp_prep is not None:
        gr, gv, gl = serve(True, grp_prep)
        grouped = {"value": gv, "wall_s": gr["wall_s"], "vs_per_run_streams": gv / value,
                   "step_latency_ms": {"p50": float(np.percentile(gl, 50)), "p99": float(np.percentile(gl, 99)),
                                       "n": int(gl.size)},
                   "peak_resident": gr["peak_resident"], "regroups": gr["regroups"]}
    seq = server.serve_sequential(specs, seq_prep, log=log)
    seq_v = seq["zone_cycles"] / seq["step_s"]
    log(f"sequential eager: stepping {seq['step_s']:.2f}s ({seq_v:.4e} zone-cycles/s), {seq['wall_s']:.2f}s "
        f"with each run's init")

    cpu = None
    if args.cpu_steps > 0:
        g = T.grid_puffy(128, 96)
        from koral_zig_amd import host
        nt = host_threads()
        k1, kt = 3 * args.cpu_steps, 12 * args.cpu_steps
        v1, s1 = cpu_baseline_fresh(g, T.physics_puffy(), host.MP_PUFFY, args.seed, False, k1)
        vt, st = cpu_baseline_fresh(g, T.physics_puffy(), host.MP_PUFFY, args.seed, False, kt, nt)
        cpu = cpu_entry(v1, s1, k1, vt, st, kt, nt, "one 128x96 run, a=0, from its own seeded init")
    sizes = "{" + ",".join(f"{a}x{b}" for a, b in server.C4_SIZES) + "}"
    out = {
        "metric": "aggregate zone-cycles/s of a PUFFY run server on 1 GPU; p50/p99 step latency",
        "value": value,
        "unit": "zone-cycles/s",
        "n_gpus": 1,
        "steps": int(sum(s.nsteps for s in specs)),
        "warmup": 0,
        "ms_per_step": float(np.percentile(lat, 50)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"run server: {args.runs} PUFFY runs, Poisson arrivals {args.rate}/s, sizes "
                               f"{sizes} uniform, spin U[0,0.95], {args.run_steps} CFL steps each, graph batches "
                               f"of {args.batch} steps, HBM budget {args.hbm_gib} GiB",
                   "seed": args.seed},
        "roofline": None,
        "cpu_baseline": cpu,
        "step_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                            "n": int(lat.size), "definition": "graph batch launch to observed completion / steps"},
        "server": {"wall_s": res["wall_s"], "zone_cycles": res["zone_cycles"], "peak_resident": res["peak_resident"],
                   "peak_gib": res["peak_bytes"] / 2**30, "last_arrival_s": specs and max(s.arrival for s in specs)},
        "sequential_eager": {"value": seq_v, "step_s": seq["step_s"], "wall_with_init_s": seq["wall_s"],
                             "speedup": value / seq_v},
        "server_mode": "one run group" if args.server_mode == "grouped" else "per-run streams",
        "grouped": grouped,
    }
    print(json.dumps(out), flush=True)


def face_units(grid):
    """Faces one k_face_flux launch computes (every active dimension)."""
    n = (grid.nx, grid.ny, grid.nz)
    tot = 0
    for d in range(3):
        if n[d] <= 1:
            continue
        a, b = (n[1], n[2]) if d == 0 else ((n[0], n[2]) if d == 1 else (n[0], n[1]))
        tot += (n[d] + 1) * (a + 2 if a > 1 else a) * (b + 2 if b > 1 else b)
    return tot


def cpu_baseline_fresh(grid, phys, mp, seed, rho_u_only, steps, threads=1):
    """The oracle over `steps` CFL steps of `grid` from its own seeded init on
    `threads` host threads: zone-cycles/s on this host."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import kzo
    from koral_zig_amd import host
    tab = host.metric_tables(grid, mp)
    kzo.set_threads(threads)
    sim = kzo.Sim(grid, phys, tab)
    host.init_puffy(sim, grid, tab, phys, seed=seed, rho_u_only=rho_u_only)
    sim.step()  # the startup step (tiny dt) is not timed
    per = []
    for _ in range(steps):
        t0 = time.perf_counter()
        sim.step()
        per.append(time.perf_counter() - t0)
    kzo.set_threads(1)
    return grid.nx * grid.ny * grid.nz / float(np.median(per)), float(np.sum(per))


def cpu_entry(v1, s1, steps1, vt, st, stepst, nt, what):
    """The cpu_baseline object: the all-cores figure as `value`, the
    single-thread one beside it; both from the median step time."""
    log(f"cpu baseline: 1 thread {v1:.4e} zone-cycles/s ({s1:.1f}s), {nt} threads {vt:.4e} ({st:.1f}s)")
    return {"value": vt, "unit": "zone-cycles/s", "cores": nt, "kind": "port",
            "sample": f"median of {stepst} CFL steps of {what} on {nt} threads ({st:.1f}s); the same oracle, "
                      f"bitwise the serial result, on 1 thread, median of {steps1} steps ({s1:.1f}s): "
                      f"{v1:.4e} zone-cycles/s",
            "value_1thread": v1}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=384)
    ap.add_argument("--ny", type=int, default=360)
    ap.add_argument("--seed", type=int, default=SEED)
    ap.add_argument("--cpu-steps", type=int, default=5,
                    help="oracle steps in the single-thread CPU baseline sample (all cores: at least 20; 0: skip)")
    ap.add_argument("--latency-steps", type=int, default=50)
    ap.add_argument("--prod-steps", type=int, default=300,
                    help="puffy: CFL steps of the production option set (radiative viscosity and the dynamo on), "
                         "timed after the main run and reported beside it (0: skip)")
    ap.add_argument("--workload", choices=["puffy", "wedge3d", "implicit-batch", "run-server"], default="puffy",
                    help="puffy: the 384x360 step (default); wedge3d: the 3D wedge; implicit-batch: the "
                         "independent-cell solve batch; run-server: many runs served concurrently")
    ap.add_argument("--runs", type=int, default=96, help="run-server: number of runs")
    ap.add_argument("--rate", type=float, default=4.0, help="run-server: Poisson arrival rate per second")
    ap.add_argument("--run-steps", type=int, default=500, help="run-server: CFL steps per run")
    ap.add_argument("--batch", type=int, default=10, help="run-server: steps per captured graph")
    ap.add_argument("--hbm-gib", type=float, default=8.0, help="run-server: device-memory budget of resident runs")
    ap.add_argument("--server-mode", choices=["streams", "grouped", "both"], default="streams",
                    help="run-server: per-run streams, one run group with cross-run implicit launches, or both "
                         "(value from per-run streams, the group's figures beside it)")
    ap.add_argument("--nx3", type=int, default=192)
    ap.add_argument("--ny3", type=int, default=192)
    ap.add_argument("--nz3", type=int, default=64)
    ap.add_argument("--cells", type=int, default=1 << 22, help="implicit-batch: number of cells")
    ap.add_argument("--cpu-cells", type=int, default=20000, help="implicit-batch: oracle sample size (0: skip)")
    ap.add_argument("--no-mixed", dest="mixed", action="store_false",
                    help="implicit-batch: skip the FP32-Jacobian comparison")
    ap.add_argument("--lib", help="load this libkz_hip.so instead of the in-tree one (A/B of two builds)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs the step on one GPU only; --gpus must be 1 "
                                   "(the multi-GPU phi-slab decomposition is a design, not built)"}))
        sys.exit(2)

    _pkg()
    if args.lib:
        from koral_zig_amd import hip
        hip.LIB_PATH = Path(args.lib).resolve()
    if args.workload == "implicit-batch":
        if "--steps" not in sys.argv:
            args.steps = 5
        if "--warmup" not in sys.argv:
            args.warmup = 1
        bench_implicit_batch(args)
        return
    if args.workload == "run-server":
        bench_run_server(args)
        return
    if args.workload == "wedge3d":
        if "--steps" not in sys.argv:
            args.steps = 200
        if "--warmup" not in sys.argv:
            args.warmup = 5
        if "--latency-steps" not in sys.argv:
            args.latency_steps = 10
    from koral_zig_amd import abi_types as T
    from koral_zig_amd import host

    if args.workload == "wedge3d":
        # the 3D wedge: 192x192x64 (+3 ghosts), periodic phi over pi/2, spin
        # 0.9375 (RMIN from rminForSpin), rho and u perturbed
        spin = 0.9375
        nx, ny, nz = args.nx3, args.ny3, args.nz3
        grid = T.grid_puffy(nx, ny, nz, rmin=T.rmin_for_spin(spin))
        phys = T.physics_puffy(bh_spin=spin)
        mp = dict(host.MP_PUFFY, a=spin)
        label = (f"PUFFY 3D wedge {nx}x{ny}x{nz} MKS2 a={spin} M=10Msun, periodic phi over pi/2, seeded 1e-4 "
                 f"perturbation of rho and u, CFL steps, radviscosity/dynamo off")
        metric = f"zone-cycles/s full RMHD+M1 step, 3D wedge {nx}x{ny}x{nz} on 1 GPU"
        cpu_grid = T.grid_puffy(48, 48, 16, rmin=T.rmin_for_spin(spin))
        run_grid(args, grid, phys, mp, metric, label, True, cpu_grid)
        return
    grid = T.grid_puffy(args.nx, args.ny)
    phys = T.physics_puffy()
    label = (f"PUFFY torus {args.nx}x{args.ny}x1 MKS2 a=0 M=10Msun, seeded 1e-4 perturbation, "
             f"CFL steps, radviscosity/dynamo off")
    metric = "zone-cycles/s full RMHD+M1 step at 384x360 on 1 GPU; p50/p99 step latency batched"
    run_grid(args, grid, phys, host.MP_PUFFY, metric, label, False, None, production=True)


def production_physics(args, grid, tab):
    """The same run with the production option set of PUFFY (radiative
    viscosity and the mimic dynamo on, koral/problems/puffy/puffy.zig:227-230):
    zone-cycles/s over args.prod_steps CFL steps after the warmup."""
    from koral_zig_amd import abi_types as T
    from koral_zig_amd import hip, host

    phys = T.physics_puffy(radviscosity=True, dynamo=True)
    sim = hip.Sim(grid, phys, tab)
    sim.set_bl_tables(host.bl_tables(grid))
    host.init_puffy(sim, grid, tab, phys, seed=args.seed)
    sim.step_many(args.warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.step_many(args.prod_steps)
    sim.sync()
    wall = time.perf_counter() - t0
    sim.close()
    zones = grid.nx * grid.ny * grid.nz
    log(f"production option set: {args.prod_steps} steps in {wall:.3f}s")
    return {"value": zones * args.prod_steps / wall, "unit": "zone-cycles/s", "steps": args.prod_steps,
            "ms_per_step": wall / args.prod_steps * 1e3,
            "options": "radiative viscosity + mimic dynamo on (PUFFY production), otherwise as config"}


def run_grid(args, grid, phys, mp, metric, label, rho_u_only, cpu_grid, production=False):
    """Time K CFL steps of a PUFFY run on the GPU and print the JSON line;
    cpu_grid (optional) is a smaller grid of the same physics for the oracle
    sample when the run itself is too large to step on one host thread."""
    from koral_zig_amd import hip, host

    t0 = time.perf_counter()
    tab = host.metric_tables(grid, mp)
    log(f"metric tables {time.perf_counter() - t0:.1f}s")
    sim = hip.Sim(grid, phys, tab)
    host.init_puffy(sim, grid, tab, phys, seed=args.seed, rho_u_only=rho_u_only)
    log(f"setup {time.perf_counter() - t0:.1f}s (metric tables, PUFFY init, perturbation seed {args.seed:#x})")

    sim.step_many(args.warmup)
    sim.sync()
    log(f"warmup {args.warmup} steps done")

    # per-class breakdown over one step each (outside the timed region)
    classes = ["face_flux", "implicit", "u2p", "update", "wavespeeds", "ct", "bc"]
    per_step = {}
    per_launch = {}
    for k in classes:
        sim.time_kernel(k)
        sim.step_many(1)
        ms, n = sim.kernel_time()
        per_step[k] = ms
        per_launch[k] = ms / max(n, 1)
    sim.time_kernel(None)
    dom = max(per_step, key=per_step.get)
    log("per-step kernel ms: " + ", ".join(f"{k} {v:.3f}" for k, v in per_step.items()) + f" -> dominant {dom}")

    c0 = sim.counters()
    sim.time_kernel(dom)
    sim.sync()
    t0 = time.perf_counter()
    sim.step_many(args.steps)
    sim.sync()
    wall = time.perf_counter() - t0
    ms_dom, n_dom = sim.kernel_time()
    sim.time_kernel(None)
    c1 = sim.counters()
    zones = grid.nx * grid.ny * grid.nz
    value = zones * args.steps / wall
    log(f"timed {args.steps} steps in {wall:.3f}s: {value:.4e} zone-cycles/s; t={c1['t']:.6e}")

    avg_ms = ms_dom / max(n_dom, 1)
    if dom in BYTES_PER_UNIT:
        if dom == "face_flux":
            units_per_launch = face_units(grid)
        else:
            units_per_launch = zones
        achieved = BYTES_PER_UNIT[dom] * units_per_launch / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    else:
        solves = c1["n_radimp_solves"] - c0["n_radimp_solves"]
        iters = c1["n_radimp_iters"] - c0["n_radimp_iters"]
        launches = 2 * args.steps if dom == "implicit" else n_dom
        flops = (IMPLICIT_FLOPS_PER_SOLVE * solves + IMPLICIT_FLOPS_PER_ITER * iters) / max(launches, 1)
        achieved = flops / (avg_ms * 1e-3) / 1e12
        # FP64 vector ALU work (no matrix instructions): the peak is the
        # chip's FP64 vector rate
        roof = {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    c1 = (grid.nx, grid.ny, grid.nz) == (384, 360, 1)
    roof["traffic"] = TRAFFIC_BYTES_PER_LAUNCH_C1.get(dom) if c1 else None
    roof["kernel"] = dom
    roof["avg_launch_ms"] = avg_ms
    roof["launches"] = n_dom
    # the HBM-bound classes against their algorithmic bytes per launch, from
    # the per-class timings of one step before the timed region
    comps = {}
    for k, units in (("face_flux", face_units(grid)), ("update", zones), ("u2p", zones)):
        ms_l = per_launch[k]
        if ms_l > 0:
            gbs = BYTES_PER_UNIT[k] * units / (ms_l * 1e-3) / 1e9
            comps[k] = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": gbs / HBM_PEAK_GBS, "avg_launch_ms": ms_l, "bytes_per_unit": BYTES_PER_UNIT[k],
                        "units_per_launch": units}
    # the whole explicit operator (two per step) against its compulsory bytes
    op_ms = sum(per_step[k] for k in ("face_flux", "update", "ct", "wavespeeds")) / 2.0
    if op_ms > 0:
        ndim = 3 if grid.nz > 1 else 2
        gbs = EXPLICIT_OP_BYTES[ndim] * zones / (op_ms * 1e-3) / 1e9
        comps["explicit_operator"] = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                      "frac": gbs / HBM_PEAK_GBS, "avg_launch_ms": op_ms,
                                      "bytes_per_unit": EXPLICIT_OP_BYTES[ndim], "units_per_launch": zones,
                                      "kernels": "wavespeeds + face_flux + emf + update"}
    roof["components"] = comps

    # step latency: individually synchronised steps
    lat = []
    for _ in range(args.latency_steps):
        a = time.perf_counter()
        sim.step_many(1)
        sim.sync()
        lat.append((time.perf_counter() - a) * 1e3)
    lat = np.array(lat)

    cpu = None
    nt = host_threads()
    kt = max(args.cpu_steps, 20)
    if args.cpu_steps > 0 and cpu_grid is None:
        p, u, _ = sim.get_state()
        c = sim.counters()
        v1, s1 = cpu_baseline(grid, phys, tab, p, u, c, args.cpu_steps)
        vt, st = cpu_baseline(grid, phys, tab, p, u, c, kt, nt)
        cpu = cpu_entry(v1, s1, args.cpu_steps, vt, st, kt, nt,
                        f"the same {grid.nx}x{grid.ny}x{grid.nz} run from the GPU's post-warmup state")
    elif args.cpu_steps > 0:
        v1, s1 = cpu_baseline_fresh(cpu_grid, phys, mp, args.seed, rho_u_only, args.cpu_steps)
        vt, st = cpu_baseline_fresh(cpu_grid, phys, mp, args.seed, rho_u_only, kt, nt)
        cpu = cpu_entry(v1, s1, args.cpu_steps, vt, st, kt, nt,
                        f"a {cpu_grid.nx}x{cpu_grid.ny}x{cpu_grid.nz} run of the same physics from its own seeded init")
    sim.close()
    prod = production_physics(args, grid, tab) if production and args.prod_steps > 0 else None

    out = {
        "metric": metric,
        "value": value,
        "unit": "zone-cycles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": label, "nx": grid.nx, "ny": grid.ny, "nz": grid.nz, "seed": args.seed},
        "roofline": roof,
        "cpu_baseline": cpu,
        "step_latency_ms": ({"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                             "n": int(lat.size)} if lat.size else None),
    }
    if prod is not None:
        out["production_physics"] = prod
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
