"""FP64 operation counts of the implicit solve (test infrastructure).

Runs the counting build of the oracle (libkzoflops.so, kzo_flops.cpp) over
the cells of a workload, checks that it reproduces the plain oracle bit for
bit (same results, iterations and fields), and fits the count per
cell-solve as  flops = A + B * newton_iterations  by least squares over the
cells that the first rung solves (A: the reference-state set-up, the
bisection warm start and the fixed part of the Newton; B: one Newton
iteration).  The means over all cells are printed as well.

    python oracle/flops/count.py c1 [--cells N]   # implicit cells of the PUFFY 384x360 step
    python oracle/flops/count.py c2 [--cells N]   # the independent-cell batch generator

c1 builds the seeded 384x360 PUFFY state with the oracle (init + one CFL
step, about 30 s) and solves every non-polar domain cell with the step's
implicit dt (dt * gamma of RK2IMEX); c2 draws N cells with the batch
generator's seed.
"""
import argparse
import ctypes as C
import importlib.util
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT / "oracle"))


def _pkg():
    name = "koral_zig_amd"
    if name not in sys.modules:
        pkg = ROOT / "koral-zig_amd"
        spec = importlib.util.spec_from_file_location(name, pkg / "__init__.py", submodule_search_locations=[str(pkg)])
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
    return sys.modules[name]


def lib():
    p = HERE / "libkzoflops.so"
    if not p.exists():
        subprocess.run(["make", "-s", "-C", str(HERE.parent), "flops/libkzoflops.so"], check=True)
    L = C.CDLL(str(p))
    L.kzo_count_implicit_batch.argtypes = [C.c_int64] + [C.c_void_p] * 4 + [C.c_double] + [C.c_void_p] * 5
    L.kzo_count_totals.argtypes = [C.c_void_p]
    return L


_L = None


def _lib():
    global _L
    if _L is None:
        _L = lib()
    return _L


def totals():
    """(operations, divisions, square roots) counted since the library loaded."""
    t = np.zeros(3, np.uint64)
    _lib().kzo_count_totals(t.ctypes.data)
    return tuple(int(x) for x in t)


def count(pp, geom, dt, phys):
    """(results, pp_out, uu_out, flops[n]) of the counting build."""
    T = _pkg().abi_types
    pp = np.array(pp, np.float64, copy=True, order="C")
    uu = np.zeros_like(pp)
    geom = np.ascontiguousarray(geom, np.float64)
    dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float64), (len(pp),)))
    res = np.zeros(len(pp), T.IMPLICIT_RESULT_DTYPE)
    fl = np.zeros(len(pp), np.uint64)
    _lib().kzo_count_implicit_batch(len(pp), uu.ctypes.data, pp.ctypes.data, geom.ctypes.data, dt.ctypes.data,
                                   phys.gam, C.addressof(phys.rad), C.addressof(phys.opac),
                                   C.addressof(phys.implicit), res.ctypes.data, fl.ctypes.data)
    return res, pp, uu, fl


def c1_cells(seed=0x4B5A4D493335355F, nx=384, ny=360):
    import kzo
    pk = _pkg()
    T, host = pk.abi_types, __import__("koral_zig_amd.host", fromlist=["host"])
    g = T.grid_puffy(nx, ny)
    ph = T.physics_puffy()
    tab = host.metric_tables(g)
    sim = kzo.Sim(g, ph, tab)
    host.init_puffy(sim, g, tab, ph, seed=seed)
    sim.step()
    p, _, _ = sim.get_state()
    dt = sim.counters()["dt"] * (1.0 - 1.0 / np.sqrt(2.0))
    ng = g.ng
    pd = p[0, ng:ng + ny, ng:ng + nx].reshape(-1, T.NV)

    def centres(lo, hi, n):
        i = np.arange(n)
        return 0.5 * ((lo + i * (hi - lo) / n) + (lo + (i + 1) * (hi - lo) / n))

    X1, X2 = np.meshgrid(centres(g.minx, g.maxx, nx), centres(g.miny, g.maxy, ny))
    x4 = np.stack([np.zeros(X1.size), X1.ravel(), X2.ravel(), np.zeros(X1.size)], 1)
    geom = host.geom_batch(3, x4)
    keep = np.ones((ny, nx), bool)
    keep[:2] = False  # the polar rows the implicit pass skips (nccorrectpolar = 2)
    keep[-2:] = False
    keep = keep.ravel()
    return pd[keep], geom[keep], dt, ph


def c2_cells(n, seed=0x4B5A4D493335355F):
    import kzo
    pk = _pkg()
    T = pk.abi_types
    workloads = __import__("koral_zig_amd.workloads", fromlist=["workloads"])
    ph = T.physics_puffy()
    pp, geom, dt = workloads.implicit_batch_inputs(n, seed, ph, lambda p, g: kzo.opac_entry(p, g, ph.gam, ph.opac)[:, 3])
    return pp, geom, dt, ph


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=["c1", "c2"])
    ap.add_argument("--cells", type=int, default=0, help="limit the number of cells (0: all of c1; c2 default 200000)")
    a = ap.parse_args()
    import kzo
    t0 = time.perf_counter()
    if a.workload == "c1":
        pp, geom, dt, ph = c1_cells()
        if a.cells:
            pp, geom = pp[:a.cells], geom[:a.cells]
    else:
        pp, geom, dt, ph = c2_cells(a.cells or 200000)
    print(f"{a.workload}: {len(pp)} cells ({time.perf_counter() - t0:.1f}s to build)", flush=True)
    t0 = totals()
    res, po, uo, fl = count(pp, geom, dt, ph)
    t1 = totals()
    r2, p2, u2 = kzo.solve_implicit(pp, geom, dt, ph.gam, ph.rad, ph.opac, ph.implicit)
    same = (np.array_equal(res, r2) and np.array_equal(po, p2, equal_nan=True)
            and np.array_equal(uo, u2, equal_nan=True))
    print("counting build reproduces the oracle bit for bit:", same)
    if not same:
        sys.exit(1)
    it = res["iters"].astype(float)
    ok = res["ok"] == 1
    first = ok & (res["rung"] == 0)
    fl = fl.astype(float)
    X = np.stack([np.ones(first.sum()), it[first]], 1)
    (A, B), *_ = np.linalg.lstsq(X, fl[first], rcond=None)
    pred = A + B * it[first]
    print(f"solves {ok.sum()} of {len(pp)}; first-rung {first.sum()}; mean iterations {it[ok].mean():.4f}")
    print(f"flops per cell: mean {fl.mean():.1f} (solved {fl[ok].mean():.1f}, first-rung {fl[first].mean():.1f}, "
          f"later rungs {fl[ok & ~first].mean() if (ok & ~first).any() else 0:.1f}, failed "
          f"{fl[~ok].mean() if (~ok).any() else 0:.1f})")
    print(f"fit over first-rung solves: flops = {A:.1f} + {B:.1f} * iterations "
          f"(rms residual {np.sqrt(np.mean((fl[first] - pred) ** 2)):.1f})")
    print(f"total {fl.sum():.4e} flops, {it[ok].sum():.0f} iterations, {ok.sum()} solves; of the operations "
          f"{(t1[1] - t0[1]) / len(pp):.1f} divisions and {(t1[2] - t0[2]) / len(pp):.1f} square roots per cell")


if __name__ == "__main__":
    main()
