"""GPU vs the oracle at the production grid sizes, and the CT divergence.

* 384x360 (the measured 2D workload: PUFFY torus, a = 0, seeded 1e-4
  perturbation 0x4b5a4d493335355f): two CFL steps, state, flags and clock
  bit for bit, the implicit pass in the launch shape chosen by size.
* 192x192x64 wedge (a = 0.9375, periodic phi, rho and u perturbed): one
  CFL step bit for bit, the implicit pass in the persistent launch shape
  its size selects.
* calcDivB (koral/sim/ct.zig:336-381) on the device equals the oracle's bit
  for bit, and after twenty GPU steps the corner divergence stays at machine
  precision: everywhere in 2D, and in 3D away from the periodic phi seam and
  the polar rows (what tests/test_oracle_divb.py pins on the oracle).

The oracle runs on every host thread (its threads split rows; the result is
bitwise the serial one).
"""
import os

import numpy as np
import pytest

import kzo
from test_gpu_step import assert_same
from koral_zig_amd import abi_types as T
from koral_zig_amd import hip, host

pytestmark = pytest.mark.gpu
SEED = 0x4b5a4d493335355f


def _threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    kzo.set_threads(max(1, min(16, n)))


def _divb_scale(grid, p):
    return (grid.maxx - grid.minx) / grid.nx / np.abs(p[..., 6:9]).max()


@pytest.mark.timeout(300)
def test_c1_384x360_steps_and_divb():
    _threads()
    grid = T.grid_puffy(384, 360)
    phys = T.physics_puffy()
    tab = host.metric_tables(grid)
    p0 = host.puffy_prepinit(grid, tab, phys, host.PUFFY_INIT)
    cpu = kzo.Sim(grid, phys, tab)
    gpu = hip.Sim(grid, phys, tab)
    for s in (cpu, gpu):
        host.init_puffy(s, grid, tab, phys, seed=SEED, p0=p0)
    print("384x360 initialised", flush=True)
    assert_same(cpu, gpu, "init")
    for k in range(2):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"step {k + 1}")
        print(f"384x360 step {k + 1} identical", flush=True)
    c = gpu.counters()
    assert c["n_radimp_solves"] > 100000
    dc, dg = cpu.calc_divb(), gpu.calc_divb()
    assert dc.tobytes() == dg.tobytes()
    p, _, _ = gpu.get_state()
    scale = _divb_scale(grid, p)
    gpu.step_many(18)
    gpu.sync()
    d = np.abs(gpu.calc_divb()) * scale
    print(f"384x360 max |divB| dx / max|B| after 20 steps: {d.max():.3e}", flush=True)
    assert d.max() <= 1e-12
    gpu.close()


@pytest.mark.timeout(300)
def test_c3_192x192x64_step_and_divb():
    _threads()
    a = 0.9375
    grid = T.grid_puffy(192, 192, 64, rmin=T.rmin_for_spin(a))
    phys = T.physics_puffy(bh_spin=a)
    tab = host.metric_tables(grid, dict(host.MP_PUFFY, a=a))
    p0 = host.puffy_prepinit(grid, tab, phys, host.PUFFY_INIT)
    cpu = kzo.Sim(grid, phys, tab)
    gpu = hip.Sim(grid, phys, tab)
    for s in (cpu, gpu):
        host.init_puffy(s, grid, tab, phys, seed=SEED, rho_u_only=True, p0=p0)
    print("192x192x64 initialised", flush=True)
    cpu.step()
    gpu.step()
    assert_same(cpu, gpu, "step 1")
    print("192x192x64 step 1 identical", flush=True)
    dc, dg = cpu.calc_divb(), gpu.calc_divb()
    assert dc.tobytes() == dg.tobytes()
    del cpu
    p, _, _ = gpu.get_state()
    scale = _divb_scale(grid, p)
    gpu.step_many(19)
    gpu.sync()
    d = np.abs(gpu.calc_divb()) * scale
    nc = phys.nccorrectpolar
    inner = d[1:, nc + 1:grid.ny - nc, :]
    print(f"192x192x64 max |divB| dx / max|B| after 20 steps: {d.max():.3e} (off the seam and polar rows "
          f"{inner.max():.3e})", flush=True)
    assert inner.max() <= 1e-12
    gpu.close()


def test_divb_identical_small():
    """Device calcDivB on a small 2D and 3D run equals the oracle's bitwise."""
    for dims in ((32, 30, 1), (16, 12, 6)):
        grid = T.grid_puffy(*dims)
        phys = T.physics_puffy()
        tab = host.metric_tables(grid)
        cpu = kzo.Sim(grid, phys, tab)
        gpu = hip.Sim(grid, phys, tab)
        for s in (cpu, gpu):
            host.init_puffy(s, grid, tab, phys, seed=7)
            s.step()
        assert cpu.calc_divb().tobytes() == gpu.calc_divb().tobytes(), dims
        gpu.close()
