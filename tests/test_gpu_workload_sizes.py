"""GPU vs the oracle at the sizes of the measured workloads that the other
tests run smaller.

* The independent-cell implicit batch at its full size, 2^22 cells of the
  workload's own generator and seed (koral_zig_amd.workloads): outcome,
  rung, iteration count and solved state of every cell bit for bit.
* The five grid sizes the run server draws from, {64x64, 128x96, 192x192,
  256x256, 384x360}, each at a spin inside the server's range U[0, 0.95]:
  two CFL steps from the seeded PUFFY torus, state, flags and clock bit for
  bit.

The oracle runs on every host thread (bitwise its serial result).
"""
import os

import numpy as np
import pytest

import kzo
from test_gpu_step import assert_same
from koral_zig_amd import abi_types as T
from koral_zig_amd import hip, host, workloads

pytestmark = pytest.mark.gpu
SEED = 0x4B5A4D493335355F


def _threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    kzo.set_threads(max(1, min(16, n)))


@pytest.mark.timeout(300)
def test_implicit_batch_full_size():
    _threads()
    phys = T.physics_puffy()
    n = 1 << 22
    pp, geom, dt = workloads.implicit_batch_inputs(
        n, SEED, phys, lambda p, g: hip.radstate(p, g, phys.gam, phys.opac, 0)[:, 9])
    print(f"implicit batch of {n} cells generated", flush=True)
    b = hip.ImplicitBatch(n, phys)
    b.set(pp, geom, dt)
    b.solve()
    uz, pz, rz, cz = b.get()
    b.close()
    rc, pc, uc = kzo.solve_implicit(pp, geom, dt, phys.gam, phys.rad, phys.opac, phys.implicit)
    kzo.set_threads(1)
    print(f"oracle solved {int((rc['ok'] == 1).sum())} of {n}", flush=True)
    for k in ("ok", "rung", "iters", "whichprim", "whicheq", "whichframe"):
        assert np.array_equal(rc[k], rz[k]), k
    ok = rc["ok"] == 1
    assert pc.tobytes() == pz.tobytes()
    assert uc[ok].tobytes() == uz[ok].tobytes()
    assert cz == ((~ok).sum(), int(rc["iters"][ok].sum()), int(ok.sum()))


@pytest.mark.timeout(300)
@pytest.mark.parametrize("nx,ny,a", [(64, 64, 0.35), (128, 96, 0.9), (192, 192, 0.6), (256, 256, 0.15),
                                     (384, 360, 0.95)])
def test_run_server_sizes(nx, ny, a):
    _threads()
    grid = T.grid_puffy(nx, ny, rmin=T.rmin_for_spin(a))
    phys = T.physics_puffy(bh_spin=a)
    tab = host.metric_tables(grid, dict(host.MP_PUFFY, a=a))
    p0 = host.puffy_prepinit(grid, tab, phys, host.PUFFY_INIT)
    cpu = kzo.Sim(grid, phys, tab)
    gpu = hip.Sim(grid, phys, tab)
    for s in (cpu, gpu):
        host.init_puffy(s, grid, tab, phys, seed=SEED, p0=p0)
    for k in range(2):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"{nx}x{ny} a={a} step {k + 1}")
    print(f"{nx}x{ny} a={a}: two steps identical", flush=True)
    kzo.set_threads(1)
    gpu.close()
