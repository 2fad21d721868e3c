"""3D (the phi wedge with periodic phi) on the GPU vs the oracle, bit for bit.

The z face pass and the 3D corner fill (bc.zig:378-520: 12 edge tubes, 8
corner cubes), the z sweep, the 3D flux-CT and update, the 3D curl of the
dynamo and the phi loops of every pass, against the CPU restatement that the
CPU suite pins to the reference's C golden step/puffy3d.kstp (32x30x4).
Spin 0.9375 (the 192x192x64 production wedge's spin, with RMIN from
rminForSpin) has no C oracle; there the CPU path is the oracle.
"""
import numpy as np
import pytest

import kzo
from goldens import read_kstp
from koral_zig_amd import abi_types as T
from koral_zig_amd import hip, host
from test_gpu_step import assert_same

pytestmark = pytest.mark.gpu


def make_pair(nx, ny, nz, seed=11, spin=0.0, full=False):
    mp = dict(host.MP_PUFFY, a=spin)
    grid = T.grid_puffy(nx, ny, nz, rmin=T.rmin_for_spin(spin))
    phys = T.physics_puffy(radviscosity=full, dynamo=full, bh_spin=spin)
    tab = host.metric_tables(grid, mp)
    cpu = kzo.Sim(grid, phys, tab)
    gpu = hip.Sim(grid, phys, tab)
    if full:
        bl = host.bl_tables(grid, mp)
        cpu.set_bl_tables(bl)
        gpu.set_bl_tables(bl)
    fc = host.init_puffy(cpu, grid, tab, phys, seed=seed)
    fg = host.init_puffy(gpu, grid, tab, phys, seed=seed)
    assert fc == fg
    return grid, phys, tab, cpu, gpu


@pytest.fixture(scope="module")
def pair3d():
    g, ph, tab, cpu, gpu = make_pair(16, 12, 6)
    yield g, ph, tab, cpu, gpu
    gpu.close()


def test_3d_init_identical(pair3d):
    _, _, _, cpu, gpu = pair3d
    assert_same(cpu, gpu, "init 3d")


def test_3d_steps_identical(pair3d):
    _, _, _, cpu, gpu = pair3d
    for k in range(3):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"3d step {k + 1}")


def test_3d_full_physics_steps_identical():
    grid, phys, tab, cpu, gpu = make_pair(16, 12, 4, seed=5, full=True)
    for k in range(2):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"3d full step {k + 1}")
    gpu.close()


def test_3d_spin_steps_identical():
    grid, phys, tab, cpu, gpu = make_pair(16, 12, 4, seed=7, spin=0.9375)
    for k in range(2):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"3d spin step {k + 1}")
    gpu.close()


def test_puffy3d_c_state_identical():
    """The C 3D step golden's post-init state and forced dt sequence."""
    k = read_kstp("step/puffy3d.kstp.gz")
    grid = T.grid_puffy(k["nx"], k["ny"], k["nz"])
    phys = T.physics_puffy(radviscosity=True, dynamo=True)
    tab = host.metric_tables(grid)
    bl = host.bl_tables(grid)
    cpu = kzo.Sim(grid, phys, tab)
    gpu = hip.Sim(grid, phys, tab)
    r0 = k["recs"][0]
    ng = 3
    dom = (slice(ng, ng + grid.nz), slice(ng, ng + grid.ny), slice(ng, ng + grid.nx))
    p, u, _ = cpu.get_state()
    p[dom] = r0["p"]
    u[dom] = r0["u"]
    for s in (cpu, gpu):
        s.set_bl_tables(bl)
        s.set_state(p=p, u=u)
        s.set_bc(True)
        s.init_timestep_guess()
    assert_same(cpu, gpu, "puffy3d set_bc")
    for i, r in enumerate(k["recs"][1:], 1):
        cpu.step(r["dt"])
        gpu.step(r["dt"])
        assert_same(cpu, gpu, f"puffy3d step {i}")
    gpu.close()
