"""GPU radiative viscosity and dynamo vs the oracle, bit for bit.

The device R^i_j (rijvisc.zig:266-311), applyDynamo (dynamo.zig:462-469)
and whole RK2IMEX steps with both switched on (the PUFFY production option
set, selfscenarios.zig:71-90) against the CPU restatement, which the CPU
suite pins to the reference's C step golden (puffy64.kstp) and Zig
self-goldens.  Same bar as the rest of the step: fields, flags, clock and
implicit counters identical.
"""
import numpy as np
import pytest

import kzo
from goldens import read_kstp
from koral_zig_amd import abi_types as T
from koral_zig_amd import hip, host
from test_gpu_step import _diff, assert_same

pytestmark = pytest.mark.gpu


def make_pair(nx, ny, seed=11, radviscosity=True, dynamo=True):
    grid = T.grid_puffy(nx, ny)
    phys = T.physics_puffy(radviscosity=radviscosity, dynamo=dynamo)
    tab = host.metric_tables(grid)
    bl = host.bl_tables(grid)
    cpu = kzo.Sim(grid, phys, tab)
    gpu = hip.Sim(grid, phys, tab)
    cpu.set_bl_tables(bl)
    gpu.set_bl_tables(bl)
    host.init_puffy(cpu, grid, tab, phys, seed=seed)
    host.init_puffy(gpu, grid, tab, phys, seed=seed)
    return grid, phys, tab, cpu, gpu


def test_rijvisc_identical():
    grid, phys, tab, cpu, gpu = make_pair(32, 30, radviscosity=True, dynamo=False)
    dt = 0.7 / cpu.counters()["tstepdenmax"]
    rc = cpu.calc_rij_visc_total(dt)
    rg = gpu.calc_rij_visc_total(dt)
    assert np.count_nonzero(rc) > 0
    assert _diff(rc, rg) is None, _diff(rc, rg)
    gpu.close()


def test_apply_dynamo_identical():
    """One applyDynamo with B3 = 10 B2 injected (the dynamo golden's set-up,
    dynamo_golden_tests.zig:55-120), so the DAMPBETA and the curl both act."""
    grid, phys, tab, cpu, gpu = make_pair(48, 44, radviscosity=False, dynamo=True)
    p, _, _ = cpu.get_state()
    ng = 3
    p[0, ng:ng + grid.ny, ng:ng + grid.nx, 8] = 10.0 * p[0, ng:ng + grid.ny, ng:ng + grid.nx, 7]
    for s in (cpu, gpu):
        s.set_state(p=p)
        s.p2u_domain()
    assert_same(cpu, gpu, "inject", clock=False)
    cpu.apply_dynamo(10.0)
    gpu.apply_dynamo(10.0)
    assert_same(cpu, gpu, "apply_dynamo", clock=False)
    pa, _, _ = gpu.get_state()
    assert not np.array_equal(pa[..., 6:9], p[..., 6:9])
    gpu.close()


def test_full_physics_steps_identical():
    """Three CFL steps of the production option set (radiative viscosity,
    dynamo, polar-axis correction)."""
    grid, phys, tab, cpu, gpu = make_pair(32, 30, seed=4)
    for k in range(3):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"step {k + 1}")
    gpu.close()


def test_puffy64_c_state_identical():
    """The C step golden's post-init state (puffy64.kstp record 0) and its
    forced dt sequence: GPU == oracle after every step (the oracle itself
    meets the reference's own gates on this golden in the CPU suite)."""
    k = read_kstp("step/puffy64.kstp.gz")
    grid = T.grid_puffy(k["nx"], k["ny"])
    phys = T.physics_puffy(radviscosity=True, dynamo=True)
    tab = host.metric_tables(grid)
    bl = host.bl_tables(grid)
    cpu = kzo.Sim(grid, phys, tab)
    gpu = hip.Sim(grid, phys, tab)
    r0 = k["recs"][0]
    ng = 3
    dom = (slice(None), slice(ng, ng + grid.ny), slice(ng, ng + grid.nx))
    p, u, _ = cpu.get_state()
    p[dom] = r0["p"]
    u[dom] = r0["u"]
    for s in (cpu, gpu):
        s.set_bl_tables(bl)
        s.set_state(p=p, u=u)
        s.set_bc(True)
        s.init_timestep_guess()
    for i, r in enumerate(k["recs"][1:], 1):
        cpu.step(r["dt"])
        gpu.step(r["dt"])
        assert_same(cpu, gpu, f"puffy64 step {i}")
    gpu.close()


def test_step_needs_bl_tables():
    grid = T.grid_puffy(16, 12)
    tab = host.metric_tables(grid)
    gpu = hip.Sim(grid, T.physics_puffy(dynamo=True), tab)
    with pytest.raises(RuntimeError, match="InvalidConfig"):
        gpu.step(1e-6)
    gpu.close()
