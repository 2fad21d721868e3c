"""GPU step vs the oracle: every pass and whole RK2IMEX steps of PUFFY runs.

The HIP path (include/kz_step.h) and the CPU restatement start from the
same initial state (the PUFFY initializer of koral/problems/puffy/puffy.zig,
with the seeded 1e-4 perturbation so the implicit solves, fixups and floors
all fire) and must agree bit-for-bit after each pass and each step:
fields (ghosts included), per-cell flags, the CFL clock and the implicit
solver's iteration counters.  A NaN compares equal to a NaN (the payload
of a NaN is not part of the arithmetic).
"""
import numpy as np
import pytest

import kzo
from goldens import DevTracker, read_kslf
from koral_zig_amd import abi_types as T
from koral_zig_amd import hip, host

pytestmark = pytest.mark.gpu

CLOCK = ("t", "dt", "own_dt", "tstepdenmax", "tstepdenmin", "nstep", "n_radimp_failures", "n_radimp_iters",
         "n_radimp_solves", "n_guard_recoveries")


def _diff(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    eq = (a == b) | (np.isnan(a) & np.isnan(b))
    if eq.all():
        return None
    idx = np.argwhere(~eq)
    i = tuple(idx[0])
    dev = np.abs(a - b) / np.maximum(1.0, np.maximum(np.abs(a), np.abs(b)))
    return dict(n=len(idx), first=i, cpu=float(a[i]), gpu=float(b[i]), max_dev=float(np.nanmax(dev)))


def assert_same(cpu, gpu, what, clock=True):
    pc, uc, fc = cpu.get_state()
    pg, ug, fg = gpu.get_state()
    assert _diff(pc, pg) is None, (what, "p", _diff(pc, pg))
    assert _diff(uc, ug) is None, (what, "u", _diff(uc, ug))
    assert np.array_equal(fc, fg), (what, "flags", np.argwhere(fc != fg)[:5])
    if clock:
        cc, cg = cpu.counters(), gpu.counters()
        for k in CLOCK:
            assert cc[k] == cg[k] or (np.isnan(cc[k]) and np.isnan(cg[k])), (what, k, cc[k], cg[k])


def make_pair(nx, ny, seed=11):
    grid = T.grid_puffy(nx, ny)
    phys = T.physics_puffy()
    tab = host.metric_tables(grid)
    cpu = kzo.Sim(grid, phys, tab)
    gpu = hip.Sim(grid, phys, tab)
    fc = host.init_puffy(cpu, grid, tab, phys, seed=seed)
    fg = host.init_puffy(gpu, grid, tab, phys, seed=seed)
    assert fc == fg
    return grid, phys, tab, cpu, gpu


@pytest.fixture(scope="module")
def pair32():
    g, ph, tab, cpu, gpu = make_pair(32, 30)
    yield g, ph, tab, cpu, gpu
    gpu.close()


def test_init_identical(pair32):
    _, _, _, cpu, gpu = pair32
    assert_same(cpu, gpu, "init")


def test_init_matches_selfgolden():
    """The GPU-built initial state vs the reference's own post-init snapshot
    (selfgolden puffy_fast, 32x30, step 0), as the host-init test does for
    the oracle: fields 1e-10 relative, flags exact, tstepdenmax 1e-12."""
    s = read_kslf("selfgolden/puffy_fast.kslf")
    grid = T.grid_puffy(s["nx"], s["ny"])
    phys = T.physics_puffy()
    tab = host.metric_tables(grid)
    gpu = hip.Sim(grid, phys, tab)
    host.init_puffy(gpu, grid, tab, phys)
    p, u, f = gpu.get_state()
    fr = s["fields"][0]
    assert DevTracker().add(fr["p"], p.reshape(-1, T.NV)).ok(1e-10)
    assert DevTracker().add(fr["u"], u.reshape(-1, T.NV)).ok(1e-10)
    c = gpu.counters()
    assert abs(c["tstepdenmax"] - s["scalars"][0, 3]) <= 1e-12 * s["scalars"][0, 3]
    gpu.close()


def test_passes_identical():
    """Each pass of Sim.step on its own, from a shared state."""
    grid, phys, tab, cpu, gpu = make_pair(24, 20, seed=3)
    dt = 1.0 / cpu.counters()["tstepdenmax"]
    gpu.set_bc(False)
    cpu.set_bc(False)
    assert_same(cpu, gpu, "set_bc", clock=False)
    cpu.op_implicit(dt * 0.29)
    gpu.op_implicit(dt * 0.29)
    assert_same(cpu, gpu, "op_implicit", clock=False)
    cpu.calc_u2p()
    gpu.calc_u2p()
    assert_same(cpu, gpu, "calc_u2p", clock=False)
    cpu.do_correct()
    gpu.do_correct()
    assert_same(cpu, gpu, "do_correct", clock=False)
    cpu.op_explicit(dt)
    gpu.op_explicit(dt)
    assert_same(cpu, gpu, "op_explicit", clock=False)
    gpu.close()


def test_steps_identical(pair32):
    """Three CFL steps (the self-golden's length): state, flags and clock
    identical after each step."""
    _, _, _, cpu, gpu = pair32
    for k in range(3):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"step {k + 1}")
    c = gpu.counters()
    assert c["n_radimp_solves"] > 0 and c["n_radimp_iters"] > c["n_radimp_solves"]


def test_forced_dt_and_step_many():
    """A forced-dt step, then step_many (the device-resident clock) equals
    repeated step() on the oracle."""
    grid, phys, tab, cpu, gpu = make_pair(20, 16, seed=5)
    dt = 0.5 / cpu.counters()["tstepdenmax"]
    cpu.step(dt)
    gpu.step(dt)
    assert_same(cpu, gpu, "forced")
    for _ in range(2):
        cpu.step()
    gpu.step_many(2)
    gpu.sync()
    assert_same(cpu, gpu, "step_many")
    gpu.close()


def test_set_time_resume():
    """kz_sim_set_time restores the clock of a resumed run."""
    grid, phys, tab, cpu, gpu = make_pair(16, 12, seed=9)
    cpu.step()
    gpu.step()
    c = cpu.counters()
    p, u, _ = cpu.get_state()
    gpu2 = hip.Sim(grid, phys, tab)
    gpu2.set_state(p, u)
    gpu2.set_time(c["t"], c["tstepdenmax"], c["tstepdenmin"])
    cpu.step()
    gpu2.step()
    pc, uc, _ = cpu.get_state()
    pg, ug, _ = gpu2.get_state()
    assert _diff(pc, pg) is None and _diff(uc, ug) is None
    assert gpu2.counters()["t"] == cpu.counters()["t"]
    gpu.close()
    gpu2.close()



@pytest.mark.parametrize("mode", [1, 2])
def test_implicit_launch_shapes(mode):
    """Both launch shapes of the implicit pass (one merged first-rung launch;
    persistent work stealing, which large grids use) give the oracle's steps
    bit for bit, with radiative viscosity and the dynamo on."""
    grid = T.grid_puffy(40, 36)
    phys = T.physics_puffy(radviscosity=True, dynamo=True)
    tab = host.metric_tables(grid)
    bl = host.bl_tables(grid)
    cpu = kzo.Sim(grid, phys, tab)
    gpu = hip.Sim(grid, phys, tab)
    gpu.set_implicit_launch(mode)
    for s in (cpu, gpu):
        s.set_bl_tables(bl)
        host.init_puffy(s, grid, tab, phys, seed=23)
    for _ in range(3):
        cpu.step()
    gpu.step_many(3)
    gpu.sync()
    assert_same(cpu, gpu, f"launch shape {mode}")
    assert cpu.counters()["n_radimp_solves"] > 0
    gpu.close()


def test_c0_64x64_cfl_step():
    """The C0 case: 64x64 PUFFY torus with the default physics (radiative
    viscosity and dynamo off), initAllWith + initTimestepGuess, the seeded
    1e-4 perturbation of every non-B primitive with ENTR and u recomputed
    and the boundaries refilled, then one step(cflDt()): GPU and oracle
    identical in fields, flags, clock and implicit counters."""
    grid, phys, tab, cpu, gpu = make_pair(64, 64, seed=0x4B5A4D493335355F)
    assert_same(cpu, gpu, "C0 init")
    cpu.step()
    gpu.step()
    assert_same(cpu, gpu, "C0 step")
    gpu.close()
