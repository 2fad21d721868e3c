"""GPU vs the oracle with the reference options that change the hot path:

* the two reconstruction-order options of the sweep (koral/sim.zig:
  1316-1337), pinned on the oracle's side in tests/test_oracle_reduce_order.py:
  the face kernel's choice of scheme per cell;
* the b^2/rho floor in the ZAMO frame (koral/solve/invert.zig:649-684),
  pinned in tests/test_oracle_zamo_floor.py: the per-record floor kernel and
  whole steps with strong-field cells planted;
* the synchrotron bridge of the opacities (koral/physics/opacities.zig:
  208-297), pinned in tests/test_oracle_synchrotron_bridge.py: the
  opacity/radiation-state kernel, the implicit batch and whole steps;
* all three in the members of a run group, whose implicit cells are solved
  by cross-run launches.

Each must give the oracle's state, flags, return codes and clock bit for
bit.
"""
import numpy as np
import pytest

import kzo
import plant
from goldens import geoms, read_kgld
from test_gpu_implicit_batch import FIELDS, same
from test_gpu_step import assert_same
from koral_zig_amd import abi_types as T
from koral_zig_amd import hip, host

pytestmark = pytest.mark.gpu


def _copy(ph, **kw):
    q = type(ph).from_buffer_copy(ph)
    for k, v in kw.items():
        setattr(q, k, v)
    return q


def _pair(grid, phys, seed):
    tab = host.metric_tables(grid, dict(host.MP_PUFFY, a=phys.bh_spin))
    cpu = kzo.Sim(grid, phys, tab)
    gpu = hip.Sim(grid, phys, tab)
    for s in (cpu, gpu):
        host.init_puffy(s, grid, tab, phys, seed=seed)
    return cpu, gpu


@pytest.mark.parametrize("dims", ["2d", "3d"])
def test_reduceorderafterfixup(dims):
    if dims == "2d":
        grid, seed, poison = T.grid_puffy(24, 20), 3, [(10, 8), (15, 12), (16, 12), (4, 5), (0, 9)]
    else:
        grid, seed, poison = T.grid_puffy(16, 12, 6), 4, [(4, 5, 2), (9, 7, 3), (10, 7, 3)]
    ph = _copy(T.physics_puffy(), reduceorderafterfixup=1)
    cpu, gpu = _pair(grid, ph, seed)
    p0, u0, _ = cpu.get_state()
    bad = plant.negate_slot(u0, grid, poison, plant.RHO)
    print(f"reduceorderafterfixup {dims}: {len(poison)} poisoned cells", flush=True)
    for s in (cpu, gpu):
        s.set_state(p=p0, u=bad)
        s.calc_u2p()
    assert np.count_nonzero(gpu.get_state()[2][..., T.FLAG_HD_FIXUP]) >= 3
    for s in (cpu, gpu):
        s.op_explicit(1e-3)
    assert_same(cpu, gpu, "op_explicit after the flagged u2p", clock=False)
    for s in (cpu, gpu):
        s.set_state(p=p0, u=bad)
    for k in range(2):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"step {k + 1}")
    gpu.close()


@pytest.mark.parametrize("a,dims", [(0.0, "2d"), (0.9375, "2d"), (0.5, "3d")])
def test_reduceorderatbh(a, dims):
    if dims == "2d":
        grid, seed = T.grid_puffy(192, 16, rmin=T.rmin_for_spin(a)), 3
    else:
        grid, seed = T.grid_puffy(128, 12, 6, rmin=T.rmin_for_spin(a)), 4
    ix_max = host.reduce_order_ix_max(grid, dict(host.MP_PUFFY, a=a))
    ph = _copy(T.physics_puffy(bh_spin=a), reduceorderatbh=1, reduce_order_ix_max=ix_max)
    cpu, gpu = _pair(grid, ph, seed)
    print(f"reduceorderatbh a={a} {dims}: ix_max {ix_max}", flush=True)
    for k in range(2):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"step {k + 1}")
    gpu.close()


def test_both_reduce_options_full_physics():
    """Both options with radiative viscosity and the dynamo on (the production
    option set), three steps from a poisoned state."""
    a = 0.5
    grid = T.grid_puffy(96, 20, rmin=T.rmin_for_spin(a))
    mp = dict(host.MP_PUFFY, a=a)
    ph = _copy(T.physics_puffy(radviscosity=True, dynamo=True, bh_spin=a), reduceorderatbh=1,
               reduceorderafterfixup=1, reduce_order_ix_max=host.reduce_order_ix_max(grid, mp))
    tab = host.metric_tables(grid, mp)
    bl = host.bl_tables(grid, mp)
    cpu = kzo.Sim(grid, ph, tab)
    gpu = hip.Sim(grid, ph, tab)
    for s in (cpu, gpu):
        s.set_bl_tables(bl)
        host.init_puffy(s, grid, tab, ph, seed=7)
    p0, u0, _ = cpu.get_state()
    bad = plant.negate_slot(u0, grid, [(30, 8), (31, 8), (60, 12)], plant.RHO)
    for s in (cpu, gpu):
        s.set_state(p=p0, u=bad)
    for k in range(3):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"step {k + 1}")
    gpu.close()


def _zamo_floors():
    fp = T.floor_params_puffy()
    fp.b2rhofloorframe = 1
    return fp


def test_zamo_floor_records():
    """checkFloorsMhd in the ZAMO frame on the strong-field records: state and
    return codes bitwise, and no drift-frame guard counted."""
    fp = _zamo_floors()
    pp, geom = plant.guard_records()
    rc, pc, nc = kzo.check_floors_mhd(pp, geom, 5.0 / 3.0, fp)
    pz, rz, nz = hip.check_floors_mhd(pp, geom, 5.0 / 3.0, fp)
    assert nc == 0 and nz == 0
    assert (rc == -1).sum() > 1000
    assert np.array_equal(rc.astype(np.int32), rz)
    assert same(pc, pz)


def _plant_strong_field(p, grid, cells):
    """rho and u lowered 1e4-fold and B raised 300-fold at the given cells, so
    that b^2/rho exceeds its limit there."""
    ngz, ngy, ngx = plant._pad(grid)
    p = p.copy()
    for c in cells:
        ix, iy, iz = c if len(c) == 3 else (*c, 0)
        q = p[iz + ngz, iy + ngy, ix + ngx]
        q[0:2] *= 1e-4
        q[6:9] *= 300.0
    return p


@pytest.mark.parametrize("dims", ["2d", "3d"])
def test_zamo_floor_steps(dims):
    if dims == "2d":
        grid, seed, cells = T.grid_puffy(24, 20), 3, [(10, 8), (15, 12), (4, 5), (20, 10)]
    else:
        grid, seed, cells = T.grid_puffy(16, 12, 6), 4, [(4, 5, 2), (9, 7, 3), (12, 3, 5)]
    ph = T.physics_puffy()
    ph.floors.b2rhofloorframe = 1
    cpu, gpu = _pair(grid, ph, seed)
    p0 = _plant_strong_field(cpu.get_state()[0], grid, cells)
    print(f"zamo floor {dims}: {len(cells)} strong-field cells", flush=True)
    for s in (cpu, gpu):
        s.set_state(p=p0)
        s.p2u_domain()
    for k in range(2):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"step {k + 1}")
    gpu.close()


def _bridge_opac():
    op = T.opac_params_puffy()
    op.synchrotron_bridge = 1
    return op


def test_bridge_radstate():
    """The opacity / radiation-state kernel with the bridge on, over the
    rad_thermo golden states with E scaled across the Trad = Te clamp."""
    inp, _ = read_kgld("rad/rad_thermo.kgld", 36, 13)
    pp0, g0 = inp[:, 23:36], geoms(inp)
    pp = np.concatenate([pp0 * np.where(np.arange(T.NV) == 9, s, 1.0) for s in 10.0 ** np.arange(-3.0, 4.0)])
    g = np.concatenate([g0] * 7)
    op = _bridge_opac()
    for full in (1, 0):
        c = kzo.radstate(pp, g, 5.0 / 3.0, op, full)
        z = hip.radstate(pp, g, 5.0 / 3.0, op, full)
        cols = [0, 1, 3, 4, 5, 6, 7, 11]
        assert same(c[:, cols], z[:, :8]), full
        if full:
            assert same(c[:, 13], z[:, 8])
    o = kzo.opac_entry(pp, g, 5.0 / 3.0, op)
    assert same(o[:, 3], z[:, 9])
    # the bridge is really on
    assert not np.array_equal(c[:, 6], kzo.radstate(pp, g, 5.0 / 3.0, T.opac_params_puffy(), 0)[:, 6])


def test_bridge_implicit_batch():
    from koral_zig_amd import workloads
    phys = T.physics_puffy()
    phys.opac.synchrotron_bridge = 1
    pp, geom, dt = workloads.implicit_batch_inputs(
        8000, 11, phys, lambda p, g: kzo.opac_entry(p, g, phys.gam, phys.opac)[:, 3])
    rc, pc, uc = kzo.solve_implicit(pp, geom, dt, phys.gam, phys.rad, phys.opac, phys.implicit)
    b = hip.ImplicitBatch(len(pp), phys)
    b.set(pp, geom, dt)
    b.solve()
    ug, pg, rg, (nfail, iters, solves) = b.get()
    b.close()
    for k in FIELDS:
        bad = np.nonzero(rc[k] != rg[k])[0]
        assert bad.size == 0, (k, bad[:5])
    ok = rc["ok"] == 1
    assert ok.mean() > 0.95
    assert same(pc[ok], pg[ok]) and same(uc[ok], ug[ok]) and same(pg[~ok], pp[~ok])
    assert (nfail, iters, solves) == (int((~ok).sum()), int(rc["iters"][ok].sum()), int(ok.sum()))


@pytest.mark.parametrize("launch", [0, 2])
@pytest.mark.parametrize("dims", ["2d", "3d"])
def test_bridge_steps(dims, launch):
    """Launch 0 takes the shape the size selects (the merged first rung at
    these sizes), 2 forces the persistent launches."""
    grid, seed = (T.grid_puffy(24, 20), 3) if dims == "2d" else (T.grid_puffy(16, 12, 6), 4)
    ph = T.physics_puffy()
    ph.opac.synchrotron_bridge = 1
    cpu, gpu = _pair(grid, ph, seed)
    gpu.set_implicit_launch(launch)
    print(f"synchrotron bridge {dims}, implicit launch {launch}", flush=True)
    for k in range(3):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"step {k + 1}")
    gpu.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_group_with_options(mode):
    """A run group whose members run the options: the bridge variants of the
    cross-run implicit kernels (both launch shapes), and on each member's
    stream the ZAMO-frame u2p and face kernels with the per-cell order choice;
    strong-field cells planted in each member."""
    pairs = []
    for nx, ny, spin, seed, cells in ((24, 20, 0.0, 61, [(10, 8), (15, 12)]),
                                      (32, 24, 0.5, 62, [(20, 10), (5, 6)])):
        grid = T.grid_puffy(nx, ny, rmin=T.rmin_for_spin(spin))
        ph = _copy(T.physics_puffy(bh_spin=spin), reduceorderafterfixup=1)
        ph.floors.b2rhofloorframe = 1
        ph.opac.synchrotron_bridge = 1
        cpu, gpu = _pair(grid, ph, seed)
        p0 = _plant_strong_field(cpu.get_state()[0], grid, cells)
        for s in (cpu, gpu):
            s.set_state(p=p0)
            s.p2u_domain()
        pairs.append((cpu, gpu))
    print(f"group with options, implicit launch mode {mode}", flush=True)
    grp = hip.Group([g for _, g in pairs])
    grp.set_implicit_launch(mode)
    grp.step_many(2)
    grp.sync()
    for cpu, _ in pairs:
        cpu.step()
        cpu.step()
    for i, (cpu, gpu) in enumerate(pairs):
        assert_same(cpu, gpu, f"run {i}")
    grp.close()
    for _, g in pairs:
        g.close()


@pytest.mark.timeout(300)
def test_c1_384x360_all_options_production_physics():
    """The headline grid with every option on and the production physics
    (radiative viscosity and the dynamo): the viscous face kernel with the
    per-cell order and the ZAMO floor, the ZAMO u2p, and the bridge variants
    of the wavespeeds and of the implicit kernels in the launch shape this
    size selects; strong-field cells planted.  Two CFL steps, bit for bit."""
    from test_gpu_fullsize import SEED, _threads
    _threads()
    grid = T.grid_puffy(384, 360)
    mp = dict(host.MP_PUFFY)
    ph = _copy(T.physics_puffy(radviscosity=True, dynamo=True), reduceorderatbh=1, reduceorderafterfixup=1,
               reduce_order_ix_max=host.reduce_order_ix_max(grid, mp))
    ph.floors.b2rhofloorframe = 1
    ph.opac.synchrotron_bridge = 1
    tab = host.metric_tables(grid)
    bl = host.bl_tables(grid)
    p0 = host.puffy_prepinit(grid, tab, ph, host.PUFFY_INIT)
    cpu = kzo.Sim(grid, ph, tab)
    gpu = hip.Sim(grid, ph, tab)
    for s in (cpu, gpu):
        s.set_bl_tables(bl)
        host.init_puffy(s, grid, tab, ph, seed=SEED, p0=p0)
    p1 = _plant_strong_field(cpu.get_state()[0], grid, [(40, 60), (200, 180), (300, 20)])
    for s in (cpu, gpu):
        s.set_state(p=p1)
        s.p2u_domain()
    print(f"384x360 with every option, ix_max {ph.reduce_order_ix_max}", flush=True)
    for k in range(2):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"step {k + 1}")
        print(f"384x360 options step {k + 1} identical", flush=True)
    gpu.close()
