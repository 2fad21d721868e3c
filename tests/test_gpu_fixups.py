"""GPU vs the oracle on cells planted to fail: the fixup paths and the floor guard ladder.

The smooth PUFFY states of the other step tests flag no cell for a fixup,
so here chosen cells are planted (tests/plant.py) where the u2p inversion
must ask for the hydro fixup (negative conserved density,
koral/tests/state_tests.zig:755-777), where u2pRad corrects the radiation
(negative radiation energy), and where the M1 implicit ladder fails every
rung (states of the implicit batch).  The oracle's fixups are pinned to a
brute-force neighbour average in tests/test_oracle_fixups.py; here the HIP
path (k_u2p's fixup list and k_u2p_fixups, k_fixup, the grid-side failed())
must match the oracle bit for bit after the pass and after whole steps.
"""
import numpy as np
import pytest

import kzo
import plant
from test_gpu_step import assert_same
from koral_zig_amd import abi_types as T
from koral_zig_amd import hip, host

pytestmark = pytest.mark.gpu


def _copy(ph, **kw):
    q = type(ph).from_buffer_copy(ph)
    for k, v in kw.items():
        setattr(q, k, v)
    return q


def _pair(grid, phys, seed, launch=0):
    tab = host.metric_tables(grid)
    cpu = kzo.Sim(grid, phys, tab)
    gpu = hip.Sim(grid, phys, tab)
    gpu.set_implicit_launch(launch)
    for s in (cpu, gpu):
        host.init_puffy(s, grid, tab, phys, seed=seed)
    return tab, cpu, gpu


def _count(f, which):
    return int(np.count_nonzero(f[..., which]))


@pytest.mark.parametrize("case", ["hd2d", "hd+rad2d", "hd3d"])
def test_u2p_fixups_planted(case):
    if case == "hd3d":
        grid, cells, seed = T.grid_puffy(16, 12, 6), plant.PLANT_3D, 4
    else:
        grid, cells, seed = T.grid_puffy(24, 20), plant.PLANT_2D, 3
    ph = T.physics_puffy()
    if case == "hd+rad2d":
        ph = _copy(ph, do_u2prad_fixups=1)
    tab, cpu, gpu = _pair(grid, ph, seed)
    p0, u0, _ = cpu.get_state()
    bad = plant.negate_slot(u0, grid, cells, plant.RHO)
    if case == "hd+rad2d":
        # (10, 8) and (15, 12) carry both flags; (12, 14), (13, 14) only RAD
        bad = plant.negate_slot(bad, grid, [(10, 8), (15, 12), (12, 14), (13, 14), (0, 9)], plant.EE)
    print(f"case {case}: {len(cells)} planted cells", flush=True)
    for s in (cpu, gpu):
        s.set_state(p=p0, u=bad)
        s.calc_u2p()
    assert_same(cpu, gpu, "calc_u2p", clock=False)
    _, _, f = gpu.get_state()
    assert _count(f, T.FLAG_HD_FIXUP) >= 6
    if case == "hd+rad2d":
        assert _count(f, T.FLAG_RAD_FIXUP) == 5
    # whole steps from the planted state (the step's first u2p meets the
    # planted conserveds after the implicit pass)
    for s in (cpu, gpu):
        s.set_state(p=p0, u=bad)
    for k in range(2):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"step {k + 1}")
    gpu.close()


@pytest.mark.parametrize("launch", [1, 2])
@pytest.mark.parametrize("fix", [0, 1])
def test_radimp_failures(launch, fix):
    """Planted cells whose implicit ladder fails: flag -1, failure counter,
    state kept (radimp fixup off) or averaged by k_fixup (on); both launch
    shapes of the implicit pass; then whole steps."""
    grid = T.grid_puffy(24, 20)
    ph = _copy(T.physics_puffy(), do_radimp_fixups=fix)
    tab, cpu, gpu = _pair(grid, ph, 3, launch)
    p0, u0, _ = cpu.get_state()
    dt = 0.29 / cpu.counters()["tstepdenmax"]
    p1 = plant.plant_prims(p0, grid, plant.radimp_fail_states(ph))
    print(f"launch {launch} radimp fixups {fix}", flush=True)
    for s in (cpu, gpu):
        s.set_state(p=p1, u=u0)
        s.p2u_domain()
    p1, u1, _ = cpu.get_state()
    for s in (cpu, gpu):
        s.op_implicit(dt)
    assert_same(cpu, gpu, "op_implicit", clock=False)
    _, _, f = gpu.get_state()
    assert _count(f, T.FLAG_RADIMP_FIXUP) == len(plant.RADIMP_FAIL_24x20)
    assert gpu.counters()["n_radimp_failures"] == cpu.counters()["n_radimp_failures"] > 0
    for s in (cpu, gpu):
        s.set_state(p=p1, u=u1)
    for k in range(2):
        cpu.step()
        gpu.step()
        assert_same(cpu, gpu, f"step {k + 1}")
    gpu.close()


def test_floors_guard_ladder():
    """checkFloorsMhd on records that reach the drift-frame guards: state,
    return codes and the guard counter bitwise."""
    fp = T.floor_params_puffy()
    pp, geom = plant.guard_records()
    rc, pc, nc = kzo.check_floors_mhd(pp, geom, 5.0 / 3.0, fp)
    pz, rz, nz = hip.check_floors_mhd(pp, geom, 5.0 / 3.0, fp)
    assert nc > 100 and nz == nc
    assert np.array_equal(rc.astype(np.int32), rz)
    assert pc.tobytes() == pz.tobytes() or np.array_equal(pc, pz, equal_nan=True)
