"""Run groups: several resident runs stepped together, the implicit cells of
all of them solved by one set of implicit launches per implicit pass
(kz_group_*).  A cell's implicit solve reads only that cell and its own run's
dt, so every member must end each step exactly where its own oracle run
ends it: fields, flags, clock and implicit counters, bit for bit.  Checked
for runs of different sizes and spins, for both implicit launch shapes, for
eager steps and graph replays (with eager work between replays), and across
a change of membership.
"""
import numpy as np
import pytest

import kzo
from koral_zig_amd import abi_types as T
from koral_zig_amd import hip, host
from test_gpu_step import assert_same

pytestmark = pytest.mark.gpu

RUNS = ((24, 20, 0.0, 41), (32, 24, 0.5, 42), (16, 12, 0.9, 43))


def make_run(nx, ny, spin, seed):
    grid = T.grid_puffy(nx, ny, rmin=T.rmin_for_spin(spin))
    phys = T.physics_puffy(bh_spin=spin)
    tab = host.metric_tables(grid, dict(host.MP_PUFFY, a=spin))
    cpu = kzo.Sim(grid, phys, tab)
    gpu = hip.Sim(grid, phys, tab)
    assert host.init_puffy(cpu, grid, tab, phys, seed=seed) == host.init_puffy(gpu, grid, tab, phys, seed=seed)
    return cpu, gpu


def check_all(pairs, what):
    for i, (cpu, gpu) in enumerate(pairs):
        assert_same(cpu, gpu, f"{what}, run {i}")


@pytest.mark.parametrize("mode", [1, 2])
def test_group_steps_and_replays_match_oracle(mode):
    pairs = [make_run(*r) for r in RUNS]
    grp = hip.Group([g for _, g in pairs])
    grp.set_implicit_launch(mode)
    grp.step_many(2)
    grp.sync()
    for cpu, _ in pairs:
        cpu.step()
        cpu.step()
    check_all(pairs, "eager")
    grp.capture_steps(2)
    for rep in range(3):
        grp.replay()
        grp.sync()
        for cpu, _ in pairs:
            cpu.step()
            cpu.step()
        check_all(pairs, f"replay {rep}")
        if rep == 0:  # eager work between replays
            grp.step_many(1)
            grp.sync()
            for cpu, _ in pairs:
                cpu.step()
            check_all(pairs, "eager between replays")
    grp.close()
    for _, g in pairs:
        g.close()


def test_group_membership_change():
    pairs = [make_run(*r) for r in RUNS]
    (ca, ga), (cb, gb), (cc, gc) = pairs
    grp = hip.Group([ga, gb])
    grp.step_many(2)
    grp.set_runs([gb, gc])  # waits for the first membership's steps
    grp.step_many(2)
    ga.step_many(1)  # a former member steps on its own
    grp.sync()
    ga.sync()
    for _ in range(2):
        ca.step()
        cb.step()
    for _ in range(2):
        cb.step()
        cc.step()
    ca.step()
    check_all(pairs, "after the change")
    grp.set_runs([])
    grp.step_many(1)  # an empty group steps nothing
    grp.sync()
    check_all(pairs, "empty group")
    grp.close()
    for _, g in pairs:
        g.close()


def test_group_rejects_mismatched_physics_and_duplicates():
    _, g1 = make_run(16, 12, 0.0, 5)
    grid = T.grid_puffy(16, 12)
    phys = T.physics_puffy()
    phys.gam = 4.0 / 3.0
    tab = host.metric_tables(grid)
    g2 = hip.Sim(grid, phys, tab)
    grp = hip.Group()
    with pytest.raises(RuntimeError):
        grp.set_runs([g1, g2])
    with pytest.raises(RuntimeError):
        grp.set_runs([g1, g1])
    with pytest.raises(RuntimeError):
        grp.replay()  # nothing captured
    grp.set_runs([g1])
    g1.time_kernel("implicit")
    with pytest.raises(RuntimeError):
        grp.step_many(1)  # event timing is per run
    g1.time_kernel(None)
    grp.close()
    g1.close()
    g2.close()


def test_grouped_server_matches_sequential():
    """The run server in group mode: runs of three sizes and spins arriving
    over time join and leave the group at batch boundaries; every final
    state equals the sequential eager run's."""
    from koral_zig_amd import server
    from test_gpu_step import _diff
    sizes = ((24, 20), (32, 24), (16, 12))
    specs = server.poisson_specs(6, 200.0, 4, seed=9, sizes=sizes, spin_max=0.9)
    prep = server.prepare(specs, workers=0)
    prep2 = dict(prep)
    finals = {}

    def keep(spec, sim):
        finals[spec.rid] = sim.get_state() + (sim.counters(),)

    one = hip.Sim(specs[0].grid(), specs[0].phys(), prep[specs[0].rid][0])
    budget = 3 * one.device_bytes()
    one.close()
    srv = server.RunServer(budget, batch_steps=2, grouped=True)
    res = srv.serve(specs, prep, on_finish=keep)
    assert res["runs"] == 6 and len(finals) == 6
    assert res["zone_cycles"] == sum(s.nx * s.ny * s.nsteps for s in specs)
    assert len(res["latency_ms"]) == sum(s.nsteps // 2 for s in specs)
    assert res["regroups"] >= 2

    def check(spec, sim):
        p, u, f = sim.get_state()
        pg, ug, fg, cg = finals[spec.rid]
        assert _diff(p, pg) is None and _diff(u, ug) is None, spec.rid
        assert np.array_equal(f, fg), spec.rid
        c = sim.counters()
        assert c["t"] == cg["t"] and c["n_radimp_iters"] == cg["n_radimp_iters"], spec.rid

    seq = server.serve_sequential(specs, prep2, on_finish=check)
    assert seq["zone_cycles"] == res["zone_cycles"]


def test_group_full_physics_matches_oracle():
    """Radiative viscosity and the dynamo on (BL sidecar loaded): the
    viscous pass in each member's step head and the dynamo in its middle and
    tail run on the member's stream around the group's implicit passes."""
    pairs = []
    for nx, ny, seed in ((24, 20, 51), (16, 12, 52)):
        grid = T.grid_puffy(nx, ny)
        phys = T.physics_puffy(radviscosity=True, dynamo=True)
        tab = host.metric_tables(grid)
        bl = host.bl_tables(grid)
        cpu = kzo.Sim(grid, phys, tab)
        gpu = hip.Sim(grid, phys, tab)
        for s in (cpu, gpu):
            s.set_bl_tables(bl)
            host.init_puffy(s, grid, tab, phys, seed=seed)
        pairs.append((cpu, gpu))
    grp = hip.Group([g for _, g in pairs])
    grp.capture_steps(1)
    for rep in range(2):
        grp.replay()
        grp.sync()
        for cpu, _ in pairs:
            cpu.step()
        check_all(pairs, f"replay {rep}")
    grp.close()
    for _, g in pairs:
        g.close()
