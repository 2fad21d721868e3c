"""HIP-graph replay of captured CFL steps, and the run server built on it.

kz_sim_capture_steps records the launches of N CFL steps once; every
kz_sim_replay must continue the run exactly as N eager steps would.  Each
replay is checked against the oracle (fields, flags, clock and implicit
counters identical), with eager device work and host synchronisation
between replays.  The run server (many runs resident at once, each stepping
by graph replays on its own stream) must leave every run in the state the
plain sequential eager loop leaves it in, bit for bit.
"""
import numpy as np
import pytest

import kzo
from koral_zig_amd import abi_types as T
from koral_zig_amd import hip, host, server
from test_gpu_step import _diff, assert_same, make_pair

pytestmark = pytest.mark.gpu


def test_replays_match_oracle():
    grid, phys, tab, cpu, gpu = make_pair(24, 20, seed=21)
    gpu.capture_steps(2)
    for rep in range(4):
        gpu.replay()
        gpu.sync()
        for _ in range(2):
            cpu.step()
        assert_same(cpu, gpu, f"replay {rep}")
        if rep == 1:  # eager work between replays
            gpu.step_many(1)
            gpu.sync()
            cpu.step()
            assert_same(cpu, gpu, "eager between replays")
    gpu.close()


def test_replays_full_physics_match_oracle():
    """Radiative viscosity and the dynamo switched on (BL sidecar loaded)."""
    grid = T.grid_puffy(24, 20)
    phys = T.physics_puffy(radviscosity=True, dynamo=True)
    tab = host.metric_tables(grid)
    bl = host.bl_tables(grid)
    cpu = kzo.Sim(grid, phys, tab)
    gpu = hip.Sim(grid, phys, tab)
    for s in (cpu, gpu):
        s.set_bl_tables(bl)
        host.init_puffy(s, grid, tab, phys, seed=4)
    gpu.capture_steps(1)
    for rep in range(3):
        gpu.replay()
        gpu.sync()
        cpu.step()
        assert_same(cpu, gpu, f"replay {rep}")
    gpu.close()


def test_capture_rejected_while_timing():
    grid, phys, tab, cpu, gpu = make_pair(16, 12, seed=3)
    gpu.time_kernel("implicit")
    with pytest.raises(RuntimeError):
        gpu.capture_steps(1)
    gpu.time_kernel(None)
    with pytest.raises(RuntimeError):
        gpu.replay()  # nothing captured
    gpu.capture_steps(1)
    gpu.replay()
    gpu.sync()
    cpu.step()
    assert_same(cpu, gpu, "after rejected capture")
    gpu.close()


def test_server_matches_sequential():
    """Five runs of three sizes and spins, arriving together, batches of two
    steps, an HBM budget that holds about two at a time: every final state
    equals the sequential eager run's."""
    sizes = ((24, 20), (32, 24), (16, 12))
    specs = server.poisson_specs(5, 1000.0, 4, seed=7, sizes=sizes, spin_max=0.9)
    prep = server.prepare(specs, workers=0)
    prep2 = dict(prep)
    finals = {}

    def keep(spec, sim):
        finals[spec.rid] = sim.get_state() + (sim.counters(),)

    one = hip.Sim(specs[0].grid(), specs[0].phys(), prep[specs[0].rid][0])
    budget = 2 * one.device_bytes()
    one.close()
    srv = server.RunServer(budget, batch_steps=2)
    res = srv.serve(specs, prep, on_finish=keep)
    assert res["runs"] == 5 and len(finals) == 5
    assert res["zone_cycles"] == sum(s.nx * s.ny * s.nsteps for s in specs)
    assert len(res["latency_ms"]) == sum(s.nsteps // 2 for s in specs)

    def check(spec, sim):
        p, u, f = sim.get_state()
        pg, ug, fg, cg = finals[spec.rid]
        assert _diff(p, pg) is None and _diff(u, ug) is None, spec.rid
        assert np.array_equal(f, fg), spec.rid
        c = sim.counters()
        assert c["t"] == cg["t"] and c["n_radimp_iters"] == cg["n_radimp_iters"], spec.rid

    seq = server.serve_sequential(specs, prep2, on_finish=check)
    assert seq["zone_cycles"] == res["zone_cycles"]


def test_capture_while_another_thread_syncs():
    """One thread captures a run's step graph while another waits on,
    reads and frees a second run: the synchronous calls are stream-ordered
    on each run's own stream, so neither disturbs the capture, and the
    captured run still matches the oracle."""
    import threading
    grid, phys, tab, cpu, gpu = make_pair(24, 20, seed=33)
    other = [hip.Sim(grid, phys, tab) for _ in range(3)]
    for o in other:
        host.init_puffy(o, grid, tab, phys, seed=34)
    err = []

    def capture():
        try:
            for _ in range(5):
                gpu.capture_steps(3)
        except Exception as e:  # reported below
            err.append(e)

    th = threading.Thread(target=capture)
    th.start()
    k = 0
    while th.is_alive():
        o = other[k % len(other)]
        o.step_many(1)
        o.sync()
        o.counters()
        o.get_state()
        k += 1
    th.join()
    for o in other:
        o.close()
    assert not err, err
    gpu.replay()
    gpu.sync()
    for _ in range(3):
        cpu.step()
    assert_same(cpu, gpu, "replay after concurrent capture")
    gpu.close()
