"""KDMP checkpoint and restart from device state (io/dump.zig:19-160,
restart_tests.zig:71-150): the bytes of kz_sim_write_kdmp equal the Python
restatement of serializePrimDump over the run's own primitives and clock; a
fresh sim loaded from them holds the same domain primitives bit for bit and
u = p2u(p) as the oracle computes it; the clock, step and frame come back;
a wrong shape is DimMismatch; and the restarted run steps exactly like the
oracle restarted the same way (Sim.finishInit: ghosts, startup dt guess)."""
import numpy as np
import pytest

import kzo
from goldens import kdmp_bytes, read_kdmp
from koral_zig_amd import abi_types as T
from koral_zig_amd import hip, host
from test_gpu_step import assert_same

pytestmark = pytest.mark.gpu


def dom(grid):
    ngz = 3 if grid.nz > 1 else 0
    return (slice(ngz, ngz + grid.nz), slice(3, 3 + grid.ny), slice(3, 3 + grid.nx))


@pytest.mark.parametrize("shape", [(24, 20, 1), (12, 10, 4)])
def test_checkpoint_restart(shape):
    grid = T.grid_puffy(*shape)
    phys = T.physics_puffy()
    tab = host.metric_tables(grid)
    src = hip.Sim(grid, phys, tab)
    host.init_puffy(src, grid, tab, phys, seed=2)
    src.step()
    src.step()
    c = src.counters()
    data = src.write_kdmp(out_idx=7)
    p, u, _ = src.get_state()
    assert data == kdmp_bytes(p[dom(grid)], c["t"], c["nstep"], 7)
    d = read_kdmp(data)
    assert (d["t"], d["nstep"], d["out_idx"]) == (c["t"], 2, 7)

    dst = hip.Sim(grid, phys, tab)
    h = dst.load_kdmp(data)
    assert h == {k: d[k] for k in h}
    pd, ud, _ = dst.get_state()
    assert pd[dom(grid)].tobytes() == p[dom(grid)].tobytes()
    # the oracle restarted the same way: primitives, p2u, ghosts, dt guess, clock
    ref = kzo.Sim(grid, phys, tab)
    pr = np.zeros_like(p)
    pr[dom(grid)] = p[dom(grid)]
    ref.set_state(p=pr)
    ref.p2u_domain()
    ref.set_bc(True)
    ref.init_timestep_guess()
    rc = ref.counters()
    ref.set_time(d["t"], rc["tstepdenmax"], rc["tstepdenmin"])
    assert ud[dom(grid)].tobytes() == ref.get_state()[1][dom(grid)].tobytes()
    cd = dst.counters()
    assert (cd["t"], cd["nstep"]) == (c["t"], 2)
    assert cd["tstepdenmax"] == rc["tstepdenmax"]
    ref.step()
    dst.step()
    assert_same(ref, dst, "restarted step", clock=False)

    other = hip.Sim(T.grid_puffy(shape[0] + 1, shape[1], shape[2]), phys,
                    host.metric_tables(T.grid_puffy(shape[0] + 1, shape[1], shape[2])))
    with pytest.raises(RuntimeError, match="DimMismatch"):
        other.load_kdmp(data)
    for s in (src, dst, other):
        s.close()
