"""Host side of the run server: the arrival process and the prepared
set-up, without a GPU."""
import numpy as np

from koral_zig_amd import abi_types as T
from koral_zig_amd import host, server


def test_poisson_specs():
    a = server.poisson_specs(96, 4.0, 500, seed=1234)
    b = server.poisson_specs(96, 4.0, 500, seed=1234)
    assert [(s.nx, s.ny, s.spin, s.seed, s.arrival) for s in a] == [(s.nx, s.ny, s.spin, s.seed, s.arrival) for s in b]
    t = np.array([s.arrival for s in a])
    assert t[0] == 0.0 and np.all(np.diff(t) > 0)
    # mean gap of a rate-4/s process: 0.25 s (95 gaps; loose 4-sigma band)
    assert abs(np.diff(t).mean() - 0.25) < 4 * 0.25 / np.sqrt(95)
    assert {(s.nx, s.ny) for s in a} <= set(server.C4_SIZES)
    assert len({(s.nx, s.ny) for s in a}) == len(server.C4_SIZES)
    assert all(0.0 <= s.spin <= 0.95 for s in a)
    assert len({s.seed for s in a}) == 96
    assert all(s.nsteps == 500 for s in a)


def test_prepare_matches_direct():
    specs = server.poisson_specs(3, 10.0, 2, seed=5, sizes=((16, 12), (20, 16)), spin_max=0.9)
    prep = server.prepare(specs, workers=2)
    assert sorted(prep) == [0, 1, 2]
    for s in specs:
        grid = T.grid_puffy(s.nx, s.ny, rmin=T.rmin_for_spin(s.spin))
        tab = host.metric_tables(grid, dict(host.MP_PUFFY, a=s.spin))
        p0 = host.puffy_prepinit(grid, tab, T.physics_puffy(bh_spin=s.spin))
        t2, q0 = prep[s.rid]
        assert np.array_equal(q0, p0)
        for k, v in tab.items():
            if isinstance(v, np.ndarray):
                assert np.array_equal(t2[k], v), k
