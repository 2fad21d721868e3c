"""The oracle's cell-parallel passes on several threads give the serial
result bit for bit (fields with ghosts, flags, clock and implicit counters),
with radiative viscosity and the dynamo on, in 2D and 3D."""
import numpy as np
import pytest

import kzo
from koral_zig_amd import abi_types as T
from koral_zig_amd import host


@pytest.mark.parametrize("shape", [(40, 36, 1), (16, 14, 6)])
def test_threads_bitwise(shape):
    nx, ny, nz = shape
    g = T.grid_puffy(nx, ny, nz)
    ph = T.physics_puffy(radviscosity=True, dynamo=True)
    tab = host.metric_tables(g)
    bl = host.bl_tables(g)
    out = []
    for nt in (1, 3):
        kzo.set_threads(nt)
        try:
            s = kzo.Sim(g, ph, tab)
            s.set_bl_tables(bl)
            host.init_puffy(s, g, tab, ph, seed=13)
            for _ in range(2):
                s.step()
            out.append((s.get_state(), s.counters()))
        finally:
            kzo.set_threads(1)
    (a, ca), (b, cb) = out
    assert a[0].tobytes() == b[0].tobytes() and a[1].tobytes() == b[1].tobytes()
    assert np.array_equal(a[2], b[2])
    assert ca == cb
    assert ca["n_radimp_solves"] > 0
