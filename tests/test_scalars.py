"""scalars.dat reductions (koral/io/scalars.zig) in the oracle, and the
driver-side pieces of koral_zig_amd.scalars.

Pinned by the reference's C golden init/puffy_scalars.kgld: mass, Mdot
through the horizon shell and H/R at r = 15 on the PUFFY 384x360 t=0 state,
gated at 1e-6 as puffystep_golden_tests.zig:94-140 gates them (the MKS2
sqrt(-g) spread between the C and Zig metrics); the luminosities are
reported there, not gated (an LTE torus has almost no flux at the outer
shell), and here they are checked for sign and finiteness.  The state is
this build's own host init of the same torus (the C t=0 primitive blob is
not in the reference tree), which the init tests pin to the reference's
post-init snapshot.  The BL shell widths are pinned by telescoping to the
theta span, the shell index by its definition.
"""
import math

import numpy as np
import pytest

import kzo
from goldens import read_kgld
from koral_zig_amd import abi_types as T
from koral_zig_amd import host, scalars


@pytest.fixture(scope="module")
def t0_state():
    grid = T.grid_puffy(384, 360)
    phys = T.physics_puffy()
    tab = host.metric_tables(grid)
    sim = kzo.Sim(grid, phys, tab)
    host.init_puffy(sim, grid, tab, phys)
    bl = host.bl_tables(grid)
    sim.set_bl_tables(bl)
    return grid, sim, bl


def test_scalars_vs_c_golden(t0_state):
    grid, sim, _ = t0_state
    _, out = read_kgld("init/puffy_scalars.kgld", 0, 5)
    mass_c, mdot_c, radlum_c, totlum_c, h_c = out[0]
    d = scalars.Diagnostics(grid)
    assert d.ix_h == scalars.radial_shell_index(host.radius_bl(grid), 2.0)
    assert d.ix_l == grid.nx - 1  # r = 5000 lies beyond RMAX
    row = d.row(sim, 0.0)
    assert math.isclose(row["mass"], mass_c, rel_tol=1e-6), (row["mass"], mass_c)
    assert math.isclose(row["mdot"], mdot_c, rel_tol=1e-6), (row["mdot"], mdot_c)
    assert math.isclose(row["scaleheight"], h_c, rel_tol=1e-6), (row["scaleheight"], h_c)
    assert row["radlum"] >= 0.0 and np.isfinite(row["totallum"])
    # the initial magnetisation is normalised to max pmag/ptot = 1/20 (postinit's maxbeta)
    assert math.isclose(row["max_pmag_ptot"], 1.0 / 20.0, rel_tol=1e-6)
    assert row["n_nan"] == 0 and row["nstep"] == 0


def test_shell_widths_telescope(t0_state):
    grid, _, _ = t0_state
    r = host.radius_bl(grid)
    assert np.all(np.diff(r) > 0)
    for ix in (0, 100, grid.nx - 1):
        dth, dph = host.shell_widths_bl(grid, ix)
        assert dth.shape == (1, grid.ny) and np.all(dth > 0)
        # the theta map is monotonic, so the widths telescope to the span
        lo = host.shell_widths_bl(T.grid_puffy(grid.nx, 1, rmin=1.85), ix)  # one cell spans [miny, maxy]
        assert math.isclose(dth.sum(), lo[0][0, 0], rel_tol=1e-12)
        assert np.allclose(dph, grid.maxz - grid.minz, rtol=1e-15)  # phi is not remapped


def test_radial_shell_index():
    r = np.array([1.5, 2.0, 2.5, 3.0])
    assert scalars.radial_shell_index(r, 2.0) == 2  # first strictly above
    assert scalars.radial_shell_index(r, 0.0) == 0
    assert scalars.radial_shell_index(r, 10.0) == 3  # none above: the last index
    assert scalars.r_horizon_bl(0.0) == 2.0


def test_row_format():
    row = dict(t=1.5, dt=2.0e-7, nstep=12, mass=1.0, mdot=-3.25e-20, radlum=0.0, totallum=float("nan"),
               scaleheight=0.25, max_pmag_ptot=1.0e-2, n_hd_fixup=3, n_radimp_fail=0, n_nan=1, n_floorguard=7)
    line = scalars.format_row(row)
    assert line == ("1.5000000000e0 2.000000e-7 12 1.0000000000e0 -3.2500000000e-20 0.0000000000e0 nan "
                    "2.500000e-1 1.000000e-2 3 0 1 7\n")
    log = scalars.ScalarsLog()
    log.append(row)
    assert log.text.startswith(scalars.HEADER) and log.text.endswith(line)
