"""cellFixup and the implicit-failure path of the oracle, pinned by brute force.

Cells are planted so that the pass under test must flag them (a negative
conserved density for u2pMhd, koral/tests/state_tests.zig:755-777; a
negative radiation energy for u2pRad; batch states whose M1 Newton fails
every rung for opImplicit, koral/sim.zig:1715-1774).  Two oracle sims run
the same pass from the same state, one with the fixup enabled and one
without; the state the fixup must produce is then computed from the
second by plant.brute_fixup, a separate restatement of the neighbour
average of koral/sim.zig:1027-1133, and compared bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

import kzo
import plant
from plant import at, domain
from koral_zig_amd import abi_types as T
from koral_zig_amd import host


def _copy(ph, **kw):
    q = type(ph).from_buffer_copy(ph)
    for k, v in kw.items():
        setattr(q, k, v)
    return q


def _start(grid, phys, seed):
    tab = host.metric_tables(grid)
    s = kzo.Sim(grid, phys, tab)
    host.init_puffy(s, grid, tab, phys, seed=seed)
    p, u, _ = s.get_state()
    return tab, s, p, u


def _p2u_of(grid, phys, tab, p):
    s = kzo.Sim(grid, phys, tab)
    s.set_state(p=p, u=np.zeros_like(p))
    s.p2u_domain()
    return s.get_state()[1]


def _nonpolar(grid, phys, cells):
    return [c for c in plant.cells3(cells) if not plant.polar(phys, grid, c[1])]


def _flagged(f, grid, which):
    d = domain(f, grid)[..., which]
    return sorted((int(ix), int(iy), int(iz)) for iz, iy, ix in np.argwhere(d != 0))


def test_negative_density_demands_fixup():
    """koral/tests/state_tests.zig:755-777: a clean conserved state inverts
    with the hot solver and no fixup; negating its density demands one."""
    grid = T.grid_puffy(24, 20)
    ph = T.physics_puffy()
    tab = host.metric_tables(grid)
    x4 = np.array([[0.0, np.log(6.0 - 0.1), 1.0 / np.pi, 0.0]])
    g = host.geom_batch(3, x4)
    pp = np.array([[1.0, 0.1, 0.1, -0.02, 0.05, 0.0, 0.1, 0.0, 0.02, 1e-10, 0.0, 0.0, 0.0]])
    gam = ph.gam
    pp[0, 5] = pp[0, 0] * np.log(((gam - 1.0) * pp[0, 1]) ** (1.0 / (gam - 1.0)) / pp[0, 0] ** (1.0 / (gam - 1.0) + 1.0))
    uu = kzo.p2u(pp, g, gam)
    r, _ = kzo.u2p_mhd(uu, pp, g, gam, ph.floors)
    assert r["ret"][0] == 0 and not r["entropy_used"][0] and not r["fixup"][0]
    bad = uu.copy()
    bad[0, 0] = -bad[0, 0]
    r, _ = kzo.u2p_mhd(bad, pp, g, gam, ph.floors)
    assert r["fixup"][0]
    del tab


@pytest.mark.parametrize("dims", ["2d", "3d"])
def test_hd_fixup_brute_force(dims):
    if dims == "2d":
        grid, cells, seed = T.grid_puffy(24, 20), plant.PLANT_2D, 3
    else:
        grid, cells, seed = T.grid_puffy(16, 12, 6), plant.PLANT_3D, 4
    ph = T.physics_puffy()
    tab, a, p0, u0 = _start(grid, ph, seed)
    b = kzo.Sim(grid, _copy(ph, do_fixups=0), tab)
    bad = plant.negate_slot(u0, grid, cells, plant.RHO)
    for s in (a, b):
        s.set_state(p=p0, u=bad)
        s.calc_u2p()
    pa, ua, fa = a.get_state()
    pb, ub, fb = b.get_state()
    assert np.array_equal(fa, fb)
    hd = T.FLAG_HD_FIXUP
    # every non-polar planted cell and nothing else asks for the fixup
    assert _flagged(fa, grid, hd) == sorted(_nonpolar(grid, ph, cells))
    pe, fixed = plant.brute_fixup(pb, fb, hd, grid, ph)
    if dims == "2d":
        # (0, 6) keeps one unflagged neighbour: too few
        assert (0, 6, 0) not in fixed and len(fixed) == len(_nonpolar(grid, ph, cells)) - 1
    else:
        # (0, 5, 0) and (0, 6, 0) keep two: too few in 3D
        assert (0, 5, 0) not in fixed and (0, 6, 0) not in fixed and (4, 5, 2) in fixed
        assert len(fixed) == len(_nonpolar(grid, ph, cells)) - 2
    assert np.array_equal(domain(pa, grid), domain(pe, grid), equal_nan=True)
    # rho and B are never averaged
    for c in fixed:
        assert at(pa, grid, c)[0] == at(pb, grid, c)[0]
        assert np.array_equal(at(pa, grid, c)[6:9], at(pb, grid, c)[6:9])
    # fixed cells carry p2u of their new primitives; the rest keep u
    uc = _p2u_of(grid, ph, tab, pa)
    for c in fixed:
        assert np.array_equal(at(ua, grid, c), at(uc, grid, c))
    keep = np.ones(domain(ua, grid).shape[:3], bool)
    for ix, iy, iz in fixed:
        keep[iz, iy, ix] = False
    assert np.array_equal(domain(ua, grid)[keep], domain(bad, grid)[keep])


def test_polar_rows_never_fixed():
    """Polar-axis rows skip the inversion (B only) and carry no fixup flag."""
    grid = T.grid_puffy(24, 20)
    ph = T.physics_puffy()
    tab, a, p0, u0 = _start(grid, ph, 3)
    polar_cells = [(5, 0), (7, 19), (3, 1), (20, 18)]
    a.set_state(p=p0, u=plant.negate_slot(u0, grid, polar_cells, plant.RHO))
    a.calc_u2p()
    pa, _, fa = a.get_state()
    assert _flagged(fa, grid, T.FLAG_HD_FIXUP) == []
    for c in plant.cells3(polar_cells):
        assert np.array_equal(at(pa, grid, c)[:6], at(p0, grid, c)[:6])


def test_rad_fixup_brute_force():
    """do_u2prad_fixups on: cells whose u2pRad corrected them take the
    neighbour mean of E..F^3."""
    grid = T.grid_puffy(24, 20)
    ph = _copy(T.physics_puffy(), do_u2prad_fixups=1)
    tab, a, p0, u0 = _start(grid, ph, 3)
    b = kzo.Sim(grid, _copy(ph, do_fixups=0), tab)
    cells = [(12, 14), (13, 14), (10, 8), (0, 9), (0, 10), (1, 10)]
    bad = plant.negate_slot(u0, grid, cells, plant.EE)
    for s in (a, b):
        s.set_state(p=p0, u=bad)
        s.calc_u2p()
    pa, ua, fa = a.get_state()
    pb, _, fb = b.get_state()
    assert np.array_equal(fa, fb)
    assert _flagged(fa, grid, T.FLAG_HD_FIXUP) == []
    assert _flagged(fa, grid, T.FLAG_RAD_FIXUP) == sorted(plant.cells3(cells))
    pe, fixed = plant.brute_fixup(pb, fb, T.FLAG_RAD_FIXUP, grid, ph)
    # (0, 10) keeps only (0, 11): too few
    assert (0, 10, 0) not in fixed and len(fixed) == len(cells) - 1
    assert np.array_equal(domain(pa, grid), domain(pe, grid), equal_nan=True)
    uc = _p2u_of(grid, ph, tab, pa)
    for c in fixed:
        assert np.array_equal(at(ua, grid, c), at(uc, grid, c))


@pytest.fixture(scope="module")
def radimp_case():
    grid = T.grid_puffy(24, 20)
    ph = T.physics_puffy()
    tab, s, p0, u0 = _start(grid, ph, 3)
    dt = 0.29 / s.counters()["tstepdenmax"]
    p1 = plant.plant_prims(p0, grid, plant.radimp_fail_states(ph))
    s.set_state(p=p1, u=u0)
    s.p2u_domain()
    p1, u1, _ = s.get_state()
    return grid, ph, tab, p1, u1, dt


def test_radimp_failure_flags(radimp_case):
    """A cell whose ladder fails keeps its state, is flagged -1 and counted;
    with the radimp fixup off nothing else changes it."""
    grid, ph, tab, p1, u1, dt = radimp_case
    s = kzo.Sim(grid, ph, tab)
    s.set_state(p=p1, u=u1)
    s.op_implicit(dt)
    p, u, f = s.get_state()
    planted = sorted((ix, iy, 0) for ix, iy in plant.RADIMP_FAIL_24x20.values())
    assert _flagged(f, grid, T.FLAG_RADIMP_FIXUP) == planted
    assert all(at(f, grid, c)[T.FLAG_RADIMP_FIXUP] == -1 for c in planted)
    assert s.counters()["n_radimp_failures"] == len(planted)
    for c in planted:
        assert np.array_equal(at(p, grid, c), at(p1, grid, c))
        assert np.array_equal(at(u, grid, c), at(u1, grid, c))


def test_radimp_fixup_brute_force(radimp_case):
    grid, ph, tab, p1, u1, dt = radimp_case
    on = kzo.Sim(grid, _copy(ph, do_radimp_fixups=1), tab)
    off = kzo.Sim(grid, ph, tab)
    for s in (on, off):
        s.set_state(p=p1, u=u1)
        s.op_implicit(dt)
    pa, ua, fa = on.get_state()
    pb, _, fb = off.get_state()
    assert np.array_equal(fa, fb)
    pe, fixed = plant.brute_fixup(pb, fb, T.FLAG_RADIMP_FIXUP, grid, ph)
    assert len(fixed) == len(plant.RADIMP_FAIL_24x20)
    assert np.array_equal(domain(pa, grid), domain(pe, grid), equal_nan=True)
    uc = _p2u_of(grid, ph, tab, pa)
    for c in fixed:
        assert np.array_equal(at(ua, grid, c), at(uc, grid, c))
        assert at(pa, grid, c)[0] == at(pb, grid, c)[0]
