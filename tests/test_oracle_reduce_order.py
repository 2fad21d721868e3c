"""The two reconstruction-order options of the sweep (koral/sim.zig:1316-1337)
in the oracle.

* REDUCEORDERATBH: cells whose radial index is at most reduce_order_ix_max,
  the largest index (ghosts included) whose centre lies inside the BL
  horizon (sim.zig:481-495), reconstruct with the linear minmod scheme
  instead of PPM.  The index from the host library is checked against the
  MKS2 radius r = exp(x1) + R0 and r_H = 1 + sqrt(1 - a^2) (metric.zig:20-23).
* REDUCEORDERAFTERFIXUP: a cell that the last u2p or implicit pass flagged
  for a fixup reconstructs one order lower.  As in the reference's own test
  (koral/tests/mhd_evolution_tests.zig:448-472): on a state with no flag
  the option changes nothing; with a poisoned cell it does.

Both pins are sharpened by locality: one explicit operator with the option
on and off differs only within two cells of the cells that reconstruct
differently (their two faces per dimension, the EMF corners around them),
and the cells that differ include them.
"""
import math

import numpy as np
import pytest

import kzo
import plant
from plant import domain
from koral_zig_amd import abi_types as T
from koral_zig_amd import host


def _copy(ph, **kw):
    q = type(ph).from_buffer_copy(ph)
    for k, v in kw.items():
        setattr(q, k, v)
    return q


def _brute_ix_max(grid, a, mksr0):
    """Largest jx in [-ng, nx + ng) with r(x1c(jx)) < r_H(a); None if none."""
    dx = (grid.maxx - grid.minx) / grid.nx
    r_h = 1.0 + math.sqrt(1.0 - a * a)
    best = None
    for jx in range(-grid.ng, grid.nx + grid.ng):
        xc = 0.5 * ((grid.minx + jx * dx) + (grid.minx + (jx + 1) * dx))
        if math.exp(xc) + mksr0 < r_h:
            best = jx
    return best


@pytest.mark.parametrize("a,nx,rmin", [(0.0, 384, None), (0.5, 64, None), (0.9375, 192, None), (0.0, 64, 2.5),
                                       (0.9, 96, 1.0)])
def test_reduce_order_ix_max(a, nx, rmin):
    grid = T.grid_puffy(nx, 32, rmin=T.rmin_for_spin(a) if rmin is None else rmin)
    mp = dict(host.MP_PUFFY, a=a)
    got = host.reduce_order_ix_max(grid, mp)
    want = _brute_ix_max(grid, a, mp["mksr0"])
    if want is None:
        assert got == -2 ** 31
    else:
        assert got == want
    # the PUFFY inner radius 1.85 r_H(a) / r_H(0) puts cells inside the horizon
    if rmin is None:
        assert want is not None and want >= 0


def _state(grid, ph, seed):
    tab = host.metric_tables(grid, dict(host.MP_PUFFY, a=ph.bh_spin))
    s = kzo.Sim(grid, ph, tab)
    host.init_puffy(s, grid, tab, ph, seed=seed)
    p, u, f = s.get_state()
    return tab, p, u


def _near(diff, cells, r):
    """True where a domain cell lies within r cells (each axis) of one of `cells`."""
    near = np.zeros_like(diff)
    nz, ny, nx = diff.shape
    for ix, iy, iz in cells:
        near[max(iz - r, 0):iz + r + 1, max(iy - r, 0):iy + r + 1, max(ix - r, 0):ix + r + 1] = True
    return near


def _explicit(grid, ph, tab, p, u, dt, poison=None):
    s = kzo.Sim(grid, ph, tab)
    s.set_state(p=p, u=u if poison is None else plant.negate_slot(u, grid, poison, plant.RHO))
    s.calc_u2p()
    flags = s.get_state()[2]
    s.op_explicit(dt)
    return s.get_state(), flags


@pytest.mark.parametrize("dims", ["2d", "3d"])
def test_reduceorderafterfixup(dims):
    if dims == "2d":
        grid, seed, poison = T.grid_puffy(24, 20), 3, [(10, 8), (15, 12), (16, 12), (4, 5)]
    else:
        grid, seed, poison = T.grid_puffy(16, 12, 6), 4, [(4, 5, 2), (9, 7, 3), (10, 7, 3)]
    ph = T.physics_puffy()
    on = _copy(ph, reduceorderafterfixup=1)
    tab, p, u = _state(grid, ph, seed)
    dt = 1e-3
    # no flag set: the option is inert, bit for bit
    (pa, ua, _), fa = _explicit(grid, on, tab, p, u, dt)
    (pb, ub, _), _ = _explicit(grid, ph, tab, p, u, dt)
    assert not np.any(domain(fa, grid)[..., T.FLAG_HD_FIXUP])
    assert np.array_equal(ua, ub) and np.array_equal(pa, pb)
    # poisoned cells: the same u2p failure and fixup in both runs; the
    # flagged cells reconstruct linearly with the option on
    (pa, ua, _), fa = _explicit(grid, on, tab, p, u, dt, poison)
    (pb, ub, _), fb = _explicit(grid, ph, tab, p, u, dt, poison)
    assert np.array_equal(fa, fb)
    flagged = [(int(ix), int(iy), int(iz)) for iz, iy, ix in
               np.argwhere(domain(fa, grid)[..., T.FLAG_HD_FIXUP] != 0)]
    assert sorted(flagged) == sorted(plant.cells3(poison))
    diff = np.any(domain(ua, grid) != domain(ub, grid), axis=-1)
    assert diff.any()
    assert not np.any(diff & ~_near(diff, flagged, 2))
    for ix, iy, iz in flagged:
        assert diff[iz, iy, ix]


@pytest.mark.parametrize("a,dims", [(0.0, "2d"), (0.9375, "2d"), (0.5, "3d")])
def test_reduceorderatbh(a, dims):
    # radially fine enough that domain cells lie inside the horizon (the inner
    # x ghosts copy cell 0, so ghost-only thresholds change nothing)
    if dims == "2d":
        grid, seed = T.grid_puffy(192, 16, rmin=T.rmin_for_spin(a)), 3
    else:
        grid, seed = T.grid_puffy(128, 12, 6, rmin=T.rmin_for_spin(a)), 4
    ph = T.physics_puffy(bh_spin=a)
    ix_max = host.reduce_order_ix_max(grid, dict(host.MP_PUFFY, a=a))
    assert 1 <= ix_max <= 2
    on = _copy(ph, reduceorderatbh=1, reduce_order_ix_max=ix_max)
    tab, p, u = _state(grid, ph, seed)
    dt = 1e-3
    (pa, ua, _), _ = _explicit(grid, on, tab, p, u, dt)
    (pb, ub, _), _ = _explicit(grid, ph, tab, p, u, dt)
    diff = np.any(domain(ua, grid) != domain(ub, grid), axis=-1)
    nz, ny, nx = diff.shape
    # cells up to ix_max (and the ghosts below) reconstruct linearly: the
    # update differs next to them and at most two cells further out
    assert diff[:, :, :max(ix_max, 0) + 1].any()
    assert not diff[:, :, ix_max + 3:].any()
    # the same threshold one index lower moves the boundary of the change
    lower = _copy(on, reduce_order_ix_max=ix_max - 1)
    (pc, uc, _), _ = _explicit(grid, lower, tab, p, u, dt)
    d2 = np.any(domain(uc, grid) != domain(ua, grid), axis=-1)
    assert d2.any() and not d2[:, :, ix_max + 3:].any()
