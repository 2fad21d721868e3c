"""Planted bad cells and a brute-force restatement of what cellFixup must do.

The PUFFY runs of the step tests never flag a cell for a fixup: the torus is
smooth. These helpers write chosen cells into a state so that the u2p
inversion or the implicit solve fails there, and state, independently of
the oracle's cell_fixup, which result koral/sim.zig:1027-1133 demands:

* a flagged cell outside the polar-axis rows takes, in each slot the fixup
  type owns, the mean of its unflagged in-domain neighbours visited in the
  order x-, x+, y-, y+, z-, z+ (summed in that order, then divided by their
  count);
* it needs at least 2 such neighbours on a 2D grid and 3 on a 3D grid,
  else it keeps what the pass left;
* HD owns every slot except rho and B1..B3 below B1 (u, v^1..v^3, entropy);
  RAD owns E..F^3; RADIMP owns both blocks, never rho or B;
* the conserved vector of a fixed cell is p2u of its new primitives.
"""
import numpy as np

from koral_zig_amd import abi_types as T

RHO, B1, EE, FZ = 0, 6, 9, 12

# (ix, iy) domain cells of a 24x20 grid (polar rows 0, 1, 18, 19):
# an isolated cell, an adjacent pair, a pair of neighbours around (0, 6) that
# leaves it with one unflagged neighbour, two polar-row cells and a cell on
# the outer radial edge next to the first non-polar row
PLANT_2D = [(10, 8), (15, 12), (16, 12), (0, 6), (1, 6), (0, 5), (5, 0), (7, 19), (23, 2)]
# (ix, iy, iz) of a 16x12x6 grid (polar rows 0, 1, 10, 11): (4, 5, 2) has
# its z-, z+ and x+ neighbours flagged, which leaves 3; (0, 5, 0) and
# (0, 6, 0) sit on the inner-radius, lowest-phi domain edge with each other
# and an x+ neighbour flagged, which leaves each 2 (enough in 2D, too few in
# 3D); (1, 5, 0) and (1, 6, 0) keep 3; (9, 0, 4) is on a polar row
PLANT_3D = [(4, 5, 2), (4, 5, 1), (4, 5, 3), (5, 5, 2), (0, 5, 0), (1, 5, 0), (0, 6, 0), (1, 6, 0), (9, 0, 4)]


def fixed_slots(which):
    """Slots a fixup of flag `which` overwrites (sim.zig:1104-1113)."""
    iv = np.arange(T.NV)
    if which == T.FLAG_HD_FIXUP:
        return (iv != RHO) & (iv < B1)
    if which == T.FLAG_RAD_FIXUP:
        return (iv >= EE) & (iv <= FZ)
    return (iv != RHO) & ((iv < B1) | ((iv >= EE) & (iv <= FZ)))


def _pad(grid):
    (_, _, _), (ngz, ngy, ngx) = T.padded_shape(grid)
    return ngz, ngy, ngx


def cells3(cells):
    return [c if len(c) == 3 else (c[0], c[1], 0) for c in cells]


def negate_slot(u, grid, cells, slot):
    """Copy of the padded conserved array with u[slot] negated at `cells`:
    a negative conserved density makes u2pMhd demand a fixup
    (koral/tests/state_tests.zig:755-777); a negative radiation energy
    makes u2pRad correct the cell."""
    ngz, ngy, ngx = _pad(grid)
    u = u.copy()
    for ix, iy, iz in cells3(cells):
        u[iz + ngz, iy + ngy, ix + ngx, slot] = -u[iz + ngz, iy + ngy, ix + ngx, slot]
    return u


def polar(phys, grid, iy):
    nc = phys.nccorrectpolar if phys.correct_polaraxis else 0
    return iy < nc or iy >= grid.ny - nc


def brute_fixup(p_pass, flags, which, grid, phys):
    """Primitives after the fixup of flag plane `which`, from the pass's
    primitives `p_pass` and flags (padded arrays).  Returns (p, fixed):
    the new padded primitives and the (ix, iy, iz) cells that changed."""
    ngz, ngy, ngx = _pad(grid)
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    slots = np.nonzero(fixed_slots(which))[0]
    need = 2 if nz == 1 else 3
    p = p_pass.copy()
    fixed = []
    for iz in range(nz):
        for iy in range(ny):
            if polar(phys, grid, iy):
                continue
            for ix in range(nx):
                if flags[iz + ngz, iy + ngy, ix + ngx, which] == 0:
                    continue
                nbs = []
                for jx, jy, jz in ((ix - 1, iy, iz), (ix + 1, iy, iz), (ix, iy - 1, iz), (ix, iy + 1, iz),
                                   (ix, iy, iz - 1), (ix, iy, iz + 1)):
                    if not (0 <= jx < nx and 0 <= jy < ny and 0 <= jz < nz):
                        continue
                    if flags[jz + ngz, jy + ngy, jx + ngx, which] != 0:
                        continue
                    nbs.append(p_pass[jz + ngz, jy + ngy, jx + ngx])
                if len(nbs) < need:
                    continue
                for iv in slots:
                    acc = 0.0
                    for q in nbs:
                        acc = acc + float(q[iv])
                    p[iz + ngz, iy + ngy, ix + ngx, iv] = acc / float(len(nbs))
                fixed.append((ix, iy, iz))
    return p, fixed


def domain(a, grid):
    ngz, ngy, ngx = _pad(grid)
    return a[ngz:ngz + grid.nz, ngy:ngy + grid.ny, ngx:ngx + grid.nx]


def at(a, grid, c):
    ngz, ngy, ngx = _pad(grid)
    ix, iy, iz = c
    return a[iz + ngz, iy + ngy, ix + ngx]


# cells of the implicit-solve batch (koral_zig_amd.workloads, 20000 cells,
# seed 11) whose M1 Newton fails every rung when planted at the listed
# domain cell of a 24x20 PUFFY grid and solved with 0.29 of the CFL dt:
# batch index -> (ix, iy); (9, 12) and (10, 12) are neighbours
RADIMP_FAIL_24x20 = {283: (9, 12), 14000: (10, 12), 956: (13, 14), 2539: (13, 3), 11480: (2, 8)}


def radimp_fail_states(phys):
    import kzo
    from koral_zig_amd import workloads
    pp, geom, dt = workloads.implicit_batch_inputs(
        20000, 11, phys, lambda p, g: kzo.opac_entry(p, g, phys.gam, phys.opac)[:, 3])
    return {k: pp[k] for k in RADIMP_FAIL_24x20}


def plant_prims(p, grid, states):
    """Copy of the padded primitives with cell (ix, iy) set to states[k]."""
    ngz, ngy, ngx = _pad(grid)
    p = p.copy()
    for k, (ix, iy) in RADIMP_FAIL_24x20.items():
        p[ngz, iy + ngy, ix + ngx] = states[k]
    return p


def guard_records(n=4000, seed=1):
    """(pp[n, 13], geom[n, 47]) for checkFloorsMhd whose b^2/rho or b^2/u
    floor fires with a strong field and a fast flow, so that the drift-frame
    velocity update of koral/solve/invert.zig:552-647 meets its guards
    (mostly the subluminal damping of vpar).  Geometries are those of the
    reference's floors golden records; four special records close the set:
    a NaN density (the field-direction guard: every comparison with NaN
    fails), an infinite B^1 (b^2 is NaN: no floor), a zero field and a
    record below the density and energy floors."""
    from goldens import geoms, read_kgld
    inp, _ = read_kgld("state/floors.kgld", 36, 14)
    g0 = geoms(inp)
    rng = np.random.default_rng(seed)
    g = g0[rng.integers(0, len(g0), n)]
    pp = np.zeros((n, T.NV))
    pp[:, 0] = 10.0 ** rng.uniform(-6, 0, n)
    pp[:, 1] = pp[:, 0] * 10.0 ** rng.uniform(-3, 0, n)
    pp[:, 2:5] = rng.normal(0, 1, (n, 3)) * 10.0 ** rng.uniform(-2, 1.5, (n, 1))
    pp[:, 5] = 1.0
    pp[:, 6:9] = rng.normal(0, 1, (n, 3)) * 10.0 ** rng.uniform(-1, 2, (n, 1))
    pp[:, 9] = 1e-10
    sp = np.tile(pp[:1], (4, 1))
    sp[0, 0] = np.nan
    sp[1, 6] = np.inf
    sp[2, 6:9] = 0.0
    sp[3, 0:2] = (-1.0, -1.0)
    return np.concatenate([pp, sp]), np.concatenate([g, g[:4]])
