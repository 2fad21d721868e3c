"""Pin the oracle's radiative shear viscosity and mean-field dynamo.

The shear/viscosity C goldens (init/puffy_{shear,visc}.kini.gz) are absent
from the reference tree, so the radiative viscosity is pinned by the
reference's own theory gates (koral/tests/radvisc_tests.zig,
koral/physics/radvisc.zig unit tests): exact limiter values, a hand-built
flat-space shear, and the kinematic invariants of sigma on a PUFFY state.
The dynamo is pinned by its law gates (koral/tests/dynamo_tests.zig) and by
the C golden of one apply_dynamo on the 384x360 PUFFY state
(koral/tests/golden/dynamo_golden_tests.zig, field-scale 3e-3).
"""
import math

import numpy as np
import pytest

import kzo
from goldens import read_kini
from koral_zig_amd import abi_types as T
from koral_zig_amd import host

PI = math.pi


# ---- radiative viscosity ----------------------------------------------------

def test_visc_coeff_limiter_chain():
    """radvisc.zig:245-270 — free, mfp-capped, horizon kill, nu cap."""
    rhor = 2.0
    rmin = 1.2 * rhor
    step_far = kzo.step_function(100.0 - rmin, 0.2 * rmin)
    assert kzo.visc_coeff(10.0, 1.0, 100.0, rhor, 0.1, 1e-30) == 0.1 * ((1.0 / 10.0) * step_far)
    assert kzo.visc_coeff(1e-100, 1.0, 100.0, rhor, 0.1, 1e-30) == 0.1 * (100.0 * step_far)
    assert kzo.visc_coeff(10.0, 1.0, rhor, rhor, 0.1, 1e-30) == 0.0
    assert kzo.visc_coeff(10.0, 0.5, 100.0, rhor, 0.1, 1e10) == 0.5 * 0.5 / 2.0 / 1e10 / 2.0


def test_step_function_closed_form():
    """math/misc.zig:9-12 — 1/(1+exp(-2kx)), k = 1.47222/x9: 1/2 at 0, the
    x9 point at ~0.95, odd symmetry f(x) + f(-x) = 1."""
    assert kzo.step_function(0.0, 1.0) == 0.5
    assert abs(kzo.step_function(1.0, 1.0) - 1.0 / (1.0 + math.exp(-2.0 * 1.47222))) < 1e-15
    for x in (0.3, 2.0, 7.0):
        assert abs(kzo.step_function(x, 1.3) + kzo.step_function(-x, 1.3) - 1.0) < 1e-15


def _mink_geom():
    row = np.zeros(T.GEOM_DOUBLES)
    gg = np.zeros((4, 5))
    gg[:, :4] = np.diag([-1.0, 1.0, 1.0, 1.0])
    row[0:20] = gg.ravel()
    row[20:40] = gg.ravel()
    row[40], row[41] = 1.0, 1.0
    return row


def test_shear_flat_space():
    """radvisc_tests.zig:190-235 — a pure shear field in Minkowski at rest gives
    sigma_12 = sigma_21 = s and zeros elsewhere; zero gradient -> zero tensor."""
    s = 0.3
    du = np.zeros((4, 4))
    du[1, 2] = du[2, 1] = s
    z = np.zeros((4, 4))
    sig = kzo.shear_from_gradients(du, z, [1, 0, 0, 0], [-1, 0, 0, 0], _mink_geom(), np.zeros(64))
    expect = np.zeros((4, 4))
    expect[1, 2] = expect[2, 1] = s
    assert np.max(np.abs(sig - expect)) <= 1e-15
    sig0 = kzo.shear_from_gradients(z, z, [1, 0, 0, 0], [-1, 0, 0, 0], _mink_geom(), np.zeros(64))
    assert np.array_equal(sig0, np.zeros((4, 4)))


@pytest.fixture(scope="module")
def puffy_32x40():
    grid = T.grid_puffy(32, 40)
    phys = T.physics_puffy(radviscosity=True, dynamo=True)
    tab = host.metric_tables(grid)
    sim = kzo.Sim(grid, phys, tab)
    host.init_puffy(sim, grid, tab, phys)
    sim.set_bl_tables(host.bl_tables(grid))
    return grid, phys, tab, sim


def test_shear_invariants(puffy_32x40):
    """radvisc_tests.zig:140-188 — sigma_mu_nu is symmetric and orthogonal to
    the radiation-frame 4-velocity on an interior torus cell."""
    grid, phys, tab, sim = puffy_32x40
    p, _, _ = sim.get_state()
    (sz, sy, sx), (ngz, ngy, ngx) = T.padded_shape(grid)
    cell = None
    for iy in range(grid.ny // 2, grid.ny):
        for ix in range(grid.nx // 3, grid.nx):
            if p[0, iy + ngy, ix + ngx, 0] > 1e-20:
                cell = (ix, iy)
                break
        if cell:
            break
    assert cell is not None
    ix, iy = cell
    sig = sim.shear_lab(ix, iy)
    smag = np.max(np.abs(sig))
    assert smag > 0
    assert np.max(np.abs(sig - sig.T)) <= 1e-12 * smag
    # u^mu of the radiation frame from the host's VELR->VEL4 restatement
    geom = host.metric_point(3, [0.0, grid.minx + (ix + 0.5) * (grid.maxx - grid.minx) / grid.nx,
                                 grid.miny + (iy + 0.5) * (grid.maxy - grid.miny) / grid.ny, 0.0])
    g = geom[0:16].reshape(4, 4)
    G = geom[16:32].reshape(4, 4)
    v = np.array([0.0, *p[0, iy + ngy, ix + ngx, 10:13]])
    qsq = v[1:] @ g[1:, 1:] @ v[1:]
    ag = math.sqrt((-1.0 / G[0, 0]) * (1.0 + qsq))
    ucon = np.array([-ag * G[0, 0], v[1] - ag * G[0, 1], v[2] - ag * G[0, 2], v[3] - ag * G[0, 3]])
    assert np.max(np.abs(sig @ ucon)) <= 1e-9 * smag


def test_rij_visc_ring_and_corners(puffy_32x40):
    """rijvisc.zig:266-311 — R^i_j filled on the domain and the one-cell ring,
    corners left zero, finite everywhere it is filled."""
    grid, phys, tab, sim = puffy_32x40
    r = sim.calc_rij_visc_total(1.0)[0]
    ng = 3
    ring = r[ng - 1:ng + grid.ny + 1, ng - 1:ng + grid.nx + 1]
    assert np.isfinite(ring).all()
    for (jy, jx) in ((0, 0), (0, -1), (-1, 0), (-1, -1)):
        assert np.array_equal(ring[jy, jx], np.zeros(16))
    assert np.count_nonzero(r[ng:ng + grid.ny, ng:ng + grid.nx]) > 0
    outside = r.copy()
    outside[ng - 1:ng + grid.ny + 1, ng - 1:ng + grid.nx + 1] = 0
    assert np.array_equal(outside, np.zeros_like(outside))


# ---- dynamo ------------------------------------------------------------------

def test_damp_bphi_formula():
    """dynamo_tests.zig:84-120 — DAMPBETA saturation, monotonicity, no overshoot."""
    dp = T.dynamo_params_puffy()
    ab, bs = dp.alphabeta, dp.betasaturated
    assert kzo.damp_bphi(ab, 1, 1, 1e-3, 0.05, bs, 2.0) == 0.0
    assert kzo.damp_bphi(ab, 1, 1, 1e-3, bs, bs, 2.0) == 0.0
    d = kzo.damp_bphi(ab, 1, 1, 1e-4, 0.3, bs, 2.0)
    assert d < 0 and abs(2.0 + d) < 2.0 and (2.0 + d) * 2.0 >= 0
    assert kzo.damp_bphi(ab, 1, 1, 1e-4, 0.3, bs, -2.0) > 0
    assert kzo.damp_bphi(ab, 1, 1, 1.0, 5.0, bs, 2.0) == -2.0
    assert kzo.damp_bphi(ab, 0, 1, 1e-3, 5.0, bs, 2.0) == 0.0
    assert kzo.damp_bphi(ab, 1, 0, 1e-3, 5.0, bs, 2.0) == 0.0
    assert abs(kzo.damp_bphi(ab, 1, 1, 1e-5, 0.5, bs, 1.0)) > abs(kzo.damp_bphi(ab, 1, 1, 1e-5, 0.2, bs, 1.0))


def _active(dp, th, bphi, r=10.0):
    return kzo.dynamo_delta_a(dp, 0.0, 1.0, 2.0, 6.0, r, th, bphi, 1.0, 0.0, 1.0, 100.0, 0.0)


def test_dynamo_law_gates():
    """dynamo_tests.zig:348-390 — equatorial sign flip, height window, horizon skip."""
    dp = T.dynamo_params_puffy()
    dp.calchronthego = 0
    dp.dampbeta = 0
    north, _ = _active(dp, PI / 2.0 - 0.1, 1e-3)
    south, _ = _active(dp, PI / 2.0 + 0.1, 1e-3)
    assert north < 0 < south
    assert abs(north + south) <= 1e-12 * abs(north)
    dp.dampbeta = 1
    hrd = dp.expectedhr * PI / 2.0
    inside, _ = _active(dp, PI / 2.0 - 0.5 * hrd, 1e-3)
    outside, b_out = _active(dp, PI / 2.0 - 1.5 * hrd, 1e-3)
    assert inside != 0.0 and outside == 0.0 and b_out == 1e-3
    aphi, b = _active(dp, PI / 2.0 - 0.1, 7e-4, r=2.0)
    assert aphi == 0.0 and b == 7e-4


def test_isco_and_horizon():
    """metric.zig:20-32 — Schwarzschild 2 and 6; Kerr a=0.9375 ISCO ~2.044."""
    assert kzo.r_horizon_bl(0.0) == 2.0 and kzo.r_isco_bl(0.0) == 6.0
    assert abs(kzo.r_isco_bl(0.9375) - 2.0442013) < 1e-6
    assert abs(kzo.r_horizon_bl(0.9375) - (1.0 + math.sqrt(1.0 - 0.9375 ** 2))) < 1e-15


def _inject_b3(sim, grid, tab, phys, factor=10.0):
    p, u, f = sim.get_state()
    ng = 3
    dom = (0, slice(ng, ng + grid.ny), slice(ng, ng + grid.nx))
    p[dom + (8,)] = factor * p[dom + (7,)]
    sim.set_state(p=p)
    sim.p2u_domain()


def test_dynamo_sign_flip_and_damping():
    """dynamo_tests.zig:181-264 — after one mimicDynamo with B3 = 10 B2
    injected: |B3| never grows nor flips sign, and Delta A_phi has the sign of
    -(pi/2 - theta) B3 wherever it acts."""
    grid = T.grid_puffy(48, 44)
    phys = T.physics_puffy(dynamo=True)
    tab = host.metric_tables(grid)
    sim = kzo.Sim(grid, phys, tab)
    host.init_puffy(sim, grid, tab, phys)
    bl = host.bl_tables(grid)
    sim.set_bl_tables(bl)
    _inject_b3(sim, grid, tab, phys)
    ng = 3
    b3_before = sim.get_state()[0][0, ng:ng + grid.ny, ng:ng + grid.nx, 8].copy()
    a = sim.mimic_dynamo(10.0)[0, ng:ng + grid.ny, ng:ng + grid.nx]
    b3_after = sim.get_state()[0][0, ng:ng + grid.ny, ng:ng + grid.nx, 8]
    assert (np.abs(b3_after) <= np.abs(b3_before) + 1e-30).all()
    assert (b3_after * b3_before >= 0).all()
    th = bl.reshape(1, grid.ny + 6, grid.nx + 6, host.BL_REC)[0, ng:ng + grid.ny, ng:ng + grid.nx, 35]
    expected = -(PI / 2.0 - th) * b3_before
    act = (np.abs(a) > 1e-40) & (np.abs(expected) > 1e-40)
    assert act.sum() > 0
    assert (a[act] * expected[act] > 0).all()


@pytest.mark.slow
def test_dynamo_golden():
    """dynamo_golden_tests.zig:55-120 — one apply_dynamo (dt = 10) on the
    384x360 PUFFY t=0 state with B3 = 10 B2 injected, B1..B3 over the domain vs
    C at field-scale 3e-3 (the C init's quadrature floor)."""
    grid = T.grid_puffy(384, 360)
    phys = T.physics_puffy(dynamo=True)
    tab = host.metric_tables(grid)
    sim = kzo.Sim(grid, phys, tab)
    host.init_puffy(sim, grid, tab, phys)
    sim.set_bl_tables(host.bl_tables(grid))
    _inject_b3(sim, grid, tab, phys)
    sim.apply_dynamo(10.0)
    p, _, _ = sim.get_state()
    k = read_kini("init/puffy_dynamo.kini.gz")
    ng = 3
    for c, slot in enumerate((6, 7, 8)):
        cval = k["data"][0, :, :, c]
        zval = p[0][np.ix_(k["cy"] + ng, k["cx"] + ng)][:, :, slot]
        dev = np.max(np.abs(cval - zval)) / np.max(np.abs(cval))
        assert dev <= 3e-3, (slot, dev)
