"""Theory gates of the implicit radiation-gas solver on the oracle.

Restated from koral/tests/implicit_tests.zig:89-431 (the reference's M9
gates): properties the mathematics fixes, independent of any golden file.

* one rung (MHD primitives, energy equation, lab frame) conserves the
  gas+radiation totals of its internal conserved vector to 1e-14: the
  constraint builds one fluid from the other (implicit_tests.zig:89-131);
* a comoving LTE state is a fixed point of the full solve (:137-160);
* many small implicit steps follow an adaptive RK45 (Cash-Karp) integral of
  dE/dtau = -G^0(E) to 1e-6 (:166-260);
* one step with dt a million relaxation times lands on the equilibrium,
  monotonically and without overshoot (:266-300);
* the RAD- and MHD-primitive rungs agree to 1e-5 where both truly converge
  (:320-387);
* random hostile states never give non-finite output, and failures leave
  the primitives untouched (:389-431).

The reference draws its states from Zig's PRNG; here numpy's generator
draws from the same distributions, so the counts differ and the gates are
the reference's.
"""
import numpy as np
import pytest

import kzo
from koral_zig_amd import abi_types as T
from koral_zig_amd import host

GAM = 5.0 / 3.0
MINK, KS = 0, 2


def _geom(coords, x):
    return host.geom_batch(coords, np.array([x], float))


def _brems(op):
    q = type(op).from_buffer_copy(op)
    q.synchrotron = 0
    q.comptonization = 0
    return q


def pp_from_temps(op, rho, tgas, v, erf, f, b):
    """implicit_tests.zig:51-75 ppFromTemps."""
    uint = tgas * rho * op.kb_over_mugas_mp / (GAM - 1.0)
    n = 1.0 / (GAM - 1.0)
    pp = np.zeros(T.NV)
    pp[0], pp[1] = rho, uint
    pp[2:5] = v
    pp[5] = rho * np.log(((GAM - 1.0) * uint) ** n / rho ** (n + 1.0))
    pp[6:9] = b
    pp[9] = erf
    pp[10:13] = f
    return pp


def lte_e(op, t):
    """thermo.zig:147-149 lteEfromT."""
    return op.four_sigmarad * t * t * t * t


@pytest.fixture(scope="module")
def P():
    ph = T.physics_puffy()
    return ph.rad, ph.opac, ph.implicit


def test_internal_conservation(P):
    rp, op, ip = P
    rng = np.random.default_rng(0x4d39636f6e)
    n = 800
    g = np.repeat(_geom(MINK, [0, 0, 0, 0]), n, axis=0)
    pp = np.zeros((n, T.NV))
    dt = np.zeros(n)
    for i in range(n):
        rho = 10.0 ** (-11.0 + 5.0 * rng.random())
        tgas = 10.0 ** (6.0 + 5.0 * rng.random())
        zeta = 10.0 ** (-1.5 + 3.0 * rng.random())
        v = [0.3 * (2 * rng.random() - 1), 0.2 * (2 * rng.random() - 1), 0.0]
        fr = [0.3 * (2 * rng.random() - 1), 0.0, 0.1 * (2 * rng.random() - 1)]
        pp[i] = pp_from_temps(op, rho, tgas, v, lte_e(op, zeta * tgas), fr, [0, 0, 0])
        dt[i] = 10.0 ** (-8.0 + 8.0 * rng.random())
    uu00 = kzo.p2u(pp, g, GAM)
    ok, _, _, uu = kzo.solve_4d_prim(pp, g, dt, GAM, rp, op, ip, kzo.WP_MHD)
    m = ok == 1
    # one rung over kappa dt spanning ten decades: many fail by design (the
    # reference's draws converge > 150 of 400; numpy's converge about 37 %)
    assert m.sum() > n // 3
    tot0 = uu00[m, 1:5] + uu00[m, 9:13]
    tot1 = uu[m, 1:5] + uu[m, 9:13]
    scale = np.maximum(np.abs(uu00[m, 1:5]), np.abs(uu00[m, 9:13]))
    worst = max((np.abs(tot1 - tot0) / scale).max(), (np.abs(uu[m, 0] - uu00[m, 0]) / uu00[m, 0]).max())
    assert worst <= 1e-14, worst


@pytest.mark.parametrize("tgas", [1.0e6, 1.0e8, 1.0e10])
def test_lte_fixed_point(P, tgas):
    rp, op, ip = P
    op = _brems(op)
    g = _geom(MINK, [0, 0, 0, 0])
    pp0 = pp_from_temps(op, 1.0e-8, tgas, [0, 0, 0], lte_e(op, tgas), [0, 0, 0], [0, 0, 0])[None]
    r, pp, _ = kzo.solve_implicit(pp0, g, 1.0, GAM, rp, op, ip)
    assert r["ok"][0] == 1
    nz = pp0[0] != 0.0
    assert (np.abs(pp[0, nz] - pp0[0, nz]) / np.abs(pp0[0, nz]) <= 1e-6).all()
    assert (np.abs(pp[0, ~nz]) <= 1e-8).all()


def _relax_rhs(op, geom, e, rho, utot):
    """dE/dtau = -G^0 in the fluid frame at rest (implicit_tests.zig:166-180)."""
    pp = pp_from_temps(op, rho, 1.0, [0, 0, 0], e, [0, 0, 0], [0, 0, 0])
    n = 1.0 / (GAM - 1.0)
    pp[1] = utot - e
    pp[5] = rho * np.log(((GAM - 1.0) * pp[1]) ** n / rho ** (n + 1.0))
    out, _ = kzo.gi(pp[None], geom, GAM, op)
    return -out[0, 4]


def _rk45(f, y0, t_end, rtol):
    """Adaptive Cash-Karp 4(5) integration of y' = f(y) to t_end."""
    a = [[], [0.2], [3 / 40, 9 / 40], [0.3, -0.9, 1.2], [-11 / 54, 2.5, -70 / 27, 35 / 27],
         [1631 / 55296, 175 / 512, 575 / 13824, 44275 / 110592, 253 / 4096]]
    b5 = [37 / 378, 0, 250 / 621, 125 / 594, 0, 512 / 1771]
    b4 = [2825 / 27648, 0, 18575 / 48384, 13525 / 55296, 277 / 14336, 0.25]
    y, t, h = y0, 0.0, t_end / 100.0
    while t < t_end:
        h = min(h, t_end - t)
        k = []
        for i in range(6):
            k.append(f(y + h * sum(a[i][j] * k[j] for j in range(i))))
        y5 = y + h * sum(b5[i] * k[i] for i in range(6))
        y4 = y + h * sum(b4[i] * k[i] for i in range(6))
        err = abs(y5 - y4) / (rtol * abs(y5))
        if err <= 1.0:
            t += h
            y = y5
        h *= min(5.0, max(0.1, 0.9 * (1.0 / max(err, 1e-10)) ** 0.2))
    return y


def test_relaxation_matches_rk45(P):
    rp, op, ip = P
    op = _brems(op)
    ip = type(ip).from_buffer_copy(ip)
    ip.conv, ip.convrel, ip.convrelerr = 1.0e-13, 1.0e-11, 1.0e-11
    g = _geom(MINK, [0, 0, 0, 0])
    rho, tgas0 = 1.0e-8, 1.0e8
    e0 = lte_e(op, 0.5 * tgas0)
    pp = pp_from_temps(op, rho, tgas0, [0, 0, 0], e0, [0, 0, 0], [0, 0, 0])[None]
    utot = pp[0, 1] + e0
    tau = abs(e0 / _relax_rhs(op, g, e0, rho, utot))
    nstep = 20000
    t_end = 0.1 * tau
    dt = t_end / nstep
    for _ in range(nstep):
        r, pp, _ = kzo.solve_implicit(pp, g, dt, GAM, rp, op, ip)
        assert r["ok"][0] == 1
    e_ref = _rk45(lambda e: _relax_rhs(op, g, e, rho, utot), e0, t_end, 1e-12)
    assert abs(pp[0, 9] - e_ref) / e_ref <= 1e-6
    assert pp[0, 9] > e0


def test_l_stability(P):
    rp, op, ip = P
    op = _brems(op)
    g = _geom(MINK, [0, 0, 0, 0])
    rho, tgas0 = 1.0e-8, 1.0e8
    e0 = lte_e(op, 0.2 * tgas0)
    pp = pp_from_temps(op, rho, tgas0, [0, 0, 0], e0, [0, 0, 0], [0, 0, 0])[None]
    utot = pp[0, 1] + e0
    lo, hi = e0, utot * 0.99
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if _relax_rhs(op, g, mid, rho, utot) > 0.0:
            lo = mid
        else:
            hi = mid
    e_eq = 0.5 * (lo + hi)
    tau = abs(e0 / _relax_rhs(op, g, e0, rho, utot))
    r, pp, _ = kzo.solve_implicit(pp, g, 1.0e6 * tau, GAM, rp, op, ip)
    assert r["ok"][0] == 1
    e1 = pp[0, 9]
    assert e0 < e1 <= e_eq * (1.0 + 1e-3)
    assert abs(e1 - e_eq) < 0.01 * abs(e0 - e_eq)


def test_branch_agreement(P):
    rp, op, ip = P
    rng = np.random.default_rng(0x4d39627261)
    n = 600
    g = np.repeat(_geom(MINK, [0, 0, 0, 0]), n, axis=0)
    pp00 = np.zeros((n, T.NV))
    dt = np.zeros(n)
    for i in range(n):
        rho = 10.0 ** (-10.0 + 4.0 * rng.random())
        tgas = 10.0 ** (7.0 + 3.0 * rng.random())
        zeta = 10.0 ** (-1.0 + 2.0 * rng.random())
        pp00[i] = pp_from_temps(op, rho, tgas, [0.1, 0, 0], lte_e(op, zeta * tgas), [0, 0.1, 0], [0, 0, 0])
        dt[i] = 10.0 ** (-6.0 + 5.0 * rng.random())
    okr, _, ppr, _ = kzo.solve_4d_prim(pp00, g, dt, GAM, rp, op, ip, kzo.WP_RAD)
    okm, _, ppm, _ = kzo.solve_4d_prim(pp00, g, dt, GAM, rp, op, ip, kzo.WP_MHD)
    er = kzo.true_residual(ppr, pp00, dt, g, GAM, op, kzo.WP_RAD, kzo.WE_ENERGY, kzo.WF_LAB)
    em = kzo.true_residual(ppm, pp00, dt, g, GAM, op, kzo.WP_MHD, kzo.WE_ENERGY, kzo.WF_LAB)
    both = (okr == 1) & (okm == 1) & (er < 1e-8) & (em < 1e-8)
    assert both.sum() >= 3
    vel = np.zeros(T.NV, bool)
    vel[2:5] = vel[10:13] = True
    worst = 0.0
    for i in np.nonzero(both)[0]:
        vmax = max(1e-14, np.abs(ppr[i, vel]).max(), np.abs(ppm[i, vel]).max())
        scale = np.where(vel, vmax, np.maximum(np.maximum(np.abs(ppr[i]), np.abs(ppm[i])), 1e-300))
        worst = max(worst, (np.abs(ppr[i] - ppm[i]) / scale).max())
    assert worst <= 1e-5, worst


def fuzz_states(op, n=3000, seed=0x4d39667a):
    """implicit_tests.zig:389-431 state sweep: (pp0, geom, dt)."""
    rng = np.random.default_rng(seed)
    geoms = [_geom(MINK, [0, 0, 0, 0])[0], _geom(KS, [0, 7.3, 1.1, 0.4])[0]]
    pp = np.zeros((n, T.NV))
    gg = np.zeros((n, T.GEOM_DOUBLES))
    dt = np.zeros(n)
    for i in range(n):
        geo = geoms[i % 2]
        gij = geo[0:20].reshape(4, 5)[1:4, 1:4]
        vecs = []
        for _ in range(2):
            d = 2.0 * rng.random(3) - 1.0
            q = d @ gij @ d
            gt = 1.0 + 2.0 * rng.random()
            vecs.append(d * np.sqrt((gt * gt - 1.0) / q))
        rho = 10.0 ** (-14.0 + 10.0 * rng.random())
        tgas = 10.0 ** (4.0 + 9.0 * rng.random())
        zeta = 10.0 ** (-3.0 + 6.0 * rng.random())
        uint = tgas * rho * op.kb_over_mugas_mp / (GAM - 1.0)
        beta = 10.0 ** (-2.0 + 4.0 * rng.random())
        bmag = np.sqrt(2.0 * (GAM - 1.0) * uint / beta)
        pp[i] = pp_from_temps(op, rho, tgas, vecs[0], lte_e(op, zeta * tgas), vecs[1], [bmag, 0, 0])
        gg[i] = geo
        dt[i] = 10.0 ** (-10.0 + 10.0 * rng.random())
    return pp, gg, dt


def test_fuzz_clean_failures(P):
    rp, op, ip = P
    pp0, g, dt = fuzz_states(op)
    r, pp, uu = kzo.solve_implicit(pp0, g, dt, GAM, rp, op, ip)
    ok = r["ok"] == 1
    assert np.isfinite(pp[ok]).all() and np.isfinite(uu[ok]).all()
    assert np.array_equal(pp[~ok], pp0[~ok])
    assert (~ok).sum() < len(pp0) // 2
