"""GPU per-point physics vs the oracle on the reference's golden inputs.

Every device function of the step (include/kz_cellops.h) is replayed on the
same records the reference's golden tests use (koral/testing/*_golden_tests
inputs under tests/golden/) and compared with the CPU restatement.  The bar
is bit-identity: both sides evaluate the same expression shapes with fused
multiply-add disabled and the same transcendental code, so any difference
is a defect.  Where the goldens carry a reference output, the GPU result is
also held to that golden test's own tolerance.
"""
import numpy as np
import pytest

import kzo
from goldens import DevTracker, geoms, read_kgld
from koral_zig_amd import abi_types as T
from koral_zig_amd import hip

pytestmark = pytest.mark.gpu
GAM = 5.0 / 3.0


def same(a, b):
    """Bit-identical arrays (NaN payloads included)."""
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


def first_diff(a, b):
    a = np.ascontiguousarray(a, float).ravel()
    b = np.ascontiguousarray(b, float).ravel()
    bad = np.nonzero(a.view(np.int64) != b.view(np.int64))[0]
    if bad.size == 0:
        return None
    i = bad[0]
    return dict(n=int(bad.size), at=int(i), cpu=float(a[i]), gpu=float(b[i]))


@pytest.fixture(scope="module")
def P():
    return dict(floors=T.floor_params_puffy(), rad=T.rad_params_puffy(), opac=T.opac_params_puffy(),
                imp=T.implicit_params_puffy())


def _math_inputs(rng, n=20000):
    x = np.concatenate([rng.uniform(-700, 700, n), 10.0 ** rng.uniform(-300, 300, n),
                        -(10.0 ** rng.uniform(-300, 300, n // 4)), rng.uniform(-1, 1, n),
                        np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 5e-324, 1.7976931348623157e308])])
    y = np.concatenate([rng.uniform(-5, 5, x.size - 9), np.array([0.5, 2.0, -1.0, 1.0 / 3.0, 3.0, 0.0, 1.0, 2.0, 0.5])])
    return x, y


@pytest.mark.parametrize("fn", ["exp", "log", "pow", "cbrt", "log1p", "sqrt", "div"])
def test_math_bitwise(fn):
    """Device transcendentals and IEEE div/sqrt round exactly like the host."""
    rng = np.random.default_rng(5)
    x, y = _math_inputs(rng)
    if fn in ("log", "cbrt", "sqrt"):
        x = np.abs(x)
    if fn == "log1p":
        x = np.abs(x) - 0.5
    yy = y if fn in ("pow", "div") else None
    c = kzo.math_fn(fn, x, yy)
    g = hip.math_fn(fn, x, yy)
    assert same(c, g), first_diff(c, g)


def test_p2u(P):
    inp, out = read_kgld("state/p2u.kgld", 36, 13)
    pp, g = inp[:, 23:36], geoms(inp)
    c, z = kzo.p2u(pp, g, GAM), hip.p2u(pp, g, GAM)
    assert same(c, z), first_diff(c, z)


def test_flux_prime():
    inp, out = read_kgld("flux/fluxprime.kgld", 37, 13)
    idim = inp[:, 23].astype(np.int32)
    fc, ec = kzo.flux_prime(inp[:, 24:37], idim, geoms(inp), GAM)
    fg, eg = hip.flux_prime(inp[:, 24:37], idim, geoms(inp), GAM)
    assert same(ec, eg)
    assert same(fc, fg), first_diff(fc, fg)


def test_reconstruct():
    """Donor, minmod and PPM on the recon golden records (flux_golden_tests.zig:18-41)."""
    inp, out = read_kgld("flux/recon.kgld", 12, 2)
    order = np.maximum(0, 2 - inp[:, 10].astype(np.int32))
    assert set(order.tolist()) == {0, 1, 2}
    lc, rc = kzo.reconstruct(order, inp[:, 11], inp[:, 0:5], inp[:, 5:10])
    lg, rg = hip.reconstruct(order, inp[:, 11], inp[:, 0:5], inp[:, 5:10])
    assert same(lc, lg) and same(rc, rg), (first_diff(lc, lg), first_diff(rc, rg))
    assert DevTracker().add(out[:, 0], lg).add(out[:, 1], rg).ok(1e-14)


def test_u2p_solver_w():
    inp, out = read_kgld("state/u2p_solver.kgld", 49, 20)
    g = geoms(inp)
    for etype in (0, 1):
        rc, pc = kzo.u2p_solver_w(inp[:, 23:36], inp[:, 36:49], g, etype, GAM)
        rz, pz = hip.u2p_solver_w(inp[:, 23:36], inp[:, 36:49], g, etype, GAM)
        assert same(rc, rz), etype
        assert same(pc, pz), (etype, first_diff(pc, pz))


def test_gas_wavespeeds():
    inp, out = read_kgld("flux/wavespeed.kgld", 36, 4)
    c = kzo.gas_wavespeeds(inp[:, 23:36], geoms(inp), GAM, (1, 1, 0))
    z = hip.gas_wavespeeds(inp[:, 23:36], geoms(inp), GAM, (1, 1, 0))
    assert same(c, z), first_diff(c, z)
    assert DevTracker().add(out, z[:, :4]).ok(1e-13)


def test_rad_wavespeeds():
    inp, out = read_kgld("rad/rad_wavespeeds.kgld", 36, 8)
    g = geoms(inp, 20)
    c = kzo.rad_wavespeeds(inp[:, 20:33], g, inp[:, 33:36], (1, 1, 0))
    z = hip.rad_wavespeeds(inp[:, 20:33], g, inp[:, 33:36], (1, 1, 0))
    assert same(c, z), first_diff(c, z)


def test_rij():
    inp, out = read_kgld("rad/rad_rij.kgld", 33, 21)
    g = geoms(inp, 20)
    c = kzo.rij(inp[:, 20:33], g)
    z = hip.rij(inp[:, 20:33], g)
    assert same(c, z), first_diff(c, z)


def test_floors_mhd(P):
    inp, out = read_kgld("state/floors.kgld", 36, 14)
    rc, pc, nc = kzo.check_floors_mhd(inp[:, 23:36], geoms(inp), GAM, P["floors"])
    pz, rz, nz = hip.check_floors_mhd(inp[:, 23:36], geoms(inp), GAM, P["floors"])
    assert same(rc.astype(np.int32), rz)
    assert nc == nz
    assert same(pc, pz), first_diff(pc, pz)


def test_u2p_mhd(P):
    """u2p on the u2p-solver golden inputs (conserved state + guess)."""
    inp, out = read_kgld("state/u2p_solver.kgld", 49, 20)
    g = geoms(inp)
    uu, guess = inp[:, 23:36], inp[:, 36:49]
    rc, pc = kzo.u2p_mhd(uu, guess, g, GAM, P["floors"])
    pz, rz = hip.u2p_mhd(uu, guess, g, GAM, P["floors"])
    for k in ("ret", "entropy_used", "fixup"):
        assert same(rc[k], rz[k]), k
    assert same(pc, pz), first_diff(pc, pz)


def test_u2p_rad_and_floors(P):
    inp, out = read_kgld("rad/rad_u2prad.kgld", 31, 6)
    n = len(inp)
    uu = np.zeros((n, 13))
    uu[:, 9:13] = inp[:, 23:27]
    pp = np.zeros((n, 13))
    pp[:, 9:13] = inp[:, 27:31]
    rc, pc = kzo.u2p_rad(uu, pp, geoms(inp), P["rad"])
    pz, rz = hip.u2p_rad(uu, pp, geoms(inp), P["rad"])
    assert same(rc["ret"], rz["ret"]) and same(rc["corrected"], rz["corrected"])
    assert same(pc, pz), first_diff(pc, pz)
    inp, out = read_kgld("rad/rad_floors.kgld", 36, 14)
    rc, pc = kzo.check_floors_rad(inp[:, 23:36], geoms(inp), P["rad"])
    pz, rz = hip.check_floors_rad(inp[:, 23:36], geoms(inp), P["rad"])
    assert same(rc.astype(np.int32), rz)
    assert same(pc, pz), first_diff(pc, pz)


def test_radstate(P):
    inp, out = read_kgld("rad/rad_thermo.kgld", 36, 13)
    pp, g = inp[:, 23:36], geoms(inp)
    for full in (1, 0):
        c = kzo.radstate(pp, g, GAM, P["opac"], full)
        z = hip.radstate(pp, g, GAM, P["opac"], full)
        cols = [0, 1, 3, 4, 5, 6, 7, 11]
        assert same(c[:, cols], z[:, :8]), (full, first_diff(c[:, cols], z[:, :8]))
        if full:
            assert same(c[:, 13], z[:, 8])
    o = kzo.opac_entry(pp, g, GAM, P["opac"])
    assert same(o[:, 3], z[:, 9]), first_diff(o[:, 3], z[:, 9])


def test_implicit(P):
    """The whole 6-rung Newton ladder: outcome, rung, iteration count and the
    solved state identical to the oracle on every golden record."""
    inp, out = read_kgld("rad/rad_implicit.kgld", 37, 18)
    g = geoms(inp)
    dt = inp[:, 23]
    pp0 = inp[:, 24:37]
    rc, pc, uc = kzo.solve_implicit(pp0, g, dt, GAM, P["rad"], P["opac"], P["imp"])
    u0 = kzo.p2u(pp0, g, GAM)
    uz, pz, rz = hip.solve_implicit(u0, pp0, g, dt, GAM, P["rad"], P["opac"], P["imp"])
    for k in ("ok", "rung", "iters", "whichprim", "whicheq", "whichframe"):
        assert same(rc[k], rz[k]), (k, first_diff(rc[k].astype(float), rz[k].astype(float)))
    ok = rc["ok"] != 0
    assert same(pc[ok], pz[ok]), first_diff(pc[ok], pz[ok])
    assert same(uc[ok], uz[ok]), first_diff(uc[ok], uz[ok])
