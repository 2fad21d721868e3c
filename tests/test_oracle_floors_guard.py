"""checkFloorsMhd's drift-frame guard ladder (koral/solve/invert.zig:552-647).

The reference's floors golden records never reach the guards, so crafted
records (plant.guard_records: strong fields, fast flows) drive the b^2/rho
and b^2/u floors into the drift-frame velocity update, where the damping of
vpar and the bail-outs fire.  What the mathematics fixes, checked with an
independent numpy evaluation of b^2 (relele velrToVel4 and mhd
bconBcovBsqFrom4vel written from their definitions):

* mass and internal energy are raised by f = b^2 / (b2rhoratiomax rho) and
  fuu = b^2 / (b2uuratiomax u) exactly when those ratios are exceeded;
* a guarded update either keeps the pre-floor velocity (a bail-out, then
  at most the gamma ceiling's uniform rescale) or ends inside the light
  cone with u^t < gammamaxhd (the damping loop's exit condition), or, when
  fifty halvings of vpar leave it outside, at the pure drift velocity
  (u + beta_par b, beta_par = -b^t / (b^2 u^t)) or the gamma ceiling;
* the Lorentz factor never ends above gammamaxhd;
* a floored record's entropy is re-synced to S(rho, u) (hydro.sFromU);
* the guard counter counts records: a batch counts what its records count
  one at a time.
"""
import numpy as np

import kzo
import plant
from koral_zig_amd import abi_types as T

GAM = 5.0 / 3.0


def _metric(geom):
    n = len(geom)
    g = geom[:, 0:20].reshape(n, 4, 5)[:, :, :4]
    G = geom[:, 20:40].reshape(n, 4, 5)[:, :, :4]
    return g, G, geom[:, 41]


def _lorentz(pp, g):
    v = pp[:, 2:5]
    return np.sqrt(1.0 + np.einsum("ni,nij,nj->n", v, g[:, 1:4, 1:4], v))


def _ucon(pp, g, G, alpha):
    """u^mu of VELR primitives: u^t = gamma / alpha, u^i = u~^i - gamma alpha G^{ti}."""
    gam = _lorentz(pp, g)
    u = np.zeros((len(pp), 4))
    u[:, 0] = gam / alpha
    u[:, 1:4] = pp[:, 2:5] - (gam * alpha)[:, None] * G[:, 0, 1:4]
    return u


def _fields(pp, geom):
    """(u^mu, b^mu, b^2): b^t = B^i u_i, b^i = (B^i + b^t u^i) / u^t."""
    g, G, alpha = _metric(geom)
    u = _ucon(pp, g, G, alpha)
    ucov = np.einsum("nij,nj->ni", g, u)
    B = pp[:, 6:9]
    bt = np.einsum("ni,ni->n", B, ucov[:, 1:4])
    b = np.zeros_like(u)
    b[:, 0] = bt
    b[:, 1:4] = (B + bt[:, None] * u[:, 1:4]) / u[:, 0:1]
    return u, b, np.einsum("ni,nij,nj->n", b, g, b)


def _bsq(pp, geom):
    return _fields(pp, geom)[2]


def _drift_vel3(pp, geom):
    """Coordinate 3-velocity of the drift frame: u + beta_par b with
    beta_par = -b^t / (b^2 u^t), the limit the damping of vpar tends to."""
    u, b, bsq = _fields(pp, geom)
    w = u + (-b[:, 0] / (bsq * u[:, 0]))[:, None] * b
    return w[:, 1:4] / w[:, 0:1]


def _vel3(pp, geom):
    g, G, alpha = _metric(geom)
    u = _ucon(pp, g, G, alpha)
    return u[:, 1:4] / u[:, 0:1]


def test_guard_ladder_pins():
    fp = T.floor_params_puffy()
    pp, geom = plant.guard_records()
    ret, po, ng = kzo.check_floors_mhd(pp, geom, GAM, fp)
    assert ng > 100
    n = len(pp) - 4
    pin, gin, out = pp[:n], geom[:n], po[:n]
    rho, uu = pin[:, 0], pin[:, 1]
    assert (rho >= fp.rhofloor).all() and (uu >= fp.uurhoratiomin * rho).all()
    uu = np.minimum(uu, fp.uurhoratiomax * rho)
    bsq = _bsq(pin, gin)
    g, _, alpha = _metric(gin)
    # b^2 of a flow of Lorentz factor gamma carries a cancellation of order
    # gamma^2 in the last place, so the two evaluations agree to that
    tol = 1e-13 * _lorentz(pin, g) ** 2
    far = (np.abs(bsq / (fp.b2rhoratiomax * rho) - 1.0) > 10 * tol) & (np.abs(bsq / (fp.b2uuratiomax * uu) - 1.0) > 10 * tol)
    f = np.where(bsq > fp.b2rhoratiomax * rho, bsq / (fp.b2rhoratiomax * rho), 1.0)
    fuu = np.where(bsq > fp.b2uuratiomax * uu, bsq / (fp.b2uuratiomax * uu), 1.0)
    assert (np.abs(out[far, 0] / (rho * f)[far] - 1.0) <= tol[far]).all()
    assert (np.abs(out[far, 1] / (uu * fuu)[far] - 1.0) <= tol[far]).all()
    floored = far & ((f > 1.0) | (fuu > 1.0))
    assert floored.sum() > n // 4
    gam_out = _lorentz(out, g)
    assert (gam_out <= fp.gammamaxhd * (1.0 + 1e-12)).all()
    vin, vout = pin[:, 2:5], out[:, 2:5]
    scale = np.einsum("ni,ni->n", vout, vin) / np.einsum("ni,ni->n", vin, vin)
    kept = np.all(np.abs(vout - scale[:, None] * vin) <= 1e-13 * np.abs(vin).max(axis=1, keepdims=True), axis=1)
    kept &= (scale > 0.0) & (scale <= 1.0 + 1e-15)
    inside = gam_out / alpha < fp.gammamaxhd
    ceiling = np.abs(gam_out / fp.gammamaxhd - 1.0) <= 1e-12
    drift = np.abs(_vel3(out, gin) - _drift_vel3(pin, gin)).max(axis=1) <= tol
    assert np.all(kept[floored] | inside[floored] | drift[floored] | ceiling[floored])
    assert (drift & ~inside & ~kept)[floored].sum() > 0
    assert inside[floored & ~kept].sum() > 100
    ncoef = 1.0 / (GAM - 1.0)
    s_exp = out[:, 0] * np.log(((GAM - 1.0) * out[:, 1]) ** ncoef / out[:, 0] ** (ncoef + 1.0))
    assert np.allclose(out[floored, 5], s_exp[floored], rtol=1e-10, atol=1e-14)
    # batch counter = sum of the records' own counts
    one = sum(kzo.check_floors_mhd(pp[i:i + 1], geom[i:i + 1], GAM, fp)[2] for i in range(300))
    assert one == kzo.check_floors_mhd(pp[:300], geom[:300], GAM, fp)[2] and one > 0


def test_guard_special_records():
    """A NaN density leaves the update through the field-direction guard
    with the pre-floor velocity; an infinite field (b^2 = NaN) and no field
    floor nothing."""
    fp = T.floor_params_puffy()
    pp, geom = plant.guard_records()
    sp, sg = pp[-4:], geom[-4:]
    for i, guarded in ((0, 1), (1, 0), (2, 0)):
        _, po, ng = kzo.check_floors_mhd(sp[i:i + 1], sg[i:i + 1], GAM, fp)
        assert ng == guarded, i
        g, _, _ = _metric(sg[i:i + 1])
        vin, vout = sp[i, 2:5], po[0, 2:5]
        gin = _lorentz(sp[i:i + 1], g)[0]
        a = np.sqrt((fp.gammamaxhd ** 2 - 1.0) / (gin * gin - 1.0)) if gin > fp.gammamaxhd else 1.0
        assert np.allclose(vout, vin * a, rtol=1e-14, atol=0), i
    _, po, _ = kzo.check_floors_mhd(sp[3:4], sg[3:4], GAM, fp)
    assert po[0, 0] >= fp.rhofloor and po[0, 1] >= fp.uurhoratiomin * fp.rhofloor
