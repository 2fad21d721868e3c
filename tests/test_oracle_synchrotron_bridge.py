"""The synchrotron bridge of the opacities (koral/physics/opacities.zig:
70-74, 208-297; USE_SYNCHROTRON_BRIDGE_FUNCTIONS) in the oracle.

The reference has no test of its own for the bridge.  Its definition ties
the bridge-on opacities to the bridge-off ones channel by channel, and
those ties are checked here from three oracle evaluations of each state
(bridge off; bridge on; synchrotron off, which leaves the free-free part)
with the non-relativistic terms written out in numpy from their formulas:

* gas absorption: the bridge drops the Terelfactor suppression
  theta^2 / (1 + theta^2), theta = Te k / (m_e c^2), and adds
  2.35869e-21 (n_e/rho) B^2 / Te^3 (cgs, times rho in code units);
* radiation absorption where no clamp applies (Trad <= Te, since the clamp
  Trad < TradBB^(4/3) / Te^(1/3) with TradBB = Trad reads Trad > Te): the
  same with Trad in place of Te in the added term;
* number opacity (no clamp): the Terelfactor is replaced by the crossover
  (Te / 5.07783e9) / (1 + Te / 5.07783e9);
* where the clamp applies, the radiation channels equal the bridge-off
  ones of the same state with its radiation energy raised so that
  Trad = TradBB^(4/3) / Te^(1/3), plus the added term at that Trad;
* thermodynamics and scattering are untouched.

The states are those of the reference's rad_thermo golden records, with the
radiation energy scaled over six decades so that both sides of the clamp
occur.
"""
import numpy as np

import kzo
from goldens import geoms, read_kgld
from koral_zig_amd import abi_types as T
from test_oracle_floors_guard import _bsq

GAM = 5.0 / 3.0
TC_N = 5.07783e9


def _op(**kw):
    op = T.opac_params_puffy()
    for k, v in kw.items():
        setattr(op, k, v)
    return op


def _states():
    inp, _ = read_kgld("rad/rad_thermo.kgld", 36, 13)
    pp, g = inp[:, 23:36], geoms(inp)
    out_p, out_g = [], []
    for s in 10.0 ** np.linspace(-3.0, 3.0, 7):
        q = pp.copy()
        q[:, 9] *= s
        out_p.append(q)
        out_g.append(g)
    return np.concatenate(out_p), np.concatenate(out_g)


def _nr(op, rho, ne, bsq, t):
    """2.35869e-21 (n_e/rho) B^2 / T^3 in cgs, as an opacity in code units."""
    rhocgs = op.rhogu2cgs * rho
    nethcgs = ne * op.numdensgu2cgs
    bsqcgs = op.fourpi * (bsq * op.enden_gu2cgs_factor)
    return op.kappacgs2gu * (2.35869e-21 * (nethcgs / rhocgs) * bsqcgs / (t * t * t)) * rho


EPS = np.finfo(float).eps


def _close(got, want, on, off, ff, terel, rtol):
    """|got - want| within rtol of want plus the rounding that the two
    subtractions carry: off - ff loses eps (|off| + |ff|) before the division
    by terel (where synchrotron is far below free-free at low Te), on - ff
    loses eps (|on| + |ff|)."""
    bound = rtol * np.abs(want) + 8 * EPS * ((np.abs(off) + np.abs(ff)) / terel + np.abs(on) + np.abs(ff))
    bad = np.abs(got - want) > bound
    assert not bad.any(), (np.argwhere(bad)[:5].ravel(), got[bad][:3], want[bad][:3])


def test_bridge_channel_by_channel():
    pp, g = _states()
    off = kzo.radstate(pp, g, GAM, _op(), 1)
    on = kzo.radstate(pp, g, GAM, _op(synchrotron_bridge=1), 1)
    ff = kzo.radstate(pp, g, GAM, _op(synchrotron=0), 1)
    # thermodynamics, Ehat, TradBB and scattering unchanged
    assert np.array_equal(on[:, :6], off[:, :6])
    rho, te, ne, trad = pp[:, 0], off[:, 1], off[:, 2], off[:, 4]
    bsq = _bsq(pp, g)
    ok = np.isfinite(bsq) & (bsq > 0)
    theta = te * T.opac_params_puffy().k_over_mecsq
    terel = theta * theta / (1.0 + theta * theta)
    op = _op()
    # gas absorption (column 6)
    want = (off[:, 6] - ff[:, 6]) / terel + _nr(op, rho, ne, bsq, te)
    c = (on[:, 6], off[:, 6], ff[:, 6])
    _close((on[:, 6] - ff[:, 6])[ok], want[ok], *(x[ok] for x in c), terel[ok], 1e-10)
    # radiation channels where the clamp does not apply
    clamp = trad > te
    assert clamp[ok].sum() > 100 and (~clamp)[ok].sum() > 100
    nc = ok & ~clamp
    cross = (te / TC_N) / (1.0 + te / TC_N)
    for col, add, fac in ((7, _nr(op, rho, ne, bsq, trad), 1.0), (9, 0.0, cross), (11, 0.0, 1.0)):
        want = (off[:, col] - ff[:, col]) / terel * fac + add
        c = (on[:, col], off[:, col], ff[:, col])
        _close((on[:, col] - ff[:, col])[nc], want[nc], *(x[nc] for x in c), terel[nc] / np.broadcast_to(fac, terel.shape)[nc], 1e-10)


def test_bridge_clamp():
    """Trad > Te: the synchrotron Trad is raised to TradBB^(4/3) / Te^(1/3);
    the same state with E scaled by (Trad_lim / Trad)^4 has that Trad."""
    pp, g = _states()
    off = kzo.radstate(pp, g, GAM, _op(), 1)
    te, trad = off[:, 1], off[:, 4]
    bsq = _bsq(pp, g)
    sel = np.isfinite(bsq) & (bsq > 0) & (trad > te) & (trad < 1e3 * te)
    assert sel.sum() > 100
    pp, g, te, trad, bsq = pp[sel], g[sel], te[sel], trad[sel], bsq[sel]
    tlim = trad ** (4.0 / 3.0) / te ** (1.0 / 3.0)
    up = pp.copy()
    up[:, 9] *= (tlim / trad) ** 4
    off_up = kzo.radstate(up, g, GAM, _op(), 1)
    ff_up = kzo.radstate(up, g, GAM, _op(synchrotron=0), 1)
    assert np.allclose(off_up[:, 4], tlim, rtol=1e-13)
    assert np.array_equal(off_up[:, 1], te)
    on = kzo.radstate(pp, g, GAM, _op(synchrotron_bridge=1), 1)
    ff = kzo.radstate(pp, g, GAM, _op(synchrotron=0), 1)
    theta = te * T.opac_params_puffy().k_over_mecsq
    terel = theta * theta / (1.0 + theta * theta)
    op = _op()
    ne = off[sel, 2]
    # the radiation absorption at the clamped Trad, plus the added term there,
    # and the Rosseland mean at the clamped Trad (the scaled state's Trad
    # matches the limit to rounding, so the tolerance is loose)
    for col, add in ((7, _nr(op, pp[:, 0], ne, bsq, tlim)), (11, 0.0)):
        want = (off_up[:, col] - ff_up[:, col]) / terel + add
        _close(on[:, col] - ff[:, col], want, on[:, col], off_up[:, col], ff_up[:, col], terel, 1e-9)
