"""Synthetic inputs of the measured workloads that are not a PUFFY run.

implicit_batch_inputs: the independent-cell implicit-solve batch — 2^22
cells at full size — with states spread over the regimes the solver meets:
rho log-uniform over [1e-25, 1e-3] code units, T_gas log-uniform over
[1e4, 1e10] K (u from thermo.zig:173-176 uFromTrho), gas VELR velocity of
Lorentz factor U[1, 2.5] and B of plasma beta log-U[1e-2, 1e2] along random
directions (the reference's implicit harness, oracle/harness_implicit.c
controlled_state), fluid-frame radiation energy E_hat / u_gas log-U[1e-6,
1e6] with |F|/E U[0, 0.99] realised as the radiation-frame speed of the M1
closure f = 4 gamma^2 beta / (4 gamma^2 - 1), optical depth log-U[1e-4,
1e4] realised as dt = tau / chi (chi = calcChiSlim, the solver's stiffness
chi*dt), at MKS2 a=0 points with r U[2.5, 50] (x1 = ln(r - r0)) and
x2 U[0.001, 0.999].  Deterministic for a seed.
"""
import numpy as np

from . import abi_types as T
from . import host

COORDS_MKS2 = 3


def _random_vec(rng, g, norm2):
    """Spatial vector of squared norm g_ij a^i a^j = norm2 along a random
    direction."""
    n = rng.uniform(-1.0, 1.0, (len(norm2), 3))
    q = np.einsum("ni,nij,nj->n", n, g[:, 1:4, 1:4], n)
    return n * np.sqrt(norm2 / q)[:, None]


def _velr_gamma(rng, g, gam):
    """VELR velocity (u~^i) of ZAMO Lorentz factor gam: |u~|^2 = gam^2 - 1."""
    return _random_vec(rng, g, gam * gam - 1.0)


def _vel4(v, g, G):
    """relele velrToVel4: u^t = gamma / alpha, u^i = u~^i - gamma alpha G^{ti}."""
    alpha = np.sqrt(-1.0 / G[:, 0, 0])
    gamma = np.sqrt(1.0 + np.einsum("ni,nij,nj->n", v, g[:, 1:4, 1:4], v))
    u = np.zeros((len(v), 4))
    u[:, 0] = gamma / alpha
    u[:, 1:4] = v - (gamma * alpha)[:, None] * G[:, 0, 1:4]
    return u


def _gamma_m1(f):
    """Radiation-frame Lorentz factor of flux ratio f under the M1 closure."""
    f2 = f * f
    b = 16.0 - 8.0 * f2
    x = (b + np.sqrt(b * b + 64.0 * f2 * (1.0 - f2))) / (32.0 * (1.0 - f2))
    return np.sqrt(x)


def implicit_batch_inputs(n, seed, phys, chi_fn):
    """(pp[n, 13], geom[n, GEOM_DOUBLES], dt[n]); chi_fn(pp, geom) -> chi."""
    rng = np.random.default_rng(seed)
    mp = host.MP_PUFFY
    r = rng.uniform(2.5, 50.0, n)
    x4 = np.zeros((n, 4))
    x4[:, 1] = np.log(r - mp["mksr0"])
    x4[:, 2] = rng.uniform(0.001, 0.999, n)
    geom = host.geom_batch(COORDS_MKS2, x4, mp)
    gg = geom[:, 0:20].reshape(n, 4, 5)[:, :, :4]
    GG = geom[:, 20:40].reshape(n, 4, 5)[:, :, :4]
    gam = phys.gam
    op = phys.opac
    rho = 10.0 ** rng.uniform(-25.0, -3.0, n)
    tgas = 10.0 ** rng.uniform(4.0, 10.0, n)
    uint = tgas * rho * op.kb_over_mugas_mp / (gam - 1.0)
    pp = np.zeros((n, T.NV))
    pp[:, 0] = rho
    pp[:, 1] = uint
    vg = _velr_gamma(rng, gg, rng.uniform(1.0, 2.5, n))
    pp[:, 2:5] = vg
    pgas = (gam - 1.0) * uint
    ncoef = 1.0 / (gam - 1.0)
    pp[:, 5] = rho * np.log(pgas ** ncoef / rho ** (ncoef + 1.0))
    beta = 10.0 ** rng.uniform(-2.0, 2.0, n)
    bsq = 2.0 * (gam - 1.0) * uint / beta
    pp[:, 6:9] = _random_vec(rng, gg, bsq)
    f = rng.uniform(0.0, 0.99, n)
    vr = _velr_gamma(rng, gg, _gamma_m1(f))
    pp[:, 10:13] = vr
    ug, ur = _vel4(vg, gg, GG), _vel4(vr, gg, GG)
    grel = -np.einsum("ni,nij,nj->n", ur, gg, ug)
    ratio = 10.0 ** rng.uniform(-6.0, 6.0, n)
    pp[:, 9] = ratio * uint * 3.0 / (4.0 * grel * grel - 1.0)
    tau = 10.0 ** rng.uniform(-4.0, 4.0, n)
    dt = tau / chi_fn(pp, geom)
    return pp, geom, dt
