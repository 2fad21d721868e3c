"""The independent-cell implicit workload generator holds the distribution
it states: u from T_gas through uFromTrho, the radiation-frame speed of the
M1 closure at the drawn |F|/E, the fluid-frame E_hat / u ratio and the
optical depth chi*dt, each recovered from the generated state with the
oracle's own routines (not the generator's formulas)."""
import numpy as np

import kzo
from koral_zig_amd import abi_types as T
from koral_zig_amd import workloads


def test_m1_gamma_closed_form():
    f = np.array([0.0, 0.3, 0.7, 0.99])
    g = workloads._gamma_m1(f)
    beta = np.sqrt(1.0 - 1.0 / (g * g))
    assert np.allclose(4 * g * g * beta / (4 * g * g - 1), f, rtol=1e-12, atol=1e-15)
    assert g[0] == 1.0


def test_generator_recovers_targets():
    phys = T.physics_puffy()
    n = 3000
    chi = lambda p, g: kzo.opac_entry(p, g, phys.gam, phys.opac)[:, 3]  # noqa: E731
    pp, geom, dt = workloads.implicit_batch_inputs(n, 5, phys, chi)
    pp2, geom2, dt2 = workloads.implicit_batch_inputs(n, 5, phys, chi)
    assert np.array_equal(pp, pp2) and np.array_equal(dt, dt2)
    rng = np.random.default_rng(5)  # replay the draws in the generator's order
    r = rng.uniform(2.5, 50.0, n)
    rng.uniform(0.001, 0.999, n)
    rho = 10.0 ** rng.uniform(-25.0, -3.0, n)
    tgas = 10.0 ** rng.uniform(4.0, 10.0, n)
    assert np.array_equal(pp[:, 0], rho)
    rs = kzo.radstate(pp, geom, phys.gam, phys.opac, 0)
    assert np.allclose(rs[:, 0], tgas, rtol=1e-12)          # T_gas of the oracle's thermo
    assert np.allclose(np.exp(geom[:, 44]) + 0.1, r, rtol=1e-13)
    tau = chi(pp, geom) * dt
    assert tau.min() >= 1e-4 * (1 - 1e-12) and tau.max() <= 1e4 * (1 + 1e-12)
    ratio = rs[:, 3] / pp[:, 1]                              # oracle E_hat / u_gas
    assert ratio.min() >= 1e-6 * (1 - 1e-9) and ratio.max() <= 1e6 * (1 + 1e-9)
    assert np.log10(ratio).std() > 3.0
