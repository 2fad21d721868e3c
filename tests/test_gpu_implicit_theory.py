"""The implicit solver's theory-gate states through the GPU ladder.

tests/test_oracle_implicit_theory.py holds the oracle to the reference's M9
gates (koral/tests/implicit_tests.zig:89-431).  Here the device ladder
(kz_solve_implicit_lab_batch) runs the same states and must equal the
oracle bit for bit: the 3000-state fuzz sweep over Minkowski and
Kerr-Schild geometries, the LTE fixed points and the stiff L-stability
step.  The device result then meets those gates by identity; the fuzz
gates (finite output where solved, untouched primitives where failed) are
also checked on the device output directly.
"""
import numpy as np
import pytest

import kzo
from koral_zig_amd import abi_types as T
from koral_zig_amd import hip
from test_oracle_implicit_theory import GAM, MINK, _brems, _geom, fuzz_states, lte_e, pp_from_temps

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()


def test_fuzz_states_bitwise():
    ph = T.physics_puffy()
    pp0, g, dt = fuzz_states(ph.opac)
    rc, pc, uc = kzo.solve_implicit(pp0, g, dt, GAM, ph.rad, ph.opac, ph.implicit)
    uz, pz, rz = hip.solve_implicit(np.zeros_like(pp0), pp0, g, dt, GAM, ph.rad, ph.opac, ph.implicit)
    for k in ("ok", "rung", "iters"):
        assert _same(rc[k], rz[k]), k
    ok = rz["ok"] == 1
    assert _same(pc, pz)
    assert _same(uc[ok], uz[ok])
    assert np.isfinite(pz[ok]).all() and np.isfinite(uz[ok]).all()
    assert np.array_equal(pz[~ok], pp0[~ok])


def test_lte_and_stiff_states_bitwise():
    ph = T.physics_puffy()
    op = _brems(ph.opac)
    g = _geom(MINK, [0, 0, 0, 0])
    rows, dts = [], []
    for tgas in (1.0e6, 1.0e8, 1.0e10):
        rows.append(pp_from_temps(op, 1.0e-8, tgas, [0, 0, 0], lte_e(op, tgas), [0, 0, 0], [0, 0, 0]))
        dts.append(1.0)
    for frac, dt in ((0.2, 1.0e3), (0.5, 1.0e-3), (0.2, 1.0e9)):
        rows.append(pp_from_temps(op, 1.0e-8, 1.0e8, [0, 0, 0], lte_e(op, frac * 1.0e8), [0, 0, 0], [0, 0, 0]))
        dts.append(dt)
    pp0 = np.array(rows)
    gg = np.repeat(g, len(pp0), axis=0)
    dt = np.array(dts)
    rc, pc, uc = kzo.solve_implicit(pp0, gg, dt, GAM, ph.rad, op, ph.implicit)
    uz, pz, rz = hip.solve_implicit(np.zeros_like(pp0), pp0, gg, dt, GAM, ph.rad, op, ph.implicit)
    assert (rz["ok"] == 1).all()
    assert _same(rc["iters"], rz["iters"]) and _same(pc, pz) and _same(uc, uz)
