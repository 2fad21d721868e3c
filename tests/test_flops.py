"""The operation-count build of the oracle and the FP32-Jacobian option.

oracle/flops compiles the oracle's implicit solve with a counting double;
its count is only meaningful if that build computes exactly what the oracle
computes, so the first test requires bit-identical results on a sample of
the independent-cell workload.  The FP32 LU of the mixed-precision Newton
is pinned to numpy's inverse, and the option is checked to change the
solve (iteration counts) while keeping its outcomes.
"""
import importlib.util
from pathlib import Path

import numpy as np

import kzo
from koral_zig_amd import abi_types as T
from koral_zig_amd import workloads

ROOT = Path(__file__).resolve().parent.parent


def _count_mod():
    spec = importlib.util.spec_from_file_location("kzo_flops_count", ROOT / "oracle" / "flops" / "count.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _cells(n, seed=3):
    ph = T.physics_puffy()
    pp, geom, dt = workloads.implicit_batch_inputs(n, seed, ph, lambda p, g: kzo.opac_entry(p, g, ph.gam, ph.opac)[:, 3])
    return ph, pp, geom, dt


def test_counting_build_reproduces_oracle():
    cm = _count_mod()
    ph, pp, geom, dt = _cells(400)
    res, po, uo, fl = cm.count(pp, geom, dt, ph)
    r2, p2, u2 = kzo.solve_implicit(pp, geom, dt, ph.gam, ph.rad, ph.opac, ph.implicit)
    assert res.tobytes() == r2.tobytes()
    assert po.tobytes() == p2.tobytes() and uo.tobytes() == u2.tobytes()
    ok = res["ok"] == 1
    first = ok & (res["rung"] == 0)
    assert (fl > 0).all()
    # more Newton iterations cost more operations; failing ladders the most
    it = res["iters"][first]
    a = np.polyfit(it, fl[first].astype(float), 1)
    assert a[0] > 5000
    if (~ok).any():
        assert fl[~ok].mean() > fl[first].mean()


def test_invert4f_matches_numpy():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(200, 4, 4)) + 4.0 * np.eye(4)
    a[0, 0, 0] = 0.0  # the first pivot must be taken from another row
    inv = kzo.invert4f(a.astype(np.float32)).reshape(-1, 4, 4).astype(np.float64)
    ref = np.linalg.inv(a.astype(np.float32).astype(np.float64))
    assert np.allclose(inv, ref, rtol=1e-4, atol=1e-5)
    assert kzo.invert4f(a.astype(np.float32)).dtype == np.float32


def test_fp32_jacobian_changes_path_not_outcome():
    ph, pp, geom, dt = _cells(1500, seed=4)
    ip = T.implicit_params_puffy()
    ip.jacobian_fp32 = 1
    r64, _, _ = kzo.solve_implicit(pp, geom, dt, ph.gam, ph.rad, ph.opac, ph.implicit)
    r32, p32, _ = kzo.solve_implicit(pp, geom, dt, ph.gam, ph.rad, ph.opac, ip)
    assert (r64["ok"] == r32["ok"]).mean() > 0.99
    assert not np.array_equal(r64["iters"], r32["iters"])
    assert np.isfinite(p32[r32["ok"] == 1]).all()
