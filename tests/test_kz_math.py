"""Accuracy of the portable transcendentals (koral-zig_amd/csrc/kz_math.h,
through the oracle's build of the same source) against the host's C library
(Python's math module and numpy's cbrt call glibc), in units in the last
place of the reference value, over arguments spread across the whole
double range the step can produce; plus the special values the routines
must return exactly (perfect cubes, exp(0), log(1), the IEEE edge cases).
The GPU build is checked bit for bit against these same routines in
tests/test_gpu_cellops.py."""
import math

import numpy as np

import kzo


def ulps(got, ref):
    got, ref = np.asarray(got, float), np.asarray(ref, float)
    sp = np.spacing(np.abs(ref))
    return np.abs(got - ref) / np.where(sp > 0, sp, 5e-324)


def _rng():
    return np.random.default_rng(20261202)


def test_exp():
    x = np.concatenate([_rng().uniform(-745.0, 709.7, 20000), _rng().uniform(-1, 1, 5000) * 1e-8])
    ref = np.array([math.exp(v) for v in x])
    assert ulps(kzo.math_fn("exp", x), ref).max() <= 1.0


def test_log():
    x = np.concatenate([10.0 ** _rng().uniform(-320, 308, 20000), 1.0 + _rng().uniform(-1e-3, 1e-3, 5000),
                        [5e-324, 2.2250738585072014e-308, 1.7976931348623157e308]])
    ref = np.array([math.log(v) for v in x])
    assert ulps(kzo.math_fn("log", x), ref).max() <= 1.0


def test_pow():
    r = _rng()
    x = 10.0 ** r.uniform(-30, 30, 20000)
    y = r.uniform(-10, 10, 20000)
    ref = np.array([math.pow(a, b) for a, b in zip(x, y)])
    assert ulps(kzo.math_fn("pow", x, y), ref).max() <= 2.0
    for e in (0.2, 0.25, 0.463, 0.69, 0.86, 1.69, -0.463, -3.5):  # the exponents the opacities use
        z = 10.0 ** r.uniform(-40, 40, 5000)
        ref = np.array([math.pow(a, e) for a in z])
        assert ulps(kzo.math_fn("pow", z, e), ref).max() <= 1.0, e


def test_cbrt():
    x = np.concatenate([10.0 ** _rng().uniform(-320, 308, 20000), -(10.0 ** _rng().uniform(-300, 300, 2000)),
                        [5e-324, 2.2250738585072014e-308, 1.7976931348623157e308]])
    assert ulps(kzo.math_fn("cbrt", x), np.cbrt(x)).max() <= 1.0
    k = np.arange(-200, 201, dtype=float)
    assert np.array_equal(kzo.math_fn("cbrt", k ** 3), k)  # perfect cubes are exact
    p = 2.0 ** np.arange(-1020, 1021, 3)
    assert np.array_equal(kzo.math_fn("cbrt", p), 2.0 ** (np.arange(-1020, 1021, 3) / 3))


def test_log1p():
    x = np.concatenate([_rng().uniform(-0.999, 10, 10000), _rng().uniform(-1e-6, 1e-6, 5000)])
    ref = np.array([math.log1p(v) for v in x])
    assert ulps(kzo.math_fn("log1p", x), ref).max() <= 2.0


def test_special_values():
    inf, nan = math.inf, math.nan
    assert kzo.math_fn("exp", np.array([0.0]))[0] == 1.0
    assert kzo.math_fn("log", np.array([1.0]))[0] == 0.0
    e = kzo.math_fn("exp", np.array([710.0, -746.0, inf, -inf, nan]))
    assert e[0] == inf and e[1] == 0.0 and e[2] == inf and e[3] == 0.0 and np.isnan(e[4])
    lg = kzo.math_fn("log", np.array([0.0, -1.0, inf, nan]))
    assert lg[0] == -inf and np.isnan(lg[1]) and lg[2] == inf and np.isnan(lg[3])
    c = kzo.math_fn("cbrt", np.array([0.0, -0.0, inf, -inf, nan]))
    assert c[0] == 0.0 and c[2] == inf and c[3] == -inf and np.isnan(c[4])
    assert np.signbit(c[1])
    pw = kzo.math_fn("pow", np.array([2.0, -8.0, 0.0, 1.0, -2.0]), np.array([10.0, 1.0 / 3.0, -1.0, nan, 3.0]))
    assert pw[0] == 1024.0 and np.isnan(pw[1]) and pw[2] == inf and pw[3] == 1.0 and pw[4] == -8.0
