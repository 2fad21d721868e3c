"""kzm_pow_from_log (kz_math.h) returns kzm_pow's bits for every finite
positive base and finite exponent, so the device opacities may take two
powers of one base from one logarithm: checked bit for bit over bases from
subnormal to huge (1.0 included) and the exponents the opacity chain uses,
plus random ones."""
import numpy as np

import kzo


def test_pow_from_log_is_pow():
    rng = np.random.default_rng(8)
    x = np.concatenate([10.0 ** rng.uniform(-310, 308, 20000), [1.0, 5e-324, 2.2250738585072014e-308,
                                                                   1.7976931348623157e308, 0.999999999, 1.0000001]])
    for y in (-0.463, 0.69, 1.69, 0.463, 0.2, 0.86, 0.0, -1.0, 3.0):
        a = kzo.math_fn("pow", x, np.full_like(x, y))
        b = kzo.math_fn("pow_from_log", x, np.full_like(x, y))
        assert a.tobytes() == b.tobytes(), y
    y = rng.uniform(-40, 40, x.size)
    assert kzo.math_fn("pow", x, y).tobytes() == kzo.math_fn("pow_from_log", x, y).tobytes()
