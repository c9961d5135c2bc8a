"""GPU: the h solve's power x**0.333333333 (all_sph_gradh.F90:253,317 with
ONETHIRD = 0.333333333, macros.h:18) is correctly rounded on the device.

The device evaluates it from a Newton-refined cube root and an atanh-series
logarithm (seren_amd/csrc/walk.h).  Each result must equal the double
nearest to the exact power x**y, y the double nearest to 0.333333333,
computed here with 60-digit decimal arithmetic,
over a log-uniform sample of [2^-1000, 2^1000], the range of m/rho in the
clouds, and edge cases (powers of two and eight, values next to 1, the
limits of the fast range and values outside it).
"""
import decimal

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _exact(xs):
    # the exponent is the double nearest to 0.333333333, as the reference's
    # FP64 ONETHIRD and libm pow take it (it differs from 0.333333333 by
    # ~1e-17, which moves x**y by up to ~1e-14 relative)
    ctx = decimal.Context(prec=60)
    y = decimal.Decimal(0.333333333)
    return np.array([float(ctx.power(decimal.Decimal(float(x)), y)) for x in xs])


def _sample():
    rng = np.random.default_rng(77)
    wide = np.exp2(rng.uniform(-1000, 1000, 12000))
    typical = np.exp(rng.uniform(np.log(1e-12), np.log(1e4), 12000))
    edges = [1.0, 2.0, 8.0, 0.125, 1.0 - 2**-53, 1.0 + 2**-52, 2.0**-1000, 2.0**1000,
             2.0**-1001, 2.0**1001, 3.0, 27.0, 1e-300, 1e300]
    edges += [np.nextafter(2.0**k, 0.0) for k in range(-30, 31)]
    edges += [2.0**k for k in range(-30, 31)]
    return np.concatenate([wide, typical, np.array(edges)])


def test_device_pow_third_correctly_rounded():
    from seren_amd.engine import ForceEngine
    print("case pow_third 24k + edges", flush=True)
    x = _sample()
    eng = ForceEngine()
    try:
        y = eng.pow_third(x)
    finally:
        eng.close()
    want = _exact(x)
    bad = np.flatnonzero(y != want)
    libm = int(np.sum(x ** 0.333333333 != want))
    print(f"  device misrounded: {len(bad)}; host libm pow misrounded: {libm} of {len(x)}",
          flush=True)
    assert len(bad) == 0, (x[bad][:5], y[bad][:5], want[bad][:5])
