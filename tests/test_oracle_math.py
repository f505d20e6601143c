"""The oracle's FP32 transcendental functions against float64 numpy.

These pin the series evaluations in oracle/oracle_math.h: a wrong
coefficient, a dropped term or a bad quadrant would show up as errors far
above the few-ulp bounds asserted here.
"""
import numpy as np
import pytest


def ulp_err(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    spacing = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    spacing = np.maximum(spacing, np.float64(np.finfo(np.float32).tiny))
    return np.abs(got - want) / spacing


def vec(fn, xs):
    return np.array([fn(float(x)) for x in xs], np.float32)


def test_acos_accuracy(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-1, 1, 20000), np.linspace(-1, 1, 2001),
                         [0.0, -0.0, 0.5, -0.5, 1.0, -1.0, 0.49999997, 0.50000006]]).astype(np.float32)
    got = vec(L.oracle_acos, xs)
    want = np.arccos(xs.astype(np.float64))
    # near |x| = 1 acos has an infinite derivative: compare absolute error there
    err = np.where(np.abs(xs) > 0.999, np.abs(got - want) / 2.4e-7, ulp_err(got, want))
    assert err.max() <= 4.0, float(err.max())
    assert L.oracle_acos(1.0) == 0.0
    assert abs(L.oracle_acos(-1.0) - np.float32(np.pi)) == 0.0
    assert np.isnan(L.oracle_acos(1.0000001))
    assert np.isnan(L.oracle_acos(float("nan")))


@pytest.mark.parametrize("name,ref", [("oracle_sin", np.sin), ("oracle_cos", np.cos), ("oracle_tan", np.tan)])
def test_trig_accuracy(oracle, name, ref):
    fn = getattr(oracle.lib(), name)
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(-4, 4, 20000), np.linspace(0, np.pi, 1001),
                         [0.0, np.pi / 4, np.pi / 2, np.float32(np.pi / 2)]]).astype(np.float32)
    if name == "oracle_tan":   # keep away from the poles where tan is ill-conditioned
        xs = xs[np.abs(np.cos(xs.astype(np.float64))) > 1e-3]
    got = vec(fn, xs)
    want = ref(xs.astype(np.float64))
    err = np.abs(got - want) / np.maximum(np.abs(want), 1e-6) / 6e-8
    if name != "oracle_tan":
        # absolute error near the zeros of sin/cos
        err = np.minimum(err, np.abs(got - want) / 1.2e-7)
    assert np.quantile(err, 0.999) <= 4.0 and err.max() <= 64, (float(np.quantile(err, 0.999)), float(err.max()))
    assert fn(0.0) == (1.0 if name == "oracle_cos" else 0.0)


def test_pow_accuracy(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(3)
    x = rng.uniform(0.0, 1.0, 20000).astype(np.float32)
    y = rng.choice(np.array([1.0, 2.0, 10.0, 63.75, 102.0, 153.0], np.float32), 20000)
    got = np.array([L.oracle_pow(float(a), float(b)) for a, b in zip(x, y)], np.float64)
    want = np.power(x.astype(np.float64), y.astype(np.float64))
    rel = np.abs(got - want) / np.maximum(want, 1e-30)
    # error grows with |y * log2 x|; the shading tolerance is 1/255
    bound = 2e-7 * (1.0 + np.abs(y * np.log2(np.maximum(x, 1e-30)))) + 1e-6
    big = want > 1e-30
    assert np.all(rel[big] <= bound[big] * 8)
    assert L.oracle_pow(0.0, 102.0) == 0.0
    assert L.oracle_pow(1.0, 153.0) == 1.0
    assert L.oracle_pow(0.5, 0.0) == 1.0
    assert L.oracle_pow(0.5, 1.0) == 0.5
    assert L.oracle_pow(0.25, 2.0) == 0.0625


def test_saturating_conversion(oracle):
    f = oracle.lib().oracle_f2u_sat
    assert f(float("nan")) == 0
    assert f(-5.0) == 0
    assert f(-0.5) == 0
    assert f(3.99) == 3
    assert f(4294967296.0) == 0xFFFFFFFF
    assert f(1e30) == 0xFFFFFFFF
    assert f(float("inf")) == 0xFFFFFFFF
