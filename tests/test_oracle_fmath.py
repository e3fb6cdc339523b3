"""include/mtx_fmath.h (shared by the oracle and the HIP kernels) against
numpy's f64 libm rounded to f32: at most 1 ulp anywhere, and correctly
rounded on almost every input; C99 special cases as the C library has them."""
import numpy as np
import pytest


def ulps(a, b):
    a = a.astype(np.float32).view(np.int32).astype(np.int64)
    b = b.astype(np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


@pytest.fixture(scope="module")
def inputs():
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-7, 7, 400000), rng.uniform(-1e4, 1e4, 50000),
                        rng.uniform(-1e-3, 1e-3, 20000)]).astype(np.float32)
    return rng, x


def test_sin_cos(orc, inputs):
    rng, x = inputs
    for name, ref in (("sin", np.sin), ("cos", np.cos)):
        got = orc.math_eval(name, x)
        want = ref(x.astype(np.float64)).astype(np.float32)
        u = ulps(got, want)
        assert u.max() <= 1, name
        assert (u == 0).mean() > 0.9999, name


def test_pow(orc, inputs):
    rng, _ = inputs
    b = np.concatenate([rng.uniform(0, 1, 200000), rng.uniform(0, 50, 100000)]).astype(np.float32)
    e = np.concatenate([rng.uniform(-6, 6, 200000), rng.uniform(-3, 3, 100000)]).astype(np.float32)
    got = orc.math_eval("pow", b, e)
    with np.errstate(all="ignore"):
        want = np.power(b.astype(np.float64), e.astype(np.float64)).astype(np.float32)
    fin = np.isfinite(want) & (want != 0)
    u = ulps(got[fin], want[fin])
    assert u.max() <= 1 and (u == 0).mean() > 0.9999


def test_atan2_acos(orc, inputs):
    rng, x = inputs
    y = rng.uniform(-7, 7, x.size).astype(np.float32)
    got = orc.math_eval("atan2", y, x)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64)).astype(np.float32)
    assert ulps(got, want).max() <= 1
    c = np.concatenate([rng.uniform(-1, 1, 300000), [-1, -0.0, 0.0, 1]]).astype(np.float32)
    got = orc.math_eval("acos", c)
    want = np.arccos(c.astype(np.float64)).astype(np.float32)
    assert ulps(got, want).max() <= 1


def test_exp(orc, inputs):
    rng, _ = inputs
    x = np.concatenate([rng.uniform(-10, 10, 300000), rng.uniform(-103, 88, 100000),
                        rng.uniform(-1e-4, 1e-4, 20000)]).astype(np.float32)
    got = orc.math_eval("exp", x)
    want = np.exp(x.astype(np.float64)).astype(np.float32)
    u = ulps(got, want)
    assert u.max() <= 1 and (u == 0).mean() > 0.9999
    sp = orc.math_eval("exp", np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 89.0, -104.0], np.float32))
    assert sp[0] == 1 and sp[1] == 1 and sp[2] == np.inf and sp[3] == 0 and np.isnan(sp[4])
    assert sp[5] == np.inf and sp[6] == 0


def test_sincos_equals_sin_and_cos(orc, inputs):
    _, x = inputs
    x = np.concatenate([x, np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e6, -3e5], np.float32)])
    for part, ref in (("sincos.s", "sin"), ("sincos.c", "cos")):
        assert np.array_equal(orc.math_eval(part, x).view(np.uint32), orc.math_eval(ref, x).view(np.uint32)), part


def test_special_cases(orc):
    inf, nan = np.float32(np.inf), np.float32(np.nan)
    z, nz = np.float32(0.0), np.float32(-0.0)
    cases = [  # (base, exponent, expected) per C99 F.9.4.4
        (z, 5.0, 0.0), (nz, 3.0, -0.0), (z, -1.0, inf), (2.0, 0.0, 1.0), (1.0, nan, 1.0), (-2.0, 3.0, -8.0),
        (-2.0, 0.5, nan), (inf, -1.0, 0.0), (0.5, inf, 0.0), (2.0, inf, inf), (0.5, -inf, inf)]
    b = np.array([c[0] for c in cases], np.float32)
    e = np.array([c[1] for c in cases], np.float32)
    got = orc.math_eval("pow", b, e)
    want = np.array([c[2] for c in cases], np.float32)
    assert np.array_equal(got, want, equal_nan=True)
    assert np.signbit(got[1])
    y = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, inf, 1.0], np.float32)
    x = np.array([1.0, 1.0, -1.0, -1.0, 0.0, 0.0, inf, -inf], np.float32)
    got = orc.math_eval("atan2", y, x)
    assert np.array_equal(got, np.arctan2(y, x)) and np.array_equal(np.signbit(got), np.signbit(np.arctan2(y, x)))
    s = orc.math_eval("sin", np.array([nz, z, inf, nan], np.float32))
    assert np.signbit(s[0]) and s[1] == 0 and np.isnan(s[2]) and np.isnan(s[3])
