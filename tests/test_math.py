"""Accuracy of the shared FP64 elementary functions (include/wq3d_math.h)
against the host C library, in units in the last place."""
import numpy as np
import pytest


def ulp_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    spacing = np.spacing(np.maximum(np.abs(b), np.finfo(np.float64).tiny))
    return np.abs(a - b) / spacing


RNG = np.random.default_rng(7)

CASES = [
    ("sin", 0, lambda: RNG.uniform(-10.0, 10.0, 200000), np.sin, 1.0),
    ("sin_small", 0, lambda: RNG.uniform(-1e-3, 1e-3, 50000), np.sin, 1.0),
    ("cos", 1, lambda: RNG.uniform(-10.0, 10.0, 200000), np.cos, 1.0),
    ("cos_colat", 1, lambda: RNG.uniform(0.0, np.pi, 100000), np.cos, 1.0),
    ("exp", 2, lambda: RNG.uniform(-50.0, 50.0, 200000), np.exp, 1.0),
    ("exp_wide", 2, lambda: RNG.uniform(-700.0, 700.0, 100000), np.exp, 1.0),
    ("log", 3, lambda: np.exp(RNG.uniform(-600.0, 600.0, 200000)), np.log, 1.0),
    ("log_near1", 3, lambda: RNG.uniform(0.9, 1.1, 100000), np.log, 1.0),
    ("log10", 4, lambda: np.exp(RNG.uniform(-60.0, 60.0, 200000)), np.log10, 2.0),
    ("asin", 6, lambda: RNG.uniform(-1.0, 1.0, 200000), np.arcsin, 3.0),
    ("pow10", 7, lambda: RNG.uniform(-30.0, 30.0, 200000), lambda x: 10.0 ** x, 2.0),
]


@pytest.mark.parametrize("name,fn,gen,ref,max_ulp", CASES, ids=[c[0] for c in CASES])
def test_unary(oracle_mod, name, fn, gen, ref, max_ulp):
    x = gen()
    got = oracle_mod.math_fn(fn, x)
    err = ulp_err(got, ref(x))
    assert err.max() <= max_ulp, f"{name}: max {err.max():.2f} ulp at x={x[np.argmax(err)]!r}"


def test_atan2(oracle_mod):
    y = RNG.normal(size=200000) * np.exp(RNG.uniform(-5, 5, 200000))
    x = RNG.normal(size=200000) * np.exp(RNG.uniform(-5, 5, 200000))
    err = ulp_err(oracle_mod.math_fn(5, y, x), np.arctan2(y, x))
    assert err.max() <= 2.0, err.max()


def test_special_values(oracle_mod):
    assert oracle_mod.math_fn(2, [0.0])[0] == 1.0
    assert oracle_mod.math_fn(3, [1.0])[0] == 0.0
    assert oracle_mod.math_fn(0, [0.0])[0] == 0.0
    assert oracle_mod.math_fn(1, [0.0])[0] == 1.0
    assert np.isnan(oracle_mod.math_fn(3, [-1.0])[0])
    a = oracle_mod.math_fn(5, [0.0, -0.0, 0.0, 1.0, -1.0], [1.0, 1.0, -1.0, 0.0, 0.0])
    assert a[0] == 0.0 and a[1] == 0.0 and np.signbit(a[1])
    assert a[2] == np.pi and a[3] == np.pi / 2 and a[4] == -np.pi / 2
    assert oracle_mod.math_fn(6, [1.0])[0] == np.pi / 2
