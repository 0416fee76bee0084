"""Pins the shared float32 transcendentals (fhtgs_math.h, used by the oracle
and the kernels) against libm in float64, and the SH colour basis of the
oracle against textbook real spherical harmonics."""
import math

import numpy as np
import pytest


def _ulp_err(got: np.ndarray, want: np.ndarray) -> np.ndarray:
    want32 = want.astype(np.float32)
    ulp = np.spacing(np.abs(want32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / ulp


def _apply(lib, name, xs):
    f = getattr(lib, name)
    return np.array([f(float(x)) for x in xs], dtype=np.float32)


def test_expf_within_2ulp(oracle_lib):
    lib = oracle_lib.load()
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-87.0, 88.0, 20000), rng.uniform(-6.0, 0.0, 20000),
                         np.linspace(-5.541, 0.0, 2001)]).astype(np.float32)
    got = _apply(lib, "orc_expf", xs)
    err = _ulp_err(got, np.exp(xs.astype(np.float64)))
    assert err.max() <= 2.0, f"expf max error {err.max()} ulp at x={xs[err.argmax()]}"
    assert lib.orc_expf(0.0) == 1.0
    assert lib.orc_expf(-200.0) == 0.0
    assert math.isinf(lib.orc_expf(100.0))
    assert math.isnan(lib.orc_expf(float("nan")))


def test_logf_within_2ulp(oracle_lib):
    lib = oracle_lib.load()
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.exp(rng.uniform(-100.0, 80.0, 20000)), rng.uniform(0.004, 255.0, 20000),
                         rng.uniform(0.5, 2.0, 5000), np.array([1e-40, 3e-39, 1e-45])]).astype(np.float32)
    xs = xs[xs > 0]
    got = _apply(lib, "orc_logf", xs)
    want = np.log(xs.astype(np.float64))
    err = np.abs(got.astype(np.float64) - want) / np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    err = np.where(np.abs(want) < 1e-30, 0.0, err)
    assert err.max() <= 2.0, f"logf max error {err.max()} ulp at x={xs[err.argmax()]}"
    assert lib.orc_logf(1.0) == 0.0
    assert lib.orc_logf(0.0) == float("-inf")
    assert math.isnan(lib.orc_logf(-1.0))


def test_atan2_asin_tan_tanh_accuracy(oracle_lib):
    lib = oracle_lib.load()
    rng = np.random.default_rng(2)
    y = rng.normal(size=20000).astype(np.float32)
    x = rng.normal(size=20000).astype(np.float32)
    got = np.array([lib.orc_atan2f(float(a), float(b)) for a, b in zip(y, x)], dtype=np.float64)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.abs(got - want).max() <= 4e-7
    for yy, xx in ((0.0, 1.0), (1.0, 0.0), (0.0, -1.0), (-1.0, 0.0), (1.0, 1.0), (-1.0, -1.0)):
        assert abs(lib.orc_atan2f(yy, xx) - math.atan2(yy, xx)) <= 3e-7
    s = rng.uniform(-1.0, 1.0, 20000).astype(np.float32)
    got = np.array([lib.orc_asinf(float(v)) for v in s], dtype=np.float64)
    assert np.abs(got - np.arcsin(s.astype(np.float64))).max() <= 5e-7
    t = rng.uniform(-1.55, 1.55, 20000).astype(np.float32)
    got = np.array([lib.orc_tanf(float(v)) for v in t], dtype=np.float64)
    want = np.tan(t.astype(np.float64))
    assert (np.abs(got - want) / np.maximum(1.0, np.abs(want))).max() <= 1e-6
    h = np.concatenate([rng.uniform(-10, 10, 10000), rng.uniform(-0.07, 0.07, 5000)]).astype(np.float32)
    got = np.array([lib.orc_tanhf(float(v)) for v in h], dtype=np.float64)
    assert np.abs(got - np.tanh(h.astype(np.float64))).max() <= 3e-7


def _real_sh_textbook(d: np.ndarray) -> np.ndarray:
    """Real SH up to degree 3 with the Condon-Shortley phase, ordered
    m = -l..l per degree, from the Cartesian closed forms."""
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    pi = math.pi
    c = lambda v: math.sqrt(v)
    Y = [np.full_like(x, 0.5 * c(1 / pi))]
    a1 = c(3 / (4 * pi))
    Y += [-a1 * y, a1 * z, -a1 * x]
    Y += [0.5 * c(15 / pi) * x * y, -0.5 * c(15 / pi) * y * z, 0.25 * c(5 / pi) * (3 * z * z - 1),
          -0.5 * c(15 / pi) * x * z, 0.25 * c(15 / pi) * (x * x - y * y)]
    Y += [-0.25 * c(35 / (2 * pi)) * y * (3 * x * x - y * y), 0.5 * c(105 / pi) * x * y * z,
          -0.25 * c(21 / (2 * pi)) * y * (5 * z * z - 1), 0.25 * c(7 / pi) * z * (5 * z * z - 3),
          -0.25 * c(21 / (2 * pi)) * x * (5 * z * z - 1), 0.25 * c(105 / pi) * z * (x * x - y * y),
          -0.25 * c(35 / (2 * pi)) * x * (x * x - 3 * y * y)]
    return np.stack(Y, axis=1)


def _oracle_basis(lib, dirs: np.ndarray) -> np.ndarray:
    out = np.zeros((len(dirs), 16))
    rgb = np.zeros(3, dtype=np.float32)
    for j in range(16):
        sh0 = np.zeros(3, dtype=np.float32)
        shr = np.zeros((15, 3), dtype=np.float32)
        if j == 0:
            sh0[0] = 1.0
        else:
            shr[j - 1, 0] = 1.0
        for i, d in enumerate(dirs.astype(np.float32)):
            d = np.ascontiguousarray(d)
            lib.orc_sh_eval(sh0.ctypes.data, shr.ctypes.data, d.ctypes.data, 16, rgb.ctypes.data)
            out[i, j] = float(rgb[0]) - 0.5
    return out


def test_sh_basis_is_textbook_real_sh(oracle_lib):
    lib = oracle_lib.load()
    rng = np.random.default_rng(3)
    d = rng.normal(size=(300, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    got = _oracle_basis(lib, d)
    want = _real_sh_textbook(d)
    assert np.abs(got - want).max() <= 5e-6


def test_sh_basis_orthonormal_on_sphere(oracle_lib):
    lib = oracle_lib.load()
    n = 4000
    k = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * k / n)
    theta = math.pi * (1 + 5 ** 0.5) * k
    d = np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)], axis=1)
    B = _oracle_basis(lib, d)
    G = (B.T @ B) * (4 * math.pi / n)
    assert np.abs(G - np.eye(16)).max() <= 5e-3


@pytest.mark.parametrize("active,deg", [(1, 0), (4, 1), (9, 2), (16, 3)])
def test_sh_active_bases_truncate_degree(oracle_lib, active, deg):
    lib = oracle_lib.load()
    rng = np.random.default_rng(4 + active)
    sh0 = rng.normal(size=3).astype(np.float32)
    shr = rng.normal(size=(15, 3)).astype(np.float32)
    d = rng.normal(size=3)
    d = (d / np.linalg.norm(d)).astype(np.float32)
    rgb = np.zeros(3, dtype=np.float32)
    lib.orc_sh_eval(sh0.ctypes.data, shr.ctypes.data, d.ctypes.data, active, rgb.ctypes.data)
    Y = _real_sh_textbook(d[None].astype(np.float64))[0]
    nb = (deg + 1) ** 2
    coeffs = np.concatenate([sh0[None], shr], axis=0).astype(np.float64)
    want = 0.5 + Y[:nb] @ coeffs[:nb]
    assert np.abs(rgb - want).max() <= 2e-5
