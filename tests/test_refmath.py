"""CPU checks of the plain path's elementary functions (refmath.h), through
their host build (librefmath_host.so; the kernels compile the same header):
- sin, cos and the sin / cos of sincos equal the host libm's bit for bit
  (glibc on x86-64: sin and cos dispatched to the build with fused
  multiply-adds, sincos without), and so does hypot;
- the table of sin / cos(i/128) double-double pairs equals a 300-bit mpmath
  evaluation rounded as documented.
The GPU test test_gpu_exact.py repeats the libm comparison on the device.
"""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "fnec-rust_amd", "librefmath_host.so")


@pytest.fixture(scope="module")
def refmath():
    if not os.path.exists(LIB):
        pytest.skip("librefmath_host.so is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB)
    L.fnec_refmath_eval.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def evaluate(L, fn, x, y=None):
    out = np.zeros_like(x)
    L.fnec_refmath_eval(fn, _p(x), _p(y), x.size, _p(out))
    return out


@pytest.mark.parametrize("fn", [0, 1, 5, 6])
def test_sin_cos_sincos_match_host_libm(refmath, oracle, fn):
    rng = np.random.default_rng(fn)
    for lo, hi in ((-0.13, 0.13), (-0.86, 0.86), (-2.5, 2.5), (-80.0, 80.0), (-1e5, 1e5)):
        x = np.ascontiguousarray(rng.uniform(lo, hi, 400_000))
        dev, host = evaluate(refmath, fn, x), oracle.libm(fn, x)
        assert np.array_equal(dev.view(np.uint64), host.view(np.uint64)), (fn, lo, hi)


def test_sincos_differs_from_separate_calls(refmath, oracle):
    """The two builds of the method disagree in the last bit on about 0.1% of
    arguments: the fill must use the sincos one (a guard that the sincos and
    the sin / cos paths are not accidentally the same code)."""
    x = np.ascontiguousarray(np.random.default_rng(9).uniform(-60.0, 60.0, 400_000))
    n_diff = int(np.count_nonzero(oracle.libm(5, x) != oracle.libm(0, x)))
    assert 0 < n_diff < x.size // 100
    assert np.array_equal(evaluate(refmath, 5, x), oracle.libm(5, x))


def test_hypot_matches_host_libm(refmath, oracle):
    rng = np.random.default_rng(3)
    for scale in (1.0, 1e-3, 1e3, 1e-160, 1e160, 1e-310):
        x = np.ascontiguousarray(rng.uniform(-1, 1, 300_000) * scale)
        y = np.ascontiguousarray(rng.uniform(-1, 1, 300_000) * scale * rng.choice([1e-9, 1e-2, 1.0], 300_000))
        assert np.array_equal(evaluate(refmath, 2, x, y), oracle.libm(2, x, y)), scale


def test_sincos_table_is_the_mpmath_double_double():
    mp = pytest.importorskip("mpmath")
    mp.mp.prec = 300
    text = open(os.path.join(ROOT, "fnec-rust_amd", "csrc", "sincos_table.h")).read()
    body = text[text.index("FNEC_SINCOS_TABLE_INIT"):].split("\n")[1:]
    vals = [float.fromhex(v.strip()) if v.strip() != "0.0" else 0.0
            for line in body for v in line.replace("\\", "").split(",") if v.strip()]
    assert len(vals) == 111 * 4
    for i in range(111):
        xi = mp.mpf(i) / 128
        s, c = mp.sin(xi), mp.cos(xi)
        assert vals[4 * i] == float(s) and vals[4 * i + 1] == float(s - float(s))
        assert vals[4 * i + 2] == float(c) and vals[4 * i + 3] == float(c - float(c))


def test_phase_table_is_the_nearest_double():
    """phase_table.h (sincos_tab of the fast fill): sin and cos of j pi/32 as
    the nearest doubles, against 300-bit mpmath, with exact zeros and ones at
    the quarter turns and the symmetries sin(j) = cos(16 - j) and
    sin(j + 32) = -sin(j) holding exactly."""
    mp = pytest.importorskip("mpmath")
    mp.mp.prec = 300
    text = open(os.path.join(ROOT, "fnec-rust_amd", "csrc", "phase_table.h")).read()
    body = text[text.index("FNEC_PHASE_TABLE_INIT"):].split("\n")[1:]
    vals = [float.fromhex(v.strip(" {}")) if v.strip(" {}") != "0.0" else 0.0
            for line in body for v in line.replace("\\", "").split(",") if v.strip(" {}")]
    assert len(vals) == 128
    for j in range(64):
        s, c = vals[2 * j], vals[2 * j + 1]
        x = mp.pi * j / 32
        assert abs(s - mp.sin(x)) <= abs(np.spacing(s)) / 2 + mp.mpf(2) ** -280
        assert abs(c - mp.cos(x)) <= abs(np.spacing(c)) / 2 + mp.mpf(2) ** -280
        assert s == -vals[2 * ((j + 32) % 64)] and c == -vals[2 * ((j + 32) % 64) + 1]
        if j <= 16:
            assert s == vals[2 * (16 - j) + 1]
    assert vals[0:2] == [0.0, 1.0] and vals[32:34] == [1.0, 0.0]
