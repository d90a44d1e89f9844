"""GPU test of fnec_hip_rp_noalloc, the far-field entry point that allocates
no device memory (its job lives in a fixed 32 MiB device array), through the
C ABI: two calls in a row, the second on a smaller job, each equal bit for bit
to fnec_hip_rp on the same inputs; a job beyond the array is refused with
RESOURCE_EXHAUSTED and leaves both entry points working."""
import ctypes as C
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F = importlib.import_module("fnec-rust_amd")
W = importlib.import_module("fnec-rust_amd.workloads")


def call(fn, p, x, th, ph):
    segs = np.ascontiguousarray(p["segs"], F.SEG)
    n = len(segs)
    cur = np.ascontiguousarray(x[:n], np.complex128)
    th = np.ascontiguousarray(th, np.float64)
    ph = np.ascontiguousarray(ph, np.float64)
    out = np.full((len(th), 4), np.nan)
    total = C.c_double(np.nan)
    st = fn(F._p(segs), F._p(cur), n, float(p["freq_hz"]), C.byref(p["ground_struct"]), F._p(th), F._p(ph),
            len(th), F._p(out), C.byref(total))
    return st, out, total.value


def test_rp_noalloc_equals_rp_on_two_calls_then_refuses_an_oversized_job():
    F.init()
    L = F.lib()
    # first call: config 2 (N = 2048 over PEC) on the full 1-degree grid
    d = F.Deck(W.config2_deck())
    p = d.problem("cli", 14.2e6)
    x = F.solve_problem(p, precision=F.PREC_FP64)
    th, ph = p["rp_theta"], p["rp_phi"]
    assert len(th) == 65341
    st_a, rows_a, tot_a = call(L.fnec_hip_rp_noalloc, p, x, th, ph)
    st_b, rows_b, tot_b = call(L.fnec_hip_rp, p, x, th, ph)
    assert st_a == st_b == F.OK
    assert np.array_equal(rows_a, rows_b) and tot_a == tot_b
    # second call, smaller: the 51-segment dipole on the 5-degree grid
    d1 = F.Deck(W.CONFIG1_DECK)
    p1 = d1.problem("cli", d1.frequencies("cli")[0])
    x1 = F.solve_problem(p1, precision=F.PREC_FP64)
    st_a, rows_a, tot_a = call(L.fnec_hip_rp_noalloc, p1, x1, p1["rp_theta"], p1["rp_phi"])
    st_b, rows_b, tot_b = call(L.fnec_hip_rp, p1, x1, p1["rp_theta"], p1["rp_phi"])
    assert st_a == st_b == F.OK
    assert np.array_equal(rows_a, rows_b) and tot_a == tot_b
    # a grid of 600k points needs about 46 MiB (80 B per point): refused,
    # nothing written
    big = np.linspace(0.0, 180.0, 600_000)
    st, rows, tot = call(L.fnec_hip_rp_noalloc, p1, x1, big, np.zeros_like(big))
    assert st == F.RESOURCE_EXHAUSTED and np.isnan(rows).all()
    assert b"device scratch" in L.fnec_hip_last_error()
    # both entry points still work afterwards
    st_a, rows_a, tot_a = call(L.fnec_hip_rp_noalloc, p1, x1, p1["rp_theta"], p1["rp_phi"])
    assert st_a == F.OK and np.array_equal(rows_a, rows_b) and tot_a == tot_b
    st_b, rows_c, _ = call(L.fnec_hip_rp, p1, x1, p1["rp_theta"], p1["rp_phi"])
    assert st_b == F.OK and np.array_equal(rows_c, rows_b)
