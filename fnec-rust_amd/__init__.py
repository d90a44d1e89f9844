"""ctypes binding of ``libfnec_hip.so`` (the C ABI in ``include/fnec_hip.h``).

This is the Python-side mirror of the reference's ``nec_accel`` entry points
(crates/nec_accel/src/wgpu_device.rs:461,832,1294): callers pass segment
geometry, Hallén right-hand sides and wire/junction topology as arrays and get
currents, impedances and pattern rows back. Every function raises
``FnecError`` on a non-zero status; ``FnecUnavailable`` when the library or a
gfx950 device is missing (the reference's ``None`` -> CPU fallback signal).
There is no CPU fallback inside this module.
"""
from __future__ import annotations

import ctypes as C  # noqa: N812
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfnec_hip.so")

OK, SINGULAR, DIM_MISMATCH, UNAVAILABLE, RESOURCE_EXHAUSTED, INTERNAL, INVALID_ARG = range(7)
PREC_FP64, PREC_MIXED, PREC_FP64_MFMA = 0, 1, 2
GROUND = {"free": 0, "pec": 1, "finite": 2}

SEG = np.dtype([("mid", np.float64, 3), ("dir", np.float64, 3), ("length", np.float64),
                ("radius", np.float64)])

EXPORTS = [
    "fnec_hip_device_count", "fnec_hip_init", "fnec_hip_last_error", "fnec_hip_fill_z",
    "fnec_hip_solve_hallen", "fnec_hip_solve_hallen_paths", "fnec_hip_rp",
    "fnec_hip_solve_batch", "fnec_hip_session_create", "fnec_hip_session_submit",
    "fnec_hip_session_poll", "fnec_hip_session_pending", "fnec_hip_session_destroy",
    "fnec_hip_request_footprint", "fnec_hip_debug_normal_eqs", "fnec_hip_debug_lu_f32", "fnec_hip_debug_mfma_probe",
    "fnec_hip_debug_mfma_probe_f32", "fnec_hip_set_tunable", "fnec_hip_get_tunable",
    "fnec_hip_debug_solve_z", "fnec_hip_debug_solve_square", "fnec_hip_debug_refmath",
    "fnec_hip_rp_noalloc", "fnec_hip_debug_fill_fast", "fnec_hip_solve_hallen_planewave",
    "fnec_hip_solve_hallen_current_source",
]


class FnecError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"fnec_hip status {status}: {msg}")
        self.status = status


class FnecUnavailable(FnecError):
    pass


class Ground(C.Structure):
    _fields_ = [("kind", C.c_int), ("eps_r", C.c_double), ("sigma", C.c_double)]


class SolveStats(C.Structure):
    _fields_ = [("abs_res", C.c_double), ("rel_res", C.c_double), ("diag_spread", C.c_double),
                ("backward_error", C.c_double), ("refine_iters", C.c_int), ("n_swaps", C.c_int),
                ("exact_panels", C.c_int), ("rows", C.c_int), ("unknowns", C.c_int), ("t_fill_us", C.c_double),
                ("t_gram_us", C.c_double), ("t_lu_us", C.c_double), ("t_solve_us", C.c_double),
                ("t_total_us", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class C64(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class Stamp(C.Structure):
    _fields_ = [("row", C.c_size_t), ("col", C.c_size_t), ("delta", C64)]


STAMP = np.dtype([("row", np.uint64), ("col", np.uint64), ("re", np.float64), ("im", np.float64)])


class Request(C.Structure):
    _fields_ = [("segs", C.c_void_p), ("n", C.c_size_t), ("rhs", C.c_void_p), ("cos_vec", C.c_void_p),
                ("wire_first", C.c_void_p), ("wire_last", C.c_void_p), ("n_wires", C.c_size_t),
                ("j_a", C.c_void_p), ("j_b", C.c_void_p), ("j_sign", C.c_void_p), ("n_junc", C.c_size_t),
                ("freq_hz", C.c_double), ("ground", Ground), ("feed_index", C.c_size_t),
                ("v_feed", C64), ("rp_theta", C.c_void_p), ("rp_phi", C.c_void_p),
                ("rp_npts", C.c_size_t), ("tag", C.c_uint64), ("x_out", C.c_void_p), ("rp_out", C.c_void_p),
                ("stamps", C.c_void_p), ("n_stamps", C.c_size_t)]


class Result(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("status", C.c_int), ("z_in", C64), ("feed_current", C64),
                ("x", C.c_void_p), ("rp", C.c_void_p), ("rp_total", C.c_double),
                ("t_submit_us", C.c_double), ("t_start_us", C.c_double), ("t_done_us", C.c_double),
                ("batch", C.c_uint64), ("stats", SolveStats)]


def build(force: bool = True) -> str:
    """Compile libfnec_hip.so for gfx950 with hipcc (no GPU needed)."""
    args = ["make", "-s", "-C", HERE, "-j4"]
    if force:
        args.insert(3, "-B")
    subprocess.run(args, check=True)
    return LIB_PATH


_lib = None


def lib():
    """Load the HIP library; raises FnecUnavailable if it was never built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FnecUnavailable(UNAVAILABLE, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, sz, dbl, i32 = C.c_void_p, C.c_size_t, C.c_double, C.c_int
        L.fnec_hip_last_error.restype = C.c_char_p
        L.fnec_hip_fill_z.argtypes = [vp, sz, dbl, C.POINTER(Ground), vp]
        L.fnec_hip_debug_fill_fast.argtypes = [vp, sz, dbl, C.POINTER(Ground), vp]
        L.fnec_hip_solve_hallen.argtypes = [vp, sz, vp, vp, vp, vp, sz, vp, vp, vp, sz, dbl,
                                            C.POINTER(Ground), i32, vp, C.POINTER(SolveStats)]
        L.fnec_hip_solve_hallen_paths.argtypes = [vp, sz, vp, vp, vp, vp, sz, dbl, C.POINTER(Ground),
                                                  i32, vp, C.POINTER(SolveStats)]
        L.fnec_hip_solve_hallen_planewave.argtypes = [vp, sz, vp, sz, vp, vp, vp, vp, sz, vp, vp, sz, vp, sz,
                                                      dbl, C.POINTER(Ground), i32, vp, C.POINTER(SolveStats)]
        L.fnec_hip_solve_hallen_current_source.argtypes = [vp, sz, vp, vp, sz, C64, vp, vp, sz, vp, vp, sz, vp,
                                                           sz, dbl, C.POINTER(Ground), i32, vp,
                                                           C.POINTER(SolveStats)]
        L.fnec_hip_rp.argtypes = [vp, vp, sz, dbl, C.POINTER(Ground), vp, vp, sz, vp, C.POINTER(dbl)]
        L.fnec_hip_rp_noalloc.argtypes = [vp, vp, sz, dbl, C.POINTER(Ground), vp, vp, sz, vp, C.POINTER(dbl)]
        L.fnec_hip_solve_batch.argtypes = [vp, sz, i32, vp]
        L.fnec_hip_session_create.restype = vp
        L.fnec_hip_session_create.argtypes = [sz, i32, i32]
        L.fnec_hip_session_submit.argtypes = [vp, C.POINTER(Request)]
        L.fnec_hip_session_poll.argtypes = [vp, vp, i32, i32]
        L.fnec_hip_session_pending.argtypes = [vp]
        L.fnec_hip_session_destroy.argtypes = [vp]
        L.fnec_hip_request_footprint.restype = sz
        L.fnec_hip_request_footprint.argtypes = [sz, sz, sz, sz, i32]
        L.fnec_hip_debug_normal_eqs.argtypes = [vp, sz, vp, vp, vp, vp, sz, vp, vp, vp, sz, dbl,
                                                C.POINTER(Ground), vp]
        L.fnec_hip_debug_lu_f32.argtypes = [vp, sz, vp, vp, vp, vp, sz, vp, vp, vp, sz, dbl, vp, vp, vp]
        L.fnec_hip_debug_mfma_probe.argtypes = [vp, vp, vp]
        L.fnec_hip_debug_solve_z.argtypes = [vp, sz, vp, vp, vp, vp, sz, vp, vp, vp, sz, vp,
                                             C.POINTER(SolveStats)]
        L.fnec_hip_debug_solve_square.argtypes = [vp, vp, sz, vp, vp, C.POINTER(SolveStats)]
        L.fnec_hip_debug_refmath.argtypes = [i32, vp, vp, sz, vp]
        L.fnec_hip_debug_mfma_probe_f32.argtypes = [vp, vp, vp]
        L.fnec_hip_set_tunable.argtypes = [C.c_char_p, C.c_long]
        L.fnec_hip_get_tunable.argtypes = [C.c_char_p]
        L.fnec_hip_get_tunable.restype = C.c_long
        _lib = L
    return _lib


def _check(st: int):
    if st != OK:
        msg = lib().fnec_hip_last_error()
        msg = msg.decode(errors="replace") if msg else ""
        if st == UNAVAILABLE:
            raise FnecUnavailable(st, msg)
        raise FnecError(st, msg)


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


def ground(kind: str = "free", eps_r: float = 0.0, sigma: float = 0.0) -> Ground:
    return Ground(GROUND[kind], float(eps_r), float(sigma))


def segments(mid, dirv, length, radius) -> np.ndarray:
    """Pack per-segment arrays into the fnec_seg layout."""
    n = len(length)
    s = np.zeros(n, SEG)
    s["mid"] = mid
    s["dir"] = dirv
    s["length"] = length
    s["radius"] = radius
    return s


def _sz(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def init() -> None:
    _check(lib().fnec_hip_init())


def device_count() -> int:
    return lib().fnec_hip_device_count()


def fill_z(segs: np.ndarray, freq_hz: float, g: Ground | None = None, fast: bool = False) -> np.ndarray:
    """Z (row-major): the plain path's fill, or with fast the mixed and
    FP64-MFMA paths' (fnec_hip_debug_fill_fast)."""
    segs = np.ascontiguousarray(segs, SEG)
    n = len(segs)
    z = np.zeros((n, n), np.complex128)
    fn = lib().fnec_hip_debug_fill_fast if fast else lib().fnec_hip_fill_z
    _check(fn(_p(segs), n, float(freq_hz), C.byref(g or ground()), _p(z)))
    return z


def solve_hallen(segs, rhs, cosv, wire_first=(), wire_last=(), ja=(), jb=(), js=(), freq_hz=14.2e6,
                 g: Ground | None = None, precision: int = PREC_FP64, with_stats: bool = False):
    """fnec_hip_solve_hallen: returns x (currents then homogeneous constants)."""
    segs = np.ascontiguousarray(segs, SEG)
    n = len(segs)
    wf, wl, a, b = _sz(wire_first), _sz(wire_last), _sz(ja), _sz(jb)
    sg = np.ascontiguousarray(np.asarray(js, np.float64))
    rhs = np.ascontiguousarray(rhs, np.complex128)
    cosv = np.ascontiguousarray(cosv, np.float64)
    x = np.zeros(n + max(len(wf), 1), np.complex128)
    st = SolveStats()
    _check(lib().fnec_hip_solve_hallen(_p(segs), n, _p(rhs), _p(cosv), _p(wf), _p(wl), len(wf),
                                       _p(a), _p(b), _p(sg), len(a), float(freq_hz),
                                       C.byref(g or ground()), int(precision), _p(x), C.byref(st)))
    return (x, st.as_dict()) if with_stats else x


def solve_hallen_paths(segs, rhs, cosv, path_of, free_ends, freq_hz, g: Ground | None = None,
                       precision: int = PREC_FP64, with_stats: bool = False):
    segs = np.ascontiguousarray(segs, SEG)
    n = len(segs)
    po, fe = _sz(path_of), _sz(free_ends)
    npaths = int(po.max()) + 1 if n else 0
    x = np.zeros(n + npaths, np.complex128)
    st = SolveStats()
    _check(lib().fnec_hip_solve_hallen_paths(_p(segs), n, _p(np.ascontiguousarray(rhs, np.complex128)),
                                             _p(np.ascontiguousarray(cosv, np.float64)), _p(po), _p(fe),
                                             len(fe), float(freq_hz), C.byref(g or ground()),
                                             int(precision), _p(x), C.byref(st)))
    return (x, st.as_dict()) if with_stats else x


def _stamp_arr(stamps):
    if stamps is None or len(stamps) == 0:
        return None, 0
    a = np.ascontiguousarray(stamps, STAMP)
    return a, len(a)


def solve_hallen_planewave(segs, rhs, cosv, sinv, freq_hz, wire_first=(), wire_last=(), path_of=None,
                           free_ends=(), g: Ground | None = None, precision: int = PREC_FP64,
                           stamps=None, with_stats: bool = False):
    """fnec_hip_solve_hallen_planewave: rhs is (n,) or (n_rhs, n); returns x of
    shape (n_rhs, n + 2 G) (or (n + 2 G,) for one incidence)."""
    segs = np.ascontiguousarray(segs, SEG)
    n = len(segs)
    r = np.ascontiguousarray(np.atleast_2d(rhs), np.complex128)
    nrhs = r.shape[0]
    wf, wl, fe = _sz(wire_first), _sz(wire_last), _sz(free_ends)
    po = None if path_of is None else _sz(path_of)
    groups = (int(po.max()) + 1 if n else 0) if po is not None else max(len(wf), 1)
    x = np.zeros((nrhs, n + 2 * groups), np.complex128)
    sa, ns = _stamp_arr(stamps)
    st = SolveStats()
    _check(lib().fnec_hip_solve_hallen_planewave(
        _p(segs), n, _p(r), nrhs, _p(np.ascontiguousarray(cosv, np.float64)),
        _p(np.ascontiguousarray(sinv, np.float64)), _p(wf), _p(wl), len(wf), _p(po) if po is not None else None,
        _p(fe), len(fe), _p(sa) if sa is not None else None, ns, float(freq_hz), C.byref(g or ground()),
        int(precision), _p(x), C.byref(st)))
    x = x if np.ndim(rhs) == 2 else x[0]
    return (x, st.as_dict()) if with_stats else x


def solve_hallen_current_source(segs, shape, cosv, src, i0, freq_hz, wire_first=(), wire_last=(), path_of=None,
                                free_ends=(), g: Ground | None = None, precision: int = PREC_FP64,
                                stamps=None, with_stats: bool = False):
    """fnec_hip_solve_hallen_current_source: x = currents, constants, port voltage."""
    segs = np.ascontiguousarray(segs, SEG)
    n = len(segs)
    wf, wl, fe = _sz(wire_first), _sz(wire_last), _sz(free_ends)
    po = None if path_of is None else _sz(path_of)
    groups = (int(po.max()) + 1 if n else 0) if po is not None else max(len(wf), 1)
    x = np.zeros(n + groups + 1, np.complex128)
    sa, ns = _stamp_arr(stamps)
    st = SolveStats()
    _check(lib().fnec_hip_solve_hallen_current_source(
        _p(segs), n, _p(np.ascontiguousarray(shape, np.complex128)), _p(np.ascontiguousarray(cosv, np.float64)),
        int(src), C64(complex(i0).real, complex(i0).imag), _p(wf), _p(wl), len(wf),
        _p(po) if po is not None else None, _p(fe), len(fe), _p(sa) if sa is not None else None, ns,
        float(freq_hz), C.byref(g or ground()), int(precision), _p(x), C.byref(st)))
    return (x, st.as_dict()) if with_stats else x


def rp(segs, currents, freq_hz, theta, phi, g: Ground | None = None):
    """fnec_hip_rp: returns (rows (npts, 4), normalisation integral)."""
    segs = np.ascontiguousarray(segs, SEG)
    cur = np.ascontiguousarray(currents, np.complex128)
    th = np.ascontiguousarray(theta, np.float64)
    ph = np.ascontiguousarray(phi, np.float64)
    out = np.zeros((len(th), 4))
    total = C.c_double(0.0)
    _check(lib().fnec_hip_rp(_p(segs), _p(cur), len(segs), float(freq_hz), C.byref(g or ground()),
                             _p(th), _p(ph), len(th), _p(out), C.byref(total)))
    return out, total.value


def debug_normal_eqs(segs, rhs, cosv, wire_first=(), wire_last=(), ja=(), jb=(), js=(),
                     freq_hz=14.2e6, g: Ground | None = None) -> np.ndarray:
    segs = np.ascontiguousarray(segs, SEG)
    n = len(segs)
    wf, wl, a, b = _sz(wire_first), _sz(wire_last), _sz(ja), _sz(jb)
    sg = np.ascontiguousarray(np.asarray(js, np.float64))
    s = n + max(len(wf), 1)
    aug = np.zeros((s, s + 1), np.complex128)
    _check(lib().fnec_hip_debug_normal_eqs(_p(segs), n, _p(np.ascontiguousarray(rhs, np.complex128)),
                                           _p(np.ascontiguousarray(cosv, np.float64)), _p(wf), _p(wl),
                                           len(wf), _p(a), _p(b), _p(sg), len(a), float(freq_hz),
                                           C.byref(g or ground()), _p(aug)))
    return aug


def debug_solve_z(z, rhs, cosv, wire_first=(), wire_last=(), ja=(), jb=(), js=()):
    """The plain path (normal equations, elimination, back substitution) on a
    caller-supplied Z (n x n): (x, stats dict)."""
    z = np.ascontiguousarray(z, np.complex128)
    n = z.shape[0]
    wf, wl, a, b = _sz(wire_first), _sz(wire_last), _sz(ja), _sz(jb)
    sg = np.ascontiguousarray(np.asarray(js, np.float64))
    x = np.zeros(n + max(len(wf), 1), np.complex128)
    st = SolveStats()
    _check(lib().fnec_hip_debug_solve_z(_p(z), n, _p(np.ascontiguousarray(rhs, np.complex128)),
                                        _p(np.ascontiguousarray(cosv, np.float64)), _p(wf), _p(wl), len(wf),
                                        _p(a), _p(b), _p(sg), len(a), _p(x), C.byref(st)))
    return x, st.as_dict()


def debug_solve_square(a, b):
    """The plain path's elimination + back substitution of a x = b:
    (status, x, ipiv, stats dict); status SINGULAR leaves x unset."""
    a = np.ascontiguousarray(a, np.complex128)
    b = np.ascontiguousarray(b, np.complex128)
    n = a.shape[0]
    x = np.zeros(n, np.complex128)
    ipiv = np.zeros(n, np.int32)
    st = SolveStats()
    rc = lib().fnec_hip_debug_solve_square(_p(a), _p(b), n, _p(x), _p(ipiv), C.byref(st))
    if rc not in (OK, SINGULAR):
        _check(rc)
    return rc, x, ipiv, st.as_dict()


def debug_refmath(fn: int, x, y=None) -> np.ndarray:
    """The plain path's device sin (0), cos (1), hypot (2), sqrt (3), x / y (4),
    and the sin (5) and cos (6) of sincos."""
    x = np.ascontiguousarray(x, np.float64)
    yy = None if y is None else np.ascontiguousarray(y, np.float64)
    out = np.zeros_like(x)
    _check(lib().fnec_hip_debug_refmath(int(fn), _p(x), _p(yy), x.size, _p(out)))
    return out


def debug_lu_f32(p: dict):
    """The mixed path's FP32 LU of Z for a Deck.problem dict: (packed factors
    L\\U as complex64 n x n, ipiv int32 n) with P Z32 = L U."""
    segs = np.ascontiguousarray(p["segs"], SEG)
    n = len(segs)
    wf, wl, a, b = _sz(p["wf"]), _sz(p["wl"]), _sz(p["ja"]), _sz(p["jb"])
    sg = np.ascontiguousarray(np.asarray(p["js"], np.float64))
    lu = np.zeros((n, n), np.complex64)
    ipiv = np.zeros(n, np.int32)
    _check(lib().fnec_hip_debug_lu_f32(_p(segs), n, _p(np.ascontiguousarray(p["rhs"], np.complex128)),
                                       _p(np.ascontiguousarray(p["cos"], np.float64)), _p(wf), _p(wl), len(wf),
                                       _p(a), _p(b), _p(sg), len(a), float(p["freq_hz"]),
                                       C.byref(p["ground_struct"]), _p(lu), _p(ipiv)))
    return lu, ipiv


def debug_mfma_probe(a16x4: np.ndarray, b4x16: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a16x4, np.float64)
    b = np.ascontiguousarray(b4x16, np.float64)
    out = np.zeros((64, 4))
    _check(lib().fnec_hip_debug_mfma_probe(_p(a), _p(b), _p(out)))
    return out


def set_tunable(name: str, value: int) -> None:
    _check(lib().fnec_hip_set_tunable(name.encode(), int(value)))


def get_tunable(name: str) -> int:
    return int(lib().fnec_hip_get_tunable(name.encode()))


def debug_mfma_probe_f32(a16x4: np.ndarray, b4x16: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a16x4, np.float32)
    b = np.ascontiguousarray(b4x16, np.float32)
    out = np.zeros((64, 4), np.float32)
    _check(lib().fnec_hip_debug_mfma_probe_f32(_p(a), _p(b), _p(out)))
    return out


# ---------------------------------------------------------------------------
# Host side (include/fnec_host.h): decks -> Hallén inputs, CLI report
# ---------------------------------------------------------------------------

SEM = {"cli": 0, "worker": 1, "py": 2}


class DeckProblem(C.Structure):
    _fields_ = [("route", C.c_int), ("req", Request), ("path_of", C.c_void_p), ("free_ends", C.c_void_p),
                ("n_free", C.c_size_t), ("n_feeds", C.c_size_t), ("feed_idx", C.c_void_p),
                ("feed_v", C.c_void_p), ("feed_tag", C.c_void_p), ("feed_seg", C.c_void_p),
                ("seg_tag", C.c_void_p), ("seg_index", C.c_void_p), ("variant", C.c_int),
                ("sin_vec", C.c_void_p), ("n_rhs", C.c_size_t), ("rhs_all", C.c_void_p), ("n_recv", C.c_size_t),
                ("recv_theta", C.c_void_p), ("recv_phi", C.c_void_p), ("src", C.c_size_t), ("i0", C64),
                ("feed_port", C.c_void_p)]


def _host_sigs(L):
    if getattr(L, "_host_ready", False):
        return L
    L.fnec_deck_parse.restype = C.c_void_p
    L.fnec_deck_parse.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t]
    L.fnec_deck_free.argtypes = [C.c_void_p]
    L.fnec_deck_frequencies.restype = C.c_size_t
    L.fnec_deck_frequencies.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
    L.fnec_deck_segment_count.restype = C.c_size_t
    L.fnec_deck_segment_count.argtypes = [C.c_void_p]
    L.fnec_deck_build.argtypes = [C.c_void_p, C.c_int, C.c_double, C.POINTER(DeckProblem)]
    L.fnec_deck_solve.argtypes = [C.POINTER(DeckProblem), C.c_int, C.c_void_p, C.c_void_p]
    L.fnec_deck_report.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.c_char_p, C.c_size_t,
                                   C.POINTER(C.c_size_t)]
    L._host_ready = True
    return L


def _arr(ptr, n, dtype):
    if not ptr or n == 0:
        return np.zeros(0, dtype)
    nbytes = n * np.dtype(dtype).itemsize
    return np.frombuffer((C.c_char * nbytes).from_address(ptr), dtype=dtype).copy()


class Deck:
    """A parsed NEC deck (nec_parser::parse) with the reference callers'
    builders: frequencies, geometry, Hallén inputs, CLI report."""

    def __init__(self, text: str):
        L = _host_sigs(lib())
        err = C.create_string_buffer(512)
        h = L.fnec_deck_parse(text.encode(), err, 512)
        if not h:
            raise FnecError(INVALID_ARG, err.value.decode(errors="replace"))
        self._h = h

    def close(self):
        if self._h:
            lib().fnec_deck_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def frequencies(self, semantics: str = "cli") -> np.ndarray:
        L = lib()
        n = L.fnec_deck_frequencies(self._h, SEM[semantics], None, 0)
        out = np.zeros(n)
        L.fnec_deck_frequencies(self._h, SEM[semantics], _p(out), n)
        return out

    def problem(self, semantics: str, freq_hz: float) -> dict:
        """Hallén inputs as numpy arrays (copies)."""
        L = lib()
        pr = DeckProblem()
        _check(L.fnec_deck_build(self._h, SEM[semantics], float(freq_hz), C.byref(pr)))
        r = pr.req
        n = r.n
        out = {
            "route": pr.route,
            "segs": _arr(r.segs, n, SEG),
            "rhs": _arr(r.rhs, n, np.complex128),
            "cos": _arr(r.cos_vec, n, np.float64),
            "wf": _arr(r.wire_first, r.n_wires, np.uint64),
            "wl": _arr(r.wire_last, r.n_wires, np.uint64),
            "ja": _arr(r.j_a, r.n_junc, np.uint64),
            "jb": _arr(r.j_b, r.n_junc, np.uint64),
            "js": _arr(r.j_sign, r.n_junc, np.float64),
            "freq_hz": r.freq_hz,
            "ground": (r.ground.kind, r.ground.eps_r, r.ground.sigma),
            "ground_struct": Ground(r.ground.kind, r.ground.eps_r, r.ground.sigma),
            "feed_index": r.feed_index,
            "v_feed": complex(r.v_feed.re, r.v_feed.im),
            "rp_theta": _arr(r.rp_theta, r.rp_npts, np.float64),
            "rp_phi": _arr(r.rp_phi, r.rp_npts, np.float64),
            "path_of": _arr(pr.path_of, n if pr.route else 0, np.uint64),
            "free_ends": _arr(pr.free_ends, pr.n_free, np.uint64),
            "feed_idx": _arr(pr.feed_idx, pr.n_feeds, np.uint64),
            "feed_v": _arr(pr.feed_v, pr.n_feeds, np.complex128),
            "feed_tag": _arr(pr.feed_tag, pr.n_feeds, np.uint32),
            "feed_seg": _arr(pr.feed_seg, pr.n_feeds, np.uint32),
            "seg_tag": _arr(pr.seg_tag, n, np.uint32),
            "seg_index": _arr(pr.seg_index, n, np.uint32),
            "stamps": _arr(r.stamps, r.n_stamps, STAMP),
            "variant": pr.variant,
        }
        if pr.variant == 1:
            out["sin"] = _arr(pr.sin_vec, n, np.float64)
            out["rhs_all"] = _arr(pr.rhs_all, n * pr.n_rhs, np.complex128).reshape(pr.n_rhs, n)
            out["recv_theta"] = _arr(pr.recv_theta, pr.n_recv, np.float64)
            out["recv_phi"] = _arr(pr.recv_phi, pr.n_recv, np.float64)
        if pr.variant == 2:
            out["src"] = pr.src
            out["i0"] = complex(pr.i0.re, pr.i0.im)
            out["feed_port"] = _arr(pr.feed_port, pr.n_feeds, np.uint8)
        return out

    def report(self, freq_index: int = 0, precision: int = PREC_FP64) -> str:
        L = lib()
        need = C.c_size_t(0)
        _check(L.fnec_deck_report(self._h, freq_index, precision, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        _check(L.fnec_deck_report(self._h, freq_index, precision, buf, need.value, C.byref(need)))
        return buf.value.decode()


def solve_problem(p: dict, precision: int = PREC_FP64, with_stats: bool = False):
    """Solve the dict returned by Deck.problem on the GPU (either route, with
    its LD/TL/NT stamps) through fnec_deck_solve."""
    L = _host_sigs(lib())
    q, keep = make_request(p)
    pr = DeckProblem()
    pr.route = int(p["route"])
    pr.req = q
    n = len(keep["segs"])
    if pr.route == 1:
        keep["po"] = _sz(p["path_of"])
        keep["fe"] = _sz(p["free_ends"])
        pr.path_of = _p(keep["po"])
        pr.free_ends = _p(keep["fe"])
        pr.n_free = len(keep["fe"])
        groups = int(keep["po"].max()) + 1 if n else 0
    else:
        groups = max(len(keep["wf"]), 1)
    pr.variant = int(p.get("variant", 0))
    shape = None
    if pr.variant == 1:
        keep["sin"] = np.ascontiguousarray(p["sin"], np.float64)
        keep["rhs_all"] = np.ascontiguousarray(p["rhs_all"], np.complex128)
        pr.sin_vec = _p(keep["sin"])
        pr.rhs_all = _p(keep["rhs_all"])
        pr.n_rhs = keep["rhs_all"].shape[0]
        shape = (pr.n_rhs, n + 2 * groups)
    elif pr.variant == 2:
        pr.src = int(p["src"])
        pr.i0 = C64(p["i0"].real, p["i0"].imag)
        shape = (n + groups + 1,)
    x = np.zeros(shape or (n + groups,), np.complex128)
    st = SolveStats()
    _check(L.fnec_deck_solve(C.byref(pr), int(precision), _p(x), C.byref(st)))
    return (x, st.as_dict()) if with_stats else x


PROFILE_CLASSES = ("fill", "gram", "lu_panel", "lu_trsm", "lu_update", "backsub", "rp", "zsolve", "matvec",
                   "pcg_ctl", "fsetup", "lu_update_narrow", "lu_update_k128",
                   # host-side phases (steady clock): problem and plan building, staging
                   # copies, far-field set-up, result processing, time blocked on the stream,
                   # kernel enqueueing up to the first wait
                   "host_prep", "host_stage", "host_rp", "host_post", "host_wait", "host_enqueue")


def profile(enable: bool | None = None, reset: bool = False, only: str | None = None) -> dict:
    """Per-kernel event timings: {class: (launches, total_us, flops, bytes)}.

    ``only`` restricts event recording to one class ("" for every class)."""
    L = lib()
    L.fnec_hip_profile_read.argtypes = [C.c_char_p, C.c_void_p]
    L.fnec_hip_profile_only.argtypes = [C.c_char_p]
    if only is not None:
        _check(L.fnec_hip_profile_only(only.encode()))
    if reset:
        L.fnec_hip_profile_reset()
    if enable is not None:
        L.fnec_hip_profile_enable(int(enable))
    out = {}
    for k in PROFILE_CLASSES:
        buf = np.zeros(4)
        _check(L.fnec_hip_profile_read(k.encode(), _p(buf)))
        out[k] = tuple(float(v) for v in buf)
    return out


def make_request(p: dict, tag: int = 0, rp_theta=None, rp_phi=None, outputs: bool = False) -> tuple:
    """fnec_request for a Deck.problem dict (worker route); returns
    (request, keepalive) — keep the second item alive while the request is used.
    With outputs, the request carries buffers for the currents (keepalive "x")
    and, with a grid, the RP rows (keepalive "rows") that a session fills."""
    keep = {k: np.ascontiguousarray(p[k]) for k in ("segs", "rhs", "cos", "wf", "wl", "ja", "jb", "js")}
    q = Request()
    q.segs = _p(keep["segs"])
    q.n = len(keep["segs"])
    q.rhs = _p(keep["rhs"])
    q.cos_vec = _p(keep["cos"])
    q.wire_first = _p(keep["wf"])
    q.wire_last = _p(keep["wl"])
    q.n_wires = len(keep["wf"])
    q.j_a = _p(keep["ja"])
    q.j_b = _p(keep["jb"])
    q.j_sign = _p(keep["js"])
    q.n_junc = len(keep["ja"])
    q.freq_hz = p["freq_hz"]
    q.ground = p["ground_struct"]
    q.feed_index = int(p["feed_index"])
    q.v_feed = C64(p["v_feed"].real, p["v_feed"].imag)
    if len(p.get("stamps", ())):
        keep["stamps"] = np.ascontiguousarray(p["stamps"], STAMP)
        q.stamps = _p(keep["stamps"])
        q.n_stamps = len(keep["stamps"])
    if rp_theta is not None:
        keep["th"] = np.ascontiguousarray(rp_theta, np.float64)
        keep["ph"] = np.ascontiguousarray(rp_phi, np.float64)
        q.rp_theta = _p(keep["th"])
        q.rp_phi = _p(keep["ph"])
        q.rp_npts = len(keep["th"])
    q.tag = tag
    if outputs:
        keep["x"] = np.zeros(len(keep["segs"]) + max(len(keep["wf"]), 1), np.complex128)
        q.x_out = _p(keep["x"])
        if rp_theta is not None:
            keep["rows"] = np.zeros((len(keep["th"]), 4))
            q.rp_out = _p(keep["rows"])
    return q, keep


def solve_batch(problems: list, precision: int = PREC_FP64, rp_grids: list | None = None):
    """fnec_hip_solve_batch over Deck.problem dicts; returns a list of dicts
    with z_in, x (currents + constants), stats and the RP rows if requested."""
    reqs = (Request * len(problems))()
    res = (Result * len(problems))()
    keeps, xs, rps = [], [], []
    for i, p in enumerate(problems):
        th, ph = (rp_grids[i] if rp_grids and rp_grids[i] is not None else (None, None))
        q, k = make_request(p, i, th, ph)
        reqs[i] = q
        keeps.append(k)
        x = np.zeros(len(p["segs"]) + max(len(p["wf"]), 1), np.complex128)
        xs.append(x)
        res[i].x = _p(x)
        if th is not None:
            rows = np.zeros((len(th), 4))
            rps.append(rows)
            res[i].rp = _p(rows)
        else:
            rps.append(None)
    _check(lib().fnec_hip_solve_batch(C.cast(reqs, C.c_void_p), len(problems), precision, C.cast(res, C.c_void_p)))
    out = []
    for i in range(len(problems)):
        r = res[i]
        out.append({"status": r.status, "z_in": complex(r.z_in.re, r.z_in.im), "x": xs[i], "rp": rps[i],
                    "rp_total": r.rp_total, "stats": r.stats.as_dict()})
    return out


def autotune(problems: list, precision: int = PREC_MIXED, **kw) -> dict:
    """Pick the engine's knobs (update tile and K = 128 passes, CG gang shape,
    solve size limit, CG read-back chunk) by timing batched solves of
    `problems`; see
    tuning.autotune for the options."""
    from .tuning import autotune as _autotune

    return _autotune(problems, precision, **kw)
