"""ORACLE — test infrastructure only.

ctypes bindings to ``libfnec_oracle.so``, the C restatement of the
reference's f64 Hallén path (see ``fnec_oracle.h`` for the file:line map).
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C  # noqa: N812
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfnec_oracle.so")

SEG_DTYPE = np.dtype(
    [
        ("tag", np.uint32),
        ("tag_index", np.uint32),
        ("start", np.float64, 3),
        ("end", np.float64, 3),
        ("mid", np.float64, 3),
        ("dir", np.float64, 3),
        ("length", np.float64),
        ("radius", np.float64),
    ],
    align=True,
)

GROUND_KIND = {"free": 0, "pec": 1, "finite": 2, "deferred": 3}
SEMANTICS = {"cli": 0, "worker": 1, "py": 2}


class Ground(C.Structure):
    _fields_ = [("kind", C.c_int), ("eps_r", C.c_double), ("sigma", C.c_double)]


class CX(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class Result(C.Structure):
    _fields_ = [
        ("status", C.c_int),
        ("err", C.c_char * 256),
        ("freq_hz", C.c_double),
        ("n", C.c_int),
        ("s", C.c_int),
        ("r", C.c_int),
        ("nw", C.c_int),
        ("used_paths", C.c_int),
        ("ground", Ground),
        ("segs", C.c_void_p),
        ("x", C.c_void_p),
        ("nfeed", C.c_int),
        ("feed_tag", C.POINTER(C.c_int)),
        ("feed_seg", C.POINTER(C.c_int)),
        ("feed_idx", C.POINTER(C.c_int)),
        ("feed_v", C.c_void_p),
        ("feed_i", C.c_void_p),
        ("feed_z", C.c_void_p),
        ("np", C.c_int),
        ("rp_theta", C.POINTER(C.c_double)),
        ("rp_phi", C.POINTER(C.c_double)),
        ("rp", C.POINTER(C.c_double)),
        ("rp_total", C.c_double),
        ("abs_res", C.c_double),
        ("rel_res", C.c_double),
        ("diag_spread", C.c_double),
        ("report", C.c_char_p),
        ("variant", C.c_int),
        ("port_v", CX),
        ("nrecv", C.c_int),
        ("recv_theta", C.POINTER(C.c_double)),
        ("recv_phi", C.POINTER(C.c_double)),
        ("recv_db", C.POINTER(C.c_double)),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -O2 -ffp-contract=off)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.c_void_p
        L.or_run_deck.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(Result)]
        L.or_result_free.argtypes = [C.POINTER(Result)]
        L.or_fill_z.argtypes = [dp, C.c_int, C.c_double, Ground, dp]
        L.or_fill_elem.argtypes = [dp, dp, C.c_double, C.c_int]
        L.or_fill_elem.restype = CX
        L.or_solve_square.argtypes = [dp, dp, C.c_int, dp]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_blocked.argtypes = [C.c_int]
        L.or_solve_square_piv.argtypes = [dp, dp, C.c_int, dp, dp]
        L.or_libm_eval.argtypes = [C.c_int, dp, dp, C.c_size_t, dp]
        L.or_solve_hallen.argtypes = [dp, C.c_int, dp, dp, dp, dp, C.c_int, dp, dp, dp, C.c_int, dp]
        L.or_hallen_normal_eqs.argtypes = [dp, C.c_int, dp, dp, dp, dp, C.c_int, dp, dp, dp, C.c_int, dp, dp, dp]
        L.or_radiation_pattern.argtypes = [dp, C.c_int, dp, C.c_double, Ground, dp, dp, C.c_int, dp]
        L.or_radiation_pattern.restype = C.c_double
        L.or_power_integral.argtypes = [dp, C.c_int, dp, C.c_double, Ground]
        L.or_power_integral.restype = C.c_double
        L.or_far_field.argtypes = [dp, C.c_int, dp, C.c_double, Ground, C.c_double, C.c_double,
                                   C.POINTER(CX), C.POINTER(CX)]
        L.or_radiation_efficiency.argtypes = [dp, C.c_int, dp, C.c_double, Ground, C.c_double]
        L.or_radiation_efficiency.restype = C.c_double
        L.or_fmt_fixed.argtypes = [C.c_char_p, C.c_int, C.c_double, C.c_int]
        L.or_fmt_exp.argtypes = [C.c_char_p, C.c_int, C.c_double, C.c_int]
        L.or_axial_ratio.argtypes = [CX, CX]
        L.or_axial_ratio.restype = C.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and len(a) else None


def ground(kind: str = "free", eps_r: float = 0.0, sigma: float = 0.0) -> Ground:
    return Ground(GROUND_KIND[kind], eps_r, sigma)


def run_deck(text: str, semantics: str = "cli", freq_index: int = 0, freq_hz: float = 0.0,
             report: bool = False) -> dict:
    """Run the whole-deck pipeline; raises RuntimeError on a reference error."""
    L = lib()
    r = Result()
    st = L.or_run_deck(text.encode(), SEMANTICS[semantics], freq_index, float(freq_hz), int(report), C.byref(r))
    try:
        if st != 0:
            raise RuntimeError(f"oracle status {st}: {r.err.decode(errors='replace')}")
        n, s = r.n, r.s
        segs = np.frombuffer((C.c_char * (n * SEG_DTYPE.itemsize)).from_address(r.segs), dtype=SEG_DTYPE).copy()
        x = np.frombuffer((C.c_double * (2 * s)).from_address(r.x), dtype=np.complex128).copy()
        cplx = lambda ptr, k: np.frombuffer((C.c_double * (2 * k)).from_address(ptr), dtype=np.complex128).copy() if k else np.zeros(0, complex)
        nf = r.nfeed
        out = {
            "freq_hz": r.freq_hz,
            "n": n,
            "s": s,
            "r": r.r,
            "nw": r.nw,
            "used_paths": bool(r.used_paths),
            "ground": (r.ground.kind, r.ground.eps_r, r.ground.sigma),
            "segs": segs,
            "x": x,
            "currents": x[:n],
            "feed_tag": [r.feed_tag[i] for i in range(nf)],
            "feed_seg": [r.feed_seg[i] for i in range(nf)],
            "feed_idx": [r.feed_idx[i] for i in range(nf)],
            "feed_v": cplx(r.feed_v, nf),
            "feed_i": cplx(r.feed_i, nf),
            "feed_z": cplx(r.feed_z, nf),
            "abs_res": r.abs_res,
            "rel_res": r.rel_res,
            "diag_spread": r.diag_spread,
            "rp_total": r.rp_total,
        }
        npt = r.np
        if npt:
            out["rp_theta"] = np.array(r.rp_theta[:npt])
            out["rp_phi"] = np.array(r.rp_phi[:npt])
            out["rp"] = np.array(r.rp[: 4 * npt]).reshape(npt, 4)
        out["variant"] = ("delta-gap", "plane-wave", "current-source")[r.variant]
        out["port_v"] = complex(r.port_v.re, r.port_v.im)
        if r.nrecv:
            k = r.nrecv
            out["recv"] = np.stack([np.array(r.recv_theta[:k]), np.array(r.recv_phi[:k]), np.array(r.recv_db[:k])], 1)
        if report and r.report:
            out["report"] = r.report.decode()
        return out
    finally:
        L.or_result_free(C.byref(r))


def fill_z(segs: np.ndarray, freq_hz: float, g: Ground | None = None) -> np.ndarray:
    segs = np.ascontiguousarray(segs, dtype=SEG_DTYPE)
    n = len(segs)
    z = np.zeros((n, n), np.complex128)
    lib().or_fill_z(_p(segs), n, float(freq_hz), g or ground(), z.ctypes.data_as(C.c_void_p))
    return z


def set_threads(n: int) -> None:
    """Threads for the fill, Gram and elimination row loops (fixtures only;
    results do not depend on the count)."""
    lib().or_set_threads(int(n))


def set_blocked(on: bool) -> None:
    """Cache-blocked Gram and elimination loops (fixtures of large systems;
    results bit for bit those of the reference-order loops)."""
    lib().or_set_blocked(int(bool(on)))


def solve_square_piv(a: np.ndarray, b: np.ndarray) -> tuple[int, np.ndarray, np.ndarray]:
    """solve_square_in_place with the pivot row chosen at each column."""
    a = np.ascontiguousarray(a, np.complex128)
    b = np.ascontiguousarray(b, np.complex128)
    x = np.zeros(len(b), np.complex128)
    piv = np.full(len(b), -1, np.int32)
    st = lib().or_solve_square_piv(_p(a.ravel()), _p(b), len(b), _p(x), _p(piv))
    return st, x, piv


def libm(fn: int, x: np.ndarray, y: np.ndarray | None = None) -> np.ndarray:
    """The host libm over an array: sin (0), cos (1), hypot(x, y) (2), and the
    sin (5) and cos (6) of sincos(x)."""
    x = np.ascontiguousarray(x, np.float64)
    yy = None if y is None else np.ascontiguousarray(y, np.float64)
    out = np.zeros_like(x)
    lib().or_libm_eval(int(fn), _p(x), _p(yy), x.size, _p(out))
    return out


def solve_square(a: np.ndarray, b: np.ndarray) -> tuple[int, np.ndarray]:
    a = np.ascontiguousarray(a, np.complex128)
    b = np.ascontiguousarray(b, np.complex128)
    x = np.zeros(len(b), np.complex128)
    st = lib().or_solve_square(_p(a.ravel()), _p(b), len(b), _p(x))
    return st, x


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def solve_hallen(z, rhs, cosv, wf, wl, ja=(), jb=(), js=()):
    z = np.ascontiguousarray(z, np.complex128)
    n = z.shape[0]
    wf, wl, ja, jb = _i32(wf), _i32(wl), _i32(ja), _i32(jb)
    js = np.ascontiguousarray(np.asarray(js, np.float64))
    nw = len(wf)
    x = np.zeros(n + max(nw, 1), np.complex128)
    st = lib().or_solve_hallen(_p(z.ravel()), n, _p(np.ascontiguousarray(rhs, np.complex128)),
                               _p(np.ascontiguousarray(cosv, np.float64)), _p(wf), _p(wl), nw,
                               _p(ja), _p(jb), _p(js), len(ja), _p(x))
    return st, x


def hallen_normal_eqs(z, rhs, cosv, wf, wl, ja=(), jb=(), js=()):
    z = np.ascontiguousarray(z, np.complex128)
    n = z.shape[0]
    wf, wl, ja, jb = _i32(wf), _i32(wl), _i32(ja), _i32(jb)
    js = np.ascontiguousarray(np.asarray(js, np.float64))
    s = n + max(len(wf), 1)
    ata = np.zeros((s, s), np.complex128)
    aty = np.zeros(s, np.complex128)
    rows = C.c_int(0)
    lib().or_hallen_normal_eqs(_p(z.ravel()), n, _p(np.ascontiguousarray(cosv, np.float64)),
                               _p(np.ascontiguousarray(rhs, np.complex128)), _p(wf), _p(wl), len(wf),
                               _p(ja), _p(jb), _p(js), len(ja), _p(ata.ravel()), _p(aty),
                               C.cast(C.pointer(rows), C.c_void_p))
    return ata, aty, rows.value


def radiation_pattern(segs, cur, freq_hz, theta, phi, g: Ground | None = None):
    segs = np.ascontiguousarray(segs, dtype=SEG_DTYPE)
    cur = np.ascontiguousarray(cur, np.complex128)
    theta = np.ascontiguousarray(theta, np.float64)
    phi = np.ascontiguousarray(phi, np.float64)
    out = np.zeros((len(theta), 4))
    total = lib().or_radiation_pattern(_p(segs), len(segs), _p(cur), float(freq_hz), g or ground(),
                                       _p(theta), _p(phi), len(theta), _p(out.ravel()))
    return out, total


def power_integral(segs, cur, k, g: Ground | None = None) -> float:
    segs = np.ascontiguousarray(segs, dtype=SEG_DTYPE)
    cur = np.ascontiguousarray(cur, np.complex128)
    return lib().or_power_integral(_p(segs), len(segs), _p(cur), float(k), g or ground())


def far_field(segs, cur, k, th_deg, ph_deg, g: Ground | None = None) -> tuple[complex, complex]:
    """(F_theta, F_phi) of far_field_components(_with_image / _fresnel)."""
    segs = np.ascontiguousarray(segs, SEG_DTYPE)
    cur = np.ascontiguousarray(cur, np.complex128)
    ft, fp = CX(), CX()
    lib().or_far_field(_p(segs), len(segs), _p(cur), float(k), g or ground(), float(th_deg),
                       float(ph_deg), C.byref(ft), C.byref(fp))
    return complex(ft.re, ft.im), complex(fp.re, fp.im)


def radiation_efficiency(segs, cur, freq_hz, p_in, g: Ground | None = None) -> float:
    segs = np.ascontiguousarray(segs, SEG_DTYPE)
    cur = np.ascontiguousarray(cur, np.complex128)
    return lib().or_radiation_efficiency(_p(segs), len(segs), _p(cur), float(freq_hz), g or ground(),
                                         float(p_in))


def fmt_fixed(v: float, prec: int) -> str:
    b = C.create_string_buffer(64)
    lib().or_fmt_fixed(b, 64, float(v), prec)
    return b.value.decode()


def fmt_exp(v: float, prec: int) -> str:
    b = C.create_string_buffer(64)
    lib().or_fmt_exp(b, 64, float(v), prec)
    return b.value.decode()


def make_segments(wires, tag0: int = 1) -> np.ndarray:
    """Expand straight wires ``(start, end, nseg, radius)`` exactly as
    geometry.rs:710-754 (expand_wire) does; tags count up from ``tag0``."""
    segs = []
    for w, (p1, p2, ns, a) in enumerate(wires):
        x1, y1, z1 = (float(v) for v in p1)
        wdx, wdy, wdz = p2[0] - x1, p2[1] - y1, p2[2] - z1
        wl = float(np.sqrt(wdx * wdx + wdy * wdy + wdz * wdz))
        d = (wdx / wl, wdy / wl, wdz / wl)
        nn = float(ns)
        for i in range(ns):
            t0 = i / nn
            t1 = (i + 1.0) / nn
            tm = (t0 + t1) * 0.5
            segs.append((tag0 + w, i + 1,
                         (x1 + wdx * t0, y1 + wdy * t0, z1 + wdz * t0),
                         (x1 + wdx * t1, y1 + wdy * t1, z1 + wdz * t1),
                         (x1 + wdx * tm, y1 + wdy * tm, z1 + wdz * tm),
                         d, wl / nn, float(a)))
    return np.array(segs, dtype=SEG_DTYPE)


def axial_ratio(et: complex, ep: complex) -> float:
    return lib().or_axial_ratio(CX(et.real, et.imag), CX(ep.real, ep.imag))


def deck_inputs(text: str, freq_hz: float, semantics: str = "worker") -> dict:
    """Geometry, Hallén rhs/cos vectors, wire ranges and junction tuples of a
    deck, as the reference's worker (raw wires, all junctions) or CLI
    (merged wires, cross-component junctions) builds them."""
    L = lib()
    err = C.create_string_buffer(256)
    L.or_deck_parse.restype = C.c_void_p
    L.or_deck_parse.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    L.or_deck_free.argtypes = [C.c_void_p]
    L.or_build_geometry.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_char_p, C.c_int]
    L.or_build_hallen_rhs.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p]
    L.or_wire_endpoints.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
    L.or_merge_collinear.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
    L.or_detect_junctions.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    L.or_ground_model.restype = Ground
    L.or_ground_model.argtypes = [C.c_void_p]
    d = L.or_deck_parse(text.encode(), err, 256)
    if not d:
        raise RuntimeError(err.value.decode())
    try:
        maxn = 1 << 16
        segs = np.zeros(maxn, SEG_DTYPE)
        n = L.or_build_geometry(d, _p(segs), maxn, err, 256)
        if n < 0:
            raise RuntimeError(err.value.decode())
        segs = segs[:n].copy()
        rhs = np.zeros(n, np.complex128)
        cosv = np.zeros(n)
        st = L.or_build_hallen_rhs(d, _p(segs), n, float(freq_hz), _p(rhs), _p(cosv))
        if st:
            raise RuntimeError(f"excitation error {st}")
        wf = np.zeros(n + 1, np.int32)
        wl = np.zeros(n + 1, np.int32)
        if semantics == "worker":
            nw = L.or_wire_endpoints(_p(segs), n, _p(wf), _p(wl), n + 1)
        else:
            nw = L.or_merge_collinear(_p(segs), n, _p(wf), _p(wl), n + 1)
        wf, wl = wf[:nw].copy(), wl[:nw].copy()
        mj = 4 * n + 4
        ja = np.zeros(mj, np.int32)
        jb = np.zeros(mj, np.int32)
        js = np.zeros(mj)
        nj = L.or_detect_junctions(_p(segs), _p(wf), _p(wl), nw, 1e-6, _p(ja), _p(jb), _p(js), mj)
        ja, jb, js = ja[:nj].copy(), jb[:nj].copy(), js[:nj].copy()
        if semantics != "worker":
            comp = np.zeros(n, int)
            for i in range(nw):
                comp[wf[i]:wl[i] + 1] = i
            keep = comp[ja] != comp[jb]
            ja, jb, js = ja[keep], jb[keep], js[keep]
        g = L.or_ground_model(d)
        out = {"segs": segs, "rhs": rhs, "cos": cosv, "wf": wf, "wl": wl, "ja": ja, "jb": jb,
               "js": js, "ground": (g.kind, g.eps_r, g.sigma), "n": n, "route": 0}
        if semantics == "cli":
            L.or_conductor_paths.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
            L.or_build_hallen_rhs_paths.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_void_p,
                                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
            path_of = np.zeros(n, np.int32)
            sign_of = np.zeros(n)
            s_of = np.zeros(n)
            fa = np.zeros(n + 1, np.int32)
            fb = np.zeros(n + 1, np.int32)
            nontriv = C.c_int(0)
            npaths = L.or_conductor_paths(_p(segs), n, _p(path_of), _p(sign_of), _p(s_of), _p(fa), _p(fb),
                                          n + 1, C.cast(C.pointer(nontriv), C.c_void_p))
            if npaths > 0 and nontriv.value:
                rhs2 = np.zeros(n, np.complex128)
                cos2 = np.zeros(n)
                st = L.or_build_hallen_rhs_paths(d, _p(segs), n, float(freq_hz), _p(path_of), _p(sign_of),
                                                  _p(s_of), _p(rhs2), _p(cos2))
                if st:
                    raise RuntimeError(f"excitation error {st}")
                fe = np.empty(2 * npaths, np.int32)
                fe[0::2] = fa[:npaths]
                fe[1::2] = fb[:npaths]
                out.update(route=1, rhs=rhs2, cos=cos2, path_of=path_of, free_ends=fe)
        return out
    finally:
        L.or_deck_free(d)


def variant_inputs(text: str, freq_hz: float) -> dict | None:
    """The CLI's plane-wave or current-source inputs for a deck
    (or_deck_variant_inputs): kind, route, rhs (the forcing or the unit
    source shape), cos, sin, src, i0, and the wire ranges or the paths."""
    L = lib()
    n_max = 1 << 16
    n, route, src, nw, nfree = (C.c_int() for _ in range(5))
    i0 = CX()
    rhs = np.zeros(n_max, np.complex128)
    cosv, sinv = np.zeros(n_max), np.zeros(n_max)
    wf, wl, po, fe = (np.zeros(n_max + 1, np.int32) for _ in range(4))
    L.or_deck_variant_inputs.argtypes = [C.c_char_p, C.c_double] + [C.c_void_p] * 13
    k = L.or_deck_variant_inputs(text.encode(), float(freq_hz), C.byref(n), C.byref(route), _p(rhs), _p(cosv),
                                 _p(sinv), C.byref(src), C.byref(i0), _p(wf), _p(wl), C.byref(nw), _p(po), _p(fe),
                                 C.byref(nfree))
    if k < 0:
        raise RuntimeError(f"oracle status {-k}")
    if k == 0:
        return None
    m = n.value
    return {"kind": ("plane-wave", "current-source")[k - 1], "route": route.value, "n": m,
            "rhs": rhs[:m].copy(), "cos": cosv[:m].copy(), "sin": sinv[:m].copy(), "src": src.value,
            "i0": complex(i0.re, i0.im), "wf": wf[:nw.value].copy(), "wl": wl[:nw.value].copy(),
            "path_of": po[:m].copy(), "free_ends": fe[:nfree.value].copy()}


def stamps(text: str, freq_hz: float, semantics: str = "worker") -> list:
    """The LD / TL (and for the CLI NT) stamps on Z of a deck at freq_hz, in
    application order: [(row, col, delta)] (or_deck_stamps)."""
    L = lib()
    sem = {"cli": 0, "worker": 1, "py": 2}[semantics]
    L.or_deck_stamps.argtypes = [C.c_char_p, C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_int]
    cnt = L.or_deck_stamps(text.encode(), sem, float(freq_hz), None, None, None, None, 0)
    if cnt < 0:
        raise RuntimeError(f"deck error {-cnt}")
    rows = np.zeros(max(cnt, 1), np.int32)
    cols = np.zeros(max(cnt, 1), np.int32)
    re = np.zeros(max(cnt, 1))
    im = np.zeros(max(cnt, 1))
    L.or_deck_stamps(text.encode(), sem, float(freq_hz), _p(rows), _p(cols), _p(re), _p(im), cnt)
    return [(int(rows[i]), int(cols[i]), complex(re[i], im[i])) for i in range(cnt)]
