"""Autotuner of the batched engine's knobs (fnec_hip_set_tunable).

The knobs pick tile and gang shapes of the mixed-precision path: the row tile
of the FP32 trailing-matrix update, the gang size of the CG solves, the size
limit of the one-workgroup solve, and the CG iterations queued before the
first convergence read-back. autotune() times one batched solve of a
representative workload per candidate value, searching one knob at a time
(coordinate search, repeated while it improves), and keeps the fastest
setting whose results stay within the accuracy contract of the starting one.
"""
from __future__ import annotations

import ctypes as C  # noqa: N812
import time

import numpy as np

SPACE = {
    "update_tile_m": (128, 64),
    "update_k128": (0, 1),
    "gang_n": (0, 384, 512, 768, 1024, 1536),
    "gang_max": (2, 4, 8),
    "solve1_max_n": (1024, 2048, 4096, 8192),
    "mix_first_chunk": (2, 3, 4, 5, 6),
}


def search(space: dict, start: dict, measure, rounds: int = 2, min_gain: float = 0.0):
    """Coordinate search. measure(knobs) -> (seconds, ok); a candidate with
    ok False is never chosen, and a candidate replaces the incumbent only when
    it is faster by more than the fraction min_gain (timing noise), so that
    knobs do not drift on noise. Returns (best knobs, best seconds, trials)
    where trials lists (knobs, seconds, ok) in the order measured."""
    best = dict(start)
    t_best, ok = measure(best)
    if not ok:
        raise RuntimeError("the starting setting fails its own check")
    trials = [(dict(best), t_best, True)]
    seen = {tuple(sorted(best.items())): t_best}
    for _ in range(rounds):
        improved = False
        for name, values in space.items():
            for v in values:
                cand = dict(best)
                cand[name] = v
                key = tuple(sorted(cand.items()))
                if key in seen:
                    continue
                t, ok = measure(cand)
                seen[key] = t if ok else float("inf")
                trials.append((dict(cand), t, ok))
                if ok and t < t_best * (1.0 - min_gain):
                    best, t_best, improved = cand, t, True
        if not improved:
            break
    return best, t_best, trials


def autotune(problems: list, precision: int | None = None, space: dict | None = None, repeats: int = 2,
             rounds: int = 2, tol: float = 1e-6, apply: bool = True, min_gain: float = 0.02) -> dict:
    """Tune the engine's knobs on `problems` (Deck.problem dicts, or prebuilt
    fnec_request structures, solved as one batch). Each candidate is timed as
    the best of `repeats` solves (the C call only: the requests are built
    once). A candidate is rejected when any
    problem's Z_in moves by more than `tol` (relative) from the starting
    setting's, or any status is not OK, and must beat the incumbent by more
    than min_gain (2% by default: one batch's timing varies by about 1%
    run to run). With apply, the best knobs are left
    set; otherwise the starting ones are restored. Returns {"knobs", "seconds",
    "start_seconds", "trials"}."""
    import importlib

    F = importlib.import_module("fnec-rust_amd")
    if precision is None:
        precision = F.PREC_MIXED
    space = SPACE if space is None else space
    L = F.lib()
    reqs = (F.Request * len(problems))()
    res = (F.Result * len(problems))()
    keep = []
    for i, p in enumerate(problems):
        if isinstance(p, F.Request):  # prebuilt (the caller keeps its arrays alive)
            reqs[i] = p
            continue
        q, k = F.make_request(p, i)
        reqs[i] = q
        keep.append(k)
    start = {k: F.get_tunable(k) for k in space}

    def run():
        t0 = time.perf_counter()
        st = L.fnec_hip_solve_batch(C.cast(reqs, C.c_void_p), len(problems), precision, C.cast(res, C.c_void_p))
        t = time.perf_counter() - t0
        F._check(st)
        ok = all(res[i].status == F.OK for i in range(len(problems)))
        z = np.array([complex(res[i].z_in.re, res[i].z_in.im) for i in range(len(problems))])
        return t, ok, z

    run()  # warm the arenas
    z_ref = None

    def measure(knobs):
        nonlocal z_ref
        for k, v in knobs.items():
            F.set_tunable(k, v)
        best_t, ok_all, z = float("inf"), True, None
        for _ in range(max(1, repeats)):
            t, ok, z = run()
            best_t = min(best_t, t)
            ok_all = ok_all and ok
        if z_ref is None:
            z_ref = z
        elif ok_all:
            ok_all = bool(np.all(np.abs(z - z_ref) <= tol * np.abs(z_ref)))
        return best_t, ok_all

    try:
        best, t_best, trials = search(space, start, measure, rounds, min_gain)
    except Exception:
        for k, v in start.items():
            F.set_tunable(k, v)
        raise
    final = best if apply else start
    for k, v in final.items():
        F.set_tunable(k, v)
    return {"knobs": best, "seconds": t_best, "start_seconds": trials[0][1],
            "trials": [{"knobs": k, "seconds": t, "ok": o} for k, t, o in trials]}
