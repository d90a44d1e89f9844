#!/usr/bin/env python3
"""Serving benchmark of the Hallén path on one MI355X.

Workload: the mixed-N request stream (20k requests in full; N log-uniform in
[51, 4096]; random parallel / junctioned wire sets; 30% with a 2x2 degree RP
grid; 10% over PEC ground), built by fnec-rust_amd/workloads.py. A "step" is
one window of requests submitted to the serving session and drained;
``--steps K`` windows are timed after ``--warmup W`` untimed ones.

Printed (one JSON line):
- served solves/s (value) and the closed-loop window latency p50/p99;
- open-loop latency: p50/p99 with the stream's own Poisson arrival times at a
  stated load (a separate slice of the stream);
- the plain FP64 path (the reference's algorithm and rounding) one request at
  a time over exactly the requests of the timed region (the reference point
  of the speedup), the speedup over it, and, for context, the FP64
  matrix-core path on the first 256 of them;
- accuracy: the max relative Z_in difference of the served results from the
  plain path over every timed request, and from the oracle (the C
  restatement of the reference CPU path) over the timed requests the CPU
  baseline solves;
- the roofline of the dominant kernel from HIP events recorded on its stream
  during the timed region (events around that kernel class only; the
  per-class breakdown comes from untimed windows profiled before it);
- a CPU baseline: the oracle on this host's cores on a bounded sample of the
  timed requests, extrapolated over the stream's N distribution.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import ctypes as C  # noqa: N812
import importlib
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

F = importlib.import_module("fnec-rust_amd")
W = importlib.import_module("fnec-rust_amd.workloads")
AT = importlib.import_module("fnec-rust_amd.tuning")

METRIC = "solves/s + p99 latency, mixed-N queue, vs plain FP64 GPU at rel ΔZin≤1e-6"
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
PEAK_F16_TFLOPS = 2500.0
# lu_update_k128 runs the FP32 algorithm on the f16 matrix cores, three f16
# products per FP32 product: its algorithmic peak is 2500 / 3 = 833 TFLOP/s,
# a ridge of 104 flops per HBM byte, above its 64 at K = 128, so HBM bounds it
HBM_BOUND = ("zsolve", "matvec", "backsub", "lu_update_k128")

RP_THETA = np.repeat(np.arange(0.0, 181.0, 2.0)[None, :], 180, axis=0).ravel()  # phi outer, theta inner
RP_PHI = np.repeat(np.arange(0.0, 360.0, 2.0), 91)


def build_requests(stream):
    """Host side: decks -> Hallén inputs (worker semantics) -> fnec_request,
    each with buffers for its currents and, with a grid, its pattern rows, so
    that a served request returns what a worker reply carries."""
    items = []
    for r in stream:
        d = F.Deck(r["deck"])
        p = d.problem("worker", r["freq_hz"])
        d.close()
        q, keep = F.make_request(p, tag=r["id"], rp_theta=RP_THETA if r["rp"] else None,
                                 rp_phi=RP_PHI if r["rp"] else None, outputs=True)
        items.append({"req": q, "keep": keep, "meta": r, "p": p})
    return items


# host clock readings (ns) at the start and end of the closed-loop timed
# region, so that a kernel trace of the same run can be cut to that region
TIMED_REGION = {}


def host_clocks():
    return {"monotonic": time.clock_gettime_ns(time.CLOCK_MONOTONIC),
            "boottime": time.clock_gettime_ns(time.CLOCK_BOOTTIME),
            "monotonic_raw": time.clock_gettime_ns(time.CLOCK_MONOTONIC_RAW)}


def serve(items, window, steps, warmup, prof_windows, precision, budget, max_batch, depth, timed_class=None):
    """Warmup windows, then ``prof_windows`` untimed windows with events on
    every kernel class (the per-class breakdown; it names the dominant class),
    then the timed windows with events around the dominant class only. In each
    phase a window of requests is submitted whenever fewer than ``depth``
    windows are outstanding, so new requests arrive while earlier ones are on
    the device."""
    L = F.lib()
    sess = L.fnec_hip_session_create(budget, precision, max_batch)
    if not sess:
        raise F.FnecError(F.INVALID_ARG, "session create failed")
    out = (F.Result * 4096)()
    results = {}
    batches = {}

    def run_phase(first, nwin):
        t_sub = {}
        lat = []
        bt = {}

        def submit(w):
            for it in items[(first + w) * window:(first + w + 1) * window]:
                t_sub[it["req"].tag] = time.perf_counter()
                st = L.fnec_hip_session_submit(sess, C.byref(it["req"]))
                if st:
                    F._check(st)

        nxt = outstanding = 0
        while nxt < min(depth, nwin):
            submit(nxt)
            nxt += 1
            outstanding += window
        while outstanding:
            k = L.fnec_hip_session_poll(sess, C.cast(out, C.c_void_p), 4096, 2000)
            if k < 0:
                F._check(-k)
            now = time.perf_counter()
            for i in range(k):
                r = out[i]
                if r.status:
                    raise F.FnecError(r.status, f"request {r.tag} failed")
                results[r.tag] = complex(r.z_in.re, r.z_in.im)
                lat.append(now - t_sub[r.tag])
                b = bt.setdefault(r.batch, [r.t_start_us, r.t_done_us, 0])
                b[2] += 1
            outstanding -= k
            while nxt < nwin and outstanding <= (depth - 1) * window:
                submit(nxt)
                nxt += 1
                outstanding += window
        return lat, bt

    run_phase(0, warmup)
    F.profile(enable=True, reset=True, only="")
    run_phase(warmup, prof_windows)
    prof_all = F.profile(enable=False)
    timed_cls = [k for k in prof_all if prof_all[k][0] > 0 and prof_all[k][2] > 0]
    dom = max(timed_cls, key=lambda k: prof_all[k][1]) if timed_cls else "lu_update"
    if timed_class:
        dom = timed_class
    F.profile(enable=True, reset=True, only=dom)
    clocks0 = host_clocks()
    t0 = time.perf_counter()
    lat, batches = run_phase(warmup + prof_windows, steps)
    elapsed = time.perf_counter() - t0
    TIMED_REGION.update(start=clocks0, end=host_clocks())
    prof_timed = F.profile(enable=False, only="")
    L.fnec_hip_session_destroy(sess)
    return elapsed, np.array(lat), results, dom, prof_all, prof_timed, batches


def serve_poisson(items, times, warmup_items, precision, budget, max_batch, prof_class):
    """Open loop: request i is submitted at its arrival time times[i] (s from
    the start), whatever is outstanding; latency runs from that arrival time
    to the poll that returns the result. ``warmup_items`` requests are served
    first (closed loop) and not timed. Events are recorded around
    ``prof_class`` only."""
    L = F.lib()
    sess = L.fnec_hip_session_create(budget, precision, max_batch)
    if not sess:
        raise F.FnecError(F.INVALID_ARG, "session create failed")
    out = (F.Result * 4096)()
    results = {}
    lat_by_tag = {}

    def drain(k_expect, t_arr, lat, bt, wait_us):
        k = L.fnec_hip_session_poll(sess, C.cast(out, C.c_void_p), 4096, wait_us)
        if k < 0:
            F._check(-k)
        now = time.perf_counter()
        for i in range(k):
            r = out[i]
            if r.status:
                raise F.FnecError(r.status, f"request {r.tag} failed")
            results[r.tag] = complex(r.z_in.re, r.z_in.im)
            if t_arr is not None:
                lat.append(now - t_arr[r.tag])
                lat_by_tag[r.tag] = now - t_arr[r.tag]
                b = bt.setdefault(r.batch, [r.t_start_us, r.t_done_us, 0])
                b[2] += 1
        return k

    left = 0
    for it in warmup_items:
        F._check(L.fnec_hip_session_submit(sess, C.byref(it["req"])))
        left += 1
    while left:
        left -= drain(left, None, None, None, 2000)
    F.profile(enable=True, reset=True, only=prof_class)
    lat, bt, t_arr = [], {}, {}
    nxt = done = 0
    t0 = time.perf_counter()
    while done < len(items):
        now = time.perf_counter() - t0
        while nxt < len(items) and times[nxt] <= now:
            it = items[nxt]
            t_arr[it["req"].tag] = t0 + times[nxt]
            F._check(L.fnec_hip_session_submit(sess, C.byref(it["req"])))
            nxt += 1
        wait = 2000 if nxt >= len(items) else max(0, min(2000, int((times[nxt] - now) * 1e6)))
        done += drain(None, t_arr, lat, bt, wait)
    elapsed = time.perf_counter() - t0
    prof = F.profile(enable=False, only="")
    L.fnec_hip_session_destroy(sess)
    LAT_BY_TAG.clear()
    LAT_BY_TAG.update(lat_by_tag)
    return elapsed, np.array(lat), results, prof, bt


# latency (s) of each request of the last open-loop phase, by tag
LAT_BY_TAG = {}


def batch_overlap(batches):
    """Batches formed in the timed region, their mean size, and the share of
    the span from the first batch start to the last batch end during which
    two or more batches were on the device."""
    if not batches:
        return None
    ev = sorted([(v[0], 1) for v in batches.values()] + [(v[1], -1) for v in batches.values()])
    span = ev[-1][0] - ev[0][0]
    busy2 = 0.0
    level = 0
    for (t, d), (t_next, _) in zip(ev, ev[1:] + [(ev[-1][0], 0)]):
        level += d
        if level >= 2:
            busy2 += t_next - t
    sizes = [v[2] for v in batches.values()]
    return {"batches": len(batches), "mean_batch": round(float(np.mean(sizes)), 1),
            "two_or_more_on_device_frac": round(busy2 / span, 3) if span > 0 else 0.0}


def plain_fp64(items, precision=F.PREC_FP64, outputs=None):
    """One request at a time through fnec_hip_solve_hallen (+ fnec_hip_rp).
    With a dict ``outputs``, the served currents and pattern rows of each item
    (its request's output buffers) are compared with this path's, outside the
    timed loop's arithmetic: worst relative current difference (2-norm) and
    worst linear-gain difference relative to the pattern's maximum."""
    zs = {}
    t_cmp = 0.0
    t0 = time.perf_counter()
    for it in items:
        k = it["keep"]
        r = it["req"]
        x = F.solve_hallen(k["segs"], k["rhs"], k["cos"], k["wf"], k["wl"], k["ja"], k["jb"], k["js"],
                           r.freq_hz, r.ground, precision)
        n = len(k["segs"])
        zs[r.tag] = complex(r.v_feed.re, r.v_feed.im) / x[r.feed_index]
        rows = None
        if it["meta"]["rp"]:
            rows, _ = F.rp(k["segs"], x[:n], r.freq_hz, RP_THETA, RP_PHI, r.ground)
        if outputs is not None:
            tc = time.perf_counter()
            dx = float(np.linalg.norm(k["x"][:n] - x[:n]) / np.linalg.norm(x[:n]))
            outputs["max_rel_dx"] = max(outputs.get("max_rel_dx", 0.0), dx)
            if rows is not None:
                live = rows[:, 0] > -999.0
                ga, gb = 10.0 ** (k["rows"][live, 0] / 10.0), 10.0 ** (rows[live, 0] / 10.0)
                dg = float(np.max(np.abs(ga - gb)) / np.max(gb))
                outputs["max_gain_diff"] = max(outputs.get("max_gain_diff", 0.0), dg)
                outputs["patterns"] = outputs.get("patterns", 0) + 1
            t_cmp += time.perf_counter() - tc
    return time.perf_counter() - t0 - t_cmp, zs


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(stream, timed, budget_s, threads):
    """Oracle (C restatement of the reference CPU path, one thread per solve
    as the reference) on a bounded sample of the timed requests with N <= 600,
    spread over ``threads`` host threads. Returns the baseline record and the
    oracle's Z_in of every sampled request (by tag)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # noqa: E402
    O.build()
    rng = np.random.default_rng(7)
    small = [r for r in timed if r["n"] <= 600]
    order = rng.permutation(len(small))

    def one(r):
        t = time.perf_counter()
        res = O.run_deck(r["deck"], "worker", 0, r["freq_hz"])
        if r["rp"]:
            kind = 1 if r["pec"] else 0
            O.radiation_pattern(res["segs"], res["currents"], r["freq_hz"], RP_THETA, RP_PHI,
                                O.Ground(kind, 0.0, 0.0))
        return r["n"], r["rp"], time.perf_counter() - t, r["id"], complex(res["feed_z"][0])

    samples = []
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        pending = set()
        it = iter(order)
        for _ in range(min(threads, len(small))):
            pending.add(ex.submit(one, small[next(it)]))
        while pending:
            done, pending = cf.wait(pending, return_when=cf.FIRST_COMPLETED)
            for f in done:
                samples.append(f.result())
                if time.perf_counter() - t0 < budget_s:
                    nxt = next(it, None)
                    if nxt is not None:
                        pending.add(ex.submit(one, small[nxt]))
    wall = time.perf_counter() - t0
    ns = np.array([s[0] for s in samples], float)
    rps = np.array([s[1] for s in samples], float)
    ts = np.array([s[2] for s in samples])
    # per-request single-thread cost model: t = a N^3 + b N * rp (RP terms ~ N)
    A = np.stack([ns ** 3, ns * rps], axis=1)
    coef, *_ = np.linalg.lstsq(A, ts, rcond=None)
    coef = np.maximum(coef, 0.0)
    N = np.array([r["n"] for r in stream], float)
    R = np.array([r["rp"] for r in stream], float)
    mean_t = float(np.mean(coef[0] * N ** 3 + coef[1] * N * R))
    value = threads / mean_t
    return {
        "value": value,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{len(samples)} of the timed requests with N<=600 solved by the oracle (worker semantics + RP), "
                   f"one thread per solve, {threads} solves at a time, in {wall:.1f} s; per-request cost fitted "
                   f"as a*N^3 + b*N*[rp] on N {int(ns.min())}..{int(ns.max())} and extrapolated over all "
                   f"{len(stream)} requests of the stream (N up to {int(N.max())})"),
        "cpu": cpu_model(),
    }, {s[3]: s[4] for s in samples}


def measured_traffic(kernel, precision):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE on this workload (scaled by the copy-probe calibration), from
    the committed summary in profiles/; None when no measurement matches."""
    path = os.path.join(ROOT, "profiles", "serving_stream_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    ring = kernel == "lu_update_k128" and F.get_tunable("update_ring") == 1
    e = d.get("kernels", {}).get(kernel + ("_ring" if ring else ""))
    if not e or e.get("precision") != ("mixed" if precision == F.PREC_MIXED else "fp64"):
        return None
    tm = F.get_tunable("update_tile_m")
    if kernel == "lu_update" and e.get("symbol") != f"k_lu_update_f32<{tm}, 64, {'true' if tm == 64 else 'false'}>":
        return None  # measured with the other tile and product form
    if kernel == "lu_update" and tm == 64 and F.get_tunable("update_k128") == 1:
        return None  # measured with one K = 64 pass per step
    if kernel == "lu_update_k128" and e.get("symbol") != ("k_lu_update_f16x2s" if ring else "k_lu_update_f16x2d"):
        return None  # measured with another form of the K = 128 pass
    return {"bytes_per_launch": e["hbm_bytes_per_launch"], "source": "profiles/serving_stream_traffic.json"}


def copy_probe():
    """A plain device-to-device copy of 256 MiB, five times: reads and writes
    exactly 256 MiB per dispatch, the reference size for scaling FETCH_SIZE /
    WRITE_SIZE counter readings into HBM bytes."""
    import torch

    x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": x.numel() * 4, "dispatches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--window", type=int, default=256, help="requests per step")
    ap.add_argument("--profile-windows", type=int, default=2,
                    help="untimed windows with events on every kernel class, before the timed ones")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--plain-sample", type=int, default=0,
                    help="timed requests the plain FP64 path solves (0: all of them)")
    ap.add_argument("--latency-rate", type=float, default=2000.0,
                    help="load of the open-loop latency phase (requests/s; 0 skips it)")
    ap.add_argument("--latency-requests", type=int, default=2048)
    ap.add_argument("--max-batch", type=int, default=256)
    ap.add_argument("--lanes", type=int, default=2, help="batches on the device at once")
    ap.add_argument("--arrival", choices=["windows", "poisson"], default="windows",
                    help="windows: closed loop at saturation (the throughput figure); poisson: open loop, "
                         "each request submitted at its arrival time (latency at a given load)")
    ap.add_argument("--rate", type=float, default=2000.0, help="requests/s of the poisson arrival mode")
    ap.add_argument("--mix-tol-exp", type=int, default=None,
                    help="CG stopping rule of the mixed path: relative residual 10^-exp (default 14)")
    ap.add_argument("--batch-split-n", type=int, default=None,
                    help="serving batches take one N class, split at this size (0: one class)")
    ap.add_argument("--profile-class", default=None,
                    help="kernel class timed in the timed windows (default: the dominant one)")
    ap.add_argument("--update-k128", type=int, choices=[0, 1], default=None,
                    help="K = 128 trailing-update passes of the FP32 LU (default 1; with --no-autotune)")
    ap.add_argument("--update-ring", type=int, choices=[0, 1], default=None,
                    help="K = 128 passes stream their operands through an LDS ring (1) or load them into "
                         "registers (0, the default)")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="set an engine tunable after autotuning (repeatable)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the engine's default knobs (no autotune on a window of the stream)")
    ap.add_argument("--depth", type=int, default=None,
                    help="windows outstanding at once (default: lanes)")
    ap.add_argument("--budget-gib", type=float, default=96.0)
    ap.add_argument("--cpu-budget-s", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU; multi-GPU runs are not supported "
              "(requests shard across GPUs with no device communication)", file=sys.stderr)
        sys.exit(2)
    F.init()
    F.set_tunable("lanes", args.lanes)
    if args.mix_tol_exp is not None:
        F.set_tunable("mix_tol_exp", args.mix_tol_exp)
    if args.batch_split_n is not None:
        F.set_tunable("batch_split_n", args.batch_split_n)
    if args.update_k128 is not None:
        F.set_tunable("update_k128", args.update_k128)
    if args.update_ring is not None:
        F.set_tunable("update_ring", args.update_ring)
    depth = args.depth or args.lanes
    precision = F.PREC_MIXED if args.precision == "mixed" else F.PREC_FP64
    total = (args.warmup + args.profile_windows + args.steps) * args.window
    n_lat = args.latency_requests if args.arrival == "windows" and args.latency_rate > 0 else 0
    t_gen = time.perf_counter()
    stream = W.config5_stream(count=total + n_lat)
    items = build_requests(stream)
    t_gen = time.perf_counter() - t_gen
    print(f"[bench] built {len(items)} requests in {t_gen:.1f}s; serving {args.steps} x {args.window}",
          file=sys.stderr, flush=True)

    tuned = None
    if not args.no_autotune:
        t_at = time.perf_counter()
        at = F.autotune([it["req"] for it in items[:args.window]], precision, repeats=1, rounds=1)
        tuned = {"knobs": at["knobs"], "window_ms": round(at["seconds"] * 1e3, 2),
                 "start_window_ms": round(at["start_seconds"] * 1e3, 2), "trials": len(at["trials"]),
                 "tuning_s": round(time.perf_counter() - t_at, 2)}
        print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)
    for kv in args.set:
        name, value = kv.split("=", 1)
        F.set_tunable(name, int(value))
    if args.arrival == "poisson":
        # open loop: the stream's Poisson arrival times, rescaled to the rate
        first = args.warmup * args.window
        timed_items = items[first:first + args.steps * args.window]
        t_first = stream[first]["t_arrival"]
        scale = 1000.0 / args.rate  # the stream is generated at 1000 requests/s
        times = [(stream[first + i]["t_arrival"] - t_first) * scale for i in range(len(timed_items))]
        k128 = F.get_tunable("update_k128") == 1 and F.get_tunable("update_tile_m") == 64
        dom = args.profile_class or ("lu_update_k128" if k128 else "lu_update")
        elapsed, lat, zs_served, prof, batches = serve_poisson(
            timed_items, times, items[:first], precision, int(args.budget_gib * 2 ** 30), args.max_batch, dom)
        prof_all = prof
        timed = len(timed_items)
        arrival = f"Poisson at {args.rate:g} requests/s, open loop (each request submitted at its arrival time)"
    else:
        elapsed, lat, zs_served, dom, prof_all, prof, batches = serve(
            items, args.window, args.steps, args.warmup, args.profile_windows, precision,
            int(args.budget_gib * 2 ** 30), args.max_batch, depth, args.profile_class)
        first = (args.warmup + args.profile_windows) * args.window
        timed = args.steps * args.window
        arrival = (f"windows of {args.window} requests; a window is submitted whenever fewer than {depth} "
                   f"are outstanding")
    value = timed / elapsed
    print(f"[bench] served {timed} requests in {elapsed:.2f}s", file=sys.stderr, flush=True)

    # open-loop latency at a stated load, on a slice of the stream after the
    # closed-loop requests (served through a fresh session)
    open_loop = None
    if n_lat:
        lat_items = items[total:total + n_lat]
        t_first = stream[total]["t_arrival"]
        scale = 1000.0 / args.latency_rate  # the stream is generated at 1000 requests/s
        times = [(stream[total + i]["t_arrival"] - t_first) * scale for i in range(n_lat)]
        el_ol, lat_ol, _, _, bt_ol = serve_poisson(lat_items, times, items[:args.window], precision,
                                                   int(args.budget_gib * 2 ** 30), args.max_batch, dom)
        open_loop = {"offered_per_s": args.latency_rate, "served_per_s": round(n_lat / el_ol, 1),
                     "requests": n_lat, "p50": round(float(np.percentile(lat_ol, 50)) * 1e3, 3),
                     "p99": round(float(np.percentile(lat_ol, 99)) * 1e3, 3),
                     "arrival": "the stream's Poisson arrival times rescaled to the load, open loop; "
                                "latency from arrival to the poll that returns the result",
                     **(batch_overlap(bt_ol) or {})}
        split = F.get_tunable("batch_split_n")
        for name, sel in (("n_small", lambda n: n <= split), ("n_large", lambda n: n > split)):
            ls = [LAT_BY_TAG[it["req"].tag] for it in lat_items if sel(it["meta"]["n"]) and it["req"].tag in LAT_BY_TAG]
            if ls:
                open_loop[name] = {"n_max" if name == "n_small" else "n_min": split if name == "n_small" else split + 1,
                                   "requests": len(ls), "p50": round(float(np.percentile(ls, 50)) * 1e3, 3),
                                   "p99": round(float(np.percentile(ls, 99)) * 1e3, 3)}
        print(f"[bench] open loop at {args.latency_rate:g}/s: p50 {open_loop['p50']:.1f} ms, "
              f"p99 {open_loop['p99']:.1f} ms", file=sys.stderr, flush=True)

    timed_items = items[first:first + timed]
    sample_items = timed_items[:args.plain_sample] if args.plain_sample else timed_items
    served_out = {}
    t_plain, zs_plain = plain_fp64(sample_items, outputs=served_out)
    plain_rate = len(sample_items) / t_plain
    dz = max(abs(zs_served[k] - zs_plain[k]) / abs(zs_plain[k]) for k in zs_plain)
    print(f"[bench] plain FP64 one-at-a-time: {plain_rate:.1f} solves/s on {len(sample_items)}",
          file=sys.stderr, flush=True)
    mfma_items = timed_items[:256]
    t_mfma, zs_mfma = plain_fp64(mfma_items, F.PREC_FP64_MFMA)
    mfma_rate = len(mfma_items) / t_mfma
    print(f"[bench] FP64 matrix-core path one-at-a-time: {mfma_rate:.1f} solves/s on {len(mfma_items)}",
          file=sys.stderr, flush=True)

    # dominant kernel class by time in the profiled pass; its events in the
    # timed region give the achieved rate, and its bound is the roofline side
    # its algorithmic intensity falls on (flops per HBM byte vs the ridge
    # peak_flops / peak_bandwidth)
    launches, tot_us, flops, nbytes = prof[dom]
    avg_us = tot_us / launches
    mfma = dom in ("gram", "lu_update")
    peak_f = PEAK_FP32_TFLOPS if precision == F.PREC_MIXED and dom == "lu_update" else PEAK_FP64_TFLOPS
    if dom == "lu_update_k128":
        peak_f = PEAK_F16_TFLOPS / 3.0
    ridge = peak_f * 1e12 / (PEAK_HBM_GBS * 1e9)
    if dom in HBM_BOUND or (nbytes and flops / nbytes < ridge):
        achieved = nbytes / launches / (avg_us * 1e-6) / 1e9
        peak, unit, bound = PEAK_HBM_GBS, "GB/s", "hbm"
    else:
        achieved = flops / launches / (avg_us * 1e-6) / 1e12
        peak, unit, bound = peak_f, "TFLOP/s", "mfma" if mfma else "valu"
    traffic = measured_traffic(dom, precision)
    roofline = {
        "kernel": dom,
        "bound": bound,
        "achieved": round(achieved, 3),
        "peak": peak,
        "unit": unit,
        "frac": round(achieved / peak, 4),
        "traffic": traffic["bytes_per_launch"] if traffic else None,
        "traffic_source": traffic["source"] if traffic else None,
        "intensity_flop_per_byte": round(flops / nbytes, 2) if nbytes else None,
        "ridge_flop_per_byte": round(ridge, 2),
        "avg_launch_us": round(avg_us, 2),
        "launches": int(launches),
        "algorithmic_bytes_per_launch": round(nbytes / launches) if nbytes else None,
        "kernels_source": (f"{args.profile_windows} untimed windows of {args.window} with events on every class"
                           if args.arrival == "windows" else f"timed region, events on {dom} only"),
        "timed_region_clock_ns": TIMED_REGION or None,
        "kernels": {k: {"launches": int(v[0]), "total_ms": round(v[1] / 1e3, 3),
                        "tflops": round(v[2] / (v[1] * 1e-6) / 1e12, 3) if v[1] and v[2] else None,
                        "gbs": round(v[3] / (v[1] * 1e-6) / 1e9, 1) if v[1] and v[3] else None,
                        "bytes_per_launch": round(v[3] / v[0]) if v[3] else None}
                    for k, v in prof_all.items() if v[0] > 0},
    }
    cpu, z_oracle = None, {}
    if not args.no_cpu_baseline:
        # the host's cores as this process sees them, at most the 16 host
        # cores a one-GPU share of the GPU machine provides
        threads = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count() or 1)
        cpu, z_oracle = cpu_baseline(stream, [stream[first + i] for i in range(timed)], args.cpu_budget_s,
                                     threads)
    acc = {"vs_plain_fp64": {"max_rel_dzin": dz, "requests": len(zs_plain),
                             "max_rel_dcurrents": served_out.get("max_rel_dx"),
                             "max_gain_diff": served_out.get("max_gain_diff"),
                             "patterns": served_out.get("patterns", 0)}}
    if z_oracle:
        acc["vs_oracle"] = {
            "max_rel_dzin_served": max(abs(zs_served[k] - z) / abs(z) for k, z in z_oracle.items()),
            "max_rel_dzin_plain_fp64": max(abs(zs_plain[k] - z) / abs(z) for k, z in z_oracle.items()
                                           if k in zs_plain),
            "requests": len(z_oracle)}

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if precision == F.PREC_FP64 else "c64+c128",
        "data": "synthetic",
        "config": {
            "workload": "mixed-N serving stream (N log-U[51,4096], 30% 2x2 deg RP, 10% PEC), "
                        f"{timed} timed requests, " + (f"Poisson arrivals at {args.rate:g}/s" if args.arrival == "poisson"
                                                       else f"in windows of {args.window}"),
            "precision": args.precision,
            "max_batch": args.max_batch,
            "tunables": tuned["knobs"] if tuned else {k: F.get_tunable(k) for k in AT.SPACE},
        },
        "autotune": tuned,
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
                       "arrival": arrival},
        "serving": dict(lanes=args.lanes, batch_split_n=F.get_tunable("batch_split_n"),
                        **(batch_overlap(batches) or {})),
        "open_loop_latency_ms": open_loop,
        "plain_fp64_gpu": {"value": round(plain_rate, 3), "unit": "solves/s",
                           "requests": len(sample_items),
                           "what": "fnec_hip_solve_hallen(FNEC_PREC_FP64) + fnec_hip_rp, one request at a "
                                   "time, over the timed requests: the reference's algorithm and rounding"},
        "speedup_vs_plain_fp64": round(value / plain_rate, 3),
        "fp64_matrix_core_gpu": {"value": round(mfma_rate, 3), "unit": "solves/s", "requests": len(mfma_items),
                                 "speedup": round(value / mfma_rate, 3),
                                 "max_rel_dzin_vs_plain_fp64": max((abs(zs_mfma[k] - zs_plain[k]) / abs(zs_plain[k])
                                                                    for k in zs_mfma if k in zs_plain), default=None)},
        "accuracy": acc,
        "max_rel_dzin_vs_plain_fp64": dz,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
