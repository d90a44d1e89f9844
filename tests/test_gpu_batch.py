"""GPU tests of batching and the serving session through the C ABI.

- a batch solves every problem exactly as a single call does (bitwise, both
  precisions): each kernel walks a work list of (problem, tile) items and no
  reduction crosses problems;
- far-field grids of a batch read the currents in place on the device and
  match the stand-alone pattern call bitwise, also when the grids are large
  next to the solves (many small problems with 2 x 2 degree grids);
- the session returns every request once, with the same result as a batch,
  whatever the admission splits (max_batch, HBM budget), and rejects a request
  whose footprint exceeds its budget;
- requests that arrive while a batch is on the device run beside it on a
  second lane, and each result is bitwise the result of the request alone.
"""
import ctypes as C  # noqa: N812
import importlib
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F = importlib.import_module("fnec-rust_amd")
W = importlib.import_module("fnec-rust_amd.workloads")

TH2 = np.repeat(np.arange(0.0, 181.0, 2.0)[None, :], 180, axis=0).ravel()
PH2 = np.repeat(np.arange(0.0, 360.0, 2.0), 91)


@pytest.fixture(scope="module", autouse=True)
def device():
    F.init()


def problems(seed, sizes, pec_every=3):
    rng = np.random.default_rng(seed)
    out = []
    for i, n in enumerate(sizes):
        txt = W.random_request_deck(rng, n, float(rng.uniform(3.0, 30.0)), pec=bool(i % pec_every == 2),
                                    junctioned=bool(i % 2))
        d = F.Deck(txt)
        out.append(d.problem("worker", d.frequencies("worker")[0]))
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_single_bitwise(precision):
    probs = problems(21, [51, 400, 97, 1200, 64, 700])
    batch = F.solve_batch(probs, precision=precision)
    for p, b in zip(probs, batch):
        assert b["status"] == F.OK
        x = F.solve_problem(p, precision=precision)
        assert np.array_equal(x, b["x"])
        fi = int(p["feed_index"])
        assert b["z_in"] == p["v_feed"] / x[fi] or abs(b["z_in"] - p["v_feed"] / x[fi]) <= 1e-15 * abs(b["z_in"])


def test_batch_rp_large_grids_on_small_solves():
    probs = problems(22, [51 + (i % 7) for i in range(40)])
    grids = [(TH2, PH2)] * len(probs)
    batch = F.solve_batch(probs, precision=F.PREC_MIXED, rp_grids=grids)
    for p, b in zip(probs, batch):
        assert b["status"] == F.OK
        n = len(p["segs"])
        rows, total = F.rp(p["segs"], b["x"][:n], p["freq_hz"], TH2, PH2, p["ground_struct"])
        assert np.array_equal(rows, b["rp"])
        assert total == b["rp_total"]


def run_session(probs, precision, budget, max_batch, order=None):
    L = F.lib()
    sess = L.fnec_hip_session_create(budget, precision, max_batch)
    assert sess
    keep = []
    try:
        idx = order if order is not None else range(len(probs))
        for i in idx:
            q, k = F.make_request(probs[i], tag=i)
            keep.append((q, k))
            st = L.fnec_hip_session_submit(sess, C.byref(q))
            if st:
                return st, None
        out = (F.Result * 64)()
        got = {}
        deadline = time.monotonic() + 120.0
        while len(got) < len(probs) and time.monotonic() < deadline:
            k = L.fnec_hip_session_poll(sess, C.cast(out, C.c_void_p), 64, 200000)
            assert k >= 0
            for j in range(k):
                r = out[j]
                assert r.tag not in got
                got[r.tag] = (r.status, complex(r.z_in.re, r.z_in.im), r.t_submit_us, r.t_done_us)
        assert len(got) == len(probs)
        assert L.fnec_hip_session_pending(sess) == 0
        return 0, got
    finally:
        L.fnec_hip_session_destroy(sess)


@pytest.mark.parametrize("precision", [0, 1])
def test_session_equals_batch_under_any_split(precision):
    probs = problems(23, [51, 300, 77, 900, 150, 64, 500, 51, 1024, 200, 333, 90])
    batch = F.solve_batch(probs, precision=precision)
    for budget, max_batch in [(1 << 40, 256), (1 << 40, 5), (1 << 40, 1)]:
        st, got = run_session(probs, precision, budget, max_batch)
        assert st == 0
        for i, b in enumerate(batch):
            status, zin, t0, t1 = got[i]
            assert status == F.OK
            assert zin == b["z_in"]
            assert t1 >= t0
    # reversed arrival order, budget that admits a few requests at a time
    fp = max(F.lib().fnec_hip_request_footprint(len(p["segs"]), len(p["wf"]), len(p["ja"]), 0, precision)
             for p in probs)
    st, got = run_session(probs, precision, 3 * fp, 256, order=list(reversed(range(len(probs)))))
    assert st == 0
    for i, b in enumerate(batch):
        assert got[i][1] == b["z_in"]


def test_session_reuses_request_storage_across_polls():
    """A session reuses a finished request's storage for later submits: rounds
    of submit + poll with requests of different shapes (junctions or none, a
    far-field grid or none) return the batch result bitwise every time."""
    probs = problems(25, [51, 300, 77, 900, 150, 64, 500, 120])
    batch = F.solve_batch(probs, precision=1)
    L = F.lib()
    sess = L.fnec_hip_session_create(1 << 40, 1, 256)
    assert sess
    try:
        out = (F.Result * 16)()
        keep = []
        for rnd, order in enumerate([[0, 1], [2, 3, 4], [5], [6, 7, 0], [3, 1]]):
            for i in order:
                grid = (TH2, PH2) if (i + rnd) % 3 == 0 else (None, None)
                q, k = F.make_request(probs[i], tag=100 * rnd + i, rp_theta=grid[0], rp_phi=grid[1])
                keep.append((q, k))
                assert L.fnec_hip_session_submit(sess, C.byref(q)) == 0
            got = {}
            deadline = time.monotonic() + 60.0
            while len(got) < len(order) and time.monotonic() < deadline:
                k = L.fnec_hip_session_poll(sess, C.cast(out, C.c_void_p), 16, 200000)
                for j in range(k):
                    got[out[j].tag] = (out[j].status, complex(out[j].z_in.re, out[j].z_in.im))
            for i in order:
                status, zin = got[100 * rnd + i]
                assert status == F.OK
                assert zin == batch[i]["z_in"]
    finally:
        L.fnec_hip_session_destroy(sess)


@pytest.mark.parametrize("precision", [0, 1])
def test_session_overlaps_batches_and_results_match_alone(precision):
    """Continuous batching: requests submitted while a batch of large solves is
    on the device form a second batch that runs beside it on another lane (its
    start precedes the first batch's end), and every result, far-field total
    included, is bitwise the result of the same request solved alone."""
    assert F.get_tunable("lanes") >= 2
    heavy = problems(26, [4096] * 6)
    light = problems(27, [51, 300, 77, 900, 150, 64, 500, 120])
    probs = heavy + light
    grids = [(TH2, PH2) if i % 3 == 1 else None for i in range(len(probs))]
    alone = [F.solve_batch([p], precision=precision, rp_grids=[g])[0] for p, g in zip(probs, grids)]
    L = F.lib()
    sess = L.fnec_hip_session_create(1 << 40, precision, 256)
    assert sess
    try:
        reqs = [F.make_request(p, tag=i, rp_theta=g[0] if g else None, rp_phi=g[1] if g else None)
                for i, (p, g) in enumerate(zip(probs, grids))]
        for i in range(len(heavy)):
            assert L.fnec_hip_session_submit(sess, C.byref(reqs[i][0])) == 0
        time.sleep(0.003)
        for i in range(len(heavy), len(probs)):
            assert L.fnec_hip_session_submit(sess, C.byref(reqs[i][0])) == 0
        out = (F.Result * 64)()
        got = {}
        deadline = time.monotonic() + 120.0
        while len(got) < len(probs) and time.monotonic() < deadline:
            k = L.fnec_hip_session_poll(sess, C.cast(out, C.c_void_p), 64, 200000)
            assert k >= 0
            for j in range(k):
                r = out[j]
                got[r.tag] = (r.status, complex(r.z_in.re, r.z_in.im), complex(r.feed_current.re, r.feed_current.im),
                              r.rp_total, r.batch, r.t_start_us, r.t_done_us)
        assert len(got) == len(probs)
    finally:
        L.fnec_hip_session_destroy(sess)
    for i, a in enumerate(alone):
        status, zin, cur, rp_total, _, _, _ = got[i]
        assert status == F.OK and a["status"] == F.OK
        assert zin == a["z_in"]
        assert cur == a["x"][int(probs[i]["feed_index"])]
        if grids[i] is not None:
            assert rp_total == a["rp_total"]
    heavy_end = max(got[i][6] for i in range(len(heavy)))
    light_start = min(got[i][5] for i in range(len(heavy), len(probs)))
    assert light_start < heavy_end


def test_session_returns_currents_and_pattern_rows():
    """Requests that carry output buffers get the currents and the RP rows
    written before their result is polled, bitwise those of a batch call."""
    probs = problems(28, [51, 400, 97, 700, 64])
    grids = [(TH2, PH2) if i % 2 == 0 else None for i in range(len(probs))]
    batch = F.solve_batch(probs, precision=1, rp_grids=grids)
    L = F.lib()
    sess = L.fnec_hip_session_create(1 << 40, 1, 256)
    assert sess
    try:
        reqs = [F.make_request(p, tag=i, rp_theta=g[0] if g else None, rp_phi=g[1] if g else None, outputs=True)
                for i, (p, g) in enumerate(zip(probs, grids))]
        for q, _ in reqs:
            assert L.fnec_hip_session_submit(sess, C.byref(q)) == 0
        out = (F.Result * 16)()
        got = {}
        deadline = time.monotonic() + 60.0
        while len(got) < len(probs) and time.monotonic() < deadline:
            k = L.fnec_hip_session_poll(sess, C.cast(out, C.c_void_p), 16, 200000)
            for j in range(k):
                got[out[j].tag] = out[j].status
        assert len(got) == len(probs)
    finally:
        L.fnec_hip_session_destroy(sess)
    for i, b in enumerate(batch):
        assert got[i] == F.OK
        assert np.array_equal(reqs[i][1]["x"], b["x"])
        if grids[i] is not None:
            assert np.array_equal(reqs[i][1]["rows"], b["rp"])


def test_session_rejects_request_over_budget():
    probs = problems(24, [400])
    st, _ = run_session(probs, 1, 1 << 20, 8)
    assert st == F.RESOURCE_EXHAUSTED


def test_direct_batches_beside_a_session_and_one_lane():
    """Engine lanes under contention: direct batch calls from another thread
    while a session serves (each call waits for a free lane), and a session
    created with one lane. Every result is bitwise that of the request alone."""
    import threading

    probs = problems(29, [51, 300, 77, 900, 150, 64, 500, 120, 700, 90])
    alone = F.solve_batch(probs, precision=1)
    L = F.lib()

    def serve(precision):
        sess = L.fnec_hip_session_create(1 << 40, precision, 4)
        assert sess
        try:
            keep = [F.make_request(p, tag=i) for i, p in enumerate(probs)]
            for q, _ in keep:
                assert L.fnec_hip_session_submit(sess, C.byref(q)) == 0
            out = (F.Result * 16)()
            got = {}
            deadline = time.monotonic() + 120.0
            while len(got) < len(probs) and time.monotonic() < deadline:
                k = L.fnec_hip_session_poll(sess, C.cast(out, C.c_void_p), 16, 200000)
                for j in range(k):
                    got[out[j].tag] = (out[j].status, complex(out[j].z_in.re, out[j].z_in.im))
            return got
        finally:
            L.fnec_hip_session_destroy(sess)

    side = []

    def direct():
        for _ in range(3):
            side.append(F.solve_batch(probs[::2], precision=1))

    th = threading.Thread(target=direct)
    th.start()
    got = serve(1)
    th.join()
    for i, a in enumerate(alone):
        assert got[i] == (F.OK, a["z_in"])
    for res in side:
        for p, a, b in zip(probs[::2], alone[::2], res):
            assert b["status"] == F.OK and np.array_equal(a["x"], b["x"])
    old = F.get_tunable("lanes")
    F.set_tunable("lanes", 1)
    try:
        got = serve(1)
    finally:
        F.set_tunable("lanes", old)
    for i, a in enumerate(alone):
        assert got[i] == (F.OK, a["z_in"])
