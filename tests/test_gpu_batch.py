"""GPU tests of batching and the serving session through the C ABI.

- a batch solves every problem exactly as a single call does (bitwise, both
  precisions): each kernel walks a work list of (problem, tile) items and no
  reduction crosses problems;
- far-field grids of a batch read the currents in place on the device and
  match the stand-alone pattern call bitwise, also when the grids are large
  next to the solves (many small problems with 2 x 2 degree grids);
- the session returns every request once, with the same result as a batch,
  whatever the admission splits (max_batch, HBM budget), and rejects a request
  whose footprint exceeds its budget.
"""
import ctypes as C  # noqa: N812
import importlib

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
        for _ in range(10 * len(probs) + 10):
            k = L.fnec_hip_session_poll(sess, C.cast(out, C.c_void_p), 64, 0)
            assert k >= 0
            for j in range(k):
                r = out[j]
                assert r.tag not in got
                got[r.tag] = (r.status, complex(r.z_in.re, r.z_in.im), r.t_submit_us, r.t_done_us)
            if len(got) == len(probs):
                break
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
            for _ in range(20):
                k = L.fnec_hip_session_poll(sess, C.cast(out, C.c_void_p), 16, 0)
                for j in range(k):
                    got[out[j].tag] = (out[j].status, complex(out[j].z_in.re, out[j].z_in.im))
                if len(got) == len(order):
                    break
            for i in order:
                status, zin = got[100 * rnd + i]
                assert status == F.OK
                assert zin == batch[i]["z_in"]
    finally:
        L.fnec_hip_session_destroy(sess)


def test_session_rejects_request_over_budget():
    probs = problems(24, [400])
    st, _ = run_session(probs, 1, 1 << 20, 8)
    assert st == F.RESOURCE_EXHAUSTED
