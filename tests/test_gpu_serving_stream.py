"""GPU test of the serving loop on the benchmark's own request stream: a seeded
512-request slice of the mixed-N stream (N log-uniform in [51, 4096], random
parallel and junctioned wire sets, 30% with the 2 x 2 degree RP grid, 10% over
PEC ground), submitted to a mixed-precision serving session at its Poisson
arrival times (rescaled to 2000 requests/s, so requests arrive while earlier
batches are on the device), each request carrying output buffers for its
currents and pattern rows.

- every request is returned once, with status OK;
- Z_in within 1e-6 of the plain FP64 path (the reference's algorithm and
  rounding) solved one request at a time, and the currents within 1e-6;
- for the requests with N <= 600, Z_in within 1e-6 of the oracle (the C
  restatement of the reference CPU path) as well;
- the returned pattern rows within 1e-6 (linear gain) of the plain path's
  pattern from its own currents, with the -999.99 sentinels in place.
"""
import concurrent.futures as cf
import ctypes as C
import importlib
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F = importlib.import_module("fnec-rust_amd")
W = importlib.import_module("fnec-rust_amd.workloads")

THETA = np.repeat(np.arange(0.0, 181.0, 2.0)[None, :], 180, axis=0).ravel()
PHI = np.repeat(np.arange(0.0, 360.0, 2.0), 91)


@pytest.fixture(scope="module", autouse=True)
def device():
    F.init()


def gain_close(a, b, tol):
    live = b[:, 0] > -999.0
    assert np.array_equal(live, a[:, 0] > -999.0)
    ga, gb = 10.0 ** (a[live, 0] / 10.0), 10.0 ** (b[live, 0] / 10.0)
    return float(np.max(np.abs(ga - gb)) / np.max(gb))


def test_config5_stream_slice_served_at_poisson_arrivals(oracle):
    stream = W.config5_stream(count=2048)[1536:2048]
    reqs, probs = [], []
    for r in stream:
        d = F.Deck(r["deck"])
        p = d.problem("worker", r["freq_hz"])
        d.close()
        q, keep = F.make_request(p, tag=r["id"], rp_theta=THETA if r["rp"] else None,
                                 rp_phi=PHI if r["rp"] else None, outputs=True)
        reqs.append((q, keep))
        probs.append(p)
    t_arr = np.array([r["t_arrival"] for r in stream])
    t_arr = (t_arr - t_arr[0]) * 0.5  # the stream runs at 1000 requests/s: 2000/s
    L = F.lib()
    sess = L.fnec_hip_session_create(1 << 40, F.PREC_MIXED, 256)
    assert sess
    out = (F.Result * 512)()
    got = {}
    try:
        t0 = time.perf_counter()
        nxt = 0
        while len(got) < len(reqs):
            now = time.perf_counter() - t0
            while nxt < len(reqs) and t_arr[nxt] <= now:
                assert L.fnec_hip_session_submit(sess, C.byref(reqs[nxt][0])) == 0
                nxt += 1
            k = L.fnec_hip_session_poll(sess, C.cast(out, C.c_void_p), 512, 2000)
            assert k >= 0
            for i in range(k):
                assert out[i].status == F.OK, out[i].tag
                assert out[i].tag not in got
                got[out[i].tag] = complex(out[i].z_in.re, out[i].z_in.im)
        assert L.fnec_hip_session_pending(sess) == 0
    finally:
        L.fnec_hip_session_destroy(sess)
    worst_z = worst_x = worst_g = 0.0
    for r, (q, keep), p in zip(stream, reqs, probs):
        x = F.solve_problem(p, precision=F.PREC_FP64)
        z_plain = p["v_feed"] / x[int(p["feed_index"])]
        dz = abs(got[r["id"]] - z_plain) / abs(z_plain)
        n = len(p["segs"])
        dx = float(np.linalg.norm(keep["x"][:n] - x[:n]) / np.linalg.norm(x[:n]))
        assert dz <= 1e-6 and dx <= 1e-6, (r["id"], r["n"], dz, dx)
        worst_z, worst_x = max(worst_z, dz), max(worst_x, dx)
        if r["rp"]:
            rows, _ = F.rp(p["segs"], x[:n], p["freq_hz"], THETA, PHI, p["ground_struct"])
            worst_g = max(worst_g, gain_close(keep["rows"], rows, 1e-6))
    assert worst_g <= 1e-6
    small = [(r, p) for r, p in zip(stream, probs) if r["n"] <= 600]

    def one(r):
        res = oracle.run_deck(r["deck"], "worker", 0, r["freq_hz"])
        return r["id"], complex(res["feed_z"][0])

    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        zo = dict(ex.map(one, [r for r, _ in small]))
    worst_o = max(abs(got[k] - z) / abs(z) for k, z in zo.items())
    print(f"config 5 slice: {len(reqs)} requests ({sum(r['rp'] for r in stream)} with RP, "
          f"{sum(r['pec'] for r in stream)} over PEC); served vs plain FP64: Zin {worst_z:.1e}, "
          f"currents {worst_x:.1e}, gains {worst_g:.1e}; vs oracle ({len(zo)} with N <= 600): {worst_o:.1e}")
    assert worst_o <= 1e-6
