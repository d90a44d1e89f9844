"""Writes tests/golden/oracle/*.npz: full-size results of the oracle (the C
restatement of the reference CPU path) that GPU parity tests compare against
when the oracle itself would take minutes. Run from the repository root:
    python tests/golden/make_oracle_fixtures.py [config2] [config3] [config4a_f0] ...
"""
import importlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

W = importlib.import_module("fnec-rust_amd.workloads")
OUT = os.path.join(ROOT, "tests", "golden", "oracle")


def exact_solution(deck, freq_hz, semantics):
    """Regularised LS solution in 80-bit extended precision: LAPACK solve of
    the FP64 normal equations refined with M^H (y - M x) - lambda x evaluated
    in np.clongdouble."""
    import scipy.linalg as sl
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from test_oracle import build_m_numpy
    inp = O.deck_inputs(deck, freq_hz, semantics)
    kind, er, sg = inp["ground"]
    z = O.fill_z(inp["segs"], freq_hz, O.Ground(kind if kind != 3 else 0, er, sg))
    M, y = build_m_numpy(z, inp["cos"], inp["rhs"], list(inp["wf"]), list(inp["wl"]), list(inp["ja"]),
                         list(inp["jb"]), list(inp["js"]))
    lu = sl.lu_factor(M.conj().T @ M + 1e-8 * np.eye(M.shape[1]))
    Ml, yl = M.astype(np.clongdouble), y.astype(np.clongdouble)
    x = sl.lu_solve(lu, M.conj().T @ y).astype(np.clongdouble)
    for _ in range(6):
        r = Ml.conj().T @ (yl - Ml @ x) - np.clongdouble(1e-8) * x
        x = x + sl.lu_solve(lu, r.astype(np.complex128)).astype(np.clongdouble)
    return x.astype(np.complex128)


def run(name, deck, freq_hz=0.0, semantics="cli", threads=1):
    # row loops and cache-blocked loop nests: the results do not depend on
    # either (each element takes the same operations in the same order)
    O.set_threads(threads)
    O.set_blocked(True)
    t0 = time.time()
    r = O.run_deck(deck, semantics, 0, freq_hz)
    dt = time.time() - t0
    O.set_threads(1)
    O.set_blocked(False)
    kw = dict(x=r["x"], feed_z=r["feed_z"], feed_idx=np.array(r["feed_idx"]), abs_res=r["abs_res"],
              rel_res=r["rel_res"], diag_spread=r["diag_spread"], seconds=dt, freq_hz=r["freq_hz"],
              x_exact=exact_solution(deck, r["freq_hz"], semantics))
    if "rp" in r:
        kw.update(rp=r["rp"], rp_theta=r["rp_theta"], rp_phi=r["rp_phi"], rp_total=r["rp_total"])
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **kw)
    print(name, "N", r["n"], "S", r["s"], "seconds %.1f" % dt, "Zin", r["feed_z"][0], flush=True)


if __name__ == "__main__":
    os.makedirs(OUT, exist_ok=True)
    O.build()
    which = sys.argv[1:] or ["config2"]
    if "config2-exact" in which:
        # add the extended-precision solution to an existing config2 fixture
        f = os.path.join(OUT, "config2.npz")
        d = dict(np.load(f))
        d["x_exact"] = exact_solution(W.config2_deck(), 14.2e6, "cli")
        np.savez_compressed(f, **d)
        n = 2048
        print("config2 reference error vs exact:",
              np.linalg.norm(d["x"][:n] - d["x_exact"][:n]) / np.linalg.norm(d["x_exact"][:n]))
    if "config2" in which:
        run("config2", W.config2_deck())
    if "config3" in which:
        # N = 8192, worker semantics (one KCL row), the 5-degree RP grid
        run("config3", W.config3_deck(), 0.0, "worker",
            threads=int(os.environ.get("FIXTURE_THREADS", os.cpu_count() or 1)))
    if "config3-rp" in which:
        # the oracle's pattern (gains and normalisation) on the deck's 5-degree
        # grid from the oracle's currents, added to an existing config3 fixture
        f = os.path.join(OUT, "config3.npz")
        d = dict(np.load(f))
        txt = W.config3_deck()
        inp = O.deck_inputs(txt, float(d["freq_hz"]), "cli")
        th = np.array([5.0 * it for ip in range(73) for it in range(37)])
        ph = np.array([5.0 * ip for ip in range(73) for it in range(37)])
        t0 = time.time()
        rows, total = O.radiation_pattern(inp["segs"], d["x"][:8192], float(d["freq_hz"]), th, ph)
        d.update(rp=rows, rp_theta=th, rp_phi=ph, rp_total=total)
        np.savez_compressed(f, **d)
        print("config3 rp", rows.shape, "seconds %.1f" % (time.time() - t0), flush=True)
    if "config4a_f0" in which:
        run("config4a_f0", W.config4a_deck(), W.config4a_frequencies()[0])
