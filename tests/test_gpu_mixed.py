"""GPU tests of the mixed-precision solve (FNEC_PREC_MIXED) through the C ABI.

The mixed path factors Z in FP32 on MFMA and iterates the regularised normal
equations of linear.rs:672-797 in FP64 (preconditioned CG). Gates:
- feedpoint impedance within 1e-6 (relative) of the FP64 reference, the
  accuracy contract of the mixed path;
- currents within 1e-7 of the exact (extended-precision) solution, or of the
  plain FP64 GPU solve where no exact solution is formed (large N);
- normwise backward error of the normal equations <= 1e-12;
- bitwise determinism run to run and alone vs inside a batch.
"""
import importlib
import os

import numpy as np
import pytest

from conftest import GOLDEN
from test_gpu_parity import exact_solution, rel

pytestmark = pytest.mark.gpu

F = importlib.import_module("fnec-rust_amd")
W = importlib.import_module("fnec-rust_amd.workloads")


@pytest.fixture(scope="module", autouse=True)
def device():
    F.init()


def deck_text(name):
    return open(os.path.join(GOLDEN, "corpus", name + ".nec")).read()


def test_mfma_f32_lane_layout():
    a = (np.arange(64, dtype=np.float32).reshape(16, 4) + 1.0).astype(np.float32)
    b = (np.arange(64, dtype=np.float32).reshape(4, 16) * 0.5 + 2.0).astype(np.float32)
    lanes = F.debug_mfma_probe_f32(a, b)
    d = np.zeros((16, 16), np.float32)
    for lane in range(64):
        for v in range(4):
            d[4 * (lane // 16) + v, lane % 16] = lanes[lane, v]
    assert np.array_equal(d, a @ b)


MIXED_CASES = [
    ("dipole-freesp-rp-large-grid", "cli"),
    ("dipole-ground-51seg", "cli"),
    ("dipole-gn0-fresnel-51seg", "cli"),
    ("yagi-5elm-51seg", "cli"),
    ("dipole-loaded", "worker"),
    ("dipole-loaded", "cli"),
    ("multi-source", "cli"),
]


@pytest.mark.parametrize("name,sem", MIXED_CASES)
def test_mixed_matches_oracle(oracle, name, sem):
    txt = deck_text(name)
    r = oracle.run_deck(txt, sem)
    n = r["n"]
    inp = oracle.deck_inputs(txt, r["freq_hz"], sem)
    kind, er, sg = inp["ground"]
    z = oracle.fill_z(inp["segs"], r["freq_hz"], oracle.Ground(kind if kind != 3 else 0, er, sg))
    xe = exact_solution(oracle, inp, z)
    p = F.Deck(txt).problem(sem, r["freq_hz"])
    x, st = F.solve_problem(p, precision=F.PREC_MIXED, with_stats=True)
    e_exact = rel(x[:n], xe[:n])
    print(f"{name}/{sem}: mixed vs exact {e_exact:.2e}, iterations {st['refine_iters']}, "
          f"backward error {st['backward_error']:.2e}")
    assert e_exact < 1e-7
    assert st["backward_error"] <= 1e-12
    for k, idx in enumerate(r["feed_idx"]):
        zin = r["feed_v"][k] / x[idx]
        assert abs(zin - r["feed_z"][k]) / abs(r["feed_z"][k]) <= 1e-6


def test_mixed_gain_matches_oracle(oracle):
    txt = deck_text("dipole-ground-rp-51seg")
    r = oracle.run_deck(txt, "cli")
    p = F.Deck(txt).problem("cli", r["freq_hz"])
    x = F.solve_problem(p, precision=F.PREC_MIXED)
    rows, total = F.rp(p["segs"], x[:r["n"]], r["freq_hz"], r["rp_theta"], r["rp_phi"], p["ground_struct"])
    live = r["rp"][:, 0] > -999.0
    gl = 10.0 ** (rows[live, 0] / 10.0)
    gl_ref = 10.0 ** (r["rp"][live, 0] / 10.0)
    assert np.max(np.abs(gl - gl_ref)) / np.max(gl_ref) <= 1e-6


def random_problems(seed, sizes, sem="worker"):
    rng = np.random.default_rng(seed)
    out = []
    for i, n in enumerate(sizes):
        txt = W.random_request_deck(rng, n, float(rng.uniform(3.0, 30.0)), pec=bool(i % 3 == 2),
                                    junctioned=bool(i % 2))
        d = F.Deck(txt)
        out.append(d.problem(sem, d.frequencies(sem)[0]))
    return out


def test_mixed_matches_plain_on_serving_requests():
    probs = random_problems(11, [51, 97, 180, 333, 640, 1024, 1500, 2300])
    for p in probs:
        x64 = F.solve_problem(p, precision=F.PREC_FP64)
        xm, st = F.solve_problem(p, precision=F.PREC_MIXED, with_stats=True)
        fi = int(p["feed_index"])
        dz = abs(x64[fi] - xm[fi]) / abs(x64[fi])
        di = rel(xm, x64)
        print(f"N {len(p['segs'])}: rel dZin {dz:.2e} rel dI {di:.2e} its {st['refine_iters']} "
              f"bwd {st['backward_error']:.1e}")
        assert dz <= 1e-6
        assert di <= 1e-7
        assert st["backward_error"] <= 1e-12


def test_mixed_batch_equals_single_bitwise():
    probs = random_problems(12, [64, 300, 129, 800, 51])
    batch = F.solve_batch(probs, precision=F.PREC_MIXED)
    for p, b in zip(probs, batch):
        assert b["status"] == F.OK
        x = F.solve_problem(p, precision=F.PREC_MIXED)
        assert np.array_equal(x, b["x"])
    again = F.solve_batch(probs, precision=F.PREC_MIXED)
    for a, b in zip(batch, again):
        assert np.array_equal(a["x"], b["x"])


@pytest.mark.parametrize("gang_n,gang_max", [(64, 16), (150, 3), (0, 16)])
def test_gang_solves_match_plain(gang_n, gang_max):
    """CG solves split over gangs of workgroups (forced onto small problems
    through the tunables; gang_n = 0 keeps one workgroup per problem): the
    result meets the mixed-path gates and a batch equals the single solves
    bitwise, since the gang size depends on n only."""
    probs = random_problems(13, [65, 130, 333, 700, 1100, 191])
    old = F.get_tunable("gang_n"), F.get_tunable("gang_max")
    try:
        F.set_tunable("gang_n", gang_n)
        F.set_tunable("gang_max", gang_max)
        singles = []
        for p in probs:
            x64 = F.solve_problem(p, precision=F.PREC_FP64)
            xm, st = F.solve_problem(p, precision=F.PREC_MIXED, with_stats=True)
            fi = int(p["feed_index"])
            assert abs(x64[fi] - xm[fi]) / abs(x64[fi]) <= 1e-6
            assert rel(xm, x64) <= 1e-7
            assert st["backward_error"] <= 1e-12
            singles.append(xm)
        batch = F.solve_batch(probs, precision=F.PREC_MIXED)
        for x, b in zip(singles, batch):
            assert b["status"] == F.OK
            assert np.array_equal(x, b["x"])
    finally:
        F.set_tunable("gang_n", old[0])
        F.set_tunable("gang_max", old[1])


def test_chained_block_solves_match_plain():
    """CG solves on the flag-chained one-block-per-workgroup kernels (the path
    for problems above solve1_max_n, forced here with solve1_max_n = 0): the
    mixed-path gates hold and a batch equals the single solves bitwise."""
    probs = random_problems(14, [65, 333, 1100, 191])
    old = F.get_tunable("solve1_max_n")
    try:
        F.set_tunable("solve1_max_n", 0)
        singles = []
        for p in probs:
            x64 = F.solve_problem(p, precision=F.PREC_FP64)
            xm, st = F.solve_problem(p, precision=F.PREC_MIXED, with_stats=True)
            fi = int(p["feed_index"])
            assert abs(x64[fi] - xm[fi]) / abs(x64[fi]) <= 1e-6
            assert rel(xm, x64) <= 1e-7
            assert st["backward_error"] <= 1e-12
            singles.append(xm)
        batch = F.solve_batch(probs, precision=F.PREC_MIXED)
        for x, b in zip(singles, batch):
            assert b["status"] == F.OK
            assert np.array_equal(x, b["x"])
    finally:
        F.set_tunable("solve1_max_n", old)


def test_config3_mixed_large_solve():
    """Configuration 3: 32 random wires x 256 segments (N = 8192, one KCL
    junction) at 7 MHz. Mixed vs the plain FP64 GPU solve of the same deck."""
    txt = W.config3_deck(rp=False)
    d = F.Deck(txt)
    f = d.frequencies("worker")[0]
    p = d.problem("worker", f)
    assert len(p["segs"]) == 8192 and len(p["ja"]) == 1
    x64 = F.solve_problem(p, precision=F.PREC_FP64)
    xm, st = F.solve_problem(p, precision=F.PREC_MIXED, with_stats=True)
    fi = int(p["feed_index"])
    dz = abs(x64[fi] - xm[fi]) / abs(x64[fi])
    print(f"config 3: rel dZin {dz:.2e}, rel dI {rel(xm, x64):.2e}, its {st['refine_iters']}, "
          f"backward error {st['backward_error']:.2e}")
    assert dz <= 1e-6
    assert rel(xm[:8192], x64[:8192]) <= 1e-7
    assert st["backward_error"] <= 1e-12


CONFIG3_FIXTURE = os.path.join(GOLDEN, "oracle", "config3.npz")


@pytest.mark.skipif(not os.path.exists(CONFIG3_FIXTURE), reason="config 3 oracle fixture not generated")
def test_config3_against_the_fp64_oracle():
    """Configuration 3 against the FP64 oracle (tests/golden/oracle/config3.npz:
    the oracle's solution, Z_in and 5-degree pattern, and the extended-precision
    solution of the same system): the mixed path's Z_in, currents and gains,
    and the plain path bit for bit with the reference-order back substitution."""
    fx = np.load(CONFIG3_FIXTURE)
    n = 8192
    d = F.Deck(W.config3_deck(rp=False))
    p = d.problem("worker", 7e6)
    fi = int(p["feed_index"])
    xm, st = F.solve_problem(p, precision=F.PREC_MIXED, with_stats=True)
    z_o = fx["feed_z"][0]
    dz = abs(p["v_feed"] / xm[fi] - z_o) / abs(z_o)
    di = rel(xm[:n], fx["x"][:n])
    e_m, e_o = rel(xm[:n], fx["x_exact"][:n]), rel(fx["x"][:n], fx["x_exact"][:n])
    rows, _ = F.rp(p["segs"], xm[:n], 7e6, fx["rp_theta"], fx["rp_phi"], p["ground_struct"])
    live = fx["rp"][:, 0] > -999.0
    assert np.array_equal(rows[:, 0] > -999.0, live)
    gl, gl_o = 10.0 ** (rows[live, 0] / 10.0), 10.0 ** (fx["rp"][live, 0] / 10.0)
    dg = float(np.max(np.abs(gl - gl_o)) / np.max(gl_o))
    print(f"config 3 vs the FP64 oracle: rel dZin {dz:.2e}, rel dI {di:.2e}, max rel dG {dg:.2e}; "
          f"vs exact: mixed {e_m:.2e}, oracle {e_o:.2e}; its {st['refine_iters']}", flush=True)
    assert dz <= 1e-6 and di <= 1e-6 and dg <= 1e-6
    F.set_tunable("exact_backsub", 1)
    try:
        x64 = F.solve_problem(p, precision=F.PREC_FP64)
    finally:
        F.set_tunable("exact_backsub", 0)
    assert np.array_equal(x64.view(np.uint64), fx["x"].view(np.uint64))


def test_mixed_with_row_interchanges():
    """Segments half a radius long: the FP32 LU of Z interchanges rows
    (kappa(Z) ~ 3e3), which exercises the factor permutation in the solves."""
    a, n = 0.01, 40
    ln = n * a * 0.5
    txt = (f"GW 1 {n} 0 0 {-ln / 2} 0 0 {ln / 2} {a}\nGW 2 {n} 0.05 0 {-ln / 2} 0.05 0 {ln / 2} {a}\n"
           f"EX 0 1 {n // 2} 0 1 0\nFR 0 1 0 0 30 0\nEN\n")
    d = F.Deck(txt)
    p = d.problem("worker", d.frequencies("worker")[0])
    x64 = F.solve_problem(p, precision=F.PREC_FP64)
    xm, st = F.solve_problem(p, precision=F.PREC_MIXED, with_stats=True)
    assert st["n_swaps"] > 0
    fi = int(p["feed_index"])
    assert abs(x64[fi] - xm[fi]) / abs(x64[fi]) <= 1e-6
    assert rel(xm, x64) <= 1e-6


def test_update_tile_forms_meet_the_contract():
    """The FP32 trailing update in its forms: 64-row tiles with three real
    products per complex multiply-add, with K = 128 passes on the f16 matrix
    cores for every other step (the default) or one K = 64 pass per step,
    and 128-row tiles with four products (K = 64 passes). The factors differ
    at rounding level; every solve meets the mixed path's gates against the
    plain FP64 solve, and each form is deterministic."""
    probs = random_problems(14, [51, 200, 700, 1300, 2100])
    ref = F.solve_batch(probs, precision=F.PREC_FP64)
    old = F.get_tunable("update_tile_m"), F.get_tunable("update_k128")
    try:
        for tm, k128 in ((64, 1), (64, 0), (128, 1)):
            F.set_tunable("update_tile_m", tm)
            F.set_tunable("update_k128", k128)
            a = F.solve_batch(probs, precision=F.PREC_MIXED)
            b = F.solve_batch(probs, precision=F.PREC_MIXED)
            for p, r, x, y in zip(probs, ref, a, b):
                assert x["status"] == F.OK
                assert np.array_equal(x["x"], y["x"])
                assert abs(x["z_in"] - r["z_in"]) <= 1e-6 * abs(r["z_in"]), (tm, k128, len(p["segs"]))
                assert rel(x["x"], r["x"]) <= 1e-7
                assert x["stats"]["backward_error"] <= 1e-12
    finally:
        F.set_tunable("update_tile_m", old[0])
        F.set_tunable("update_k128", old[1])


def interchange_deck(n):
    """Two parallel wires of n segments half a radius long: kappa(Z) ~ 3e3,
    so the FP32 LU of Z interchanges rows."""
    a = 0.01
    ln = n * a * 0.5
    return (f"GW 1 {n} 0 0 {-ln / 2} 0 0 {ln / 2} {a}\nGW 2 {n} 0.05 0 {-ln / 2} 0.05 0 {ln / 2} {a}\n"
            f"EX 0 1 {n // 2} 0 1 0\nFR 0 1 0 0 30 0\nEN\n")


def test_k128_deferred_update_with_row_interchanges():
    """Deferred K = 128 updates under row interchanges. With N = 400 (7 outer
    steps: the last one alone) and N = 460 (8: the last one combining), the
    interchanges of a combining step move rows that still lack the deferred
    step's update into its U12 block and out of it. A factor that mishandled
    them would cost the CG its rate or its convergence. Both forms meet the
    gates against plain FP64 with a few iterations, and a batch of both
    equals the single solves bitwise."""
    probs = []
    for n in (200, 230):
        d = F.Deck(interchange_deck(n))
        probs.append(d.problem("worker", d.frequencies("worker")[0]))
    ref = [F.solve_problem(p, precision=F.PREC_FP64) for p in probs]
    old = F.get_tunable("update_k128")
    try:
        its = {}
        for k128 in (1, 0):
            F.set_tunable("update_k128", k128)
            batch = F.solve_batch(probs, precision=F.PREC_MIXED)
            for p, x64, b in zip(probs, ref, batch):
                n = len(p["segs"])
                xm, st = F.solve_problem(p, precision=F.PREC_MIXED, with_stats=True)
                assert np.array_equal(xm, b["x"]), (k128, n)
                assert st["n_swaps"] > 0 and st["exact_panels"] > 0, (k128, n)
                fi = int(p["feed_index"])
                assert abs(x64[fi] - xm[fi]) / abs(x64[fi]) <= 1e-6, (k128, n)
                assert rel(xm, x64) <= 1e-6, (k128, n)
                assert st["backward_error"] <= 1e-12, (k128, n)
                its[(k128, n)] = st["refine_iters"]
                print(f"K=128 {k128}, N {n}: {st['n_swaps']} interchanges over {st['exact_panels']} panels, "
                      f"{st['refine_iters']} CG iterations")
        for n in (400, 460):
            assert its[(1, n)] <= its[(0, n)] + 1 and its[(1, n)] <= 8, its
    finally:
        F.set_tunable("update_k128", old)


def test_autotune_keeps_results_within_contract():
    """autotune() on a mixed-N batch: the chosen knobs are no slower than the
    start and the batch solved with them is within 1e-6 (Z_in) of the start."""
    A = importlib.import_module("fnec-rust_amd.tuning")
    probs = random_problems(13, [51, 300, 700, 1200, 2000, 90, 450])
    start = {k: F.get_tunable(k) for k in A.SPACE}
    try:
        before = F.solve_batch(probs, precision=F.PREC_MIXED)
        out = F.autotune(probs, repeats=1, rounds=1)
        print("autotune:", out["knobs"], f"{out['start_seconds'] * 1e3:.1f} -> {out['seconds'] * 1e3:.1f} ms")
        assert out["seconds"] <= out["start_seconds"]
        assert {k: F.get_tunable(k) for k in A.SPACE} == out["knobs"]
        after = F.solve_batch(probs, precision=F.PREC_MIXED)
        for a, b in zip(before, after):
            assert b["status"] == F.OK
            assert abs(b["z_in"] - a["z_in"]) <= 1e-6 * abs(a["z_in"])
    finally:
        for k, v in start.items():
            F.set_tunable(k, v)


def lu_residual(p, lu, ipiv):
    """||P Z32 - L U||_F / ||Z32||_F for packed factors, with P from the
    interchanges in order and Z32 the FP32 rounding of the fill."""
    z = F.fill_z(p["segs"], p["freq_hz"], p["ground_struct"]).astype(np.complex64).astype(np.complex128)
    n = len(z)
    perm = np.arange(n)
    for i, r in enumerate(ipiv):
        perm[i], perm[r] = perm[r], perm[i]
    f = lu.astype(np.complex128)
    low = np.tril(f, -1) + np.eye(n)
    return np.linalg.norm(z[perm] - low @ np.triu(f)) / np.linalg.norm(z)


def test_fp32_factor_reconstructs_z():
    """The mixed path's FP32 factorisation itself, independent of the CG it
    preconditions: P Z32 = L U to FP32 accuracy for each update form, K = 128
    passes on the f16 matrix cores (default) and one K = 64 FP32 pass per
    step. Decks whose factorisation takes interchanges (N = 400, an odd step
    count; N = 460, even) and a random geometry without (N = 700)."""
    probs = []
    for n in (200, 230):
        d = F.Deck(interchange_deck(n))
        probs.append(d.problem("worker", d.frequencies("worker")[0]))
    probs += random_problems(21, [700])
    old = F.get_tunable("update_k128")
    try:
        for i, p in enumerate(probs):
            n = len(p["segs"])
            res = {}
            for k128 in (1, 0):
                F.set_tunable("update_k128", k128)
                lu, ipiv = F.debug_lu_f32(p)
                swaps = int((ipiv != np.arange(n)).sum())
                res[k128] = lu_residual(p, lu, ipiv)
                print(f"N {n}, K=128 {k128}: {swaps} interchanges, relative residual {res[k128]:.2e}")
                # about 1e-7 measured; a missed or doubled update would leave O(1e-2)
                assert res[k128] <= 1e-6, (n, k128, res[k128])
                if i < 2:
                    assert swaps > 0, n
            assert res[1] <= 4.0 * res[0] + 1e-7, (n, res)
    finally:
        F.set_tunable("update_k128", old)
