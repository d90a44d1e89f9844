"""Pins the oracle's plane-wave and current-source Hallén solves (CPU), the
restatement of planewave.rs, excitation.rs:350-447, linear.rs:910-1319 and
the CLI routing of solve_session.rs:262-448:

- the reference's own checks: the induced-current shapes of its nec2c tables
  (tests/golden/planewave_nec2c.json, from planewave_nec2c.rs) within its 5 /
  7 / 12 % gates; broadside symmetry; elliptic waves that reduce to linear;
  Rayleigh-Carson reciprocity against the transmit pattern; the start-to-
  start split dipole reproducing the single wire through the path solver
  (planewave_junction.rs); the current-source impedance equal to the
  voltage-source one, the forced current honoured, currents linear in i0
  (current_source.rs); the corpus's EX 4 gates;
- mathematics they leave open: the oracle's solution against an
  independently assembled M (numpy) and its regularised normal equations
  solved in extended precision;
- the receive sweep (peak response 0 dB, the first incidence's currents equal
  to the single-incidence solve) and the report sections.
"""
import json
import math

import numpy as np
import pytest

from conftest import read_golden

NEC2C = json.loads(read_golden("planewave_nec2c.json"))
DIPOLE = "CE\nGW 1 51 0 0 -5.282 0 0 5.282 0.001\nGE\n{ex}\nFR 0 1 0 0 14.2 0\nEN\n"


def deck(geom, ex, mhz=14.2):
    return f"CE\n{geom}\nGE\n{ex}\nFR 0 1 0 0 {mhz!r} 0\nEN\n"


def currents(oracle, txt):
    r = oracle.run_deck(txt, "cli")
    return r["currents"], r


def shape_error(cur, ref, lo, hi, pk):
    ref = np.array([complex(a, b) for a, b in ref])
    align = ref[pk] / cur[pk]
    return float(np.max(np.abs(cur[lo:hi] * align - ref[lo:hi])) / abs(ref[pk]))


@pytest.mark.parametrize("case", ["dipole_linear", "tilted_elliptic"])
def test_nec2c_shapes(oracle, case):
    c = NEC2C[case]
    cur, r = currents(oracle, deck(c["deck"], c["ex"], c["freq_hz"] / 1e6))
    assert r["variant"] == "plane-wave" and len(cur) == len(c["currents"])
    e = shape_error(cur, c["currents"], 0, len(cur), c["peak_index"])
    assert e < c["tolerance"], e


def test_nec2c_two_wire_shape(oracle):
    c = NEC2C["two_wire"]
    cur, _ = currents(oracle, deck(c["deck"], c["ex"], c["freq_hz"] / 1e6))
    for lo, hi, pk in c["wires"]:
        assert shape_error(cur, c["currents"], lo, hi, pk) < c["tolerance"]


def test_broadside_symmetry(oracle):
    cur, _ = currents(oracle, DIPOLE.format(ex="EX 1 1 1 0 90 0 0"))
    assert np.max(np.abs(cur - cur[::-1])) / abs(cur[25]) < 1e-6
    two = "GW 1 21 0 0 -2.5 0 0 2.5 0.001\nGW 2 21 1 0 -2.5 1 0 2.5 0.001"
    cur, _ = currents(oracle, deck(two, "EX 1 1 1 0 90 90 0", 28.0))
    assert np.max(np.abs(cur[:21] - cur[21:])) / np.max(np.abs(cur)) < 1e-6


def test_elliptic_reductions(oracle):
    lin, _ = currents(oracle, DIPOLE.format(ex="EX 1 1 1 0 30 0 0"))
    ell, _ = currents(oracle, DIPOLE.format(ex="EX 2 1 1 0 30 0 0 0 0 0.5"))
    assert np.max(np.abs(lin - ell)) / abs(lin[25]) < 1e-9  # z wire: only theta couples
    tilt = "GW 1 41 -3 -3 0 3 3 0 0.001"
    lin, _ = currents(oracle, deck(tilt, "EX 1 1 1 0 45 0 0"))
    e0, _ = currents(oracle, deck(tilt, "EX 3 1 1 0 45 0 0 0 0 0"))
    right, _ = currents(oracle, deck(tilt, "EX 2 1 1 0 45 0 0 0 0 0.5"))
    left, _ = currents(oracle, deck(tilt, "EX 3 1 1 0 45 0 0 0 0 0.5"))
    assert np.max(np.abs(lin - e0)) / abs(lin[20]) < 1e-9
    # opposite hands differ on a tilted wire, and are each other's image under
    # the swap of the minor axis' sign: right(AR) - lin = -(left(AR) - lin)
    assert np.max(np.abs(right - left)) / abs(lin[20]) > 1e-3
    assert np.max(np.abs((right - lin) + (left - lin))) / abs(lin[20]) < 1e-9


def test_reciprocity_with_the_transmit_pattern(oracle):
    tx = oracle.run_deck(DIPOLE.format(ex="EX 0 1 26 0 1 0"), "cli")
    ratios = []
    for th in (40.0, 55.0, 70.0, 90.0):
        rows, _ = oracle.radiation_pattern(tx["segs"], tx["currents"], 14.2e6, [th], [0.0])
        cur, _ = currents(oracle, DIPOLE.format(ex=f"EX 1 1 1 0 {th!r} 0 0"))
        ratios.append(abs(cur[25]) ** 2 / 10.0 ** (rows[0, 1] / 10.0))
    ratios = np.array(ratios)
    assert np.max(np.abs(ratios - ratios.mean())) / ratios.mean() < 0.05


def test_split_dipole_receive_reproduces_the_single_wire(oracle):
    single = "GW 1 52 0 0 -5.282 0 0 5.282 0.001"
    split = "GW 1 26 0 0 0 0 0 5.282 0.001\nGW 2 26 0 0 0 0 0 -5.282 0.001"
    for th in (35.0, 55.0, 75.0, 90.0):
        ex = f"EX 1 1 1 0 {th!r} 0 0"
        a, _ = currents(oracle, deck(single, ex))
        b, rb = currents(oracle, deck(split, ex))
        assert oracle.variant_inputs(deck(split, ex), 14.2e6)["route"] == 1
        assert abs(np.max(np.abs(b)) - np.max(np.abs(a))) / np.max(np.abs(a)) < 1e-8


def test_current_source_matches_the_voltage_source(oracle):
    zv = oracle.run_deck(DIPOLE.format(ex="EX 0 1 26 0 1 0"), "cli")["feed_z"][0]
    r = oracle.run_deck(DIPOLE.format(ex="EX 4 1 26 0 1 0"), "cli")
    assert r["variant"] == "current-source"
    assert abs(r["currents"][25] - 1.0) < 1e-6  # the forced current
    zi = r["port_v"] / 1.0
    assert abs(zi - zv) / abs(zv) < 5e-3 and abs(r["feed_z"][0] - zi) / abs(zi) < 1e-6
    # linear in i0
    r2 = oracle.run_deck(DIPOLE.format(ex="EX 4 1 26 0 2 -1"), "cli")
    assert np.max(np.abs(r2["currents"] - (2 - 1j) * r["currents"])) / abs(r2["currents"][25]) < 1e-9
    assert abs(r2["port_v"] - (2 - 1j) * r["port_v"]) / abs(r2["port_v"]) < 1e-9
    # offset feed and a two-wire array (current_source.rs:133-175)
    for geom, feed in (("GW 1 51 0 0 -5.282 0 0 5.282 0.001", "1 15"),
                       ("GW 1 21 0 0 -2.5 0 0 2.5 0.001\nGW 2 21 1 0 -2.5 1 0 2.5 0.001", "1 11")):
        zv = oracle.run_deck(deck(geom, f"EX 0 {feed} 0 1 0", 28.0 if "GW 2" in geom else 14.2), "cli")["feed_z"][0]
        ri = oracle.run_deck(deck(geom, f"EX 4 {feed} 0 1 0", 28.0 if "GW 2" in geom else 14.2), "cli")
        assert abs(ri["port_v"] - zv) / abs(zv) < 5e-3


def test_current_source_corpus_gate(oracle):
    c = json.loads(read_golden("corpus", "reference-results.json"))["cases"]["dipole-ex4-freesp-51seg"]
    z = oracle.run_deck(DIPOLE.format(ex="EX 4 1 26 0 1.0 0.0"), "cli")["feed_z"][0]
    g, fz = c["tolerance_gates"], c["feedpoint_impedance"]
    assert abs(z.real - fz["real_ohm"]) <= max(g["R_absolute_ohm"], g["R_percent_rel"] / 100 * fz["real_ohm"])
    assert abs(z.imag - fz["imag_ohm"]) <= max(g["X_absolute_ohm"], g["X_percent_rel"] / 100 * fz["imag_ohm"])


def dense_system(oracle, txt, v):
    """M and y assembled here from the oracle's inputs and Z, as the
    reference's linear.rs:910-1319 lay them out."""
    inp = oracle.deck_inputs(txt, 14.2e6 if "14.2" in txt else 28e6, "cli")
    f = 14.2e6 if "14.2" in txt else 28e6
    z = oracle.fill_z(inp["segs"], f)
    n = len(z)
    if v["route"] == 0:
        groups = [list(range(a, b + 1)) for a, b in zip(v["wf"], v["wl"])] or [list(range(n))]
        ends = [(g[0], g[-1]) for g in groups]
    else:
        npth = int(v["path_of"].max()) + 1
        groups = [[m for m in range(n) if v["path_of"][m] == p] for p in range(npth)]
        ends = [tuple(v["free_ends"][2 * p:2 * p + 2]) for p in range(npth)]
    hom = 2 if v["kind"] == "plane-wave" else 1
    cols = n + hom * len(groups) + (v["kind"] == "current-source")
    w8 = 1e6 if v["kind"] == "current-source" else 1.0
    rows = n + 2 * len(groups) + (v["kind"] == "current-source")
    M = np.zeros((rows, cols), complex)
    y = np.zeros(rows, complex)
    M[:n, :n] = z
    for gi, g in enumerate(groups):
        for m in g:
            M[m, n + hom * gi] = -v["cos"][m]
            if hom == 2:
                M[m, n + 2 * gi + 1] = -v["sin"][m]
        M[n + 2 * gi, ends[gi][0]] = w8
        M[n + 2 * gi + 1, ends[gi][1]] = w8
    if v["kind"] == "current-source":
        M[:n, cols - 1] = -v["rhs"]
        M[rows - 1, v["src"]] = w8
        y[rows - 1] = w8 * v["i0"]
    else:
        y[:n] = v["rhs"]
    return M, y


@pytest.mark.parametrize("txt", [DIPOLE.format(ex="EX 1 1 1 0 30 0 0"),
                                 deck("GW 1 26 0 0 0 0 0 5.282 0.001\nGW 2 26 0 0 0 0 0 -5.282 0.001",
                                      "EX 1 1 1 0 55 0 0"),
                                 DIPOLE.format(ex="EX 4 1 20 0 1 0.5"),
                                 deck("GW 1 10 0 0 5 0 3 1 0.001\nGW 2 10 0 0 5 0 -3 1 0.001", "EX 4 1 1 0 1 0")])
def test_solution_solves_the_independently_assembled_system(oracle, txt):
    v = oracle.variant_inputs(txt, 14.2e6)
    r = oracle.run_deck(txt, "cli")
    M, y = dense_system(oracle, txt, v)
    G = M.conj().T @ M + 1e-8 * np.eye(M.shape[1])
    b = M.conj().T @ y
    Gl, bl = G.astype(np.clongdouble), b.astype(np.clongdouble)
    x = np.linalg.solve(G, b).astype(np.clongdouble)
    for _ in range(6):
        x = x + np.linalg.solve(G, (bl - Gl @ x).astype(complex)).astype(np.clongdouble)
    x = x.astype(complex)
    n = v["n"]
    assert len(r["x"]) == M.shape[1]
    # measured 1e-14 .. 1.2e-10 (kappa(G) 5e6 .. 3e15)
    assert np.linalg.norm(r["currents"] - x[:n]) / np.linalg.norm(x[:n]) < 1e-8


def test_receive_sweep(oracle):
    txt = DIPOLE.format(ex="EX 1 7 3 0 0 0 0 15 45")
    r = oracle.run_deck(txt, "cli", report=True)
    rec = r["recv"]
    assert rec.shape == (21, 3)
    assert list(rec[:3, 0]) == [0.0, 0.0, 0.0] and list(rec[:3, 1]) == [0.0, 45.0, 90.0]
    assert np.max(rec[:, 2]) == 0.0
    # a z dipole: the response depends on theta only, peaks broadside and
    # vanishes towards the axis
    by_theta = rec[:, 2].reshape(7, 3)
    assert np.max(np.abs(by_theta - by_theta[:, :1])) < 1e-9
    assert by_theta[6, 0] == 0.0 and by_theta[0, 0] < -40
    # the currents the report prints are the first incidence's
    first, _ = currents(oracle, DIPOLE.format(ex="EX 1 1 1 0 0 0 0"))
    assert np.array_equal(r["currents"], first)
    lines = r["report"].splitlines()
    assert "SOLVER_MODE hallen-planewave" in lines and "FEEDPOINTS" in lines
    i = lines.index("RECEIVE_PATTERN")
    assert lines[i + 1:i + 3] == ["N_POINTS 21", "THETA PHI RESPONSE_DB"]
    assert lines[i + 3].split()[:2] == ["0.0000", "0.0000"]


def test_current_source_report(oracle):
    r = oracle.run_deck(DIPOLE.format(ex="EX 4 1 26 0 1 0"), "cli", report=True)
    lines = r["report"].splitlines()
    assert "SOLVER_MODE hallen-current-source" in lines
    row = lines[lines.index("TAG SEG V_RE V_IM I_RE I_IM Z_RE Z_IM") + 1].split()
    assert row[:2] == ["1", "26"] and row[4:6] == ["1.000000", "0.000000"]
    assert float(row[2]) == pytest.approx(r["port_v"].real, abs=1e-6)
