"""GPU parity of the plane-wave and current-source Hallén solves (the CLI's
EX 1/2/3 and EX 4 routes, solve_session.rs:262-448) against the oracle, on
the plain FP64 path:

- with the tunable exact_backsub = 1 the solution (currents, the cos / sin
  constants, the port voltage) equals the oracle's bit for bit on every deck of
  the host-builder test (linear, elliptic, per wire, on conductor paths, offset
  and loaded current sources), through fnec_deck_solve and through the direct
  entry points fnec_hip_solve_hallen_planewave / _current_source; with the
  default blocked back substitution it stays within 1e-13;
- the receive sweep (solve_session.rs:313-371) is one elimination with one
  right-hand side per incidence: each incidence's solution equals the oracle's
  single-incidence solve bit for bit and the same incidence solved alone, and
  the report's RECEIVE_PATTERN section equals the oracle's;
- the report contract v1 of such decks (SOLVER_MODE, the current source's
  port voltage in FEEDPOINTS) matches the oracle's;
- the other precisions refuse these systems (FnecUnavailable).
"""
import importlib

import numpy as np
import pytest

from test_host_abi import DIPOLE_G, TILT_G, VARIANT_DECKS, _vdeck

pytestmark = pytest.mark.gpu

F = importlib.import_module("fnec-rust_amd")


@pytest.fixture(scope="module", autouse=True)
def device():
    F.init()


@pytest.fixture
def exact_backsub():
    F.set_tunable("exact_backsub", 1)
    yield
    F.set_tunable("exact_backsub", 0)


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def deck_solve(txt, precision=F.PREC_FP64):
    d = F.Deck(txt)
    f = d.frequencies("cli")[0]
    p = d.problem("cli", f)
    x, st = F.solve_problem(p, precision, with_stats=True)
    return p, (x[0] if p["variant"] == 1 else x), st


@pytest.mark.parametrize("txt", VARIANT_DECKS)
def test_variant_solution_equals_the_oracle_bitwise(oracle, exact_backsub, txt):
    p, x, st = deck_solve(txt)
    ref = oracle.run_deck(txt, "cli")
    print(f"variant {p['variant']} route {p['route']} n {len(p['segs'])} swaps {st['n_swaps']}", flush=True)
    assert same_bits(x, ref["x"]), rel(x, ref["x"])
    if p["variant"] == 2:
        assert x[-1] == ref["port_v"]


@pytest.mark.parametrize("txt", VARIANT_DECKS)
def test_variant_solution_blocked_back_substitution(oracle, txt):
    _, x, _ = deck_solve(txt)
    ref = oracle.run_deck(txt, "cli")
    n = len(ref["currents"])
    assert rel(x[:n], ref["currents"]) < 1e-13
    assert rel(x, ref["x"]) < 1e-13


def test_direct_entry_points_equal_the_deck_route(oracle, exact_backsub):
    for txt in (VARIANT_DECKS[2], VARIANT_DECKS[6], VARIANT_DECKS[9], VARIANT_DECKS[10]):
        d = F.Deck(txt)
        f = d.frequencies("cli")[0]
        p = d.problem("cli", f)
        kw = dict(wire_first=p["wf"], wire_last=p["wl"]) if p["route"] == 0 else \
            dict(path_of=p["path_of"], free_ends=p["free_ends"])
        if p["variant"] == 1:
            x = F.solve_hallen_planewave(p["segs"], p["rhs"], p["cos"], p["sin"], f, **kw)
        else:
            x = F.solve_hallen_current_source(p["segs"], p["rhs"], p["cos"], p["src"], p["i0"], f,
                                              stamps=p["stamps"], **kw)
        assert same_bits(x, oracle.run_deck(txt, "cli")["x"])


def test_receive_sweep_is_one_elimination_with_many_right_hand_sides(oracle, exact_backsub):
    txt = _vdeck(TILT_G, "EX 2 4 3 0 10 5 20 25 40 0.3")
    d = F.Deck(txt)
    p = d.problem("cli", 14.2e6)
    xs = F.solve_problem(p, F.PREC_FP64)
    assert xs.shape == (13, 41 + 2)
    ref = oracle.run_deck(txt, "cli")
    assert same_bits(xs[0], ref["x"])
    for q in range(12):
        th, ph = float(p["recv_theta"][q]), float(p["recv_phi"][q])
        one = _vdeck(TILT_G, f"EX 2 1 1 0 {th!r} {ph!r} 20 0 0 0.3")
        assert same_bits(xs[q + 1], oracle.run_deck(one, "cli")["x"]), q
        # the same incidence solved alone
        alone = F.solve_hallen_planewave(p["segs"], p["rhs_all"][q + 1], p["cos"], p["sin"], 14.2e6,
                                         wire_first=p["wf"], wire_last=p["wl"])
        assert same_bits(alone, xs[q + 1]), q


def report_lines_match(got, want):
    """Report text: identical except pattern gains that differ by one unit
    in their last printed place (the device far field sums in another order)."""
    assert len(got) == len(want)
    for a, b in zip(got, want):
        if a == b:
            continue
        ta, tb = a.split(), b.split()
        assert len(ta) == len(tb), (a, b)
        for u, v in zip(ta, tb):
            if u != v:
                assert abs(float(u) - float(v)) <= 1.5e-4, (a, b)


@pytest.mark.parametrize("ex", ["EX 1 7 3 0 0 0 0 15 45", "EX 4 1 26 0 1 0", "EX 1 1 1 0 30 0 0"])
def test_variant_report_matches_the_oracle(oracle, exact_backsub, ex):
    txt = _vdeck(DIPOLE_G, ex, 14.2, "").replace("FR 0 1 0 0 14.2 0", "FR 0 1 0 0 14.2 0\nRP 0 19 1 1000 0 0 10 0")
    got = F.Deck(txt).report(0).splitlines()
    want = oracle.run_deck(txt, "cli", report=True)["report"].splitlines()
    report_lines_match(got, want)
    mode = "hallen-current-source" if ex.startswith("EX 4") else "hallen-planewave"
    assert f"SOLVER_MODE {mode}" in got
    if "7 3" in ex:
        i = got.index("RECEIVE_PATTERN")
        j = want.index("RECEIVE_PATTERN")
        assert got[i:i + 24] == want[j:j + 24]


def test_larger_plane_wave_sweep_on_parallel_wires(oracle, exact_backsub):
    """600 segments on 6 parallel wires, 9 incidences in one elimination:
    every incidence bitwise the oracle's single solve."""
    geom = "\n".join(f"GW {w + 1} 100 {0.3 * w!r} 0 -4.9 {0.3 * w!r} 0 4.9 0.002" for w in range(6))
    txt = _vdeck(geom, "EX 1 3 3 0 30 0 10 30 60", 14.0)
    p = F.Deck(txt).problem("cli", 14.0e6)
    xs, st = F.solve_problem(p, F.PREC_FP64, with_stats=True)
    assert xs.shape == (10, 600 + 12)
    for q in (0, 4, 8):
        th, ph = float(p["recv_theta"][q]), float(p["recv_phi"][q])
        one = _vdeck(geom, f"EX 1 1 1 0 {th!r} {ph!r} 10 0 0", 14.0)
        assert same_bits(xs[q + 1], oracle.run_deck(one, "cli")["x"]), q


@pytest.mark.parametrize("precision", [F.PREC_MIXED, F.PREC_FP64_MFMA])
def test_other_precisions_refuse_variant_systems(precision):
    d = F.Deck(VARIANT_DECKS[0])
    p = d.problem("cli", 14.2e6)
    with pytest.raises(F.FnecUnavailable):
        F.solve_problem(p, precision)
