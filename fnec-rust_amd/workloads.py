"""Seeded synthetic workloads, emitted as NEC deck text (the reference's input
format), so that the product and any checker parse identical inputs.

Frozen choices (each generator's docstring states its own):
- config 1: the reference's corpus deck dipole-freesp-rp-large-grid.nec.
- config 2: 16 y-directed wires at z = 10 m over PEC (GN 1), x spacing 0.2 lambda
  at 14.2 MHz, 128 segments each, lengths U[0.44,0.52] lambda, radii
  log-U[0.5,5] mm, fed at wire 1 segment 64 with 1+0j V, RP 181 x 361 at 1 deg.
- config 3: 32 random straight wires x 256 segments, endpoints U in a 40 m cube,
  radius U[1,3] mm, 7 MHz, free space; wire 2 starts bit-identically at wire 1's
  end (one KCL junction row); wires are rejection-sampled so that no two
  non-adjacent wires come within 0.25 m; fed at wire 1 segment 128.
- config 4a: 8 z-directed half-wave (10.564 m) dipoles x 128 segments, 2 m x
  spacing, radius 1 mm, free space, fed at element 1 segment 64, 256
  frequencies f_i = 3 * 10**(i/255) MHz.
- config 4b / 5: a request is 1-8 wires whose segment counts split N as evenly
  as possible (the first N mod W wires get one more); a set is either parallel
  (x-offset wires) or junctioned (a chain of bent wires sharing endpoints);
  segment length lambda/U[10,40]; centre-fed on wire 1; radius U[0.5,3] mm.
- config 5: 20k requests, N log-U[51,4096], Poisson arrivals, 30% with a
  2 x 2 deg RP grid (theta 0..180 x phi 0..358: 91 x 180 points), 10% over
  PEC ground (GN 1) with every wire above z = 1 m.
"""
from __future__ import annotations

import math

import numpy as np

C0 = 299792458.0

CONFIG1_DECK = (
    "CE\nGW 1 51 0 0 -5.282 0 0 5.282 0.001\nGE\nEX 0 1 26 0 1.0 0.0\n"
    "FR 0 1 0 0 14.2 0.0\nRP 0 37 73 0.0 0.0 5.0 5.0\nEN\n"
)


def _gw(tag, ns, p1, p2, a):
    return "GW %d %d %r %r %r %r %r %r %r" % (tag, ns, float(p1[0]), float(p1[1]), float(p1[2]),
                                               float(p2[0]), float(p2[1]), float(p2[2]), float(a))


def config2_deck(seed: int = 2, rp: bool = True) -> str:
    rng = np.random.default_rng(seed)
    lam = C0 / 14.2e6
    cards = ["CM config 2: 16 parallel wires over PEC, N = 2048"]
    for w in range(16):
        L = rng.uniform(0.44, 0.52) * lam
        a = math.exp(rng.uniform(math.log(0.5e-3), math.log(5e-3)))
        x = w * 0.2 * lam
        cards.append(_gw(w + 1, 128, (x, -L / 2, 10.0), (x, L / 2, 10.0), a))
    cards += ["GN 1", "EX 0 1 64 0 1.0 0.0", "FR 0 1 0 0 14.2 0"]
    if rp:
        cards.append("RP 0 181 361 0.0 0.0 1.0 1.0")
    cards.append("EN")
    return "\n".join(cards) + "\n"


def _seg_dist(p1, p2, q1, q2):
    """Closest distance between segments p1p2 and q1q2."""
    d1, d2, r = p2 - p1, q2 - q1, p1 - q1
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c = d1 @ r
    b = d1 @ d2
    den = a * e - b * b
    s = np.clip((b * f - c * e) / den, 0, 1) if den > 1e-12 else 0.0
    t = (b * s + f) / e
    if t < 0:
        t, s = 0.0, np.clip(-c / a, 0, 1)
    elif t > 1:
        t, s = 1.0, np.clip((b - c) / a, 0, 1)
    return float(np.linalg.norm(p1 + d1 * s - (q1 + d2 * t)))


def config3_deck(seed: int = 3, rp: bool = True, nseg: int = 256, nwires: int = 32) -> str:
    rng = np.random.default_rng(seed)
    ends = []
    while len(ends) < nwires:
        if len(ends) == 1:
            p1 = ends[0][1].copy()  # bit-identical shared endpoint -> one KCL row
        else:
            p1 = rng.uniform(-20.0, 20.0, 3)
        p2 = rng.uniform(-20.0, 20.0, 3)
        if np.linalg.norm(p2 - p1) < 5.0:
            continue
        ok = True
        for j, (q1, q2) in enumerate(ends):
            if len(ends) == 1 and j == 0:
                continue  # the junctioned pair touches by construction
            if _seg_dist(p1, p2, q1, q2) < 0.25:
                ok = False
                break
        if ok:
            ends.append((p1, p2))
    cards = ["CM config 3: 32 random wires x 256 segments, one KCL junction"]
    for w, (p1, p2) in enumerate(ends):
        a = rng.uniform(1e-3, 3e-3)
        cards.append(_gw(w + 1, nseg, p1, p2, a))
    cards += ["EX 0 1 %d 0 1.0 0.0" % (nseg // 2), "FR 0 1 0 0 7.0 0"]
    if rp:
        cards.append("RP 0 37 73 0.0 0.0 5.0 5.0")
    cards.append("EN")
    return "\n".join(cards) + "\n"


def config4a_deck() -> str:
    cards = ["CM config 4a: 8 parallel half-wave dipoles x 128 segments"]
    for w in range(8):
        x = 2.0 * w
        cards.append(_gw(w + 1, 128, (x, 0.0, -5.282), (x, 0.0, 5.282), 0.001))
    cards += ["EX 0 1 64 0 1.0 0.0", "FR 0 1 0 0 14.2 0", "EN"]
    return "\n".join(cards) + "\n"


def config4a_frequencies() -> np.ndarray:
    return np.array([3e6 * 10.0 ** (i / 255.0) for i in range(256)])


def random_request_deck(rng: np.random.Generator, n: int, freq_mhz: float, pec: bool = False,
                        rp_grid: bool = False, junctioned: bool | None = None) -> str:
    """One request: 1-8 wires with n segments in total, parallel or junctioned.

    Segment length follows NEC practice, lambda/U[10,40] (so at least ~110
    radii: the thin-wire kernel stays valid and the reference CLI's
    source length/radius guard is met); a wire's length is its segment count
    times that length, so large N means long wires or big arrays."""
    w = int(rng.integers(1, 9))
    w = max(1, min(w, n // 3))
    counts = [n // w + (1 if i < n % w else 0) for i in range(w)]
    lam = C0 / (freq_mhz * 1e6)
    seg = lam / rng.uniform(10.0, 40.0)
    if junctioned is None:
        junctioned = w > 1 and rng.random() < 0.5
    z0 = 1.0 + rng.uniform(0.0, 0.5) * lam if pec else 0.0
    cards = []
    if junctioned:
        # a chain of bent wires: wire i+1 starts at wire i's end
        p = np.array([0.0, 0.0, z0])
        for i, ns in enumerate(counts):
            L = ns * seg
            th = rng.uniform(0.2, 1.3)
            ph = rng.uniform(0, 2 * math.pi)
            d = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
            if pec and p[2] + L * d[2] < 0.5:
                d[2] = abs(d[2])
            q = p + L * d
            a = rng.uniform(0.5e-3, 3e-3)
            cards.append(_gw(i + 1, ns, p, q, a))
            p = q
    else:
        x = 0.0
        for i, ns in enumerate(counts):
            L = ns * seg
            a = rng.uniform(0.5e-3, 3e-3)
            zc = z0 + L / 2 + 0.5 if pec else 0.0
            cards.append(_gw(i + 1, ns, (x, 0.0, zc - L / 2), (x, 0.0, zc + L / 2), a))
            x += rng.uniform(0.1, 0.3) * lam
    feed = max(1, (counts[0] + 1) // 2)
    if pec:
        cards.append("GN 1")
    cards += ["EX 0 1 %d 0 1.0 0.0" % feed, "FR 0 1 0 0 %r 0" % float(freq_mhz)]
    if rp_grid:
        cards.append("RP 0 91 180 0.0 0.0 2.0 2.0")
    cards.append("EN")
    return "\n".join(cards) + "\n"


def config4b_decks(seed: int = 41, count: int = 512):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(51, 1025))
        f = float(rng.uniform(3.0, 30.0))
        out.append(random_request_deck(rng, n, f))
    return out


def config5_stream(seed: int = 5, count: int = 20000, rate_hz: float = 1000.0, n_min: int = 51,
                   n_max: int = 4096):
    """Requests of the serving benchmark: list of dicts with the deck text,
    N, arrival time (s), RP flag and ground flag."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate_hz))
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        f = float(rng.uniform(3.0, 30.0))
        pec = bool(rng.random() < 0.10)
        rp = bool(rng.random() < 0.30)
        deck = random_request_deck(rng, n, f, pec=pec, rp_grid=rp)
        reqs.append({"id": i, "n": n, "t_arrival": t, "pec": pec, "rp": rp, "deck": deck, "freq_hz": f * 1e6})
    return reqs
