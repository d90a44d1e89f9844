"""Writes planewave_nec2c.json: the nec2c induced-current tables the
reference's plane-wave tests hold (crates/nec_solver/tests/planewave_nec2c.rs:
NEC2C_CURRENTS, NEC2C_TILT_ELLIPTIC, NEC2C_TWO_WIRE), read from that file's
text as data, with the geometry and incidence each table belongs to.
Run once in a container that has the reference checkout:
    python tests/golden/make_planewave_fixtures.py"""
import json
import os
import re

SRC = "/root/reference/crates/nec_solver/tests/planewave_nec2c.rs"
text = open(SRC).read()


def table(name):
    m = re.search(name + r": \[\(f64, f64\); (\d+)\] = \[(.*?)\];", text, re.S)
    rows = [(float(a), float(b)) for a, b in re.findall(r"\(\s*([-0-9.eE+]+),\s*([-0-9.eE+]+)\)", m.group(2))]
    assert len(rows) == int(m.group(1)), name
    return rows


out = {
    "source": "crates/nec_solver/tests/planewave_nec2c.rs",
    "dipole_linear": {"deck": "GW 1 51 0 0 -5.282 0 0 5.282 0.001", "freq_hz": 14.2e6,
                      "ex": "EX 1 1 1 0 30 0 0", "peak_index": 25, "tolerance": 0.05,
                      "currents": table("NEC2C_CURRENTS")},
    "tilted_elliptic": {"deck": "GW 1 41 -3 -3 0 3 3 0 0.001", "freq_hz": 14.2e6,
                        "ex": "EX 2 1 1 0 45 0 0 0 0 0.5", "peak_index": 20, "tolerance": 0.07,
                        "currents": table("NEC2C_TILT_ELLIPTIC")},
    "two_wire": {"deck": "GW 1 21 0 0 -2.5 0 0 2.5 0.001\nGW 2 21 1 0 -2.5 1 0 2.5 0.001", "freq_hz": 28e6,
                 "ex": "EX 1 1 1 0 30 0 0", "wires": [[0, 21, 10], [21, 42, 31]], "tolerance": 0.12,
                 "currents": table("NEC2C_TWO_WIRE")},
}
json.dump(out, open(os.path.join(os.path.dirname(__file__), "planewave_nec2c.json"), "w"), indent=1)
