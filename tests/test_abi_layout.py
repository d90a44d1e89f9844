"""The ctypes mirrors of the C ABI's structs have the header's layout: a small
C program built with gcc against include/fnec_hip.h prints sizeof and the
offset of every field, compared with the ctypes Structures (CPU)."""
import importlib
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

F = importlib.import_module("fnec-rust_amd")

STRUCTS = {
    "fnec_request": F.Request,
    "fnec_result": F.Result,
    "fnec_solve_stats": F.SolveStats,
    "fnec_ground": F.Ground,
    "fnec_deck_problem": F.DeckProblem,
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_ctypes_structs_match_the_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fnec_hip.h"', '#include "fnec_host.h"',
             "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        c, f, v = line.split()
        got[(c, f)] = int(v)
    for cname, py in STRUCTS.items():
        assert got[(cname, "size")] == C_sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def C_sizeof(t):  # noqa: N802
    import ctypes

    return ctypes.sizeof(t)
