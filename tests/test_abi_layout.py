"""The ctypes mirrors of the C-ABI structures (atmospheric_flows_amd/_abi.py) have the layout the
header gives them: every field at the offset gcc assigns it in include/afl_types.h, and the same
total size. A mismatch would make the library read and write the wrong bytes of a caller's struct."""
import ctypes
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PAIRS = [("MeshDesc", "afl_mesh_desc"), ("SolverParams", "afl_solver_params"), ("StepInfo", "afl_step_info"),
         ("AdjointStats", "afl_adjoint_stats"), ("HillState", "afl_hill_state"), ("Sponge", "afl_sponge"),
         ("Diagnostics", "afl_diagnostics")]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_ctypes_structs_match_the_header(tmp_path):
    from atmospheric_flows_amd import _abi
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "afl_types.h"', "int main(void) {"]
    expect = []
    for py, c in PAIRS:
        cls = getattr(_abi, py)
        lines.append(f'  printf("{c} size %zu\\n", sizeof({c}));')
        expect.append(f"{c} size {ctypes.sizeof(cls)}")
        for name, _ in cls._fields_:
            lines.append(f'  printf("{c}.{name} %zu\\n", offsetof({c}, {name}));')
            expect.append(f"{c}.{name} {getattr(cls, name).offset}")
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = [g for g in got if g]
    assert got == expect
