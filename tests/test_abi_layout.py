"""The ctypes mirrors of the C ABI structs (sparsesolv_amd._lib) against the
header: a small C program compiled with the host compiler prints every
field's offset and each struct's size."""
import ctypes
import shutil
import subprocess
from pathlib import Path

import pytest

from sparsesolv_amd import _lib

ROOT = Path(__file__).resolve().parents[1]
STRUCTS = [("ssa_config", _lib.Config), ("ssa_result", _lib.Result), ("ssa_ic_info", _lib.IcInfo)]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no host C compiler")
def test_struct_layouts_match_header(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "sparsesolv_amd.h"', "int main(void) {"]
    for cname, py in STRUCTS:
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'    printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "abi.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in STRUCTS:
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)
