"""The ctypes mirrors of the C-ABI structs have the header's layout.

A small C program compiled with gcc against include/*.h prints sizeof and
offsetof of every field; the ctypes Structure of the Python binding must
agree field by field (a field added to one side only, a reordering or a
width change fails here).  CPU only.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _c_name(field):
    return "lambda" if field == "lambda_" else field


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not on PATH")
def test_ctypes_structs_match_header(ct, tmp_path):
    import ctypes
    structs = {"ctbb_geom": ct.Geom, "ctbb_run_params": ct.RunParams,
               "ctbb_icd_opts": ct.IcdOpts, "ctbb_icd_stats": ct.IcdStats}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "ctbb_host.h"',
             "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu %zu\\n", offsetof({cname}, {_c_name(f)}),'
                         f' sizeof((({cname}*)0)->{_c_name(f)}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", str(ROOT / "include"),
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        parts = line.split()
        got[(parts[0], parts[1])] = tuple(int(p) for p in parts[2:])
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == (ctypes.sizeof(py),), cname
        for f, _ in py._fields_:
            desc = getattr(py, f)
            assert got[(cname, f)] == (desc.offset, desc.size), (cname, f)
