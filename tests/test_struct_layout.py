"""The ctypes structures in snapatac2_amd._lib have the layout of the C structs
that include/snapgpu.h declares: a small C program compiled against the header
prints sizeof and every field's offsetof, which must equal ctypes', and the
field names, in order, are the header's (no GPU)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from snapatac2_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "snapgpu_panel_layout": _lib.PanelLayout,
    "snapgpu_block_layout": _lib.BlockLayout,
    "snapgpu_window_layout": _lib.WindowLayout,
    "snapgpu_peak_index": _lib.PeakIndex,
    "snapgpu_frag_options": _lib.FragOptions,
    "snapgpu_frag_summary": _lib.FragSummary,
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_structs_match_header(tmp_path):
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "snapgpu.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, key, value = line.split()
        got[(cname, key)] = int(value)
    for cname, cls in STRUCTS.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, (cname, field)


def header_fields(cname):
    src = open(os.path.join(ROOT, "include", "snapgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    body = re.search(r"typedef struct " + cname + r" \{(.*?)\} " + cname + ";", src, flags=re.S).group(1)
    # each declaration ends with its field name: "const uint64_t* start", "int32_t n_chrom"
    return [re.search(r"(\w+)\s*$", d).group(1) for d in body.split(";") if d.strip()]


def test_ctypes_fields_are_the_header_fields():
    for cname, cls in STRUCTS.items():
        assert header_fields(cname) == [f for f, _ in cls._fields_], cname
