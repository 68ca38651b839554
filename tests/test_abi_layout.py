"""The ctypes mirrors of the C-ABI structs (scisim_amd/__init__.py) against include/scx.h:
a C program built with gcc from the header prints each struct's size and every field's
offset, which must equal what ctypes lays out."""

import ctypes
import os
import subprocess

import scisim_amd as scx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MIRRORS = {
    "scx_config": scx.ScxConfig,
    "scx_step_params": scx.ScxStepParams,
    "scx_step_stats": scx.ScxStepStats,
    "scx_sweep_counters": scx.ScxSweepCounters,
    "scx_contact_set": scx.ScxContactSet,
}


def test_ctypes_mirrors_match_the_header(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "scx.h"', "int main(void) {"]
    for cname, py in MIRRORS.items():
        lines.append('  printf("%s sizeof %%zu\\n", sizeof(%s));' % (cname, cname))
        for field, _ in py._fields_:
            lines.append('  printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, field, cname, field))
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cname, py in MIRRORS.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)
