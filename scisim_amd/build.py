"""Builds libscx.so (HIP, gfx950) and libscx_scenes.so (host-only scene generators) in-tree.

Usage: python -m scisim_amd.build   (or scisim_amd.build.build())
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libscx.so")
SCENES_LIB = os.path.join(HERE, "libscx_scenes.so")

HIP_SOURCES = ["scx_prims.hip", "scx_broadphase.hip", "scx_contacts.hip", "scx_solver.hip", "scx_mlp.hip", "scx_api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SCX_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code; FMA contraction is chosen
# per file with '#pragma clang fp contract' (off wherever results must round like the
# FP64 CPU restatement, per expression in the sweeps).
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-value",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             ]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_scx(jobs=None):
    objs = []
    procs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    for src in HIP_SOURCES:
        obj = os.path.join(HERE, "_build", src.replace(".hip", ".o"))
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), src))
        objs.append(obj)
    failed = [src for p, src in procs if p.wait() != 0]
    if failed:
        raise RuntimeError("hipcc failed for " + ", ".join(failed))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)


def build_scenes():
    tmp = SCENES_LIB + ".tmp"
    _run(["g++", "-std=c++17", "-O3", "-fPIC", "-shared", "-Wall", "-o", tmp, os.path.join(CSRC, "scenes.cpp")])
    os.replace(tmp, SCENES_LIB)


def build_cpp_tests():
    """C++ programs written against include/scx_scisim.hpp (the reference-shaped host
    interface), linked to libscx.so; run by tests/test_gpu_cpp_interface.py."""
    root = os.path.dirname(HERE)
    src_dir = os.path.join(root, "tests", "cpp")
    out_dir = os.path.join(src_dir, "bin")
    os.makedirs(out_dir, exist_ok=True)
    for name in ("spatial_grid_tests", "slope_tests", "scene_xml_tests"):
        _run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(root, "include"),
              os.path.join(src_dir, name + ".cpp"), "-o", os.path.join(out_dir, name), "-L", HERE, "-lscx",
              "-Wl,-rpath," + HERE])
    # the autotuner's search, host code only (tests/test_tuner_search.py)
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-I", CSRC, os.path.join(src_dir, "tuner_tests.cpp"), "-o",
          os.path.join(out_dir, "tuner_tests")])


def build():
    build_scx()
    build_scenes()
    build_cpp_tests()


if __name__ == "__main__":
    build()
    sys.exit(0)
