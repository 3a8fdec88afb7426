"""Build recipe for libggks.so: hipcc for gfx950, in-tree output.

Usage: python ggks-multiselect_amd/build.py
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = [os.path.join(HERE, "csrc", "ggks_select.hip"), os.path.join(HERE, "csrc", "ggks_sort.hip")]
HEADERS = [os.path.join(HERE, "csrc", h) for h in
           ("ggks_keys.hpp", "ggks_types.hpp", "ggks_kernels.hpp", "ggks_plan.hpp",
            "ggks_host.hpp", "ggks_batch.hpp", "ggks_serve.hpp", "ggks_gen.hpp")]
OUT = os.path.join(HERE, "libggks.so")
# C++ harness of the reference-named wrappers (include/ggks_compat.hpp):
# host-only g++, linked against libggks.so
HARNESS_SRC = os.path.join(ROOT, "tests", "cpp", "compat_harness.cpp")
HARNESS = os.path.join(ROOT, "tests", "cpp", "compat_harness")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "csrc"),
]


def build(verbose: bool = False) -> str:
    tmp = OUT + ".tmp"
    cmd = [HIPCC, *FLAGS, *SOURCES, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass does not know the packed-fp32 feature: harmless warning
    noise = "packed-fp32-ops' is not a recognized feature"
    err = "\n".join(l for l in r.stderr.splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, OUT)
    hcmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), HARNESS_SRC,
            "-L" + HERE, "-lggks", "-Wl,-rpath,$ORIGIN/../../ggks-multiselect_amd",
            "-o", HARNESS + ".tmp"]
    if verbose:
        print(" ".join(hcmd), flush=True)
    r = subprocess.run(hcmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for the compat harness:\n{r.stderr}")
    os.replace(HARNESS + ".tmp", HARNESS)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
