"""The numpy record dtypes the device-invariant tests read back match the C
layouts of ggks_types.hpp (compiled with g++ on the host), and the library
exports the inspection entry point."""
import os
import shutil
import subprocess

import pytest

from invariants import CHILD, CHUNK, QUERY, SEGDESC, SEGSTATE

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "ggks-multiselect_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_record_layouts(tmp_path):
    exe = str(tmp_path / "layout_probe")
    subprocess.run(["g++", "-std=c++17", "-I", CSRC, "-o", exe,
                    os.path.join(HERE, "cpp", "layout_probe.cpp")], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    dts = {"SegDesc": SEGDESC, "SegState": SEGSTATE, "Chunk": CHUNK, "Child": CHILD, "Query": QUERY}
    n = 0
    for line in out.splitlines():
        name, field, val = line.split()
        dt = dts[name]
        if field == "itemsize":
            assert dt.itemsize == int(val), name
        else:
            assert dt.fields[field][1] == int(val), f"{name}.{field}"
        n += 1
    assert n == sum(len(d.names) + 1 for d in dts.values())


def test_debug_regions_exported(pkg):
    L = pkg._lib
    assert hasattr(L.lib(), "ggks_ctx_debug_regions")
    assert L.DebugRegions.segs.offset == 0
