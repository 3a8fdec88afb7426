"""CPU: the library's host logic (planner, the server's scratch allocator and
admission decision, from csrc/ggks_host.hpp and csrc/ggks_plan.hpp) built
with g++ -fsanitize=address,undefined and driven through randomised and
edge-case inputs by tests/cpp/host_sanitize.cpp."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_logic_under_sanitizers(tmp_path):
    exe = str(tmp_path / "host_sanitize")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
           "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "ggks-multiselect_amd", "csrc"),
           os.path.join(ROOT, "tests", "cpp", "host_sanitize.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout[-3000:] + r.stderr[-3000:]
