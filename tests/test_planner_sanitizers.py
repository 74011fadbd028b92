"""The host-only block planner (qiskit-blackwell_amd/csrc/host_plan.cpp, the
code the serving loop runs on its planner threads) built with the host
compiler under AddressSanitizer and UndefinedBehaviorSanitizer, driven by
tests/sanitize/plan_driver.cpp: random plans replayed exactly in FP64 and
malformed arguments rejected, with every sanitizer report fatal."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "qiskit-blackwell_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="host C++ compiler not found")
def test_planner_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "plan_driver")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           os.path.join(CSRC, "host_plan.cpp"), os.path.join(ROOT, "tests", "sanitize", "plan_driver.cpp"),
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "planner ok" in r.stdout, r.stdout
