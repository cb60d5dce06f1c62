"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
library's host physics (schedule generation, validation, grid parsing,
budgets) and the CPU oracle, driven with ordinary and hostile inputs by
tests/sanitize_driver.cpp.  Any out-of-bounds access, leak or undefined
operation makes the driver exit non-zero with the sanitizer's report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not installed")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize_driver")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "bubble-simulation_amd", "csrc"),
           os.path.join(ROOT, "tests", "sanitize_driver.cpp"),
           os.path.join(ROOT, "bubble-simulation_amd", "csrc", "physics.cpp"),
           os.path.join(ROOT, "oracle", "oracle.cpp"), "-o", exe]
    build = subprocess.run(cmd, capture_output=True, text=True)
    if build.returncode != 0 and "asan" in build.stderr.lower():
        pytest.skip("the sanitizer runtime is not available: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-2000:]
    # verify_asan_link_order=0: tolerate libraries the environment preloads
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, (run.stdout + run.stderr)[-4000:]
    assert "sanitized host run ok" in run.stdout
