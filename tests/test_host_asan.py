"""The host graph code (xbfs_amd/csrc/graph_host.cpp: generators, binary
loader and writer, text edge-list reader, CSR validation, concatenation,
degree statistics) built with AddressSanitizer and UndefinedBehaviorSanitizer
and run over valid and malformed inputs (tests/asan_host_driver.cpp).  The
reference has no sanitizer run; its loader only asserts (graph.hpp:32,37,61-62).
CPU only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_graph_host_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "asan_host_driver")
    build = subprocess.run(
        ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
         "-fopenmp", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "xbfs_amd", "csrc"),
         os.path.join(ROOT, "tests", "asan_host_driver.cpp"),
         os.path.join(ROOT, "xbfs_amd", "csrc", "graph_host.cpp"), "-o", exe],
        capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-4000:]
    work = tmp_path / "files"
    work.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4")
    run = subprocess.run([exe, str(work)], capture_output=True, text=True, timeout=300, env=env)
    out = run.stdout + run.stderr
    if "LeakSanitizer does not work under ptrace" in out:  # some sandboxes trace processes
        env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0"
        run = subprocess.run([exe, str(work)], capture_output=True, text=True, timeout=300, env=env)
        out = run.stdout + run.stderr
    assert run.returncode == 0, out[-4000:]
    assert "OK (0 failures)" in run.stdout
    assert "runtime error" not in out and "AddressSanitizer" not in out
