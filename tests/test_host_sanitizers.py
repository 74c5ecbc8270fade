"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only,
no GPU code is instrumented): the solver's planning logic (line plans for
every length 1..4096, tile configurations, tuning table, twiddle, Bluestein
and finite-difference tables; fishfry_amd/csrc/ff_plan.cpp), the serving
arena (ff_arena.hpp) and the CPU oracle (oracle/oracle.c).  The serving
scheduler (ff_server.cpp) runs under ThreadSanitizer, and under ASan/UBSan,
with client, waiter and statistics threads, on a CPU stand-in device
(tests/cpp/fake_device.cpp: HIP runtime calls and the solver entry points
it makes, events that complete asynchronously).  Each program exits
non-zero on a failed check, and the sanitizers abort on the first finding
(-fno-sanitize-recover, halt_on_error), so a clean exit means no report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fishfry_amd", "csrc")
CPP = os.path.join(ROOT, "tests", "cpp")
HIPCC = "/opt/rocm/bin/hipcc"
# host-side instrumentation only (the planner is compiled host-only)
HOST_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-Xarch_host", "-fno-sanitize-recover=all"]
GCC_SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
           TSAN_OPTIONS="halt_on_error=1")
HOST_TSAN = ["-Xarch_host", "-fsanitize=thread"]


def run(exe, expect):
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=ENV)
    assert r.returncode == 0, r.stdout + r.stderr
    assert expect in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert "ThreadSanitizer" not in r.stderr


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_planner_under_sanitizers(tmp_path):
    exe = tmp_path / "planner_test"
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer"]
                   + HOST_SAN + ["-I", CSRC, os.path.join(CPP, "planner_test.cpp"), os.path.join(CSRC, "ff_plan.cpp"),
                                 "-o", str(exe), "-lpthread"], check=True, capture_output=True)
    run(exe, "planner ok")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_arena_under_sanitizers(tmp_path):
    exe = tmp_path / "arena_test"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer"] + GCC_SAN
                   + ["-I", CSRC, os.path.join(CPP, "arena_test.cpp"), "-o", str(exe)], check=True)
    run(exe, "arena ok")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not installed")
def test_oracle_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_sanitize"
    subprocess.run(["gcc", "-std=gnu11", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off"] + GCC_SAN
                   + ["-I", os.path.join(ROOT, "oracle"), os.path.join(CPP, "oracle_sanitize.c"),
                      os.path.join(ROOT, "oracle", "oracle.c"), "-o", str(exe), "-lm"], check=True)
    run(exe, "oracle ok")


def build_server(tmp_path, san, name):
    exe = tmp_path / name
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer"]
                   + san + ["-I", CSRC, os.path.join(CPP, "server_tsan.cpp"), os.path.join(CSRC, "ff_server.cpp"),
                            os.path.join(CPP, "fake_device.cpp"), "-o", str(exe), "-lpthread"],
                   check=True, capture_output=True)
    return exe


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_server_scheduler_under_thread_sanitizer(tmp_path):
    run(build_server(tmp_path, HOST_TSAN, "server_tsan"), "server ok")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_server_scheduler_under_address_sanitizer(tmp_path):
    run(build_server(tmp_path, HOST_SAN, "server_asan"), "server ok")
