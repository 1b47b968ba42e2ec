"""CPU: the host code under the sanitizers.

- The CPU oracle (oracle/contour3d_oracle.c, the checker every parity test
  leans on) built with AddressSanitizer and UndefinedBehaviorSanitizer,
  driven through both of its entry points on small fields, a one-cell-thick
  lattice, a lattice with no crossing, two contours and lines
  (tests/oracle_sanitize/driver.c): no out-of-bounds access, leak or
  undefined operation.
- The serving worker's wait (contour3d_amd/csrc/c3d_wait.h) under
  ThreadSanitizer with the worker's predicate and a submitting thread
  (tests/server_wait/check_wait.cpp): no data race."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "oracle_asan"
    subprocess.run(["gcc", "-std=c99", "-O1", "-g", "-ffp-contract=off", "-fno-omit-frame-pointer",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "contour3d_amd", "csrc", "single_rank"),
                    os.path.join(ROOT, "oracle", "contour3d_oracle.c"),
                    os.path.join(ROOT, "tests", "oracle_sanitize", "driver.c"), "-o", str(exe), "-lm"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout, r.stderr[-3000:])
    assert r.returncode == 0, r.stderr[-3000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    t, cov = r.stdout.split()[1], r.stdout.split()[3]
    assert int(t) > 100 and int(cov) > 100


# gcc 11's ThreadSanitizer does not intercept pthread_cond_clockwait, which
# libstdc++'s condition_variable::wait_until(steady_clock) calls, and reports
# a "double lock" and a data race for any such wait (a ten-line program with
# one correct timed wait shows it); LLVM's runtime intercepts it
CLANG = "/opt/rocm/llvm/bin/clang++"


def test_server_wait_under_tsan(tmp_path):
    if not os.path.exists(CLANG):
        pytest.skip("no clang++ with a ThreadSanitizer runtime")
    exe = tmp_path / "check_wait_tsan"
    r = subprocess.run([CLANG, "-O1", "-g", "-std=c++17", "-pthread", "-fsanitize=thread",
                        "-I", os.path.join(ROOT, "contour3d_amd", "csrc"),
                        os.path.join(ROOT, "tests", "server_wait", "check_wait.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip(f"no ThreadSanitizer runtime for this compiler: {r.stderr[-300:]}")
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout, r.stderr[-3000:])
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[-3000:]
    assert r.returncode == 0, r.stdout
