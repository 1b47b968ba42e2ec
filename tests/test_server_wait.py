"""CPU: the serving worker's wait (c3d::wait_work in
contour3d_amd/csrc/c3d_wait.h).  With a batching window (max_wait_ms > 0),
a request that arrives while the idle worker waits with no time bound must
start when its window ends, without another notification.
tests/server_wait/check_wait.cpp drives the header with the worker's
predicate; a build that picks the wait's form once, before waiting, must
fail the same check (such a worker never started the last request of a
burst)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "server_wait", "check_wait.cpp")
INC = os.path.join(ROOT, "contour3d_amd", "csrc")


def build(tmp_path, *defs):
    exe = tmp_path / ("check_wait" + "".join(d.replace("-D", "_") for d in defs))
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I", INC, *defs, SRC, "-o", str(exe)], check=True)
    return exe


def test_request_starts_when_its_window_ends(tmp_path):
    r = subprocess.run([str(build(tmp_path))], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.count("started after") == 2


def test_check_sees_a_wait_decided_once(tmp_path):
    r = subprocess.run([str(build(tmp_path, "-DDECIDE_ONCE"))], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 1
    assert "MISS n=1" in r.stdout
