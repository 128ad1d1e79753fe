"""The resident serving protocol on the CPU, under ThreadSanitizer.

tests/model/resident_model.cpp runs the server's own host-side core
(syndrome_amd/csrc/serve_core.h) against host threads that execute the
resident kernel's protocol steps (syndrome_amd/csrc/resident_protocol.h),
with random delays between steps, short grid lifetimes (every lull ends the
grid, so relaunches race with publication), a 64-entry ring and 16
transport-block records that wrap constantly.  It stops at its first
failure: a block or transport block delivered twice or not at all, a wrong
result, a wrong transport-block decision, per-block CRC flag or payload.
ThreadSanitizer checks every cross-thread access the protocol makes for a
happens-before edge.  With --fail-launch one relaunch fails: every pending
block and transport block must then be delivered once, as failed, and new
submissions refused (the server's failure path).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "model", "resident_model.cpp")


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no C++ compiler")
    exe = tmp_path_factory.mktemp("model") / "resident_model"
    cmd = [cxx, "-std=c++20", "-O1", "-g", "-fsanitize=thread", "-pthread", "-Wall", "-Werror",
           "-I", os.path.join(ROOT, "syndrome_amd", "csrc"), SRC, "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _run(exe, *args):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66 second_deadlock_stack=1")
    r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=240, env=env)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stderr[-4000:]
    assert "model ok" in r.stdout
    return r.stdout


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_every_block_and_transport_block_delivered_once(model, seed):
    out = _run(model, "--seconds", "2", "--seed", str(seed), "--clients", "3")
    launches = int(out.split(" grid launches")[0].rsplit(" ", 1)[1])
    assert launches > 20  # grids left and were relaunched many times


@pytest.mark.parametrize("seed", [4, 5])
def test_failed_launch_delivers_everything_as_failed(model, seed):
    out = _run(model, "--seconds", "1.5", "--seed", str(seed), "--fail-launch")
    assert "sticky error -7" in out and "submits refused" in out
