"""Host code under AddressSanitizer, UndefinedBehaviorSanitizer and
ThreadSanitizer (CPU only).

The request loop (minigw_amd/csrc/mgw_serve.cpp) and the gw_input readers and
output writers (minigw_amd/csrc/mgw_io.cpp) are compiled with g++
-fsanitize=address,undefined (errors fatal, leak checking on) together with
host-only harnesses under tests/native/: the request loop over a stand-in
device layer that forces its failure conditions from the host (a device
budget cut, cancellation of queued and running jobs, a judge that reports the
FP64 fallback, destruction with jobs queued), and the I/O code over good,
truncated, malformed and missing files.  The request loop's harness is also
built with -fsanitize=thread: its workers, the submitting thread and the
waiting threads share the job table, the budget and the cancel flags.
"""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(REPO, "tests", "native")
CSRC = os.path.join(REPO, "minigw_amd", "csrc")
H5 = "/root/reference/regression_tests/h2o_serial/gw_input.h5"
FLAGS = ["-std=c++17", "-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-fno-omit-frame-pointer", "-pthread", "-I" + os.path.join(REPO, "include")]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")


def _build(tmp_path, name, sources, libs=(), flags=FLAGS):
    exe = str(tmp_path / name)
    res = subprocess.run(["g++"] + flags + sources + ["-o", exe] + list(libs), capture_output=True,
                         text=True)
    assert res.returncode == 0, res.stderr
    return exe


def _run(cmd):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               TSAN_OPTIONS="halt_on_error=1:second_deadlock_stack=1")
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=300, env=env)
    assert res.returncode == 0, res.stdout + res.stderr
    return res.stdout


def test_request_loop_under_sanitizers(tmp_path):
    exe = _build(tmp_path, "serve_harness",
                 [os.path.join(NATIVE, "serve_harness.cpp"), os.path.join(CSRC, "mgw_serve.cpp")])
    assert "all scenarios passed" in _run([exe])


def test_request_loop_under_thread_sanitizer(tmp_path):
    flags = [f for f in FLAGS if not f.startswith("-fsanitize")] + ["-fsanitize=thread"]
    exe = _build(tmp_path, "serve_harness_tsan",
                 [os.path.join(NATIVE, "serve_harness.cpp"), os.path.join(CSRC, "mgw_serve.cpp")],
                 flags=flags)
    assert "all scenarios passed" in _run([exe])


def test_io_under_sanitizers(tmp_path):
    exe = _build(tmp_path, "io_harness",
                 [os.path.join(NATIVE, "io_harness.cpp"), os.path.join(CSRC, "mgw_io.cpp")], ["-ldl"])
    scratch = tmp_path / "files"
    scratch.mkdir()
    cmd = [exe, str(scratch)] + ([H5] if os.path.exists(H5) else [])
    assert "all checks passed" in _run(cmd)
