"""The server's scheduling core on the CPU, under sanitizers.

hera_gpu_amd/csrc/hv_sched.h holds everything of the serving loop that is not
a HIP call: queueing, admission by footprint, instances per shape bucket
(reuse and LRU eviction), completion, result copies, output pre-faulting,
cancellation and shutdown.  tests/sched/fake_server.cpp runs it against a
fake device (a thread per launch that writes a byte pattern) through a
concurrent mix with injected failures at admission, launch and on the
device, cancels, an over-budget job, a budget that forces evictions and a
shutdown with queued jobs.  It is built twice for the host: with
ThreadSanitizer (ROCm's clang++, whose runtime intercepts
pthread_cond_clockwait; GCC 11's libtsan does not and then misreads every
std::condition_variable::wait_for as a held mutex), and with
AddressSanitizer + UndefinedBehaviorSanitizer (g++).
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "sched" / "fake_server.cpp"
INC = [f"-I{ROOT / 'include'}", f"-I{ROOT / 'hera_gpu_amd' / 'csrc'}"]

CLANG = "/opt/rocm/llvm/bin/clang++"
SANITIZERS = {
    "thread": (CLANG, ["-fsanitize=thread"]),
    "address_undefined": ("g++", ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]),
}


def _env():
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = "halt_on_error=1 exitcode=66 second_deadlock_stack=1"
    env["ASAN_OPTIONS"] = "halt_on_error=1 detect_leaks=1 verify_asan_link_order=0 exitcode=67"
    env["UBSAN_OPTIONS"] = "halt_on_error=1 print_stacktrace=1"
    return env


@pytest.mark.parametrize("san", sorted(SANITIZERS))
def test_scheduler_under_sanitizer(san, tmp_path):
    cxx, flags = SANITIZERS[san]
    if shutil.which(cxx) is None:
        pytest.skip(f"{cxx} not available")
    exe = tmp_path / f"fake_server_{san}"
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-pthread", *flags, *INC, str(SRC), "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    r = subprocess.run([str(exe), "300"], capture_output=True, text=True, timeout=240, env=_env())
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    assert "all scheduling checks passed" in r.stdout
    assert "WARNING: ThreadSanitizer" not in out and "ERROR: AddressSanitizer" not in out, out[-6000:]
    assert "runtime error" not in out, out[-6000:]
