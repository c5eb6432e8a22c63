"""The shard engine's scheduler (apogee-vm_amd/csrc/jobqueue.h) built on the
host alone with the thread sanitizer and the address sanitizer, and run over
mock workers (tests/native/sched_stress.cpp): outputs per job, the HBM budget
with retained arenas, the fast lane, refused ids, a worker that fails to
start, draining on destruction. The reference tests determinism under
concurrency the same way (crates/prover/tests/streaming.rs:54-110)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "sched_stress.cpp")
INC = os.path.join(ROOT, "apogee-vm_amd", "csrc")


@pytest.mark.parametrize("sanitizer", ["thread", "address"])
def test_scheduler_under_sanitizer(tmp_path, sanitizer):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = tmp_path / f"sched_{sanitizer}"
    subprocess.run([gxx, "-std=c++17", "-O1", "-g", f"-fsanitize={sanitizer}", "-fno-omit-frame-pointer",
                    "-I", INC, SRC, "-o", str(exe), "-lpthread"], check=True, capture_output=True)
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = "halt_on_error=1"
    env["ASAN_OPTIONS"] = "detect_leaks=1:halt_on_error=1"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
