"""The serving pool's host scheduler on the CPU, through a mock device that
follows the kernels' queue rules (tests/native/srv_mock_main.cpp): 240
requests of uneven size, Kerr and gridded, with FP64 hand-overs, submitted
from one thread while a second steps the pool and a third collects and
checks every result; more Kerr classes than fit at once, an output budget
smaller than the stream, a device table smaller than the stream, launches
with fewer lanes than entries to resume (carried over).  Built and
run plainly, under ThreadSanitizer and under AddressSanitizer with
UndefinedBehaviorSanitizer."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "native", "srv_mock_main.cpp"),
       os.path.join(ROOT, "gray_amd", "csrc", "srv_sched.cpp")]
INC = ["-I", os.path.join(ROOT, "gray_amd", "csrc"), "-I", os.path.join(ROOT, "include")]

VARIANTS = {
    "plain": ["-O2"],
    "tsan": ["-O1", "-g", "-fsanitize=thread"],
    "asan": ["-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
             "-fno-sanitize-recover=undefined"],
}


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_scheduler_with_mock_device(tmp_path, variant):
    exe = str(tmp_path / f"srv_mock_{variant}")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-pthread", *VARIANTS[variant],
                        *INC, *SRC, "-o", exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = "halt_on_error=1 second_deadlock_stack=1"
    env["ASAN_OPTIONS"] = "detect_leaks=1 halt_on_error=1"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK" in r.stdout and "WARNING: ThreadSanitizer" not in r.stderr, r.stdout + r.stderr
    # the run carried unclaimed entries of every list over (launches with
    # fewer lanes than entries to resume)
    line = [l for l in r.stdout.splitlines() if l.startswith("carried")][0].split()
    assert all(int(line[k]) > 0 for k in (2, 4, 6)), line
