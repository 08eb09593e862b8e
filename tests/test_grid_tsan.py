"""The work grid's concurrency under the thread sanitizer, on the CPU: tests/native/grid_race.cpp (CPU workers claiming
one chunk at a time beside a GPU claimant of 97 and a thread updating the cells' scores; a coordinator appending while
workers claim) is built with g++ -fsanitize=thread together with the grid (x2_grid.cpp) and run; any data race the
sanitizer finds fails the test, as does a chunk claimed twice or not at all."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not installed")
def test_grid_claims_race_free_under_tsan(tmp_path):
    exe = str(tmp_path / "grid_race")
    src = [os.path.join(ROOT, "tests", "native", "grid_race.cpp"),
           os.path.join(ROOT, "xpu2-foundation_amd", "csrc", "x2_grid.cpp")]
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread", "-o", exe] + src, check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66 second_deadlock_stack=1")
    run = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    print(run.stdout[-2000:], run.stderr[-4000:])
    assert "WARNING: ThreadSanitizer" not in run.stderr
    assert run.returncode == 0 and run.stdout.strip().endswith("ok")
