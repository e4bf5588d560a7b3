"""The served-batch planner (which queued requests form the next batch,
with and without per-table ordering) checked on the CPU by its C++ unit
test on random queues and limits."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_plan")


def test_batch_planner_invariants():
    if not os.path.exists(BIN):
        from poggers_amd import build
        build.build_cpp_tests()
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "plan ok" in r.stdout
