"""The serving staging ring (HBM-budget admission) checked on the CPU by its
C++ unit test against a byte-occupancy model."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_ring")


def test_ring_allocator_model():
    if not os.path.exists(BIN):
        from poggers_amd import build
        build.build_cpp_tests()
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ring ok" in r.stdout
