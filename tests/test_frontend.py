"""The C++ templated front-end (include/poggers_amd/tables.hpp) and the
many-table wrappers (include/poggers_amd/serving.hpp) driven by their own
test program: host-only checks on the CPU; bulk ops, a filter batch and the
request server on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_frontend")


def _binary():
    if not os.path.exists(BIN):
        from poggers_amd import build
        build.build_cpp_tests()
    return BIN


def test_frontend_host_checks():
    r = subprocess.run([_binary(), "cpu"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "cpu checks ok" in r.stdout


@pytest.mark.gpu
def test_frontend_bulk_ops_on_gpu():
    r = subprocess.run([_binary(), "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    for name in ("del_tcf ok", "flat tcf ok", "fast_flat_tcf ok", "p2_table ok", "fast_tcf ok", "hash_table_64 ok", "small_tcf ok", "tcqf_twelve_16 ok",
                 "cms_free_table ok", "tcf3 ok", "filter_batch ok", "request_server ok"):
        assert name in r.stdout
