"""CPU checks of the Python binding's argument checks: a tensor reaches a
kernel only if it is contiguous, has the element width the table's
representation needs, holds at least as many elements as there are keys, and
(for the table and batch calls) lives on the GPU.  No GPU calls."""
import pytest
import torch

import poggers_amd as pg


def test_value_widths_per_representation():
    assert pg.value_bytes(pg.REP_KV64) == 8
    assert pg.value_bytes(pg.REP_FP16_V16) == 2 and pg.value_bytes(pg.REP_KV16) == 2
    assert pg.value_bytes(pg.REP_FP8_V8) == 1 and pg.value_bytes(pg.REP_FP10_V6) == 1


def test_arg_accepts_a_matching_host_tensor_and_none():
    t = torch.zeros(8, dtype=torch.int16)
    assert pg._arg(t, "vals", 8, 2, device=False).value == t.data_ptr()
    assert pg._arg(None, "vals", 8, 2) is None


@pytest.mark.parametrize("t,n,size,device,msg", [
    (torch.zeros(8, dtype=torch.int16), 8, 2, True, "GPU tensor"),            # host tensor for a table call
    (torch.zeros(4, dtype=torch.int16), 5, 2, False, "holds 4 elements"),     # too short: the kernel would overrun
    (torch.zeros(8, dtype=torch.int16), 8, 8, False, "8-byte elements"),      # 16-bit values for a 64-bit table
    (torch.zeros(16, dtype=torch.int64)[::2], 8, 8, False, "contiguous"),     # strided view
    ([1, 2, 3], 3, 8, False, "torch tensor"),
])
def test_arg_refuses(t, n, size, device, msg):
    with pytest.raises(ValueError, match=msg):
        pg._arg(t, "x", n, size, device=device)


def test_server_request_checks_widths_against_the_table():
    class T:
        def __init__(self, rep):
            self.desc = pg.TableDesc(rep, [])
    tables = [T(pg.REP_KV64), T(pg.REP_FP16_V16)]
    keys = torch.zeros(10, dtype=torch.int64)
    pg.Server._request(tables, 0, pg.OP_INSERT, keys, torch.zeros(10, dtype=torch.int64), None, None)
    with pytest.raises(ValueError, match="8-byte"):
        pg.Server._request(tables, 0, pg.OP_INSERT, keys, torch.zeros(10, dtype=torch.int16), None, None)
    with pytest.raises(ValueError, match="2-byte"):
        pg.Server._request(tables, 1, pg.OP_QUERY, keys, None, torch.zeros(10, dtype=torch.int64), None)
    with pytest.raises(ValueError, match="holds 9"):
        pg.Server._request(tables, 1, pg.OP_QUERY, keys, None, None, torch.zeros(9, dtype=torch.uint8))
    with pytest.raises(ValueError, match="out of range"):
        pg.Server._request(tables, 2, pg.OP_QUERY, keys, None, None, None)
