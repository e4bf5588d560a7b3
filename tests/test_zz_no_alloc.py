"""pg_fill_noalloc (the fill count through a static device scratch array, no
allocation) against pg_fill, through the C ABI: two calls in a row, the
second on a smaller table, then tables of one and of four levels, each call
checked against pg_fill on the same table and the second call checked to
start from a cleared scratch array."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
import poggers_amd as pg

pytestmark = pytest.mark.gpu


def fills(t):
    lib = pg.lib()
    a, b = ctypes.c_uint64(), ctypes.c_uint64()
    s = pg._stream(None)
    pg.check(lib.pg_fill(t._h, ctypes.byref(a), s), "pg_fill")
    pg.check(lib.pg_fill_noalloc(t._h, ctypes.byref(b), s), "pg_fill_noalloc")
    return a.value, b.value


def filled(desc, n, seed):
    t = pg.Table(desc)
    keys = torch.from_numpy(oracle.gen_keys(seed, 0, n).view(np.int64)).cuda()
    st = torch.zeros(n, dtype=torch.uint8, device="cuda")
    t.bulk_insert(keys, None, st)
    torch.cuda.synchronize()
    return t, int((st > 0).sum().item())


def test_fill_noalloc_equals_fill_on_two_calls_second_smaller():
    big, nbig = filled(pg.tcf_desc(1 << 20), 900_000, 1)
    small, nsmall = filled(pg.kv_desc(1 << 12, bucket_size=4, tile=4, prober=pg.PROBE_DOUBLE), 3000, 2)
    a, b = fills(big)
    assert a == b == nbig
    a, b = fills(small)          # smaller input: the scratch array holds the big count until cleared
    assert a == b == nsmall
    one, n1 = filled(pg.key_only_desc(1 << 10), 700, 3)
    a, b = fills(one)
    assert a == b == n1
    four = pg.tcf_desc(1 << 12, 1 << 8)
    four.levels += [pg.Level(1 << 7, pg.SCHEME_LINEAR_BUCKETS, 8, 4, pg.PROBE_LINEAR, pg.HASH_FMIX64, 4000),
                    pg.Level(1 << 6, pg.SCHEME_LINEAR_BUCKETS, 8, 4, pg.PROBE_LINEAR, pg.HASH_FMIX64, 4000)]
    t4, n4 = filled(four, 4300, 4)
    a, b = fills(t4)
    assert a == b == n4
    a, b = fills(big)            # and back to the big table
    assert a == b == nbig


def test_fill_noalloc_rejects_missing_arguments():
    lib = pg.lib()
    out = ctypes.c_uint64()
    assert lib.pg_fill_noalloc(None, ctypes.byref(out), pg._stream(None)) == 1  # PG_ERR_INVALID
    t = pg.Table(pg.tcf_desc(1 << 10))
    assert lib.pg_fill_noalloc(t._h, None, pg._stream(None)) == 1
