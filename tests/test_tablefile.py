"""Raw bucket-array table files (pg_save / pg_load / pg_file_info)."""
import ctypes
import os

import numpy as np
import pytest

import poggers_amd as pg


class FileHeader(ctypes.Structure):
    _fields_ = [("magic", ctypes.c_char * 8), ("version", ctypes.c_uint32), ("header_bytes", ctypes.c_uint32),
                ("desc", pg.Desc), ("seed", ctypes.c_uint64), ("level_bytes", ctypes.c_uint64 * pg.MAX_LEVELS)]


def test_file_header_is_read_on_the_host(tmp_path):
    desc = pg.tcf_desc(1 << 20, 1 << 10, tile=2)
    h = FileHeader()
    h.magic = b"PGAMDTB1"
    h.version = 3
    h.header_bytes = ctypes.sizeof(FileHeader)
    h.desc = desc.to_c()
    h.seed = 4242
    h.level_bytes[0] = 4 << 20
    h.level_bytes[1] = 4 << 10
    p = str(tmp_path / "t.pgt")
    with open(p, "wb") as f:
        f.write(bytes(h))
    d, seed = pg.file_info(p)
    assert seed == 4242 and d.rep == desc.rep
    assert [lv.as_dict() for lv in d.levels] == [lv.as_dict() for lv in desc.levels]
    bad = bytearray(bytes(h))
    bad[0:8] = b"NOTATABL"
    with open(p, "wb") as f:
        f.write(bad)
    with pytest.raises(pg.PgError):
        pg.file_info(p)


@pytest.mark.gpu
def test_save_load_roundtrip(tmp_path):
    import torch
    import oracle
    from tests.gpu_util import Ops
    for desc in (pg.tcf_desc(1 << 16, 1 << 9, tile=4),
                 pg.kv_desc(1 << 16, bucket_size=16, tile=4, prober=pg.PROBE_DOUBLE, max_probes=20),
                 pg.flat_tcf_desc(1 << 16, 1 << 9),
                 pg.TableDesc(pg.REP_KV64, [pg.Level(1 << 16, pg.SCHEME_POWER_OF_N, 8, 8, pg.PROBE_DOUBLE,
                                                     pg.HASH_MURMUR64A, 3)])):
        t = pg.Table(desc, seed=99)
        ops = Ops(t)
        keys = oracle.gen_keys(5, 0, 50000)
        ops.insert(keys, keys if t.is_kv else (keys & np.uint64(0xFFFF)).astype(np.uint16))
        ops.remove(keys[::7])
        p = str(tmp_path / "t.pgt")
        t.save(p)
        u = pg.Table.load(p)
        assert u.seed == 99
        for lvl in range(len(desc.levels)):
            assert t.level_bytes(lvl) == u.level_bytes(lvl)
        q = np.concatenate([keys, oracle.gen_keys(5, 1 << 40, 20000)])
        f1, v1 = ops.query(q)
        f2, v2 = Ops(u).query(q)
        assert (f1 == f2).all() and (v1 == v2).all()
        assert u.get_fill() == t.get_fill()
