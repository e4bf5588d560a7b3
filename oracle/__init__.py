"""ctypes front of the CPU oracle (``pg_oracle.c``).

Test infrastructure: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker or the timed CPU baseline -- never as the product path.

The oracle restates poggers' sequential semantics (see the citations in
``pg_oracle.c``).  Descriptors use the same enum values as the product C-ABI
so a test can describe one table once and build it on both sides.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

REP_KV64, REP_FP16_V16, REP_FP8_V8, REP_FP10_V6, REP_FP12, REP_KEY64, REP_KV16 = 0, 1, 2, 3, 4, 5, 6
SCHEME_BUCKET, SCHEME_SINGLE_SLOT, SCHEME_SHORTCUT, SCHEME_LINEAR_BUCKETS, SCHEME_GROUPED = 0, 1, 2, 3, 4
SCHEME_POWER_OF_N, SCHEME_POWER_OF_N_SHORTCUT_SLOTS, SCHEME_POWER_OF_N_SHORTCUT_REPLACE = 5, 6, 7
PROBE_LINEAR, PROBE_DOUBLE, PROBE_P2, PROBE_XOR_P2, PROBE_SPLIT64 = 0, 1, 2, 3, 4
HASH_FMIX64, HASH_MURMUR64A = 0, 1


class LevelDesc(ctypes.Structure):
    _fields_ = [
        ("nslots", ctypes.c_uint64),
        ("scheme", ctypes.c_uint32),
        ("bucket_size", ctypes.c_uint32),
        ("tile", ctypes.c_uint32),
        ("prober", ctypes.c_uint32),
        ("hasher", ctypes.c_uint32),
        ("max_probes", ctypes.c_uint32),
        ("group", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


def build() -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u64, u32, vp, i32 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int
        L.or_murmur64a.restype = u64
        L.or_murmur64a.argtypes = [vp, i32, u64]
        L.or_fmix64.restype = u64
        L.or_fmix64.argtypes = [u64]
        L.or_murmur3_x64_128.restype = None
        L.or_murmur3_x64_128.argtypes = [vp, i32, u32, vp]
        L.or_hash.restype = u64
        L.or_hash.argtypes = [u32, u64, u64]
        L.or_probes.restype = i32
        L.or_probes.argtypes = [u32, u32, u32, u64, u64, u64, vp, i32]
        L.or_pack.restype = u32
        L.or_pack.argtypes = [u32, u32, u32, u32]
        L.or_unpack_fp.restype = u32
        L.or_unpack_fp.argtypes = [u32, u32]
        L.or_unpack_val.restype = u32
        L.or_unpack_val.argtypes = [u32, u32, u32]
        L.or_create.restype = vp
        L.or_create.argtypes = [u32, u32, ctypes.POINTER(LevelDesc), u64]
        L.or_destroy.restype = None
        L.or_destroy.argtypes = [vp]
        L.or_clear.restype = None
        L.or_clear.argtypes = [vp]
        L.or_insert.restype = i32
        L.or_insert.argtypes = [vp, u64, u64, i32]
        L.or_query.restype = i32
        L.or_query.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_uint64)]
        L.or_remove.restype = i32
        L.or_remove.argtypes = [vp, u64]
        L.or_bulk_insert.restype = None
        L.or_bulk_insert.argtypes = [vp, vp, vp, u32, u64, vp, i32]
        L.or_bulk_query.restype = None
        L.or_bulk_query.argtypes = [vp, vp, u64, vp, u32, vp]
        L.or_bulk_remove.restype = None
        L.or_bulk_remove.argtypes = [vp, vp, u64, vp]
        L.or_bulk_insert_mt.restype = i32
        L.or_bulk_insert_mt.argtypes = [vp, vp, vp, u64, vp, u32]
        L.or_bulk_query_mt.restype = i32
        L.or_bulk_query_mt.argtypes = [vp, vp, u64, vp, vp, u32]
        L.or_hw_threads.restype = u32
        L.or_hw_threads.argtypes = []
        L.or_insert_if_not_exists.restype = i32
        L.or_insert_if_not_exists.argtypes = [vp, u64, u64, i32, ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_uint64)]
        L.or_bulk_insert_if_not_exists.restype = None
        L.or_bulk_insert_if_not_exists.argtypes = [vp, vp, vp, u64, i32, vp, vp, vp]
        L.or_fill.restype = u64
        L.or_fill.argtypes = [vp]
        L.or_level_bytes.restype = u64
        L.or_level_bytes.argtypes = [vp, u32]
        L.or_level_data.restype = vp
        L.or_level_data.argtypes = [vp, u32]
        L.or_level_buckets.restype = u64
        L.or_level_buckets.argtypes = [vp, u32]
        L.or_stats.restype = None
        L.or_stats.argtypes = [vp, vp]
        L.or_stats_reset.restype = None
        L.or_stats_reset.argtypes = [vp]
        L.or_key_at.restype = u64
        L.or_key_at.argtypes = [u64, u64]
        L.or_gen_keys.restype = None
        L.or_gen_keys.argtypes = [u64, u64, u64, vp]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def murmur64a(data: bytes, seed: int) -> int:
    buf = ctypes.create_string_buffer(data, len(data))
    return lib().or_murmur64a(buf, len(data), seed)


def murmur64a_u64(key: int, seed: int) -> int:
    return murmur64a(int(key).to_bytes(8, "little"), seed)


def fmix64(k: int) -> int:
    return lib().or_fmix64(k)


def murmur3_x64_128(data: bytes, seed: int) -> bytes:
    buf = ctypes.create_string_buffer(data, len(data))
    out = ctypes.create_string_buffer(16)
    lib().or_murmur3_x64_128(buf, len(data), seed, out)
    return out.raw


def hash64(hasher: int, key: int, seed: int) -> int:
    return lib().or_hash(hasher, key, seed)


def probes(prober: int, hasher: int, max_probes: int, seed: int, key: int, tag: int, cap: int = 64):
    out = (ctypes.c_uint64 * cap)()
    n = lib().or_probes(prober, hasher, max_probes, seed, key, tag, out, cap)
    return [out[i] for i in range(n)]


def hw_threads() -> int:
    """Host threads online (sysconf)."""
    return int(lib().or_hw_threads())


def gen_keys(seed: int, start: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    if n:
        lib().or_gen_keys(seed, start, n, _ptr(out))
    return out


def level(nslots, scheme, bucket_size, tile=4, prober=PROBE_SPLIT64, hasher=HASH_FMIX64, max_probes=2,
          group=0):
    return dict(nslots=int(nslots), scheme=scheme, bucket_size=bucket_size, tile=tile,
                prober=prober, hasher=hasher, max_probes=max_probes, group=group)


class Table:
    """A sequential oracle table; ``levels`` is a list of ``level(...)`` dicts."""

    def __init__(self, rep: int, levels, seed: int = 42):
        arr = (LevelDesc * len(levels))()
        for i, lv in enumerate(levels):
            for k, v in lv.items():
                setattr(arr[i], k, v)
        self.rep = rep
        self.levels = levels
        self._h = lib().or_create(rep, len(levels), arr, seed)
        if not self._h:
            raise ValueError("oracle table creation failed")
        self.val_bytes = 8 if rep == REP_KV64 else (1 if rep in (REP_FP8_V8, REP_FP10_V6) else 2)
        # REP_FP12 and REP_KEY64 store no values: queries return 0
        self.val_dtype = {8: np.uint64, 2: np.uint16, 1: np.uint8}[self.val_bytes]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().or_destroy(h)
            self._h = None

    def clear(self):
        lib().or_clear(self._h)

    def insert(self, key, val=0, with_delete=False) -> int:
        return lib().or_insert(self._h, int(key), int(val), int(with_delete))

    def query(self, key):
        v = ctypes.c_uint64(0)
        r = lib().or_query(self._h, int(key), ctypes.byref(v))
        return bool(r), v.value

    def remove(self, key) -> int:
        return lib().or_remove(self._h, int(key))

    def bulk_insert(self, keys: np.ndarray, vals: np.ndarray | None = None, with_delete=False) -> np.ndarray:
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        st = np.zeros(len(keys), dtype=np.uint8)
        if vals is not None:
            vals = np.ascontiguousarray(vals, dtype=self.val_dtype)
        lib().or_bulk_insert(self._h, _ptr(keys), _ptr(vals), self.val_bytes, len(keys), _ptr(st),
                             int(with_delete))
        return st

    def bulk_query(self, keys: np.ndarray):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        found = np.zeros(len(keys), dtype=np.uint8)
        vals = np.zeros(len(keys), dtype=self.val_dtype)
        lib().or_bulk_query(self._h, _ptr(keys), len(keys), _ptr(vals), self.val_bytes, _ptr(found))
        return found, vals

    def bulk_insert_mt(self, keys: np.ndarray, vals: np.ndarray | None = None, threads: int = 0):
        """The same filter insert on `threads` host threads (0 = every core)
        with atomic slot claims; 16/16 slots, shortcut level 0 and optional
        linear backing only.  Returns (status, threads used)."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        st = np.zeros(len(keys), dtype=np.uint8)
        v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint16)
        used = lib().or_bulk_insert_mt(self._h, _ptr(keys), None if v is None else _ptr(v), len(keys),
                                       _ptr(st), threads)
        if used < 0:
            raise ValueError("the multi-threaded run covers 16/16 slots with a shortcut level 0 "
                             "and an optional linear backing level")
        return st, used

    def bulk_query_mt(self, keys: np.ndarray, threads: int = 0):
        """bulk_query on `threads` host threads; returns (found, vals, threads used)."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        found = np.zeros(len(keys), dtype=np.uint8)
        vals = np.zeros(len(keys), dtype=np.uint16)
        used = lib().or_bulk_query_mt(self._h, _ptr(keys), len(keys), _ptr(vals), _ptr(found), threads)
        if used < 0:
            raise ValueError("the multi-threaded run covers 16/16 slots with a shortcut level 0 "
                             "and an optional linear backing level")
        return found, vals, used

    def insert_if_not_exists(self, key, val=0, delete=False):
        f = ctypes.c_int(0)
        v = ctypes.c_uint64(0)
        r = lib().or_insert_if_not_exists(self._h, int(key), int(val), int(delete), ctypes.byref(f),
                                          ctypes.byref(v))
        return r, bool(f.value), v.value

    def bulk_insert_if_not_exists(self, keys: np.ndarray, vals: np.ndarray | None = None, delete=False):
        """Returns (status, found, vals_out) per key."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        n = len(keys)
        st = np.zeros(n, dtype=np.uint8)
        found = np.zeros(n, dtype=np.uint8)
        out = np.zeros(n, dtype=np.uint16)
        if vals is not None:
            vals = np.ascontiguousarray(vals, dtype=np.uint16)
        lib().or_bulk_insert_if_not_exists(self._h, _ptr(keys), _ptr(vals), n, int(delete), _ptr(out),
                                           _ptr(found), _ptr(st))
        return st, found, out

    def bulk_remove(self, keys: np.ndarray) -> np.ndarray:
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        found = np.zeros(len(keys), dtype=np.uint8)
        lib().or_bulk_remove(self._h, _ptr(keys), len(keys), _ptr(found))
        return found

    def fill(self) -> int:
        return lib().or_fill(self._h)

    def level_bytes(self, lvl: int) -> bytes:
        n = lib().or_level_bytes(self._h, lvl)
        p = lib().or_level_data(self._h, lvl)
        return ctypes.string_at(p, n)

    def set_level_bytes(self, lvl: int, data: bytes):
        """Overwrite one level's slot array with a raw image (for instance a
        concurrent GPU layout), so later operations run the sequential
        algorithm over exactly those slots."""
        n = lib().or_level_bytes(self._h, lvl)
        if len(data) != n:
            raise ValueError(f"image is {len(data)} bytes, level {lvl} holds {n}")
        ctypes.memmove(lib().or_level_data(self._h, lvl), data, n)

    def level_array(self, lvl: int) -> np.ndarray:
        dt = {REP_KV64: np.uint64, REP_FP16_V16: np.uint32, REP_FP8_V8: np.uint16,
              REP_FP10_V6: np.uint16, REP_FP12: np.uint16, REP_KEY64: np.uint64, REP_KV16: np.uint32}[self.rep]
        return np.frombuffer(self.level_bytes(lvl), dtype=dt).copy()

    def num_buckets(self, lvl: int) -> int:
        return lib().or_level_buckets(self._h, lvl)

    def stats(self):
        out = np.zeros(4, dtype=np.uint64)
        lib().or_stats(self._h, _ptr(out))
        return dict(sectors_read=int(out[0]), sectors_written=int(out[1]),
                    buckets_read_l0=int(out[2]), buckets_read_l1=int(out[3]))

    def stats_reset(self):
        lib().or_stats_reset(self._h)
