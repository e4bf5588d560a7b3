"""poggers_amd -- MI355X-native bulk hash tables and Two-Choice Filter.

A thin ctypes binding over the C ABI in ``include/poggers_amd.h``
(``libpoggers_amd.so``, hand-written HIP for gfx950).  PyTorch is only used
for device buffers and streams.  There is no CPU fallback: if the shared
library is missing every entry point raises.

The naming follows the reference's table API (tables/bucketed_table.cuh:
``generate_on_device``, ``insert``, ``query``, ``remove``,
``insert_with_delete``, ``get_fill``, ``get_num_slots``,
``host_bytes_in_use``) with the bulk forms the C ABI provides.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpoggers_amd.so")

# enum values of include/poggers_amd.h
PG_OK = 0
REP_KV64, REP_FP16_V16, REP_FP8_V8, REP_FP10_V6, REP_FP12, REP_KEY64, REP_KV16 = 0, 1, 2, 3, 4, 5, 6
SCHEME_BUCKET, SCHEME_SINGLE_SLOT, SCHEME_SHORTCUT, SCHEME_LINEAR_BUCKETS, SCHEME_GROUPED = 0, 1, 2, 3, 4
SCHEME_POWER_OF_N, SCHEME_POWER_OF_N_SHORTCUT_SLOTS, SCHEME_POWER_OF_N_SHORTCUT_REPLACE = 5, 6, 7
PROBE_LINEAR, PROBE_DOUBLE, PROBE_P2, PROBE_XOR_P2, PROBE_SPLIT64 = 0, 1, 2, 3, 4
HASH_FMIX64, HASH_MURMUR64A = 0, 1
CONCURRENT, DETERMINISTIC = 0, 1
MAX_LEVELS = 4

# Every symbol the header declares (checked by the CPU test suite).
EXPORTS = [
    "pg_create", "pg_destroy", "pg_clear", "pg_set_mode", "pg_bulk_insert",
    "pg_bulk_insert_with_delete", "pg_bulk_query", "pg_bulk_remove", "pg_fill",
    "pg_bytes", "pg_num_slots", "pg_level_ptr", "pg_level_sync", "pg_error_string", "pg_gen_keys",
    "pg_random_sector_bench", "pg_fastmod_host", "pg_set_stats", "pg_set_items",
    "pg_bulk_insert_if_not_exists", "pg_bulk_insert_if_not_exists_delete",
    "pg_batch_create", "pg_batch_destroy", "pg_batch_set_mode", "pg_batch_set_launch", "pg_batch_set_items",
    "pg_batch_clear",
    "pg_batch_insert", "pg_batch_query", "pg_batch_capture", "pg_batch_replay", "pg_batch_fill",
    "pg_batch_level_ptr", "pg_batch_bytes",
    "pg_server_create", "pg_server_destroy", "pg_server_submit", "pg_server_progress", "pg_server_pending",
    "pg_server_run_plain", "pg_server_set_batch", "pg_server_set_order", "pg_server_submit_many", "pg_save",
    "pg_load", "pg_file_info", "pg_fill_noalloc",
]
OP_INSERT, OP_QUERY, OP_REMOVE = 0, 1, 2
ORDER_NONE, ORDER_TABLE = 0, 1


class PgError(RuntimeError):
    pass


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


class Request(ctypes.Structure):
    """pg_request (include/poggers_amd.h)."""
    _fields_ = [
        ("table", ctypes.c_uint32),
        ("op", ctypes.c_uint32),
        ("keys", ctypes.c_void_p),
        ("vals", ctypes.c_void_p),
        ("n", ctypes.c_uint64),
        ("vals_out", ctypes.c_void_p),
        ("flags", ctypes.c_void_p),
    ]


class Desc(ctypes.Structure):
    _fields_ = [
        ("rep", ctypes.c_uint32),
        ("num_levels", ctypes.c_uint32),
        ("level", LevelDesc * MAX_LEVELS),
    ]


_lib = None


def lib():
    """Load libpoggers_amd.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PgError(f"{LIB_PATH} is missing: run `python -m poggers_amd.build` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    u64, u32, vp, i32 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int
    L.pg_create.argtypes = [ctypes.POINTER(Desc), u64, ctypes.POINTER(vp)]
    L.pg_destroy.argtypes = [vp]
    L.pg_clear.argtypes = [vp, vp]
    L.pg_set_mode.argtypes = [vp, u32]
    L.pg_bulk_insert.argtypes = [vp, vp, vp, u64, vp, vp, vp]
    L.pg_bulk_insert_with_delete.argtypes = [vp, vp, vp, u64, vp, vp, vp]
    L.pg_bulk_query.argtypes = [vp, vp, u64, vp, vp, vp]
    L.pg_bulk_remove.argtypes = [vp, vp, u64, vp, vp]
    L.pg_fill.argtypes = [vp, ctypes.POINTER(u64), vp]
    L.pg_fill_noalloc.argtypes = [vp, ctypes.POINTER(u64), vp]
    L.pg_bytes.argtypes = [vp, ctypes.POINTER(u64)]
    L.pg_num_slots.argtypes = [vp, ctypes.POINTER(u64)]
    L.pg_level_ptr.argtypes = [vp, u32, ctypes.POINTER(vp), ctypes.POINTER(u64)]
    L.pg_level_sync.argtypes = [vp, u32, vp]
    L.pg_error_string.argtypes = [i32]
    L.pg_error_string.restype = ctypes.c_char_p
    L.pg_gen_keys.argtypes = [u64, u64, u64, vp, vp]
    L.pg_random_sector_bench.argtypes = [vp, u64, u64, u32, u64, vp, vp]
    L.pg_set_stats.argtypes = [vp, vp]
    L.pg_set_items.argtypes = [vp, u32, u32]
    L.pg_batch_create.argtypes = [ctypes.POINTER(Desc), u32, vp, vp, u64, ctypes.POINTER(vp)]
    L.pg_batch_destroy.argtypes = [vp]
    L.pg_batch_set_mode.argtypes = [vp, u32]
    L.pg_batch_set_launch.argtypes = [vp, u32, u32]
    L.pg_batch_set_items.argtypes = [vp, u32]
    L.pg_batch_clear.argtypes = [vp, vp]
    L.pg_batch_insert.argtypes = [vp, vp, vp, u64, vp, vp, vp, vp]
    L.pg_batch_query.argtypes = [vp, vp, vp, u64, vp, vp, vp]
    L.pg_batch_capture.argtypes = [vp, vp, vp, u64, vp, vp, u64, vp, vp, vp, u32, vp]
    L.pg_batch_replay.argtypes = [vp, vp]
    L.pg_batch_fill.argtypes = [vp, u32, ctypes.POINTER(u64), vp]
    L.pg_batch_level_ptr.argtypes = [vp, u32, u32, ctypes.POINTER(vp), ctypes.POINTER(u64)]
    L.pg_batch_bytes.argtypes = [vp, ctypes.POINTER(u64)]
    L.pg_server_create.argtypes = [vp, u32, u64, u32, ctypes.POINTER(vp)]
    L.pg_server_destroy.argtypes = [vp]
    L.pg_server_submit.argtypes = [vp, u32, u32, vp, vp, u64, vp, vp, ctypes.POINTER(u64)]
    L.pg_server_progress.argtypes = [vp, vp, u32, ctypes.POINTER(u32)]
    L.pg_server_pending.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    L.pg_server_run_plain.argtypes = [vp, u32, u32, vp, vp, u64, vp, vp]
    L.pg_server_set_batch.argtypes = [vp, u64, u32]
    L.pg_server_set_order.argtypes = [vp, u32]
    L.pg_server_submit_many.argtypes = [vp, vp, u32, vp, vp]
    L.pg_save.argtypes = [vp, ctypes.c_char_p]
    L.pg_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
    L.pg_file_info.argtypes = [ctypes.c_char_p, ctypes.POINTER(Desc), ctypes.POINTER(u64)]
    L.pg_bulk_insert_if_not_exists.argtypes = [vp, vp, vp, u64, vp, vp, vp, vp, vp]
    L.pg_bulk_insert_if_not_exists_delete.argtypes = [vp, vp, vp, u64, vp, vp, vp, vp, vp]
    L.pg_fastmod_host.argtypes = [u64, u64]
    L.pg_fastmod_host.restype = u64
    for name in EXPORTS:
        if name not in ("pg_error_string", "pg_fastmod_host"):
            getattr(L, name).restype = i32
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != PG_OK:
        msg = lib().pg_error_string(rc).decode()
        raise PgError(f"{what}: {msg} (status {rc})")


def _ptr(t):
    """Device pointer of a tensor (or None)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def value_bytes(rep: int) -> int:
    """Bytes per element of a representation's value buffers (vals, vals_out)."""
    if rep == REP_KV64:
        return 8
    if rep in (REP_FP8_V8, REP_FP10_V6):
        return 1
    return 2


def _arg(t, what: str, n: int, itemsize: int, device: bool = True):
    """Check one tensor argument of a bulk call before its pointer reaches a
    kernel: a torch tensor, contiguous, `itemsize` bytes per element, at least
    `n` elements, and on the GPU when `device` (else on the GPU or in host
    memory).  Returns its pointer (None for None)."""
    if t is None:
        return None
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{what}: expected a torch tensor, got {type(t).__name__}")
    if device and not t.is_cuda:
        raise ValueError(f"{what}: expected a GPU tensor, got one on {t.device}")
    if not device and t.device.type not in ("cuda", "cpu"):
        raise ValueError(f"{what}: expected a GPU or host tensor, got one on {t.device}")
    if not t.is_contiguous():
        raise ValueError(f"{what}: expected a contiguous tensor")
    if t.element_size() != itemsize:
        raise ValueError(f"{what}: expected {itemsize}-byte elements, got {t.dtype}")
    if t.numel() < n:
        raise ValueError(f"{what}: holds {t.numel()} elements, the call needs {n}")
    return ctypes.c_void_p(t.data_ptr())


def _keys(keys, device: bool = True):
    """Pointer and count of a key tensor (8-byte elements)."""
    p = _arg(keys, "keys", 0, 8, device)
    if p is None:
        raise ValueError("keys: a tensor is required")
    return p, keys.numel()


def _stream(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


@dataclass
class Level:
    nslots: int
    scheme: int
    bucket_size: int
    tile: int = 4
    prober: int = PROBE_SPLIT64
    hasher: int = HASH_FMIX64
    max_probes: int = 2
    group: int = 0  # SCHEME_GROUPED: buckets per 128-byte line

    def as_dict(self):
        return dict(nslots=int(self.nslots), scheme=self.scheme, bucket_size=self.bucket_size,
                    tile=self.tile, prober=self.prober, hasher=self.hasher, max_probes=self.max_probes,
                    group=self.group)


@dataclass
class TableDesc:
    rep: int
    levels: list = field(default_factory=list)

    def to_c(self) -> Desc:
        d = Desc()
        d.rep = self.rep
        d.num_levels = len(self.levels)
        for i, lv in enumerate(self.levels):
            for k, v in lv.as_dict().items():
                setattr(d.level[i], k, v)
        return d


def tcf_desc(nslots: int, backing_slots: int | None = None, bucket_size: int = 16, tile: int = 4,
             prober: int = PROBE_SPLIT64, hasher: int = HASH_FMIX64, backing_bucket: int = 8,
             backing_prober: int = PROBE_LINEAR, backing_probes: int = 4000,
             rep: int = REP_FP16_V16) -> TableDesc:
    """Two-Choice Filter: a POWER_OF_N_SHORTCUT level plus a LINEAR_BUCKETS
    backing level (the reference's del_TCF composition,
    tests/src/delete_tests.cu:138-152).  ``rep`` picks the slot: 16-bit
    fingerprints and values (default), 8/8 or 10/6.  By default the backing
    level holds 1 % of the primary's slots."""
    if backing_slots is None:
        backing_slots = nslots // 100
    levels = [Level(nslots, SCHEME_SHORTCUT, bucket_size, tile, prober, hasher, 2)]
    if backing_slots:
        levels.append(Level(backing_slots, SCHEME_LINEAR_BUCKETS, backing_bucket, tile, backing_prober,
                            hasher, backing_probes))
    return TableDesc(rep, levels)


def kv_desc(nslots: int, bucket_size: int = 1, tile: int = 4, prober: int = PROBE_LINEAR,
            hasher: int = HASH_FMIX64, max_probes: int | None = None, scheme: int = SCHEME_BUCKET,
            backing: Level | None = None) -> TableDesc:
    """Key-value table of key_val_pair<uint64,uint64> slots (bucket_insert)."""
    if max_probes is None:
        nb = (nslots + bucket_size - 1) // bucket_size
        max_probes = nb if prober == PROBE_LINEAR else 20
    levels = [Level(nslots, scheme, bucket_size, tile, prober, hasher, int(min(max_probes, 2**32 - 1)))]
    if backing is not None:
        levels.append(backing)
    return TableDesc(REP_KV64, levels)


def key_only_desc(nslots: int, bucket_size: int = 4, tile: int = 1, prober: int = PROBE_DOUBLE,
                  hasher: int = HASH_MURMUR64A, max_probes: int = 20) -> TableDesc:
    """Key-only table of 8-byte keys (dynamic_container<key_container, uint64_t>);
    the defaults are the CMS allocator's free table (allocators/cms.cuh:203:
    tile 1, 4-slot buckets, bucket_insert, double hashing, 20 probes)."""
    return TableDesc(REP_KEY64, [Level(nslots, SCHEME_BUCKET, bucket_size, tile, prober, hasher, max_probes)])


def flat_tcf_desc(nslots: int, backing_slots: int | None = None, tile: int = 4) -> TableDesc:
    """The reference's flat TCF (data_structs/tcf.cuh:29-30): 16/16
    key_val_pair slots (REP_KV16) in 16-slot buckets under
    power_of_n_insert_shortcut_scheme with two double-hashing probes, backed
    by tiny_static_table_4 (4-slot buckets, bucket_insert, 20 double-hashing
    probes).  Without backing_slots the backing level gets nslots / 100 slots
    (at least one bucket)."""
    if backing_slots is None:
        backing_slots = max(4, nslots // 100)
    return TableDesc(REP_KV16, [
        Level(nslots, SCHEME_POWER_OF_N_SHORTCUT_SLOTS, 16, tile, PROBE_DOUBLE, HASH_MURMUR64A, 2),
        Level(backing_slots, SCHEME_BUCKET, 4, tile, PROBE_DOUBLE, HASH_MURMUR64A, 20)])


class Table:
    """A device-resident table (``pg_create``).  Key/value/result arguments
    are torch CUDA tensors: keys int64 (uint64 bit patterns), values int64
    (KV), int16 (16/16 filters) or uint8 (8/8 and 10/6 filters), flags uint8."""

    def __init__(self, desc: TableDesc, seed: int = 42):
        self.desc = desc
        self.seed = seed
        self._cdesc = desc.to_c()
        h = ctypes.c_void_p()
        check(lib().pg_create(ctypes.byref(self._cdesc), seed, ctypes.byref(h)), "pg_create")
        self._h = h

    # reference-style constructor name (tables/bucketed_table.cuh:132)
    @classmethod
    def generate_on_device(cls, desc: TableDesc, seed: int = 42) -> "Table":
        return cls(desc, seed)

    def free_on_device(self):
        if getattr(self, "_h", None):
            check(lib().pg_destroy(self._h), "pg_destroy")
            self._h = None

    def __del__(self):
        try:
            self.free_on_device()
        except Exception:
            pass

    @property
    def is_kv(self) -> bool:
        """True for key_val_pair<u64,u64> tables, whose value buffers are
        64-bit; the other tables (the 16/16 key-value slots included) take
        16- or 8-bit value buffers."""
        return self.desc.rep == REP_KV64

    def set_mode(self, mode: int):
        check(lib().pg_set_mode(self._h, mode), "pg_set_mode")

    def clear(self, stream=None):
        check(lib().pg_clear(self._h, _stream(stream)), "pg_clear")

    def _vb(self):
        return value_bytes(self.desc.rep)

    def bulk_insert(self, keys, vals=None, status=None, fail_count=None, stream=None, with_delete=False):
        fn = lib().pg_bulk_insert_with_delete if with_delete else lib().pg_bulk_insert
        k, n = _keys(keys)
        check(fn(self._h, k, _arg(vals, "vals", n, self._vb()), n, _arg(status, "status", n, 1),
                 _arg(fail_count, "fail_count", 1, 8), _stream(stream)), "pg_bulk_insert")

    def bulk_insert_with_delete(self, keys, vals=None, status=None, fail_count=None, stream=None):
        self.bulk_insert(keys, vals, status, fail_count, stream, with_delete=True)

    def bulk_insert_if_not_exists(self, keys, vals=None, vals_out=None, found=None, status=None,
                                  fail_count=None, stream=None, delete=False):
        fn = lib().pg_bulk_insert_if_not_exists_delete if delete else lib().pg_bulk_insert_if_not_exists
        k, n = _keys(keys)
        vb = self._vb()
        check(fn(self._h, k, _arg(vals, "vals", n, vb), n, _arg(vals_out, "vals_out", n, vb),
                 _arg(found, "found", n, 1), _arg(status, "status", n, 1), _arg(fail_count, "fail_count", 1, 8),
                 _stream(stream)), "pg_bulk_insert_if_not_exists")

    def bulk_query(self, keys, vals_out=None, found=None, stream=None):
        k, n = _keys(keys)
        check(lib().pg_bulk_query(self._h, k, n, _arg(vals_out, "vals_out", n, self._vb()),
                                  _arg(found, "found", n, 1), _stream(stream)), "pg_bulk_query")

    def bulk_remove(self, keys, found=None, stream=None):
        k, n = _keys(keys)
        check(lib().pg_bulk_remove(self._h, k, n, _arg(found, "found", n, 1), _stream(stream)),
              "pg_bulk_remove")

    def set_stats(self, stats=None):
        """Accumulate algorithmic counters into a 3-element int64 CUDA tensor
        (None disables)."""
        check(lib().pg_set_stats(self._h, _arg(stats, "stats", 3, 8)), "pg_set_stats")

    def set_items(self, insert_items: int = 0, query_items: int = 0):
        """Keys per tile for concurrent inserts / queries (1, 2, 4; 0 = the
        default, 1).  Results do not depend on it."""
        check(lib().pg_set_items(self._h, insert_items, query_items), "pg_set_items")

    def get_fill(self, stream=None) -> int:
        out = ctypes.c_uint64()
        check(lib().pg_fill(self._h, ctypes.byref(out), _stream(stream)), "pg_fill")
        return out.value

    def host_bytes_in_use(self) -> int:
        out = ctypes.c_uint64()
        check(lib().pg_bytes(self._h, ctypes.byref(out)), "pg_bytes")
        return out.value

    def get_num_slots(self) -> int:
        out = ctypes.c_uint64()
        check(lib().pg_num_slots(self._h, ctypes.byref(out)), "pg_num_slots")
        return out.value

    def level_ptr(self, level: int):
        p = ctypes.c_void_p()
        n = ctypes.c_uint64()
        check(lib().pg_level_ptr(self._h, level, ctypes.byref(p), ctypes.byref(n)), "pg_level_ptr")
        return p.value, n.value

    def level_bytes(self, level: int) -> bytes:
        """Raw copy of one level's slot array (device -> host)."""
        import torch
        ptr, n = self.level_ptr(level)
        torch.cuda.synchronize()
        buf = ctypes.create_string_buffer(n)
        _hip_memcpy(buf, ptr, n, 2)  # hipMemcpyDeviceToHost
        return buf.raw

    def save(self, path: str):
        """Write the table (descriptor, seed, raw slot arrays) to `path`."""
        check(lib().pg_save(self._h, path.encode()), "pg_save")

    @classmethod
    def load(cls, path: str) -> "Table":
        """Recreate a table saved with save()."""
        desc, seed = file_info(path)
        t = cls.__new__(cls)
        h = ctypes.c_void_p()
        check(lib().pg_load(path.encode(), ctypes.byref(h)), "pg_load")
        t.desc, t.seed, t._h = desc, seed, h
        t._cdesc = desc.to_c()
        return t

    def level_upload(self, level: int, data: bytes):
        """Overwrite one level's slot array with a raw image (host -> device)."""
        import torch
        ptr, n = self.level_ptr(level)
        if len(data) != n:
            raise PgError(f"image is {len(data)} bytes, level {level} holds {n}")
        torch.cuda.synchronize()
        buf = ctypes.create_string_buffer(data, n)
        _hip_memcpy(ptr, ctypes.addressof(buf), n, 1)  # hipMemcpyHostToDevice
        check(lib().pg_level_sync(self._h, level, None), "pg_level_sync")
        torch.cuda.synchronize()


class Batch:
    """Many independent filters sharing one shape, driven by one persistent
    kernel per bulk call (``pg_batch_*``).  ``proto`` is a TableDesc whose
    level sizes are ignored; ``nslots`` lists each table's level-0 slots."""

    def __init__(self, proto: TableDesc, nslots, backing_slots=None, seed: int = 42):
        import numpy as np
        self.proto = proto
        self.ntables = len(nslots)
        self._ns = np.ascontiguousarray(nslots, dtype=np.uint64)
        self._bs = None if backing_slots is None else np.ascontiguousarray(backing_slots, dtype=np.uint64)
        self._cdesc = proto.to_c()
        h = ctypes.c_void_p()
        check(lib().pg_batch_create(ctypes.byref(self._cdesc), self.ntables,
                                    self._ns.ctypes.data_as(ctypes.c_void_p),
                                    None if self._bs is None else self._bs.ctypes.data_as(ctypes.c_void_p),
                                    seed, ctypes.byref(h)), "pg_batch_create")
        self._h = h

    def free(self):
        if getattr(self, "_h", None):
            check(lib().pg_batch_destroy(self._h), "pg_batch_destroy")
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def set_mode(self, mode: int):
        check(lib().pg_batch_set_mode(self._h, mode), "pg_batch_set_mode")

    def set_launch(self, chunk: int, blocks: int = 0):
        check(lib().pg_batch_set_launch(self._h, chunk, blocks), "pg_batch_set_launch")

    def set_items(self, items: int):
        """Keys each tile keeps in flight in the persistent kernels (1, 2, 4)."""
        check(lib().pg_batch_set_items(self._h, items), "pg_batch_set_items")

    def clear(self, stream=None):
        check(lib().pg_batch_clear(self._h, _stream(stream)), "pg_batch_clear")

    def _offsets(self, offsets, what):
        p = _arg(offsets, what, self.ntables + 1, 8)
        if p is None:
            raise ValueError(f"{what}: a tensor is required")
        return p

    def insert(self, keys, offsets, vals=None, status=None, fail_count=None, stream=None):
        k, n = _keys(keys)
        vb = value_bytes(self.proto.rep)
        check(lib().pg_batch_insert(self._h, k, self._offsets(offsets, "offsets"), n, _arg(vals, "vals", n, vb),
                                    _arg(status, "status", n, 1), _arg(fail_count, "fail_count", 1, 8),
                                    _stream(stream)), "pg_batch_insert")

    def query(self, keys, offsets, vals_out=None, found=None, stream=None):
        k, n = _keys(keys)
        vb = value_bytes(self.proto.rep)
        check(lib().pg_batch_query(self._h, k, self._offsets(offsets, "offsets"), n,
                                   _arg(vals_out, "vals_out", n, vb), _arg(found, "found", n, 1),
                                   _stream(stream)), "pg_batch_query")

    def capture(self, ins_keys, ins_offsets, q_keys, q_offsets, status=None, found=None, fail_count=None,
                clear_first=True, stream=None):
        """Record clear (optional) + insert + query in a hipGraph that
        replay() relaunches.  The graph records the tensors' addresses, so
        the Batch keeps them until the next capture; refill them in place
        between replays."""
        ik, ni = _keys(ins_keys)
        qk, nq = _keys(q_keys)
        check(lib().pg_batch_capture(self._h, ik, self._offsets(ins_offsets, "ins_offsets"), ni, qk,
                                     self._offsets(q_offsets, "q_offsets"), nq, _arg(status, "status", ni, 1),
                                     _arg(found, "found", nq, 1), _arg(fail_count, "fail_count", 1, 8),
                                     int(clear_first), _stream(stream)), "pg_batch_capture")
        # the graph holds these tensors' addresses: keep them for its replays
        self._graph_tensors = (ins_keys, ins_offsets, q_keys, q_offsets, status, found, fail_count)

    def replay(self, stream=None):
        check(lib().pg_batch_replay(self._h, _stream(stream)), "pg_batch_replay")

    def fill(self, table: int, stream=None) -> int:
        out = ctypes.c_uint64()
        check(lib().pg_batch_fill(self._h, table, ctypes.byref(out), _stream(stream)), "pg_batch_fill")
        return out.value

    def level_bytes(self, table: int, level: int) -> bytes:
        import torch
        p = ctypes.c_void_p()
        n = ctypes.c_uint64()
        check(lib().pg_batch_level_ptr(self._h, table, level, ctypes.byref(p), ctypes.byref(n)), "level_ptr")
        torch.cuda.synchronize()
        buf = ctypes.create_string_buffer(n.value)
        _hip_memcpy(buf, p.value, n.value, 2)
        return buf.raw

    def bytes(self) -> int:
        out = ctypes.c_uint64()
        check(lib().pg_batch_bytes(self._h, ctypes.byref(out)), "pg_batch_bytes")
        return out.value


class Server:
    """Continuous-batching request loop over resident tables (``pg_server_*``).
    Request buffers are host tensors, ideally pinned (``pin_memory=True``)."""

    def __init__(self, tables, budget_bytes: int = 1 << 30, nstreams: int = 4):
        self.tables = list(tables)
        arr = (ctypes.c_void_p * len(self.tables))(*[t._h.value for t in self.tables])
        h = ctypes.c_void_p()
        check(lib().pg_server_create(arr, len(self.tables), budget_bytes, nstreams, ctypes.byref(h)),
              "pg_server_create")
        self._h = h
        self._done = (ctypes.c_uint64 * 65536)()
        # ticket -> the request's tensors (or its packed array), held until
        # progress() reports the ticket: the server reads and writes them
        # in place, so they must outlive the request
        self._live = {}

    def free(self):
        if getattr(self, "_h", None):
            try:
                check(lib().pg_server_destroy(self._h), "pg_server_destroy")  # waits for in-flight work
            finally:
                self._h = None
                self._live.clear()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @staticmethod
    def _request(tables, table, op, keys, vals, vals_out, flags):
        """Pointers of one request's tensors, checked (GPU or host memory,
        contiguous, element widths of the table's representation when
        `tables` is known, at least as many elements as keys)."""
        k, n = _keys(keys, device=False)
        if tables is not None:
            if not 0 <= table < len(tables):
                raise ValueError(f"table {table} out of range")
            vb = value_bytes(tables[table].desc.rep)
        else:
            vb = None
        def val(t, what):
            if t is None:
                return None
            return _arg(t, what, n, vb if vb is not None else t.element_size(), device=False)
        return k, n, val(vals, "vals"), val(vals_out, "vals_out"), _arg(flags, "flags", n, 1, device=False)

    def submit(self, table: int, op: int, keys, vals=None, vals_out=None, flags=None) -> int:
        t = ctypes.c_uint64()
        k, n, v, o, f = self._request(self.tables, table, op, keys, vals, vals_out, flags)
        check(lib().pg_server_submit(self._h, table, op, k, v, n, o, f, ctypes.byref(t)), "pg_server_submit")
        self._live[t.value] = (keys, vals, vals_out, flags)
        return t.value

    @staticmethod
    def pack(requests, tables=None):
        """A pg_request array for submit_many from (table, op, keys, vals,
        vals_out, flags) tuples of tensors (None where absent).  Each tensor is
        checked; pass the server's `tables` to check value widths too."""
        arr = (Request * len(requests))()
        arr._tensors = [tuple(r[2:]) for r in requests]  # kept alive with the array
        for i, (table, op, keys, vals, out, flags) in enumerate(requests):
            k, n, v, o, f = Server._request(tables, table, op, keys, vals, out, flags)
            r = arr[i]
            r.table, r.op, r.n = table, op, n
            r.keys = k.value
            r.vals = v.value if v is not None else None
            r.vals_out = o.value if o is not None else None
            r.flags = f.value if f is not None else None
        return arr

    def submit_many(self, packed, start: int = 0, count: int | None = None):
        """Queue packed[start:start+count] in order; returns their tickets."""
        if count is None:
            count = len(packed) - start
        if count <= 0:
            return []
        tickets = (ctypes.c_uint64 * count)()
        nq = ctypes.c_uint32()
        ptr = ctypes.cast(ctypes.byref(packed, start * ctypes.sizeof(Request)), ctypes.c_void_p)
        rc = lib().pg_server_submit_many(self._h, ptr, count, tickets, ctypes.byref(nq))
        for tk in list(tickets)[:nq.value]:   # the requests queued (all of them on success)
            self._live[tk] = packed
        check(rc, "pg_server_submit_many")
        return list(tickets)

    def progress(self, cap: int = 65536):
        """One engine step; returns up to `cap` tickets that finished (the
        rest are returned by later steps)."""
        nd = ctypes.c_uint32()
        check(lib().pg_server_progress(self._h, self._done, min(cap, 65536), ctypes.byref(nd)),
              "pg_server_progress")
        done = [self._done[i] for i in range(nd.value)]
        for tk in done:
            self._live.pop(tk, None)
        return done

    def set_batch(self, max_batch_keys: int, max_batch_requests: int = 4096):
        check(lib().pg_server_set_batch(self._h, max_batch_keys, max_batch_requests), "pg_server_set_batch")

    def set_order(self, order: int):
        """ORDER_TABLE: on each table, writes take effect in submission order
        relative to every other request of that table; ORDER_NONE (default):
        requests in flight together run concurrently."""
        check(lib().pg_server_set_order(self._h, order), "pg_server_set_order")

    def pending(self):
        q, f = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().pg_server_pending(self._h, ctypes.byref(q), ctypes.byref(f)), "pg_server_pending")
        return q.value, f.value

    def run_plain(self, table: int, op: int, keys, vals=None, vals_out=None, flags=None):
        k, n, v, o, f = self._request(self.tables, table, op, keys, vals, vals_out, flags)
        check(lib().pg_server_run_plain(self._h, table, op, k, v, n, o, f), "pg_server_run_plain")


_hip = None


def _hip_memcpy(dst, src, n: int, kind: int):
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
    d = dst if isinstance(dst, ctypes.Array) else ctypes.c_void_p(dst)
    rc = _hip.hipMemcpy(d, ctypes.c_void_p(src), n, kind)
    if rc != 0:
        raise PgError(f"hipMemcpy failed ({rc})")


def file_info(path: str):
    """(TableDesc, seed) of a table file, read on the host only."""
    d = Desc()
    seed = ctypes.c_uint64()
    check(lib().pg_file_info(path.encode(), ctypes.byref(d), ctypes.byref(seed)), "pg_file_info")
    levels = [Level(d.level[i].nslots, d.level[i].scheme, d.level[i].bucket_size, d.level[i].tile,
                    d.level[i].prober, d.level[i].hasher, d.level[i].max_probes, d.level[i].group)
              for i in range(d.num_levels)]
    return TableDesc(d.rep, levels), seed.value


def gen_keys(seed: int, start: int, n: int, out, stream=None):
    """Fill ``out`` (int64 CUDA tensor) with splitmix64 keys start..start+n-1."""
    check(lib().pg_gen_keys(seed, start, n, _arg(out, "out", n, 8), _stream(stream)), "pg_gen_keys")


def random_sector_bench(buf, n_reads: int, op: int, seed: int, sink, stream=None):
    p = _arg(buf, "buf", 1, buf.element_size())
    check(lib().pg_random_sector_bench(p, buf.numel() * buf.element_size(), n_reads, op, seed,
                                       _arg(sink, "sink", 1, 8), _stream(stream)), "pg_random_sector_bench")


def fastmod_host(n: int, d: int) -> int:
    return lib().pg_fastmod_host(n, d)
