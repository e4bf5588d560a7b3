"""CPU checks of the C ABI library: it loads, exports every symbol the header
declares, and its host-side range reduction is exact.  No GPU calls."""
import os
import random
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "poggers_amd.h")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pg_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_export_list():
    import poggers_amd
    assert header_symbols() == sorted(poggers_amd.EXPORTS)


def test_library_exports_every_symbol():
    import poggers_amd
    lib = poggers_amd.lib()
    for name in header_symbols():
        assert hasattr(lib, name), name


def test_error_strings():
    import poggers_amd
    lib = poggers_amd.lib()
    assert lib.pg_error_string(0) == b"ok"
    assert lib.pg_error_string(2) == b"unsupported by this scheme"


def test_create_rejects_bad_descriptors_without_gpu_work():
    # Validation happens before any device call.
    import ctypes
    import poggers_amd as pg
    bad = [
        pg.TableDesc(pg.REP_FP16_V16, [pg.Level(1024, pg.SCHEME_SHORTCUT, 16, tile=3)]),
        pg.TableDesc(pg.REP_FP16_V16, [pg.Level(1024, pg.SCHEME_SHORTCUT, 16, prober=pg.PROBE_XOR_P2,
                                                max_probes=3)]),
        pg.TableDesc(pg.REP_KV64, [pg.Level(1024, pg.SCHEME_SHORTCUT, 16)]),
        pg.TableDesc(pg.REP_FP16_V16, [pg.Level(0, pg.SCHEME_SHORTCUT, 16)]),
        pg.TableDesc(pg.REP_FP16_V16, []),
        pg.TableDesc(pg.REP_FP16_V16, [pg.Level(1024, pg.SCHEME_SHORTCUT, 16, tile=4),
                                       pg.Level(64, pg.SCHEME_LINEAR_BUCKETS, 8, tile=8,
                                                prober=pg.PROBE_LINEAR, max_probes=100)]),
        # grouped levels: K buckets must fit one 128-byte line, K only with GROUPED
        pg.TableDesc(pg.REP_FP16_V16, [pg.Level(1024, pg.SCHEME_GROUPED, 16, group=3)]),   # 3 * 64 B
        pg.TableDesc(pg.REP_FP12, [pg.Level(1024, pg.SCHEME_GROUPED, 16, group=6)]),      # 6 * 24 B
        pg.TableDesc(pg.REP_FP16_V16, [pg.Level(1024, pg.SCHEME_GROUPED, 16, group=0)]),
        pg.TableDesc(pg.REP_FP16_V16, [pg.Level(1024, pg.SCHEME_SHORTCUT, 16, group=2)]),
        pg.TableDesc(pg.REP_KV64, [pg.Level(1024, pg.SCHEME_GROUPED, 4, group=2)]),
        # twelve_bucket::tag() returns nothing: no tag-driven probing for it
        pg.TableDesc(pg.REP_FP12, [pg.Level(1024, pg.SCHEME_SHORTCUT, 16, prober=pg.PROBE_XOR_P2)]),
        pg.TableDesc(5, [pg.Level(1024, pg.SCHEME_SHORTCUT, 16)]),
        pg.TableDesc(7, [pg.Level(1024, pg.SCHEME_BUCKET, 16)]),
        # the flat power-of-n schemes take key-value slots and at most 4 probes;
        # the 16/16 key-value slot takes no packed-bucket scheme
        pg.TableDesc(pg.REP_FP16_V16, [pg.Level(1024, pg.SCHEME_POWER_OF_N, 16, prober=pg.PROBE_DOUBLE,
                                                max_probes=2)]),
        pg.TableDesc(pg.REP_KV64, [pg.Level(1024, pg.SCHEME_POWER_OF_N, 8, prober=pg.PROBE_DOUBLE,
                                            max_probes=5)]),
        pg.TableDesc(pg.REP_KV16, [pg.Level(1024, pg.SCHEME_POWER_OF_N_SHORTCUT_SLOTS, 16,
                                            prober=pg.PROBE_DOUBLE, max_probes=8)]),
        pg.TableDesc(pg.REP_KV16, [pg.Level(1024, pg.SCHEME_SHORTCUT, 16, prober=pg.PROBE_DOUBLE)]),
        pg.TableDesc(pg.REP_KV16, [pg.Level(1024, pg.SCHEME_POWER_OF_N, 0, prober=pg.PROBE_DOUBLE)]),
        # the replace scheme is over key-value slots only, with at most 4 probes
        pg.TableDesc(pg.REP_FP16_V16, [pg.Level(1024, pg.SCHEME_POWER_OF_N_SHORTCUT_REPLACE, 16,
                                                prober=pg.PROBE_DOUBLE, max_probes=2)]),
        pg.TableDesc(pg.REP_KV64, [pg.Level(1024, pg.SCHEME_POWER_OF_N_SHORTCUT_REPLACE, 16,
                                            prober=pg.PROBE_DOUBLE, max_probes=5)]),
        pg.TableDesc(pg.REP_KV64, [pg.Level(1024, 8, 16, prober=pg.PROBE_DOUBLE, max_probes=2)]),
        # single_slot_insert needs Partition_Size == Bucket_Size (single_slot_insert.cuh:43)
        pg.TableDesc(pg.REP_KV64, [pg.Level(1024, pg.SCHEME_SINGLE_SLOT, 1, tile=4, prober=pg.PROBE_DOUBLE,
                                            max_probes=5)]),
        pg.TableDesc(pg.REP_KEY64, [pg.Level(1024, pg.SCHEME_SINGLE_SLOT, 8, tile=4, prober=pg.PROBE_DOUBLE,
                                             max_probes=5)]),
        # SPLIT64 probes are 32-bit: no level of more than 2^32 buckets
        pg.TableDesc(pg.REP_FP16_V16, [pg.Level((1 << 36) + 16, pg.SCHEME_SHORTCUT, 16,
                                                prober=pg.PROBE_SPLIT64, max_probes=2)]),
    ]
    for d in bad:
        h = ctypes.c_void_p()
        rc = pg.lib().pg_create(ctypes.byref(d.to_c()), 42, ctypes.byref(h))
        assert rc in (1, 2), rc
        assert not h.value


def test_fastmod_host_exact():
    import poggers_amd
    rng = random.Random(17)
    M = (1 << 64) - 1
    divisors = [1, 2, 3, 5, 7, 12, 335545, 1 << 24, (1 << 24) + 1, 1_000_003, (1 << 32) - 1, (1 << 32) + 15,
                (1 << 63) - 25, (1 << 63) + 1, M, M - 1]
    divisors += [rng.getrandbits(rng.randint(1, 64)) or 1 for _ in range(300)]
    for d in divisors:
        ns = [0, 1, d - 1, d, d + 1, M, M - 1, (M // d) * d, (M // d) * d - 1]
        ns += [rng.getrandbits(64) for _ in range(40)]
        for n in ns:
            n &= M
            assert poggers_amd.fastmod_host(n, d) == n % d, (n, d)
