"""Insert/delete churn on the TCF, concurrent kernels against the oracle.

* The reference's own protocols, run with its compositions:
  - sawtooth (tests/src/sawtooth_test.cu:623-803, main :859-862): del_TCF_small
    (8/8 slots, 16-slot buckets, XOR power-of-two probes, MurmurHash64A, an
    8-slot linear backing level of 400 probes sized 2 % of the primary),
    0.9 * total() keys, then 2 rounds of: delete the first half, query the
    other half, replace the first half with new keys inserted with
    insert_with_delete;
  - delete_tests (tests/src/delete_tests.cu:622-759, main :846-858): del_TCF
    (16/16 slots) and del_TCF_small with a 1 % backing level; insert
    0.9 * total(), delete the first half, query the second half, re-add the
    first half with insert_with_delete, reading the fill before, after the
    delete and at the end.
* A longer stress of this build (not a reference protocol): 10 % churn per
  round for 8 rounds on a 16/16 filter with a 1 % backing level.

Deterministic mode reproduces the oracle byte for byte.  Concurrent kernels
run a batch's ops in no fixed order, and under churn the order alone moves
the sequential algorithm's outcome a lot: after 8 rounds of the 10 % stress
the oracle replaying each batch in array order holds 35.4 K keys in its
backing level (false-positive rate 8.5e-4), replaying each batch in a
shuffled order 27.9 K (5.5e-4).  A filter delete removes the lowest slot
holding the key's fingerprint among its buckets, so with shared fingerprints
(often, with 8-bit ones) which key loses its slot depends on the order too:
the sawtooth's second round has 0 delete misses in array order and about 250
in a shuffled order.  So each count is checked against both sequential
replays of the same batches, run side by side: the GPU's must lie between
them (within sampling noise, and 5 % for false-positive rates); and, as
bounds against the array-order replay, backing occupancy at most 1.2x and
the false-positive rate at most 5 % above it."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
import poggers_amd as pg
from tests.gpu_util import Ops, pair

pytestmark = pytest.mark.gpu


def vals_of(keys, bits=16):
    v = (keys >> np.uint64(24)) & np.uint64((1 << bits) - 1)
    return v.astype(np.uint16 if bits == 16 else np.uint8)


def live(words, kmask):
    k = words & kmask
    return int(((k != 0) & (k != kmask)).sum())


def within(g, a, b, rel=0.05):
    """A GPU count inside the span of two sequential replays' counts, with
    `rel` of the larger and four standard deviations of sampling noise."""
    lo, hi = min(a, b), max(a, b)
    m = rel * hi + 4.0 * np.sqrt(hi + 1.0) + 2
    return lo - m <= g <= hi + m


def del_tcf(nslots, backing, small):
    """del_TCF / del_TCF_small (tests/src/delete_tests.cu:138-179)."""
    return pg.tcf_desc(nslots, backing, bucket_size=16, tile=4, prober=pg.PROBE_XOR_P2,
                       hasher=pg.HASH_MURMUR64A, backing_bucket=8, backing_prober=pg.PROBE_LINEAR,
                       backing_probes=400 if small else 4000,
                       rep=pg.REP_FP8_V8 if small else pg.REP_FP16_V16)


class Both:
    """The same ops on the GPU (concurrent) and on two oracles: one applying
    each batch in array order, one in a shuffled order (fixed seed).  The
    oracles run on two host threads (ctypes releases the GIL) while the GPU
    works.  Every method returns (gpu, array order, shuffled order)."""

    def __init__(self, desc, vbits):
        self.o, self.g = pair(desc)
        self.s = oracle.Table(desc.rep, [lv.as_dict() for lv in desc.levels], seed=42)
        self.ops = Ops(self.g)
        self.vbits = vbits
        self.kmask = 0xFF if desc.rep == pg.REP_FP8_V8 else 0xFFFF
        self.rng = np.random.default_rng(1234)
        self.pool = ThreadPoolExecutor(2)

    def _both(self, fo, fs, keys, *args):
        """fo on the batch as given and fs on it shuffled, on the two threads;
        fs's per-key results come back in the batch's order."""
        perm = self.rng.permutation(len(keys))
        inv = np.empty_like(perm)
        inv[perm] = np.arange(len(perm))

        def shuffled():
            r = fs(keys[perm], *[x[perm] for x in args])
            return tuple(x[inv] for x in r) if isinstance(r, tuple) else r[inv]
        return self.pool.submit(fo, keys, *args), self.pool.submit(shuffled)

    def insert(self, keys, with_delete=False):
        v = vals_of(keys, self.vbits)
        a, b = self._both(lambda k, vv: self.o.bulk_insert(k, vv, with_delete=with_delete),
                          lambda k, vv: self.s.bulk_insert(k, vv, with_delete=with_delete), keys, v)
        st_g, _ = self.ops.insert(keys, v, with_delete=with_delete)
        return int((st_g == 0).sum()), int((a.result() == 0).sum()), int((b.result() == 0).sum())

    def remove(self, keys):
        a, b = self._both(self.o.bulk_remove, self.s.bulk_remove, keys)
        g = self.ops.remove(keys)
        return int((g == 0).sum()), int((a.result() == 0).sum()), int((b.result() == 0).sum())

    def query(self, keys):
        """(false negatives, value mismatches) as (gpu, array, shuffled) each."""
        a, b = self._both(self.o.bulk_query, self.s.bulk_query, keys)
        f_g, v_g = self.ops.query(keys)
        want = vals_of(keys, self.vbits)
        out = []
        for f, v in ((f_g, v_g), a.result(), b.result()):
            out.append((int((f == 0).sum()), int(((f == 1) & (v != want)).sum())))
        return tuple(x[0] for x in out), tuple(x[1] for x in out)

    def found(self, keys):
        a, b = self._both(self.o.bulk_query, self.s.bulk_query, keys)
        g = int(self.ops.query(keys)[0].sum())
        return g, int(a.result()[0].sum()), int(b.result()[0].sum())

    def backing(self):
        return (live(self.ops.level(1), self.kmask), live(self.o.level_array(1), self.kmask),
                live(self.s.level_array(1), self.kmask))

    def fills(self):
        return self.g.get_fill(), self.o.fill(), self.s.fill()


def check_occupancy_and_fp(back, fp, what, backing_slots):
    """Backing occupancy and false positives of the GPU against the two
    replays: occupancy within 1.2x of the array-order replay's, and false
    positives inside the replays' span and not above the array-order one by
    more than 5 %.

    Past 80 % occupancy of the backing level the false positives are set by
    the lengths of the linear-probe runs an absent key walks through a
    nearly full level, which vary by 60 % between the two replays' key
    orders and, at the same occupancy, by up to 18 % between GPU runs
    (7 157 to 8 440 at 86 %, against 7 091 for the array-order replay).
    There the bound is 1.25x the array-order replay."""
    (bg, bo, bs), (fg, fo, fs) = back, fp
    assert bg <= 1.2 * bo + 20 and min(bo, bs) / 1.2 - 20 <= bg, (what, back)
    rel = 0.05 if bo <= 0.8 * backing_slots else 0.25
    assert fg <= (1.0 + rel) * fo + 4.0 * np.sqrt(fo + 1.0), (what, fp, rel)
    assert within(fg, fo, fs, rel=rel), (what, fp)


def test_sawtooth_deterministic_bit_exact():
    nslots = 1 << 14
    desc = pg.tcf_desc(nslots, nslots // 100, bucket_size=16, tile=4, prober=pg.PROBE_XOR_P2,
                       hasher=pg.HASH_MURMUR64A)
    o, g = pair(desc)
    g.set_mode(pg.DETERMINISTIC)
    ops = Ops(g)
    n = int(0.9 * nslots)
    keys = oracle.gen_keys(81, 0, n)
    assert (ops.insert(keys, vals_of(keys))[0] == o.bulk_insert(keys, vals_of(keys))).all()
    step = n // 10
    fresh = 0
    for rnd in range(6):
        a = (rnd * step) % (n - step)
        dels = keys[a:a + step].copy()
        assert (ops.remove(dels) == o.bulk_remove(dels)).all(), rnd
        kept = np.concatenate([keys[:a], keys[a + step:]])
        f_o, v_o = o.bulk_query(kept)
        f_g, v_g = ops.query(kept)
        assert (f_g == f_o).all() and (v_g == v_o).all(), rnd
        new = oracle.gen_keys(82, fresh, step)
        fresh += step
        st_o = o.bulk_insert(new, vals_of(new), with_delete=True)
        st_g, _ = ops.insert(new, vals_of(new), with_delete=True)
        assert (st_g == st_o).all(), rnd
        keys[a:a + step] = new
        for lvl in range(2):
            assert g.level_bytes(lvl) == o.level_bytes(lvl), (rnd, lvl)
        assert g.get_fill() == o.fill()


def test_reference_sawtooth_protocol():
    """sawtooth_test<del_TCF_small>(size (2^n, 2 * 2^n / 100), 50 % deletes, 2 rounds) at n = 22."""
    nbits = 22
    primary, backing = 1 << nbits, 2 * (1 << nbits) // 100
    b = Both(del_tcf(primary, backing, small=True), 8)
    nitems = int((primary + backing) * 0.9)
    ndel = nitems * 50 // 100
    keys = oracle.gen_keys(91, 0, nitems)
    absent = oracle.gen_keys(91, 1 << 40, 1 << 21)
    fails = b.insert(keys)
    print(f"fill: insert failures {fails} (gpu, array order, shuffled order)", flush=True)
    assert within(*fails)
    fresh = 0
    for rnd in range(2):
        dels = keys[:ndel].copy()
        miss = b.remove(dels)
        still = b.found(dels)          # "delete false matching"
        fn, bad = b.query(keys[ndel:])
        new = oracle.gen_keys(92, fresh, ndel)
        fresh += ndel
        keys[:ndel] = new
        readd = b.insert(new, with_delete=True)
        back = b.backing()
        fp = b.found(absent)
        print(f"round {rnd} (gpu, array order, shuffled order): delete misses {miss}, still found {still}, "
              f"false negatives {fn}, value mismatches {bad}, re-add failures {readd}, backing {back}, "
              f"false positives {fp}", flush=True)
        for what, c in (("delete misses", miss), ("still found", still), ("false negatives", fn),
                        ("value mismatches", bad), ("re-add failures", readd)):
            assert within(*c), (rnd, what, c)
        check_occupancy_and_fp(back, fp, rnd, backing)


@pytest.mark.parametrize("small", [False, True], ids=["del_TCF", "del_TCF_small"])
def test_reference_delete_tests_protocol(small):
    """delete_tests<del_TCF / del_TCF_small>(size (2^n, 2^n / 100)) at n = 22."""
    nbits = 22
    primary, backing = 1 << nbits, (1 << nbits) // 100
    b = Both(del_tcf(primary, backing, small), 8 if small else 16)
    nitems = int((primary + backing) * 0.9)
    half = nitems // 2
    keys = oracle.gen_keys(93 + small, 0, nitems)
    absent = oracle.gen_keys(93 + small, 1 << 40, 1 << 21)
    fails = b.insert(keys)
    fill_before = b.fills()
    miss = b.remove(keys[:half])
    fill_after = b.fills()
    fn, bad = b.query(keys[half:])
    readd = b.insert(keys[:half], with_delete=True)
    final_fill = b.fills()
    back = b.backing()
    fp = b.found(absent)
    print(f"(gpu, array order, shuffled order): insert failures {fails}, delete misses {miss}, false negatives "
          f"{fn}, value mismatches {bad}, re-add failures {readd}; fill before {fill_before}, after delete "
          f"{fill_after}, final {final_fill}; backing {back}, false positives {fp}", flush=True)
    for what, c in (("insert failures", fails), ("delete misses", miss), ("false negatives", fn),
                    ("value mismatches", bad), ("re-add failures", readd), ("fill before", fill_before),
                    ("fill after delete", fill_after), ("final fill", final_fill)):
        assert within(*c, rel=1e-4 if "fill" in what else 0.05), (what, c)
    check_occupancy_and_fp(back, fp, "end", backing)


@pytest.mark.parametrize("backing_pct", [1, 2])
def test_churn_ten_percent_eight_rounds(backing_pct):
    """Longer churn than the reference runs: 10 % of the keys replaced per
    round, 8 rounds, 16/16 slots, 2^22 slots, a backing level of 1 % (the
    reference's size) or 2 % of the slots.  The backing level fills round
    by round in any order (a key that spilled leaves it only when that key
    is deleted), and the false-positive rate rises as absent-key walks
    through it lengthen; the concurrent kernels must stay between the
    array-order and shuffled-order replays every round.  The 1 % level
    reaches 86 % occupancy by round 8; the 2 % level stays below half full,
    so its false positives are held to 5 % of the array-order replay's in
    every round."""
    nslots = 1 << 22
    backing = nslots * backing_pct // 100
    b = Both(pg.tcf_desc(nslots, backing, bucket_size=16, tile=4), 16)
    n = int(0.9 * nslots)
    keys = oracle.gen_keys(83, 0, n)
    assert b.insert(keys) == (0, 0, 0)
    absent = oracle.gen_keys(83, 1 << 40, 1 << 23)
    step = n // 10
    fresh = 0
    for rnd in range(9):
        if rnd:
            a = ((rnd - 1) * step) % (n - step)
            miss = b.remove(keys[a:a + step].copy())
            assert within(*miss), (rnd, miss)
            kept = np.concatenate([keys[:a], keys[a + step:]])
            fn, bad = b.query(kept)
            assert within(*fn) and within(*bad), (rnd, fn, bad)
            new = oracle.gen_keys(84, fresh, step)
            fresh += step
            assert b.insert(new, with_delete=True) == (0, 0, 0), rnd
            keys[a:a + step] = new
            fills = b.fills()
            assert within(*fills, rel=1e-5), (rnd, fills)
        back = b.backing()
        fp = b.found(absent)
        print(f"round {rnd} (gpu, array order, shuffled order): backing {back} of {backing} slots, "
              f"false positives {fp}", flush=True)
        check_occupancy_and_fp(back, fp, rnd, backing)
