"""Binned counting (batched form): onesweep passes on the top 24 key bits,
then per window of whole bins a wave-level count in LDS (hash count, rank of
the distinct keys), larger windows per workgroup.  Bit-exact
against the oracle and against the full LSD sort (count_mode 1), including
the edges of the tiling: bins straddling tile windows, ranges of one key or
of one repeated key, ranges too large for LDS (sorted and counted in chunks
through global memory), digits straddling bit 64 of 128-bit keys, and
thresholds."""
from __future__ import annotations

import numpy as np
import pytest

from gkc_amd import capi
from gkc_amd import workloads as W
from oracle import oracle as O
from tests.helpers import assert_tables_equal

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)


def _count(reads, offs, k, p, P, mode, hpc=False, min_count=1, max_count=0, batches=2, caps=None, bin_passes=0,
           count_window=0, digit_bits=0, sort_items=0):
    n = len(offs) - 1
    with capi.Counter(k, p, P, hpc=hpc, min_count=min_count, max_count=max_count) as c:
        c.set_tuning(sort_items, count_mode=mode, bin_passes=bin_passes, count_window=count_window,
                     digit_bits=digit_bits)
        bounds = np.linspace(0, n, batches + 1).astype(int)
        for a, b in zip(bounds[:-1], bounds[1:]):
            if b > a:
                c.partition(reads[int(offs[a]):int(offs[b])], offs[a:b + 1] - offs[a], flags=capi.BATCHED)
        flags = capi.BATCHED
        if caps is not None:
            c.set_table_caps(caps)
            flags |= capi.PRESIZED
        c.count(flags)
        return c.table(), c.stats()


def _reads(seqs):
    reads = np.frombuffer(b"".join(seqs), np.uint8).copy()
    offs = np.concatenate([[0], np.cumsum([len(s) for s in seqs])]).astype(np.uint64)
    return reads, offs


@pytest.mark.parametrize("cid,scale,k,p,P", [
    (1, 0.05, 31, 11, 2),       # ~300 k keys per segment: many tiles, bins straddling windows
    (2, 0.004, 31, 11, 64),
    (1, 0.02, 8, 4, 3),         # bins = the whole key (2k - 16 = 0)
    (1, 0.02, 9, 5, 5),
    (1, 0.02, 32, 16, 4),       # all 64 bits used
    (1, 0.02, 33, 17, 4),       # 128-bit keys, bin digits straddle bit 64
    (1, 0.02, 40, 13, 8),
    (4, 0.002, 63, 19, 16),
])
def test_binned_matches_oracle_and_lsd(gpu, cid, scale, k, p, P):
    w = W.workload(cid, scale=scale)
    rs = W.generate(w)
    ot = O.count(rs.reads, rs.offs, k, p, P)
    _say("binned", w.name, k, p, P)
    gt, st = _count(rs.reads, rs.offs, k, p, P, mode=0)
    assert st["binned"] == 1 and st["n_pass"] == min(3, (2 * k + 7) // 8)
    assert_tables_equal(gt, ot, P, k)
    _say("lsd", w.name, k, p, P)
    lt, sl = _count(rs.reads, rs.offs, k, p, P, mode=1)
    assert sl["binned"] == 0 and sl["n_pass"] == (2 * k + 7) // 8
    assert_tables_equal(lt, ot, P, k)


@pytest.mark.parametrize("k,p", [(31, 11), (45, 15)])
def test_bin_passes(gpu, k, p):
    """Every number of onesweep passes before the count gives the oracle's
    table: fewer passes leave larger bins, so more windows are heavy."""
    w = W.workload(1, scale=0.03)
    rs = W.generate(w)
    ot = O.count(rs.reads, rs.offs, k, p, 4)
    heavy = []
    for npss in (1, 2, 3):
        _say("bin_passes", npss, k)
        gt, st = _count(rs.reads, rs.offs, k, p, 4, mode=0, bin_passes=npss)
        assert st["binned"] == 1 and st["n_pass"] == npss
        assert_tables_equal(gt, ot, 4, k)
        heavy.append(st["heavy_ranges"])
    assert heavy[0] >= heavy[1] >= heavy[2]


@pytest.mark.parametrize("k,p,P,db,npss,items", [(31, 11, 4, 7, 2, 20), (31, 11, 4, 6, 3, 16), (31, 11, 64, 6, 2, 20),
                                                (21, 7, 8, 7, 3, 16), (9, 5, 5, 6, 3, 20), (33, 17, 4, 7, 3, 8),
                                                (45, 15, 8, 6, 2, 12), (63, 19, 16, 7, 3, 12)])
def test_digit_widths(gpu, k, p, P, db, npss, items):
    """Onesweep passes of 6- or 7-bit digits (bins of npss x db key bits, a
    digit straddling bit 64 of 128-bit keys for k = 33) give the oracle's
    table."""
    cid, scale = (4, 0.002) if k > 50 else (1, 0.03)
    rs = W.generate(W.workload(cid, scale=scale))
    ot = O.count(rs.reads, rs.offs, k, p, P)
    _say("digit bits", k, p, P, db, npss, items)
    gt, st = _count(rs.reads, rs.offs, k, p, P, mode=0, bin_passes=npss, digit_bits=db, sort_items=items)
    assert st["binned"] == 1 and st["n_pass"] == min(npss, (2 * k + db - 1) // db)
    assert_tables_equal(gt, ot, P, k)


@pytest.mark.parametrize("minc,maxc", [(2, 0), (1, 3), (3, 40)])
def test_binned_thresholds(gpu, minc, maxc):
    w = W.workload(1, scale=0.03)
    rs = W.generate(w)
    ot = O.count(rs.reads, rs.offs, 31, 11, 4, min_count=minc, max_count=maxc)
    _say("thresholds", minc, maxc)
    gt, st = _count(rs.reads, rs.offs, 31, 11, 4, mode=0, min_count=minc, max_count=maxc)
    assert st["binned"] == 1
    assert_tables_equal(gt, ot, 4, 31)


def test_binned_hpc_long_reads(gpu):
    w = W.workload(3, scale=0.002)
    rs = W.generate(w)
    ot = O.count(rs.reads, rs.offs, w.k, w.p, 16, hpc=True)
    _say("binned hpc")
    gt, st = _count(rs.reads, rs.offs, w.k, w.p, 16, mode=0, hpc=True)
    assert st["binned"] == 1
    assert_tables_equal(gt, ot, 16, w.k)


@pytest.mark.parametrize("k,p,copies", [(31, 11, 300), (40, 13, 200)])
def test_heavy_range_repeated_kmer(gpu, k, p, copies):
    """A k-mer occurring more often than LDS holds (poly-A: k-mer 0) makes its
    bin's range too large for LDS; it is sorted and counted in chunks through
    global memory (runs crossing chunk ends)."""
    rng = np.random.default_rng(5)
    seqs = [b"A" * 150] * copies
    seqs += [bytes(rng.choice(list(b"ACGT"), 150).tolist()) for _ in range(400)]
    reads, offs = _reads(seqs)
    ot = O.count(reads, offs, k, p, 4)
    _say("heavy repeated", k, copies)
    gt, st = _count(reads, offs, k, p, 4, mode=0)
    assert st["binned"] == 1 and st["heavy_ranges"] >= 1
    assert_tables_equal(gt, ot, 4, k)


@pytest.mark.parametrize("minc,nrand", [(1, 22), (2, 22), (1, 11), (2, 12)])
def test_heavy_range_distinct_kmers(gpu, minc, nrand):
    """One bin holding ~40 k keys of ~20 k distinct k-mers (fixed prefix, then
    nrand random bases, then A, so canonical = forward; one k-mer per read): a
    multi-chunk global sort over 2*nrand varying bits (11 bases: 3 passes,
    the odd case), counted twice (count pass, then emit pass)."""
    rng = np.random.default_rng(11)
    body = rng.choice(list(b"ACGT"), (20000, nrand)).astype(np.uint8)
    prefix = (b"ACGTACGTACGTACGTACGTACGTACGTACGT")[:30 - nrand]
    seqs = []
    for r in range(len(body)):
        s = prefix + bytes(body[r].tolist()) + b"A"
        seqs += [s] * (1 + r % 3)
    reads, offs = _reads(seqs)
    ot = O.count(reads, offs, 31, 11, 1, min_count=minc)
    _say("heavy distinct", minc)
    gt, st = _count(reads, offs, 31, 11, 1, mode=0, min_count=minc)
    assert st["binned"] == 1
    if len(prefix) >= 12:       # the whole 24-bit bin is the fixed prefix: one range of ~40 k keys
        assert st["heavy_ranges"] >= 1
    assert_tables_equal(gt, ot, 1, 31)


def test_binned_small_ranges(gpu):
    """Ranges of a single key, of one repeated key (no radix pass), empty
    partitions, reads of exactly k bases."""
    seqs = [b"ACGTTGCAAGGCTTAACCGGTTAAGGCCTTA", b"A" * 60, b"C" * 45, b"GATTACA" * 12,
            b"", b"ACGTTGCAAGGCTTAACCGGTTAAGGCCTTA", b"AC" * 40]
    reads, offs = _reads(seqs)
    ot = O.count(reads, offs, 31, 11, 9)
    _say("small ranges")
    gt, st = _count(reads, offs, 31, 11, 9, mode=0, batches=1)
    assert st["binned"] == 1
    assert_tables_equal(gt, ot, 9, 31)
    bf = O.brute_force_table(reads, offs, 31)
    assert dict(zip(gt[0].tolist(), gt[2].tolist())) == bf


def test_binned_presized_fallback(gpu):
    w = W.workload(1, scale=0.03)
    rs = W.generate(w)
    ot = O.count(rs.reads, rs.offs, 31, 11, 8)
    exact = np.diff(ot.part_off).astype(np.uint32)
    # the binned count lays the table out densely from its exact counts: only
    # the sum of the caps must cover the table (caps moved between partitions
    # do not fall back); the full LSD count gives each partition its own cap
    misplaced = (np.roll(exact, 3) + 1).astype(np.uint32)
    assert (misplaced < exact).any() and misplaced.sum() >= exact.sum()
    for name, caps, mode, fb in [("exact", exact, 0, 0), ("under", np.maximum(exact // 3, 1), 0, 1),
                                 ("misplaced", misplaced, 0, 0), ("misplaced, LSD count", misplaced, 1, 1)]:
        _say("binned presized", name)
        gt, st = _count(rs.reads, rs.offs, 31, 11, 8, mode=mode, caps=caps)
        assert st["binned"] == (1 - mode) and st["fallbacks"] == fb and st["host_syncs"] == 1 + fb, name
        assert_tables_equal(gt, ot, 8, 31)


@pytest.mark.parametrize("n_distinct,copies,minc", [
    (200, 6, 1), (120, 12, 3), (300, 8, 2), (384, 4, 1), (385, 4, 1),   # <= 384 distinct: the wave table
    (600, 1, 1), (768, 1, 1), (500, 2, 2),                     # 512 < n <= 768: one wave, 1024-slot table
    (769, 1, 1), (400, 3, 1), (600, 2, 2), (1000, 1, 1), (1536, 1, 1),   # <= 1536: one wave, 2048 slots
    (1537, 1, 1), (1500, 2, 1), (1500, 3, 3), (3000, 2, 1), (3000, 2, 2),   # larger: a workgroup (k_bincount)
    (300, 20, 1), (380, 40, 5), (3000, 4, 2), (500, 40, 1),    # up to 8192 keys tried in the wave table
])
def test_heavy_window_in_lds(gpu, n_distinct, copies, minc):
    """One 24-bit bin of n > 512 keys (n = the sum over the distinct keys of
    1 + r % copies): a heavy window counted in LDS.  Up to 1536 keys the
    wave-level count tries it in its 512-slot table, 64 keys at a time, and
    hands it on once more than 384 slots are taken, to one wave with a larger
    table (1024 slots up to 768 keys, else 2048); up to 8192 keys the same
    try, then a workgroup (hash count, distinct keys bucketed by their highest
    varying bits and ranked within the bucket), which counts larger windows
    from the start."""
    rng = np.random.default_rng(n_distinct + copies)
    body = rng.choice(list(b"ACGT"), (n_distinct, 11)).astype(np.uint8)
    seqs = []
    for r in range(n_distinct):
        s = b"ACGTACGTACGTACGTACG" + bytes(body[r].tolist()) + b"A"     # canonical = forward
        seqs += [s] * (1 + r % copies)
    seqs += [bytes(rng.choice(list(b"ACGT"), 150).tolist()) for _ in range(300)]
    reads, offs = _reads(seqs)
    ot = O.count(reads, offs, 31, 11, 1, min_count=minc)
    _say("heavy window in LDS", n_distinct, copies, minc)
    gt, st = _count(reads, offs, 31, 11, 1, mode=0, min_count=minc)
    assert st["binned"] == 1 and st["heavy_ranges"] >= 1
    assert_tables_equal(gt, ot, 1, 31)


@pytest.mark.parametrize("k,p,P,cw", [(31, 11, 4, 128), (31, 11, 4, 384), (31, 11, 4, 1024), (45, 15, 4, 384)])
def test_count_window(gpu, k, p, P, cw):
    """Every window size of the binned count gives the oracle's table (larger
    windows: more ranges of over 512 keys, counted 64 keys at a time)."""
    w = W.workload(1, scale=0.03)
    rs = W.generate(w)
    ot = O.count(rs.reads, rs.offs, k, p, P)
    for npss in (2, 3):
        _say("count window", cw, npss, k)
        gt, st = _count(rs.reads, rs.offs, k, p, P, mode=0, bin_passes=npss, count_window=cw)
        assert st["binned"] == 1 and st["n_pass"] == npss
        assert_tables_equal(gt, ot, P, k)


@pytest.mark.parametrize("npss,minc", [(3, 1), (2, 1), (3, 2)])
def test_wave_table_full_after_small(gpu, npss, minc):
    """Random reads (almost every k-mer distinct, a few repeated): windows of
    up to 512 keys hold more than 191 distinct keys, so the wave-level count
    gives up its 256-slot table and counts them again in all 512 slots; both
    paths meet in one table (mixed with windows that fit the small table)."""
    rng = np.random.default_rng(23 + npss)
    seqs = [bytes(rng.choice(list(b"ACGT"), 150).tolist()) for _ in range(12000)]
    seqs += seqs[:3000]                                   # some k-mers twice
    reads, offs = _reads(seqs)
    ot = O.count(reads, offs, 31, 11, 2, min_count=minc)
    _say("wave table full after small", npss, minc)
    gt, st = _count(reads, offs, 31, 11, 2, mode=0, min_count=minc, bin_passes=npss)
    assert st["binned"] == 1 and st["n_pass"] == npss
    assert_tables_equal(gt, ot, 2, 31)


def test_tuning_changes_between_counts(gpu):
    """One context counted with 8-, 6- and again 8-bit digits, 4 and 8 waves
    per onesweep workgroup and a changed window: each count (its captured
    graph rebuilt when the shape changes) gives the oracle's table; values
    outside the allowed sets are refused."""
    rs = W.generate(W.workload(1, scale=0.02))
    k, p, P = 31, 11, 16
    ot = O.count(rs.reads, rs.offs, k, p, P)
    with capi.Counter(k, p, P) as c:
        c.partition(rs.reads, rs.offs, flags=capi.BATCHED)
        for items, waves, db, npss, cw in [(20, 4, 8, 2, 384), (16, 4, 6, 3, 256), (10, 8, 8, 2, 256), (20, 4, 8, 2, 384)]:
            _say("retune", items, waves, db, npss, cw)
            c.set_tuning(items, bin_passes=npss, count_window=cw, sort_waves=waves, digit_bits=db)
            t = c.get_tuning()
            assert (t["sort_items"], t["sort_waves"], t["digit_bits"]) == (items, waves, db)
            for _ in range(2):
                c.count(capi.BATCHED)
                assert_tables_equal(c.table(), ot, P, k)
        for bad in ({"digit_bits": 5}, {"digit_bits": 9}, {"sort_waves": 6}, {"bin_passes": 4}, {"count_window": 100}):
            with pytest.raises(capi.GkcError):
                c.set_tuning(**bad)
