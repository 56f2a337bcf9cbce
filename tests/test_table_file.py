"""The table file (gkc_table_write / gkc_amd.table_file): the reference keeps
no table (kmer_counting.cu:187-190), so the file is pinned by its own
invariants: the numpy writer and reader round-trip the oracle's table, bad
files are rejected, and the file the device path streams out is byte for
byte the file the numpy writer makes from the same table, which equals the
oracle's."""
from __future__ import annotations

import numpy as np
import pytest

from gkc_amd import table_file as TF
from gkc_amd import workloads as W
from oracle import oracle as O
from tests.helpers import assert_tables_equal


def _oracle_file(tmp_path, w, rs, n_parts, min_count=1, max_count=0):
    ot = O.count(rs.reads, rs.offs, w.k, w.p, n_parts, min_count=min_count, max_count=max_count)
    path = tmp_path / "oracle.gkct"
    hi = ot.key_hi if w.k > 32 else None
    n = TF.write_table(path, w.k, w.p, 0, min_count, max_count, ot.part_off, ot.part_distinct, ot.part_kmers,
                       ot.key_lo, hi, ot.cnt)
    return ot, path, n


@pytest.mark.parametrize("cfg,n_parts,mn,mx", [(1, 16, 1, 0), (4, 8, 2, 0), (1, 64, 2, 3)])
def test_round_trip_oracle_table(tmp_path, cfg, n_parts, mn, mx):
    w = W.workload(cfg, scale=0.002)
    rs = W.generate(w)
    ot, path, n = _oracle_file(tmp_path, w, rs, n_parts, mn, mx)
    assert path.stat().st_size == n
    t = TF.read_table(path)
    assert (t.k, t.p, t.n_parts, t.key_words, t.min_count, t.max_count) == (w.k, w.p, n_parts, 1 if w.k <= 32 else 2,
                                                                           mn, mx)
    assert t.n_kmers == int(ot.part_kmers.sum())
    assert_tables_equal(t.as_table(), ot, n_parts, w.k)
    for q in range(n_parts):                            # keys strictly ascending within each partition
        a, b = int(t.part_off[q]), int(t.part_off[q + 1])
        lo = np.asarray(t.keys[a:b, 0])
        hi = np.asarray(t.keys[a:b, 1]) if t.key_words == 2 else np.zeros_like(lo)
        assert np.all((hi[1:] > hi[:-1]) | ((hi[1:] == hi[:-1]) & (lo[1:] > lo[:-1])))


def test_empty_table(tmp_path):
    path = tmp_path / "e.gkct"
    TF.write_table(path, 31, 11, 0, 1, 0, np.zeros(5, np.uint64), np.zeros(4, np.uint64), np.zeros(4, np.uint64),
                   np.zeros(0, np.uint64), None, np.zeros(0, np.uint32))
    t = TF.read_table(path)
    assert t.keys.shape == (0, 1) and len(t.counts) == 0 and t.n_parts == 4


def test_rejects_bad_files(tmp_path):
    w = W.workload(1, scale=0.001)
    rs = W.generate(w)
    _, path, n = _oracle_file(tmp_path, w, rs, 8)
    raw = path.read_bytes()
    bad = tmp_path / "bad.gkct"
    bad.write_bytes(b"XKCTAB1\x00" + raw[8:])
    with pytest.raises(TF.TableFileError):
        TF.read_table(bad)
    bad.write_bytes(raw[:-4])                           # cut short
    with pytest.raises(TF.TableFileError):
        TF.read_table(bad)
    h = bytearray(raw)
    h[8:12] = (2).to_bytes(4, "little")                 # unknown version
    bad.write_bytes(bytes(h))
    with pytest.raises(TF.TableFileError):
        TF.read_table(bad)


# ---------------------------------------------------------------- GPU path

@pytest.mark.gpu
@pytest.mark.parametrize("cfg,n_parts,flags_name", [(1, 16, "plain"), (4, 8, "plain"), (2, 64, "batched"),
                                                    (1, 32, "presized")])
def test_device_file_matches(gpu, tmp_path, cfg, n_parts, flags_name):
    """Plain and batched counts (a presized full-LSD count, its caps set above
    the exact sizes, leaves gaps between partitions on the device: the file
    is written run by run)."""
    from gkc_amd import capi
    w = W.workload(cfg, scale=0.004)
    rs = W.generate(w)
    print("[gpu-case] table file", cfg, n_parts, flags_name, flush=True)
    flags = {"plain": 0, "batched": capi.BATCHED, "presized": capi.BATCHED | capi.PRESIZED}[flags_name]
    ot = O.count(rs.reads, rs.offs, w.k, w.p, n_parts)
    path = tmp_path / "dev.gkct"
    with capi.Counter(w.k, w.p, n_parts) as c:
        if flags_name == "presized":
            c.set_tuning(count_mode=1)       # per-partition capacities (the binned count lays out densely)
        c.partition(rs.reads, rs.offs, flags=flags & capi.BATCHED)
        if flags_name == "presized":
            exact = np.diff(ot.part_off)
            c.set_table_caps((exact + exact // 4 + 3).astype(np.uint32))
        c.count(flags)
        nbytes = c.write_table(path)
        lo, hi, cnt, po, pd, pk = c.table()
        starts = c.table_starts()
        st = c.stats()
    if flags_name == "presized":
        assert st["fallbacks"] == 0
        assert np.all(np.diff(starts.astype(np.int64)) > np.diff(po.astype(np.int64))[:-1])    # gaps between partitions
    assert path.stat().st_size == nbytes
    assert_tables_equal(TF.read_table(path).as_table(), ot, n_parts, w.k)
    twin = tmp_path / "host.gkct"
    TF.write_table(twin, w.k, w.p, 0, 1, 0, po, pd, pk, lo, hi, cnt)
    assert path.read_bytes() == twin.read_bytes()


@pytest.mark.gpu
def test_device_file_errors(gpu, tmp_path):
    from gkc_amd import capi
    w = W.workload(1, scale=0.001)
    rs = W.generate(w)
    with capi.Counter(w.k, w.p, 8) as c:
        with pytest.raises(capi.GkcError):               # nothing counted yet
            c.write_table(tmp_path / "x.gkct")
        c.partition(rs.reads, rs.offs)
        c.count()
        with pytest.raises(capi.GkcError, match="cannot create"):
            c.write_table(tmp_path / "no_such_dir" / "x.gkct")
        assert c.write_table(tmp_path / "x.gkct") > 0    # the context still works


@pytest.mark.gpu
def test_device_file_many_chunks(gpu, tmp_path):
    """A presized table of tens of millions of entries: several 64 MiB
    staging chunks (both buffers used in turn, runs split across chunks)."""
    from gkc_amd import capi
    w = W.workload(2, scale=0.15)
    rs = W.generate(w)
    print("[gpu-case] table file, many chunks", rs.n, flush=True)
    path = tmp_path / "big.gkct"
    with capi.Counter(w.k, w.p, w.n_parts) as c:
        c.set_tuning(count_mode=1)           # per-partition capacities: gaps between partitions
        c.partition(rs.reads, rs.offs, flags=capi.BATCHED)
        pk0 = c.part_kmers()
        c.set_table_caps(np.minimum(pk0 // 3 + 64, pk0).astype(np.uint32))   # above the distinct counts here
        c.count(capi.BATCHED | capi.PRESIZED)
        nbytes = c.write_table(path)
        lo, hi, cnt, po, pd, pk = c.table()
        starts = c.table_starts()
        assert c.stats()["fallbacks"] == 0
    assert np.any(starts[1:] != po[1:-1])              # the device table has gaps: written run by run
    assert len(cnt) * 12 > 3 * (64 << 20)
    twin = tmp_path / "host.gkct"
    assert TF.write_table(twin, w.k, w.p, 0, 1, 0, po, pd, pk, lo, hi, cnt) == nbytes
    with open(path, "rb") as a, open(twin, "rb") as b:
        while True:
            x, y = a.read(1 << 24), b.read(1 << 24)
            assert x == y
            if not x:
                break
