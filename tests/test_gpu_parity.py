"""GPU parity of the HIP counting path against the CPU oracle (bit-exact)."""
from __future__ import annotations

import types

import numpy as np
import pytest

from gkc_amd import capi
from gkc_amd import workloads as W
from oracle import oracle as O
from tests.helpers import assert_tables_equal

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)


def _run_gpu(rs, k, p, P, hpc=False, fmode=1, pmode=0, min_count=1, max_count=0, batches=1, flags=0):
    with capi.Counter(k, p, P, hpc=hpc, filter_mode=fmode, part_mode=pmode, min_count=min_count,
                      max_count=max_count) as c:
        bounds = np.linspace(0, rs.n, batches + 1).astype(int)
        for a, b in zip(bounds[:-1], bounds[1:]):
            if b > a:
                c.partition(rs.reads[int(rs.offs[a]):int(rs.offs[b])], rs.offs[a:b + 1] - rs.offs[a], flags=flags)
        c.count(flags)
        return c.table(), c.stats()


@pytest.mark.parametrize("cid,scale,batches", [(1, 0.05, 1), (1, 0.05, 3), (2, 0.004, 2)])
def test_table_matches_oracle(gpu, cid, scale, batches):
    w = W.workload(cid, scale=scale)
    rs = W.generate(w)
    _say(w.name, "reads", rs.n, "batches", batches)
    gt, st = _run_gpu(rs, w.k, w.p, w.n_parts, batches=batches)
    ot = O.count(rs.reads, rs.offs, w.k, w.p, w.n_parts)
    assert_tables_equal(gt, ot, w.n_parts, w.k)
    assert st["n_kmers"] == rs.kmers(w.k)


@pytest.mark.parametrize("fmode,pmode", [(0, 0), (0, 1), (1, 1)])
def test_filter_and_partition_modes(gpu, fmode, pmode):
    w = W.workload(1, scale=0.02)
    rs = W.generate(w)
    _say("modes", fmode, pmode)
    gt, _ = _run_gpu(rs, 31, 11, 64, fmode=fmode, pmode=pmode)
    ot = O.count(rs.reads, rs.offs, 31, 11, 64, filter_mode=fmode, part_mode=pmode)
    assert_tables_equal(gt, ot, 64, 31)


@pytest.mark.parametrize("k,p,P", [(21, 7, 31), (17, 5, 8), (32, 16, 16), (25, 15, 512), (12, 4, 3)])
def test_other_k(gpu, k, p, P):
    w = W.workload(1, scale=0.02)
    rs = W.generate(w)
    _say("k", k, "p", p, "P", P)
    gt, _ = _run_gpu(rs, k, p, P)
    ot = O.count(rs.reads, rs.offs, k, p, P)
    assert_tables_equal(gt, ot, P, k)


@pytest.mark.parametrize("k,p,P,minc", [(63, 19, 64, 2), (33, 17, 16, 1), (64, 32, 8, 1), (45, 13, 32, 3)])
def test_128bit_keys(gpu, k, p, P, minc):
    w = W.workload(4, scale=0.002)
    rs = W.generate(w)
    _say("k", k, "p", p, "P", P, "reads", rs.n)
    gt, _ = _run_gpu(rs, k, p, P, min_count=minc)
    ot = O.count(rs.reads, rs.offs, k, p, P, min_count=minc)
    assert_tables_equal(gt, ot, P, k)


def test_hpc_long_reads(gpu):
    w = W.workload(3, scale=0.002)
    rs = W.generate(w)
    _say("hpc", rs.n, rs.bases)
    gt, _ = _run_gpu(rs, w.k, w.p, w.n_parts, hpc=True)
    ot = O.count(rs.reads, rs.offs, w.k, w.p, w.n_parts, hpc=True)
    assert_tables_equal(gt, ot, w.n_parts, w.k)


def test_count_filter_bounds(gpu):
    w = W.workload(1, scale=0.05)
    rs = W.generate(w)
    _say("filter bounds")
    gt, _ = _run_gpu(rs, 31, 11, 64, min_count=3, max_count=9)
    ot = O.count(rs.reads, rs.offs, 31, 11, 64, min_count=3, max_count=9)
    assert_tables_equal(gt, ot, 64, 31)


def test_edge_reads(gpu):
    """Empty reads, reads shorter than k, exactly k, poly-A, lowercase."""
    seqs = [b"", b"ACGT", b"A" * 30, b"A" * 31, b"A" * 200, b"acgtACGTTTGACCA" * 9, b"", b"C" * 33,
            b"GATTACA" * 40, b"T" * 64]
    reads = np.frombuffer(b"".join(seqs), np.uint8).copy()
    offs = np.concatenate([[0], np.cumsum([len(s) for s in seqs])]).astype(np.uint64)
    _say("edge reads")
    with capi.Counter(31, 11, 7) as c:
        c.partition(reads, offs)
        c.count()
        gt = c.table()
    ot = O.count(reads, offs, 31, 11, 7)
    assert_tables_equal(gt, ot, 7, 31)
    bf = O.brute_force_table(reads, offs, 31)
    lo, _, cnt, *_ = gt
    assert dict(zip(lo.tolist(), cnt.tolist())) == bf


def test_superkmer_arrays_match_reference_layout(gpu):
    w = W.workload(1, scale=0.01)
    rs = W.generate(w)
    _say("superkmer arrays")
    with capi.Counter(31, 11, 64) as c:
        out = c.superkmer_arrays(rs.reads, rs.offs)
    mm, _, _ = O.window_minimizers(rs.reads, rs.offs, 31, 11)
    skm = O.superkmer_offs(rs.offs, 31, mm)
    lens = np.diff(rs.offs).astype(np.int64)
    valid = np.zeros(rs.bases, bool)
    for r in range(rs.n):
        valid[int(rs.offs[r]):int(rs.offs[r]) + max(0, int(lens[r]) - 31 + 1)] = True
    np.testing.assert_array_equal(out["minimizers"][valid], mm[valid].astype(np.uint32))
    # superkmer lists: compare each read's list up to its terminating 0
    for r in range(rs.n):
        s = int(rs.offs[r])
        ref = skm[s:s + int(lens[r])]
        got = out["superkmer_offs"][s:s + int(lens[r])]
        end = 1 + int(np.argmax(ref[1:] == 0)) + 1
        np.testing.assert_array_equal(got[:end], ref[:end])


def test_superkmer_arrays_hpc(gpu):
    w = W.workload(3, scale=0.001)
    rs = W.generate(w)
    _say("superkmer arrays hpc")
    with capi.Counter(25, 15, 512, hpc=True) as c:
        out = c.superkmer_arrays(rs.reads, rs.offs)
    hr, hl, hp = O.hpc(rs.reads, rs.offs)
    np.testing.assert_array_equal(out["read_len"], hl)
    for r in range(rs.n):
        s = int(rs.offs[r])
        L = int(hl[r])
        np.testing.assert_array_equal(out["reads"][s:s + L], hr[s:s + L])
        np.testing.assert_array_equal(out["hpc_orig_pos"][s:s + L], hp[s:s + L])


def _hpc_edge_reads():
    """A/C/G/T reads whose homopolymer runs cross the HPC pass's 16-byte lane
    and 1024-byte round boundaries, of every length class around them (0, 1,
    15-17, 1023-1025 bytes, a few thousand), one run of a single base, and
    random reads in between so that read starts take every alignment."""
    rng = np.random.default_rng(11)
    seqs = [b"", b"A", b"AAAA", b"ACGT" * 4, b"C" * 15, b"G" * 16 + b"T", b"T" * 3000]
    for L in (15, 16, 17, 31, 33, 1023, 1024, 1025, 2047, 2049, 3001):
        runs = []
        while sum(len(r) for r in runs) < L:
            runs.append(bytes([b"ACGT"[int(rng.integers(4))]]) * int(rng.integers(1, 40)))
        seqs.append(b"".join(runs)[:L])
    for _ in range(200):
        seqs.append(bytes(rng.choice(list(b"ACGT"), size=int(rng.integers(0, 300))).tolist()))
    reads = np.frombuffer(b"".join(seqs), np.uint8).copy()
    offs = np.concatenate([[0], np.cumsum([len(x) for x in seqs])]).astype(np.uint64)
    return reads, offs


def test_hpc_edge_reads(gpu):
    reads, offs = _hpc_edge_reads()
    hr, hl, hp = O.hpc(reads, offs)
    _say("hpc edge reads", len(offs) - 1)
    with capi.Counter(25, 15, 16, hpc=True) as c:
        out = c.superkmer_arrays(reads, offs)
    np.testing.assert_array_equal(out["read_len"], hl)
    for r in range(len(offs) - 1):
        s, L = int(offs[r]), int(hl[r])
        np.testing.assert_array_equal(out["reads"][s:s + L], hr[s:s + L])
        np.testing.assert_array_equal(out["hpc_orig_pos"][s:s + L], hp[s:s + L])
    rs = types.SimpleNamespace(reads=reads, offs=offs, n=len(offs) - 1)
    ot = O.count(reads, offs, 13, 5, 16, hpc=True)
    for flags in (0, capi.BATCHED):
        gt, _ = _run_gpu(rs, 13, 5, 16, hpc=True, flags=flags, batches=2)
        assert_tables_equal(gt, ot, 16, 13)


def test_count_packed_single_partition(gpu):
    rng = np.random.default_rng(7)
    seqs = [bytes(rng.choice(list(b"ACGT"), size=int(rng.integers(31, 90))).tolist()) for _ in range(500)]
    seqs += [seqs[3]] * 5
    words, lens = capi.pack_superkmers(seqs)
    _say("count_packed")
    with capi.Counter(31, 11, 1) as c:
        lo, _, cnt, nd = c.count_packed(words, lens, cap=100000)
    reads = np.frombuffer(b"".join(seqs), np.uint8).copy()
    offs = np.concatenate([[0], np.cumsum([len(s) for s in seqs])]).astype(np.uint64)
    bf = O.brute_force_table(reads, offs, 31)
    assert dict(zip(lo.tolist(), cnt.tolist())) == bf
    assert nd == len(bf)
    assert np.all(np.diff(lo.astype(np.float64)) > 0) or len(lo) < 2


@pytest.mark.parametrize("cid,scale,P", [(1, 0.05, 64), (4, 0.002, 64), (3, 0.002, 512)])
def test_batched_form_matches_oracle(gpu, cid, scale, P):
    w = W.workload(cid, scale=scale)
    rs = W.generate(w)
    _say("batched", w.name)
    gt, st = _run_gpu(rs, w.k, w.p, P, hpc=w.hpc, min_count=w.min_count, batches=2, flags=capi.BATCHED)
    ot = O.count(rs.reads, rs.offs, w.k, w.p, P, hpc=w.hpc, min_count=w.min_count)
    assert_tables_equal(gt, ot, P, w.k)
    assert st["host_syncs"] == 2 and st["fallbacks"] == 0


def test_batched_presized_and_fallback(gpu):
    """Exact caps, generous caps and under-predicted caps (device overflow flag ->
    count-then-allocate fallback) all give the bit-exact table."""
    w = W.workload(1, scale=0.05)
    rs = W.generate(w)
    ot = O.count(rs.reads, rs.offs, 31, 11, 64)
    exact = np.diff(ot.part_off).astype(np.uint32)
    for name, caps, fb in [("exact", exact, 0), ("generous", exact * 2 + 5, 0),
                           ("under", np.maximum(exact // 2, 1), 1)]:
        _say("presized", name)
        with capi.Counter(31, 11, 64) as c:
            c.partition(rs.reads, rs.offs)
            c.set_table_caps(caps)
            c.count(capi.BATCHED | capi.PRESIZED)
            gt = c.table()
            st = c.stats()
        assert_tables_equal(gt, ot, 64, 31)
        assert st["fallbacks"] == fb, name
        assert st["host_syncs"] == 1 + fb, name


def test_batched_graph_replay_stable(gpu):
    """The captured graph replayed several times (eager work in between) gives
    the same table every time."""
    w = W.workload(1, scale=0.03)
    rs = W.generate(w)
    ot = O.count(rs.reads, rs.offs, 31, 11, 64)
    with capi.Counter(31, 11, 64) as c:
        for rep in range(4):
            _say("replay", rep)
            c.reset()
            c.partition(rs.reads, rs.offs)
            c.count(capi.BATCHED)
            assert_tables_equal(c.table(), ot, 64, 31)


@pytest.mark.parametrize("cid,scale,k,p,hpc", [(1, 0.01, 31, 11, False), (3, 0.001, 25, 15, True),
                                               (4, 0.0005, 45, 13, False), (1, 0.01, 16, 16, False)])
def test_window_records(gpu, cid, scale, k, p, hpc):
    """(minimizer, position, strand) per window against the oracle's statement
    of the reference's update order."""
    rs = W.generate(W.workload(cid, scale=scale))
    _say("windows", cid, k, p, hpc)
    with capi.Counter(k, p, 64, hpc=hpc) as c:
        out = c.window_minimizers(rs.reads, rs.offs)
        arr = c.superkmer_arrays(rs.reads, rs.offs) if p <= 16 else None
    src, lens = rs.reads, None
    if hpc:
        src, lens, _ = O.hpc(rs.reads, rs.offs)
        np.testing.assert_array_equal(arr["read_len"], lens)
    mm, pos, strand = O.window_minimizers(src, rs.offs, k, p, 1, lens=lens)
    L = lens if hpc else np.diff(rs.offs).astype(np.int64)
    valid = np.zeros(rs.bases, bool)
    for r in range(rs.n):
        valid[int(rs.offs[r]):int(rs.offs[r]) + max(0, int(L[r]) - k + 1)] = True
    np.testing.assert_array_equal(out["minimizers"][valid], mm[valid].astype(np.uint32))
    np.testing.assert_array_equal(out["pos"][valid], pos[valid])
    np.testing.assert_array_equal(out["strand"][valid], strand[valid])
