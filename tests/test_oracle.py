"""Pins the CPU oracle (oracle/gkc_oracle.c) without trusting it.

The reference holds no tests, golden vectors or fixtures for this path, so
the oracle is checked against
  * an independent brute force (string canonical form, dict counting),
  * closed forms and special cases (filter table, poly-A reads, HPC of
    hand-made reads, superkmer lists of a hand-made minimizer sequence),
  * invariants the mathematics fixes: a window's minimizer equals that of
    its reverse complement (so every canonical k-mer lands in one
    partition), the union of partition tables does not depend on P, the
    filter or the partition mode, counts sum to the number of k-mers,
  * an alternative algorithm for the minimizer (sliding-window minimum over
    per-position candidates, numpy) against the oracle's O(k) rescan,
  * the committed golden fixtures in tests/golden/.
"""
from __future__ import annotations

import numpy as np
import pytest

from gkc_amd import workloads as W
from oracle import oracle as O

COMP = bytes.maketrans(b"ACGT", b"TGCA")


def _csr(seqs):
    reads = np.frombuffer(b"".join(seqs), np.uint8).copy() if seqs else np.zeros(0, np.uint8)
    offs = np.concatenate([[0], np.cumsum([len(s) for s in seqs])]).astype(np.uint64)
    return reads, offs


def _enc(s: bytes) -> int:
    v = 0
    for ch in s:
        v = (v << 2) | b"ACGT".index(bytes([ch]).upper())
    return v


# ---------------------------------------------------------------- closed forms

def test_filter_closed_form():
    L = O.lib()
    rejected = [v for v in range(64) if not L.or_filter(v, 3, 1)]
    assert rejected == [_enc(b"AAA"), _enc(b"ACA"), _enc(b"AGA")] == [0, 4, 8]
    assert all(L.or_filter(v, 3, 0) for v in range(64))
    # p=11: exactly 3/64 of the values are rejected
    rng = np.random.default_rng(0)
    vs = rng.integers(0, 4 ** 11, 20000)
    rej = np.array([not L.or_filter(int(v), 11, 1) for v in vs])
    assert rej.sum() == sum(int(v) >> 16 in (0, 4, 8) for v in vs)


def test_code_table():
    L = O.lib()
    for ch, v in zip(b"ACGTacgt", [0, 1, 2, 3, 0, 1, 2, 3]):
        assert L.or_code(ch) == v
    assert L.or_code(ord("N")) == 255


@pytest.mark.parametrize("k", [1, 5, 21, 31, 32])
def test_canonical64_matches_strings(k):
    rng = np.random.default_rng(k)
    L = O.lib()
    for _ in range(200):
        s = bytes(rng.choice(list(b"ACGT"), size=k).tolist())
        rc = s.translate(COMP)[::-1]
        assert L.or_canonical64(_enc(s), k) == min(_enc(s), _enc(rc))
    assert L.or_canonical64(_enc(b"T" * k), k) == 0


def test_poly_a_read():
    """Every forward p-mer AAA.. fails new_filter2, the reverse complement TTT..
    passes with value 4^p-1: minimizer M; a single k-mer 0 counted L-k+1 times."""
    reads, offs = _csr([b"A" * 100])
    mm, pos, strand = O.window_minimizers(reads, offs, 31, 11, 1)
    assert np.all(mm[:70] == 4 ** 11 - 1)
    assert np.all(strand[:70] == 1)
    mm0, _, _ = O.window_minimizers(reads, offs, 31, 11, 0)
    assert np.all(mm0[:70] == 0)
    t = O.count(reads, offs, 31, 11, 64)
    assert t.key_lo.tolist() == [0] and t.cnt.tolist() == [70]
    assert int(t.part_off[(4 ** 11 - 1) % 64 + 1] - t.part_off[(4 ** 11 - 1) % 64]) == 1


def test_hpc_closed_form():
    reads, offs = _csr([b"AAACCGTT", b"", b"G", b"ACACAC", b"TTTTT"])
    out, rl, pos = O.hpc(reads, offs)
    assert rl.tolist() == [4, 0, 1, 6, 1]
    assert bytes(out[0:4]) == b"ACGT" and pos[0:4].tolist() == [0, 3, 5, 6]
    assert bytes(out[8:9]) == b"G"
    assert bytes(out[9:15]) == b"ACACAC" and pos[9:15].tolist() == list(range(6))
    assert bytes(out[15:16]) == b"T" and pos[15] == 0


def test_superkmer_lists_closed_form():
    # hand-made minimizer sequence for one read of length 10 with k=3: windows 0..7
    offs = np.array([0, 10], np.uint64)
    mm = np.array([5, 5, 7, 7, 7, 2, 2, 9, 0, 0], np.uint64)
    skm = O.superkmer_offs(offs, 3, mm)
    assert skm[:6].tolist() == [0, 2, 5, 7, 8, 0]


# ---------------------------------------------------------------- brute force

def _union(t):
    return dict(zip(t.key_lo.tolist(), t.cnt.tolist())) if t.key_hi is None else \
        {(int(h) << 64) | int(l): int(c) for l, h, c in zip(t.key_lo, t.key_hi, t.cnt)}


@pytest.mark.parametrize("cid,scale,k,p,P", [(1, 0.004, 31, 11, 64), (1, 0.004, 21, 7, 31), (2, 0.0004, 31, 11, 256),
                                             (4, 0.0003, 63, 19, 64), (4, 0.0003, 40, 12, 16), (1, 0.004, 9, 3, 5)])
def test_oracle_equals_brute_force(cid, scale, k, p, P):
    rs = W.generate(W.workload(cid, scale=scale))
    t = O.count(rs.reads, rs.offs, k, p, P)
    assert _union(t) == O.brute_force_table(rs.reads, rs.offs, k)
    assert int(t.cnt.sum()) == rs.kmers(k)
    assert int(t.part_kmers.sum()) == rs.kmers(k)
    # sorted within every partition
    for q in range(P):
        lo, hi, _ = t.part(q)
        keys = lo.astype(object) if hi is None else [(int(h) << 64) | int(l) for l, h in zip(lo, hi)]
        assert list(keys) == sorted(keys)


def test_oracle_hpc_equals_brute_force():
    rs = W.generate(W.workload(3, scale=0.0003))
    t = O.count(rs.reads, rs.offs, 25, 15, 512, hpc=True)
    assert _union(t) == O.brute_force_table(rs.reads, rs.offs, 25, hpc_on=True)


def test_count_filter():
    rs = W.generate(W.workload(1, scale=0.004))
    full = _union(O.count(rs.reads, rs.offs, 31, 11, 64))
    t = O.count(rs.reads, rs.offs, 31, 11, 64, min_count=2, max_count=5)
    assert _union(t) == {k: c for k, c in full.items() if 2 <= c <= 5}
    assert int(t.part_distinct.sum()) == len(full)


# ---------------------------------------------------------------- invariants

@pytest.mark.parametrize("k,p,fmode", [(31, 11, 1), (31, 11, 0), (25, 15, 1), (63, 19, 1), (20, 20, 1)])
def test_minimizer_reverse_complement_invariant(k, p, fmode):
    rng = np.random.default_rng(k * 100 + p)
    seqs = [bytes(rng.choice(list(b"ACGT"), size=k).tolist()) for _ in range(300)]
    seqs += [b"A" * k, b"T" * k, (b"AC" * k)[:k]]
    rc = [s.translate(COMP)[::-1] for s in seqs]
    r1, o1 = _csr(seqs)
    r2, o2 = _csr(rc)
    m1, _, _ = O.window_minimizers(r1, o1, k, p, fmode)
    m2, _, _ = O.window_minimizers(r2, o2, k, p, fmode)
    np.testing.assert_array_equal(m1[o1[:-1].astype(np.int64)], m2[o2[:-1].astype(np.int64)])


def test_minimizer_sliding_window_formulation():
    """mm(i) = min over j in [i, i+k-p] of v_j, v_j = min(M, passing f_j, passing r_j)."""
    rs = W.generate(W.workload(1, scale=0.002))
    k, p = 31, 11
    mm, _, _ = O.window_minimizers(rs.reads, rs.offs, k, p, 1)
    code = np.full(256, 0, np.int64)
    for ch, v in zip(b"ACGT", range(4)):
        code[ch] = v
    M = 4 ** p - 1
    for r in range(0, rs.n, 7):
        s = rs.reads[int(rs.offs[r]):int(rs.offs[r + 1])]
        c = code[s]
        n = len(c) - p + 1
        f = np.zeros(n, np.int64)
        rc = np.zeros(n, np.int64)
        for t in range(p):
            f = f * 4 + c[t:t + n]
            rc = rc * 4 + (3 - c[p - 1 - t:p - 1 - t + n])
        ok = lambda v: ~np.isin(v >> (2 * (p - 3)), [0, 4, 8])
        v = np.minimum(np.where(ok(f), f, M), np.where(ok(rc), rc, M))
        w = k - p + 1
        win = np.lib.stride_tricks.sliding_window_view(v, w).min(axis=1)
        np.testing.assert_array_equal(mm[int(rs.offs[r]):int(rs.offs[r]) + len(win)], win.astype(np.uint64))


def test_union_independent_of_partitioning():
    rs = W.generate(W.workload(1, scale=0.004))
    base = _union(O.count(rs.reads, rs.offs, 31, 11, 64))
    for P, fm, pm in [(1, 1, 0), (7, 1, 0), (256, 1, 1), (64, 0, 0), (13, 0, 1)]:
        t = O.count(rs.reads, rs.offs, 31, 11, P, filter_mode=fm, part_mode=pm)
        assert _union(t) == base
        # a canonical k-mer never lands in two partitions
        assert len(t.key_lo) == len(base)


def test_union_independent_of_read_order_and_strand():
    rs = W.generate(W.workload(1, scale=0.004))
    base = _union(O.count(rs.reads, rs.offs, 31, 11, 64))
    seqs = [bytes(rs.reads[int(rs.offs[i]):int(rs.offs[i + 1])]) for i in range(rs.n)]
    rng = np.random.default_rng(3)
    rng.shuffle(seqs)
    seqs = [s.translate(COMP)[::-1] if i % 2 else s for i, s in enumerate(seqs)]
    r2, o2 = _csr(seqs)
    assert _union(O.count(r2, o2, 31, 11, 64)) == base


def test_partition_rule_per_window():
    """Per-partition table = {x in union : the superkmer minimizer of x % P == part}:
    check with the window minimizer of every k-mer occurrence."""
    rs = W.generate(W.workload(1, scale=0.002))
    k, p, P = 31, 11, 64
    t = O.count(rs.reads, rs.offs, k, p, P)
    mm, _, _ = O.window_minimizers(rs.reads, rs.offs, k, p, 1)
    L = O.lib()
    where = {}
    for q in range(P):
        lo, _, _ = t.part(q)
        for key in lo.tolist():
            where[key] = q
    for r in range(0, rs.n, 5):
        a = int(rs.offs[r])
        s = bytes(rs.reads[a:int(rs.offs[r + 1])])
        for i in range(len(s) - k + 1):
            key = L.or_canonical64(_enc(s[i:i + k]), k)
            assert where[key] == int(mm[a + i]) % P
