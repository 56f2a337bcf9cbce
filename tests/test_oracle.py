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


def _with_other_bytes(rs, rng, every=97):
    """The workload's reads with bytes other than A/C/G/T written in: single
    N, runs of N, lowercase n and IUPAC / gap characters."""
    reads = rs.reads.copy()
    junk = np.frombuffer(b"NNNNnRYKM-.*X", np.uint8)
    at = rng.choice(len(reads), size=len(reads) // every, replace=False)
    reads[at] = junk[rng.integers(len(junk), size=len(at))]
    for a in rng.choice(len(reads) - 40, size=len(reads) // 3000, replace=False):
        reads[a:a + int(rng.integers(2, 40))] = ord("N")
    return reads


def _split_frags(reads, offs):
    """Every read cut into its runs of A/C/G/T bytes."""
    import re
    frags = []
    for r in range(len(offs) - 1):
        s = bytes(reads[int(offs[r]):int(offs[r + 1])])
        frags += [f for f in re.split(rb"[^ACGTacgt]+", s) if f]
    return frags


def _split_at_other_bytes(reads, offs):
    """The fragments of _split_frags as a read CSR."""
    return _csr(_split_frags(reads, offs))


@pytest.mark.parametrize("hpc,k,p,P", [(False, 31, 11, 64), (True, 25, 15, 512), (False, 63, 19, 64)])
def test_non_acgt_bytes_split_reads(hpc, k, p, P):
    """A byte other than A/C/G/T splits its read: the oracle's table equals
    the brute force that skips k-mers holding one, and equals the oracle's
    table of the reads cut at those bytes beforehand (after HPC, which
    collapses runs of N like any run)."""
    rng = np.random.default_rng(7 + k)
    rs = W.generate(W.workload(1, scale=0.01))
    reads = _with_other_bytes(rs, rng)
    t = O.count(reads, rs.offs, k, p, P, hpc=hpc)
    if hpc:
        src, lens, _ = O.hpc(reads, rs.offs)
        seqs = [bytes(src[int(rs.offs[r]):int(rs.offs[r]) + int(lens[r])]) for r in range(rs.n)]
        hr, ho = _csr(seqs)
        fr, fo = _split_at_other_bytes(hr, ho)
    else:
        fr, fo = _split_at_other_bytes(reads, rs.offs)
    t2 = O.count(fr, fo, k, p, P)
    for a in ("key_lo", "cnt", "part_off", "part_kmers"):
        np.testing.assert_array_equal(getattr(t, a), getattr(t2, a), err_msg=a)
    if k <= 32:
        bf = O.brute_force_table(reads, rs.offs, k, hpc_on=hpc)
        assert _union(t) == bf
    assert int(t.part_kmers.sum()) < rs.kmers(k)


def test_non_acgt_window_records():
    """Per-window records are written for windows of k A/C/G/T bytes only;
    those equal the records of the same window in the split reads."""
    k, p = 25, 15
    reads, offs = _csr([b"ACGTTGCAAGGCTTACGATCGATCGGATNCCGATTAGCATGCATCGATCGACTAGCTAGCAT", b"N" * 30,
                        b"ACGGTACGTAGCTAGCTAGTCGATCGTAGCTAGCTAGnTCGATCGATCGA"])
    got, gpos, gst = O.window_minimizers(reads, offs, k, p, 1)      # unwritten: strand -2
    written = 0
    for r in range(3):
        s = bytes(reads[int(offs[r]):int(offs[r + 1])])
        for i in range(len(s) - k + 1):
            w = s[i:i + k]
            if w.upper().strip(b"ACGT"):
                assert gst[int(offs[r]) + i] == -2
                continue
            written += 1
            wr, wo = _csr([w])
            m1, p1, s1 = O.window_minimizers(wr, wo, k, p, 1)
            a = int(offs[r]) + i
            assert (got[a], gpos[a], gst[a]) == (m1[0], p1[0], s1[0])
    assert written == sum(max(0, len(f) - k + 1) for f in _split_frags(reads, offs))


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


# ---------------------------------------------------------------- per-window (minimizer, position, strand)

def _window_rule(w: bytes, p: int, fmode: int):
    """Brute force over all 2(k-p+1) candidates of one window: the smallest
    passing value; among the candidates holding it, the largest start; at
    that start the reverse complement (strand 1) over the forward p-mer
    (strand 0), -1 when both hold it (f == rc); (4^p-1, 255, -1) when no
    candidate passes."""
    M = 4 ** p - 1
    cands = []
    for j in range(len(w) - p + 1):
        sub = w[j:j + p]
        f, r = _enc(sub), _enc(sub.translate(COMP)[::-1])
        for v, st in ((f, 0), (r, 1)):
            top = v >> (2 * (p - 3))
            if fmode == 0 or top not in (0, 4, 8):
                cands.append((v, j, st))
    if not cands:
        return M, 255, -1
    v = min(c[0] for c in cands)
    J = max(c[1] for c in cands if c[0] == v)
    sts = {c[2] for c in cands if c[0] == v and c[1] == J}
    return v, J, (-1 if len(sts) == 2 else sts.pop())


def _windows(rng, k, n, alphabet):
    out = []
    for _ in range(n):
        a = alphabet[rng.integers(len(alphabet))]
        out.append(bytes(rng.choice(list(a), size=k).tolist()))
    out += [b"A" * k, b"T" * k, (b"ACGT" * k)[:k], (b"AT" * k)[:k], (b"GC" * k)[:k], (b"AAC" * k)[:k]]
    return out


@pytest.mark.parametrize("k,p,fmode", [(9, 3, 1), (9, 4, 1), (10, 4, 0), (12, 5, 1), (14, 6, 0),
                                       (25, 15, 1), (31, 11, 1), (20, 8, 1)])
def test_window_position_strand_brute_force(k, p, fmode):
    """The oracle's per-window records equal an independent brute force of
    the rule over every candidate, on random windows over small alphabets
    (many ties of value and of strand) and periodic and homopolymer ones;
    even p gives palindromic p-mers (f == rc, strand -1)."""
    rng = np.random.default_rng(1000 * k + p)
    seqs = _windows(rng, k, 400, [b"ACGT", b"AC", b"AT", b"GC", b"ACG", b"AAT"])
    reads, offs = _csr(seqs)
    mm, pos, strand = O.window_minimizers(reads, offs, k, p, fmode)
    seen_tie = seen_pal = False
    for i, w in enumerate(seqs):
        a = int(offs[i])
        want = _window_rule(w, p, fmode)
        assert (int(mm[a]), int(pos[a]), int(strand[a])) == want, (w, want)
        seen_pal |= want[2] == -1 and want[1] != 255
        seen_tie |= sum(1 for j in range(k - p + 1) if want[0] in (_enc(w[j:j + p]), _enc(w[j:j + p].translate(COMP)[::-1]))) > 1
    assert seen_tie
    if p % 2 == 0:
        assert seen_pal


@pytest.mark.parametrize("k,p", [(25, 15), (31, 11), (12, 4)])
def test_window_record_invariants(k, p):
    """Closed-form checks of the records on generated reads: the p-mer at
    (pos, strand) has the reported value; for a window whose minimum sits at
    one start only, its reverse complement reports the mirrored start
    k-p-pos and the other strand."""
    rs = W.generate(W.workload(1, scale=0.002))
    mm, pos, strand = O.window_minimizers(rs.reads, rs.offs, k, p, 1)
    checked = mirrored = 0
    for r in range(0, rs.n, 5):
        s = bytes(rs.reads[int(rs.offs[r]):int(rs.offs[r + 1])])
        for i in range(0, len(s) - k + 1, 3):
            a = int(rs.offs[r]) + i
            w = s[i:i + k]
            if pos[a] == 255:
                continue
            sub = w[pos[a]:pos[a] + p]
            f, rc = _enc(sub), _enc(sub.translate(COMP)[::-1])
            want = {0: f, 1: rc, -1: f}[int(strand[a])]
            assert want == int(mm[a]) and (strand[a] != -1 or f == rc)
            checked += 1
            starts = [j for j in range(k - p + 1)
                      if int(mm[a]) in (_enc(w[j:j + p]), _enc(w[j:j + p].translate(COMP)[::-1]))]
            if len(starts) == 1:
                rw = w.translate(COMP)[::-1]
                r2, o2 = _csr([rw])
                m2, p2, s2 = O.window_minimizers(r2, o2, k, p, 1)
                assert int(m2[0]) == int(mm[a]) and int(p2[0]) == k - p - int(pos[a])
                assert int(s2[0]) == {0: 1, 1: 0, -1: -1}[int(strand[a])]
                mirrored += 1
    assert checked > 1000 and mirrored > 500


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


# ---------------------------------------------------------------- k-mer positions


def test_mix64_matches_partition_hash():
    """The sampling hash is the one partition mode 1 applies in the C oracle
    (mix64(mm) % P): equal residues for several moduli pin the 64-bit value."""
    L = O.lib()
    rng = np.random.default_rng(11)
    for v in [0, 1, 2, 3, (1 << 62) - 1] + rng.integers(0, 1 << 62, size=40).tolist():
        for P in (7, 64, 1021, 65535, (1 << 32) - 5):
            assert L.or_partition(v, P, 1) == O.mix64(v) % P


@pytest.mark.parametrize("k,hpc", [(21, False), (31, False), (25, True), (32, False)])
def test_positions_brute_force_pins(k, hpc):
    """Every listed occurrence spells its k-mer: the bases at (read, offset)
    have the listed canonical value, the strand says which form was read;
    the occurrences per k-mer are the brute-force counts; sampling keeps the
    residue class exactly."""
    rng = np.random.default_rng(k)
    rs = W.generate(W.workload(1, scale=0.002))
    reads = _with_other_bytes(rs, rng)
    L = O.lib()
    tab = O.brute_force_table(reads, rs.offs, k, hpc_on=hpc)
    pos = O.brute_force_positions(reads, rs.offs, k, hpc_on=hpc)
    assert {v: len(x) for v, x in pos.items()} == tab
    if hpc:
        src, lens, _ = O.hpc(reads, rs.offs)
        seqs = [bytes(src[int(rs.offs[r]):int(rs.offs[r]) + int(lens[r])]) for r in range(rs.n)]
    else:
        seqs = [bytes(reads[int(rs.offs[r]):int(rs.offs[r + 1])]) for r in range(rs.n)]
    for v, lst in list(pos.items())[::7]:
        for e in lst:
            r, i, st = e >> 32, (e & 0xFFFFFFFF) >> 1, e & 1
            fw = _enc(seqs[r][i:i + k])
            assert L.or_canonical64(fw, k) == v
            assert (fw != v) == bool(st)
    sub = O.brute_force_positions(reads, rs.offs, k, hpc_on=hpc, sample_mod=5)
    assert sub == {v: x for v, x in pos.items() if O.mix64(v) % 5 == 0}
    assert 0 < len(sub) < len(pos)


def test_positions_reverse_complement_mirror():
    """Reverse-complementing every read maps each occurrence (r, i, strand)
    of a non-palindromic k-mer to (r, L - k - i, 1 - strand)."""
    k = 21
    rs = W.generate(W.workload(1, scale=0.002))
    seqs = [bytes(rs.reads[int(rs.offs[r]):int(rs.offs[r + 1])]) for r in range(rs.n)]
    r2, o2 = _csr([s.translate(COMP)[::-1] for s in seqs])
    a = O.brute_force_positions(rs.reads, rs.offs, k)
    b = O.brute_force_positions(r2, o2, k)
    assert a.keys() == b.keys()
    lens = np.diff(rs.offs).astype(np.int64)
    for v, lst in a.items():
        want = sorted(e >> 32 << 32 | (int(lens[e >> 32]) - k - ((e & 0xFFFFFFFF) >> 1)) << 1 | (1 - (e & 1))
                      for e in lst)
        assert b[v] == want
