"""Position-carrying partitioning (GKC_WITH_POS) on the GPU.

The reference declares the task (CountTask::SKMPartWithPos, gkc_cuda.hpp:50,
"count a portion of k-mers and store their positions", README.md:15-18) and
never implements it, so these tests pin the device listing against what its
description fixes:
  * superkmer positions: the (read, offset, k-mers) records of all
    partitions tile the reads' k-mer positions exactly once (windows of
    A/C/G/T bytes only, HPC coordinates with HPC), every record lies in one
    run of equal window minimizer (the oracle's per-window records) whose
    residue names its partition, and records are cut only at those runs'
    ends or where a 2048-window tile of a batch ends;
  * k-mer positions: equal to the pure-Python brute force over the read
    strings (oracle.brute_force_positions), sampled or not, over several
    batches (reads numbered across them), with a count filter, in the plain
    and batched forms;
  * the entry points' refusals.
"""
from __future__ import annotations

import numpy as np
import pytest

from gkc_amd import capi
from gkc_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FILL = np.iinfo(np.uint64).max


def _say(*a):
    print("[gpu-case]", *a, flush=True)


def _with_other_bytes(rs, seed, every=97):
    rng = np.random.default_rng(seed)
    reads = rs.reads.copy()
    junk = np.frombuffer(b"NNNNnRYKM-.*X", np.uint8)
    at = rng.choice(len(reads), size=len(reads) // every, replace=False)
    reads[at] = junk[rng.integers(len(junk), size=len(at))]
    for a in rng.choice(len(reads) - 40, size=max(1, len(reads) // 3000), replace=False):
        reads[a:a + int(rng.integers(2, 40))] = ord("N")
    return reads


def _partition(c, reads, offs, batches, flags):
    n = len(offs) - 1
    b = np.linspace(0, n, batches + 1).astype(int)
    for x, y in zip(b[:-1], b[1:]):
        if y > x:
            c.partition(reads[int(offs[x]):int(offs[y])], offs[x:y + 1] - offs[x], flags=flags | capi.WITH_POS)


def _windows(reads, offs, k, p, hpc):
    """Per read: window minimizers (FILL where the window holds another byte)."""
    if hpc:
        src, lens, _ = O.hpc(reads, offs)
        mm, _, _ = O.window_minimizers(src, offs, k, p, 1, lens=lens)
    else:
        lens = np.diff(offs).astype(np.int32)
        mm, _, _ = O.window_minimizers(reads, offs, k, p, 1)
    out = []
    for r in range(len(offs) - 1):
        a, L = int(offs[r] - offs[0]), int(lens[r])
        out.append(mm[a:a + max(L - k + 1, 0)])
    return out


@pytest.mark.parametrize("k,p,P,hpc,flags,batches", [(31, 11, 64, False, 0, 1), (31, 11, 64, False, capi.BATCHED, 3),
                                                     (25, 15, 512, True, 0, 2), (21, 9, 31, False, capi.BATCHED, 1)])
def test_superkmer_positions_tile_the_reads(gpu, k, p, P, hpc, flags, batches):
    rs = W.generate(W.workload(1, scale=0.01))
    reads = _with_other_bytes(rs, 5 + k)
    _say("superkmer positions", k, p, P, hpc, flags, batches)
    wins = _windows(reads, rs.offs, k, p, hpc)
    seen = [np.zeros(len(w), np.int32) for w in wins]
    n_rec = 0
    with capi.Counter(k, p, P, hpc=hpc) as c:
        _partition(c, reads, rs.offs, batches, flags)
        for q in range(P):
            rd, of, nk = c.superkmer_positions(q)
            n_rec += len(rd)
            for r, o, n in zip(rd.tolist(), of.tolist(), nk.tolist()):
                assert n >= 1
                w = wins[r][o:o + n]
                assert len(w) == n and w[0] != FILL and (w == w[0]).all(), (r, o, n)
                assert int(w[0]) % P == q
                seen[r][o:o + n] += 1
        c.count()
        assert n_rec == c.stats()["n_superkmers"]
    n_runs = 0
    for w, s in zip(wins, seen):
        valid = w != FILL
        np.testing.assert_array_equal(s, valid.astype(np.int32))
        starts = valid & np.concatenate([[True], (w[1:] != w[:-1]) | ~valid[:-1]]) if len(w) else valid
        n_runs += int(starts.sum())
    # cuts beyond the runs' ends: at most one per 2048-window tile of a batch
    n_bases = int(rs.offs[-1])
    assert n_runs <= n_rec <= n_runs + n_bases // 2048 + 2 * batches


def _positions(c, P, mod):
    got = {}
    for q in range(P):
        keys, cnts, off, pos = c.kmer_positions(q, mod)
        assert (np.diff(keys.astype(np.uint64)) > 0).all()
        assert int(off[-1]) == len(pos) and (np.diff(off.astype(np.int64)) == cnts).all()
        for e, v in enumerate(keys.tolist()):
            lst = pos[int(off[e]):int(off[e + 1])].tolist()
            assert lst == sorted(lst)
            got[v] = lst
    return got


@pytest.mark.parametrize("k,p,P,hpc,flags,batches,mod,min_count",
                         [(31, 11, 64, False, 0, 1, 1, 1), (31, 11, 64, False, capi.BATCHED, 3, 1, 1),
                          (25, 15, 128, True, 0, 2, 3, 1), (21, 7, 31, False, capi.BATCHED, 2, 7, 2),
                          (32, 13, 16, False, 0, 1, 1, 1)])
def test_kmer_positions_match_brute_force(gpu, k, p, P, hpc, flags, batches, mod, min_count):
    rs = W.generate(W.workload(1, scale=0.004))
    reads = _with_other_bytes(rs, 9 + k)
    _say("kmer positions", k, p, P, hpc, flags, batches, mod, min_count)
    want = O.brute_force_positions(reads, rs.offs, k, hpc_on=hpc, sample_mod=mod)
    want = {v: x for v, x in want.items() if len(x) >= min_count}
    with capi.Counter(k, p, P, hpc=hpc, min_count=min_count) as c:
        _partition(c, reads, rs.offs, batches, flags)
        c.count(flags)
        got = _positions(c, P, mod)
        lo, _, cnt, _, _, _ = c.table()
        tab = dict(zip(lo.tolist(), cnt.tolist()))
    assert got == want
    assert all(tab[v] == len(x) for v, x in got.items())


def test_positions_refusals(gpu):
    rs = W.generate(W.workload(1, scale=0.001))
    n = rs.n
    h = n // 2
    with capi.Counter(31, 11, 16) as c:
        with pytest.raises(capi.GkcError):           # nothing partitioned with positions
            c.superkmer_positions(0)
        c.partition(rs.reads[:int(rs.offs[h])], rs.offs[:h + 1], flags=capi.WITH_POS)
        with pytest.raises(capi.GkcError):           # batches disagree
            c.partition(rs.reads[int(rs.offs[h]):], rs.offs[h:] - rs.offs[h])
        with pytest.raises(capi.GkcError):           # not counted yet
            c.kmer_positions(0)
        with pytest.raises(capi.GkcError):
            c.superkmer_positions(16)
        c.count()
        with pytest.raises(capi.GkcError):
            c.kmer_positions(0, 0)
        c.reset()                                    # a reset lifts the mode
        c.partition(rs.reads, rs.offs)
        c.count()
        with pytest.raises(capi.GkcError):
            c.kmer_positions(0)
    with capi.Counter(41, 11, 16) as c:
        c.partition(rs.reads, rs.offs, flags=capi.WITH_POS)
        c.count()
        with pytest.raises(capi.GkcError):           # k > 32
            c.kmer_positions(0)
