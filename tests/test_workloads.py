"""Host helpers of the bench: k-mer totals of a read set (the unit of work),
with homopolymer compression counted like the oracle counts it."""
from __future__ import annotations

import numpy as np

from gkc_amd import workloads as W
from oracle import oracle as O


def test_kmer_total_plain_and_hpc():
    w = W.workload(3, scale=0.002)
    rs = W.generate(w)
    for hpc in (False, True):
        t = O.count(rs.reads, rs.offs, w.k, w.p, 8, hpc=hpc)
        assert rs.kmers(w.k, hpc=hpc) == int(t.part_kmers.sum())


def test_hpc_lengths_closed_form():
    seqs = [b"", b"A", b"AAAA", b"ACGT", b"AACCGGTT", b"aaAA", b"ACCCA" * 3]
    reads = np.frombuffer(b"".join(seqs), np.uint8).copy()
    offs = np.concatenate([[0], np.cumsum([len(s) for s in seqs])]).astype(np.uint64)
    rs = W.ReadSet(reads, offs)
    expect = [0, 1, 1, 4, 4, 2, 7]            # runs of identical bytes (case-sensitive)
    assert rs.hpc_lengths().tolist() == expect
