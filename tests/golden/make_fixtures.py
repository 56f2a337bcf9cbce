"""Writes the small golden fixtures of tests/golden/ (run from the repo root:
`python tests/golden/make_fixtures.py`).

The reference ships no golden vectors for this path, so these are produced by
the oracle and accepted only after an independent brute-force check (string
canonical form + dict counting) of every table.  Each fixture holds the
FASTA-equivalent CSR input, the per-window minimizers, the superkmer offset
lists in the reference's layout and the per-partition (k-mer, count) tables.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gkc_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    # name: (workload id, genome bp, n_reads or None, k, p, P, hpc, min_count)
    "c1_k31_p11_P64": (1, 10_000, 200, 31, 11, 64, False, 1),
    "c3_hpc_k25_p15_P512": (3, 20_000, None, 25, 15, 512, True, 1),
    "c4_k63_p19_P64_min2": (4, 20_000, None, 63, 19, 64, False, 2),
}


def make(name: str):
    cid, G, nreads, k, p, P, hpc, minc = CASES[name]
    w = W.workload(cid)
    w.genome["genome_len"] = G
    w.genome["seg_min"] = min(w.genome["seg_min"], G // 20)
    w.genome["seg_max"] = min(w.genome["seg_max"], G // 20)
    if nreads:
        w.reads["n_reads"] = nreads
    else:
        w.reads["coverage"] = min(w.reads["coverage"], 8.0)
    if cid == 3:
        w.reads["len_mean"], w.reads["len_sd"] = 1500.0, 900.0
    rs = W.generate(w)
    t = O.count(rs.reads, rs.offs, k, p, P, hpc=hpc, min_count=minc)
    bf = O.brute_force_table(rs.reads, rs.offs, k, hpc_on=hpc)
    keys = t.key_lo.astype(object) if t.key_hi is None else \
        [(int(h) << 64) | int(l) for l, h in zip(t.key_lo, t.key_hi)]
    got = dict(zip([int(x) for x in keys], t.cnt.tolist()))
    want = {key: c for key, c in bf.items() if c >= minc}
    assert got == want, f"{name}: oracle disagrees with brute force"
    lens = None
    src = rs.reads
    if hpc:
        src, lens, _ = O.hpc(rs.reads, rs.offs)
    mm, pos, strand = O.window_minimizers(src, rs.offs, k, p, 1, lens=lens)
    skm = O.superkmer_offs(rs.offs, k, mm, lens=lens)
    return dict(reads=rs.reads, offs=rs.offs, k=k, p=p, P=P, hpc=int(hpc), min_count=minc,
                minimizers=mm, mm_pos=pos, mm_strand=strand, superkmer_offs=skm,
                key_lo=t.key_lo, key_hi=(t.key_hi if t.key_hi is not None else np.zeros(0, np.uint64)),
                cnt=t.cnt, part_off=t.part_off, part_kmers=t.part_kmers, part_distinct=t.part_distinct)


def main():
    for name in CASES:
        d = make(name)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print(name, "reads", len(d["offs"]) - 1, "entries", len(d["cnt"]))


if __name__ == "__main__":
    main()
