"""Seeded synthetic workloads for the k-mer counting path.

The five workloads are fixed here once (the reference ships no data):

    C1  2 Mbp uniform genome, 100 000 x 150 bp reads, 0.5 % substitutions,
        k=31 p=11, 64 partitions                          (CPU-runnable case)
    C2  50 Mbp, 41 % GC, 10 % segmental repeats (1-5 kb, 1 % divergence),
        30x 150 bp, 0.3 % substitutions, k=31 p=11, 256 partitions, chunked
    C3  100 Mbp, 45 % GC, 3 % tandem repeats, log-normal 10 kb +- 6 kb reads,
        15x, 5 % errors (50/30/20 sub/ins/del), HPC, k=25 p=15, 512 partitions
    C4  100 Mbp uniform, 5 % segmental duplications (10-50 kb, 0.5 %),
        40x 250 bp, 0.2 % substitutions, k=63 p=19, 1024 partitions, 128-bit
        keys, count threshold 2
    C5  200 Mbp, 1 Mbp blocks alternating 35 % / 60 % GC, 8 % repeats (half
        tandem, half segmental), 30x 150 bp, 0.3 %, k=31 p=11, 1024 partitions

Seeds: genome seed = 1000 * config + g, reads seed = genome seed + 500000.
`scale` shrinks the genome (and the read count with it) while keeping the
structure, for tests.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import os

import numpy as np

from . import _lib


class GenomeCfg(ctypes.Structure):
    _fields_ = [
        ("genome_len", ctypes.c_uint64), ("gc", ctypes.c_double), ("gc_alt", ctypes.c_double),
        ("gc_block", ctypes.c_uint64), ("seg_frac", ctypes.c_double),
        ("seg_min", ctypes.c_uint32), ("seg_max", ctypes.c_uint32), ("seg_div", ctypes.c_double),
        ("tandem_frac", ctypes.c_double), ("unit_min", ctypes.c_uint32), ("unit_max", ctypes.c_uint32),
        ("copies_min", ctypes.c_uint32), ("copies_max", ctypes.c_uint32),
    ]


class ReadsCfg(ctypes.Structure):
    _fields_ = [
        ("n_reads", ctypes.c_uint64), ("coverage", ctypes.c_double), ("len_mode", ctypes.c_int),
        ("len_mean", ctypes.c_double), ("len_sd", ctypes.c_double), ("len_min", ctypes.c_uint32),
        ("rc_prob", ctypes.c_double), ("err_rate", ctypes.c_double),
        ("sub_w", ctypes.c_double), ("ins_w", ctypes.c_double), ("del_w", ctypes.c_double),
    ]


@dataclasses.dataclass
class Workload:
    name: str
    genome: dict
    reads: dict
    k: int
    p: int
    n_parts: int
    hpc: bool = False
    min_count: int = 1
    max_count: int = 0          # 0: unbounded
    filter_mode: int = 1        # 1: new_filter2 (reference default), 0: accept all
    part_mode: int = 0          # 0: minimizer % P (reference), 1: hashed
    chunked: bool = False
    emit_windows: bool = False
    config_id: int = 0

    def describe(self) -> dict:
        return {"workload": self.name, "k": self.k, "p": self.p, "partitions": self.n_parts,
                "hpc": self.hpc, "min_count": self.min_count,
                "genome_bp": self.genome["genome_len"]}


def _wl(cid: int) -> Workload:
    G = dict(gc=0.5, gc_alt=0.5, gc_block=0, seg_frac=0.0, seg_min=0, seg_max=0, seg_div=0.0,
             tandem_frac=0.0, unit_min=0, unit_max=0, copies_min=0, copies_max=0)
    R = dict(n_reads=0, coverage=0.0, len_mode=0, len_mean=150.0, len_sd=0.0, len_min=1,
             rc_prob=0.5, err_rate=0.0, sub_w=1.0, ins_w=0.0, del_w=0.0)
    if cid == 1:
        G.update(genome_len=2_000_000)
        R.update(n_reads=100_000, err_rate=0.005)
        return Workload("C1 2Mbp 100k x 150bp k31 p11 P64", G, R, 31, 11, 64, config_id=1)
    if cid == 2:
        G.update(genome_len=50_000_000, gc=0.41, gc_alt=0.41, seg_frac=0.10, seg_min=1000,
                 seg_max=5000, seg_div=0.01)
        R.update(coverage=30.0, err_rate=0.003)
        return Workload("C2 50Mbp 30x 150bp k31 p11 P256", G, R, 31, 11, 256, chunked=True,
                        config_id=2)
    if cid == 3:
        G.update(genome_len=100_000_000, gc=0.45, gc_alt=0.45, tandem_frac=0.03, unit_min=2,
                 unit_max=50, copies_min=5, copies_max=200)
        R.update(coverage=15.0, len_mode=1, len_mean=10000.0, len_sd=6000.0, len_min=25,
                 err_rate=0.05, sub_w=0.5, ins_w=0.3, del_w=0.2)
        return Workload("C3 100Mbp 15x longreads HPC k25 p15 P512", G, R, 25, 15, 512, hpc=True,
                        emit_windows=True, config_id=3)
    if cid == 4:
        G.update(genome_len=100_000_000, seg_frac=0.05, seg_min=10_000, seg_max=50_000,
                 seg_div=0.005)
        R.update(coverage=40.0, len_mean=250.0, err_rate=0.002)
        return Workload("C4 100Mbp 40x 250bp k63 p19 P1024 u128", G, R, 63, 19, 1024,
                        min_count=2, config_id=4)
    if cid == 5:
        G.update(genome_len=200_000_000, gc=0.35, gc_alt=0.60, gc_block=1_000_000,
                 seg_frac=0.04, seg_min=1000, seg_max=5000, seg_div=0.01,
                 tandem_frac=0.04, unit_min=2, unit_max=50, copies_min=5, copies_max=200)
        R.update(coverage=30.0, err_rate=0.003)
        return Workload("C5 200Mbp skewedGC 30x 150bp k31 p11 P1024", G, R, 31, 11, 1024,
                        config_id=5)
    raise ValueError(f"unknown workload {cid}")


def workload(cid: int, scale: float = 1.0, **over) -> Workload:
    """Workload `cid` (1-5); `scale` < 1 shrinks the genome and read count."""
    w = _wl(cid)
    if scale != 1.0:
        w.genome["genome_len"] = max(int(w.genome["genome_len"] * scale), 20_000)
        if w.reads["n_reads"]:
            w.reads["n_reads"] = max(int(w.reads["n_reads"] * scale), 100)
        # keep repeat units inside the smaller genome
        for key in ("seg_min", "seg_max"):
            w.genome[key] = min(w.genome[key], w.genome["genome_len"] // 20)
        if scale < 1.0:
            w.name += f" (scale {scale:g})"
    for key, val in over.items():
        setattr(w, key, val)
    return w


def genome_seed(cid: int, g: int = 0) -> int:
    return 1000 * cid + g


@dataclasses.dataclass
class ReadSet:
    reads: np.ndarray        # uint8 bases (ASCII)
    offs: np.ndarray         # uint64, n+1 (absolute, offs[0] == 0)
    genome: np.ndarray | None = None

    @property
    def n(self) -> int:
        return len(self.offs) - 1

    @property
    def bases(self) -> int:
        return int(self.offs[-1] - self.offs[0])

    def kmers(self, k: int, hpc: bool = False) -> int:
        """k-mer occurrences: sum over reads of max(0, len - k + 1), with len
        the homopolymer-compressed length (number of runs of equal bytes)
        when hpc."""
        lens = np.diff(self.offs).astype(np.int64)
        if hpc:
            lens = self.hpc_lengths()
        return int(np.maximum(lens - k + 1, 0).sum())

    def hpc_lengths(self) -> np.ndarray:
        """Runs of identical bytes per read (the compressed read length)."""
        n = self.n
        lens = np.diff(self.offs).astype(np.int64)
        out = np.zeros(n, np.int64)
        step = 1 << 26                      # bounded memory: blocks of whole reads
        r = 0
        while r < n:
            r1 = int(np.searchsorted(self.offs, self.offs[r] + step, side="right")) - 1
            r1 = max(r1, r + 1)
            a, b = int(self.offs[r]), int(self.offs[r1])
            seg = self.reads[a:b]
            change = np.empty(b - a, np.bool_)
            if b > a:
                change[0] = True
                np.not_equal(seg[1:], seg[:-1], out=change[1:])
                starts = (self.offs[r:r1] - a).astype(np.int64)
                nz = lens[r:r1] > 0
                change[starts[nz]] = True      # a read's first byte starts a run
                cs = np.concatenate([[0], np.cumsum(change, dtype=np.int64)])
                ends = (self.offs[r + 1:r1 + 1] - a).astype(np.int64)
                out[r:r1] = cs[ends] - cs[starts]
            r = r1
        return out


def generate(w: Workload, g: int = 0, threads: int | None = None, keep_genome: bool = False) -> ReadSet:
    """Generate the genome and reads of workload `w` for genome index `g`."""
    lib = _lib.synth()
    gc = GenomeCfg(**w.genome)
    rc = ReadsCfg(**w.reads)
    gseed = genome_seed(w.config_id, g)
    rseed = gseed + 500_000
    G = w.genome["genome_len"]
    genome = np.empty(G, dtype=np.uint8)
    if lib.gkc_synth_genome(ctypes.byref(gc), ctypes.c_uint64(gseed), genome.ctypes.data_as(ctypes.c_char_p)) != 0:
        raise RuntimeError("genome generation failed")
    n = lib.gkc_synth_read_count(ctypes.byref(rc), ctypes.c_uint64(G))
    bound = lib.gkc_synth_bytes_bound(ctypes.byref(rc), ctypes.c_uint64(G), ctypes.c_uint64(rseed))
    reads = np.empty(max(bound, 1), dtype=np.uint8)
    offs = np.empty(n + 1, dtype=np.uint64)
    th = threads or min(os.cpu_count() or 1, 16)
    tot = lib.gkc_synth_reads(ctypes.byref(rc), genome.ctypes.data_as(ctypes.c_char_p), ctypes.c_uint64(G),
                              ctypes.c_uint64(rseed), th, reads.ctypes.data_as(ctypes.c_char_p),
                              offs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return ReadSet(reads[:tot], offs, genome if keep_genome else None)


def chunk_bounds(rs: ReadSet, seed: int, lo_bytes: int = 64 << 20, hi_bytes: int = 1 << 30,
                 fasta_overhead: bool = True) -> list[tuple[int, int]]:
    """Cut a read set at record boundaries into chunks whose FASTA sizes are
    log-uniform in [lo_bytes, hi_bytes] (drawn from `seed`).  Returns
    [(first_read, end_read)].  The FASTA size of read i is its length plus
    the header and newlines."""
    rng = np.random.default_rng(seed)
    lens = np.diff(rs.offs).astype(np.int64)
    rec = lens + (len(str(rs.n)) + 3 if fasta_overhead else 0)
    csum = np.concatenate([[0], np.cumsum(rec)])
    out, r = [], 0
    while r < rs.n:
        size = math.exp(rng.uniform(math.log(lo_bytes), math.log(hi_bytes)))
        e = int(np.searchsorted(csum, csum[r] + size, side="right")) - 1
        e = min(max(e, r + 1), rs.n)
        out.append((r, e))
        r = e
    return out


def write_fasta(path: str, rs: ReadSet) -> None:
    lib = _lib.synth()
    if lib.gkc_write_fasta(path.encode(), rs.reads.ctypes.data_as(ctypes.c_char_p),
                           rs.offs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.c_uint64(rs.n)) != 0:
        raise OSError(f"cannot write {path}")


def read_fasta(path: str) -> ReadSet:
    lib = _lib.synth()
    tot = ctypes.c_uint64(0)
    n = lib.gkc_read_fasta(path.encode(), None, None, ctypes.byref(tot))
    if n < 0:
        raise OSError(f"cannot read {path}")
    reads = np.empty(max(tot.value, 1), dtype=np.uint8)
    offs = np.empty(n + 1, dtype=np.uint64)
    lib.gkc_read_fasta(path.encode(), reads.ctypes.data_as(ctypes.c_char_p),
                       offs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(tot))
    return ReadSet(reads[:tot.value], offs)


def reads_text(reads: np.ndarray, offs: np.ndarray, fmt: str = "fastq", crlf: bool = False,
               final_newline: bool = True) -> np.ndarray:
    """FASTQ ('@r' header, the sequence, '+', a quality line of printable
    bytes derived from the bases, which include '@' and '+') or single-line
    FASTA ('>r' header, the sequence) text of a read CSR, as uint8."""
    reads = np.ascontiguousarray(reads, np.uint8)
    offs = np.ascontiguousarray(offs, np.uint64).astype(np.int64)
    lens = np.diff(offs)
    nl = b"\r\n" if crlf else b"\n"
    hdr = (b"@r" if fmt == "fastq" else b">r") + nl
    sep = b"+" + nl
    fq = fmt == "fastq"
    rec = len(hdr) + lens + len(nl) + ((len(sep) + lens + len(nl)) if fq else 0)
    start = np.zeros(len(lens), np.int64)
    np.cumsum(rec[:-1], out=start[1:])
    total = int(rec.sum())
    out = np.empty(total, np.uint8)
    ar = np.arange(int(offs[-1] - offs[0]), dtype=np.int64)
    base0 = offs[:-1] - offs[0]

    def put(pos, bs):
        for j, ch in enumerate(bs):
            out[pos + j] = ch
    put(start, hdr)
    s0 = start + len(hdr)
    out[np.repeat(s0 - base0, lens) + ar] = reads[offs[0]:offs[-1]]
    put(s0 + lens, nl)
    if fq:
        p = s0 + lens + len(nl)
        put(p, sep)
        q0 = p + len(sep)
        out[np.repeat(q0 - base0, lens) + ar] = (reads[offs[0]:offs[-1]] % 40) + 33
        put(q0 + lens, nl)
    if not final_newline and total:
        out = out[:total - len(nl)]
    return out
