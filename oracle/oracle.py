"""Python bindings of the CPU oracle (oracle/gkc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.  See the header
of gkc_oracle.c for the reference file:line each function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i8p = ctypes.POINTER(ctypes.c_int8)
_f64p = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.or_code.restype = ctypes.c_uint8
        L.or_code.argtypes = [ctypes.c_uint8]
        L.or_filter.restype = ctypes.c_int
        L.or_filter.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
        L.or_partition.restype = ctypes.c_uint32
        L.or_partition.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int]
        L.or_canonical64.restype = ctypes.c_uint64
        L.or_canonical64.argtypes = [ctypes.c_uint64, ctypes.c_int]
        L.or_hpc.restype = None
        L.or_hpc.argtypes = [_u8p, _u64p, ctypes.c_int64, _u8p, _i32p, _i32p]
        L.or_window_minimizers.restype = None
        L.or_window_minimizers.argtypes = [_u8p, _u64p, _i32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, _u64p, _u8p, _i8p]
        L.or_superkmer_offs.restype = None
        L.or_superkmer_offs.argtypes = [_u64p, _i32p, ctypes.c_int64, ctypes.c_int, _u64p, _i32p]
        L.or_count.restype = ctypes.c_int64
        L.or_count.argtypes = [_u8p, _u64p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                               ctypes.c_int, _u64p, _u64p, _u32p, ctypes.c_uint64, _u64p, _u64p, _u64p,
                               _u64p, _f64p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def hpc(reads: np.ndarray, offs: np.ndarray):
    """-> (hpc_reads uint8 [B], read_len int32 [n], orig_pos int32 [B])"""
    n = len(offs) - 1
    B = int(offs[-1] - offs[0])
    out = np.zeros(max(B, 1), np.uint8)
    rl = np.zeros(max(n, 1), np.int32)
    pos = np.zeros(max(B, 1), np.int32)
    lib().or_hpc(_p(reads, _u8p), _p(offs, _u64p), n, _p(out, _u8p), _p(rl, _i32p), _p(pos, _i32p))
    return out[:B], rl[:n], pos[:B]


def window_minimizers(reads, offs, k, p, filter_mode=1, lens=None):
    """Per-window minimizer / position / strand at batch-local base indices
    (entries outside [0, len-k] of each read are left at their fill value)."""
    n = len(offs) - 1
    B = int(offs[-1] - offs[0])
    mm = np.full(max(B, 1), np.iinfo(np.uint64).max, np.uint64)
    pos = np.full(max(B, 1), 255, np.uint8)
    strand = np.full(max(B, 1), -2, np.int8)
    lp = _p(np.ascontiguousarray(lens, np.int32), _i32p) if lens is not None else None
    lib().or_window_minimizers(_p(reads, _u8p), _p(offs, _u64p), lp, n, k, p, filter_mode,
                               _p(mm, _u64p), _p(pos, _u8p), _p(strand, _i8p))
    return mm[:B], pos[:B], strand[:B]


def superkmer_offs(offs, k, mm, lens=None):
    n = len(offs) - 1
    B = int(offs[-1] - offs[0])
    skm = np.full(max(B, 1), -7, np.int32)
    lp = _p(np.ascontiguousarray(lens, np.int32), _i32p) if lens is not None else None
    lib().or_superkmer_offs(_p(offs, _u64p), lp, n, k, _p(mm, _u64p), _p(skm, _i32p))
    return skm[:B]


class Table:
    """Per-partition (k-mer, count) tables concatenated in partition order."""

    def __init__(self, key_lo, key_hi, cnt, part_off, part_kmers, part_distinct, part_skms, times):
        self.key_lo, self.key_hi, self.cnt = key_lo, key_hi, cnt
        self.part_off, self.part_kmers = part_off, part_kmers
        self.part_distinct, self.part_skms, self.times = part_distinct, part_skms, times

    def part(self, i):
        a, b = int(self.part_off[i]), int(self.part_off[i + 1])
        hi = None if self.key_hi is None else self.key_hi[a:b]
        return self.key_lo[a:b], hi, self.cnt[a:b]


def count(reads, offs, k, p, n_parts, hpc=False, filter_mode=1, part_mode=0, min_count=1,
          max_count=0, threads=None) -> Table:
    n = len(offs) - 1
    lens = np.diff(offs).astype(np.int64)
    cap = int(np.maximum(lens - k + 1, 0).sum())
    key_lo = np.zeros(max(cap, 1), np.uint64)
    key_hi = np.zeros(max(cap, 1), np.uint64) if k > 32 else None
    cnt = np.zeros(max(cap, 1), np.uint32)
    part_off = np.zeros(n_parts + 1, np.uint64)
    pk = np.zeros(n_parts, np.uint64)
    pd = np.zeros(n_parts, np.uint64)
    ps = np.zeros(n_parts, np.uint64)
    t = np.zeros(2, np.float64)
    th = threads or min(os.cpu_count() or 1, 16)
    r = lib().or_count(_p(reads, _u8p), _p(offs, _u64p), n, k, p, int(hpc), n_parts, filter_mode, part_mode,
                       min_count, max_count, th, _p(key_lo, _u64p),
                       _p(key_hi, _u64p) if key_hi is not None else None, _p(cnt, _u32p), max(cap, 1),
                       _p(part_off, _u64p), _p(pk, _u64p), _p(pd, _u64p), _p(ps, _u64p), _p(t, _f64p))
    if r < 0:
        raise ValueError("oracle rejected the arguments")
    return Table(key_lo[:r], None if key_hi is None else key_hi[:r], cnt[:r], part_off, pk, pd, ps, t)


# ---- independent brute force (pure Python; tiny inputs only) ----

_COMP = bytes.maketrans(b"ACGTacgt", b"TGCAtgca")
_CODE = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3, ord("a"): 0, ord("c"): 1, ord("g"): 2, ord("t"): 3}


def brute_force_table(reads: np.ndarray, offs: np.ndarray, k: int, hpc_on: bool = False) -> dict:
    """Canonical k-mer -> count with a dict over the read strings: the
    canonical form is taken as the lexicographically smaller of the k-mer
    string and its reverse complement string (A<C<G<T), an independent
    statement of min(x, rc(x)) on 2-bit values.  A k-mer holding a byte
    other than A/C/G/T (either case) is not counted."""
    table: dict[int, int] = {}
    for r in range(len(offs) - 1):
        s = bytes(reads[int(offs[r] - offs[0]):int(offs[r + 1] - offs[0])])
        if hpc_on and s:
            s = bytes([s[0]] + [s[i] for i in range(1, len(s)) if s[i] != s[i - 1]])
        for i in range(len(s) - k + 1):
            fw = s[i:i + k].upper()
            if fw.strip(b"ACGT"):
                continue
            rc = fw.translate(_COMP)[::-1]
            c = min(fw, rc)
            v = 0
            for ch in c:
                v = (v << 2) | _CODE[ch]
            table[v] = table.get(v, 0) + 1
    return table


def mix64(v: int) -> int:
    """murmur3's 64-bit finaliser (fmix64), the sampling hash of the
    position-carrying listing."""
    m = (1 << 64) - 1
    v ^= v >> 33
    v = (v * 0xFF51AFD7ED558CCD) & m
    v ^= v >> 33
    v = (v * 0xC4CEB9FE1A85EC53) & m
    v ^= v >> 33
    return v


def brute_force_positions(reads: np.ndarray, offs: np.ndarray, k: int, hpc_on: bool = False,
                          sample_mod: int = 1) -> dict:
    """Canonical k-mer -> ascending list of read << 32 | offset << 1 | strand
    for every occurrence, over the read strings as brute_force_table reads
    them (offsets in HPC coordinates when hpc_on; strand 1 when the reverse
    complement string is the smaller), keeping the k-mers with
    mix64(value) % sample_mod == 0.  The reference declares this task
    (CountTask::SKMPartWithPos, gkc_cuda.hpp:50) without implementing it, so
    the listing is stated here from its description alone."""
    out: dict[int, list[int]] = {}
    for r in range(len(offs) - 1):
        s = bytes(reads[int(offs[r] - offs[0]):int(offs[r + 1] - offs[0])])
        if hpc_on and s:
            s = bytes([s[0]] + [s[i] for i in range(1, len(s)) if s[i] != s[i - 1]])
        for i in range(len(s) - k + 1):
            fw = s[i:i + k].upper()
            if fw.strip(b"ACGT"):
                continue
            rc = fw.translate(_COMP)[::-1]
            c = min(fw, rc)
            v = 0
            for ch in c:
                v = (v << 2) | _CODE[ch]
            if mix64(v) % sample_mod:
                continue
            out.setdefault(v, []).append(r << 32 | i << 1 | int(rc < fw))
    for v in out:
        out[v].sort()
    return out


class ParseError(ValueError):
    pass


def parse_reads(text, fmt: int = 0):
    """Reads of FASTQ / single-line FASTA text as a CSR (bases uint8, offsets
    uint64), restating the reference's record parsing:
    read_loader.hpp:44-121 (_load_fastq: 4-line records, '@' header,
    sequence, '+', quality; the read is the sequence line) and
    read_loader.hpp:123-196 (_load_fasta: '>' header and one sequence line;
    the last line may lack its newline); a '\\r' ending a sequence line is
    dropped (:103, :177).  Text that is not whole records of the format raises
    ParseError.  fmt: 0 from the first byte, 1 FASTA, 2 FASTQ."""
    a = np.frombuffer(bytes(text), np.uint8) if not isinstance(text, np.ndarray) else np.ascontiguousarray(text, np.uint8)
    n = len(a)
    if n == 0:
        return np.zeros(0, np.uint8), np.zeros(1, np.uint64)
    if fmt == 0:
        fmt = 1 if a[0] == ord(">") else 2 if a[0] == ord("@") else 0
        if fmt == 0:
            raise ParseError("text starts with neither '>' nor '@'")
    L = 4 if fmt == 2 else 2
    nl = np.flatnonzero(a == 10).astype(np.int64)
    starts = np.concatenate([[0], nl + 1])
    ends = np.concatenate([nl, [n]])
    if a[-1] == 10:                      # no line after the final newline
        starts, ends = starts[:-1], ends[:-1]
    if len(starts) % L:
        raise ParseError("text does not hold whole records")
    hs, he = starts[0::L], ends[0::L]
    ok = (hs < he) & (a[np.minimum(hs, n - 1)] == (ord("@") if fmt == 2 else ord(">")))
    if fmt == 2:
        ps, pe = starts[2::L], ends[2::L]
        ok &= (ps < pe) & (a[np.minimum(ps, n - 1)] == ord("+"))
    if not ok.all():
        raise ParseError("malformed record")
    ss, se = starts[1::L], ends[1::L].copy()
    cr = (se > ss) & (a[np.maximum(se - 1, 0)] == 13)
    se -= cr
    lens = se - ss
    offs = np.zeros(len(lens) + 1, np.uint64)
    np.cumsum(lens, out=offs[1:])
    idx = np.repeat(ss - offs[:-1].astype(np.int64), lens) + np.arange(int(offs[-1]), dtype=np.int64)
    return a[idx], offs
