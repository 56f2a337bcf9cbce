"""The k-mer table file written by ``gkc_table_write`` (declared in
include/gkc_hip.h), read back with numpy and no GPU.

The reference keeps no table: ``kmc_counting_GPU`` computes the (k-mer,
count) entries of a partition and drops them (kmer_counting.cu:187-190).
This file is the persistent form of the device table, little-endian:

* 64-byte header: ``char magic[8] = "GKCTAB1\\0"``; u32 ``version`` (1), ``k``,
  ``p``, ``hpc``, ``n_parts``, ``key_words`` (1 for k <= 32, 2 above),
  ``min_count``, ``max_count``; u64 ``n_entries``, ``n_kmers`` (k-mer
  occurrences counted), reserved 0;
* u64 ``part_off[n_parts + 1]``, ``part_distinct[n_parts]``,
  ``part_kmers[n_parts]``;
* u64 ``keys[n_entries][key_words]`` (low 64 bits first), partition by
  partition, ascending within each;
* u32 ``counts[n_entries]``.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

MAGIC = b"GKCTAB1\x00"
VERSION = 1
HEADER = 64


class TableFileError(ValueError):
    pass


@dataclass
class TableFile:
    k: int
    p: int
    hpc: int
    n_parts: int
    key_words: int
    min_count: int
    max_count: int
    n_kmers: int
    part_off: np.ndarray
    part_distinct: np.ndarray
    part_kmers: np.ndarray
    keys: np.ndarray          # (n_entries, key_words) uint64
    counts: np.ndarray        # (n_entries,) uint32

    def as_table(self):
        """(key_lo, key_hi or None, counts, part_off, part_distinct, part_kmers),
        the shape of ``Counter.table()``."""
        hi = self.keys[:, 1] if self.key_words == 2 else None
        return self.keys[:, 0], hi, self.counts, self.part_off, self.part_distinct, self.part_kmers


def read_table(path, mmap: bool = True) -> TableFile:
    """Parse and check a table file; the entry arrays are memory-mapped
    (read-only) unless ``mmap`` is False."""
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        hdr = f.read(HEADER)
    if len(hdr) < HEADER or hdr[:8] != MAGIC:
        raise TableFileError(f"{path}: not a GKC table file")
    h32 = np.frombuffer(hdr, "<u4", 8, 8)
    version, k, p, hpc, P, kw, mn, mx = (int(x) for x in h32)
    n, nk, _ = (int(x) for x in np.frombuffer(hdr, "<u8", 3, 40))
    if version != VERSION:
        raise TableFileError(f"{path}: version {version}, expected {VERSION}")
    if kw not in (1, 2) or kw != (1 if k <= 32 else 2):
        raise TableFileError(f"{path}: key_words {kw} does not fit k = {k}")
    o_keys = HEADER + 8 * (3 * P + 1)
    o_cnts = o_keys + 8 * kw * n
    if size != o_cnts + 4 * n:
        raise TableFileError(f"{path}: {size} bytes, the header implies {o_cnts + 4 * n}")
    info = np.fromfile(path, "<u8", 3 * P + 1, offset=HEADER)
    part_off, part_distinct, part_kmers = info[:P + 1], info[P + 1:2 * P + 1], info[2 * P + 1:]
    if int(part_off[0]) != 0 or int(part_off[-1]) != n or np.any(np.diff(part_off.astype(np.int64)) < 0):
        raise TableFileError(f"{path}: partition offsets do not cover the {n} entries")
    if mmap and n:
        keys = np.memmap(path, "<u8", "r", o_keys, (n, kw))
        counts = np.memmap(path, "<u4", "r", o_cnts, (n,))
    else:
        keys = np.fromfile(path, "<u8", n * kw, offset=o_keys).reshape(n, kw)
        counts = np.fromfile(path, "<u4", n, offset=o_cnts)
    return TableFile(k, p, hpc, P, kw, mn, mx, nk, part_off, part_distinct, part_kmers, keys, counts)


def write_table(path, k: int, p: int, hpc: int, min_count: int, max_count: int, part_off, part_distinct,
                part_kmers, key_lo, key_hi, counts) -> int:
    """The same layout from host arrays (a table copied with
    ``Counter.table()``, or the oracle's); returns the bytes written."""
    P = len(part_distinct)
    n = len(counts)
    kw = 1 if k <= 32 else 2
    h = bytearray(HEADER)
    h[:8] = MAGIC
    h[8:40] = np.array([VERSION, k, p, hpc, P, kw, min_count, max_count], "<u4").tobytes()
    h[40:56] = np.array([n, int(np.sum(np.asarray(part_kmers, np.uint64)))], "<u8").tobytes()
    keys = np.empty((n, kw), "<u8")
    keys[:, 0] = key_lo
    if kw == 2:
        keys[:, 1] = key_hi
    with open(path, "wb") as f:
        f.write(bytes(h))
        f.write(np.asarray(part_off, "<u8").tobytes())
        f.write(np.asarray(part_distinct, "<u8").tobytes())
        f.write(np.asarray(part_kmers, "<u8").tobytes())
        f.write(keys.tobytes())
        f.write(np.asarray(counts, "<u4").tobytes())
    return HEADER + 8 * (3 * P + 1) + 8 * kw * n + 4 * n
