"""ctypes binding of the C ABI in include/gkc_hip.h (libgkc_hip.so).

Host-side mirror of the reference's two counting calls:
  Counter.partition  ~ GenSuperkmerGPU + SKMStore::add_skms (gkc_cuda.hpp:58-61,
                       superkmers.hpp:167-186), bins kept on the device
  Counter.count      ~ kmc_counting_GPU over every partition (kmer_counting.hpp:9-12)
  Counter.superkmer_arrays  the per-sub-batch arrays GenSuperkmerGPU hands to its
                       callback (T_h_data, types.h:32-46)
Errors raise GkcError with the library's message (the reference prints and
exits, gkc_cuda.hpp:4-9).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib

DEVICE_INPUT = 1
BATCHED = 2
PRESIZED = 4
LEARNED = 16
EMIT_WINDOWS = 32
WITH_POS = 64
GKC_ECAP = -5

_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_u64p = C.POINTER(C.c_uint64)
_f32p = C.POINTER(C.c_float)


class Params(C.Structure):
    _fields_ = [("k", C.c_int32), ("p", C.c_int32), ("hpc", C.c_int32), ("n_parts", C.c_uint32),
                ("filter_mode", C.c_int32), ("part_mode", C.c_int32), ("min_count", C.c_uint32),
                ("max_count", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("sort_items", C.c_int32), ("sort_blocks_per_cu", C.c_int32), ("single_list", C.c_int32),
                ("count_mode", C.c_int32),
                ("bin_passes", C.c_int32), ("count_window", C.c_int32), ("sort_waves", C.c_int32),
                ("digit_bits", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("ms_partition", C.c_double), ("ms_count", C.c_double), ("ms_extract", C.c_double),
                ("ms_sort", C.c_double), ("ms_runlength", C.c_double), ("ms_sort_pass_avg", C.c_double),
                ("n_kmers", C.c_uint64), ("n_superkmers", C.c_uint64), ("n_words", C.c_uint64),
                ("n_distinct", C.c_uint64), ("n_entries", C.c_uint64), ("n_pass", C.c_uint32),
                ("fallbacks", C.c_uint32), ("host_syncs", C.c_uint32), ("partition_fallbacks", C.c_uint32),
                ("heavy_ranges", C.c_uint32), ("binned", C.c_uint32)]

    def asdict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


# exported symbols and their signatures (every symbol include/gkc_hip.h declares)
SIGNATURES = {
    "gkc_last_error": (C.c_char_p, []),
    "gkc_version": (C.c_int, []),
    "gkc_create": (C.c_void_p, [C.c_int, C.POINTER(Params)]),
    "gkc_destroy": (None, [C.c_void_p]),
    "gkc_stream": (C.c_void_p, [C.c_void_p]),
    "gkc_synchronize": (C.c_int, [C.c_void_p]),
    "gkc_superkmer_arrays": (C.c_int, [C.c_void_p, C.c_void_p, _u64p, C.c_int64, _u32p, _i32p, C.c_void_p, _i32p, _i32p]),
    "gkc_window_minimizers": (C.c_int, [C.c_void_p, C.c_void_p, _u64p, C.c_int64, _u32p, _u8p, C.POINTER(C.c_int8)]),
    "gkc_windows_device": (C.c_int, [C.c_void_p, C.POINTER(_u32p), C.POINTER(_u8p), C.POINTER(C.POINTER(C.c_int8)), _u64p]),
    "gkc_partition": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_uint32]),
    "gkc_reset": (C.c_int, [C.c_void_p]),
    "gkc_count": (C.c_int, [C.c_void_p, C.c_uint32]),
    "gkc_table_info": (C.c_int, [C.c_void_p, _u64p, _u64p, _u64p]),
    "gkc_table_copy": (C.c_int64, [C.c_void_p, C.c_int32, _u64p, _u64p, _u32p, C.c_uint64]),
    "gkc_table_write": (C.c_int64, [C.c_void_p, C.c_char_p]),
    "gkc_table_starts": (C.c_int, [C.c_void_p, _u64p]),
    "gkc_table_device": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(_u32p), _u64p]),
    "gkc_count_packed": (C.c_int, [C.c_void_p, _u32p, C.c_uint64, _u32p, C.c_uint64, _u64p, _u64p, _u32p, C.c_uint64,
                                   _u64p, _u64p]),
    "gkc_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "gkc_host_alloc": (C.c_void_p, [C.c_uint64]),
    "gkc_host_free": (None, [C.c_void_p]),
    "gkc_device_select": (C.c_int, [C.c_int]),
    "gkc_mlp_n_params": (C.c_uint64, []),
    "gkc_mlp_last_error": (C.c_char_p, []),
    "gkc_mlp_create": (C.c_void_p, [C.c_int, C.c_uint64]),
    "gkc_mlp_destroy": (None, [C.c_void_p]),
    "gkc_mlp_get_params": (C.c_int, [C.c_void_p, _f32p]),
    "gkc_mlp_set_params": (C.c_int, [C.c_void_p, _f32p]),
    "gkc_mlp_forward": (C.c_int, [C.c_void_p, _f32p, C.c_int64, _f32p]),
    "gkc_mlp_forward_bf16": (C.c_int, [C.c_void_p, _f32p, C.c_int64, _f32p]),
    "gkc_mlp_loss_grad": (C.c_int, [C.c_void_p, _f32p, _f32p, C.c_int64, _f32p, _f32p]),
    "gkc_mlp_train": (C.c_int, [C.c_void_p, _f32p, _f32p, C.c_int64, C.c_int, C.c_int, C.c_float, C.c_float,
                                C.c_uint64, _f32p]),
    "gkc_bf16_matmul32": (C.c_int, [C.c_int, _f32p, _f32p, C.c_int, C.c_int, _f32p]),
    "gkc_sketch": (C.c_int, [C.c_void_p, _f32p]),
    "gkc_set_table_caps": (C.c_int, [C.c_void_p, _u32p]),
    "gkc_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "gkc_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "gkc_part_kmers": (C.c_int, [C.c_void_p, _u64p]),
    "gkc_parse_reads": (C.c_int64, [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gkc_parse_reads_noalloc": (C.c_int64, [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "gkc_stream_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "gkc_scatter_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p]),
    "gkc_vendor_sort_u64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.c_void_p, _f32p]),
    "gkc_set_predictor": (C.c_int, [C.c_void_p, C.c_void_p, _f32p, _f32p, _f32p, _f32p, C.c_float]),
    "gkc_learned_caps": (C.c_int, [C.c_void_p, _u32p]),
    "gkc_superkmer_positions": (C.c_int64, [C.c_void_p, C.c_int32, _u32p, _u32p, _u32p, C.c_uint64]),
    "gkc_kmer_positions": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, _u64p, _u32p, _u64p, _u64p, C.c_uint64,
                                     C.c_uint64, _u64p, _u64p]),
}

_bound = None


def lib() -> C.CDLL:
    global _bound
    if _bound is None:
        L = _lib.hip()
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = L
    return _bound


class GkcError(RuntimeError):
    pass


def _check(rc: int) -> int:
    if rc < 0:
        raise GkcError(f"gkc error {rc}: {lib().gkc_last_error().decode()}")
    return rc


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


class Counter:
    """One counting context on one GPU."""

    def __init__(self, k: int, p: int, n_parts: int, hpc: bool = False, filter_mode: int = 1, part_mode: int = 0,
                 min_count: int = 1, max_count: int = 0, device: int = 0):
        self.k, self.p, self.n_parts = k, p, n_parts
        self.prm = Params(k, p, int(hpc), n_parts, filter_mode, part_mode, min_count, max_count)
        h = lib().gkc_create(device, C.byref(self.prm))
        if not h:
            raise GkcError(lib().gkc_last_error().decode())
        self.h = h
        self._destroy = lib().gkc_destroy      # held here: module globals may be gone when __del__ runs at exit

    def close(self) -> None:
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- phase 1
    def partition(self, reads, offs, n: int | None = None, flags: int = 0) -> None:
        """reads/offs: host numpy arrays, or device addresses (int) with DEVICE_INPUT."""
        if flags & DEVICE_INPUT:
            _check(lib().gkc_partition(self.h, C.c_void_p(int(reads)), C.c_void_p(int(offs)), int(n), flags))
        else:
            reads = np.ascontiguousarray(reads, np.uint8)
            offs = np.ascontiguousarray(offs, np.uint64)
            _check(lib().gkc_partition(self.h, reads.ctypes.data, offs.ctypes.data, len(offs) - 1, flags))

    def reset(self) -> None:
        _check(lib().gkc_reset(self.h))

    # -- phase 2
    def count(self, flags: int = 0) -> None:
        _check(lib().gkc_count(self.h, flags))

    def table_info(self):
        P = self.n_parts
        po = np.zeros(P + 1, np.uint64)
        pd = np.zeros(P, np.uint64)
        pk = np.zeros(P, np.uint64)
        _check(lib().gkc_table_info(self.h, _p(po, _u64p), _p(pd, _u64p), _p(pk, _u64p)))
        return po, pd, pk

    def table(self, part: int = -1):
        po, pd, pk = self.table_info()
        n = int(po[-1]) if part < 0 else int(po[part + 1] - po[part])
        lo = np.zeros(max(n, 1), np.uint64)
        hi = np.zeros(max(n, 1), np.uint64) if self.k > 32 else None
        cnt = np.zeros(max(n, 1), np.uint32)
        got = _check(lib().gkc_table_copy(self.h, part, _p(lo, _u64p), _p(hi, _u64p) if hi is not None else None,
                                          _p(cnt, _u32p), max(n, 1)))
        return lo[:got], (None if hi is None else hi[:got]), cnt[:got], po, pd, pk

    def write_table(self, path) -> int:
        """The table to a file (gkc_table_write); returns the bytes written.
        Read it back with gkc_amd.table_file.read_table."""
        return _check(lib().gkc_table_write(self.h, os.fsencode(path)))

    def stats(self) -> dict:
        s = Stats()
        _check(lib().gkc_get_stats(self.h, C.byref(s)))
        return s.asdict()

    def sketch(self) -> np.ndarray:
        out = np.zeros((self.n_parts, 256), np.float32)
        _check(lib().gkc_sketch(self.h, _p(out, _f32p)))
        return out

    def table_starts(self) -> np.ndarray:
        """Physical start of each partition's entries in the device table."""
        out = np.zeros(self.n_parts, np.uint64)
        _check(lib().gkc_table_starts(self.h, _p(out, _u64p)))
        return out

    def set_predictor(self, mlp_handle, x_mean, x_std, t_mean, t_std, margin: float, keep=None) -> None:
        """Attach a learned predictor (a gkc_mlp handle and its normalisation)
        for count(BATCHED | PRESIZED | LEARNED); mlp_handle None detaches."""
        if mlp_handle is None:
            _check(lib().gkc_set_predictor(self.h, None, None, None, None, None, 1.0))
            self._pred_keep = None
            return
        arrs = [np.ascontiguousarray(a, np.float32) for a in (x_mean, x_std, t_mean, t_std)]
        _check(lib().gkc_set_predictor(self.h, mlp_handle, *[_p(a, _f32p) for a in arrs], float(margin)))
        self._pred_keep = keep                      # the MLP must outlive its use

    def learned_caps(self) -> np.ndarray:
        out = np.zeros(self.n_parts, np.uint32)
        _check(lib().gkc_learned_caps(self.h, _p(out, _u32p)))
        return out

    def set_table_caps(self, caps) -> None:
        caps = np.ascontiguousarray(caps, np.uint32)
        _check(lib().gkc_set_table_caps(self.h, _p(caps, _u32p)))

    def set_tuning(self, sort_items: int = 0, sort_blocks_per_cu: int = 0, single_list: int = 0,
                   count_mode: int = 0, bin_passes: int = 0, count_window: int = 0, sort_waves: int = 0,
                   digit_bits: int = 0) -> None:
        """count_mode (batched form): 0 binned counting, 1 full LSD sort + run-length;
        bin_passes: onesweep passes before the binned count (0 automatic);
        count_window: keys per window of the binned count (0: 256);
        sort_waves: waves per onesweep workgroup, 4 or 8 (0: 4);
        digit_bits: digit width of the binned form's passes, 6..8 (0: 8)."""
        t = Tuning(sort_items, sort_blocks_per_cu, single_list, count_mode, bin_passes, count_window, sort_waves,
                   digit_bits)
        _check(lib().gkc_set_tuning(self.h, C.byref(t)))

    def get_tuning(self) -> dict:
        t = Tuning()
        _check(lib().gkc_get_tuning(self.h, C.byref(t)))
        return {"sort_items": t.sort_items, "sort_blocks_per_cu": t.sort_blocks_per_cu, "single_list": t.single_list,
                "count_mode": t.count_mode, "bin_passes": t.bin_passes, "count_window": t.count_window,
                "sort_waves": t.sort_waves, "digit_bits": t.digit_bits}

    def part_kmers(self) -> np.ndarray:
        out = np.zeros(self.n_parts, np.uint64)
        _check(lib().gkc_part_kmers(self.h, _p(out, _u64p)))
        return out

    def superkmer_positions(self, part: int):
        """Partition `part`'s superkmers after partition(..., WITH_POS):
        (read, offset of the first base in its read, k-mers), reads numbered
        from 0 at the last reset."""
        n = _check(lib().gkc_superkmer_positions(self.h, part, None, None, None, 0))
        rd, of, nk = (np.zeros(max(n, 1), np.uint32) for _ in range(3))
        _check(lib().gkc_superkmer_positions(self.h, part, _p(rd, _u32p), _p(of, _u32p), _p(nk, _u32p), n))
        return rd[:n], of[:n], nk[:n]

    def kmer_positions(self, part: int, sample_mod: int = 1):
        """After count() on WITH_POS bins (k <= 32): the partition's sampled
        k-mers (mix64(key) % sample_mod == 0) as (keys, counts, pos_off,
        positions); positions[pos_off[e]:pos_off[e + 1]] are k-mer e's
        read << 32 | offset << 1 | strand, ascending."""
        ne, npos = np.zeros(1, np.uint64), np.zeros(1, np.uint64)
        r = lib().gkc_kmer_positions(self.h, part, sample_mod, None, None, None, None, 0, 0, _p(ne, _u64p),
                                     _p(npos, _u64p))
        if r != GKC_ECAP:
            _check(r)
        e, q = int(ne[0]), int(npos[0])
        keys, cnts = np.zeros(max(e, 1), np.uint64), np.zeros(max(e, 1), np.uint32)
        off, pos = np.zeros(e + 1, np.uint64), np.zeros(max(q, 1), np.uint64)
        _check(lib().gkc_kmer_positions(self.h, part, sample_mod, _p(keys, _u64p), _p(cnts, _u32p), _p(off, _u64p),
                                        _p(pos, _u64p), e, q, _p(ne, _u64p), _p(npos, _u64p)))
        return keys[:e], cnts[:e], off, pos[:q]

    def synchronize(self) -> None:
        _check(lib().gkc_synchronize(self.h))

    def window_minimizers(self, reads, offs) -> dict:
        reads = np.ascontiguousarray(reads, np.uint8)
        offs = np.ascontiguousarray(offs, np.uint64)
        B = int(offs[-1] - offs[0])
        mm = np.zeros(max(B, 1), np.uint32)
        pos = np.zeros(max(B, 1), np.uint8)
        st = np.zeros(max(B, 1), np.int8)
        _check(lib().gkc_window_minimizers(self.h, reads.ctypes.data, _p(offs, _u64p), len(offs) - 1, _p(mm, _u32p),
                                           _p(pos, _u8p), st.ctypes.data_as(C.POINTER(C.c_int8))))
        return {"minimizers": mm[:B], "pos": pos[:B], "strand": st[:B]}

    def windows_device(self) -> tuple[int, int, int, int]:
        """Device addresses of the (minimizer u32, position u8, strand i8)
        records the last partition(..., EMIT_WINDOWS) wrote, and their count
        (batch-local base index; valid until the next such call)."""
        mm, pos, st, n = _u32p(), _u8p(), C.POINTER(C.c_int8)(), C.c_uint64(0)
        _check(lib().gkc_windows_device(self.h, C.byref(mm), C.byref(pos), C.byref(st), C.byref(n)))
        addr = lambda q: C.cast(q, C.c_void_p).value or 0
        return addr(mm), addr(pos), addr(st), n.value

    # -- drop-in #1 arrays
    def superkmer_arrays(self, reads, offs) -> dict:
        reads = np.ascontiguousarray(reads, np.uint8)
        offs = np.ascontiguousarray(offs, np.uint64)
        n = len(offs) - 1
        B = int(offs[-1] - offs[0])
        mm = np.zeros(max(B, 1), np.uint32)
        skm = np.zeros(max(B, 1), np.int32)
        ro = np.zeros(max(B, 1), np.uint8)
        rl = np.zeros(max(n, 1), np.int32)
        hp = np.zeros(max(B, 1), np.int32)
        _check(lib().gkc_superkmer_arrays(self.h, reads.ctypes.data, _p(offs, _u64p), n, _p(mm, _u32p),
                                          _p(skm, _i32p), ro.ctypes.data, _p(rl, _i32p), _p(hp, _i32p)))
        return {"minimizers": mm[:B], "superkmer_offs": skm[:B], "reads": ro[:B], "read_len": rl[:n],
                "hpc_orig_pos": hp[:B]}

    # -- drop-in #2 on packed host superkmers
    def count_packed(self, words: np.ndarray, lens: np.ndarray, cap: int):
        words = np.ascontiguousarray(words, np.uint32)
        lens = np.ascontiguousarray(lens, np.uint32)
        lo = np.zeros(max(cap, 1), np.uint64)
        hi = np.zeros(max(cap, 1), np.uint64) if self.k > 32 else None
        cnt = np.zeros(max(cap, 1), np.uint32)
        ne = C.c_uint64(0)
        nd = C.c_uint64(0)
        _check(lib().gkc_count_packed(self.h, _p(words, _u32p), len(words), _p(lens, _u32p), len(lens),
                                      _p(lo, _u64p), _p(hi, _u64p) if hi is not None else None, _p(cnt, _u32p),
                                      max(cap, 1), C.byref(ne), C.byref(nd)))
        n = ne.value
        return lo[:n], (None if hi is None else hi[:n]), cnt[:n], nd.value


def parse_reads(text_ptr: int, n_bytes: int, bases_ptr: int, offs_ptr: int, fmt: int = 0, stream: int = 0) -> int:
    """FASTQ / single-line FASTA text in device memory to the read CSR (device
    buffers: bases of n_bytes bytes, offs of n_bytes // 2 + 2 entries), the
    reference's ReadLoader record rules (read_loader.hpp:44-196) on the GPU;
    returns the number of reads.  fmt: 0 from the first byte, 1 FASTA, 2 FASTQ."""
    return _check(lib().gkc_parse_reads(C.c_void_p(text_ptr), n_bytes, fmt, C.c_void_p(bases_ptr),
                                        C.c_void_p(offs_ptr), C.c_void_p(stream)))


def stream_copy(dst_ptr: int, src_ptr: int, nbytes: int, stream: int = 0) -> None:
    """The library's streaming copy (device memory, 16-byte aligned sizes and
    addresses), asynchronous on `stream`: the practical HBM ceiling."""
    _check(lib().gkc_stream_copy(C.c_void_p(dst_ptr), C.c_void_p(src_ptr), nbytes, C.c_void_p(stream)))


def scatter_copy(dst_ptr: int, src_ptr: int, n_keys: int, seg_keys: int, runs: int, workgroups: int,
                 stream: int = 0) -> None:
    """Copy u64 keys with one onesweep pass's write pattern and none of its
    work (gkc_scatter_copy), asynchronous on `stream`: that pattern's ceiling."""
    _check(lib().gkc_scatter_copy(C.c_void_p(dst_ptr), C.c_void_p(src_ptr), n_keys, seg_keys, runs, workgroups,
                                  C.c_void_p(stream)))


def vendor_sort_u64(in_ptr: int, out_ptr: int, n: int, begin_bit: int, end_bit: int, stream: int = 0) -> float:
    """rocPRIM's radix sort of n device u64 keys on bits [begin_bit, end_bit)
    (a baseline for the onesweep passes); returns its milliseconds."""
    ms = np.zeros(1, np.float32)
    _check(lib().gkc_vendor_sort_u64(C.c_void_p(in_ptr), C.c_void_p(out_ptr), n, begin_bit, end_bit,
                                     C.c_void_p(stream), _p(ms, _f32p)))
    return float(ms[0])


# bytes of gkc_parse_reads_noalloc's fixed device scratch (GKC_PARSE_NOALLOC_SCRATCH_BYTES)
PARSE_NOALLOC_SCRATCH_BYTES = 32 << 20


def parse_reads_noalloc(text_ptr: int, n_bytes: int, bases_ptr: int, offs_ptr: int, fmt: int = 0,
                        stream: int = 0) -> int:
    """parse_reads with no device allocation: the library's fixed per-GPU
    scratch array instead of allocated scratch; raises GkcError (GKC_ECAP)
    for a text too large for it."""
    return _check(lib().gkc_parse_reads_noalloc(C.c_void_p(text_ptr), n_bytes, fmt, C.c_void_p(bases_ptr),
                                                C.c_void_p(offs_ptr), C.c_void_p(stream)))


def pack_superkmers(seqs: list[bytes]) -> tuple[np.ndarray, np.ndarray]:
    """2-bit MSB-first words (16 bases each), every superkmer word-aligned."""
    code = np.zeros(256, np.uint32)
    for ch, v in zip(b"ACGTacgt", [0, 1, 2, 3, 0, 1, 2, 3]):
        code[ch] = v
    words, lens = [], []
    for s in seqs:
        a = code[np.frombuffer(s, np.uint8)]
        pad = (-len(a)) % 16
        a = np.concatenate([a, np.zeros(pad, np.uint32)]).reshape(-1, 16)
        sh = (30 - 2 * np.arange(16)).astype(np.uint32)
        words.append((a << sh).sum(axis=1, dtype=np.uint64).astype(np.uint32))
        lens.append(len(s))
    w = np.concatenate(words) if words else np.zeros(0, np.uint32)
    return w, np.array(lens, np.uint32)
