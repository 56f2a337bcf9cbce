"""Loaders for the package's native libraries (built in-tree by gkc_amd.build).

Every loader fails loudly when its library is missing: there is no fallback
path for the HIP library.
"""
from __future__ import annotations

import ctypes
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

_synth = None
_hip = None

c_u64p = ctypes.POINTER(ctypes.c_uint64)


def _path(name: str) -> str:
    p = os.path.join(LIB_DIR, name)
    if not os.path.exists(p):
        raise RuntimeError(f"{name} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                           f"(or `python -m gkc_amd.build`) first")
    return p


def synth() -> ctypes.CDLL:
    global _synth
    if _synth is None:
        lib = ctypes.CDLL(_path("libgkc_synth.so"))
        lib.gkc_synth_genome.restype = ctypes.c_int
        lib.gkc_synth_genome.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_char_p]
        lib.gkc_synth_read_count.restype = ctypes.c_uint64
        lib.gkc_synth_read_count.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        lib.gkc_synth_bytes_bound.restype = ctypes.c_uint64
        lib.gkc_synth_bytes_bound.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64]
        lib.gkc_synth_reads.restype = ctypes.c_uint64
        lib.gkc_synth_reads.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint64,
                                        ctypes.c_int, ctypes.c_char_p, c_u64p]
        lib.gkc_write_fasta.restype = ctypes.c_int
        lib.gkc_write_fasta.argtypes = [ctypes.c_char_p, ctypes.c_char_p, c_u64p, ctypes.c_uint64]
        lib.gkc_read_fasta.restype = ctypes.c_int64
        lib.gkc_read_fasta.argtypes = [ctypes.c_char_p, ctypes.c_char_p, c_u64p, c_u64p]
        _synth = lib
    return _synth


def hip_path() -> str:
    return _path("libgkc_hip.so")


def hip() -> ctypes.CDLL:
    """The C-ABI HIP library (argtypes are declared by gkc_amd.capi)."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL(hip_path())
    return _hip
