"""gkc_parse_reads_noalloc: the on-device read ingest with no device
allocation (its scratch is a fixed per-GPU array in the library) against
the allocating gkc_parse_reads on the same text: byte-equal bases and equal
offsets, on consecutive calls that reuse the array (the second smaller than
the first, after a call that left the malformed flag set), and a text too
large for the array rejected with GKC_ECAP."""
from __future__ import annotations

import numpy as np
import pytest

from gkc_amd import workloads as W
from oracle import oracle as O


def _say(*a):
    print("[gpu-case]", *a, flush=True)


def _parse(fn, text: np.ndarray, shift: int = 0, fmt: int = 0):
    """Run one parse entry point on `text` placed `shift` bytes into a device
    buffer; guard bytes around the output must stay untouched."""
    import torch
    n = len(text)
    buf = torch.zeros(n + shift + 16, dtype=torch.uint8, device="cuda")
    if n:
        buf[shift:shift + n] = torch.from_numpy(np.ascontiguousarray(text)).cuda()
    guard = 64
    bases = torch.full((max(n, 1) + guard,), 0xEE, dtype=torch.uint8, device="cuda")
    offs = torch.full((n // 2 + 2,), -1, dtype=torch.int64, device="cuda")
    nr = fn(buf.data_ptr() + shift, n, bases.data_ptr(), offs.data_ptr(), fmt)
    o = offs[:nr + 1].cpu().numpy().view(np.uint64)
    b = bases.cpu().numpy()
    T = int(o[-1])
    assert (b[T:] == 0xEE).all(), "bytes written past the reads"
    return b[:T], o


@pytest.mark.gpu
def test_noalloc_equals_allocating_on_two_calls(gpu):
    from gkc_amd import capi
    rs = W.generate(W.workload(1, scale=0.05))
    big = W.reads_text(rs.reads, rs.offs, "fastq")
    n_small = (len(rs.offs) - 1) // 3
    small = W.reads_text(rs.reads, rs.offs[:n_small + 1], "fasta", True, False)
    for i, (t, shift) in enumerate([(big, 0), (small, 5)]):
        _say("noalloc call", i, len(t), shift)
        r_na, o_na = _parse(capi.parse_reads_noalloc, t, shift)
        r_a, o_a = _parse(capi.parse_reads, t, shift)
        assert np.array_equal(o_na, o_a) and np.array_equal(r_na, r_a)
        orr, oo = O.parse_reads(t)
        assert np.array_equal(o_na, oo) and np.array_equal(r_na, orr)
    assert len(o_na) - 1 == n_small


@pytest.mark.gpu
def test_noalloc_after_malformed_text(gpu):
    """A malformed text leaves the array's malformed flag set; the next call
    clears it before use."""
    from gkc_amd import capi
    _say("noalloc malformed then good")
    with pytest.raises(capi.GkcError):
        _parse(capi.parse_reads_noalloc, np.frombuffer(b"@r\nAC\n-\nII\n@s\nGG\n+\nII\n", np.uint8))
    t = np.frombuffer(b"@r\nACGT\n+\nIIII\n@s\nGG\n+\nII\n", np.uint8)
    r, o = _parse(capi.parse_reads_noalloc, t)
    assert bytes(r) == b"ACGTGG" and list(o) == [0, 4, 6]


@pytest.mark.gpu
def test_noalloc_rejects_text_too_large(gpu):
    """4.2 M one-base FASTA records need more than the 32 MiB array: rejected
    with GKC_ECAP, and the entry point still works afterwards."""
    from gkc_amd import capi
    n_rec = 4_200_000
    t = np.frombuffer(b">\nA\n" * n_rec, np.uint8)
    assert n_rec * 8 > capi.PARSE_NOALLOC_SCRATCH_BYTES
    _say("noalloc too large", len(t))
    with pytest.raises(capi.GkcError, match="error -5"):
        _parse(capi.parse_reads_noalloc, t)
    r, o = _parse(capi.parse_reads_noalloc, np.frombuffer(b">a\nACG\n", np.uint8))
    assert bytes(r) == b"ACG" and list(o) == [0, 3]
