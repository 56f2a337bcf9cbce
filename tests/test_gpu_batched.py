"""Batched form end to end: chunks of uneven size partitioned into presized
bins (graph replay per chunk, one sync per chunk), a forced bin overflow that
must fall back to count-then-allocate, device-resident input at unaligned
addresses, and the learned table presizing."""
from __future__ import annotations

import numpy as np
import pytest

from gkc_amd import capi
from gkc_amd import workloads as W
from oracle import oracle as O
from tests.helpers import assert_tables_equal

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)


def _chunks(n, sizes):
    b = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    b[-1] = n
    return list(zip(b[:-1], b[1:]))


def test_uneven_chunks_presized_bins(gpu):
    w = W.workload(2, scale=0.004)
    rs = W.generate(w)
    ot = O.count(rs.reads, rs.offs, w.k, w.p, w.n_parts)
    parts = _chunks(rs.n, [rs.n // 10, rs.n // 3, rs.n // 7, rs.n])
    _say("uneven chunks", [b - a for a, b in parts])
    with capi.Counter(w.k, w.p, w.n_parts) as c:
        for a, b in parts:
            c.partition(rs.reads[int(rs.offs[a]):int(rs.offs[b])], rs.offs[a:b + 1], flags=capi.BATCHED)
        c.count(capi.BATCHED)
        st = c.stats()
        assert_tables_equal(c.table(), ot, w.n_parts, w.k)
    assert st["partition_fallbacks"] == 0


def test_bin_overflow_falls_back(gpu):
    """Chunk 2 concentrates on partitions chunk 1 barely touched: the
    presized bins overflow on the device and the chunk is redone exactly."""
    w = W.workload(1, scale=0.02)
    rs = W.generate(w)
    rng = np.random.default_rng(5)
    unit = bytes(rng.choice(list(b"ACGT"), size=40).tolist())
    skew = [(unit * 5)[:150]] * 3000
    reads = np.concatenate([rs.reads, np.frombuffer(b"".join(skew), np.uint8)])
    offs = np.concatenate([rs.offs, rs.offs[-1] + np.cumsum([150] * len(skew)).astype(np.uint64)])
    ot = O.count(reads, offs, 31, 11, 64)
    _say("bin overflow")
    with capi.Counter(31, 11, 64) as c:
        c.partition(reads[:int(offs[rs.n])], offs[:rs.n + 1], flags=capi.BATCHED)
        c.partition(reads[int(offs[rs.n]):], offs[rs.n:], flags=capi.BATCHED)
        c.count(capi.BATCHED)
        st = c.stats()
        assert_tables_equal(c.table(), ot, 64, 31)
    assert st["partition_fallbacks"] >= 1


def test_device_input_unaligned(gpu):
    import torch
    w = W.workload(1, scale=0.02)
    rs = W.generate(w)
    ot = O.count(rs.reads, rs.offs, 31, 11, 64)
    d_reads = torch.from_numpy(rs.reads).cuda()
    d_offs = torch.from_numpy(rs.offs.view(np.int64)).cuda()
    torch.cuda.synchronize()
    parts = _chunks(rs.n, [137, 1001, 555, rs.n])
    _say("device input", parts)
    for flags in (0, capi.BATCHED):
        with capi.Counter(31, 11, 64) as c:
            for a, b in parts:
                c.partition(d_reads.data_ptr() + int(rs.offs[a]), d_offs.data_ptr() + 8 * a, b - a,
                            flags=capi.DEVICE_INPUT | flags)
            c.count(flags)
            assert_tables_equal(c.table(), ot, 64, 31)


def test_device_input_unaligned_hpc(gpu):
    """HPC on device reads that start off a 16-byte boundary (the compressed
    reads keep the input's layout, lead included)."""
    import torch
    w = W.workload(3, scale=0.001)
    rs = W.generate(w)
    ot = O.count(rs.reads, rs.offs, w.k, w.p, 32, hpc=True)
    d_reads = torch.from_numpy(rs.reads).cuda()
    d_offs = torch.from_numpy(rs.offs.view(np.int64)).cuda()
    torch.cuda.synchronize()
    parts = _chunks(rs.n, [3, 40, rs.n])
    _say("device input hpc", parts)
    for flags in (0, capi.BATCHED):
        with capi.Counter(w.k, w.p, 32, hpc=True) as c:
            for a, b in parts:
                c.partition(d_reads.data_ptr() + int(rs.offs[a]), d_offs.data_ptr() + 8 * a, b - a,
                            flags=capi.DEVICE_INPUT | flags)
            c.count(flags)
            assert_tables_equal(c.table(), ot, 32, w.k)
