"""GPU parity of the cases the other suites leave open: reads holding bytes
other than A/C/G/T (split there, KMC-style), per-window records written by
gkc_partition(GKC_EMIT_WINDOWS) on the long-read HPC configuration, 1024
partitions for 64- and 128-bit keys in every form, and chunks partitioned in
shuffled order (the batched bins are sized from the first chunk seen)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from gkc_amd import _lib, capi
from gkc_amd import workloads as W
from oracle import oracle as O
from tests.helpers import assert_tables_equal, device_to_host

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, flush=True)


def _with_other_bytes(rs, seed, every=97):
    """The reads with single N, runs of N, lowercase n and IUPAC / gap bytes
    written in (the same rule as the oracle tests use)."""
    rng = np.random.default_rng(seed)
    reads = rs.reads.copy()
    junk = np.frombuffer(b"NNNNnRYKM-.*X", np.uint8)
    at = rng.choice(len(reads), size=len(reads) // every, replace=False)
    reads[at] = junk[rng.integers(len(junk), size=len(at))]
    for a in rng.choice(len(reads) - 40, size=max(1, len(reads) // 3000), replace=False):
        reads[a:a + int(rng.integers(2, 40))] = ord("N")
    return reads


def _count(reads, offs, k, p, P, flags=0, hpc=False, min_count=1, batches=1, caps=None):
    with capi.Counter(k, p, P, hpc=hpc, min_count=min_count) as c:
        n = len(offs) - 1
        b = np.linspace(0, n, batches + 1).astype(int)
        for x, y in zip(b[:-1], b[1:]):
            if y > x:
                c.partition(reads[int(offs[x]):int(offs[y])], offs[x:y + 1] - offs[x], flags=flags & ~capi.PRESIZED)
        if caps is not None:
            c.set_table_caps(caps)
        c.count(flags)
        return c.table(), c.stats()


# ---------------------------------------------------------------- bytes other than A/C/G/T

@pytest.mark.parametrize("k,p,P,hpc,flags", [(31, 11, 64, False, 0), (31, 11, 64, False, capi.BATCHED),
                                             (25, 15, 512, True, 0), (25, 15, 512, True, capi.BATCHED),
                                             (63, 19, 64, False, capi.BATCHED)])
def test_non_acgt_bytes_split_reads(gpu, k, p, P, hpc, flags):
    rs = W.generate(W.workload(1, scale=0.02))
    reads = _with_other_bytes(rs, 11 + k)
    _say("non-ACGT", k, p, P, hpc, flags)
    gt, st = _count(reads, rs.offs, k, p, P, flags=flags, hpc=hpc, batches=2)
    ot = O.count(reads, rs.offs, k, p, P, hpc=hpc)
    assert_tables_equal(gt, ot, P, k)
    assert st["n_kmers"] == int(ot.part_kmers.sum()) < rs.kmers(k, hpc=hpc)


def test_non_acgt_window_records(gpu):
    rs = W.generate(W.workload(1, scale=0.01))
    reads = _with_other_bytes(rs, 5)
    k, p = 31, 11
    _say("non-ACGT windows")
    with capi.Counter(k, p, 64) as c:
        out = c.window_minimizers(reads, rs.offs)
    mm, pos, strand = O.window_minimizers(reads, rs.offs, k, p, 1)
    written = strand != -2                         # the oracle writes windows of k A/C/G/T bytes only
    assert 0 < written.sum() < rs.kmers(k)
    np.testing.assert_array_equal(out["minimizers"][written], mm[written].astype(np.uint32))
    np.testing.assert_array_equal(out["pos"][written], pos[written])
    np.testing.assert_array_equal(out["strand"][written], strand[written])


def test_superkmer_arrays_reject_non_acgt(gpu):
    rs = W.generate(W.workload(1, scale=0.002))
    reads = rs.reads.copy()
    reads[len(reads) // 2] = ord("N")
    with capi.Counter(31, 11, 64) as c:
        with pytest.raises(capi.GkcError, match="A/C/G/T"):
            c.superkmer_arrays(reads, rs.offs)
        arr = c.superkmer_arrays(rs.reads, rs.offs)      # the context stays usable
    assert arr["superkmer_offs"][0] == 0


def test_dropin_driver_non_acgt(gpu, tmp_path):
    """GenSuperkmerGPU hands the callback the reads cut at non-ACGT bytes;
    the drop-in driver's tables equal the oracle's."""
    rs = W.generate(W.workload(1, scale=0.01))
    reads = _with_other_bytes(rs, 3)
    fa = str(tmp_path / "reads.fa")
    out = str(tmp_path / "table.bin")
    W.write_fasta(fa, W.ReadSet(reads, rs.offs))
    _say("drop-in non-ACGT")
    exe = os.path.join(_lib.LIB_DIR, "gkc_count")
    r = subprocess.run([exe, "-read", fa, "-k", "31", "-p", "11", "-skm", "32", "-t", "2", "-batch", "500",
                        "-out", out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ot = O.count(reads, rs.offs, 31, 11, 32)
    assert "Total number of distinct kmers: %d" % int(ot.part_distinct.sum()) in r.stderr
    from tests.test_dropin import _read_table
    P2, k2, parts = _read_table(out)
    for q in range(32):
        _, lo, cnt, d = parts[q]
        olo, _, oc = ot.part(q)
        np.testing.assert_array_equal(lo, olo)
        np.testing.assert_array_equal(cnt, oc)


# ---------------------------------------------------------------- per-window records through gkc_partition

@pytest.mark.parametrize("batches", [1, 3])
def test_emit_windows_through_partition(gpu, batches):
    """gkc_partition(GKC_EMIT_WINDOWS) + gkc_windows_device on the long-read
    configuration's parameters (HPC, k 25, p 15, 512 partitions): the records
    of the last batch equal the oracle's; the table counted from the same bins
    equals the oracle's."""
    w = W.workload(3, scale=0.0004)
    rs = W.generate(w)
    assert (w.k, w.p, w.n_parts, w.hpc) == (25, 15, 512, True)
    b = np.linspace(0, rs.n, batches + 1).astype(int)
    _say("emit windows", rs.n, "batches", batches)
    with capi.Counter(w.k, w.p, w.n_parts, hpc=True) as c:
        for x, y in zip(b[:-1], b[1:]):
            sub_r = rs.reads[int(rs.offs[x]):int(rs.offs[y])]
            sub_o = rs.offs[x:y + 1] - rs.offs[x]
            c.partition(sub_r, sub_o, flags=capi.EMIT_WINDOWS)
            pm, pp, ps, n = c.windows_device()
            assert n == int(sub_o[-1])
            mm_d = device_to_host(pm, 4 * n, np.uint32)
            pos_d = device_to_host(pp, n, np.uint8)
            st_d = device_to_host(ps, n, np.int8)
            src, lens, _ = O.hpc(sub_r, sub_o)
            mm, pos, st = O.window_minimizers(src, sub_o, w.k, w.p, 1, lens=lens)
            written = st != -2
            assert written.sum() > 0
            np.testing.assert_array_equal(mm_d[written], mm[written].astype(np.uint32))
            np.testing.assert_array_equal(pos_d[written], pos[written])
            np.testing.assert_array_equal(st_d[written], st[written])
        c.count(0)
        gt = c.table()
    assert_tables_equal(gt, O.count(rs.reads, rs.offs, w.k, w.p, w.n_parts, hpc=True), w.n_parts, w.k)


# ---------------------------------------------------------------- 1024 partitions

@pytest.mark.parametrize("cid,scale,minc", [(5, 0.0008, 1), (4, 0.0006, 2)])
@pytest.mark.parametrize("form", ["plain", "batched", "presized"])
def test_1024_partitions(gpu, cid, scale, minc, form):
    """The skewed-GC k31 and the k63 128-bit-key configurations at their own
    1024 partitions (threshold 2 for k63), plain, batched and batched with
    presized tables (caps of exact + 10 %, a third of them halved: their sum
    is short, which falls back)."""
    w = W.workload(cid, scale=scale)
    rs = W.generate(w)
    assert w.n_parts == 1024 and w.min_count == minc
    ot = O.count(rs.reads, rs.offs, w.k, w.p, 1024, min_count=minc)
    _say("P1024", w.name, form, rs.n)
    flags = {"plain": 0, "batched": capi.BATCHED, "presized": capi.BATCHED | capi.PRESIZED}[form]
    caps = None
    if form == "presized":
        exact = np.diff(ot.part_off).astype(np.uint32)
        caps = (exact * 11 // 10 + 1).astype(np.uint32)
        caps[::3] = np.maximum(caps[::3] // 2, 1)
    gt, st = _count(rs.reads, rs.offs, w.k, w.p, 1024, flags=flags, min_count=minc, batches=2, caps=caps)
    assert_tables_equal(gt, ot, 1024, w.k)
    if form == "presized":
        # dense binned layout: a fallback only when the caps' sum is short
        assert st["fallbacks"] == (1 if int(np.diff(ot.part_off).sum()) > int(caps.sum()) else 0)


# ---------------------------------------------------------------- chunk order

@pytest.mark.parametrize("seed", [1, 2])
def test_shuffled_chunk_order(gpu, seed):
    """Uneven chunks of the short-read workload partitioned in a shuffled
    order (the smallest first for one seed: the bins are sized from the first
    chunk's sample and must regrow) give the oracle's table."""
    w = W.workload(2, scale=0.003)
    rs = W.generate(w)
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.choice(np.arange(1, rs.n), size=5, replace=False))
    bounds = list(zip(np.r_[0, cuts], np.r_[cuts, rs.n]))
    order = rng.permutation(len(bounds))
    if seed == 1:
        order = np.argsort([b - a for a, b in bounds])           # smallest chunk first
    ot = O.count(rs.reads, rs.offs, w.k, w.p, w.n_parts)
    _say("shuffled chunks", [int(bounds[i][1] - bounds[i][0]) for i in order])
    with capi.Counter(w.k, w.p, w.n_parts) as c:
        for i in order:
            a, b = bounds[i]
            c.partition(rs.reads[int(rs.offs[a]):int(rs.offs[b])], rs.offs[a:b + 1], flags=capi.BATCHED)
        c.count(capi.BATCHED)
        assert_tables_equal(c.table(), ot, w.n_parts, w.k)


# ---------------------------------------------------------------- 4-mer sketch

def _sketch_cpu(src, offs, lens, k, p, P, tile=2048):
    """The sketch rule recomputed from the oracle's window minimizers: the
    superkmers of each read (a new one where the minimizer changes, and at
    every batch position that starts a 2048-window tile), one in eight kept
    by the multiplicative hash of its first min(k, 16) bases, its leading
    4-mer counted in its partition (minimizer % P)."""
    mm, _, _ = O.window_minimizers(src, offs, k, p, 1, lens=lens)
    code = np.zeros(256, np.int64)
    for ch, v in zip(b"ACGTacgt", [0, 1, 2, 3, 0, 1, 2, 3]):
        code[ch] = v
    m = min(k, 16)
    out = np.zeros((P, 256), np.int64)
    for r in range(len(offs) - 1):
        s = int(offs[r])
        L = int(lens[r]) if lens is not None else int(offs[r + 1] - offs[r])
        W = L - k + 1
        if W <= 0:
            continue
        g = np.arange(s, s + W)
        start = np.ones(W, bool)
        start[1:] = (mm[s + 1:s + W] != mm[s:s + W - 1]) | (g[1:] % tile == 0)
        for i in np.nonzero(start)[0]:
            lead = 0
            for t in range(m):
                lead = (lead << 2) | int(code[src[s + i + t]])
            if ((lead * 0x9E3779B1) & 0xFFFFFFFF) >> 29 == 0:
                out[int(mm[s + i]) % P, lead >> (2 * (m - 4))] += 1
    return out


@pytest.mark.parametrize("cid,scale,k,p,P,hpc,flags", [(1, 0.01, 31, 11, 64, False, 0),
                                                     (1, 0.01, 31, 11, 64, False, capi.BATCHED),
                                                     (1, 0.01, 12, 5, 16, False, capi.BATCHED),
                                                     (3, 0.0004, 25, 15, 512, True, capi.BATCHED)])
def test_sketch_matches_cpu_recomputation(gpu, cid, scale, k, p, P, hpc, flags):
    w = W.workload(cid, scale=scale)
    rs = W.generate(w)
    _say("sketch", cid, k, p, P, hpc, flags)
    with capi.Counter(k, p, P, hpc=hpc) as c:
        c.partition(rs.reads, rs.offs, flags=flags)
        sk = c.sketch()
    got = np.rint(np.expm1(sk.astype(np.float64))).astype(np.int64)
    src, lens = rs.reads, None
    if hpc:
        src, lens, _ = O.hpc(rs.reads, rs.offs)
    want = _sketch_cpu(src, rs.offs, lens, k, p, P)
    assert want.sum() > 100
    np.testing.assert_array_equal(got, want)


# ---------------------------------------------------------------- measurement helpers

def test_stream_copy_and_vendor_sort(gpu):
    """The two baselines the bench reports beside the onesweep pass: the
    library's streaming copy copies, and the vendor sort sorts (on the bit
    range asked for, and entirely)."""
    import torch
    _say("stream copy + vendor sort")
    n = 1 << 20
    a = torch.randint(0, 2 ** 62, (n,), dtype=torch.int64, device="cuda")
    b = torch.zeros_like(a)
    st = torch.cuda.current_stream().cuda_stream
    capi.stream_copy(b.data_ptr(), a.data_ptr(), n * 8, st)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    with pytest.raises(capi.GkcError):
        capi.stream_copy(b.data_ptr() + 4, a.data_ptr(), 64, st)     # 16-byte alignment required
    host = a.cpu().numpy().astype(np.uint64)
    ms = capi.vendor_sort_u64(a.data_ptr(), b.data_ptr(), n, 0, 64, st)
    assert ms > 0
    np.testing.assert_array_equal(b.cpu().numpy().astype(np.uint64), np.sort(host))
    capi.vendor_sort_u64(a.data_ptr(), b.data_ptr(), n, 46, 62, st)
    got = b.cpu().numpy().astype(np.uint64)
    d = (got >> np.uint64(46)) & np.uint64(0xFFFF)
    assert (np.diff(d.astype(np.int64)) >= 0).all()
    np.testing.assert_array_equal(np.sort(got), np.sort(host))
