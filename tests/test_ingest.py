"""On-device read ingest (gkc_parse_reads): FASTQ / single-line FASTA text
to the read CSR, against the oracle's restatement of the reference's
ReadLoader record rules (read_loader.hpp:44-196).  The oracle itself is
pinned by hand-written fixtures with their expected reads and by an
independent line-by-line parse; the GPU path is compared with it byte for
byte, then end to end (the table counted from parsed text equals the table
counted from the same reads)."""
from __future__ import annotations

import numpy as np
import pytest

from gkc_amd import workloads as W
from oracle import oracle as O

FIXTURES = [
    # (text, expected reads)
    (b"@r1\nACGT\n+\nIIII\n", [b"ACGT"]),
    (b"@r1\nACGT\n+r1\n@III\n@r2\nGG\n+\n@@\n", [b"ACGT", b"GG"]),          # quality lines starting with '@'
    (b"@a\r\nACGT\r\n+\r\nIIII\r\n", [b"ACGT"]),                             # CRLF
    (b"@a\nAC\n+\nII", [b"AC"]),                                             # no final newline
    (b"@a\n\n+\n\n@b\nT\n+\nI\n", [b"", b"T"]),                              # empty read
    (b">s1\nACGTACGT\n>s2\nNNA\n", [b"ACGTACGT", b"NNA"]),
    (b">s1\nACGT\n>s2\nGG", [b"ACGT", b"GG"]),                               # last line without newline
    (b">s1\r\nAC\r\n>s2\r\n\r\n", [b"AC", b""]),
    (b">x\nacgtN\n", [b"acgtN"]),                                            # bytes kept as they are
    (b">a\n\n>b\n\n", [b"", b""]),                                          # no bases at all
]
MALFORMED = [
    b"@r\nAC\n+\n",              # 3 lines
    b"@r\nAC\n-\nII\n",          # separator without '+'
    b"@r\nAC\n+\nII\nr2\nG\n+\nI\n",   # header without '@'
    b">a\nAC\nb\nTT\n",          # header without '>'
    b"ACGT\n",                   # neither format
    b">a\nAC\n\n",               # an extra line
]


def _naive(text: bytes) -> list[bytes]:
    """Line-by-line parse (a second, independent restatement of the rules)."""
    lines = text.split(b"\n")
    if text.endswith(b"\n"):
        lines = lines[:-1]
    L = 4 if text[:1] == b"@" else 2
    out = []
    for r in range(len(lines) // L):
        s = lines[L * r + 1]
        out.append(s[:-1] if s.endswith(b"\r") else s)
    return out


@pytest.mark.parametrize("text,want", FIXTURES)
def test_oracle_fixtures(text, want):
    r, o = O.parse_reads(text)
    got = [bytes(r[int(o[i]):int(o[i + 1])]) for i in range(len(o) - 1)]
    assert got == want
    assert _naive(text) == want


@pytest.mark.parametrize("text", MALFORMED)
def test_oracle_rejects(text):
    with pytest.raises(O.ParseError):
        O.parse_reads(text)


@pytest.mark.parametrize("fmt,crlf,final", [("fastq", False, True), ("fastq", True, False), ("fasta", False, False),
                                            ("fasta", True, True)])
def test_oracle_round_trip(fmt, crlf, final):
    rs = W.generate(W.workload(1, scale=0.01))
    t = W.reads_text(rs.reads, rs.offs, fmt, crlf, final)
    r, o = O.parse_reads(t)
    assert np.array_equal(o, rs.offs - rs.offs[0]) and np.array_equal(r, rs.reads[int(rs.offs[0]):int(rs.offs[-1])])
    small = W.reads_text(rs.reads, rs.offs[:41], fmt, crlf, final).tobytes()     # the first 40 reads
    sr, so = O.parse_reads(small)
    assert _naive(small) == [bytes(sr[int(so[i]):int(so[i + 1])]) for i in range(40)]


def _ragged(seed: int, n: int, fmt: str, crlf: bool, final: bool):
    """Reads of mixed lengths: empty (with a run of 150 in a row), 1-3 bases,
    short, ~150 and 5-20 kb; the last read non-empty (an empty last FASTQ
    quality line without newline is not a line)."""
    rng = np.random.default_rng(seed)
    kind = rng.integers(0, 5, n)
    lens = np.select([kind == 0, kind == 1, kind == 2, kind == 3],
                     [0, rng.integers(1, 4, n), rng.integers(100, 200, n), rng.integers(5000, 20000, n)],
                     rng.integers(4, 40, n))
    lens[n // 3:n // 3 + 150] = 0
    lens[-1] = max(int(lens[-1]), 1)
    offs = np.zeros(n + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    reads = np.frombuffer(b"ACGTN", np.uint8)[rng.integers(0, 5, int(offs[-1]))]
    return W.reads_text(reads, offs, fmt, crlf, final), reads, offs


@pytest.mark.parametrize("fmt,crlf,final", [("fastq", False, False), ("fasta", True, True)])
def test_oracle_ragged(fmt, crlf, final):
    t, reads, offs = _ragged(7, 400, fmt, crlf, final)
    r, o = O.parse_reads(t)
    assert np.array_equal(o, offs) and np.array_equal(r, reads)


# ---------------------------------------------------------------- GPU path

def _gpu_parse(text: np.ndarray, fmt: int = 0, shift: int = 0, out_shift: int = 0):
    import torch
    from gkc_amd import capi
    n = len(text)
    buf = torch.zeros(n + shift + 16, dtype=torch.uint8, device="cuda")
    if n:
        buf[shift:shift + n] = torch.from_numpy(np.ascontiguousarray(text)).cuda()
    guard = 64
    bases = torch.full((max(n, 1) + out_shift + guard,), 0xEE, dtype=torch.uint8, device="cuda")
    offs = torch.empty(n // 2 + 2, dtype=torch.int64, device="cuda")
    nr = capi.parse_reads(buf.data_ptr() + shift, n, bases.data_ptr() + out_shift, offs.data_ptr(), fmt)
    o = offs[:nr + 1].cpu().numpy().view(np.uint64)
    b = bases.cpu().numpy()
    T = int(o[-1])
    assert (b[:out_shift] == 0xEE).all() and (b[out_shift + T:] == 0xEE).all()     # nothing written outside
    return b[out_shift:out_shift + T], o


def _say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("i", range(len(FIXTURES)))
@pytest.mark.parametrize("shift", [0, 3])
def test_gpu_fixtures(gpu, i, shift):
    text, want = FIXTURES[i]
    _say("ingest fixture", i, shift)
    r, o = _gpu_parse(np.frombuffer(text, np.uint8), shift=shift)
    got = [bytes(r[int(o[j]):int(o[j + 1])]) for j in range(len(o) - 1)]
    assert got == want


@pytest.mark.gpu
@pytest.mark.parametrize("i", range(len(MALFORMED)))
def test_gpu_rejects(gpu, i):
    from gkc_amd import capi
    _say("ingest malformed", i)
    with pytest.raises(capi.GkcError):
        _gpu_parse(np.frombuffer(MALFORMED[i], np.uint8))


@pytest.mark.gpu
@pytest.mark.parametrize("fmt,crlf,final,shift", [("fastq", False, True, 0), ("fastq", True, False, 5),
                                                  ("fasta", False, False, 1), ("fasta", True, True, 0)])
def test_gpu_matches_oracle(gpu, fmt, crlf, final, shift):
    """Tens of thousands of records across many 4 KiB blocks (records
    straddling block boundaries), every line ending and alignment."""
    rs = W.generate(W.workload(1, scale=0.05))
    t = W.reads_text(rs.reads, rs.offs, fmt, crlf, final)
    _say("ingest synthetic", fmt, crlf, final, shift, len(t))
    r, o = _gpu_parse(t, shift=shift)
    orr, oo = O.parse_reads(t)
    assert np.array_equal(o, oo) and np.array_equal(r, orr)


@pytest.mark.gpu
@pytest.mark.parametrize("fmt,crlf,final,shift,out_shift", [("fastq", False, False, 0, 0), ("fastq", True, True, 7, 1),
                                                            ("fasta", False, True, 2, 2), ("fasta", True, False, 13, 3)])
def test_gpu_ragged(gpu, fmt, crlf, final, shift, out_shift):
    """Empty runs longer than the gather's 64-record window, reads of 1-3
    bases (dwords split between records) and reads longer than its 8 KiB
    output chunk, at every output alignment."""
    t, reads, offs = _ragged(11 + out_shift, 500, fmt, crlf, final)
    _say("ingest ragged", fmt, crlf, final, shift, out_shift, len(t))
    r, o = _gpu_parse(t, shift=shift, out_shift=out_shift)
    assert np.array_equal(o, offs) and np.array_equal(r, reads)


@pytest.mark.gpu
def test_gpu_empty_and_format_flag(gpu):
    from gkc_amd import capi
    r, o = _gpu_parse(np.zeros(0, np.uint8))
    assert len(r) == 0 and list(o) == [0]
    with pytest.raises(capi.GkcError):                       # FASTQ text read as FASTA
        _gpu_parse(np.frombuffer(FIXTURES[0][0], np.uint8), fmt=1)


@pytest.mark.gpu
def test_gpu_parsed_reads_count_like_the_reads(gpu):
    """Partition + count from the parsed text's CSR gives the oracle's table
    of the original reads."""
    import torch
    from gkc_amd import capi
    from tests.helpers import assert_tables_equal
    w = W.workload(1, scale=0.03)
    rs = W.generate(w)
    t = W.reads_text(rs.reads, rs.offs, "fastq")
    n = len(t)
    text = torch.from_numpy(t).cuda()
    bases = torch.empty(n, dtype=torch.uint8, device="cuda")
    offs = torch.empty(n // 2 + 2, dtype=torch.int64, device="cuda")
    nr = capi.parse_reads(text.data_ptr(), n, bases.data_ptr(), offs.data_ptr())
    assert nr == len(rs.offs) - 1
    _say("ingest then count", nr)
    with capi.Counter(w.k, w.p, 8) as c:
        c.partition(bases.data_ptr(), offs.data_ptr(), nr, flags=capi.DEVICE_INPUT | capi.BATCHED)
        c.count(capi.BATCHED)
        gt = c.table()
    assert_tables_equal(gt, O.count(rs.reads, rs.offs, w.k, w.p, 8), 8, w.k)
