"""The C++ drop-in entry points (GenSuperkmerGPU, SKMStore::add_skms,
kmc_counting_GPU with the reference's signatures) driven by gkc_count, a
two-phase program shaped like the reference's own driver; its per-partition
tables must equal the oracle's."""
from __future__ import annotations

import os
import struct
import subprocess

import numpy as np
import pytest

from gkc_amd import _lib
from gkc_amd import workloads as W
from oracle import oracle as O

EXE = os.path.join(_lib.LIB_DIR, "gkc_count")


def _read_table(path):
    with open(path, "rb") as f:
        P, k = struct.unpack("<II", f.read(8))
        parts = []
        for _ in range(P):
            n, d = struct.unpack("<QQ", f.read(16))
            rec = np.frombuffer(f.read(20 * n), dtype=np.dtype([("hi", "<u8"), ("lo", "<u8"), ("c", "<u4")]))
            parts.append((rec["hi"].copy(), rec["lo"].copy(), rec["c"].copy(), d))
    return P, k, parts


def test_driver_binary_built():
    assert os.path.exists(EXE), "gkc_count not built"


@pytest.mark.gpu
@pytest.mark.parametrize("cid,scale,k,p,P,extra", [(1, 0.02, 31, 11, 64, []), (1, 0.02, 21, 7, 31, ["-kmin", "2"]),
                                                   (3, 0.0005, 25, 15, 64, ["-hpc"]),
                                                   (4, 0.0005, 45, 13, 16, ["-kmin", "2"])])
def test_dropin_driver_matches_oracle(gpu, tmp_path, cid, scale, k, p, P, extra):
    rs = W.generate(W.workload(cid, scale=scale))
    fa = str(tmp_path / "reads.fa")
    out = str(tmp_path / "table.bin")
    W.write_fasta(fa, rs)
    print("[gpu-case] gkc_count", cid, k, p, P, extra, flush=True)
    r = subprocess.run([EXE, "-read", fa, "-k", str(k), "-p", str(p), "-skm", str(P), "-t", "3", "-batch", "700",
                        "-out", out, *extra], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    hpc = "-hpc" in extra
    minc = int(extra[extra.index("-kmin") + 1]) if "-kmin" in extra else 1
    ot = O.count(rs.reads, rs.offs, k, p, P, hpc=hpc, min_count=minc)
    P2, k2, parts = _read_table(out)
    assert (P2, k2) == (P, k)
    for q in range(P):
        hi, lo, c, d = parts[q]
        olo, ohi, oc = ot.part(q)
        np.testing.assert_array_equal(lo, olo)
        np.testing.assert_array_equal(c, oc)
        if k > 32:
            np.testing.assert_array_equal(hi, ohi)
        assert d == int(ot.part_distinct[q])
    assert "Total number of distinct kmers: %d" % int(ot.part_distinct.sum()) in r.stderr


MMPOS_PROG = r'''
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "gkc/gkc_cuda.hpp"

int main(int argc, char** argv) {    // reads.bin offs.bin k p hpc out
    FILE* f = std::fopen(argv[1], "rb");
    std::vector<char> reads;
    for (int ch; (ch = std::fgetc(f)) != EOF;) reads.push_back((char)ch);
    std::fclose(f);
    f = std::fopen(argv[2], "rb");
    std::vector<unsigned long long> offs;
    for (unsigned long long x; std::fread(&x, 8, 1, f) == 1;) offs.push_back(x);
    std::fclose(f);
    std::vector<ReadPtr> rp;
    for (size_t i = 0; i + 1 < offs.size(); i++) rp.push_back({reads.data() + offs[i], (T_read_len)(offs[i + 1] - offs[i])});
    PinnedCSR pinned(rp);
    CUDAParams g{};
    g.device_id = 0; g.NUM_BLOCKS_PER_GRID = 4; g.NUM_THREADS_PER_BLOCK = 64;     // 256 reads per sub-batch
    FILE* o = std::fopen(argv[6], "wb");
    int fails = 0;
    GenSuperkmerGPU(pinned, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]) != 0, g, StoreMinimizerPos, 16,
                    [&](T_h_data hd) {
                        if (!hd.mm_pos || !hd.mm_strand) fails++;
                        const size_t B = hd.reads_offs[hd.reads_cnt] - hd.reads_offs[0];
                        unsigned long long nb = B;
                        std::fwrite(&nb, 8, 1, o);
                        std::fwrite(hd.minimizers, 4, B, o);
                        std::fwrite(hd.mm_pos, 1, B, o);
                        std::fwrite(hd.mm_strand, 1, B, o);
                    });
    std::fclose(o);
    return fails;
}
'''


@pytest.mark.gpu
@pytest.mark.parametrize("k,p,hpc", [(31, 11, False), (25, 15, True)])
def test_store_minimizer_pos(gpu, tmp_path, k, p, hpc):
    """GenSuperkmerGPU(..., StoreMinimizerPos, ...) hands the callback each
    window's minimizer offset and strand beside the minimizers: per
    sub-batch, equal to the oracle's per-window records of those reads."""
    import shutil
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    src = tmp_path / "t.cpp"
    src.write_text(MMPOS_PROG)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(root, "include"), str(src), "-L", _lib.LIB_DIR,
                        "-lgkc_dropin", "-lgkc_hip", f"-Wl,-rpath,{_lib.LIB_DIR}", "-o", str(tmp_path / "t")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rs = W.generate(W.workload(1, scale=0.002))
    rs.reads.tofile(tmp_path / "reads.bin")
    rs.offs.astype(np.uint64).tofile(tmp_path / "offs.bin")
    print("[gpu-case] StoreMinimizerPos", k, p, hpc, flush=True)
    r = subprocess.run([str(tmp_path / "t"), str(tmp_path / "reads.bin"), str(tmp_path / "offs.bin"), str(k), str(p),
                        str(int(hpc)), str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    raw = (tmp_path / "out.bin").read_bytes()
    at, r0 = 0, 0
    while at < len(raw):
        B = int(np.frombuffer(raw, np.uint64, 1, at)[0]); at += 8
        mm = np.frombuffer(raw, np.uint32, B, at); at += 4 * B
        pos = np.frombuffer(raw, np.uint8, B, at); at += B
        st = np.frombuffer(raw, np.int8, B, at); at += B
        n = 0
        while n < rs.n - r0 and int(rs.offs[r0 + n + 1] - rs.offs[r0]) <= B:
            n += 1
        o = (rs.offs[r0:r0 + n + 1] - rs.offs[r0]).astype(np.uint64)
        rd = rs.reads[int(rs.offs[r0]):int(rs.offs[r0 + n])]
        if hpc:
            hr, lens, _ = O.hpc(rd, o)
            omm, opos, ost = O.window_minimizers(hr, o, k, p, 1, lens=lens)
        else:
            lens = np.diff(o).astype(np.int32)
            omm, opos, ost = O.window_minimizers(rd, o, k, p, 1)
        for j in range(n):
            a, L = int(o[j]), int(lens[j])
            w = slice(a, a + max(L - k + 1, 0))
            np.testing.assert_array_equal(mm[w], omm[w].astype(np.uint32))
            np.testing.assert_array_equal(pos[w], opos[w])
            np.testing.assert_array_equal(st[w], ost[w])
        r0 += n
    assert r0 == rs.n
