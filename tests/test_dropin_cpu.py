"""Host-only parts of the C++ drop-in, on the CPU.

CalcSKMPartSize_instream (reference gkc_cuda.cu:350-365) over the oracle's
per-window minimizers and superkmer offset lists: the bases it assigns to
each partition must equal the oracle's per-partition k-mers plus k - 1 per
superkmer (a superkmer of n k-mers has n + k - 1 bases), a closed form of the
same scatter.  SKMStore::add_skm splits a superkmer at bytes other than
A/C/G/T: the packed runs it keeps are checked against the runs themselves.
"""
from __future__ import annotations

import os
import shutil
import subprocess

import numpy as np
import pytest

from gkc_amd import workloads as W
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gkc_amd", "lib")

PROG = r'''
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "gkc/gkc_cuda.hpp"
#include "gkc/superkmers.hpp"

template <typename T> std::vector<T> load(const char* path) {
    FILE* f = std::fopen(path, "rb");
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<T> v(n / sizeof(T));
    if (std::fread(v.data(), sizeof(T), v.size(), f) != v.size()) std::exit(3);
    std::fclose(f);
    return v;
}

int main(int argc, char** argv) {
    if (argv[1][0] == 'p') {            // p offs skm mm P k out
        auto offs = load<T_CSR_cap>(argv[2]);
        auto skm = load<T_read_len>(argv[3]);
        auto mm = load<T_minimizer>(argv[4]);
        const int P = std::atoi(argv[5]), k = std::atoi(argv[6]);
        std::vector<std::atomic<size_t>> sizes(P);
        for (auto& s : sizes) s = 0;
        CalcSKMPartSize_instream((T_read_cnt)(offs.size() - 1), skm.data(), offs.data(), mm.data(), P, k, sizes.data());
        FILE* f = std::fopen(argv[7], "wb");
        for (auto& s : sizes) { size_t x = s; std::fwrite(&x, 8, 1, f); }
        std::fclose(f);
    } else {                            // s text out: add_skm of argv[2], packed words and run lengths out
        SKMStore st;
        const char* t = argv[2];
        size_t n = 0;
        while (t[n]) n++;
        st.add_skm(t, (uint32_t)n);
        std::vector<uint32_t> words, lens;
        st.take(words, lens);
        FILE* f = std::fopen(argv[3], "wb");
        uint32_t nl = (uint32_t)lens.size();
        std::fwrite(&nl, 4, 1, f);
        std::fwrite(lens.data(), 4, lens.size(), f);
        std::fwrite(words.data(), 4, words.size(), f);
        std::fclose(f);
    }
    return 0;
}
'''


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    if shutil.which("g++") is None or not os.path.exists(os.path.join(LIB, "libgkc_dropin.so")):
        pytest.skip("needs g++ and the built drop-in library")
    d = tmp_path_factory.mktemp("dropin_cpu")
    (d / "t.cpp").write_text(PROG)
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(d / "t.cpp"), "-L", LIB,
                        "-lgkc_dropin", "-lgkc_hip", f"-Wl,-rpath,{LIB}", "-o", str(d / "t")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return d


@pytest.mark.parametrize("k,p,P", [(31, 11, 64), (21, 7, 31), (25, 15, 512)])
def test_calc_skm_part_size(prog, k, p, P):
    rs = W.generate(W.workload(1, scale=0.01))
    mm, _, _ = O.window_minimizers(rs.reads, rs.offs, k, p, 1)
    skm = O.superkmer_offs(rs.offs, k, mm)
    rs.offs.astype(np.uint64).tofile(prog / "offs.bin")
    skm.astype(np.int32).tofile(prog / "skm.bin")
    mm.astype(np.uint32).tofile(prog / "mm.bin")
    r = subprocess.run([str(prog / "t"), "p", str(prog / "offs.bin"), str(prog / "skm.bin"), str(prog / "mm.bin"),
                        str(P), str(k), str(prog / "out.bin")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    sizes = np.fromfile(prog / "out.bin", np.uint64)
    t = O.count(rs.reads, rs.offs, k, p, P)
    np.testing.assert_array_equal(sizes, t.part_kmers + (k - 1) * t.part_skms)


@pytest.mark.parametrize("text,runs", [(b"ACGTACGTAC", [b"ACGTACGTAC"]), (b"ACGNNTTGCAnGG", [b"ACG", b"TTGCA", b"GG"]),
                                       (b"NNN", []), (b"-ACGT.", [b"ACGT"]), (b"acgtRYacg", [b"acgt", b"acg"])])
def test_add_skm_splits_at_other_bytes(prog, text, runs):
    r = subprocess.run([str(prog / "t"), "s", text.decode(), str(prog / "s.bin")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(prog / "s.bin", np.uint32)
    nl = int(raw[0])
    lens, words = raw[1:1 + nl], raw[1 + nl:]
    assert lens.tolist() == [len(x) for x in runs]
    from gkc_amd.capi import pack_superkmers
    want, _ = pack_superkmers(runs)
    np.testing.assert_array_equal(words, want)
