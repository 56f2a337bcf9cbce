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
