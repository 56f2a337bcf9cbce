"""The committed golden fixtures (tests/golden/*.npz, made by
tests/golden/make_fixtures.py and checked there against brute force)
against the oracle (CPU) and the HIP path (GPU)."""
from __future__ import annotations

import glob
import os

import numpy as np
import pytest

from gkc_amd import capi
from oracle import oracle as O

FIX = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(p) for p in FIX])
def test_oracle_reproduces_fixture(path):
    d = _load(path)
    k, p, P = int(d["k"]), int(d["p"]), int(d["P"])
    t = O.count(d["reads"], d["offs"], k, p, P, hpc=bool(d["hpc"]), min_count=int(d["min_count"]))
    np.testing.assert_array_equal(t.part_off, d["part_off"])
    np.testing.assert_array_equal(t.key_lo, d["key_lo"])
    np.testing.assert_array_equal(t.cnt, d["cnt"])
    if k > 32:
        np.testing.assert_array_equal(t.key_hi, d["key_hi"])
    np.testing.assert_array_equal(t.part_distinct, d["part_distinct"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(p) for p in FIX])
def test_hip_reproduces_fixture(gpu, path):
    d = _load(path)
    k, p, P = int(d["k"]), int(d["p"]), int(d["P"])
    print("[gpu-case] fixture", os.path.basename(path), flush=True)
    with capi.Counter(k, p, P, hpc=bool(d["hpc"]), min_count=int(d["min_count"])) as c:
        c.partition(d["reads"], d["offs"])
        c.count()
        lo, hi, cnt, po, pd, pk = c.table()
        np.testing.assert_array_equal(po, d["part_off"])
        np.testing.assert_array_equal(lo, d["key_lo"])
        np.testing.assert_array_equal(cnt, d["cnt"])
        if k > 32:
            np.testing.assert_array_equal(hi, d["key_hi"])
        np.testing.assert_array_equal(pd, d["part_distinct"])
        np.testing.assert_array_equal(pk, d["part_kmers"])
        if p <= 16:
            arr = c.superkmer_arrays(d["reads"], d["offs"])
            offs = d["offs"]
            lens = arr["read_len"] if d["hpc"] else np.diff(offs).astype(np.int64)
            for r in range(len(offs) - 1):
                s, L = int(offs[r]), int(lens[r])
                W = max(0, L - k + 1)
                np.testing.assert_array_equal(arr["minimizers"][s:s + W], d["minimizers"][s:s + W].astype(np.uint32))
                ref = d["superkmer_offs"][s:s + L]
                if W:
                    end = 1 + int(np.argmax(ref[1:] == 0)) + 1 if L > 2 else L
                    np.testing.assert_array_equal(arr["superkmer_offs"][s:s + end], ref[:end])
