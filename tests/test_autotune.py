"""Autotuner: bucket keys and cache lookup on the CPU; on the GPU every tuned
variant of the onesweep pass gives the same (bit-exact) table."""
from __future__ import annotations

import pytest

from gkc_amd import autotune, capi
from gkc_amd import workloads as W
from oracle import oracle as O
from tests.helpers import assert_tables_equal


def test_bucket_keys():
    assert autotune.bucket_key(31, 150, 4.7e6) == "u64/rl7/ps22"
    assert autotune.bucket_key(31, 150, 4.7e6, 256) == "u64/rl7/ps22/P256"
    assert autotune.bucket_key(63, 250, 3e6) == "u128/rl8/ps22"
    assert autotune.bucket_key(25, 10000, 2.1e6).startswith("u64/rl13/")


def test_nearest_prefers_exact_then_partition_size(tmp_path, monkeypatch):
    import json
    cache = {"u64/rl7/ps22/P256": {"sort_items": 20, "sort_blocks_per_cu": 0, "bin_passes": 2},
             "u64/rl7/ps22/P1024": {"sort_items": 20, "sort_blocks_per_cu": 0, "bin_passes": 3},
             "u64/rl7/ps18/P64": {"sort_items": 12, "sort_blocks_per_cu": 0, "bin_passes": 2},
             "u128/rl8/ps22/P1024": {"sort_items": 8, "sort_blocks_per_cu": 0}}
    f = tmp_path / "cache.json"
    f.write_text(json.dumps(cache))
    monkeypatch.setattr(autotune, "CACHE", str(f))
    assert autotune.nearest("u64/rl7/ps22/P1024")["bin_passes"] == 3
    assert autotune.nearest("u64/rl7/ps22/P512")["sort_items"] == 20      # same size bucket, other P
    assert autotune.nearest("u64/rl7/ps18/P128")["sort_items"] == 12      # closest size bucket
    assert autotune.nearest("u128/rl7/ps20/P16")["sort_items"] == 8       # key width first
    assert autotune.nearest("u64/rl7/ps22") is not None                   # keys without P still resolve


def test_cache_entries_are_valid_variants():
    for key, v in autotune.load_cache().items():
        cands = autotune.CANDIDATES_64 if key.startswith("u64") else autotune.CANDIDATES_128
        assert (v["sort_items"], v["sort_blocks_per_cu"], v.get("sort_waves", 4) or 4) in cands, key
        assert v.get("bin_passes", 0) in (0, 2, 3), key
        assert v.get("count_window", 0) in (0,) + autotune.COUNT_WINDOWS, key
        assert v.get("digit_bits", 0) in (0,) + autotune.DIGIT_BITS, key


@pytest.mark.gpu
@pytest.mark.parametrize("cid,scale,k,p,P", [(1, 0.03, 31, 11, 64), (4, 0.001, 63, 19, 32)])
def test_every_variant_bit_exact(gpu, cid, scale, k, p, P):
    rs = W.generate(W.workload(cid, scale=scale))
    ot = O.count(rs.reads, rs.offs, k, p, P)
    cands = autotune.CANDIDATES_64 if k <= 32 else autotune.CANDIDATES_128
    with capi.Counter(k, p, P) as c:
        c.partition(rs.reads, rs.offs)
        for items, bpc, waves in cands:
            print("[gpu-case] variant", items, bpc, waves, flush=True)
            c.set_tuning(items, bpc, sort_waves=waves)
            assert c.get_tuning()["sort_waves"] == waves
            for flags in (0, capi.BATCHED):
                c.count(flags)
                assert_tables_equal(c.table(), ot, P, k)
