"""Autotuner of the counting kernels.

Picks the onesweep tile size (keys per thread), the waves per workgroup (4,
or 8 for 64-bit keys: the same tile LDS with twice the waves to hide the
loads behind, or tiles of 8192 / 9216 keys, two workgroups per CU, whose
longer digit runs write more whole lines) and the number of persistent
workgroups per CU of the sort passes by timing every candidate on the data
already in a counter's bins (the batched form, instrumented pass timing),
then the digit width and number of onesweep passes before the binned count
(8-, 7- or 6-bit digits, 2 or 3 passes: fewer or narrower passes leave
larger bins, which pays when the bins stay small enough for the wave-level
counts) and the count window, and caches the winners keyed on
(key width, read-length bucket, partition-size bucket, partitions) in
gkc_amd/tune_cache.json.  Every candidate produces the same table
(bit-exactness across variants is checked in tests/test_autotune.py and
tests/test_gpu_binned.py); only speed differs.
"""
from __future__ import annotations

import json
import math
import os

from . import capi

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_cache.json")
# (keys per thread, workgroups per CU (0: occupancy limit), waves per workgroup)
CANDIDATES_64 = [(i, b, 4) for i in (8, 12, 16, 20, 24) for b in (0, 2)] + [(i, 0, 8) for i in (8, 10, 12, 16, 18)]
CANDIDATES_128 = [(i, b, 4) for i in (4, 8, 12) for b in (0, 2)]


def bucket_key(k: int, mean_read_len: float, keys_per_partition: float, n_parts: int | None = None) -> str:
    kw = "u64" if k <= 32 else "u128"
    rl = int(round(math.log2(max(mean_read_len, 1.0))))
    ps = int(round(math.log2(max(keys_per_partition, 1.0))))
    key = f"{kw}/rl{rl}/ps{ps}"
    return key if n_parts is None else f"{key}/P{n_parts}"


def load_cache() -> dict:
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            return json.load(f)
    return {}


def save_cache(cache: dict) -> None:
    with open(CACHE, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)


def lookup(key: str) -> dict | None:
    return load_cache().get(key)


def nearest(key: str) -> dict | None:
    """Exact bucket, else the entry with the same key width and the closest
    partition-size bucket (a differing partition count or read-length
    bucket counting as one step further)."""
    cache = load_cache()
    if key in cache:
        return cache[key]
    f = key.split("/")
    best, dist = None, None
    for k2, v in cache.items():
        f2 = k2.split("/")
        if f2[0] != f[0]:
            continue
        d = 2 * abs(int(f2[2][2:]) - int(f[2][2:])) + (f2[1] != f[1]) + ((f2[3:] or None) != (f[3:] or None))
        if dist is None or d < dist:
            best, dist = v, d
    return best


def apply(counter: "capi.Counter", key: str) -> dict | None:
    t = nearest(key)
    if t:
        counter.set_tuning(t["sort_items"], t["sort_blocks_per_cu"], bin_passes=t.get("bin_passes", 0),
                           count_window=t.get("count_window", 0), sort_waves=t.get("sort_waves", 0),
                           digit_bits=t.get("digit_bits", 0))
    return t


def tune_sort(counter: "capi.Counter", reps: int = 2, log=None) -> dict:
    """Time every candidate on the counter's current bins; returns the best
    {sort_items, sort_blocks_per_cu, sort_waves, ms_pass} and the full sweep."""
    cands = CANDIDATES_64 if counter.k <= 32 else CANDIDATES_128
    sweep = []
    prev = os.environ.get("GKC_NO_GRAPH")
    os.environ["GKC_NO_GRAPH"] = "1"
    try:
        for items, bpc, waves in cands:
            counter.set_tuning(items, bpc, sort_waves=waves)
            best = None
            for _ in range(reps):
                counter.count(capi.BATCHED)
                ms = counter.stats()["ms_sort_pass_avg"]
                best = ms if best is None else min(best, ms)
            sweep.append({"sort_items": items, "sort_blocks_per_cu": bpc, "sort_waves": waves, "ms_pass": best})
            if log:
                log(f"tune items={items} blocks/CU={bpc or 'max'} waves={waves}: {best:.3f} ms/pass")
    finally:
        if prev is None:
            os.environ.pop("GKC_NO_GRAPH", None)
        else:
            os.environ["GKC_NO_GRAPH"] = prev
    win = min(sweep, key=lambda r: r["ms_pass"])
    counter.set_tuning(win["sort_items"], win["sort_blocks_per_cu"], sort_waves=win["sort_waves"])
    return {**win, "sweep": sweep}


COUNT_WINDOWS = (128, 256, 384)
DIGIT_BITS = (8, 7, 6)


def tune_bin_passes(counter: "capi.Counter", reps: int = 3, log=None) -> dict:
    """Time whole batched counts (graph replay) of the counter's current bins
    with digits of 8, 7 or 6 bits, 2 or 3 onesweep passes before the binned
    count (bins of 12 to 24 key bits: narrower digits rank with fewer ballots
    and smaller histograms but leave larger bins) and windows of 128, 256 and
    384 keys (larger windows spread each window's fixed work over more keys,
    but hold more distinct keys, each ranked against all of them); returns the best {digit_bits, bin_passes,
    count_window, ms_count} and the sweep.  The sort tuning already set is
    kept (6- and 7-bit passes use the nearest tile they are built for)."""
    t0 = counter.get_tuning()
    sweep = []
    for db in DIGIT_BITS:
        for npass in (3, 2):
            for cw in COUNT_WINDOWS:
                counter.set_tuning(t0["sort_items"], t0["sort_blocks_per_cu"], t0["single_list"], bin_passes=npass,
                                   count_window=cw, sort_waves=t0["sort_waves"], digit_bits=db)
                best = None
                for _ in range(reps):
                    counter.count(capi.BATCHED)
                    ms = counter.stats()["ms_count"]
                    best = ms if best is None else min(best, ms)
                sweep.append({"digit_bits": db, "bin_passes": npass, "count_window": cw, "ms_count": best})
                if log:
                    log(f"tune digit_bits={db} bin_passes={npass} count_window={cw}: {best:.3f} ms per count")
    win = min(sweep, key=lambda r: r["ms_count"])
    counter.set_tuning(t0["sort_items"], t0["sort_blocks_per_cu"], t0["single_list"], bin_passes=win["bin_passes"],
                       count_window=win["count_window"], sort_waves=t0["sort_waves"], digit_bits=win["digit_bits"])
    return {**win, "sweep": sweep}
