"""Benchmark of the MI355X k-mer counting path.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--config 2] [--scale 1.0]

Workload (default): the 50 Mbp, 30x, 150 bp, k=31 p=11, 256-partition short-read
set, reads resident in HBM before timing, cut into log-uniform 64 MB-1 GB
chunks.  One step = the whole hot path over the whole read set: every chunk
through encode -> minimizer -> superkmer -> partition scatter, then every
partition through extraction -> onesweep radix sort -> run-length count into
the (k-mer, count) table.

Two forms are timed: the plain single-launch counter (count-then-allocate,
partition by partition) and the batched + learned form (all partitions per
launch, hipGraph replay, tables presized by the learned predictor with the
on-device overflow fallback).  `value` is the batched + learned k-mers/s.
Rank 0 prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "exact k-mers/sec, batched+learned vs plain, 1 MI355X, bit-exact; % HBM peak"


def _log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def measured_traffic(workload: str, items: int):
    """HBM bytes per onesweep launch from the committed rocprofv3 counter runs
    (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel's readings) of the same
    workload and kernel variant, or (None, None)."""
    import glob
    import re
    here = os.path.dirname(os.path.abspath(__file__))
    for f in sorted(glob.glob(os.path.join(here, "profiles", "*traffic*.md"))):
        m = re.search(r"<!-- traffic-json: (\{.*\}) -->", open(f).read())
        if m:
            d = json.loads(m.group(1))
            if d.get("workload") == workload and d.get("items") == items:
                return int(d["traffic_bytes"]), os.path.relpath(f, here)
    return None, None


def learned_report(predictor, ctx, part_distinct, part_kmers) -> dict:
    """Predicted vs exact per-partition counts of the last batched run (the
    bins still hold its superkmers and its sketch)."""
    import numpy as np

    from gkc_amd import learned
    tt = part_kmers.astype(np.float64)
    tot_p, dist_p = predictor.predict(ctx.sketch(), tt)
    td = part_distinct.astype(np.float64)
    caps = predictor.caps(ctx).astype(np.float64)
    return {"mape_total_pct": learned.mape(tot_p, tt), "mape_distinct_pct": learned.mape(dist_p, td),
            "under_predicted_partitions": int(np.sum(caps < td)), "partitions": int(len(td)),
            "table_capacity_over_exact": float(caps.sum() / max(td.sum(), 1)),
            "table_capacity_over_total_bound": float(caps.sum() / max(tt.sum(), 1)),
            "margin": predictor.margin, "trained_on": predictor.info.get("pairs")}


def ingest_report(rs, max_reads, with_cpu, np, torch, capi, W) -> dict:
    """On-device FASTQ ingest (gkc_parse_reads) of the first reads of the
    workload: the text resident in HBM, the call timed by wall clock (it
    returns when its work is done; scratch allocation included), its output
    checked against the reads.  Roofline: text read once, bases and offsets
    written once, over the call's time."""
    nr = min(max_reads, rs.n)
    t = W.reads_text(rs.reads, rs.offs[:nr + 1], "fastq")
    n = len(t)
    nb = int(rs.offs[nr] - rs.offs[0])
    text = torch.from_numpy(t).cuda()
    bases = torch.empty(n, dtype=torch.uint8, device="cuda")
    offs = torch.empty(n // 2 + 2, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    args = (text.data_ptr(), n, bases.data_ptr(), offs.data_ptr())
    got = capi.parse_reads(*args)
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        got = capi.parse_reads(*args)
        times.append(time.perf_counter() - t0)
    ms = sorted(times)[len(times) // 2] * 1e3
    want = torch.from_numpy(np.ascontiguousarray(rs.reads[int(rs.offs[0]):int(rs.offs[nr])])).cuda()
    ok = got == nr and int(offs[nr].item()) == nb and bool(torch.equal(bases[:nb], want))
    alg = n + nb + 8 * (nr + 1)
    rep = {"format": "fastq", "text_bytes": n, "reads": nr, "bases": nb, "ms": ms, "text_GB_per_s": n / ms / 1e6,
           "reads_per_s": nr / ms * 1e3, "output_equal": ok,
           "roofline": {"bound": "hbm", "achieved": alg / ms / 1e6, "peak": 8000.0, "unit": "GB/s",
                        "frac": alg / ms / 1e6 / 8000.0, "algorithmic_bytes": alg}}
    if with_cpu:
        from oracle import oracle as O
        ns = min(200_000, nr)
        ts = W.reads_text(rs.reads, rs.offs[:ns + 1], "fastq")
        t0 = time.perf_counter()
        O.parse_reads(ts)
        dt = time.perf_counter() - t0
        rep["cpu_baseline"] = {"value": len(ts) / dt / 1e9, "unit": "text GB/s", "cores": 1, "kind": "port",
                               "sample": f"FASTQ text of the first {ns} reads ({len(ts)} bytes), oracle parse"}
    del text, bases, offs, want
    return rep


def table_file_report(ctx, directory, np) -> dict:
    """gkc_table_write of the step's table (the last count) to a file in
    `directory`, deleted afterwards: wall clock of the call (device-to-host
    copies through pinned staging overlapping the file writes), bytes, and the
    file read back: header and partition offsets equal to the device table's,
    the first partition's entries equal to gkc_table_copy's."""
    import shutil
    import tempfile
    from gkc_amd import table_file as TF
    po, pd, pk = ctx.table_info()
    kw = 1 if ctx.k <= 32 else 2
    need = 64 + 8 * (3 * len(pd) + 1) + (8 * kw + 4) * int(po[-1])
    free = shutil.disk_usage(directory).free
    if free < 2 * need:
        return {"skipped": f"{directory}: {free} bytes free, the file needs {need}"}
    t0 = time.perf_counter()
    ctx.write_table(os.devnull)                 # the same call with no file system behind it
    dt_null = time.perf_counter() - t0
    fd, path = tempfile.mkstemp(suffix=".gkct", dir=directory)
    os.close(fd)
    try:
        t0 = time.perf_counter()
        n = ctx.write_table(path)
        dt = time.perf_counter() - t0
        t = TF.read_table(path)
        lo, hi, cnt, _, _, _ = ctx.table(0)
        m = len(cnt)
        ok = (n == need and bool(np.array_equal(t.part_off, po)) and bool(np.array_equal(t.part_kmers, pk))
              and bool(np.array_equal(t.keys[:m, 0], lo)) and bool(np.array_equal(t.counts[:m], cnt))
              and (hi is None or bool(np.array_equal(t.keys[:m, 1], hi))))
        del t
        return {"dir": directory, "bytes": n, "entries": int(po[-1]), "ms": dt * 1e3, "GB_per_s": n / dt / 1e9,
                "to_devnull_ms": dt_null * 1e3, "to_devnull_GB_per_s": n / dt_null / 1e9, "read_back_equal": ok}
    finally:
        os.unlink(path)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-reads", type=int, default=1_500_000, help="reads in the CPU baseline sample")
    ap.add_argument("--forms", default="plain,batched")
    ap.add_argument("--tune", action="store_true", help="run the autotuner on this workload and cache the result")
    ap.add_argument("--no-ingest", action="store_true", help="skip the on-device FASTQ ingest measurement")
    ap.add_argument("--ingest-reads", type=int, default=2_000_000, help="reads in the ingest sample")
    ap.add_argument("--no-table-file", action="store_true", help="skip the table file write measurement")
    ap.add_argument("--table-dir", default=None, help="directory of the table file measurement (default: temp dir)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="end with a 1 GiB device copy (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding (partition ownership + one "
                                   "all-to-all of superkmer bins) is designed but not run",
                          "n_gpus": args.gpus}))
        return 2

    import numpy as np
    import torch

    from gkc_amd import autotune, capi, learned
    from gkc_amd import workloads as W

    w = W.workload(args.config, scale=args.scale)
    t0 = time.time()
    rs = W.generate(w)
    N = rs.kmers(w.k, hpc=w.hpc)
    _log(f"{w.name}: {rs.n} reads, {rs.bases} bases, {N} k-mers (generated in {time.time() - t0:.1f}s)")
    if w.chunked:
        chunks = W.chunk_bounds(rs, seed=W.genome_seed(w.config_id) + 777)
    else:   # one batch holds < 2^31 bases: equal chunks of at most 1 GiB
        chunks = W.chunk_bounds(rs, seed=0, lo_bytes=1 << 30, hi_bytes=1 << 30, fasta_overhead=False)
    _log(f"{len(chunks)} chunks: " + ", ".join(f"{int(rs.offs[b] - rs.offs[a]) >> 20}MB" for a, b in chunks))

    dev = torch.device("cuda:0")
    d_reads = torch.from_numpy(rs.reads).to(dev)
    d_offs = torch.from_numpy(rs.offs.view(np.int64)).to(dev)
    torch.cuda.synchronize()
    rp, op = d_reads.data_ptr(), d_offs.data_ptr()

    def run_step(ctx, flags_p, flags_c, caps_fn=None):
        ctx.reset()
        for a, b in chunks:
            ctx.partition(rp + int(rs.offs[a]), op + 8 * a, b - a, flags=capi.DEVICE_INPUT | flags_p)
        if caps_fn is not None:
            ctx.set_table_caps(caps_fn(ctx))
        ctx.count(flags_c)

    ctx = capi.Counter(w.k, w.p, w.n_parts, hpc=w.hpc, filter_mode=w.filter_mode, part_mode=w.part_mode,
                       min_count=w.min_count, max_count=w.max_count)
    forms = args.forms.split(",")
    tkey = autotune.bucket_key(w.k, rs.bases / max(rs.n, 1), N / w.n_parts, w.n_parts)
    if args.tune:
        run_step(ctx, capi.BATCHED, capi.BATCHED)
        res = autotune.tune_sort(ctx, log=_log)
        resp = autotune.tune_bin_passes(ctx, log=_log)
        cache = autotune.load_cache()
        cache[tkey] = {"sort_items": res["sort_items"], "sort_blocks_per_cu": res["sort_blocks_per_cu"],
                       "ms_pass": res["ms_pass"], "bin_passes": resp["bin_passes"],
                       "count_window": resp["count_window"], "ms_count": resp["ms_count"], "workload": w.name}
        autotune.save_cache(cache)
    tuning = autotune.apply(ctx, tkey)
    _log(f"tuning {tkey}: {tuning or 'defaults'}")
    results = {}
    predictor = learned.Predictor.load_default(w.n_parts) if "batched" in forms else None

    def timed(name, flags_p, flags_c, caps_fn=None):
        for _ in range(args.warmup):
            run_step(ctx, flags_p, flags_c, caps_fn)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            run_step(ctx, flags_p, flags_c, caps_fn)
        ctx.synchronize()
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        st = ctx.stats()
        ms = el / args.steps * 1e3
        results[name] = {"kmers_per_s": N / (ms / 1e3), "bases_per_s": rs.bases / (ms / 1e3), "ms_per_step": ms,
                         "ms_count": st["ms_count"], "host_syncs": st["host_syncs"], "fallbacks": st["fallbacks"],
                         "partition_fallbacks": st["partition_fallbacks"], "binned": st["binned"],
                         "heavy_ranges": st["heavy_ranges"], "n_distinct": st["n_distinct"],
                         "n_entries": st["n_entries"]}
        _log(f"{name}: {ms:.2f} ms/step, {N / ms / 1e6:.2f} G k-mers/s, count {st['ms_count']:.2f} ms, "
             f"syncs {st['host_syncs']}, fallbacks {st['fallbacks']}")
        return st

    if "plain" in forms:
        timed("plain", 0, 0)
    if "batched" in forms:
        caps_fn = (lambda c: predictor.caps(c)) if predictor is not None else None
        timed("batched_learned", capi.BATCHED, capi.BATCHED | (capi.PRESIZED if predictor else 0), caps_fn)

    # full-size invariants of the last table: counts sum to the k-mer total
    lo, hi, cnt, po, pd, pk = ctx.table()
    full_ok = bool(int(pk.sum()) == N and (w.min_count > 1 or int(cnt.astype(np.uint64).sum()) == N)
                   and int(pd.sum()) >= len(cnt))
    for q in range(0, w.n_parts, max(1, w.n_parts // 16)):
        a, b = int(po[q]), int(po[q + 1])
        keys = lo[a:b] if hi is None else (hi[a:b].astype(object) * (1 << 64) + lo[a:b].astype(object))
        full_ok &= bool(np.all(keys[1:] > keys[:-1]))

    # roofline of the dominant kernel (onesweep pass), measured with HIP events
    # on the library's stream in an instrumented (un-graphed) replay
    os.environ["GKC_NO_GRAPH"] = "1"
    run_step(ctx, capi.BATCHED, capi.BATCHED)
    st = ctx.stats()
    os.environ.pop("GKC_NO_GRAPH")
    stages = {"ms_partition_last_chunk": st["ms_partition"], "ms_extract": st["ms_extract"],
              "ms_sort": st["ms_sort"], "ms_sort_pass_avg": st["ms_sort_pass_avg"], "n_pass": st["n_pass"],
              "ms_count_emit": st["ms_runlength"], "ms_count_total": st["ms_count"], "binned": st["binned"],
              "heavy_ranges": st["heavy_ranges"]}
    _log(f"instrumented batched count: {stages}")
    W_key = 8 if w.k <= 32 else 16
    pass_bytes = 2 * W_key * N                   # key read + key write per pass
    pass_ms = st["ms_sort_pass_avg"]
    achieved = pass_bytes / (pass_ms / 1e3) / 1e9 if pass_ms > 0 else 0.0
    # whole-path algorithmic bytes per k-mer (read ASCII + 2-bit bins + keys + passes + table)
    r = rs.bases / N
    s_w = 4.0 * st["n_words"] / N
    h_b = 16.0 * st["n_superkmers"] / N
    delta = st["n_distinct"] / N
    # partition (ASCII in, 2-bit superkmers + headers out), extraction (bins in,
    # keys out), onesweep passes (read + write), count (keys in, entries out)
    # and, binned, the copy of the staged entries into the table
    A = (r + (s_w + h_b) + (s_w + h_b + W_key) + st["n_pass"] * 2 * W_key + (W_key + (W_key + 4) * delta)
         + (2 * (W_key + 4) * delta if st["binned"] else 0.0))
    roof_kmers = HBM_PEAK_GBS * 1e9 / A
    traffic, traffic_src = measured_traffic(w.name, ctx.get_tuning()["sort_items"])

    line = {
        "metric": METRIC,
        "value": results.get("batched_learned", results.get("plain", {})).get("kmers_per_s", 0.0),
        "unit": "k-mers/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": results.get("batched_learned", results.get("plain", {})).get("ms_per_step", 0.0),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64" if w.k <= 32 else "u128", "data": "synthetic (seeded generator, reads resident in HBM)",
        "config": {**w.describe(), "reads": rs.n, "bases": rs.bases, "kmers": N, "chunks": len(chunks),
                   "parallelism": "single GPU"},
        "forms": results,
        "speedup_batched_vs_plain": (results["batched_learned"]["kmers_per_s"] / results["plain"]["kmers_per_s"]
                                     if "plain" in results and "batched_learned" in results else None),
        "path_roofline": {"bytes_per_kmer": A, "roofline_kmers_per_s": roof_kmers,
                          "frac": (results.get("batched_learned", results.get("plain", {})).get("kmers_per_s", 0.0)
                                   / roof_kmers)},
        "roofline": {"kernel": "k_onesweep<u64>" if W_key == 8 else "k_onesweep<u128k>", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": traffic_src, "bytes_per_launch": pass_bytes,
                     "avg_launch_ms": pass_ms},
        "stages_instrumented": stages,
        "full_size_invariants_ok": full_ok,
        "learned": learned_report(predictor, ctx, pd, pk) if predictor is not None else None,
        "tuning": {"key": tkey, **ctx.get_tuning()},
    }

    if not args.no_cpu:
        from oracle import oracle as O
        nr = min(args.cpu_reads, rs.n)
        sub_reads = rs.reads[:int(rs.offs[nr])]
        sub_offs = rs.offs[:nr + 1].copy()
        threads = min(os.cpu_count() or 1, 16)
        t = time.perf_counter()
        ot = O.count(sub_reads, sub_offs, w.k, w.p, w.n_parts, hpc=w.hpc, filter_mode=w.filter_mode,
                     part_mode=w.part_mode, min_count=w.min_count, max_count=w.max_count, threads=threads)
        cpu_s = time.perf_counter() - t
        n_sub = int(ot.part_kmers.sum())
        # bit-exact check of the GPU table on the same sample, batched form
        with capi.Counter(w.k, w.p, w.n_parts, hpc=w.hpc, filter_mode=w.filter_mode, part_mode=w.part_mode,
                          min_count=w.min_count, max_count=w.max_count) as c2:
            c2.partition(sub_reads, sub_offs)
            if predictor is not None:
                c2.set_table_caps(predictor.caps(c2))
            c2.count(capi.BATCHED | (capi.PRESIZED if predictor else 0))
            glo, ghi, gcnt, gpo, _, _ = c2.table()
        exact = bool(np.array_equal(gpo, ot.part_off) and np.array_equal(glo, ot.key_lo)
                     and np.array_equal(gcnt, ot.cnt) and (ghi is None or np.array_equal(ghi, ot.key_hi)))
        line["cpu_baseline"] = {"value": n_sub / cpu_s, "unit": "k-mers/s", "cores": threads, "kind": "port",
                                "sample": f"first {nr} reads of the same workload ({n_sub} k-mers), "
                                          f"oracle phase1 {ot.times[0]:.2f}s + phase2 {ot.times[1]:.2f}s"}
        line["bit_exact_sample"] = exact
    if not args.no_table_file:
        import tempfile
        line["table_file"] = table_file_report(ctx, args.table_dir or tempfile.gettempdir(), np)
    ctx.close()
    if not args.no_ingest:
        line["ingest"] = ingest_report(rs, args.ingest_reads, not args.no_cpu, np, torch, capi, W)
    if args.calib_copy:
        a = torch.ones(1 << 28, dtype=torch.int32, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        line["calib_copy_bytes"] = {"read": a.numel() * 4, "write": b.numel() * 4}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
