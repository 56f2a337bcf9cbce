"""Benchmark of the MI355X k-mer counting path.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--config 2] [--scale 1.0]

Workload (default): the 50 Mbp, 30x, 150 bp, k=31 p=11, 256-partition short-read
set, reads resident in HBM before timing, cut into log-uniform 64 MB-1 GB
chunks.  One step = the whole hot path over the whole read set: every chunk
through encode -> minimizer -> superkmer -> partition scatter, then every
partition through extraction -> onesweep radix sort -> run-length count into
the (k-mer, count) table.

Forms timed on the headline workload:
  plain              the single-launch counter: count-then-allocate, partition
                     by partition
  batched            all partitions per launch, hipGraph replay, tables sized
                     by a counting pass and a host sync (count-then-allocate)
  batched_learned    the same with tables presized by the learned predictor,
                     run on the device inside the count (no counting pass, no
                     host round trip before the table is written; the
                     on-device check falls back)
  h2d_batched_learned  the reads start in pinned host memory: each chunk's
                     upload runs on a copy stream while the previous chunk is
                     partitioned (PCIe included; never `value`)
  h2d_pipelined_batched_learned  jobs back to back from pinned host memory:
                     as above, and the next job's uploads run during this
                     job's count (the steady state of a stream of jobs; PCIe
                     included; never `value`)
`value` is the batched + learned k-mers/s with the reads resident in HBM.
The other four workloads are measured too (plain, batched and batched +
learned, fewer steps) under `configs`.  One JSON line is printed.
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


def measured_traffic(workload: str, items: int, waves: int = 4):
    """HBM bytes per onesweep launch from the committed rocprofv3 counter runs
    (FETCH_SIZE and WRITE_SIZE scaled by a copy kernel's readings) of the same
    workload and kernel variant (keys per thread, waves per workgroup), or
    (None, None)."""
    import glob
    import re
    here = os.path.dirname(os.path.abspath(__file__))
    for f in sorted(glob.glob(os.path.join(here, "profiles", "*traffic*.md"))):
        m = re.search(r"<!-- traffic-json: (\{.*\}) -->", open(f).read())
        if m:
            d = json.loads(m.group(1))
            if d.get("workload") == workload and d.get("items") == items and d.get("waves", 4) == waves:
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
    try:
        dev = ctx.learned_caps().astype(np.float64)          # what the last learned count used
    except Exception:
        dev = None
    return {"mape_total_pct": learned.mape(tot_p, tt), "mape_distinct_pct": learned.mape(dist_p, td),
            "device_caps_over_exact": None if dev is None else float(dev.sum() / max(td.sum(), 1)),
            "table_short": None if dev is None else bool(dev.sum() < td.sum()),
            "under_predicted_partitions": int(np.sum(caps < td)), "partitions": int(len(td)),
            "table_capacity_over_exact": float(caps.sum() / max(td.sum(), 1)),
            "table_capacity_over_total_bound": float(caps.sum() / max(tt.sum(), 1)),
            "margin": predictor.margin, "trained_on": predictor.info.get("pairs")}


def ingest_report(rs, max_reads, with_cpu, np, torch, capi, W) -> dict:
    """On-device FASTQ ingest (gkc_parse_reads) of the first reads of the
    workload: the text resident in HBM, the call timed by wall clock (it
    returns when its work is done; an untimed first call allocates the
    scratch, which the timed calls reuse), its output checked against the
    reads.  Roofline: text read once, bases and offsets
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


def run_config(cid: int, scale: float, steps: int, warmup: int, forms: list[str], cpu_reads: int, tune: bool,
               headline: bool, np, torch, capi, learned, autotune, W) -> tuple[dict, dict]:
    """Generate workload `cid`, time the requested forms, check the last table's
    invariants and (on a bounded sample) its bit-exactness against the oracle,
    and measure the onesweep pass and the whole-path roofline.  Returns the
    report and a handle dict (context and data) for the headline's extras."""
    w = W.workload(cid, scale=scale)
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

    ctx = capi.Counter(w.k, w.p, w.n_parts, hpc=w.hpc, filter_mode=w.filter_mode, part_mode=w.part_mode,
                       min_count=w.min_count, max_count=w.max_count)
    predictor = learned.Predictor.load_default(w.n_parts)

    if predictor is not None:
        predictor.attach(ctx)                   # learned capacities computed on the device in the count

    def count(flags_c, use_caps):
        if use_caps and predictor is not None:
            flags_c |= capi.PRESIZED | capi.LEARNED
        ctx.count(flags_c)

    def step_resident(flags_p, flags_c, use_caps=False):
        ctx.reset()
        for a, b in chunks:
            ctx.partition(rp + int(rs.offs[a]), op + 8 * a, b - a, flags=capi.DEVICE_INPUT | flags_p)
        count(flags_c, use_caps)

    # pinned host copy of the reads and double-buffered device staging for the
    # PCIe-inclusive form (set up outside the timed region)
    h2d = None
    if "h2d_batched_learned" in forms or "h2d_pipelined_batched_learned" in forms:
        h_reads = torch.empty(rs.bases, dtype=torch.uint8, pin_memory=True)
        h_reads.numpy()[:] = rs.reads
        h_offs = torch.empty(rs.n + 1, dtype=torch.int64, pin_memory=True)
        h_offs.numpy()[:] = rs.offs.view(np.int64)
        max_b = max(int(rs.offs[b] - rs.offs[a]) for a, b in chunks)
        max_r = max(b - a for a, b in chunks)
        dbuf = [torch.empty(max_b + 64, dtype=torch.uint8, device=dev) for _ in range(2)]
        doff = [torch.empty(max_r + 1, dtype=torch.int64, device=dev) for _ in range(2)]
        s_cp = torch.cuda.Stream()
        h2d = (h_reads, h_offs, dbuf, doff, s_cp)
    pipe = None
    if "h2d_pipelined_batched_learned" in forms:   # two device copies of the whole job's reads
        pipe = {"buf": [torch.empty(rs.bases + 64, dtype=torch.uint8, device=dev) for _ in range(2)],
                "off": [torch.empty(rs.n + 1, dtype=torch.int64, device=dev) for _ in range(2)],
                "par": 0, "next": None}

    def step_h2d():
        h_reads, h_offs, dbuf, doff, s_cp = h2d
        ctx.reset()

        def upload(j):
            a, b = chunks[j]
            oa, ob = int(rs.offs[a]), int(rs.offs[b])
            with torch.cuda.stream(s_cp):
                dbuf[j % 2][:ob - oa].copy_(h_reads[oa:ob], non_blocking=True)
                doff[j % 2][:b - a + 1].copy_(h_offs[a:b + 1], non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(s_cp)
            return ev
        ev = upload(0)
        for j, (a, b) in enumerate(chunks):
            nxt = upload(j + 1) if j + 1 < len(chunks) else None   # its buffer's previous chunk is done
            ev.synchronize()
            ctx.partition(dbuf[j % 2].data_ptr(), doff[j % 2].data_ptr(), b - a,
                          flags=capi.DEVICE_INPUT | capi.BATCHED)
            ev = nxt
        count(capi.BATCHED, True)

    def upload_job(par):
        """Every chunk of the job to device copy `par` on the copy stream;
        one event per chunk."""
        h_reads, h_offs, _, _, s_cp = h2d
        evs = []
        with torch.cuda.stream(s_cp):
            for a, b in chunks:
                oa, ob = int(rs.offs[a]), int(rs.offs[b])
                pipe["buf"][par][oa:ob].copy_(h_reads[oa:ob], non_blocking=True)
                pipe["off"][par][a:b + 1].copy_(h_offs[a:b + 1], non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(s_cp)
                evs.append(ev)
        return evs

    def step_h2d_pipelined():
        """Jobs back to back, reads in pinned host memory: each chunk's
        partitioning waits for its upload, and the next job's uploads run
        during this job's count (steady state of a stream of jobs)."""
        par = pipe["par"]
        evs = pipe["next"] or upload_job(par)
        ctx.reset()
        bp, opp = pipe["buf"][par].data_ptr(), pipe["off"][par].data_ptr()
        for (a, b), ev in zip(chunks, evs):
            ev.synchronize()
            ctx.partition(bp + int(rs.offs[a]), opp + 8 * a, b - a, flags=capi.DEVICE_INPUT | capi.BATCHED)
        pipe["par"] = 1 - par
        pipe["next"] = upload_job(1 - par)
        count(capi.BATCHED, True)

    runners = {
        "plain": lambda: step_resident(0, 0),
        "batched": lambda: step_resident(capi.BATCHED, capi.BATCHED),
        "batched_learned": lambda: step_resident(capi.BATCHED, capi.BATCHED, True),
        "h2d_batched_learned": step_h2d,
        "h2d_pipelined_batched_learned": step_h2d_pipelined,
    }
    tkey = autotune.bucket_key(w.k, rs.bases / max(rs.n, 1), N / w.n_parts, w.n_parts)
    if tune:
        step_resident(capi.BATCHED, capi.BATCHED)
        res = autotune.tune_sort(ctx, log=_log)
        resp = autotune.tune_bin_passes(ctx, log=_log)
        cache = autotune.load_cache()
        cache[tkey] = {"sort_items": res["sort_items"], "sort_blocks_per_cu": res["sort_blocks_per_cu"],
                       "sort_waves": res["sort_waves"], "ms_pass": res["ms_pass"], "bin_passes": resp["bin_passes"],
                       "digit_bits": resp["digit_bits"],
                       "count_window": resp["count_window"], "ms_count": resp["ms_count"], "workload": w.name}
        autotune.save_cache(cache)
    tuning = autotune.apply(ctx, tkey)
    _log(f"tuning {tkey}: {tuning or 'defaults'}")

    results = {}
    for name in forms:
        run = runners[name]
        for _ in range(warmup):
            run()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            run()
        ctx.synchronize()
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        st = ctx.stats()
        ms = el / steps * 1e3
        results[name] = {"kmers_per_s": N / (ms / 1e3), "bases_per_s": rs.bases / (ms / 1e3), "ms_per_step": ms,
                         "ms_count": st["ms_count"], "host_syncs": st["host_syncs"], "fallbacks": st["fallbacks"],
                         "partition_fallbacks": st["partition_fallbacks"], "binned": st["binned"],
                         "heavy_ranges": st["heavy_ranges"], "n_distinct": st["n_distinct"],
                         "n_entries": st["n_entries"]}
        _log(f"{w.name} {name}: {ms:.2f} ms/step, {N / ms / 1e6:.2f} G k-mers/s, count {st['ms_count']:.2f} ms, "
             f"syncs {st['host_syncs']}, fallbacks {st['fallbacks']}")
    if h2d is not None and "h2d_batched_learned" in results:     # the uploads alone, for the PCIe rate
        h_reads, h_offs, dbuf, doff, s_cp = h2d
        torch.cuda.synchronize()
        t = time.perf_counter()
        for j, (a, b) in enumerate(chunks):
            oa, ob = int(rs.offs[a]), int(rs.offs[b])
            dbuf[j % 2][:ob - oa].copy_(h_reads[oa:ob], non_blocking=True)
        torch.cuda.synchronize()
        results["h2d_batched_learned"]["upload_only_GB_per_s"] = rs.bases / (time.perf_counter() - t) / 1e9

    # full-size invariants of the last table: counts sum to the k-mer total
    lo, hi, cnt, po, pd, pk = ctx.table()
    full_ok = bool(int(pk.sum()) == N and (w.min_count > 1 or int(cnt.astype(np.uint64).sum()) == N)
                   and int(pd.sum()) >= len(cnt))
    for q in range(0, w.n_parts, max(1, w.n_parts // 16)):
        a, b = int(po[q]), int(po[q + 1])
        keys = lo[a:b] if hi is None else (hi[a:b].astype(object) * (1 << 64) + lo[a:b].astype(object))
        full_ok &= bool(np.all(keys[1:] > keys[:-1]))
    del lo, hi, cnt

    # roofline of the dominant kernel (onesweep pass), measured with HIP events
    # on the library's stream in an instrumented (un-graphed) replay
    os.environ["GKC_NO_GRAPH"] = "1"
    step_resident(capi.BATCHED, capi.BATCHED)
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
    A = (r + (s_w + h_b) + (s_w + h_b + W_key) + st["n_pass"] * 2 * W_key + (W_key + (W_key + 4) * delta)
         + (2 * (W_key + 4) * delta if st["binned"] else 0.0))
    roof_kmers = HBM_PEAK_GBS * 1e9 / A
    tn = ctx.get_tuning()
    traffic, traffic_src = measured_traffic(w.name, tn["sort_items"], tn["sort_waves"] or 4)
    best = results.get("batched_learned", results.get("plain", {}))

    rep = {
        "workload": w.name, "config": {**w.describe(), "reads": rs.n, "bases": rs.bases, "kmers": N,
                                       "chunks": len(chunks)},
        "forms": results,
        "speedup_batched_learned_vs_plain": (results["batched_learned"]["kmers_per_s"] / results["plain"]["kmers_per_s"]
                                             if "plain" in results and "batched_learned" in results else None),
        "learned_gain_vs_batched_count_then_allocate": (
            results["batched_learned"]["kmers_per_s"] / results["batched"]["kmers_per_s"]
            if "batched" in results and "batched_learned" in results else None),
        "path_roofline": {"bytes_per_kmer": A, "roofline_kmers_per_s": roof_kmers,
                          "frac": best.get("kmers_per_s", 0.0) / roof_kmers},
        "roofline": {"kernel": "k_onesweep<u64>" if W_key == 8 else "k_onesweep<u128k>", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_source": traffic_src, "bytes_per_launch": pass_bytes,
                     "avg_launch_ms": pass_ms},
        "stages_instrumented": stages,
        "full_size_invariants_ok": full_ok,
        "learned": learned_report(predictor, ctx, pd, pk) if predictor is not None else None,
        "tuning": {"key": tkey, **ctx.get_tuning()},
    }

    if cpu_reads > 0:
        from oracle import oracle as O
        nr = min(cpu_reads, rs.n)
        sub_reads = rs.reads[:int(rs.offs[nr])]
        sub_offs = rs.offs[:nr + 1].copy()
        threads = cpu_cores()
        t = time.perf_counter()
        ot = O.count(sub_reads, sub_offs, w.k, w.p, w.n_parts, hpc=w.hpc, filter_mode=w.filter_mode,
                     part_mode=w.part_mode, min_count=w.min_count, max_count=w.max_count, threads=threads)
        cpu_s = time.perf_counter() - t
        n_sub = int(ot.part_kmers.sum())
        # bit-exact check of the GPU table on the same sample, batched + learned form
        with capi.Counter(w.k, w.p, w.n_parts, hpc=w.hpc, filter_mode=w.filter_mode, part_mode=w.part_mode,
                          min_count=w.min_count, max_count=w.max_count) as c2:
            c2.partition(sub_reads, sub_offs, flags=capi.BATCHED)
            if predictor is not None:
                predictor.attach(c2)
            c2.count(capi.BATCHED | (capi.PRESIZED | capi.LEARNED if predictor else 0))
            glo, ghi, gcnt, gpo, _, _ = c2.table()
        exact = bool(np.array_equal(gpo, ot.part_off) and np.array_equal(glo, ot.key_lo)
                     and np.array_equal(gcnt, ot.cnt) and (ghi is None or np.array_equal(ghi, ot.key_hi)))
        rep["cpu_baseline"] = {"value": n_sub / cpu_s, "unit": "k-mers/s", "cores": threads, "kind": "port",
                               "sample": f"first {nr} reads of the same workload ({n_sub} k-mers), "
                                         f"oracle phase1 {ot.times[0]:.2f}s + phase2 {ot.times[1]:.2f}s",
                               "cores_note": "the host cores this process may run on, at most the 16 a "
                                             "one-GPU node shares out per GPU"}
        rep["bit_exact_sample"] = exact
    handle = {"ctx": ctx, "rs": rs, "N": N, "W_key": W_key, "pass_ms": pass_ms, "d_reads": d_reads,
              "d_offs": d_offs, "h2d": h2d, "k": w.k, "P": w.n_parts}
    return rep, handle


def cpu_cores() -> int:
    """Host cores for the CPU baseline: the cores this process may run on,
    capped at 16, the host cores a one-GPU node shares out per GPU."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def copy_ceiling(torch, capi, nbytes: int) -> dict:
    """The library's streaming copy over nbytes (read + write), as the
    practical ceiling for a kernel that reads and writes nbytes."""
    nbytes -= nbytes % 16
    a = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    a.fill_(1)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream().cuda_stream
    capi.stream_copy(b.data_ptr(), a.data_ptr(), nbytes, st)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    best = 1e9
    for _ in range(5):
        ev[0].record()
        capi.stream_copy(b.data_ptr(), a.data_ptr(), nbytes, st)
        ev[1].record()
        ev[1].synchronize()
        best = min(best, ev[0].elapsed_time(ev[1]))
    ok = bool(torch.equal(a[-4096:], b[-4096:]))
    del a, b
    return {"bytes_each_way": nbytes, "ms": best, "GB_per_s": 2 * nbytes / best / 1e6, "output_equal": ok}


def scatter_ceiling(torch, capi, n_keys: int, n_parts: int, runs_list=(256, 128)) -> dict:
    """The library's scatter copy (one onesweep pass's write pattern: 5120-key
    tiles, each written as `runs` runs into its segment's buckets, segments the
    size of an average partition on 8 per-die lists, 3 workgroups per CU, no
    ranking, histogram or look-back) over about as many keys as the pass: the
    ceiling of that access pattern.  Best of five after a warm-up."""
    tile = 5120
    seg = max(tile, (n_keys // n_parts) // tile * tile)
    n = seg * max(1, n_keys // seg)
    a = torch.ones(n, dtype=torch.int64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream().cuda_stream
    props = torch.cuda.get_device_properties(0)
    wg = 3 * props.multi_processor_count
    wg -= wg % 8
    out = {"keys": n, "seg_keys": seg, "tile_keys": tile, "workgroups": wg}
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    for runs in runs_list:
        capi.scatter_copy(b.data_ptr(), a.data_ptr(), n, seg, runs, wg, st)
        best = 1e9
        for _ in range(5):
            ev[0].record()
            capi.scatter_copy(b.data_ptr(), a.data_ptr(), n, seg, runs, wg, st)
            ev[1].record()
            ev[1].synchronize()
            best = min(best, ev[0].elapsed_time(ev[1]))
        out[f"runs_{runs}"] = {"keys_per_run": tile // runs, "ms": best, "GB_per_s": 16 * n / best / 1e6}
    del a, b
    return out


def vendor_sort(torch, capi, n: int, begin_bit: int, end_bit: int) -> dict:
    """rocPRIM's radix sort of n random 64-bit keys on the bits the onesweep
    passes sort, timed on the device (best of three after a warm-up)."""
    a = torch.randint(0, 2 ** 62, (n,), dtype=torch.int64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream().cuda_stream
    capi.vendor_sort_u64(a.data_ptr(), b.data_ptr(), n, begin_bit, end_bit, st)
    best = min(capi.vendor_sort_u64(a.data_ptr(), b.data_ptr(), n, begin_bit, end_bit, st) for _ in range(3))
    d = (b[:1 << 20] >> begin_bit) & ((1 << (end_bit - begin_bit)) - 1)
    ok = bool((d[1:] >= d[:-1]).all())
    del a, b
    passes = (end_bit - begin_bit) / 8
    return {"library": "rocPRIM radix_sort_keys", "keys": n, "bits": [begin_bit, end_bit], "ms": best,
            "ms_per_8bit_digit": best / passes, "GB_per_s_per_8bit_digit": 2 * 8 * n / (best / passes) / 1e6,
            "sorted": ok}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-reads", type=int, default=1_500_000, help="reads in the CPU baseline sample")
    ap.add_argument("--forms", default="plain,batched,batched_learned,h2d_batched_learned,h2d_pipelined_batched_learned")
    ap.add_argument("--other-configs", default="1,3,4,5",
                    help="workloads also measured (fewer steps), comma-separated; empty for none")
    ap.add_argument("--other-cpu-reads", type=int, default=200_000, help="reads in their CPU baseline samples")
    ap.add_argument("--tune", action="store_true", help="run the autotuner on this workload and cache the result")
    ap.add_argument("--no-ingest", action="store_true", help="skip the on-device FASTQ ingest measurement")
    ap.add_argument("--ingest-reads", type=int, default=2_000_000, help="reads in the ingest sample")
    ap.add_argument("--no-table-file", action="store_true", help="skip the table file write measurement")
    ap.add_argument("--table-dir", default=None, help="directory of the table file measurement (default: temp dir)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="end with a 1 GiB library stream copy (reference kernel for HBM byte counters)")
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

    mods = dict(np=np, torch=torch, capi=capi, learned=learned, autotune=autotune, W=W)
    forms = [f for f in args.forms.split(",") if f]
    rep, hd = run_config(args.config, args.scale, args.steps, args.warmup, forms,
                         0 if args.no_cpu else args.cpu_reads, args.tune, True, **mods)
    ctx = hd["ctx"]
    main_form = rep["forms"].get("batched_learned", rep["forms"].get("plain", {}))
    # the practical ceiling of the onesweep pass: a streaming copy of its traffic
    ceil = copy_ceiling(torch, capi, hd["W_key"] * hd["N"])
    rep["roofline"]["copy_ceiling"] = ceil
    rep["roofline"]["frac_of_copy_ceiling"] = rep["roofline"]["achieved"] / ceil["GB_per_s"]
    if hd["W_key"] == 8:    # the ceiling of the pass's write pattern (runs of a 5120-key tile's 256 digits)
        sc = scatter_ceiling(torch, capi, hd["N"], hd["P"])
        rep["roofline"]["scatter_ceiling"] = sc
        rep["roofline"]["frac_of_scatter_ceiling"] = rep["roofline"]["achieved"] / sc["runs_256"]["GB_per_s"]
        _log(f"scatter copy with the pass's write pattern ({sc['seg_keys']}-key segments, 5120-key tiles): "
             + ", ".join(f"{v['keys_per_run']}-key runs {v['GB_per_s']:.0f} GB/s" for k_, v in sc.items()
                         if k_.startswith("runs_")))
    if hd["W_key"] == 8:    # the vendor library's sort of as many keys on the same 16 bits
        kb = 2 * hd["k"]
        vs = vendor_sort(torch, capi, hd["N"], max(kb - 16, 0), kb)
        vs["onesweep_ms_per_pass"] = rep["roofline"]["avg_launch_ms"]
        rep["roofline"]["vendor_sort"] = vs
        _log(f"rocPRIM radix sort of {hd['N']} keys on bits {vs['bits']}: {vs['ms']:.2f} ms "
             f"({vs['ms_per_8bit_digit']:.2f} ms per 8-bit digit; onesweep pass {vs['onesweep_ms_per_pass']:.2f} ms)")

    line = {
        "metric": METRIC,
        "value": main_form.get("kmers_per_s", 0.0),
        "unit": "k-mers/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": main_form.get("ms_per_step", 0.0),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64" if hd["W_key"] == 8 else "u128",
        "data": "synthetic (seeded generator, reads resident in HBM)",
        "config": {**rep["config"], "parallelism": "single GPU"},
    }
    line.update({k: v for k, v in rep.items() if k not in ("config", "workload")})
    if not args.no_table_file:
        import tempfile
        line["table_file"] = table_file_report(ctx, args.table_dir or tempfile.gettempdir(), np)
    ctx.close()
    rs = hd["rs"]
    del hd
    torch.cuda.empty_cache()
    if not args.no_ingest:
        line["ingest"] = ingest_report(rs, args.ingest_reads, not args.no_cpu, np, torch, capi, W)
    del rs
    others = {}
    for cid in [int(x) for x in args.other_configs.split(",") if x.strip()]:
        if cid == args.config:
            continue
        orep, ohd = run_config(cid, args.scale, min(args.steps, 3), min(args.warmup, 1),
                               ["plain", "batched", "batched_learned"],
                               0 if args.no_cpu else args.other_cpu_reads, False, False, **mods)
        ohd["ctx"].close()
        del ohd
        torch.cuda.empty_cache()
        others[orep["workload"]] = orep
    if others:
        line["configs"] = others
    if args.calib_copy:
        a = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        capi.stream_copy(b.data_ptr(), a.data_ptr(), a.numel() * 4, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        line["calib_copy_bytes"] = {"read": a.numel() * 4, "write": b.numel() * 4}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
