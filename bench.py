#!/usr/bin/env python3
"""Benchmark: MB of text per second to a bit-exact suffix array on one MI355X.

Workload (default): the 4 GiB byte-alphabet text -- 2^32 characters, Zipf(s = 1.1) over 256 byte values with a
64 MiB period-17 block at a seeded offset -- DC39 with character packing (9-bit characters), 64-bit SA.
A step is one full SA construction of that text, with the text already resident in HBM and the SA left in
HBM.  After the timed steps the SA is validated on the device (sa_check.hpp restated: permutation + order);
a suffix array is unique, so a valid SA is bit-exact.  The same line carries:

* the same text at DC21 and DC133, the 1 GiB genome-like text at DC39 with 40-bit SA and the 8 MiB
  i.i.d. DNA text at DC39 with 64-bit SA (``other_configs``, each timed and validated the same way);
* the host-to-host rate of the default configuration (text in pageable host memory, SA back into
  pageable host memory, PCIe included), its SA compared with the device SA byte for byte;
* ``roofline``: the dominant kernel, the LSD radix pass k_onesweep over 64-bit keys.  A launch moves N
  (u64 key, u32 value) pairs, 24 B per pair of algorithmic traffic (a first pass keyed by the character
  stream reads the stream and no identity values; a last pass that keeps only the low key words writes 4 B
  keys; passes with split positions move a byte instead of a 4-byte value); the engine sums these bytes per
  launch, and the launches are timed with hipEvents on the engine's stream during the timed steps.
  ``traffic`` is the HBM bytes per launch measured by rocprofv3 (FETCH_SIZE, WRITE_SIZE) in a separate
  profiled run, committed under profiles/ (named in ``traffic_source``);
* ``cpu_baseline``: the multi-threaded host restatement of the same algorithm (oracle/host_dcx.cpp, every
  host core) on a prefix of the same text, median of 5 runs after 1 warm-up;
* ``batched_acgtn``: the batched workload (10,000 ACGTN texts, log-uniform 1 KiB - 4 MiB, autotuned X and
  sub-batch size) in SAs/s against the plain per-text path, and ``serving_poisson``: the serving workload
  (2,000 requests, Poisson 200/s) with continuous batching against plain one-at-a-time serving, p50/p99.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--workload zipf_4gib_dc39]
    python bench.py --workload batch_acgtn | serve_poisson
    python bench.py --phase-marks ...          (profiling: phase-boundary marker kernels, a 1 GiB copy first)
    python bench.py --phase-table READ_DIR WRITE_DIR TRACE_DIR     (per-phase table from rocprofv3 csv output)
    python bench.py --radix-traffic READ_DIR WRITE_DIR TRACE_DIR   (HBM bytes per launch of the radix pass)
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MB text/s to bit-exact SA vs host DCX; batched SAs/s + p99 latency vs plain"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join("profiles", "onesweep_traffic.json")

# level-0 phases between the engine's marker kernels (k_phase_mark<P>), named after the reference's timers
# (lightweight_dcx/src/pdcx/*.hpp: phase_01_02_sort_samples, phase_02_ranks, phase_03_recursion,
# phase_04_merge); the single global sort of this engine orders the samples and the merge tuples at once
PHASES = ["compress_alphabet", "phase_01_02_sort_samples+phase_04_sort_merge_samples", "phase_02_ranks",
          "phase_03_recursion", "phase_04_merge (DC comparator on tied groups)", "output (SA width)"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def load_traffic():
    """HBM bytes per onesweep launch from the committed rocprofv3 --pmc summary, if present."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f)
    except Exception:
        return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(text, sample_bytes, X, reps=5):
    """The host DCX (oracle/host_dcx.cpp) with one thread per host core on a prefix of the text: median of
    `reps` runs after one warm-up."""
    import oracle
    s = text[:sample_bytes].copy()
    T = len(os.sched_getaffinity(0))
    oracle.host_dcx_sa(s, X=X, threads=T)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        oracle.host_dcx_sa(s, X=X, threads=T)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {"value": round(s.size / 1e6 / med, 3), "unit": "MB/s", "cores": T, "kind": "port",
            "sample": f"first {sample_bytes >> 20} MiB of the same text, multi-threaded C++ restatement of PDCX "
                      f"(oracle/host_dcx.cpp: packed sample records, sample sort over {T} threads, names, "
                      f"recursion with discarding, merge records sorted by the DC comparator), DC{X} packed, "
                      f"median of {reps} runs after 1 warm-up",
            "runs_s": [round(t, 3) for t in times], "cpu_model": cpu_model(), "nproc": os.cpu_count()}


def build_steps(eng, dev_text, n, dev_sa, width, X, steps, warmup, profile):
    import torch
    kw = dict(X=X, pack_samples=True, pack_merging=True)
    for i in range(warmup):
        st = eng.build_sa_device(dev_text.data_ptr(), n, dev_sa.data_ptr(), width, **kw)
        log(f"[bench] dc{X} n={n} warmup {i} device_ms={st['ms_device']:.2f}")
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(eng.build_sa_device(dev_text.data_ptr(), n, dev_sa.data_ptr(), width, profile_kernels=profile, **kw))
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    return ms, stats


def config_summary(name, n, X, width, ms, stats, valid):
    last = stats[-1]
    return {"workload": name, "n": n, "dcx": f"dc{X}", "sa_width_bytes": width, "ms_per_step": round(ms, 3),
            "mb_s": round(n / 1e6 / (ms / 1000.0), 2), "device_ms": round(last["ms_device"], 3),
            "phases_ms": {k: round(v, 3) for k, v in last["ms_phase"].items()}, "sa_valid": valid,
            "alg_bytes": last["alg_bytes"],
            "pipeline_alg_frac": round(last["alg_bytes"] / (ms / 1000.0) / 1e9 / HBM_PEAK_GBS, 4),
            "peak_device_bytes": last["peak_device_bytes"],
            "levels": [{k: l[k] for k in ("n", "samples", "names", "reduced", "discarding")} for l in last["levels"]]}


def bench_batch(args):
    """Batched mode: many independent texts, lengths log-uniform 1 KiB - 4 MiB over {A,C,G,T,N} (N at 0.1%),
    per-text 32-bit SAs; the batched call against the plain per-text path on a sample of the same texts."""
    import numpy as np
    import torch
    from distributed_suffix_sorting_evaluation_amd import Engine, textgen
    count = args.batch_count
    lens = textgen.lengths(count, 1024, 4 << 20, 4)
    off = np.zeros(count + 1, dtype=np.uint64)
    off[1:] = np.cumsum(lens)
    total = int(off[-1])
    log(f"[bench] batch: {count} texts, {total / 2 ** 30:.2f} GiB")
    text = textgen.generate("acgtn", total, 4)
    dt = torch.from_numpy(text).cuda()
    dsa = torch.empty(total, dtype=torch.int32, device="cuda")
    eng = Engine(0)
    if args.batch_x:
        # fixed parameters (profiling runs): no tuning builds in the trace
        x, size, x_ms, size_ms = args.batch_x, args.batch_chars, None, None
    else:
        # X and the sub-batch size by measurement on the first 10% of the texts (the SAs depend on neither)
        from distributed_suffix_sorting_evaluation_amd.autotune import choose_batch_params
        k = max(1, count // 10)
        x, size, x_ms, size_ms = choose_batch_params(eng, dt.data_ptr(), off[:k + 1], dsa.data_ptr(),
                                                     pack_samples=True, pack_merging=True)
        log(f"[bench] autotuned X={x} sub-batch={size} on {k} texts: {x_ms} {size_ms}")
    kw = dict(X=x, pack_samples=True, pack_merging=True)
    for _ in range(args.warmup):
        eng.build_sa_batch_device(dt.data_ptr(), off, dsa.data_ptr(), batch_max_chars=size, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.build_sa_batch_device(dt.data_ptr(), off, dsa.data_ptr(), batch_max_chars=size, **kw)
    torch.cuda.synchronize()
    batched_s = (time.perf_counter() - t0) / args.steps
    # plain path on an evenly spaced sample of the texts (--plain-sample 0: not measured)
    sample = np.linspace(0, count - 1, min(count, args.plain_sample)).astype(np.int64)
    out8 = torch.empty(int(lens.max()) * 4, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in sample:
        a, b = int(off[k]), int(off[k + 1])
        eng.build_sa_device(dt.data_ptr() + a, b - a, out8.data_ptr(), 4, **kw)
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    plain_rate = len(sample) / plain_s if len(sample) else None
    sample_chars = int(sum(int(lens[k]) for k in sample))
    # validate a few texts of the batched output on the device
    check = sample[:8] if len(sample) else np.linspace(0, count - 1, 8).astype(np.int64)
    valid = all(eng.check_sa_device(dt.data_ptr() + int(off[k]), int(lens[k]), dsa.data_ptr() + 4 * int(off[k]), 4)
                for k in check)
    eng.close()
    rate = count / batched_s
    return {"metric": METRIC, "value": round(rate, 1), "unit": "SAs/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(batched_s * 1e3, 2), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
            "config": {"workload": "batch_acgtn", "texts": count, "total_bytes": total, "lengths": "log-uniform 1 KiB-4 MiB",
                       "alphabet": "ACGT + N at 0.1%", "dcx": f"dc{x}" + ("" if args.batch_x else " (autotuned)"), "sa_width_bytes": 4,
                       "streams": 2},
            "autotune_ms_per_x": x_ms, "autotune_ms_per_sub_batch_chars": size_ms, "sub_batch_chars": size,
            "batched_mb_s": round(total / 1e6 / batched_s, 1),
            "plain": {"texts_per_s": round(plain_rate, 1), "sample_texts": int(len(sample)),
                      "mb_s": round(sample_chars / 1e6 / plain_s, 1)} if plain_rate else None,
            "batched_over_plain": round(rate / plain_rate, 2) if plain_rate else None, "sa_valid_sample": valid}


def bench_serve(args):
    """Serving: Poisson arrivals, sizes log-uniform 64 KiB - 256 MiB, alphabet 4 or 256 (50/50), 80 GB
    admission budget, 64-bit SAs; continuous batching (2 workers, a batch window) against the plain
    one-at-a-time path."""
    import numpy as np
    import torch
    from distributed_suffix_sorting_evaluation_amd import textgen
    from distributed_suffix_sorting_evaluation_amd.serve import (SAServer, log_uniform_sizes, poisson_arrivals,
                                                                 run_open_loop)
    count, rate = args.serve_count, args.serve_rate
    sizes = log_uniform_sizes(count, 64 << 10, 256 << 20, 5)
    rng = np.random.default_rng(5)
    sig256 = rng.random(count) < 0.5
    pool_n = 1 << 30
    pools = {False: torch.from_numpy(textgen.generate("uniform_acgt", pool_n, 5)).cuda(),
             True: torch.from_numpy(textgen.generate("uniform_sigma", pool_n, 6, sigma=256)).cuda()}
    starts = [int(rng.integers(0, pool_n - int(s) + 1)) for s in sizes]
    texts = [pools[bool(b)][o:o + int(s)] for b, o, s in zip(sig256, starts, sizes)]
    arrivals = poisson_arrivals(count, rate, 5)
    classes = np.where(sig256, 3, 0)  # the client knows its alphabet: 256 letters or ACGT
    log(f"[bench] serve: {count} requests at {rate}/s, {int(sizes.sum()) / 2 ** 30:.2f} GiB offered")
    common = dict(hbm_budget=80 * 10 ** 9, X=39, warm_chars=256 << 20, reserve_bytes=40 << 30, sa_width=8)
    # an untimed pass of the whole request set first: process-wide first-use costs (allocator growth, code
    # loading) are paid here and not by whichever timed run comes first
    srv = SAServer(**common)
    try:
        run_open_loop(srv, texts, np.zeros(count), keep_results=False, sigma_classes=classes)
    finally:
        srv.close()
    out = {}
    runs = (("continuous_batching", dict(workers=2, batching=True, batch_window_ms=args.batch_window_ms), arrivals),
            ("plain", dict(workers=1, batching=False), arrivals),
            ("continuous_batching_burst", dict(workers=2, batching=True), np.zeros(count)),
            ("plain_burst", dict(workers=1, batching=False), np.zeros(count)))
    if args.serve_runs == "cb":  # profiling: the continuous-batching Poisson run alone
        runs = runs[:1]
    elif args.serve_runs == "poisson":  # the Poisson runs, without the bursts
        runs = runs[:2]
    for name, kwargs, arr in runs:
        srv = SAServer(**common, **kwargs)
        try:
            # results are released as requests complete (2,000 64-bit SAs would not fit HBM together),
            # except every keep-th, which is validated after the run
            keep = max(1, count // 16)
            res = run_open_loop(srv, texts, arr, keep_results=False, sigma_classes=classes, keep_every=keep)
            reqs = res.pop("requests_list")
            from distributed_suffix_sorting_evaluation_amd import Engine
            with Engine(0) as chk:
                picks = [r for r in reqs if r.error == 0 and r.keep]
                res["sa_valid_sample"] = all(chk.check_sa_device(r.text.data_ptr(), r.n, r.sa.data_ptr(), 8)
                                             for r in picks)
            res["batched_requests"] = sum(1 for r in reqs if not r.large and r.error == 0 and
                                          sum(1 for q in reqs if q.batch == r.batch) > 1)
            res["large_requests"] = sum(1 for r in reqs if r.large)
            for r in reqs:
                r.sa = None
        finally:
            srv.close()
        torch.cuda.synchronize()
        out[name] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in res.items()}
        log(f"[bench] {name}: {out[name]}")
    if args.serve_runs == "cb":
        print(json.dumps({"continuous_batching": out["continuous_batching"]}), flush=True)
        sys.exit(0)
    cb, pl = out["continuous_batching"], out["plain"]
    return {"metric": METRIC, "value": cb["texts_per_s"], "unit": "texts/s", "n_gpus": 1, "steps": 1, "warmup": 0,
            "ms_per_step": round(cb["wall_s"] * 1e3, 1), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u8", "data": "synthetic",
            "config": {"workload": "serve_poisson", "requests": count, "rate_per_s": rate,
                       "sizes": "log-uniform 64 KiB-256 MiB", "alphabet": "4 or 256 (50/50)", "hbm_budget_bytes": 80 * 10 ** 9,
                       "dcx": "dc39", "sa_width_bytes": 8, "batch_window_ms": args.batch_window_ms},
            "p50_ms": cb["p50_ms"], "p99_ms": cb["p99_ms"], "continuous_batching": cb, "plain": pl,
            "burst": ({"continuous_batching": out["continuous_batching_burst"], "plain": out["plain_burst"]}
                      if "plain_burst" in out else None),
            "texts_per_s_over_plain": round(cb["texts_per_s"] / max(pl["texts_per_s"], 1e-9), 2),
            "burst_texts_per_s_over_plain": (round(out["continuous_batching_burst"]["texts_per_s"] /
                                                   max(out["plain_burst"]["texts_per_s"], 1e-9), 2)
                                             if "plain_burst" in out else None),
            "p99_over_plain": round(cb["p99_ms"] / max(pl["p99_ms"], 1e-9), 3)}


def _rocprof_rows(d, suffix):
    import csv
    import glob
    files = sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))
    if not files:
        raise SystemExit(f"no *{suffix} under {d}")
    with open(files[0]) as f:
        return list(csv.DictReader(f))


READ_COUNTERS = "TCC_EA0_RDREQ_sum TCC_EA0_RDREQ_32B_sum TCC_EA0_RDREQ_128B_sum TCC_BUBBLE_sum"
WRITE_COUNTERS = "TCC_EA0_WRREQ_sum TCC_EA0_WRREQ_64B_sum"


def _dispatch_counters(d):
    """dispatch id -> (kernel name, {counter: value}) from a rocprofv3 --pmc csv run"""
    out = {}
    for r in _rocprof_rows(d, "counter_collection.csv"):
        did = int(r["Dispatch_Id"])
        name, vals = out.setdefault(did, (r["Kernel_Name"], {}))
        vals[r["Counter_Name"]] = vals.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return out


def _read_bytes(c):
    """HBM bytes read by a dispatch from the L2's memory requests: 32-byte and 128-byte requests counted
    apart, the others 64 bytes"""
    r, r32, r128 = c["TCC_EA0_RDREQ_sum"], c["TCC_EA0_RDREQ_32B_sum"], c["TCC_EA0_RDREQ_128B_sum"]
    return 32.0 * r32 + 128.0 * r128 + 64.0 * (r - r32 - r128)


def _write_bytes(c):
    w, w64 = c["TCC_EA0_WRREQ_sum"], c["TCC_EA0_WRREQ_64B_sum"]
    return 32.0 * (w - w64) + 64.0 * w64


def phase_table(read_dir, write_dir, trace_dir):
    """Per-phase HBM traffic and kernel time of one build from three rocprofv3 runs of `bench.py --phase-marks`
    (csv output): --pmc READ_COUNTERS, --pmc WRITE_COUNTERS and --kernel-trace.  A dispatch's bytes come
    from the L2's memory request counts by request size (the derived FETCH_SIZE of this ROCm prices the
    128-byte reads of gfx950 at 64 bytes); the 1 GiB copy that --phase-marks runs first, which reads and
    writes exactly 2^30 bytes, checks them.  Dispatches are assigned to phases by the marker kernels.
    Prints a markdown table and a JSON line."""
    rd = _dispatch_counters(read_dir)
    wr = _dispatch_counters(write_dir)
    trace = _rocprof_rows(trace_dir, "kernel_trace.csv")
    tnames = {int(r["Dispatch_Id"]): r["Kernel_Name"] for r in trace}
    tdur = {int(r["Dispatch_Id"]): (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6 for r in trace}

    def phases_of(names):
        """dispatch id -> phase index from the marker kernels in dispatch order: after mark P, phase P (0: from
        the engine's first kernel to the first mark; the calibration copy before it is left out)"""
        ph, cur, started = {}, 0, False
        for did in sorted(names):
            name = names[did]
            if "k_phase_mark<" in name:
                cur = int(name.split("k_phase_mark<")[1].split(">")[0])
                continue
            started = started or "dcx::" in name
            if started:
                ph[did] = cur
        return ph

    def copy_check(items, fn):
        # the largest dispatch before the engine's first kernel: the 1 GiB fill (writes) and copy (reads and writes)
        best = (None, None, -1.0)
        for did in sorted(items):
            name, c = items[did]
            if "dcx::" in name:
                break
            v = fn(c)
            if v > best[2]:
                best = (v / float(1 << 30), name.split("(")[0], v)
        return best[0], best[1]

    rcheck, rname = copy_check(rd, _read_bytes)
    wcheck, wname = copy_check(wr, _write_bytes)
    pr = phases_of({k: v[0] for k, v in rd.items()})
    pw = phases_of({k: v[0] for k, v in wr.items()})
    pt = phases_of(tnames)
    agg = {p: {"read": 0.0, "write": 0.0, "ms": 0.0, "launches": 0, "kernels": {}} for p in range(len(PHASES))}
    for did, p in pr.items():
        if p in agg:
            agg[p]["read"] += _read_bytes(rd[did][1])
    for did, p in pw.items():
        if p in agg:
            agg[p]["write"] += _write_bytes(wr[did][1])
    for did, p in pt.items():
        if p in agg:
            agg[p]["ms"] += tdur[did]
            agg[p]["launches"] += 1
            k = tnames[did].split("(")[0]
            agg[p]["kernels"][k] = agg[p]["kernels"].get(k, 0.0) + tdur[did]
    print(f"check on the 1 GiB calibration dispatches: read by `{rname}` / 2^30 = {rcheck:.6f}, "
          f"written by `{wname}` / 2^30 = {wcheck:.6f}")
    print()
    print("| phase (reference timer) | launches | kernel ms | read GB | written GB | measured GB/s | of 8 TB/s | top kernels (ms) |")
    print("|---|---:|---:|---:|---:|---:|---:|---|")
    tot = {"read": 0.0, "write": 0.0, "ms": 0.0, "launches": 0}
    res = {}
    for p in range(len(PHASES)):
        a = agg[p]
        gbs = (a["read"] + a["write"]) / (a["ms"] / 1e3) / 1e9 if a["ms"] else 0.0
        top = sorted(a["kernels"].items(), key=lambda kv: -kv[1])[:3]
        tops = "; ".join(f"`{k.replace('dcx::', '')[:60]}` {v:.2f}" for k, v in top)
        print(f"| {PHASES[p]} | {a['launches']} | {a['ms']:.3f} | {a['read'] / 1e9:.3f} | {a['write'] / 1e9:.3f} | "
              f"{gbs:.0f} | {gbs / HBM_PEAK_GBS:.3f} | {tops} |")
        for k in tot:
            tot[k] += a[k]
        res[PHASES[p]] = {"launches": a["launches"], "kernel_ms": round(a["ms"], 3),
                              "read_bytes": int(a["read"]), "write_bytes": int(a["write"]),
                              "measured_gb_s": round(gbs, 1)}
    gbs = (tot["read"] + tot["write"]) / (tot["ms"] / 1e3) / 1e9 if tot["ms"] else 0.0
    print(f"| all phases | {tot['launches']} | {tot['ms']:.3f} | {tot['read'] / 1e9:.3f} | {tot['write'] / 1e9:.3f} | "
          f"{gbs:.0f} | {gbs / HBM_PEAK_GBS:.3f} | |")
    print()
    print(json.dumps({"phases": res, "copy_read_check": rcheck, "copy_write_check": wcheck}))


def radix_traffic(read_dir, write_dir, trace_dir):
    """HBM bytes per launch of the 64-bit radix pass (k_onesweep with 64-bit keys: the key sources
    ArrayKeys<unsigned long>, StreamKeys and CodedKeys) from the same three rocprofv3 runs as --phase-table,
    with the launch count and the mean launch time of the trace.  Prints a JSON line."""
    def radix(name):
        return "k_onesweep<" in name and ("ArrayKeys<unsigned long>" in name or "StreamKeys<" in name
                                          or "CodedKeys" in name)
    rd = {d: v for d, v in _dispatch_counters(read_dir).items() if radix(v[0])}
    wr = {d: v for d, v in _dispatch_counters(write_dir).items() if radix(v[0])}
    trace = [r for r in _rocprof_rows(trace_dir, "kernel_trace.csv") if radix(r["Kernel_Name"])]
    ms = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6 for r in trace]
    rb = sum(_read_bytes(c) for _, c in rd.values()) / max(len(rd), 1)
    wb = sum(_write_bytes(c) for _, c in wr.values()) / max(len(wr), 1)
    print(json.dumps({"launches": len(trace), "pmc_launches": [len(rd), len(wr)],
                      "read_bytes_per_launch": int(rb), "write_bytes_per_launch": int(wb),
                      "bytes_per_launch": int(rb + wb),
                      "trace_avg_launch_us": round(1e3 * sum(ms) / max(len(ms), 1), 1)}))


def overlap_table(trace_dir):
    """Device overlap of the serving workers from a rocprofv3 --kernel-trace csv run of `--workload
    serve_poisson`: per stream the kernel count and busy time, and the time during which kernels of two or
    more streams were in flight together (a sweep over the dispatch intervals)."""
    rows = _rocprof_rows(trace_dir, "kernel_trace.csv")
    key = "Stream_Id" if rows and "Stream_Id" in rows[0] else "Queue_Id"
    ev, per = [], {}
    for r in rows:
        if "dcx::" not in r["Kernel_Name"]:
            continue
        a, b, q = int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r[key]
        ev += [(a, 1, q), (b, -1, q)]
        st = per.setdefault(q, [0, 0])
        st[0] += 1
        st[1] += b - a
    ev.sort()
    active, both, busy, last = {}, 0, 0, None
    for t, d, q in ev:
        if last is not None:
            n_streams = sum(1 for v in active.values() if v > 0)
            if n_streams >= 2:
                both += t - last
            if n_streams >= 1:
                busy += t - last
        active[q] = active.get(q, 0) + d
        last = t
    print(f"streams ({key}): " + ", ".join(f"{q}: {c} kernels, {ns / 1e6:.1f} ms busy" for q, (c, ns) in sorted(per.items())))
    print(f"device busy with engine kernels: {busy / 1e6:.1f} ms; kernels of two or more streams in flight together: "
          f"{both / 1e6:.1f} ms ({both / max(busy, 1):.3f} of the busy time)")
    print(json.dumps({"streams": {q: {"kernels": c, "busy_ms": round(ns / 1e6, 3)} for q, (c, ns) in per.items()},
                      "busy_ms": round(busy / 1e6, 3), "overlap_ms": round(both / 1e6, 3)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="zipf_4gib_dc39")
    ap.add_argument("--cpu-sample-mib", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-validate", action="store_true")
    ap.add_argument("--no-host-to-host", action="store_true")
    ap.add_argument("--no-other-configs", action="store_true")
    ap.add_argument("--batch-count", type=int, default=10000)
    ap.add_argument("--plain-sample", type=int, default=500)
    ap.add_argument("--batch-x", type=int, default=0, help="batch_acgtn: this X instead of the autotuned one")
    ap.add_argument("--batch-chars", type=int, default=0,
                    help="batch_acgtn with --batch-x: characters per sub-batch (0: the engine's default)")
    ap.add_argument("--serve-count", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--batch-window-ms", type=float, default=2.0)
    ap.add_argument("--no-batch-serve", action="store_true",
                    help="default workload: leave out the batched and serving measurements "
                         "(--no-other-configs leaves them out too)")
    ap.add_argument("--serve-runs", choices=("all", "cb", "poisson"), default="all",
                    help="serve_poisson: all four runs (default), only the continuous-batching Poisson run, "
                         "or the two Poisson runs")
    ap.add_argument("--phase-marks", action="store_true",
                    help="profiling: the engine launches marker kernels at its level-0 phase boundaries, and a "
                         "1 GiB device copy runs first (the reference dispatch for the counter scale)")
    ap.add_argument("--phase-table", nargs=3, metavar=("READ_DIR", "WRITE_DIR", "TRACE_DIR"),
                    help="print the per-phase table of rocprofv3 csv output of --phase-marks runs")
    ap.add_argument("--radix-traffic", nargs=3, metavar=("READ_DIR", "WRITE_DIR", "TRACE_DIR"),
                    help="HBM bytes per launch of the 64-bit radix pass from the --phase-table runs")
    ap.add_argument("--overlap-table", metavar="TRACE_DIR",
                    help="print the stream overlap of a rocprofv3 --kernel-trace csv run of serve_poisson")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy three times (reference kernel for the "
                         "FETCH_SIZE / WRITE_SIZE counter scale)")
    args = ap.parse_args()

    if args.phase_table:
        phase_table(*args.phase_table)
        return
    if args.overlap_table:
        overlap_table(args.overlap_table)
        return
    if args.radix_traffic:
        radix_traffic(*args.radix_traffic)
        return

    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
        a.fill_(1)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 1 << 30, "copies": 3}))
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: multi-GPU sharding is design-only "
                                   "(each GPU box here has one MI355X); run with --gpus 1"}))
        sys.exit(2)

    if args.phase_marks:
        os.environ["DCX_PHASE_MARKS"] = "1"

    import numpy as np
    import torch
    from distributed_suffix_sorting_evaluation_amd import Engine, textgen

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)

    if args.workload == "batch_acgtn":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.workload == "serve_poisson":
        print(json.dumps(bench_serve(args)), flush=True)
        return

    if args.phase_marks:
        a = torch.ones(1 << 30, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    wl = textgen.WORKLOADS[args.workload]
    n, X, width = wl["n"], wl["X"], wl["sa_width"]
    log(f"[bench] generating {args.workload} n={n}")
    text = textgen.generate(wl["kind"], n, wl["seed"])
    dev_text = torch.from_numpy(text).to("cuda")
    dev_sa = torch.empty(n * width, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()

    eng = Engine(0)
    ms_per_step, steps = build_steps(eng, dev_text, n, dev_sa, width, X, args.steps, args.warmup, True)
    value = n / 1e6 / (ms_per_step / 1000.0)
    for i, st in enumerate(steps):
        log(f"[bench] step {i} device_ms={st['ms_device']:.2f} phases={st['ms_phase']}")
    valid = None
    if not args.no_validate:
        valid = eng.check_sa_device(dev_text.data_ptr(), n, dev_sa.data_ptr(), width)
        log(f"[bench] device SA check: {valid}")

    passes = sum(s["sort_passes"] for s in steps)
    items = sum(s["sort_pass_items"] for s in steps)
    pass_bytes = sum(s["sort_pass_bytes"] for s in steps)
    pass_ms = sum(s["sort_pass_ms"] for s in steps)
    avg_ms = pass_ms / max(passes, 1)
    avg_items = items / max(passes, 1)
    avg_bytes = pass_bytes / max(passes, 1)
    achieved = avg_bytes / (avg_ms / 1000.0) / 1e9 if passes else 0.0
    traffic = load_traffic()
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic.get("bytes_per_launch") if traffic else None,
                "traffic_source": (f"{TRAFFIC_FILE}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per launch on "
                                   f"{traffic.get('workload')} (committed, not measured in this run)") if traffic else None,
                "kernel": "k_onesweep (LSD radix pass over 64-bit keys: u32 values, or split indices with byte values)",
                "launches": passes, "avg_launch_ms": round(avg_ms, 4), "avg_pairs_per_launch": int(avg_items),
                "alg_bytes_per_launch": int(avg_bytes)}
    summary = config_summary(args.workload, n, X, width, ms_per_step, steps, valid)
    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "MB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": args.workload, "n": n, "text": wl["kind"], "seed": wl["seed"], "dcx": f"dc{X}",
                   "packing": True, "sa_width_bytes": width, "input": "text resident in HBM, SA left in HBM"},
        "roofline": roofline,
        "sa_valid": valid,
        "device_ms": summary["device_ms"],
        "phases_ms": summary["phases_ms"],
        "levels": summary["levels"],
        "alg_bytes": summary["alg_bytes"],
        "pipeline_alg_frac": summary["pipeline_alg_frac"],
        "peak_device_bytes": summary["peak_device_bytes"],
    }

    if not args.no_host_to_host:
        # text from pageable host memory, SA back into pageable host memory; compared with the device SA
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        sa_host = eng.build_sa(text, sa_width=width, X=X)
        t3 = time.perf_counter()
        host_bytes = sa_host.reshape(-1).view(np.uint8)
        chunk = 1 << 30
        same = host_bytes.size == dev_sa.numel() and all(
            torch.equal(torch.from_numpy(host_bytes[i:i + chunk]).cuda(), dev_sa[i:i + chunk])
            for i in range(0, host_bytes.size, chunk))
        result["host_to_host_mb_s"] = round(n / 1e6 / (t3 - t2), 2)
        result["host_to_host_sa_equals_device_sa"] = same
        log(f"[bench] host to host {result['host_to_host_mb_s']} MB/s, same SA: {same}")
        del sa_host
        torch.cuda.empty_cache()

    gpu_sample = None
    if not args.no_cpu_baseline:
        # the GPU on the same prefix the CPU baseline builds
        sb = args.cpu_sample_mib << 20
        out = torch.empty(sb * width, dtype=torch.uint8, device="cuda")
        ms_s, _ = build_steps(eng, dev_text, sb, out, width, X, 3, 1, False)
        gpu_sample = round(sb / 1e6 / (ms_s / 1000.0), 2)
        del out

    others = {}
    if not args.no_other_configs:
        # the other difference covers on the same text, then the genome and DNA configurations
        for Xo in (21, 133):
            ms, sts = build_steps(eng, dev_text, n, dev_sa, width, Xo, 2, 1, False)
            ok = None if args.no_validate else eng.check_sa_device(dev_text.data_ptr(), n, dev_sa.data_ptr(), width)
            name = f"{args.workload.rsplit('_dc', 1)[0]}_dc{Xo}"
            others[name] = config_summary(name, n, Xo, width, ms, sts, ok)
            log(f"[bench] dc{Xo}: {ms:.1f} ms")
        del dev_sa
        torch.cuda.empty_cache()
        for name in ("genome_1gib_dc39", "dna_8mib_dc39"):
            w2 = textgen.WORKLOADS[name]
            t2 = torch.from_numpy(textgen.generate(w2["kind"], w2["n"], w2["seed"])).cuda()
            s2 = torch.empty(w2["n"] * w2["sa_width"], dtype=torch.uint8, device="cuda")
            ms, sts = build_steps(eng, t2, w2["n"], s2, w2["sa_width"], w2["X"], 3, 1, False)
            ok = None if args.no_validate else eng.check_sa_device(t2.data_ptr(), w2["n"], s2.data_ptr(), w2["sa_width"])
            others[name] = config_summary(name, w2["n"], w2["X"], w2["sa_width"], ms, sts, ok)
            log(f"[bench] {name}: {ms:.2f} ms")
            del t2, s2
        result["other_configs"] = others
    eng.close()

    if not args.no_cpu_baseline:
        log(f"[bench] cpu baseline on {args.cpu_sample_mib} MiB")
        cb = cpu_baseline(text, args.cpu_sample_mib << 20, X)
        cb["gpu_same_sample_mb_s"] = gpu_sample
        result["cpu_baseline"] = cb
        result["gpu_over_cpu_baseline"] = round(value / cb["value"], 1)
        result["gpu_over_cpu_same_sample"] = round(gpu_sample / cb["value"], 1) if gpu_sample else None
    else:
        result["cpu_baseline"] = None

    if not (args.no_batch_serve or args.no_other_configs):
        # the batched and the serving workloads in the same run (each also has a line of its own:
        # --workload batch_acgtn, serve_poisson), after the 4 GiB text's buffers are released
        dev_text = dev_sa = text = None
        torch.cuda.empty_cache()
        b = bench_batch(args)
        result["batched_acgtn"] = {k: b[k] for k in ("value", "unit", "ms_per_step", "batched_mb_s", "plain",
                                                     "batched_over_plain", "sa_valid_sample", "sub_batch_chars")}
        result["batched_acgtn"]["config"] = b["config"]
        torch.cuda.empty_cache()
        sv = bench_serve(argparse.Namespace(**{**vars(args), "serve_runs": "poisson"}))
        result["serving_poisson"] = {k: sv[k] for k in ("value", "unit", "p50_ms", "p99_ms", "texts_per_s_over_plain",
                                                        "p99_over_plain", "continuous_batching", "plain")}
        result["serving_poisson"]["config"] = sv["config"]
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
