#!/usr/bin/env python3
"""Benchmark of the sort-based groupby hot path on one MI355X.

Workload (one step): groupby-sum of 100M rows, int32 keys uniform in
[0, 1000), float32 values N(0, 1), synthetic and seeded, generated on the
device before the timed region. A step is one full `groupby_aggregate(keys,
{'v': vals}, {'v': 'sum'})` as `JaxFrame.groupby('k').agg({'v': 'sum'})` runs
it: integer keys whose range turns out below 1024 take the one-pass kernel
(one read of keys and values into wave-private LDS tables indexed by key mod
1024, with the row counts and the key min / max), then the in-order merge of
the block tables that checks the range and writes the groups in key order,
and the read-back of the group count. (Wider ranges fall back to the key
statistics pass and the sort- or table-based paths.)

Beside the step: the batched groupby (4096 frames in one HIP-graph replay)
and the served stream (request loop over resident frames, Poisson arrivals)
against their plain per-query paths, and the 2e8-key hash partition.

Prints one JSON line with rows/s, per-step p50/p99, the roofline of the
dominant kernel (HIP events recorded around that kernel's launches only, on
its launch stream, inside the timed region; the per-kernel breakdown comes
from an untimed pass with events around every launch), and a CPU baseline
(the numpy restatement timed on a bounded sample on this host).
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

PEAK_HBM_GBPS = 8000.0  # MI355X HBM3E
PROF_EVERY = 4     # dominant-kernel launches per timed one in the timed region


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _kernel_bytes(name: str, n: int, kb: int, vb: int) -> float | None:
    """Algorithmic HBM bytes of one launch of a groupby-path kernel."""
    if name in ("onepass_table",) or name.startswith("wave_table") or name.startswith("tile_reduce"):
        return n * (kb + vb)                 # read keys + values once (folded in LDS)
    if name.startswith("chunk_scatter"):
        return n * (kb + vb) + n * vb        # read keys + values, write values
    if name == "segreduce_tiles":
        return n * vb                        # read sorted values (known groups)
    if name == "sort_stats":
        return n * kb                        # read keys
    return None


def _traffic(kernel: str):
    """Measured HBM bytes per launch from the committed rocprof PMC summary,
    if one exists for this kernel and workload."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "groupby_sum_100M_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(kernel)
    except (OSError, ValueError):
        return None


def cpu_baseline(n_sample: int, seed: int):
    """The numpy restatement of the reference groupby (oracle) on a bounded
    sample of the same distribution, single-threaded numpy."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 1000, n_sample, dtype=np.int32)
    v = rng.standard_normal(n_sample, dtype=np.float32)
    t0 = time.perf_counter()
    oracle.groupby_aggregate(k, {"v": v}, {"v": "sum"})
    dt = time.perf_counter() - t0
    return n_sample / dt, dt


def batched_vs_plain(total_rows: int | None, reps: int = 5, plain_frames: int = 256):
    """Batched groupby of 4096 frames (row counts log-uniform in [1K, 4M]
    scaled to total_rows, or unscaled when it is None; int32 keys with
    per-frame cardinality log-uniform in [10, 1M], float32 N(0,1) values) as
    HIP-graph replays (one per pipeline of at most 2^30 rows), against the
    plain path (one groupby_aggregate call per frame) on the first
    plain_frames frames."""
    import torch
    from jaxframes_amd import batched, datagen
    keys, vals, offs, cards = datagen.batched_frames(total_rows=total_rows)
    n = int(keys.shape[0])
    plan = batched.BatchedGroupBy(keys, vals, offs, key_range=(0, int(cards.max()) - 1),
                                  aggs=("sum", "count"))
    plan.run(graph=True)
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        a = time.perf_counter()
        plan.launch(graph=True)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
    dt = (time.perf_counter() - t0) / reps
    groups = int(plan.result().keys.shape[0])
    digit_bits, pipelines = plan.digit_bits, (1 if plan.parts is None else len(plan.parts))
    # the plain path is timed without the plan's buffers held by the allocator
    del plan
    torch.cuda.empty_cache()
    # warmed over the whole sample (the allocator grows its cached blocks to
    # every frame's workspace size once), then the median of 3 passes
    rows_plain = int(offs[plain_frames])
    batched.plain_per_frame(keys, vals, offs, "sum", frames=range(plain_frames))
    torch.cuda.synchronize()
    plain_t = []
    for _ in range(3):
        t1 = time.perf_counter()
        batched.plain_per_frame(keys, vals, offs, "sum", frames=range(plain_frames))
        torch.cuda.synchronize()
        plain_t.append(time.perf_counter() - t1)
    plain_rps = rows_plain / float(np.median(plain_t))
    out = {"workload": f"4096 frames, {n} rows, int32 keys (per-frame cardinality log-uniform "
                       f"[10, 1M]), float32 values; sum + count per frame",
           "rows": n, "frames": int(offs.shape[0] - 1), "groups": groups,
           "ms_per_batch": round(dt * 1e3, 3), "rows_per_s": round(n / dt, 1),
           "p50_ms": round(float(np.percentile(lat, 50)), 3),
           "p99_ms": round(float(np.percentile(lat, 99)), 3),
           "plain_rows_per_s": round(plain_rps, 1),
           "plain_sample": f"first {plain_frames} frames, {rows_plain} rows",
           "speedup_vs_plain": round(n / dt / plain_rps, 2),
           "hip_graph": True, "digit_bits": digit_bits, "pipelines": pipelines}
    del keys, vals
    torch.cuda.empty_cache()
    return out


def served_vs_plain(nq: int = 200, qps_levels=(2000.0, 1e6), slots: int = 4,
                    budget_bytes: int = 64 << 30, reps: int = 5):
    """The served stream: 8 resident frames (1M..50M rows, int64 keys of
    cardinality 10..1M, float32 values), 70% filter -> groupby(sum, count,
    mean) and 30% inner joins with 1e5..1e6-row dimension tables, Poisson
    arrivals at each rate in qps_levels through the request loop (slots
    workers, HBM budget), against the plain path (the same queries one at a
    time with no resident index and no batching). Also the hash partition of
    2e8 uniform int64 keys into 256 buckets."""
    import torch
    from jaxframes_amd import datagen, partition, serving
    frames = datagen.served_frames()
    store = serving.FrameStore(budget_bytes)
    for name, k, v, _ in frames:
        store.load(name, k, v)
    qs = datagen.served_queries(frames, nq)
    rows = sum(store[q.frame].n for q in qs)
    warm = qs[:8]
    for q in warm:
        serving.run_plain(store, q)
    torch.cuda.synchronize()
    plain_runs = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for q in qs:
            serving.run_plain(store, q)
        torch.cuda.synchronize()
        plain_runs.append(time.perf_counter() - t0)
    plain_s = sorted(plain_runs)[len(plain_runs) // 2]
    levels = []
    for qps in qps_levels:
        runs = []
        for rep in range(reps):   # the replay with the median wall time is reported
            srv = serving.Server(store, slots=slots)
            srv.warmup(qs)   # every query once per slot stream: allocator pools are per stream
            arr = serving.poisson_arrivals(len(qs), qps, seed=3 + rep)
            t0 = time.perf_counter()
            res = serving.replay(srv, qs, arr)
            wall = time.perf_counter() - t0
            srv.close()
            lat = np.array([l for _, l in res]) * 1e3
            runs.append((wall, lat, srv))
        runs.sort(key=lambda r: r[0])
        wall, lat, srv = runs[len(runs) // 2]
        levels.append({"offered_qps": qps, "wall_s": round(wall, 4),
                       "wall_s_all": [round(r[0], 4) for r in runs],
                       "rows_per_s": round(rows / wall, 1), "qps": round(len(qs) / wall, 1),
                       "p50_ms": round(float(np.percentile(lat, 50)), 3),
                       "p99_ms": round(float(np.percentile(lat, 99)), 3),
                       "peak_running": srv.peak_running,
                       "mean_batch": round(float(np.mean(srv.batches)), 2),
                       "speedup_vs_plain": round(plain_s / wall, 2)})
    pk = datagen.partition_keys(200_000_000)
    partition.hash_partition(pk, 256, return_keys=True, return_rows=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        partition.hash_partition(pk, 256, return_keys=True, return_rows=False)
    torch.cuda.synchronize()
    pdt = (time.perf_counter() - t0) / 5
    del pk, store
    torch.cuda.empty_cache()
    return {"workload": "8 resident frames (rows log-uniform [1e6, 5e7], int64 keys, "
                        "cardinality log-uniform [10, 1e6], float32 values); 70% filter(v > t) "
                        "-> groupby(k).agg(sum, count, mean), 30% inner join with a "
                        "1e5..1e6-row dimension table; Poisson arrivals",
            "queries": len(qs), "rows": rows, "slots": slots, "hbm_budget_bytes": budget_bytes,
            "plain_rows_per_s": round(rows / plain_s, 1),
            "plain_qps": round(len(qs) / plain_s, 1), "levels": levels,
            "hash_partition": {"workload": "2e8 uniform int64 keys -> 256 buckets (fmix32)",
                               "ms": round(pdt * 1e3, 3),
                               "keys_per_s": round(2e8 / pdt, 1)}}


def only_workload(name: str, reps: int, batched_rows: int = 400_000_000) -> dict:
    """One workload by itself, for profiling: the kernels it launches, with
    their HIP-event time per call, and the wall time per call."""
    import torch
    from jaxframes_amd import _lib, datagen, joining, partition, sorting
    if name == "batched":
        return {"only": name, **batched_vs_plain(batched_rows or None)}
    if name == "served":
        return {"only": name, **served_vs_plain()}
    if name == "sort":
        keys = datagen.sort_keys(100_000_000)
        desc = "radix sort of 100M uniform uint64 keys with the uint32 row index"
        step = lambda: sorting.radix_sort(keys, None, True, return_index=True,
                                          index_dtype=torch.int32)
    elif name == "join":
        lk, lv, rk, rv, m = datagen.join_tables()
        desc = (f"inner join of 1e8 x 1e7 int64 keys (right keys twice each, 80% of left rows "
                f"match by Zipf(1.2)), 2 float32 payloads per side, {m} pairs")
        step = lambda: joining.sort_merge_join(lk, lv, rk, rv, "inner")
    elif name == "partition":
        keys = datagen.partition_keys(200_000_000)
        desc = "hash partition of 2e8 uniform int64 keys into 256 buckets (fmix32), keys only"
        step = lambda: partition.hash_partition(keys, 256, return_keys=True, return_rows=False)
    else:
        raise SystemExit(f"unknown workload {name}")
    step()
    torch.cuda.synchronize()
    with _lib.KernelTimer() as kt:
        t0 = time.perf_counter()
        for _ in range(reps):
            step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
    ks = kt.stats()
    return {"only": name, "workload": desc, "ms_per_call": round(dt * 1e3, 3), "reps": reps,
            "kernels": {k: {"us_per_call": round(v[0] / reps * 1e3, 1),
                            "launches_per_call": v[1] // reps}
                        for k, v in sorted(ks.items(), key=lambda kv: -kv[1][0])}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=100_000_000)
    ap.add_argument("--cpu-sample", type=int, default=100_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-rows", type=int, default=400_000_000,
                    help="rows of the batched groupby (the 4096 frames' row counts are scaled "
                         "to it); 0 = the unscaled distribution (about 2e9 rows)")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--only", choices=("sort", "join", "partition", "batched", "served"),
                    help="run one other workload by itself (for profiling) and print its line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 3 device copies of 400 MB (float32[1e8]) so that PMC byte "
                         "counters can be scaled against a transfer of known size")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(this build has no multi-GPU path)", file=sys.stderr)
        return 2

    import torch
    from jaxframes_amd import _lib, grouping

    _lib.require_gpu()
    dev = torch.device("cuda", 0)
    if args.only:
        print(json.dumps(only_workload(args.only, max(args.steps // 4, 3), args.batched_rows)),
              flush=True)
        return 0
    if args.copy_calibration:
        a = torch.ones(100_000_000, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read": a.numel() * 4,
                                               "bytes_written": b.numel() * 4, "launches": 3}}))
        return 0
    n = args.rows
    g = torch.Generator(device=dev)
    g.manual_seed(1234)
    keys = torch.randint(0, 1000, (n,), dtype=torch.int32, device=dev, generator=g)
    vals = torch.randn(n, dtype=torch.float32, device=dev, generator=g)

    def step():
        return grouping.groupby_aggregate(keys, {"v": vals}, {"v": "sum"})

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # per-kernel breakdown from an untimed pass with events around every launch
    nb = max(3, args.warmup)
    with _lib.KernelTimer() as kt_all:
        for _ in range(nb):
            step()
        torch.cuda.synchronize()
    breakdown = kt_all.stats()
    dname = max(breakdown.items(), key=lambda kv: kv[1][0])[0]

    # timed region: events around one launch in four of the dominant kernel
    # (a pair of events on the stream lengthens the step they bracket)
    with _lib.KernelTimer(prefix=dname, every=PROF_EVERY) as kt:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            uk, res = step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    kstats = kt.stats()
    elapsed = t1 - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = n * args.steps / elapsed

    # per-step latency distribution (each step synchronised on its own)
    lat = []
    for _ in range(max(5, min(args.steps, 20))):
        torch.cuda.synchronize()
        a = time.perf_counter()
        step()
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
    p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
    assert _lib.device_error() == 0
    ngroups = int(uk.shape[0])

    # roofline of the dominant kernel, from its events in the timed region
    dms, dcount = kstats[dname]
    avg_s = dms / 1e3 / dcount
    dbytes = _kernel_bytes(dname, n, 4, 4)
    roof = None
    if dbytes is not None:
        ach = dbytes / avg_s / 1e9
        roof = {"bound": "hbm", "kernel": dname, "achieved": round(ach, 1),
                "peak": PEAK_HBM_GBPS, "unit": "GB/s", "frac": round(ach / PEAK_HBM_GBPS, 4),
                "traffic": _traffic(dname), "avg_launch_us": round(avg_s * 1e6, 1),
                "timed_launches": dcount,
                "alg_bytes_per_launch": dbytes}
    kernels = {k: {"us_per_step": round(v[0] / nb * 1e3, 1),
                   "launches_per_step": v[1] // nb} for k, v in breakdown.items()}
    # both small-range kernels are launched and the unselected one returns at
    # once: count the bytes of the one that ran (the longer)
    ran = dict(breakdown)
    alts = [k for k in ran if k.startswith("wave_table") or k.startswith("tile_reduce")]
    for k in alts:
        if k != max(alts, key=lambda a: ran[a][0]):
            ran.pop(k)
    step_bytes = sum((_kernel_bytes(k, n, 4, 4) or 0) * (v[1] // nb) for k, v in ran.items())

    sort_based = n * 4 + 2 * n * 8 + n * 8 + ngroups * (4 + 16)
    batched_line = None if args.no_batched else batched_vs_plain(args.batched_rows or None)
    served_line = None if args.no_served else served_vs_plain()

    cpu = None
    if not args.no_cpu_baseline:
        sample = min(args.cpu_sample, n)
        rps, dt = cpu_baseline(sample, seed=1234)
        cpu = {"value": round(rps, 1), "unit": "rows/s", "cores": 1, "kind": "port",
               "sample": f"{sample} rows, same distribution, numpy restatement of "
                         f"SortBasedGroupBy.groupby_aggregate ({dt:.1f} s)",
               "host_cpu": _cpu_model(), "os_cpu_count": os.cpu_count()}

    line = {
        "metric": "groupby-sum rows/s + % HBM peak at 100M rows; batched vs plain; p50/p99 ms",
        "value": round(value, 1),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "p50_ms": round(p50, 4),
        "p99_ms": round(p99, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (device-generated, seeded)",
        "config": {"workload": "groupby-sum, 100M rows, int32 keys uniform [0,1000), "
                               "float32 N(0,1) values",
                   "rows": n, "groups": ngroups, "agg": "sum",
                   "path": "one pass: wave-private LDS tables indexed by key mod 1024, "
                           "key range checked after"},
        "step_alg_bytes": step_bytes,
        "step_hbm_frac": round(step_bytes / (ms_per_step / 1e3) / 1e9 / PEAK_HBM_GBPS, 4),
        # the same step priced at the sort-based groupby's algorithmic bytes:
        # key histogram N*kb + one LSD pass 2N(kb+vb) + segmented-reduce read
        # N(kb+vb) + group outputs G*(kb+16); this path reads N(kb+vb) once
        "step_bytes_sort_based": sort_based,
        "step_frac_sort_based": round(sort_based / (ms_per_step / 1e3) / 1e9 / PEAK_HBM_GBPS, 4),
        "roofline": roof,
        "batched": batched_line,
        "served": served_line,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
