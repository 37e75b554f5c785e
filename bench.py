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

Beside them, the other workloads at their full sizes, each with the
algorithmic bytes and HIP-event time of every kernel it launches (achieved
GB/s against the 8 TB/s peak per pass) and an on-device check of its result:
the 100M-key uint64 radix sorts (uniform, and Zipf(1.1) over 10^6 distinct
keys), the 1e8 x 1e7 inner join and the 2e8-key hash partition.

Prints one JSON line with rows/s, per-step p50/p99, the roofline of the
dominant kernel (HIP events recorded around that kernel's launches only, on
its launch stream, inside the timed region; the per-kernel breakdown comes
from an untimed pass with events around every launch), the error of the
timed step's result against the oracle on the same inputs, and a CPU
baseline (the numpy restatement and pandas, timed on those inputs on this
host).
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


def _traffic(kernel: str, workload: str = "groupby_sum_100M"):
    """Measured HBM bytes per launch from the committed rocprof PMC summary
    profiles/<workload>_traffic.json, if one exists for this kernel."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"{workload}_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(kernel)
    except (OSError, ValueError):
        return None


def cpu_baseline(k: np.ndarray, v: np.ndarray, gpu_keys: np.ndarray, gpu_sums: np.ndarray):
    """The numpy restatement of the reference groupby (the oracle) and pandas,
    each timed on the host over the same rows the GPU step aggregated
    (single-threaded numpy semantics). The oracle's output is also the
    checker of the GPU result: group keys must be equal and the float32 sums
    are compared with its float64 sums (max relative error per group, and
    max error over the group's sum of |x|)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle
    import pandas as pd
    n = k.shape[0]
    t0 = time.perf_counter()
    ek, eres = oracle.groupby_aggregate(k, {"v": v}, {"v": "sum"})
    dt = time.perf_counter() - t0
    df = pd.DataFrame({"k": k, "v": v})
    t1 = time.perf_counter()
    ps = df.groupby("k", sort=True)["v"].sum()
    pdt = time.perf_counter() - t1
    keys_equal = bool(ek.shape == gpu_keys.shape and np.array_equal(ek, gpu_keys))
    ref = np.asarray(eres["v"], np.float64)
    kmin = int(ek[0]) if ek.shape[0] else 0
    absx = np.bincount(k.astype(np.int64) - kmin, weights=np.abs(v.astype(np.float64)))
    absx = absx[ek.astype(np.int64) - kmin]
    err = np.abs(gpu_sums.astype(np.float64) - ref) if keys_equal else np.array([np.inf])
    rel = err / np.maximum(np.abs(ref), np.finfo(np.float64).tiny) if keys_equal else err
    nrm = err / np.maximum(absx, np.finfo(np.float64).tiny) if keys_equal else err
    pandas_keys_equal = bool(np.array_equal(ps.index.to_numpy(), ek))
    check = {"vs": "oracle (numpy restatement, float64 sums) on the timed inputs",
             "keys_equal": keys_equal, "groups": int(ek.shape[0]),
             "max_rel_err": float(rel.max()), "max_norm_err": float(nrm.max()),
             "tolerance": {"max_norm_err": 2e-7, "max_rel_err_where_not_cancelling": 1e-6}}
    base = {"value": round(n / dt, 1), "unit": "rows/s", "cores": 1, "kind": "port",
            "sample": f"the {n} timed rows (int32 keys in [0,1000), float32 values), "
                      f"numpy restatement of SortBasedGroupBy.groupby_aggregate ({dt:.1f} s)",
            "pandas": {"value": round(n / pdt, 1), "unit": "rows/s",
                       "what": "DataFrame.groupby('k')['v'].sum() on the same rows",
                       "seconds": round(pdt, 3), "keys_equal_oracle": pandas_keys_equal},
            "versions": {"numpy": np.__version__, "pandas": pd.__version__,
                         "python": platform.python_version()},
            "threads": {v_: os.environ.get(v_) for v_ in
                        ("OMP_NUM_THREADS", "MKL_NUM_THREADS", "OPENBLAS_NUM_THREADS")},
            "host_cpu": _cpu_model(), "os_cpu_count": os.cpu_count()}
    return base, check


def _tuned(plan) -> dict:
    """The sort configuration the autotuner chose for a batched plan, with
    the timings of every candidate of its size class."""
    from jaxframes_amd import batched
    p0 = (plan.parts or [plan])[0]
    waves, kpt = batched.SORT_SHAPES[plan.sort_shape][0 if p0.total_bits <= 32 else 1]
    key = batched.Autotuner.key(p0.n, p0.total_bits)
    return {"digit_bits": plan.digit_bits, "sort_shape": plan.sort_shape,
            "waves_per_tile": waves, "keys_per_lane": kpt, "tile_keys": waves * 64 * kpt,
            "size_class": key, "candidates_ms": batched.TUNER.timings.get(key)}


def batched_vs_plain(total_rows: int | None, reps: int = 5, plain_frames: int = 256,
                     plain: bool = True):
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
    res = plan.result()
    groups = int(res.keys.shape[0])
    digit_bits, pipelines = plan.digit_bits, (1 if plan.parts is None else len(plan.parts))
    parts = plan.parts or [plan]
    cb = 4 if parts[0].total_bits <= 32 else 8
    # rows and groups of the frames on the one-pass tables (routed plans) and
    # of the sorted frames
    gcount = torch.diff(res.offsets).cpu().numpy()
    routed = [p for p in parts if p.route is not None]
    n_narrow = sum(p.narrow_rows for p in routed)
    f0 = np.cumsum([0] + [p.nframes for p in parts])
    g_narrow = sum(int(gcount[f0[i] + p.narrow_ids].sum()) for i, p in enumerate(parts)
                   if p.route is not None)
    n_sorted, g_sorted = n - n_narrow, groups - g_narrow
    routing = {"routed_pipelines": len(routed), "narrow_frames": sum(p.narrow_frames for p in routed),
               "narrow_rows": n_narrow, "narrow_groups": g_narrow, "sorted_rows": n_sorted,
               "sorted_groups": g_sorted}
    del res

    def bytes_of(name, launches):
        # int32 keys and float32 values; composite keys of cb bytes
        if name == "batch_compose":
            # routed: keys and values of the sorted frames read, composites and
            # values written compacted; otherwise keys read, composites written
            return n_sorted * (4 + 4 + cb + 4) if routed else n * (4 + cb)
        if name.startswith("onesweep"):
            return launches * 2 * n_sorted * (cb + 4) // pipelines
        if name == "segreduce_tiles":
            return n_sorted * (cb + 4) + g_sorted * (cb + 4 + 8)
        if name == "bn_table":
            return n_narrow * 8
        if name == "bw_copy":
            return g_sorted * 2 * (cb + 4 + 8)
        return None

    # per-kernel HIP-event times of eager launches of the same pipeline; the
    # committed PMC summary covers the 4.0e8-row scaling
    eager = _measure(lambda: plan.launch(graph=False), 3, bytes_of,
                     "batched_groupby_4096" if total_rows == 400_000_000 else None)
    # the plain path is timed without the plan's buffers held by the allocator
    tuned = _tuned(plan)
    del plan
    torch.cuda.empty_cache()
    if not plain:
        return {"workload": f"4096 frames, {n} rows, int32 keys (per-frame cardinality "
                            f"log-uniform [10, 1M]), float32 values; sum + count per frame",
                "rows": n, "groups": groups, "ms_per_batch": round(dt * 1e3, 3),
                "rows_per_s": round(n / dt, 1), "tuned": tuned, "pipelines": pipelines,
                "routing": routing,
                "eager_ms_per_batch": eager["ms_per_call"],
                "alg_bytes_per_batch": eager["alg_bytes_per_call"],
                "batch_hbm_frac": round(eager["alg_bytes_per_call"] / dt / 1e9 / PEAK_HBM_GBPS, 4),
                "roofline": eager["roofline"], "kernels": eager["kernels"]}
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
           "hip_graph": True, "tuned": tuned, "pipelines": pipelines, "routing": routing,
           "eager_ms_per_batch": eager["ms_per_call"],
           "alg_bytes_per_batch": eager["alg_bytes_per_call"],
           "batch_hbm_frac": round(eager["alg_bytes_per_call"] / dt / 1e9 / PEAK_HBM_GBPS, 4),
           "roofline": eager["roofline"], "kernels": eager["kernels"]}
    del keys, vals
    torch.cuda.empty_cache()
    return out


def served_vs_plain(nq: int = 200, qps_levels=(2000.0, 10000.0, 1e6), slots: int = 4,
                    budget_bytes: int = 64 << 30, reps: int = 5,
                    switch_interval: float | None = None):
    """The served stream: 8 resident frames (1M..50M rows, int64 keys of
    cardinality 10..1M, float32 values), 70% filter -> groupby(sum, count,
    mean) and 30% inner joins with 1e5..1e6-row dimension tables, Poisson
    arrivals at each rate in qps_levels through the request loop (slots
    workers, HBM budget), against the plain path (the same queries one at a
    time with no resident index and no batching) and against the served
    executors run one query at a time on the resident indexes (so the gain
    of the indexes and the gain of batching and overlap are reported
    apart), with each frame's index build time. switch_interval, when set,
    is the interpreter's GIL switch interval while the server is open (a
    process-wide setting, restored when it closes; by default untouched).
    Also the hash partition of 2e8 uniform int64 keys into 256 buckets."""
    import torch
    from jaxframes_amd import datagen, partition, serving
    frames = datagen.served_frames()
    store = serving.FrameStore(budget_bytes)
    index_build = []
    for name, k, v, _ in frames:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        store.load(name, k, v)
        torch.cuda.synchronize()
        index_build.append({"frame": name, "rows": int(k.shape[0]),
                            "groups": store[name].ngroups,
                            "ms": round((time.perf_counter() - t0) * 1e3, 2)})
    qs = datagen.served_queries(frames, nq)
    rows = sum(store[q.frame].n for q in qs)
    warm = qs[:8]
    for q in warm:
        serving.run_plain(store, q)
        serving.execute(store, q)
    torch.cuda.synchronize()

    def median_pass(run):
        runs = []
        for _ in range(reps):
            t0 = time.perf_counter()
            for q in qs:
                run(store, q)
            torch.cuda.synchronize()
            runs.append(time.perf_counter() - t0)
        return sorted(runs)[len(runs) // 2]
    # the plain path: no resident index, one query at a time
    plain_s = median_pass(serving.run_plain)
    # the served executors on the resident indexes, still one query at a
    # time on one stream: the gain of the indexes alone
    seq_s = median_pass(serving.execute)
    levels = []
    for qps in qps_levels:
        # one long-lived server per offered rate, warmed once (allocator pools
        # are per stream), then replayed: a new server per replay brings new
        # streams and pools, and its replays slowed one after another
        srv = serving.Server(store, slots=slots, switch_interval=switch_interval)
        srv.warmup(qs)   # every query once per slot stream
        runs = []
        for rep in range(reps):   # the replay with the median wall time is reported
            arr = serving.poisson_arrivals(len(qs), qps, seed=3 + rep)
            b0 = len(srv.batches)
            t0 = time.perf_counter()
            res = serving.replay(srv, qs, arr)
            wall = time.perf_counter() - t0
            lat = np.array([l for _, l in res]) * 1e3
            runs.append((wall, lat, srv.batches[b0:]))
        srv.close()
        runs.sort(key=lambda r: r[0])
        wall, lat, batches = runs[len(runs) // 2]
        levels.append({"offered_qps": qps, "wall_s": round(wall, 4),
                       "wall_s_all": [round(r[0], 4) for r in runs],
                       "rows_per_s": round(rows / wall, 1), "qps": round(len(qs) / wall, 1),
                       "p50_ms": round(float(np.percentile(lat, 50)), 3),
                       "p99_ms": round(float(np.percentile(lat, 99)), 3),
                       "peak_running": srv.peak_running,
                       "mean_batch": round(float(np.mean(batches)), 2),
                       "speedup_vs_plain": round(plain_s / wall, 2),
                       "speedup_vs_indexed_sequential": round(seq_s / wall, 2)})
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
            "gil_switch_interval_s": switch_interval,
            "index_build": index_build,
            "plain_rows_per_s": round(rows / plain_s, 1),
            "plain_qps": round(len(qs) / plain_s, 1),
            "indexed_sequential_rows_per_s": round(rows / seq_s, 1),
            "index_gain": round(plain_s / seq_s, 2),
            "levels": levels,
            "hash_partition": {"workload": "2e8 uniform int64 keys -> 256 buckets (fmix32)",
                               "ms": round(pdt * 1e3, 3),
                               "keys_per_s": round(2e8 / pdt, 1)}}


def _ordered_host(t) -> np.ndarray:
    """Host copy of an integer key column as unsigned integers in key order."""
    import torch
    if t.dtype in (torch.int64, torch.uint64):
        return t.view(torch.int64).cpu().numpy().view(np.uint64) ^ np.uint64(
            0 if t.dtype == torch.uint64 else 1 << 63)
    return t.view(torch.int32).cpu().numpy().view(np.uint32) ^ np.uint32(
        0 if t.dtype == torch.uint32 else 1 << 31)


def _varying_digits(t, bits: int = 8) -> int:
    """Digits of `bits` bits whose value differs between keys: the radix
    sort runs one pass per such digit (a digit equal in every key is
    skipped), so this is the pass count of the sort of t."""
    o = _ordered_host(t)
    x = int(np.bitwise_or.reduce(o) ^ np.bitwise_and.reduce(o))
    nb = o.dtype.itemsize * 8
    return sum(1 for d in range(0, nb, bits) if (x >> d) & ((1 << bits) - 1))


def _sort_pass_bytes(n: int, kb: int, pb: int, passes: int) -> int:
    """Algorithmic bytes of `passes` scatter passes sorting n keys of kb bytes
    that carry a pb-byte payload the first pass creates (the row index):
    the first pass reads the keys, every pass writes keys and payload, and
    each later pass reads both back."""
    if passes <= 0:
        return 0
    return n * kb + passes * n * (kb + pb) + (passes - 1) * n * (kb + pb)


def _sort_pass_bytes_pay(n: int, kb: int, pb: int, passes: int) -> int:
    """Algorithmic bytes of `passes` scatter passes over n keys of kb bytes
    carrying an existing pb-byte payload: every pass reads and writes both."""
    return passes * 2 * n * (kb + pb)


def _measure(step, reps: int, bytes_of, traffic_key: str | None = None) -> dict:
    """Wall time per call of step (no events), then the HIP-event time of
    every kernel launch over reps more calls; bytes_of(name, launches) gives
    a kernel's algorithmic bytes per call (None: not priced)."""
    import torch
    from jaxframes_amd import _lib
    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    with _lib.KernelTimer() as kt:
        for _ in range(reps):
            step()
        torch.cuda.synchronize()
    kernels = {}
    for name, (ms, cnt) in sorted(kt.stats().items(), key=lambda kv: -kv[1][0]):
        us = ms / reps * 1e3
        launches = max(cnt // reps, 1)
        e = {"us_per_call": round(us, 1), "launches_per_call": launches}
        b = bytes_of(name, launches)
        if b:
            gbps = b / (us * 1e-6) / 1e9
            e.update(alg_bytes_per_call=int(b), achieved_GBps=round(gbps, 1),
                     frac=round(gbps / PEAK_HBM_GBPS, 4))
        kernels[name] = e
    dom = max(kernels, key=lambda k: kernels[k]["us_per_call"])
    de = kernels[dom]
    roof = None
    if "alg_bytes_per_call" in de:
        roof = {"bound": "hbm", "kernel": dom, "achieved": de["achieved_GBps"],
                "peak": PEAK_HBM_GBPS, "unit": "GB/s", "frac": de["frac"],
                "traffic": _traffic(dom, traffic_key) if traffic_key else None,
                "avg_launch_us": round(de["us_per_call"] / de["launches_per_call"], 1),
                "alg_bytes_per_launch": de["alg_bytes_per_call"] // de["launches_per_call"]}
    total = sum(e.get("alg_bytes_per_call", 0) for e in kernels.values())
    return {"ms_per_call": round(wall * 1e3, 3), "reps": reps, "alg_bytes_per_call": total,
            "call_hbm_frac": round(total / wall / 1e9 / PEAK_HBM_GBPS, 4),
            "roofline": roof, "kernels": kernels}


def workload_sort(kind: str, reps: int, n: int = 100_000_000) -> dict:
    """Radix sort of n uint64 keys (uniform, or Zipf(1.1) over 10^6 distinct
    values) carrying the uint32 row index; the result is checked on the
    device: the index is a permutation, keys_out = keys_in[index], keys_out
    is non-decreasing and equal keys keep increasing row indices."""
    import torch
    from jaxframes_amd import datagen, sorting
    keys = datagen.sort_keys(n, kind=kind)
    passes = _varying_digits(keys)

    def step():
        return sorting.radix_sort(keys, None, True, return_index=True, index_dtype=torch.int32)

    def bytes_of(name, launches):
        if name.startswith("sort_stats") or name == "sort_hist":
            return n * 8 * launches
        if name.startswith("onesweep"):
            return _sort_pass_bytes(n, 8, 4, launches)
        return None

    out = {"workload": f"radix sort of {n} {kind} uint64 keys with the uint32 row index"
                       + (" (Zipf s=1.1 over 10^6 distinct values)" if kind == "zipf" else ""),
           "rows": n, "passes": passes, **_measure(step, reps, bytes_of, f"sort_u64_{kind}_100M")}
    out["rows_per_s"] = round(n / (out["ms_per_call"] / 1e3), 1)
    sk, _, idx = step()
    sign = -(1 << 63)
    ok = sk.view(torch.int64) ^ sign
    il = idx.long()
    ties = ok[1:] == ok[:-1]
    out["check"] = {
        "sorted": bool((ok[1:] >= ok[:-1]).all()),
        "permutation": bool(torch.bincount(il, minlength=n).max().item() == 1),
        "gather_equal": bool(torch.equal(keys.view(torch.int64)[il], sk.view(torch.int64))),
        "ties": int(ties.sum()),
        "ties_in_row_order": bool((il[1:][ties] > il[:-1][ties]).all())}
    del ok, il, ties
    out["rocprim"] = rocprim_sort(keys, sk, idx, reps)
    return out


def rocprim_sort(keys, sk, idx, reps: int) -> dict | None:
    """The vendor-library baseline: rocPRIM's radix_sort_pairs (stable, all
    64 key bits) on the same device-resident uint64 keys carrying the uint32
    row index, timed per call like jfx_sort; its output must equal jfx_sort's
    (keys and row order: both sorts are stable)."""
    import ctypes

    import torch
    from jaxframes_amd import _lib
    from jaxframes_amd.build import CMP_LIB
    if not CMP_LIB.exists():
        return None
    cmp = ctypes.CDLL(str(CMP_LIB))
    f = cmp.jfx_cmp_rocprim_sort_u64_u32
    f.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_uint64, ctypes.c_void_p,
                                          ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p]
    n = int(keys.shape[0])
    dev = keys.device
    rows = torch.arange(n, dtype=torch.int32, device=dev)
    kout = torch.empty_like(keys)
    vout = torch.empty_like(rows)
    tb = ctypes.c_uint64(0)
    stream = ctypes.c_void_p(_lib.stream_handle(dev))
    args = (keys.data_ptr(), kout.data_ptr(), rows.data_ptr(), vout.data_ptr(), n)
    if f(*args, None, ctypes.byref(tb), stream) != 0:
        return None
    temp = torch.empty(tb.value, dtype=torch.uint8, device=dev)

    def call():
        st = f(*args, temp.data_ptr(), ctypes.byref(tb), stream)
        if st != 0:
            raise RuntimeError(f"rocprim radix_sort_pairs failed: {st}")
    call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    same = bool(torch.equal(kout.view(torch.int64), sk.view(torch.int64))
                and torch.equal(vout, idx.to(torch.int32)))
    del temp, kout, vout, rows
    torch.cuda.empty_cache()
    return {"what": "rocprim::radix_sort_pairs (ROCm 7.2 rocPRIM, bits 0-64) on the same keys "
                    "with the uint32 row index", "ms_per_call": round(dt * 1e3, 3),
            "rows_per_s": round(n / dt, 1), "temp_bytes": int(tb.value),
            "output_equal_jfx_sort": same}


def workload_join(reps: int) -> dict:
    """Inner join of 1e8 x 1e7 int64 keys (right keys twice each, 80% of the
    left rows matching by Zipf(1.2) rank), two float32 payloads per side;
    checked: the pair count and non-decreasing output keys."""
    import torch
    from jaxframes_amd import datagen, joining
    lk, lv, rk, rv, m = datagen.join_tables()
    L, R = lk.shape[0], rk.shape[0]
    # the right side (the smaller) is sorted first; left rows outside its key
    # range are dropped before the left sort (they have no pair): Lk rows
    # remain, whose keys vary in pl digits
    inr = (lk >= rk.min()) & (lk <= rk.max())
    Lk = int(inr.sum())
    pl, pr = _varying_digits(lk[inr]), _varying_digits(rk)
    del inr

    def step():
        return joining.sort_merge_join(lk, lv, rk, rv, "inner")

    def bytes_of(name, launches):
        # the right side's two float32 payloads are packed into one 8-byte
        # record that rides through its sort (read and written once by the
        # packing); the left side's are packed by the pre-filter's scatter
        if name == "pack_cols":
            return R * (8 + 8)
        if name == "pf_onepass":
            # every key read once; the kept rows' payloads read, key and record written
            return L * 8 + Lk * (8 + 8 + 8)
        if name.startswith("sort_stats") or name == "sort_hist":
            return (Lk + R) * 8
        if name.startswith("onesweep"):
            if launches != pl + pr:
                return None
            return _sort_pass_bytes_pay(Lk, 8, 8, pl) + _sort_pass_bytes_pay(R, 8, 8, pr)
        if name == "join_count":
            return (Lk + R) * 8
        if name == "join_fill":
            # sorted keys and payload records of both sides read once; the key
            # and four payload columns written per pair
            return (Lk + R) * (8 + 8) + m * (8 + 4 * 4)
        if name == "gather":
            return launches * m * (4 + 4 + 4)
        return None

    out = {"workload": f"inner join of {L} x {R} int64 keys (right keys twice each, 80% of left "
                       f"rows match by Zipf(1.2)), 2 float32 payloads per side, {m} pairs",
           "rows": L + R, "pairs": m, "left_rows_in_right_range": Lk, "passes": [pl, pr],
           **_measure(step, reps, bytes_of, "join_1e8x1e7")}
    out["rows_per_s"] = round((L + R) / (out["ms_per_call"] / 1e3), 1)
    keys, cols = step()
    out["check"] = {"pairs": int(keys.shape[0]), "pairs_expected": m,
                    "keys_sorted": bool((keys[1:] >= keys[:-1]).all()),
                    "columns": sorted(cols)}
    return out


def _fmix32_bucket(k, nb: int):
    """fmix32 of the low 32 bits of int64 keys, mod nb (int64 arithmetic)."""
    m = 0xFFFFFFFF
    h = k & m
    h = h ^ (h >> 16)
    h = (h * 0x85EBCA6B) & m
    h = h ^ (h >> 13)
    h = (h * 0xC2B2AE35) & m
    h = h ^ (h >> 16)
    return h % nb


def workload_partition(reps: int, n: int = 200_000_000, nb: int = 256) -> dict:
    """Hash partition of n uniform int64 keys into nb buckets (keys only);
    checked on the device: the offsets are the bucket counts of fmix32(low 32
    bits) mod nb, and the partitioned keys' buckets are non-decreasing."""
    import torch
    from jaxframes_amd import datagen, partition
    keys = datagen.partition_keys(n)

    def step():
        return partition.hash_partition(keys, nb, return_keys=True, return_rows=False)

    def bytes_of(name, launches):
        if name == "part_hist":
            return n * 8
        if name.startswith("part_scatter") or name.startswith("part_onesweep"):
            return n * 16
        return None

    out = {"workload": f"hash partition of {n} uniform int64 keys into {nb} buckets "
                       "(fmix32), keys only",
           "rows": n, **_measure(step, reps, bytes_of, "hash_partition_2e8")}
    out["rows_per_s"] = round(n / (out["ms_per_call"] / 1e3), 1)
    r = step()
    b = _fmix32_bucket(r["keys"], nb)
    cnt = torch.bincount(_fmix32_bucket(keys, nb), minlength=nb)
    out["check"] = {"offsets_equal_counts": bool(torch.equal(r["offsets"][1:] - r["offsets"][:-1], cnt)),
                    "buckets_sorted": bool((b[1:] >= b[:-1]).all())}
    return out


def only_workload(name: str, reps: int, batched_rows: int = 400_000_000) -> dict:
    """One workload by itself (for profiling)."""
    if name == "batched":
        # the batched pipeline alone (its plain per-frame comparison launches
        # the same kernels thousands of times)
        return {"only": name, **batched_vs_plain(batched_rows or None, plain=False)}
    if name == "served":
        return {"only": name, **served_vs_plain()}
    if name in ("sort", "sort_zipf"):
        return {"only": name, **workload_sort("zipf" if name == "sort_zipf" else "uniform", reps)}
    if name == "join":
        return {"only": name, **workload_join(reps)}
    if name == "partition":
        return {"only": name, **workload_partition(reps)}
    raise SystemExit(f"unknown workload {name}")


def bandwidth_probe(nbytes: int = 800_000_000, reps: int = 5) -> dict:
    """HIP-event time of jfx_bandwidth_probe's read-only, copy and write-only
    kernels over nbytes, for every grid / loads-in-flight / non-temporal setting,
    and of its scatter copy (a radix pass's writes without the ranking) over
    run lengths; returns the best of each mode (GB/s counts each byte read or
    written once) and the whole sweep."""
    import ctypes
    import torch
    from jaxframes_amd import _lib
    dev = torch.device("cuda", 0)
    src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    stream = ctypes.c_void_p(_lib.stream_handle(dev))
    sweep = []
    for mode in (0, 1, 4):
        for blocks in (1024, 2048, 4096, 8192, 16384):
            for unroll in (1, 2, 4, 8):
                for nt in (0, 1):
                    args = (mode, src.data_ptr(), dst.data_ptr(), nbytes, blocks, unroll, nt, stream)
                    _lib.call("jfx_bandwidth_probe", *args)
                    with _lib.KernelTimer() as kt:
                        for _ in range(reps):
                            _lib.call("jfx_bandwidth_probe", *args)
                        torch.cuda.synchronize()
                    ms, cnt = next(iter(kt.stats().values()))
                    us = ms / cnt * 1e3
                    moved = nbytes * (2 if mode == 1 else 1)
                    sweep.append({"mode": {0: "read", 1: "copy", 4: "write"}[mode], "blocks": blocks,
                                  "unroll": unroll, "nontemporal": nt, "us": round(us, 1),
                                  "GBps": round(moved / (us * 1e-6) / 1e9, 1)})
    # a radix pass's scattered writes without its ranking: 100M uint64 keys
    # (with and without uint32 payloads) in 16384-key tiles written as runs
    # of 16..256 keys (64 = an 8-bit digit of a uniform key at 16384-key
    # tiles, 32 = 9 bits, 16 = 10 bits)
    tiles = 6104
    n = tiles * 16384
    del src, dst
    ssrc = torch.ones(n * 3, dtype=torch.int32, device=dev)
    sdst = torch.empty_like(ssrc)
    for mode in (2, 3):
        moved = n * (12 if mode == 2 else 8)
        for runlen in (16, 32, 64, 128, 256):
            for nt in (0, 1):
                args = (mode, ssrc.data_ptr(), sdst.data_ptr(), moved, 0, runlen, nt, stream)
                _lib.call("jfx_bandwidth_probe", *args)
                with _lib.KernelTimer() as kt:
                    for _ in range(reps):
                        _lib.call("jfx_bandwidth_probe", *args)
                    torch.cuda.synchronize()
                ms, cnt = next(iter(kt.stats().values()))
                us = ms / cnt * 1e3
                sweep.append({"mode": "scatter_pay" if mode == 2 else "scatter_keys",
                              "keys": n, "run_keys": runlen, "nontemporal": nt,
                              "us": round(us, 1),
                              "GBps": round(2 * moved / (us * 1e-6) / 1e9, 1)})
    best = {}
    for m in ("read", "copy", "write", "scatter_pay", "scatter_keys"):
        b = max((e for e in sweep if e["mode"] == m), key=lambda e: e["GBps"])
        best[m] = {**b, "frac_of_peak": round(b["GBps"] / PEAK_HBM_GBPS, 4)}
    for m in ("scatter_pay", "scatter_keys"):
        b = max((e for e in sweep if e["mode"] == m and e["run_keys"] == 64),
                key=lambda e: e["GBps"])
        best[m + "_run64"] = {**b, "frac_of_peak": round(b["GBps"] / PEAK_HBM_GBPS, 4)}
    return {"bytes": nbytes, "best": best, "sweep": sweep}


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
    ap.add_argument("--no-batched-unscaled", action="store_true",
                    help="skip the second batched entry, the unscaled distribution (about "
                         "1.9e9 rows in two pipelines of at most 2^30 rows)")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the sort, join and partition workloads")
    ap.add_argument("--only", choices=("sort", "sort_zipf", "join", "partition", "batched",
                                       "served"),
                    help="run one other workload by itself (for profiling) and print its line")
    ap.add_argument("--bandwidth-probe", action="store_true",
                    help="only sweep the hand-written read-only, copy and write-only kernels "
                         "over 800 MB (grid, loads in flight, non-temporal accesses) and the "
                         "scatter copy (a radix pass's writes without its ranking) and print "
                         "the best achieved GB/s of each: the ceilings the kernels are read "
                         "against")
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
    if args.bandwidth_probe:
        print(json.dumps({"bandwidth_probe": bandwidth_probe()}), flush=True)
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

    # the timed step's inputs and result, copied to the host after the timed
    # region for the CPU baseline and the check against the oracle
    cpu, check = None, None
    if not args.no_cpu_baseline:
        sample = min(args.cpu_sample, n)
        hk = keys[:sample].cpu().numpy()
        hv = vals[:sample].cpu().numpy()
        if sample == n:
            gk, gs = uk.cpu().numpy(), res["v"].cpu().numpy()
        else:
            uk2, res2 = grouping.groupby_aggregate(keys[:sample], {"v": vals[:sample]},
                                                   {"v": "sum"})
            gk, gs = uk2.cpu().numpy(), res2["v"].cpu().numpy()
        cpu, check = cpu_baseline(hk, hv, gk, gs)
        del hk, hv

    workloads = None
    if not args.no_workloads:
        reps = max(3, min(args.steps // 4, 10))
        workloads = {"sort_u64_uniform": workload_sort("uniform", reps),
                     "sort_u64_zipf": workload_sort("zipf", reps),
                     "join": workload_join(reps),
                     "hash_partition": workload_partition(reps)}
        torch.cuda.empty_cache()
    batched_line = None if args.no_batched else batched_vs_plain(args.batched_rows or None)
    batched_unscaled = None
    if not (args.no_batched or args.no_batched_unscaled or args.batched_rows == 0):
        batched_unscaled = batched_vs_plain(None)
    served_line = None if args.no_served else served_vs_plain()

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
        "roofline": roof,
        "check": check,
        "workloads": workloads,
        "batched": batched_line,
        "batched_unscaled": batched_unscaled,
        "served": served_line,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
