#!/usr/bin/env python3
"""bench.py -- headline benchmark of the vectis MI355X engine.

Default workload (one MI355X): the sort-based GROUP BY over 100M rows
  key INT64 Zipf(s=1.1) over 1M distinct values (randomly permuted), v1 INT64 uniform [-1e6, 1e6],
  v2 DOUBLE normal(0, 1);
  SELECT key, COUNT(*), SUM(v1), MIN(v1), MAX(v1), SUM(v2), AVG(v2) GROUP BY key
with every column resident in HBM before the timed region. One step = one full GROUP BY through the C ABI
(key-range pass, histogram, radix passes, segmented reduction, group-count readback).

Prints ONE JSON line (rank 0). `roofline` is the dominant kernel's algorithmic bytes / its average launch
time (HIP events on its own stream) against 8 TB/s; `cpu_baseline` is the multithreaded C restatement
(oracle/, all host threads) timed on the same data on this host. `operators` carries the per-operator
lines (fused filter/project, join, cross product) measured the same way.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "rows/s + % HBM roofline per operator; batched vs per-query QPS at p99"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=100_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-rows", type=int, default=100_000_000, help="rows in the CPU baseline sample")
    ap.add_argument("--only", default="", help="comma list of extra operator lines (none by default)")
    ap.add_argument("--sweep", action="store_true", help="also time each sort tuning candidate (stderr)")
    ap.add_argument("--no-autotune", action="store_true", help="use the static radix-pass variant choice")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (PMC calibration of FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def calibrate_copy():
    """A plain copy kernel of known size: 1 GiB read + 1 GiB written per launch. Profiled with the same
    rocprofv3 --pmc passes as the engine's kernels, it gives the factor between counter units and bytes."""
    import torch
    n = 1 << 27  # int64 elements = 1 GiB
    src = torch.arange(n, dtype=torch.int64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8,
                      "launches": 5}), flush=True)


def measure_copy_gbs(torch, dev, gib=4, reps=10):
    """Achievable stream-copy bandwidth on this GPU: a device-to-device copy of `gib` GiB through the
    library's block-chunked copy kernel (read + write bytes / time, HIP events on the launch stream), the
    practical ceiling next to the 8 TB/s peak. Also returns the same figure for torch's copy_ (the runtime's
    copy kernel), which is slower."""
    from vectisdb_amd import capi
    n = gib * (1 << 30) // 8
    a = torch.ones(n, dtype=torch.int64, device=dev)
    b = torch.empty_like(a)
    out = {}
    for name, fn in (("engine", lambda: capi.device_copy(b, a)), ("torch", lambda: b.copy_(a))):
        for _ in range(2):
            fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        out[name] = 2 * gib * (1 << 30) / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9
    if not torch.equal(a[:1024], b[:1024]) or not torch.equal(a[-1024:], b[-1024:]):
        raise RuntimeError("device copy check failed")
    del a, b
    torch.cuda.empty_cache()
    return out["engine"], out["torch"]


def committed_traffic(kernel):
    """HBM traffic per launch of `kernel` measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by
    the copy-kernel calibration) and committed under profiles/; None when no measurement is committed."""
    path = ROOT / "profiles" / "traffic.json"
    if not path.exists():
        return None
    try:
        data = json.loads(path.read_text())
    except ValueError:
        return None
    entry = data.get("kernels", {}).get(kernel)
    return None if entry is None else entry.get("traffic_bytes_per_launch")


def zipf_tables(torch, dev, distinct, s, seed):
    import numpy as np
    ranks = np.arange(1, distinct + 1, dtype=np.float64)
    w = ranks ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    cdf[-1] = 1.0 + 1e-12  # every u in [0,1) lands inside the table
    perm = np.random.RandomState(seed).permutation(distinct).astype(np.int64)
    return torch.from_numpy(cdf).to(dev), torch.from_numpy(perm).to(dev)


def make_groupby_table(torch, capi, dev, rows, seed=1):
    cdf, perm = zipf_tables(torch, dev, 1_000_000, 1.1, seed)
    key = torch.empty(rows, dtype=torch.int64, device=dev)
    v1 = torch.empty(rows, dtype=torch.int64, device=dev)
    v2 = torch.empty(rows, dtype=torch.float64, device=dev)
    capi.generate_column(key, capi.INT64, 3, seed=seed, stream_id=1, cdf=cdf, perm=perm)
    capi.generate_column(v1, capi.INT64, 0, lo=-1_000_000, hi=1_000_000, seed=seed, stream_id=2)
    capi.generate_column(v2, capi.FLOAT64, 2, seed=seed, stream_id=3)
    torch.cuda.synchronize()
    return key, v1, v2


def kernel_stats(prof, name):
    for e in prof:
        if e["name"] == name:
            avg = e["ms"] / max(e["launches"], 1)
            b = e["bytes"] / max(e["launches"], 1)
            return avg, b
    return None, None


def bench_filter_project(torch, capi, dev, steps, copies=16, rows=1_048_576):
    """Fused scan-filter-project over 1,048,576 rows (SELECT a*2+b, c*1.5+a WHERE a < 500000 AND c > 0.25); `copies` table copies are rotated so the
    27 MB working set is not served from the 256 MB Infinity Cache."""
    sys.path.insert(0, str(ROOT / "tests"))
    from exprs import cfg0_exprs
    preds, projs = cfg0_exprs()
    tables = []
    for c in range(copies):
        a = torch.empty(rows, dtype=torch.int32, device=dev)
        b = torch.empty(rows, dtype=torch.int64, device=dev)
        x = torch.empty(rows, dtype=torch.float64, device=dev)
        capi.generate_column(a, capi.INT32, 0, lo=0, hi=999_999, seed=c, stream_id=1)
        capi.generate_column(b, capi.INT64, 0, lo=-(1 << 40), hi=(1 << 40) - 1, seed=c, stream_id=2)
        capi.generate_column(x, capi.FLOAT64, 1, seed=c, stream_id=3)
        tables.append((a, b, x))
    torch.cuda.synchronize()
    for t in tables[:2]:
        capi.filter_project(list(t), preds, projs)
    capi.profile_reset()
    capi.profile_enable(True)
    kept = 0
    t0 = time.perf_counter()
    for i in range(steps):
        _, _, _, kept = capi.filter_project(list(tables[i % copies]), preds, projs)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    capi.profile_enable(False)
    avg, inb = kernel_stats(capi.profile_read(), "fused_filter_project")
    # algorithmic bytes: read a, b, c (20 B/row) + write kept rows' two 8-byte outputs
    alg = rows * 20 + kept * 16
    gbs = alg / (avg * 1e-3) / 1e9
    return {"operator": "fused filter/project, 1,048,576 rows", "rows": rows, "kept": kept,
            "kernel_ms": round(avg, 5), "rows_per_s": round(rows / (avg * 1e-3), 1),
            "call_ms": round((t1 - t0) * 1e3 / steps, 4), "algorithmic_bytes": alg, "gbs": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


def bench_filter_project_large(torch, capi, dev, steps, rows=1 << 28):
    """The config-0 query (same columns, distributions and expressions) over 2^28 rows: the same generated
    kernel when it is not launch-bound, i.e. its streaming efficiency."""
    sys.path.insert(0, str(ROOT / "tests"))
    from exprs import cfg0_exprs
    preds, projs = cfg0_exprs()
    a = torch.empty(rows, dtype=torch.int32, device=dev)
    b = torch.empty(rows, dtype=torch.int64, device=dev)
    x = torch.empty(rows, dtype=torch.float64, device=dev)
    capi.generate_column(a, capi.INT32, 0, lo=0, hi=999_999, seed=5, stream_id=1)
    capi.generate_column(b, capi.INT64, 0, lo=-(1 << 40), hi=(1 << 40) - 1, seed=5, stream_id=2)
    capi.generate_column(x, capi.FLOAT64, 1, seed=5, stream_id=3)
    torch.cuda.synchronize()
    capi.filter_project([a, b, x], preds, projs)
    capi.profile_reset()
    capi.profile_enable(True)
    kept = 0
    for _ in range(steps):
        _, _, _, kept = capi.filter_project([a, b, x], preds, projs)
    torch.cuda.synchronize()
    capi.profile_enable(False)
    avg, _ = kernel_stats(capi.profile_read(), "fused_filter_project")
    alg = rows * 20 + kept * 16
    gbs = alg / (avg * 1e-3) / 1e9
    return {"operator": "fused filter/project, config-0 query over 2^28 rows (streaming efficiency)",
            "rows": rows, "kept": kept, "kernel_ms": round(avg, 4), "rows_per_s": round(rows / (avg * 1e-3), 1),
            "algorithmic_bytes": alg, "gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


def bench_filter_gt(torch, capi, dev, steps, rows=1 << 28):
    """FilterGreaterThanInt32 (vgpu_filter_gt_i32) over 2^28 INT32 rows uniform in [0, 2^30) with the threshold
    at the middle: about half the rows kept, in order, by one look-back compaction kernel."""
    x = torch.empty(rows, dtype=torch.int32, device=dev)
    capi.generate_column(x, capi.INT32, 0, lo=0, hi=(1 << 30) - 1, seed=13, stream_id=1)
    torch.cuda.synchronize()
    threshold = 1 << 29
    kept = capi.filter_gt_i32(x, threshold).numel()
    capi.profile_reset()
    capi.profile_enable(True)
    for _ in range(steps):
        capi.filter_gt_i32(x, threshold)
    torch.cuda.synchronize()
    capi.profile_enable(False)
    avg, _ = kernel_stats(capi.profile_read(), "filter_gt_i32")
    alg = rows * 4 + kept * 4
    gbs = alg / (avg * 1e-3) / 1e9
    return {"operator": "FilterGreaterThanInt32, 2^28 INT32 rows, half kept", "rows": rows, "kept": kept,
            "kernel_ms": round(avg, 4), "rows_per_s": round(rows / (avg * 1e-3), 1), "algorithmic_bytes": alg,
            "gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


def bench_stream_pipeline():
    """GpuStreamPipeline (H2D, generated projection kernel, D2H; three batches in flight) run by its own C++
    program (tests/cpp/bench_stream_pipeline, built with the C++ suites) in a child process."""
    import subprocess
    exe = ROOT / "tests" / "cpp" / "_bin" / "bench_stream_pipeline"
    if not exe.exists():
        return {"operator": "stream_pipeline", "error": f"{exe} is not built"}
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=180)
    lines = [l for l in res.stdout.splitlines() if l.startswith("{")]
    if res.returncode != 0 or not lines:
        return {"operator": "stream_pipeline", "error": (res.stderr or res.stdout)[-300:]}
    return json.loads(lines[-1])


def bench_join(torch, capi, dev, steps, nb=1 << 24, nprobe=1 << 28):
    """Join + aggregate: build 2^24 unique keys (permutation) + INT32 payload, probe 2^28 fk (90% hits);
    SELECT SUM(payload), COUNT(*) FROM probe JOIN build ON fk = key."""
    import numpy as np
    build = torch.from_numpy(np.random.RandomState(21).permutation(nb).astype(np.int64)).to(dev)
    payload = torch.empty(nb, dtype=torch.int32, device=dev)
    capi.generate_column(payload, capi.INT32, 0, lo=0, hi=999, seed=21, stream_id=1)
    hit = torch.empty(nprobe, dtype=torch.int64, device=dev)
    miss = torch.empty(nprobe, dtype=torch.int64, device=dev)
    sel = torch.empty(nprobe, dtype=torch.int64, device=dev)
    capi.generate_column(hit, capi.INT64, 0, lo=0, hi=nb - 1, seed=21, stream_id=2)
    capi.generate_column(miss, capi.INT64, 0, lo=nb, hi=2 * nb - 1, seed=21, stream_id=3)
    capi.generate_column(sel, capi.INT64, 0, lo=0, hi=9, seed=21, stream_id=4)
    body = ("((long long*)out)[idx] = (((const long long*)inputs[2])[row] == 0LL) ? "
            "((const long long*)inputs[1])[row] : ((const long long*)inputs[0])[row];")
    outs, _, _, _ = capi.filter_project([hit, miss, sel], [], [(body, [0, 1, 2], capi.INT64)])
    fk = outs[0]
    del hit, miss, sel
    torch.cuda.synchronize()
    aggs = [(capi.SUM, 1, 0, capi.HUGEINT), (capi.COUNT_STAR, -1, -1, capi.INT64)]
    res = capi.join_aggregate(fk, [], build, [payload], aggs)
    capi.profile_reset()
    capi.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        res = capi.join_aggregate(fk, [], build, [payload], aggs)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    capi.profile_enable(False)
    prof = capi.profile_read()
    ms = (t1 - t0) * 1e3 / steps
    kernels = {e["name"]: round(e["ms"] / max(e["launches"], 1) * (e["launches"] / steps), 4) for e in prof}
    alg = sum(e["bytes"] for e in prof) / steps
    compulsory = 12 * nb + 8 * nprobe  # read build key + payload and probe key once
    # The fact-table shape: SUM of an INT32 column of the probe side as well (it travels through the
    # partition with its key).
    fpay = torch.empty(nprobe, dtype=torch.int32, device=dev)
    capi.generate_column(fpay, capi.INT32, 0, lo=-1000, hi=1000, seed=21, stream_id=5)
    aggs2 = [(capi.SUM, 0, 0, capi.HUGEINT), (capi.SUM, 1, 0, capi.HUGEINT), (capi.COUNT_STAR, -1, -1, capi.INT64)]
    res2 = capi.join_aggregate(fk, [fpay], build, [payload], aggs2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        res2 = capi.join_aggregate(fk, [fpay], build, [payload], aggs2)
    torch.cuda.synchronize()
    ms2 = (time.perf_counter() - t0) * 1e3 / steps
    if int(res2[1]) != int(res[0]) or int(res2[2]) != int(res[1]):
        raise RuntimeError("join: the two query shapes disagree on SUM(build payload) / COUNT(*)")
    del fpay
    return {"operator": "join + aggregate (build-side table, key-range partitioned probe), 2^24 build x 2^28 probe", "probe_rows": nprobe, "build_rows": nb,
            "compulsory_bytes": compulsory,
            "compulsory_frac": round(compulsory / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "sum_payload": int(res[0]), "count": int(res[1]), "ms": round(ms, 3),
            "probe_side_sum_ms": round(ms2, 3), "probe_side_sum": int(res2[0]),
            "rows_per_s": round((nprobe + nb) / (ms * 1e-3), 1), "algorithmic_bytes": alg,
            "gbs": round(alg / (ms * 1e-3) / 1e9, 1), "kernel_ms_per_step": kernels}


def bench_execute_plan(torch, capi, dev, reps, n=1 << 26):
    """GpuEngine::ExecutePlan of the config-0 query (SCAN a, b, c -> FILTER -> PROJECTION) from HOST columns:
    the PCIe-inclusive rate of the boundary the extension calls (upload of 20 B/row through the pinned
    staging pool, the fused kernel, download of the kept rows into the result's host buffers). The headline
    and the other operator lines time device-resident data instead."""
    import ctypes as C
    import numpy as np
    sys.path.insert(0, str(ROOT / "tests"))
    from exprs import CFG0_PRED_A, CFG0_PRED_C, CFG0_PROJ_1, CFG0_PROJ_2
    from vectisdb_amd.engine import GpuExpr, GpuOpType as Op, GpuPlanNode, GpuTableRef, GpuValueType as T
    rng = np.random.default_rng(5)
    a = rng.integers(0, 1_000_000, n, dtype=np.int32)
    b = rng.integers(-(1 << 40), 1 << 40, n, dtype=np.int64)
    c = rng.random(n)
    scan = GpuPlanNode(Op.SCAN, table=GpuTableRef("memory", "main", "t"), column_names=["a", "b", "c"])
    filt = GpuPlanNode(Op.FILTER, children=[scan], expressions=[GpuExpr(CFG0_PRED_A, ["a"], T.BOOLEAN),
                                                               GpuExpr(CFG0_PRED_C, ["c"], T.BOOLEAN)])
    proj = GpuPlanNode(Op.PROJECTION, children=[filt], expressions=[GpuExpr(CFG0_PROJ_1, ["a", "b"], T.INT64),
                                                                   GpuExpr(CFG0_PROJ_2, ["c", "a"], T.FLOAT64)])
    plan = proj.to_json().encode()
    hosts = (capi.HostColumn * 3)()
    names = [b"a", b"b", b"c"]
    for i, (arr, vt) in enumerate(((a, T.INT32), (b, T.INT64), (c, T.FLOAT64))):
        hosts[i].name = names[i]
        hosts[i].type = int(vt)
        hosts[i].data = arr.ctypes.data
        hosts[i].validity = None
        hosts[i].rows = n
    lib = capi.load()

    def once():
        h = C.c_void_p()
        capi.check(lib.vgpu_execute_plan(plan, hosts, 3, None, C.byref(h)))
        try:
            if not lib.vgpu_result_success(h):
                raise RuntimeError((lib.vgpu_result_error(h) or b"").decode())
            hc = capi.HostColumn()
            capi.check(lib.vgpu_result_column(h, 0, C.byref(hc)))
            return int(hc.rows)
        finally:
            lib.vgpu_result_free(h)

    kept = once()
    expected = int(((a < 500000) & (c > 0.25)).sum())
    if kept != expected:
        raise RuntimeError(f"execute_plan: {kept} rows kept, expected {expected}")
    t0 = time.perf_counter()
    for _ in range(reps):
        once()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    moved = n * 20 + kept * 16
    return {"operator": f"ExecutePlan config-0 query from host columns, {n} rows (PCIe-inclusive)", "rows": n,
            "kept": kept, "ms": round(ms, 3), "rows_per_s": round(n / (ms * 1e-3), 1),
            "host_bytes_moved": moved, "host_gbs": round(moved / (ms * 1e-3) / 1e9, 1)}


def bench_cross(torch, capi, dev, steps):
    """Cross product + filter: 4,096 x 65,536 INT32 pairs WHERE x + y < 2^30, materialized (x, y)."""
    sys.path.insert(0, str(ROOT / "tests"))
    from exprs import col, copy_body, out
    x = torch.empty(4096, dtype=torch.int32, device=dev)
    y = torch.empty(65536, dtype=torch.int32, device=dev)
    capi.generate_column(x, capi.INT32, 0, lo=0, hi=(1 << 30) - 1, seed=11, stream_id=1)
    capi.generate_column(y, capi.INT32, 0, lo=0, hi=(1 << 30) - 1, seed=11, stream_id=2)
    pred = (out(f"(({col(0, capi.INT32)} + {col(1, capi.INT32)}) < 1073741824)", capi.BOOLEAN), [0, 1],
            capi.BOOLEAN)
    projs = [(copy_body(capi.INT32), [0], capi.INT32), (copy_body(capi.INT32), [1], capi.INT32)]
    capi.cross_filter([x], [y], [pred], projs)
    capi.profile_reset()
    capi.profile_enable(True)
    kept = 0
    for i in range(steps):
        _, kept = capi.cross_filter([x], [y], [pred], projs)
    torch.cuda.synchronize()
    capi.profile_enable(False)
    avg, _ = kernel_stats(capi.profile_read(), "fused_cross_filter")
    pairs = 4096 * 65536
    alg = 4 * (4096 + 65536) + 8 * kept
    gbs = alg / (avg * 1e-3) / 1e9
    return {"operator": "cross product + filter, 4,096 x 65,536", "pairs": pairs, "kept": kept,
            "kernel_ms": round(avg, 4), "rows_per_s": round(pairs / (avg * 1e-3), 1),
            "algorithmic_bytes": alg, "gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}


def make_batched_table(torch, capi, dev, nq=512, seed=7):
    """512 row groups of one table, sizes log-uniform in [2K, 4M] rows; a INT32 uniform [0,1e6), c DOUBLE
    uniform [0,1), v1 INT64 uniform [-1e6,1e6], v2 DOUBLE normal(0,1) as in the plain configs; the group
    count of query g is min(rows, log-uniform in [16, 1M]) and its keys are Zipf(1.1) ranks over that count
    mapped through a seeded permutation of [0, count) (inverse CDF on the device)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(2048), np.log(4 << 20), nq)).astype(np.int64)
    cards = np.minimum(sizes, np.exp(rng.uniform(np.log(16), np.log(1_000_000), nq)).astype(np.int64))
    rows = int(sizes.sum())
    a = torch.empty(rows, dtype=torch.int32, device=dev)
    c = torch.empty(rows, dtype=torch.float64, device=dev)
    capi.generate_column(a, capi.INT32, 0, lo=0, hi=999_999, seed=seed, stream_id=1)
    capi.generate_column(c, capi.FLOAT64, 1, seed=seed, stream_id=2)
    _, v1, v2 = make_groupby_table(torch, capi, dev, rows, seed=seed)
    key = torch.empty(rows, dtype=torch.int64, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    for q in range(nq):
        g = int(cards[q])
        w = torch.arange(1, g + 1, dtype=torch.float64, device=dev).pow_(-1.1)
        cdf = torch.cumsum(w, 0)
        cdf /= cdf[-1].clone()
        u = torch.rand(int(sizes[q]), dtype=torch.float64, device=dev, generator=gen)
        rank = torch.searchsorted(cdf, u).clamp_(max=g - 1)
        perm = torch.randperm(g, device=dev, generator=gen)
        key[offsets[q]:offsets[q + 1]] = perm[rank]
    offsets_t = torch.from_numpy(offsets).to(dev)
    torch.cuda.synchronize()
    return [a, c, key, v1, v2], offsets_t, sizes, cards


def bench_batched(torch, capi, dev, steps, nq=512):
    """Batched form against the per-query path: the same 512 filter + GROUP BY queries (WHERE a < 500000 AND
    c > 0.25, GROUP BY key with COUNT/SUM/MIN/MAX/SUM/AVG) run as one batched pipeline, and one query after
    another. Every query arrives at t=0, so each query's latency in the batch is the batch time, and in the
    sequential loop its completion time."""
    import numpy as np
    cols, offsets, sizes, cards = make_batched_table(torch, capi, dev, nq)
    rows = int(sizes.sum())
    preds = [("((bool*)out)[idx] = (((const int*)inputs[0])[row]) < 500000;", [0], capi.BOOLEAN),
             ("((bool*)out)[idx] = (((const double*)inputs[0])[row]) > 0.25;", [1], capi.BOOLEAN)]
    aggs = [(capi.COUNT_STAR, -1, capi.INT64), (capi.SUM, 0, capi.HUGEINT), (capi.MIN, 0, capi.INT64),
            (capi.MAX, 0, capi.INT64), (capi.SUM, 1, capi.FLOAT64), (capi.AVG, 1, capi.FLOAT64)]
    run = lambda pq: capi.batched_filter_group_by(cols, offsets, preds, 2, [3, 4], aggs, per_query=pq)
    b = run(False)
    p = run(True)
    same = b[3] == p[3] and bool(torch.equal(b[0], p[0])) and bool(torch.equal(b[1], p[1]))
    for i, a in enumerate(aggs):
        if a[2] == capi.FLOAT64:
            same = same and bool(torch.allclose(b[2][i], p[2][i], rtol=1e-12, atol=0))
        else:
            same = same and bool(torch.equal(b[2][i], p[2][i]))
    del b, p
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run(False)
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / steps
    psteps = max(1, steps // 4)
    t0 = time.perf_counter()
    for _ in range(psteps):
        run(True)
    torch.cuda.synchronize()
    tp = (time.perf_counter() - t0) / psteps
    return {"operator": f"batched filter + GROUP BY, {nq} row groups (log-uniform 2K..4M rows, "
                        f"16..1M groups, Zipf(1.1) keys per group), {rows} rows",
            "rows": rows, "queries": nq, "batched_ms": round(tb * 1e3, 3), "per_query_ms": round(tp * 1e3, 3),
            "batched_qps": round(nq / tb, 1), "per_query_qps": round(nq / tp, 1),
            "qps_ratio": round(tp / tb, 2), "rows_per_s": round(rows / tb, 1),
            "results_identical": same}


def make_served_tables(torch, capi, dev, scale=1.0, seed=31):
    """Device-resident tables of the served workload (~40 GB at scale 1): the filter table (a INT32 uniform
    [0,1e6), b INT64 uniform [-2^40,2^40), c DOUBLE uniform [0,1)), the GROUP BY table (key INT64 Zipf(1.1)
    over 1M values, v1 INT64 uniform [-1e6,1e6], v2 DOUBLE normal(0,1)), the probe table (fk INT64, 90%
    uniform over the build keys, 10% uniform over [2^24, 2^25)) and the 2^24-row build side (unique keys, a
    permutation, payload INT32 uniform [0,1000))."""
    import numpy as np
    nf = int(600_000_000 * scale)
    ng = int(600_000_000 * scale)
    nprobe = int(1_600_000_000 * scale)
    nb = 1 << 24
    a = torch.empty(nf, dtype=torch.int32, device=dev)
    b = torch.empty(nf, dtype=torch.int64, device=dev)
    c = torch.empty(nf, dtype=torch.float64, device=dev)
    capi.generate_column(a, capi.INT32, 0, lo=0, hi=999_999, seed=seed, stream_id=1)
    capi.generate_column(b, capi.INT64, 0, lo=-(1 << 40), hi=(1 << 40) - 1, seed=seed, stream_id=2)
    capi.generate_column(c, capi.FLOAT64, 1, seed=seed, stream_id=3)
    key, v1, v2 = make_groupby_table(torch, capi, dev, ng, seed=seed)
    build = torch.from_numpy(np.random.RandomState(seed).permutation(nb).astype(np.int64)).to(dev)
    payload = torch.empty(nb, dtype=torch.int32, device=dev)
    capi.generate_column(payload, capi.INT32, 0, lo=0, hi=999, seed=seed, stream_id=4)
    fk = torch.empty(nprobe, dtype=torch.int64, device=dev)
    chunk = 1 << 27
    body = ("((long long*)out)[idx] = (((const long long*)inputs[2])[row] == 0LL) ? "
            "((const long long*)inputs[1])[row] : ((const long long*)inputs[0])[row];")
    hit = torch.empty(chunk, dtype=torch.int64, device=dev)
    miss = torch.empty(chunk, dtype=torch.int64, device=dev)
    sel = torch.empty(chunk, dtype=torch.int64, device=dev)
    for i, lo in enumerate(range(0, nprobe, chunk)):
        m = min(chunk, nprobe - lo)
        capi.generate_column(hit[:m], capi.INT64, 0, lo=0, hi=nb - 1, seed=seed + i, stream_id=5)
        capi.generate_column(miss[:m], capi.INT64, 0, lo=nb, hi=2 * nb - 1, seed=seed + i, stream_id=6)
        capi.generate_column(sel[:m], capi.INT64, 0, lo=0, hi=9, seed=seed + i, stream_id=7)
        outs, _, _, _ = capi.filter_project([hit[:m], miss[:m], sel[:m]], [], [(body, [0, 1, 2], capi.INT64)])
        fk[lo:lo + m].copy_(outs[0][:m])
        del outs
    del hit, miss, sel
    x = torch.empty(4096, dtype=torch.int32, device=dev)
    y = torch.empty(65536, dtype=torch.int32, device=dev)
    capi.generate_column(x, capi.INT32, 0, lo=0, hi=(1 << 30) - 1, seed=seed, stream_id=8)
    capi.generate_column(y, capi.INT32, 0, lo=0, hi=(1 << 30) - 1, seed=seed, stream_id=9)
    torch.cuda.synchronize()
    tables = {"filter": [a, b, c], "group": [key, v1, v2], "probe": fk, "build": build, "payload": payload,
              "left": [x], "right": [y]}
    total = sum(t.numel() * t.element_size() for t in [a, b, c, key, v1, v2, fk, build, payload, x, y])
    return tables, total


def served_queries(n, rows, seed, lo=10_000, hi=50_000_000, left_rows=4096):
    """n queries drawn uniformly from the four templates (filter/project, GROUP BY, join + aggregate, cross
    product + filter), input sizes log-uniform in [lo, hi] rows, each over a uniformly placed row range of its
    template's table. A cross query's input size counts pairs: it takes size / 4,096 rows of the right side
    against the whole 4,096-row left side."""
    import numpy as np
    rng = np.random.default_rng(seed)
    kinds = rng.integers(0, 4, n)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(np.uint64)
    sizes = np.where(kinds == 3, np.maximum(sizes // left_rows, 1), sizes).astype(np.uint64)
    starts = np.array([rng.integers(0, rows[k] - int(s) + 1) for k, s in zip(kinds, sizes)], dtype=np.uint64)
    return kinds, starts, sizes


def bench_served(torch, capi, dev, nq=400, scale=1.0):
    """Served form against the plain per-query path on the same GPU, same query stream.

    Plain path: one query at a time (batching off, one executor), each query on the per-query operators
    (fused filter kernel; filter + sort-based GROUP BY; sort-merge join of its probe range with the build
    side). Served form: 4 executors on 4 HIP streams, continuous batching, footprint admission, the build
    side reduced once and resident. Each form first runs the stream as a burst (capacity = queries / time);
    the plain path is then run open-loop (Poisson arrivals) at 70% of its capacity, which gives its p99; the
    served form is run open-loop at rising rates, and the highest rate whose p99 stays at or below the
    plain path's p99 is reported beside it."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "tests"))
    from exprs import cfg0_exprs
    tables, total_bytes = make_served_tables(torch, capi, dev, scale)
    rows = [tables["filter"][0].numel(), tables["group"][0].numel(), tables["probe"].numel(),
            tables["right"][0].numel()]
    cross_pred = ("((bool*)out)[idx] = (((const int*)inputs[0])[row] + ((const int*)inputs[1])[row]) < "
                  "1073741824;", [0, 1], capi.BOOLEAN)
    cross_projs = [("((int*)out)[idx] = ((const int*)inputs[0])[row];", [0], capi.INT32),
                   ("((int*)out)[idx] = ((const int*)inputs[0])[row];", [1], capi.INT32)]
    kinds, starts, sizes = served_queries(nq, rows, seed=3)
    preds, projs = cfg0_exprs()
    aggs = [(capi.COUNT_STAR, -1, capi.INT64), (capi.SUM, 0, capi.HUGEINT), (capi.MIN, 0, capi.INT64),
            (capi.MAX, 0, capi.INT64), (capi.SUM, 1, capi.FLOAT64), (capi.AVG, 1, capi.FLOAT64)]

    def server(batching):
        s = capi.Server(slots=4 if batching else 1, batching=batching, max_batch=512,
                        max_batch_rows=int(os.environ.get("VGPU_SERVED_BATCH_ROWS", 256 << 20)))
        t = [s.add_filter_project(tables["filter"], preds, projs),
             s.add_group_by(tables["group"], [], 0, [1, 2], aggs),
             s.add_join_agg(tables["probe"], tables["build"], tables["payload"]),
             s.add_cross(tables["left"], tables["right"], [cross_pred], cross_projs)]
        return s, np.array([t[k] for k in kinds], dtype=np.int32)

    def arrivals(rate, seed):
        return np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, nq))

    def point(srv, tmpl, rate, seed):
        lat, bs, el = srv.run_open_loop(tmpl, starts, sizes, arrivals(rate, seed))
        per = {}
        for k, name in enumerate(("filter", "group_by", "join", "cross")):
            m = kinds == k
            per[name] = [round(float(np.percentile(lat[m], 50)) * 1e3, 3),
                         round(float(np.percentile(lat[m], 99)) * 1e3, 3)]
        return {"offered_qps": round(rate, 1), "achieved_qps": round(nq / el, 1),
                "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                "mean_batch": round(float(bs.mean()), 2), "p50_p99_ms_by_template": per}

    out = {"operator": f"served form, {nq} queries (filter/project, GROUP BY, join + aggregate, cross product "
                       f"templates, 10K..50M rows each) over {total_bytes / 1e9:.1f} GB resident",
           "queries": nq, "resident_bytes": total_bytes}
    plain, ptm = server(False)
    zero = np.zeros(nq)
    # Warm-up: the whole stream once (kernels compiled, the device pool grown to the working set), then
    # the burst that is timed.
    plain.run_open_loop(ptm, starts, sizes, zero)
    _, _, el = plain.run_open_loop(ptm, starts, sizes, zero)
    cap_plain = nq / el
    p_pt = point(plain, ptm, 0.7 * cap_plain, 11)
    plain.close()
    print(f"[bench] served: plain capacity {cap_plain:.1f} QPS, at 70%: {p_pt}", file=sys.stderr, flush=True)
    srv, stm = server(True)
    srv.run_open_loop(stm, starts, sizes, zero)
    _, _, el = srv.run_open_loop(stm, starts, sizes, zero)
    cap_served = nq / el
    print(f"[bench] served: batched capacity {cap_served:.1f} QPS", file=sys.stderr, flush=True)
    ladder = []
    best = None
    for m in (1, 1.5, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20):
        rate = m * p_pt["offered_qps"]
        if rate > 1.2 * cap_served:
            break
        pt = point(srv, stm, rate, 11)
        ladder.append(pt)
        print(f"[bench] served: {pt}", file=sys.stderr, flush=True)
        if pt["p99_ms"] <= p_pt["p99_ms"]:
            best = pt
        else:
            break
    out["stats"] = srv.stats()
    srv.close()
    out.update({"plain_capacity_qps": round(cap_plain, 1), "served_capacity_qps": round(cap_served, 1),
                "capacity_ratio": round(cap_served / cap_plain, 2), "plain_at_70pct": p_pt, "served_ladder": ladder,
                "served_qps_at_plain_p99": best["offered_qps"] if best else None,
                "qps_ratio_at_equal_p99": round(best["offered_qps"] / p_pt["offered_qps"], 2) if best else None})
    return out


def roofline_from_profile(prof, prefer=None):
    """Dominant kernel (most total time) -> roofline dict."""
    if not prof:
        return None, None
    top = max(prof, key=lambda e: e["ms"])
    if prefer:
        cands = [e for e in prof if e["name"] == prefer]
        if cands:
            top = cands[0]
    avg_ms = top["ms"] / max(top["launches"], 1)
    bytes_per_launch = top["bytes"] / max(top["launches"], 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    return top, {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": committed_traffic(top["name"]),
                 "kernel": top["name"], "avg_launch_ms": round(avg_ms, 4),
                 "bytes_per_launch": bytes_per_launch}


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus N > 1 is not supported: GROUP BY "
                                   "would shard by key range with one all-to-all of per-GPU partial "
                                   "aggregates, which is not built"}))
        sys.exit(2)
    import numpy as np
    import torch
    from vectisdb_amd import capi

    capi.check(capi.load().vgpu_device_init(0))
    dev = torch.device("cuda:0")
    rows = args.rows
    copy_gbs, torch_copy_gbs = measure_copy_gbs(torch, dev)
    print(f"[bench] device copy ceiling {copy_gbs:.1f} GB/s (torch copy_ {torch_copy_gbs:.1f} GB/s)",
          file=sys.stderr, flush=True)
    print(f"[bench] generating {rows} rows on device", file=sys.stderr, flush=True)
    key, v1, v2 = make_groupby_table(torch, capi, dev, rows)
    aggs = [(capi.COUNT_STAR, -1, capi.INT64), (capi.SUM, 0, capi.HUGEINT), (capi.MIN, 0, capi.INT64),
            (capi.MAX, 0, capi.INT64), (capi.SUM, 1, capi.FLOAT64), (capi.AVG, 1, capi.FLOAT64)]
    cap = capi.group_by_bound([capi.column(key)], rows)
    groups = 0
    # The sort autotuner picks the radix-pass variant per shape on the first (warm-up) sort of that shape.
    capi.sort_autotune(not args.no_autotune)
    warmup = max(args.warmup, 0 if args.no_autotune else 1)
    for i in range(warmup):
        _, _, groups = capi.group_by([key], [v1, v2], aggs, rows, capacity=cap)
    torch.cuda.synchronize()
    print(f"[bench] warmup done, groups={groups}", file=sys.stderr, flush=True)

    capi.profile_reset()
    capi.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, _, groups = capi.group_by([key], [v1, v2], aggs, rows, capacity=cap)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    capi.profile_enable(False)
    prof = capi.profile_read()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = rows / ((t1 - t0) / args.steps)
    top, roof = roofline_from_profile(prof)
    print(f"[bench] {ms_per_step:.3f} ms/step, {value/1e9:.3f} G rows/s", file=sys.stderr, flush=True)
    for e in sorted(prof, key=lambda e: -e["ms"]):
        avg = e["ms"] / max(e["launches"], 1)
        gbs = (e["bytes"] / max(e["launches"], 1)) / (avg * 1e-3) / 1e9 if avg > 0 else 0
        print(f"[bench]   {e['name']:<20} {avg:8.4f} ms/launch x{e['launches']:<4} {gbs:8.1f} GB/s",
              file=sys.stderr, flush=True)
    # algorithmic bytes for the whole GROUP BY step (sum of every kernel's attributed bytes)
    step_bytes = sum(e["bytes"] for e in prof) / args.steps

    if args.sweep:
        for rb in ("8", "11"):
            os.environ["VGPU_RADIX_BITS"] = rb
            capi.group_by([key], [v1, v2], aggs, rows, capacity=cap)
            capi.profile_reset()
            capi.profile_enable(True)
            torch.cuda.synchronize()
            s0 = time.perf_counter()
            for i in range(3):
                capi.group_by([key], [v1, v2], aggs, rows, capacity=cap)
            torch.cuda.synchronize()
            s1 = time.perf_counter()
            capi.profile_enable(False)
            sp = capi.profile_read()
            print(f"[sweep] radix_bits={rb}: {(s1 - s0) * 1e3 / 3:.3f} ms/step", file=sys.stderr, flush=True)
            for e in sorted(sp, key=lambda e: -e["ms"]):
                avg = e["ms"] / max(e["launches"], 1)
                gbs = (e["bytes"] / max(e["launches"], 1)) / (avg * 1e-3) / 1e9 if avg > 0 else 0
                print(f"[sweep]   {e['name']:<20} {avg:8.4f} ms/launch x{e['launches']:<4} {gbs:8.1f} GB/s",
                      file=sys.stderr, flush=True)
        os.environ.pop("VGPU_RADIX_BITS", None)

    operators = []
    extra = [s for s in args.only.split(",") if s] if args.only else ["filter", "filter_large", "filter_gt",
                                                                       "execute_plan", "stream_pipeline", "cross",
                                                                       "join", "batched", "served"]
    for name in extra:
        try:
            if name == "filter":
                operators.append(bench_filter_project(torch, capi, dev, 50))
            elif name == "filter_large":
                operators.append(bench_filter_project_large(torch, capi, dev, 10))
            elif name == "filter_gt":
                operators.append(bench_filter_gt(torch, capi, dev, 10))
            elif name == "stream_pipeline":
                operators.append(bench_stream_pipeline())
            elif name == "execute_plan":
                operators.append(bench_execute_plan(torch, capi, dev, 3))
            elif name == "cross":
                operators.append(bench_cross(torch, capi, dev, 10))
            elif name == "join":
                operators.append(bench_join(torch, capi, dev, 3))
            elif name == "batched":
                operators.append(bench_batched(torch, capi, dev, 4))
            elif name == "served":
                # served queries span many sort shapes: keep tuning runs out of their latencies
                capi.sort_autotune(False)
                operators.append(bench_served(torch, capi, dev))
            else:
                continue
        except Exception as exc:  # an operator line must not hide the headline number
            operators.append({"operator": name, "error": str(exc)[:300]})
        torch.cuda.empty_cache()
        print(f"[bench] operator {operators[-1]}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        n = min(args.cpu_rows, rows)
        k = key[:n].cpu().numpy()
        a = v1[:n].cpu().numpy()
        b = v2[:n].cpu().numpy()
        oracle.time_group_by_i64(k[:100_000], a[:100_000], b[:100_000], threads)
        sec, cpu_groups = oracle.time_group_by_i64(k, a, b, threads)
        cpu = {"value": round(n / sec, 1), "unit": "rows/s", "cores": threads, "kind": "port",
               "sample": f"{'all' if n == rows else 'first'} {n} rows of the same device-generated table; CPU "
                         f"restatement (not DuckDB, which is not installed): oracle/oracle.c vo_group_by_i64_mt, "
                         f"stable counting sort + exact per-group sums on {threads} threads, {cpu_groups} groups, "
                         f"{sec:.2f} s"}
        print(f"[bench] cpu baseline {cpu['value'] / 1e6:.1f} M rows/s on {threads} threads", file=sys.stderr,
              flush=True)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64+f64",
        "data": "synthetic (seeded device generator: Zipf(1.1) keys over 1M values, uniform int64, normal f64)",
        "config": {"workload": "sort-based GROUP BY, 100M rows, Zipf(1.1) keys over 1M values, 7 aggregates", "rows": rows,
                   "groups": groups, "step_algorithmic_bytes": step_bytes,
                   "step_gbs": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1),
                   # compulsory traffic: read key, v1, v2 once (24 B/row), write each group once (64 B)
                   "compulsory_bytes": 24 * rows + 64 * groups,
                   "compulsory_frac": round((24 * rows + 64 * groups) / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                            4)},
        "roofline": dict(roof, copy_gbs=round(copy_gbs, 1), torch_copy_gbs=round(torch_copy_gbs, 1),
                         frac_of_copy=round(roof["achieved"] / copy_gbs, 4))
        if roof else roof,
        "sort_variants": capi.sort_autotune_report(),
        "cpu_baseline": cpu,
        "operators": operators,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
