"""Benchmark of the onesweep LSD radix sort on one MI355X.

Default workload (the line's `value`): sort_pairs of n = 2^28 uint32 keys
(uniform over [0, 2^32)) with uint32 values = original index, ascending.  A
step = one complete sort of a fresh copy of that input (inputs are resident in
HBM before the timed region; each step sorts its own copy in place).  The same
run also measures, each in an object of the line:
  key_only       the key-only sort: 2^28 u32 keys, with its pass roofline
  hbm_ceiling    hand-written read / write / copy kernels, and a scatter copy with
                 one radix pass's write pattern (the practical ceilings)
  configs        plain 2^20 u32 (mt19937 seed 0) latency, 2^27 u64 low-16 and
                 Zipf, 2^27 f32 with +-0/+-inf, 2^26 u128/u64 pairs
  size_sweep     u32 keys and u32/u32 pairs at n = 2^20, 2^22, ..., 2^30
  byte_skip      2^27 u32 from the reference generator, every other byte constant
  cold_vs_steady 2^27 u32: the process's first sort and cache-flushed sorts
  segmented      100k lognormal segments (~2^28 u32/u32 keys) in one graph
                 replay, against one sort call per segment from host C++
  served         Poisson 5k req/s through the serving loop: p50 / p99 latency
  cpu_baseline   stable sorts of the same 2^28 records on the host cores
Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--workload kv|keys|...] [--log2n 28]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "cuda-fast-sort_amd")
sys.path.insert(0, PKG)
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E peak, GB/s
METRIC = "GKeys/s + % HBM roofline, 2^28 u32; batched-segment keys/s vs per-call"


# Other workload configurations: name -> (description, key kind, key bytes, value bytes, log2 n)
CONFIG_WORKLOADS = {
    "plain20": ("sort u32 n=2^20, keys = raw mt19937(seed 0) outputs", 0, 4, 0, 20),
    "u64low16": ("sort u64 n=2^27, low 16 bits uniform, upper 48 bits zero", 0, 8, 0, 27),
    "zipf": ("sort u64 n=2^27, Zipf(s=1.1) over 2^20 ranks, rank -> fmix64 hash", 0, 8, 0, 27),
    "f32": ("sort f32 n=2^27, N(0,1) with 1% +-0.0 and 0.1% +-inf", 2, 4, 0, 27),
    "u128kv": ("sort_pairs u128/u64 n=2^26, hi uniform over 2^20 values, lo uniform over 2^64, values = index",
               0, 16, 8, 26),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["kv", "keys", "copy", "segments", "serve", "sweep", *CONFIG_WORKLOADS],
                    default="kv",
                    help="kv: sort_pairs u32/u32 (the full line); keys: sort u32; copy: the streaming kernels "
                         "(bandwidth ceiling, and the known-size kernel that calibrates HBM counters); "
                         "segments: 100k lognormal-sized u32/u32 segments, graph replay vs one call per segment; "
                         "serve: Poisson request stream through the serving loop; "
                         "sweep: u32 keys and pairs at n = 2^20..2^30")
    ap.add_argument("--segments", type=int, default=100_000)
    ap.add_argument("--per-call-sample", type=int, default=20_000,
                    help="segments sorted one rsort_sort call each for the per-call comparison")
    ap.add_argument("--rate", type=float, default=5000.0, help="serve: Poisson arrival rate, requests/s")
    ap.add_argument("--duration", type=float, default=1.0, help="serve: seconds of arrivals")
    ap.add_argument("--window-us", type=int, default=0,
                    help="serve: batching window (0: launch as soon as a batch slot is free)")
    ap.add_argument("--serve-mode", choices=["direct", "graph", "batched"], default="direct")
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--descending", action="store_true")
    ap.add_argument("--graph", action="store_true", help="sort with RSORT_FLAG_GRAPH (replayed hipGraph)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="kv: only the headline sort (no side measurements)")
    ap.add_argument("--cpu-log2n", type=int, default=28, help="CPU baseline sample size")
    return ap.parse_args()


def gen_u32(n, device, seed):
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    hi = torch.randint(0, 1 << 16, (n,), dtype=torch.int32, device=device, generator=g)
    lo = torch.randint(0, 1 << 16, (n,), dtype=torch.int32, device=device, generator=g)
    return (hi << 16) | lo  # uniform 32-bit patterns


def timed_sorts(rs, desc, ks, vs, ws, tb, n, warmup, steps, profiled):
    """warmup untimed sorts, then `steps` plain sorts back to back between two
    synchronizations (wall ms per sort; device ms per sort from events between
    the calls), then `profiled` sorts with per-stage events for the stage
    breakdown.  Each sort takes its own fresh copy of the input."""
    import torch

    stream = torch.cuda.current_stream()
    val = lambda i: vs[i].data_ptr() if vs[i] is not None else None
    for i in range(warmup):
        rs.sort_raw(desc, ks[i].data_ptr(), val(i), ws.data_ptr(), tb, n, stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for j in range(steps):
        i = warmup + j
        rs.sort_raw(desc, ks[i].data_ptr(), val(i), ws.data_ptr(), tb, n, stream)
        ev[j + 1].record(stream)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    dev_ms = [ev[j].elapsed_time(ev[j + 1]) for j in range(steps)]
    stages = [rs.sort_profiled(desc, ks[i].data_ptr(), val(i), ws.data_ptr(), tb, n, stream)
              for i in range(warmup + steps, warmup + steps + profiled)]
    torch.cuda.synchronize()
    return wall_ms, dev_ms, stages


def pass_roofline(stages, kb, vb, n, passes):
    """The pass kernel's roofline from per-stage events: algorithmic bytes per
    launch 2*n*(kb+vb) over the mean duration of the executed passes."""
    pass_ms = [sum(s[2 + p] for s in stages) / len(stages) for p in range(passes)]
    longest = max(pass_ms)
    executed = [p for p in range(passes) if pass_ms[p] > 0.05 * longest]
    avg = sum(pass_ms[p] for p in executed) / len(executed)
    pass_bytes = 2.0 * n * (kb + vb)
    achieved = pass_bytes / (avg * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": "k_onesweep (one executed LSD pass)", "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
            "bytes_per_launch": pass_bytes, "avg_launch_ms": round(avg, 5)}, executed


def run_sort_bench(rs, n, with_values, descending, steps, warmup, seed=1, flags=0):
    """Sorts warmup + steps + profiled fresh copies of 2^k u32 keys (and
    index values); returns wall ms per timed sort, device ms per timed sort,
    the stage ms of the profiled sorts, and the device checks of the last
    result (order, digit histograms, value pairing)."""
    import torch

    dev = torch.device("cuda:0")
    keys0 = gen_u32(n, dev, seed)
    vals0 = torch.arange(n, dtype=torch.int32, device=dev) if with_values else None
    vb = 4 if with_values else 0
    desc = rs.make_desc(rs.KEY_UNSIGNED, 4, vb, descending, flags)
    tb = rs.sort_raw(desc, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device=dev)
    before = rs.digit_counts(keys0, desc, n)
    profiled = min(steps, 5)
    total = warmup + steps + profiled
    ks = [keys0.clone() for _ in range(total)]
    vs = [vals0.clone() for _ in range(total)] if with_values else [None] * total
    torch.cuda.synchronize()
    wall_ms, dev_ms, stages = timed_sorts(rs, desc, ks, vs, ws, tb, n, warmup, steps, profiled)
    ok = rs.check_sorted(ks[-1], desc, n) == 0 and torch.equal(rs.digit_counts(ks[-1], desc, n), before)
    if with_values:
        ok = ok and rs.check_pairings(keys0, ks[-1], vs[-1], 4, 4, n) == 0
    del ks, vs
    torch.cuda.empty_cache()
    return wall_ms, dev_ms, stages, bool(ok), keys0


def _fmix64(x):
    """MurmurHash3's 64-bit finaliser on an int64 tensor (a bijection of 64-bit words)."""
    m1 = 0xFF51AFD7ED558CCD - (1 << 64)
    m2 = 0xC4CEB9FE1A85EC53 - (1 << 64)
    for mul in (m1, m2, None):
        x = x ^ ((x >> 33) & ((1 << 31) - 1))
        if mul is not None:
            x = x * mul
    return x


def gen_config(name, n, dev, seed):
    """Keys (and values) of one workload configuration on the device: a uint8
    tensor of n * key_bytes bytes, and an index tensor or None."""
    import numpy as np
    import torch

    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    if name == "plain20":
        # numpy's legacy MT19937 with 32-bit draws gives the raw mt19937 outputs
        keys = np.random.RandomState(0).randint(0, 2**32, size=n, dtype=np.uint32)
        return torch.from_numpy(keys.view(np.uint8)).to(dev), None
    if name == "u64low16":
        k = torch.randint(0, 1 << 16, (n,), dtype=torch.int64, device=dev, generator=g)
        return k.view(torch.uint8), None
    if name == "zipf":
        ranks = 1 << 20
        w = torch.arange(1, ranks + 1, dtype=torch.float64, device=dev).pow(-1.1)
        cdf = torch.cumsum(w, 0)
        cdf /= cdf[-1].clone()
        u = torch.rand(n, dtype=torch.float64, device=dev, generator=g)
        r = torch.searchsorted(cdf, u, right=True).clamp_(max=ranks - 1)
        return _fmix64(r + 1).view(torch.uint8), None
    if name == "f32":
        x = torch.randn(n, dtype=torch.float32, device=dev, generator=g)
        u = torch.rand(n, dtype=torch.float32, device=dev, generator=g)
        sign = torch.rand(n, dtype=torch.float32, device=dev, generator=g) < 0.5
        x = torch.where(u < 0.011, torch.where(sign, -0.0, 0.0), x)
        x = torch.where(u < 0.001, torch.where(sign, float("-inf"), float("inf")), x)
        return x.view(torch.uint8), None
    if name == "u128kv":
        kw = torch.empty((n, 2), dtype=torch.int64, device=dev)
        lo = torch.randint(-(2**63), 2**63 - 1, (n,), dtype=torch.int64, device=dev, generator=g)
        kw[:, 0] = lo
        kw[:, 1] = torch.randint(0, 1 << 20, (n,), dtype=torch.int64, device=dev, generator=g)
        return kw.view(torch.uint8).reshape(-1), torch.arange(n, dtype=torch.int64, device=dev)
    raise ValueError(name)


def bench_config(rs, name, steps, warmup, descending=False, flags=0):
    """One workload configuration: GKeys/s, the formula roofline (every pass
    counted, also those the device skips), the pass roofline, device checks
    (order, digit histograms, value pairing)."""
    import torch

    desc_s, kind, kb, vb, log2n = CONFIG_WORKLOADS[name]
    n = 1 << log2n
    dev = torch.device("cuda:0")
    keys0, vals0 = gen_config(name, n, dev, 21)
    d = rs.make_desc(kind, kb, vb, descending, flags)
    tb = rs.sort_raw(d, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device=dev)
    before = rs.digit_counts(keys0, d, n)
    profiled = min(steps, 5)
    total = warmup + steps + profiled
    ks = [keys0.clone() for _ in range(total)]
    vs = [vals0.clone() for _ in range(total)] if vb else [None] * total
    torch.cuda.synchronize()
    wall_ms, dev_steps, stages = timed_sorts(rs, d, ks, vs, ws, tb, n, warmup, steps, profiled)
    ok = rs.check_sorted(ks[-1], d, n) == 0 and torch.equal(rs.digit_counts(ks[-1], d, n), before)
    if vb:
        ok = ok and rs.check_pairings(keys0, ks[-1], vs[-1], kb, vb, n) == 0
    del ks, vs
    torch.cuda.empty_cache()
    dev_ms = sum(dev_steps) / len(dev_steps)
    roof, executed = pass_roofline(stages, kb, vb, n, kb)
    formula_bytes = (1 + kb) * 2.0 * n * (kb + vb)
    return {
        "description": desc_s, "n": n, "gkeys_per_s": round(n / (dev_ms * 1e-3) / 1e9, 3),
        "ms_per_sort": round(dev_ms, 4), "us_per_sort": round(dev_ms * 1e3, 2), "wall_ms_per_sort": round(wall_ms, 4),
        "roofline": roof,
        "formula_frac": round(formula_bytes / (dev_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
        "executed_passes": executed, "checked": bool(ok),
    }


def bench_plain20(rs, steps, warmup):
    """Plain 2^20 u32 keys (raw mt19937 seed 0): microseconds per sort,
    direct calls and graph replays (RSORT_FLAG_GRAPH), each call timed alone
    with events around it (the reference's method) and back to back."""
    import torch

    n = 1 << 20
    dev = torch.device("cuda:0")
    keys0, _ = gen_config("plain20", n, dev, 0)
    out = {"description": CONFIG_WORKLOADS["plain20"][0], "n": n}
    stream = torch.cuda.current_stream()
    for label, flags in (("direct", 0), ("graph", rs.FLAG_GRAPH)):
        d = rs.make_desc(0, 4, 0, False, flags)
        tb = rs.sort_raw(d, None, None, None, 0, n)
        ws = torch.empty(tb, dtype=torch.uint8, device=dev)
        k = keys0.clone()
        for _ in range(warmup + 2):
            k.copy_(keys0)
            rs.sort_raw(d, k.data_ptr(), None, ws.data_ptr(), tb, n, stream)
        torch.cuda.synchronize()
        alone = []
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for _ in range(max(steps, 10)):
            k.copy_(keys0)
            torch.cuda.synchronize()
            e0.record(stream)
            rs.sort_raw(d, k.data_ptr(), None, ws.data_ptr(), tb, n, stream)
            e1.record(stream)
            torch.cuda.synchronize()
            alone.append(e0.elapsed_time(e1) * 1e3)
        ok = rs.check_sorted(k, d, n) == 0 and torch.equal(rs.digit_counts(k, d, n), rs.digit_counts(keys0, d, n))
        # back to back: K sorts of separate copies between two events
        copies = [keys0.clone() for _ in range(max(steps, 10))]
        torch.cuda.synchronize()
        e0.record(stream)
        for c in copies:
            rs.sort_raw(d, c.data_ptr(), None, ws.data_ptr(), tb, n, stream)
        e1.record(stream)
        torch.cuda.synchronize()
        b2b = e0.elapsed_time(e1) * 1e3 / len(copies)
        alone.sort()
        out[label] = {"us_per_sort_alone_median": round(alone[len(alone) // 2], 2),
                      "us_per_sort_back_to_back": round(b2b, 2), "checked": bool(ok)}
    out["us_per_sort"] = out["direct"]["us_per_sort_alone_median"]
    return out


SWEEP_LOG2N = (20, 22, 24, 26, 28, 30)


def bench_size_sweep(rs, sizes=SWEEP_LOG2N):
    """Throughput over n = 2^20..2^30, u32 keys and u32/u32 pairs (index
    values), uniform keys: each call timed alone with events around it (the
    input restored from a copy outside the events), median over the calls;
    GKeys/s and the formula roofline (1 + passes) * 2 * n * (key + value
    bytes) over 8 TB/s.  Every size's last result is checked on the device
    (order, digit histograms, value pairing)."""
    import torch

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    out = {"description": "u32 keys uniform over [0, 2^32), values = index; whole call, device time, median of "
                          "calls timed alone", "keys": {}, "pairs": {}}
    for label, vb in (("keys", 0), ("pairs", 4)):
        for log2n in sizes:
            n = 1 << log2n
            keys0 = gen_u32(n, dev, log2n)
            vals0 = torch.arange(n, dtype=torch.int32, device=dev) if vb else None
            d = rs.make_desc(0, 4, vb)
            tb = rs.sort_raw(d, None, None, None, 0, n)
            ws = torch.empty(tb, dtype=torch.uint8, device=dev)
            k = keys0.clone()
            v = vals0.clone() if vb else None
            calls = max(3, min(20, (1 << 26) // n))
            times = []
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            for c in range(calls + 2):  # two warm-up calls
                k.copy_(keys0)
                if vb:
                    v.copy_(vals0)
                torch.cuda.synchronize()
                e0.record(stream)
                rs.sort_raw(d, k.data_ptr(), v.data_ptr() if vb else None, ws.data_ptr(), tb, n, stream)
                e1.record(stream)
                torch.cuda.synchronize()
                if c >= 2:
                    times.append(e0.elapsed_time(e1))
            ok = rs.check_sorted(k, d, n) == 0 and torch.equal(rs.digit_counts(k, d, n), rs.digit_counts(keys0, d, n))
            if vb:
                ok = ok and rs.check_pairings(keys0, k, v, 4, 4, n) == 0
            times.sort()
            ms = times[len(times) // 2]
            formula = 5 * 2.0 * n * (4 + vb)
            out[label][f"2^{log2n}"] = {"gkeys_per_s": round(n / (ms * 1e-3) / 1e9, 3), "us_per_sort": round(ms * 1e3, 2),
                                        "formula_frac": round(formula / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                                        "calls": calls, "checked": bool(ok)}
            del keys0, vals0, k, v, ws
            torch.cuda.empty_cache()
    return out


def bench_init_keys_mode(rs, mode, n, steps, warmup):
    """u32 keys from the reference generator (mode 1 random, 2 byte_skip),
    steady state: GKeys/s and the passes the device skipped."""
    import torch

    dev = torch.device("cuda:0")
    keys0 = torch.empty(n, dtype=torch.int32, device=dev)
    rs.init_keys(keys0, 4, mode, rs.set_seed_radix(0), n)
    d = rs.make_desc(0, 4, 0)
    tb = rs.sort_raw(d, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device=dev)
    before = rs.digit_counts(keys0, d, n)
    profiled = 2
    ks = [keys0.clone() for _ in range(warmup + steps + profiled)]
    torch.cuda.synchronize()
    wall_ms, dev_steps, stages = timed_sorts(rs, d, ks, [None] * len(ks), ws, tb, n, warmup, steps, profiled)
    ok = rs.check_sorted(ks[-1], d, n) == 0 and torch.equal(rs.digit_counts(ks[-1], d, n), before)
    ms = sum(dev_steps) / len(dev_steps)
    _, executed = pass_roofline(stages, 4, 0, n, 4)
    del ks
    torch.cuda.empty_cache()
    return {"gkeys_per_s": round(n / (ms * 1e-3) / 1e9, 3), "ms_per_sort": round(ms, 4),
            "executed_passes": executed, "checked": bool(ok)}


def bench_cold(rs):
    """The process's first sort (no warm-up of any kind): 2^27 u32 keys from
    the reference generator, timed with events around the call (device) and
    the host clock (first-use costs included); then five sorts each after the
    caches are flushed (a 1 GiB write) and the GPU idles for 50 ms."""
    import torch

    n = 1 << 27
    dev = torch.device("cuda:0")
    keys = torch.empty(n, dtype=torch.int32, device=dev)
    rs.init_keys(keys, 4, 1, rs.set_seed_radix(0), n)
    keys0 = keys.clone()
    d = rs.make_desc(0, 4, 0)
    tb = rs.sort_raw(d, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device=dev)
    flush = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    rs.sort_raw(d, keys.data_ptr(), None, ws.data_ptr(), tb, n, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    first_wall = (time.perf_counter() - t0) * 1e3
    first_dev = e0.elapsed_time(e1)
    ok = rs.check_sorted(keys, d, n) == 0
    flushed = []
    for _ in range(5):
        keys.copy_(keys0)
        flush.fill_(1)
        torch.cuda.synchronize()
        time.sleep(0.05)
        e0.record(stream)
        rs.sort_raw(d, keys.data_ptr(), None, ws.data_ptr(), tb, n, stream)
        e1.record(stream)
        torch.cuda.synchronize()
        flushed.append(e0.elapsed_time(e1))
    del flush
    torch.cuda.empty_cache()
    fl = sum(flushed) / len(flushed)
    return {"n": n, "first_call_device_ms": round(first_dev, 4), "first_call_wall_ms": round(first_wall, 3),
            "first_call_gkeys_per_s": round(n / (first_dev * 1e-3) / 1e9, 3),
            "cache_flushed_ms": round(fl, 4), "cache_flushed_gkeys_per_s": round(n / (fl * 1e-3) / 1e9, 3),
            "checked": bool(ok)}


STREAM_FORMS = {0: "strided", 4: "strided, non-temporal", 8: "chunked", 12: "chunked, non-temporal"}


def stream_ceiling(rs, nbytes=1 << 31, reps=5):
    """Hand-written streaming kernels (rsort_stream_bandwidth): GB/s of a read,
    a write, a copy and a scatter copy (the write pattern of one radix pass:
    64 KiB tiles of 4-byte keys written as 256 runs of 64 keys to 256
    regions) of nbytes, best over the access forms and workgroups per CU in
    {2, 4, 8, 16, 32}."""
    import torch

    dev = torch.device("cuda:0")
    src = torch.ones(nbytes, dtype=torch.uint8, device=dev)
    dst = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    sink = torch.empty(4096, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    out = {"bytes": nbytes}
    for name, op, moved in (("read", 0, nbytes), ("write", 1, nbytes), ("copy", 2, 2 * nbytes),
                            ("scatter", 3, 2 * nbytes)):
        best = (0.0, 0, 0)
        for form in (STREAM_FORMS if op != 3 else (0, 4)):
            for bpc in (2, 4, 8, 16, 32):
                dptr = sink.data_ptr() if op == 0 else dst.data_ptr()
                rs.stream_bandwidth(op | form, dptr, src.data_ptr(), nbytes, bpc, stream)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(reps):
                    rs.stream_bandwidth(op | form, dptr, src.data_ptr(), nbytes, bpc, stream)
                e1.record(stream)
                torch.cuda.synchronize()
                gbs = moved * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
                if gbs > best[0]:
                    best = (gbs, form, bpc)
                if op == 3 and form == 0:  # plain stores, as the sort's scatter uses
                    out["scatter_plain_gbs"] = round(max(out.get("scatter_plain_gbs", 0.0), gbs), 1)
        out[name + "_gbs"] = round(best[0], 1)
        out[name + "_form"] = (f"{STREAM_FORMS[best[1]]}, {best[2]} workgroups per CU" if op != 3 else
                               f"{'non-temporal' if best[1] else 'plain'}, {best[2]} workgroups per CU")
    del src, dst
    torch.cuda.empty_cache()
    return out


# The sources a plain sort is built from: its kernels, their launch and the
# host dispatch (not the segmented sort, the server, the device checkers or
# the streaming kernels).
SORT_SOURCES = ("rsort_host.hip", "rsort_kernels.hpp", "rsort_launch.hpp", "rsort_launch_impl.hpp",
                "rsort_onesweep_tile.inc", "rsort_protocol.hpp", "rsort_tuning.hpp")


def kernel_sources_sha256():
    """Hash of the sources a plain sort is built from (SORT_SOURCES and the
    sort_kb*.hip instantiation units under csrc/): a traffic figure measured
    on other sources is not reported."""
    h = hashlib.sha256()
    csrc = os.path.join(PKG, "csrc")
    for f in sorted(os.listdir(csrc)):
        if f in SORT_SOURCES or (f.startswith("sort_kb") and f.endswith(".hip")):
            with open(os.path.join(csrc, f), "rb") as fh:
                h.update(f.encode() + b"\0" + fh.read())
    return h.hexdigest()


def measured_traffic(profile, kernel_prefix, n):
    """HBM bytes per launch of the pass kernel from a committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE, calibrated against a copy of known
    size), when it was taken on this workload and on these very sources."""
    path = os.path.join(ROOT, "profiles", profile)
    try:
        text = open(path).read()
        block = text[text.index("```json") + 7:]
        d = json.loads(block[:block.index("```")])
    except (OSError, ValueError):
        return None
    if d.get("n") == n and d.get("kernel", "").startswith(kernel_prefix) and \
            d.get("source_sha256") == kernel_sources_sha256():
        return d["traffic_bytes"]
    return None


def lognormal_sizes(count, seed=5, median=2048.0, sigma=0.7356, lo=32, hi=1 << 22):
    """Request / segment sizes: lognormal with median 2048 and sigma 0.7356
    (mean ~2684 keys), rounded and clipped to [32, 2^22]."""
    import numpy as np

    rng = np.random.Generator(np.random.MT19937(seed))
    v = np.floor(median * np.exp(sigma * rng.standard_normal(count)) + 0.5)
    return np.clip(v, lo, hi).astype(np.int64)


def bench_segments(rs, a, steps, warmup):
    """Segmented batch: one segmented sort of 100k lognormal segments replayed
    from its graph (device time per replay, device checks), and from host C++
    (cli/seg_percall) the same batch's graph replay against one rsort_sort
    call per segment over the first per_call_sample segments."""
    import numpy as np
    import torch

    dev = torch.device("cuda:0")
    sizes = lognormal_sizes(a.segments)
    n = int(sizes.sum())
    offs_np = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    offsets = torch.from_numpy(offs_np).to(dev)
    keys0 = gen_u32(n, dev, 7)
    vals0 = torch.arange(n, dtype=torch.int32, device=dev)
    desc = rs.make_desc(rs.KEY_UNSIGNED, 4, 4, False, rs.FLAG_GRAPH)
    tb = rs.segmented_sort_raw(desc, None, None, n, None, 8, a.segments, None, 0)
    ws = torch.empty(tb, dtype=torch.uint8, device=dev)
    k, v = keys0.clone(), vals0.clone()
    stream = torch.cuda.current_stream()
    call = lambda: rs.segmented_sort_raw(desc, k.data_ptr(), v.data_ptr(), n, offsets.data_ptr(), 8, a.segments,
                                         ws.data_ptr(), tb, stream)
    for _ in range(warmup):
        k.copy_(keys0)
        v.copy_(vals0)
        call()
    torch.cuda.synchronize()
    ms = []
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(steps):
        k.copy_(keys0)
        v.copy_(vals0)
        e0.record(stream)
        call()
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    dev_ms = sum(ms) / len(ms)
    # every segment sorted, values point at their keys and stay in their segment
    ku = k ^ (-(2**31))
    seg_start = torch.zeros(n, dtype=torch.bool, device=dev)
    seg_start[offsets[1:-1]] = True
    ok = bool(((ku[1:] >= ku[:-1]) | seg_start[1:]).all().item()) and \
        rs.check_pairings(keys0, k, v, 4, 4, n) == 0
    seg_of = torch.repeat_interleave(torch.arange(a.segments, dtype=torch.int32, device=dev),
                                     torch.from_numpy(sizes).to(dev))
    ok = ok and bool((seg_of[v.long()] == seg_of).all().item())
    lds_cap = rs.segmented_plan(desc, n, 8, a.segments).bin_capacity[-1]
    small = int(sizes[sizes <= lds_cap].sum())
    big = n - small
    bytes_ = 2.0 * small * 8 + 5 * 2.0 * big * 8  # LDS segments: one read + one write; long: the formula
    del k, v, ku, seg_start, seg_of, ws
    torch.cuda.empty_cache()
    batched = n / (dev_ms * 1e-3) / 1e9
    res = {
        "description": f"{a.segments} segments of u32/u32, lognormal sizes median 2048 sigma 0.7356 clipped "
                       f"[32, 2^22], keys uniform; one segmented sort replayed from its hipGraph",
        "n": n, "segments": a.segments, "gkeys_per_s": round(batched, 3), "ms_per_batch": round(dev_ms, 4),
        "keys_in_lds_segments": small, "keys_in_long_segments": big,
        "roofline": {"bound": "hbm", "kernel": "whole segmented sort (device time)",
                     "achieved": round(bytes_ / (dev_ms * 1e-3) / 1e9, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(bytes_ / (dev_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4), "bytes": bytes_},
        "checked": ok,
    }
    exe = os.path.join(PKG, "cli", "seg_percall")
    with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as f:
        f.write(sizes.astype("<u8").tobytes())
        path = f.name
    try:
        r = subprocess.run([exe, path, "--calls", str(min(a.per_call_sample, a.segments)), "--replays", "10"],
                           capture_output=True, text=True, timeout=300)
        cpp = json.loads(r.stdout.strip().splitlines()[-1]) if r.returncode == 0 else {"error": r.stderr[-500:]}
    finally:
        os.unlink(path)
    if "per_call" in cpp:
        res["per_call"] = dict(cpp["per_call"], note="one rsort_sort (sort_pairs u32/u32) per segment from host C++ "
                                                     "over the C ABI, back to back on one stream")
        res["cpp_batched"] = cpp["batched"]
        res["batched_over_per_call"] = cpp["batched_over_per_call"]
        res["checked"] = res["checked"] and cpp["batched"]["violations"] == 0 and cpp["per_call"]["violations"] == 0
    else:
        res["per_call"] = cpp
    return res


def _submit_raw(rs, srv, kptr, vptr, n):
    import ctypes

    rid = ctypes.c_uint64(0)
    rs._check(rs.lib().rsort_serve_submit(srv._h, kptr, vptr, n, ctypes.byref(rid)))
    return int(rid.value)


def bench_serve(rs, a):
    """Served stream: Poisson arrivals through the serving loop, latency
    percentiles and throughput; then the same requests submitted at once
    (saturation)."""
    import numpy as np
    import torch

    dev = torch.device("cuda:0")
    count = max(1, int(a.rate * a.duration))
    sizes = lognormal_sizes(count, seed=11)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(offs[-1])
    keys = gen_u32(n, dev, 13)
    vals = torch.arange(n, dtype=torch.int32, device=dev)
    keys_ref = keys.clone()
    kp, vp = keys.data_ptr(), vals.data_ptr()
    torch.cuda.synchronize()
    srv = rs.Server(rs.KeyType(rs.KEY_UNSIGNED, 4), value_bytes=4, batch_keys=1 << 24, batch_requests=8192,
                    window_us=a.window_us, depth=3, mode=a.serve_mode)
    wk, wv = keys_ref.clone(), vals.clone()
    warm = [_submit_raw(rs, srv, wk.data_ptr() + 4 * int(offs[i]), wv.data_ptr() + 4 * int(offs[i]), int(sizes[i]))
            for i in range(min(50, count))]
    for w in warm:
        srv.wait(w)
    del wk, wv
    torch.cuda.synchronize()
    st0 = srv.stats()
    rng = np.random.default_rng(17)
    gaps = rng.exponential(1.0 / a.rate, size=count)
    ids = []
    t_start = time.perf_counter()
    t_next = t_start
    for i in range(count):
        t_next += gaps[i]
        while time.perf_counter() < t_next:
            pass
        o = int(offs[i])
        ids.append(_submit_raw(rs, srv, kp + 4 * o, vp + 4 * o, int(sizes[i])))
    infos = [srv.wait(i) for i in ids]
    t_end = time.perf_counter()
    lat = np.array([x.latency_us for x in infos], dtype=np.float64)
    q = np.array([x.queue_us for x in infos], dtype=np.float64)
    st = srv.stats()
    k = keys.view(torch.int32) ^ (-(2**31))
    starts = torch.zeros(n, dtype=torch.bool, device=dev)
    starts[torch.from_numpy(offs[1:-1]).to(dev)] = True
    ok = bool(((k[1:] >= k[:-1]) | starts[1:]).all().item()) and rs.check_pairings(keys_ref, keys, vals, 4, 4, n) == 0
    kptrs = [kp + 4 * int(offs[i]) for i in range(count)]
    vptrs = [vp + 4 * int(offs[i]) for i in range(count)]
    # all submitted at once, three times (median): the server finished the
    # stream when its last request completed (latency = submit -> completion,
    # timed by the server)
    t_srvs, t_sats, sat_batches = [], [], []
    for _ in range(3):
        keys.copy_(keys_ref)
        vals.copy_(torch.arange(n, dtype=torch.int32, device=dev))
        torch.cuda.synchronize()
        st1 = srv.stats()
        t0 = time.perf_counter()
        ids2 = srv.submit_many(kptrs, vptrs, [int(x) for x in sizes])
        infos2 = srv.wait_many(ids2)
        t_sats.append(time.perf_counter() - t0)
        t_srvs.append(max(x.latency_us for x in infos2) * 1e-6)
        sat_batches.append(int(srv.stats().batches - st1.batches))
    t_srv, t_sat = float(np.median(t_srvs)), float(np.median(t_sats))
    srv.close()
    return {
        "description": f"sort_pairs u32/u32 requests, Poisson {a.rate:g} req/s for {a.duration:g} s, sizes lognormal "
                       "median 2048 sigma 0.7356 clipped [32, 2^22]",
        "requests": count, "keys": n, "mode": a.serve_mode, "depth": 3, "window_us": a.window_us,
        "latency_us": {"p50": round(float(np.percentile(lat, 50)), 1), "p90": round(float(np.percentile(lat, 90)), 1),
                       "p99": round(float(np.percentile(lat, 99)), 1), "max": round(float(lat.max()), 1),
                       "queue_p50": round(float(np.percentile(q, 50)), 1)},
        "offered_gkeys_per_s": round(n / a.duration / 1e9, 5),
        "achieved_gkeys_per_s": round(n / (t_end - t_start) / 1e9, 5),
        "batches": int(st.batches - st0.batches),
        "saturation_gkeys_per_s": round(n / t_srv / 1e9, 3),
        "saturation_note": "every request submitted at once (submit_many), median of 3; keys / the server's "
                           "submit-to-last-completion time; saturation_wall_gkeys_per_s: keys / the Python "
                           "caller's wall time from submit_many to the return of wait_many",
        "saturation_wall_gkeys_per_s": round(n / t_sat / 1e9, 3), "saturation_batches": max(sat_batches),
        "checked": ok,
    }


def host_cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(keys0, with_values, log2n):
    """Stable sorts of the same (key, index) records on the host cores:
    std::stable_sort on one core, and the oracle's multi-threaded stable sort
    (sorted runs merged in parallel) on every core; std::sort(par_unseq) too."""
    import numpy as np

    from oracle import oracle

    m = 1 << log2n
    sample_k = keys0[:m].cpu().numpy().view("uint32")
    sample_v = np.arange(m, dtype=np.uint32) if with_values else None
    cores = os.cpu_count() or 1
    t_par = oracle.time_stable_sort_par_u32(sample_k, sample_v, cores, 3)
    t_one = oracle.time_stable_sort_u32(sample_k, sample_v, 1)
    what = "(u32 key, u32 value) records" if with_values else "u32 keys"
    out = {
        "value": round(m / t_par / 1e9, 5), "unit": "GKeys/s", "cores": cores, "kind": "port",
        "sample": f"multi-threaded stable sort (runs std::stable_sort-ed and merged in parallel) of 2^{log2n} {what}, "
                  "the same input as the GPU sort, median of 3 runs",
        "single_core": {"value": round(m / t_one / 1e9, 5), "unit": "GKeys/s", "cores": 1,
                        "sample": f"std::stable_sort of the same 2^{log2n} {what}, one run"},
        "host_cores_available": cores,
        "host_cpu": host_cpu_model(),
        "pstl_backend": "parallel (TBB)" if oracle.pstl_parallel_backend() else "serial (libstdc++ without TBB)",
    }
    if oracle.pstl_parallel_backend():
        t_pu = oracle.time_sort_par_u32(sample_k, sample_v, 1)
        out["par_unseq"] = {"value": round(m / t_pu / 1e9, 5), "unit": "GKeys/s", "cores": cores,
                            "note": "std::sort(std::execution::par_unseq), not stable"}
    else:
        out["par_unseq"] = {"note": "libstdc++ without TBB runs std::sort(par_unseq) serially: not timed"}
    return out


def headline(rs, a, n, kv):
    kb, vb, passes = 4, (4 if kv else 0), 4
    flags = rs.FLAG_GRAPH if a.graph else 0
    ms, dev_steps, stages, ok, keys0 = run_sort_bench(rs, n, kv, a.descending, a.steps, a.warmup, flags=flags)
    sort_dev_ms = sum(dev_steps) / len(dev_steps)
    roof, _ = pass_roofline(stages, kb, vb, n, passes)
    roof["traffic"] = measured_traffic("onesweep_kv_u32_2p28_hbm_traffic.md" if kv else
                                       "onesweep_keys_u32_2p28_hbm_traffic.md",
                                       f"k_onesweep<4,{vb},", n)
    formula_bytes = (1 + passes) * 2.0 * n * (kb + vb)
    out = {
        "metric": METRIC if kv else "GKeys/s + % HBM roofline, 2^28 u32 key-only",
        "value": round(n / (ms * 1e-3) / 1e9, 3),
        "unit": "GKeys/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": ("sort_pairs" if kv else "sort") + f" u32{'/u32' if kv else ''} n=2^{a.log2n} "
                        + ("descending" if a.descending else "ascending")
                        + ", keys uniform over [0,2^32)" + (", values = index" if kv else ""),
            "n": n, "key": "u32", "value": "u32" if kv else None,
            "order": "descending" if a.descending else "ascending",
        },
        "roofline": roof,
        "formula_roofline": {
            "bytes": formula_bytes, "device_ms": round(sort_dev_ms, 4),
            "frac_device": round(formula_bytes / (sort_dev_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
            "frac_wall": round(formula_bytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
        },
        "stages_ms_mean": [round(sum(s[i] for s in stages) / len(stages), 5) for i in range(len(stages[0]))],
        "histogram_ms": round(sum(s[1] for s in stages) / len(stages), 5),
        "checked": ok,
    }
    return out, keys0


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch

    import rsort_amd as rs

    rs.lib()
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n = 1 << a.log2n
    if a.workload == "segments":
        print(json.dumps(bench_segments(rs, a, a.steps, a.warmup)), flush=True)
        return
    if a.workload == "serve":
        print(json.dumps(bench_serve(rs, a)), flush=True)
        return
    if a.workload == "sweep":
        print(json.dumps(bench_size_sweep(rs)), flush=True)
        return
    if a.workload in CONFIG_WORKLOADS:
        r = bench_plain20(rs, a.steps, a.warmup) if a.workload == "plain20" else \
            bench_config(rs, a.workload, a.steps, a.warmup, a.descending, rs.FLAG_GRAPH if a.graph else 0)
        print(json.dumps(r), flush=True)
        return
    if a.workload == "copy":
        print(json.dumps({"metric": "streaming kernels: read / write / copy bandwidth", "unit": "GB/s",
                          **stream_ceiling(rs, n * 4, a.steps)}), flush=True)
        return
    kv = a.workload == "kv"
    extra = kv and not a.no_extra
    cold = bench_cold(rs) if extra else None  # first: the process's first sort
    out, keys0 = headline(rs, a, n, kv)
    if extra:
        print("[bench] headline done", file=sys.stderr, flush=True)
        kms, kdevs, kst, kok, _ = run_sort_bench(rs, n, False, a.descending, max(3, a.steps // 2), 2, seed=2)
        kdev = sum(kdevs) / len(kdevs)
        kroof, _ = pass_roofline(kst, 4, 0, n, 4)
        kroof["traffic"] = measured_traffic("onesweep_keys_u32_2p28_hbm_traffic.md", "k_onesweep<4,0,", n)
        kbytes = 5 * 2.0 * n * 4
        out["key_only"] = {
            "gkeys_per_s": round(n / (kms * 1e-3) / 1e9, 3), "ms_per_step": round(kms, 4),
            "device_ms": round(kdev, 4), "formula_frac_device": round(kbytes / (kdev * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
            "formula_frac_wall": round(kbytes / (kms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4), "roofline": kroof,
            "stages_ms_mean": [round(sum(s[i] for s in kst) / len(kst), 5) for i in range(len(kst[0]))],
            "checked": kok,
        }
        print("[bench] key-only done", file=sys.stderr, flush=True)
        out["hbm_ceiling"] = stream_ceiling(rs)
        cp = out["hbm_ceiling"]["copy_gbs"]
        sc = out["hbm_ceiling"]["scatter_gbs"]
        out["roofline"]["frac_of_copy_ceiling"] = round(out["roofline"]["achieved"] / cp, 4)
        out["key_only"]["roofline"]["frac_of_copy_ceiling"] = round(kroof["achieved"] / cp, 4)
        scp = out["hbm_ceiling"]["scatter_plain_gbs"]
        out["roofline"]["frac_of_scatter_ceiling"] = round(out["roofline"]["achieved"] / sc, 4)
        out["key_only"]["roofline"]["frac_of_scatter_ceiling"] = round(kroof["achieved"] / sc, 4)
        out["roofline"]["frac_of_plain_scatter_ceiling"] = round(out["roofline"]["achieved"] / scp, 4)
        out["key_only"]["roofline"]["frac_of_plain_scatter_ceiling"] = round(kroof["achieved"] / scp, 4)
        print("[bench] ceiling done", file=sys.stderr, flush=True)
        out["configs"] = {"plain20": bench_plain20(rs, a.steps, a.warmup)}
        for name in ("u64low16", "zipf", "f32", "u128kv"):
            out["configs"][name] = bench_config(rs, name, max(3, a.steps // 2), 2)
            print(f"[bench] {name} done", file=sys.stderr, flush=True)
        out["plain20_us"] = out["configs"]["plain20"]["us_per_sort"]
        out["size_sweep"] = bench_size_sweep(rs)
        print("[bench] size sweep done", file=sys.stderr, flush=True)
        rnd = bench_init_keys_mode(rs, 1, 1 << 27, max(3, a.steps // 2), 2)
        bsk = bench_init_keys_mode(rs, 2, 1 << 27, max(3, a.steps // 2), 2)
        out["byte_skip"] = {"description": "u32 n=2^27 from the reference generator: random vs byte_skip (every "
                                           "other byte constant), steady state", "random": rnd, "byte_skip": bsk}
        out["cold_vs_steady"] = dict(cold, steady_gkeys_per_s=rnd["gkeys_per_s"])
        out["segmented"] = bench_segments(rs, a, max(3, a.steps // 2), 2)
        print("[bench] segmented done", file=sys.stderr, flush=True)
        out["served"] = bench_serve(rs, a)
        print("[bench] served done", file=sys.stderr, flush=True)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(keys0, kv, min(a.cpu_log2n, a.log2n))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
