"""Benchmark of the onesweep LSD radix sort on one MI355X.

Default workload: sort_pairs of n = 2^28 uint32 keys (uniform over [0, 2^32))
with uint32 values = original index, ascending -- the key-value configuration
the headline metric is quoted on.  A step = one complete sort of a fresh copy
of that input (inputs are resident in HBM before the timed region; each step
sorts its own copy in place).  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--workload kv|keys] [--log2n 28]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cuda-fast-sort_amd"))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E peak, GB/s
METRIC = "GKeys/s + % HBM roofline, 2^28 u32; batched-segment keys/s vs per-call"


# Other workload configurations, each a bench line of its own:
# name -> (description, key kind, key bytes, value bytes, log2 n)
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
    ap.add_argument("--workload", choices=["kv", "keys", "copy", "segments", "serve", *CONFIG_WORKLOADS], default="kv",
                    help="kv: sort_pairs u32/u32; keys: sort u32; copy: device copy of the keys "
                         "(stream-copy bandwidth, and the known-size kernel that calibrates HBM counters); "
                         "segments: one segmented sort of 100k lognormal-sized u32/u32 segments vs one "
                         "sort_pairs call per segment; serve: Poisson request stream through the serving loop")
    ap.add_argument("--segments", type=int, default=100_000)
    ap.add_argument("--per-call-sample", type=int, default=5000,
                    help="segments sorted one sort_pairs call each for the per-call comparison")
    ap.add_argument("--rate", type=float, default=5000.0, help="serve: Poisson arrival rate, requests/s")
    ap.add_argument("--duration", type=float, default=2.0, help="serve: seconds of arrivals")
    ap.add_argument("--window-us", type=int, default=0,
                    help="serve: batching window (0: launch as soon as a batch slot is free)")
    ap.add_argument("--serve-mode", choices=["direct", "graph", "batched"], default="direct")
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--descending", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the key-only side measurement")
    ap.add_argument("--cpu-log2n", type=int, default=24, help="CPU baseline sample size")
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


def run_sort_bench(rs, n, with_values, descending, steps, warmup, seed=1):
    """Sorts warmup + steps + profiled fresh copies; returns (wall ms per timed
    sort, stage ms lists of the profiled sorts, check, keys0, vals0, device ms
    per timed sort)."""
    import torch

    dev = torch.device("cuda:0")
    keys0 = gen_u32(n, dev, seed)
    vals0 = torch.arange(n, dtype=torch.int32, device=dev) if with_values else None
    vb = 4 if with_values else 0
    desc = rs.make_desc(rs.KEY_UNSIGNED, 4, vb, descending)
    tb = rs.sort_raw(desc, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device=dev)
    profiled = min(steps, 5)
    total = warmup + steps + profiled
    ks = [keys0.clone() for _ in range(total)]
    vs = [vals0.clone() for _ in range(total)] if with_values else [None] * total
    torch.cuda.synchronize()
    wall_ms, dev_ms, stages = timed_sorts(rs, desc, ks, vs, ws, tb, n, warmup, steps, profiled)
    # spot check of the last result (sortedness + pairing) on the device
    k = ks[-1].view(torch.int32) ^ (-(2**31))  # signed order == unsigned order of the bits
    ok = bool((k[1:] <= k[:-1]).all().item()) if descending else bool((k[1:] >= k[:-1]).all().item())
    if with_values:
        idx = vs[-1].long()
        ok = ok and bool((keys0[idx] == ks[-1]).all().item())
        eq = ks[-1][1:] == ks[-1][:-1]
        ok = ok and bool((vs[-1][1:][eq] > vs[-1][:-1][eq]).all().item())
    del ks, vs
    torch.cuda.empty_cache()
    return wall_ms, stages, ok, keys0, vals0, dev_ms


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
        keys = np.random.RandomState(0).randint(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
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


def bench_config(rs, a):
    """One of the other workload configurations: GKeys/s, the formula roofline
    (every pass counted, also those the device skips), per-stage device times,
    device-side checks (order, digit histograms, value pairing)."""
    import torch

    desc_s, kind, kb, vb, log2n = CONFIG_WORKLOADS[a.workload]
    n = 1 << log2n
    dev = torch.device("cuda:0")
    keys0, vals0 = gen_config(a.workload, n, dev, 21)
    d = rs.make_desc(kind, kb, vb, a.descending)
    tb = rs.sort_raw(d, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device=dev)
    before = rs.digit_counts(keys0, d, n)
    profiled = min(a.steps, 5)
    total = a.warmup + a.steps + profiled
    ks = [keys0.clone() for _ in range(total)]
    vs = [vals0.clone() for _ in range(total)] if vb else [None] * total
    torch.cuda.synchronize()
    wall_ms, dev_steps, stages = timed_sorts(rs, d, ks, vs, ws, tb, n, a.warmup, a.steps, profiled)
    ok = rs.check_sorted(ks[-1], d, n) == 0 and torch.equal(rs.digit_counts(ks[-1], d, n), before)
    if vb:
        ok = ok and rs.check_pairings(keys0, ks[-1], vs[-1], kb, vb, n) == 0
    del ks, vs
    dev_ms = sum(dev_steps) / len(dev_steps)
    pass_ms = [sum(s[2 + p] for s in stages) / len(stages) for p in range(kb)]
    longest = max(pass_ms)
    executed = [p for p in range(kb) if pass_ms[p] > 0.05 * longest]
    exec_avg = sum(pass_ms[p] for p in executed) / len(executed)
    pass_bytes = 2.0 * n * (kb + vb)
    formula_bytes = (1 + kb) * 2.0 * n * (kb + vb)
    return {
        "metric": f"GKeys/s, {desc_s}", "value": round(n / (dev_ms * 1e-3) / 1e9, 3), "unit": "GKeys/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(dev_ms, 4),
        "us_per_sort": round(dev_ms * 1e3, 2), "wall_ms_per_step": round(wall_ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": {4: "u32", 8: "u64", 16: "u128"}[kb] if kind != 2 else "f32",
        "data": "synthetic",
        "config": {"workload": a.workload, "description": desc_s, "n": n, "key_kind": kind, "key_bytes": kb,
                   "value_bytes": vb, "order": "descending" if a.descending else "ascending"},
        "roofline": {"bound": "hbm", "kernel": "k_onesweep (one executed LSD pass)",
                     "achieved": round(pass_bytes / (exec_avg * 1e-3) / 1e9, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(pass_bytes / (exec_avg * 1e-3) / 1e9 / PEAK_HBM_GBS, 4), "traffic": None,
                     "bytes_per_launch": pass_bytes, "avg_launch_ms": round(exec_avg, 5)},
        "formula_roofline": {"bytes": formula_bytes, "passes_counted": kb, "device_ms": round(dev_ms, 4),
                             "frac_device": round(formula_bytes / (dev_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)},
        "executed_passes": executed,
        "stages_ms_mean": [round(sum(s[i] for s in stages) / len(stages), 5) for i in range(len(stages[0]))],
        "checked": bool(ok),
    }


def measured_traffic(kernel_prefix, n):
    """HBM bytes per launch of the pass kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE and WRITE_SIZE, calibrated against a copy of known
    size), when it was taken on this workload; else None."""
    path = os.path.join(ROOT, "profiles", "onesweep_kv_u32_2p28_hbm_traffic.md")
    try:
        text = open(path).read()
        block = text[text.index("```json") + 7:]
        d = json.loads(block[:block.index("```")])
    except (OSError, ValueError):
        return None
    if d.get("n") == n and d.get("kernel", "").startswith(kernel_prefix):
        return d["traffic_bytes"]
    return None


def copy_bandwidth(n, steps):
    """Device-to-device copy of n u32 (read n*4 + write n*4 bytes); returns GB/s."""
    import torch

    src = gen_u32(n, torch.device("cuda:0"), 3)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    return 2.0 * n * 4 / (ms * 1e-3) / 1e9, ms


def lognormal_sizes(count, seed=5, median=2048.0, sigma=0.7356, lo=32, hi=1 << 22):
    """Request / segment sizes: lognormal with median 2048 and sigma 0.7356
    (mean ~2684 keys), rounded and clipped to [32, 2^22]."""
    import numpy as np

    rng = np.random.Generator(np.random.MT19937(seed))
    v = np.floor(median * np.exp(sigma * rng.standard_normal(count)) + 0.5)
    return np.clip(v, lo, hi).astype(np.int64)


def bench_segments(rs, a):
    """One segmented sort of the whole batch per step vs one sort_pairs call per segment."""
    import numpy as np
    import torch

    dev = torch.device("cuda:0")
    sizes = lognormal_sizes(a.segments)
    n = int(sizes.sum())
    offs_np = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    offsets = torch.from_numpy(offs_np).to(dev)
    keys0 = gen_u32(n, dev, 7)
    vals0 = torch.arange(n, dtype=torch.int32, device=dev)
    desc = rs.make_desc(rs.KEY_UNSIGNED, 4, 4, a.descending)
    tb = rs.segmented_sort_raw(desc, None, None, n, None, 8, a.segments, None, 0)
    ws = torch.empty(tb, dtype=torch.uint8, device=dev)
    total = a.steps + a.warmup
    ks = [keys0.clone() for _ in range(total)]
    vs = [vals0.clone() for _ in range(total)]
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    call = lambda i: rs.segmented_sort_raw(desc, ks[i].data_ptr(), vs[i].data_ptr(), n, offsets.data_ptr(), 8,
                                           a.segments, ws.data_ptr(), tb, stream)
    for i in range(a.warmup):
        call(i)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(a.steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for j, i in enumerate(range(a.warmup, total)):
        call(i)
        ev[j + 1].record(stream)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / a.steps
    dev_ms = [ev[j].elapsed_time(ev[j + 1]) for j in range(a.steps)]
    ms = sum(dev_ms) / len(dev_ms)
    # every segment sorted, and values still pair with their keys
    k = ks[-1].view(torch.int32) ^ (-(2**31))
    up = k[1:] >= k[:-1]
    seg_start = torch.zeros(n, dtype=torch.bool, device=dev)
    seg_start[offsets[1:-1]] = True
    ok = bool((up | seg_start[1:]).all().item()) and bool((keys0[vs[-1].long()] == ks[-1]).all().item())
    lds_cap = rs.segmented_plan(desc, n, 8, a.segments).bin_capacity[4]
    small = sizes[sizes <= lds_cap].sum()
    big = sizes[sizes > lds_cap].sum()
    kvb = 8
    bytes_ = 2.0 * small * kvb + 5 * 2.0 * big * kvb  # LDS segments: one read + one write; long: formula
    batched_k, batched_v = ks[-1], vs[-1]
    del ks, vs
    # per-call comparison: the same segments, one sort_pairs call each
    m = min(a.per_call_sample, a.segments)
    pk, pv = keys0.clone(), vals0.clone()
    fdesc = rs.make_desc(rs.KEY_UNSIGNED, 4, 4, a.descending)
    ftb = rs.sort_raw(fdesc, None, None, None, 0, int(sizes[:m].max()))
    fws = torch.empty(ftb, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kp, vp = pk.data_ptr(), pv.data_ptr()
    for i in range(m):
        o, c = int(offs_np[i]), int(sizes[i])
        rs.sort_raw(fdesc, kp + 4 * o, vp + 4 * o, fws.data_ptr(), ftb, c, stream)
    torch.cuda.synchronize()
    per_call_s = time.perf_counter() - t0
    per_call_keys = int(sizes[:m].sum())
    same = bool((pk[:per_call_keys] == batched_k[:per_call_keys]).all().item()) and bool(
        (pv[:per_call_keys] == batched_v[:per_call_keys]).all().item())
    batched = n / (ms * 1e-3) / 1e9
    per_call = per_call_keys / per_call_s / 1e9
    return {
        "metric": "segmented sort keys/s, 100k lognormal segments vs per-call sort_pairs",
        "value": round(batched, 3), "unit": "GKeys/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 4), "wall_ms_per_step": round(wall_ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"segmented sort_pairs u32/u32, {a.segments} segments, lognormal sizes median 2048 "
                               f"sigma 0.7356 clipped [32, 2^22], n={n}", "segments": a.segments, "n": n,
                   "lds_capacity": int(lds_cap), "keys_in_lds_segments": int(small), "keys_in_long_segments": int(big)},
        "roofline": {"bound": "hbm", "kernel": "whole segmented sort (device time)",
                     "achieved": round(bytes_ / (ms * 1e-3) / 1e9, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(bytes_ / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4), "traffic": None,
                     "bytes": bytes_},
        "per_call": {"gkeys_per_s": round(per_call, 5), "calls": m, "keys": per_call_keys,
                     "us_per_call": round(per_call_s * 1e6 / m, 2),
                     "note": "one rsort_sort call (sort_pairs) per segment from Python via ctypes, same stream"},
        "batched_over_per_call": round(batched / per_call, 1),
        "checked": ok,
        "per_call_equals_batched": same,
    }


def bench_serve(rs, a):
    """Poisson request stream through the serving loop: latency percentiles and
    throughput; then the same requests submitted at once (saturation)."""
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
    # warm-up requests (not timed)
    wk, wv = keys_ref.clone(), vals.clone()
    warm = [_submit_raw(rs, srv, wk.data_ptr() + 4 * int(offs[i]), wv.data_ptr() + 4 * int(offs[i]), int(sizes[i]))
            for i in range(min(50, count))]
    for w in warm:
        srv.wait(w)
    del wk, wv
    keys.copy_(keys_ref)
    vals.copy_(torch.arange(n, dtype=torch.int32, device=dev))
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
    # correctness: every request sorted with its values
    k = keys.view(torch.int32) ^ (-(2**31))
    starts = torch.zeros(n, dtype=torch.bool, device=dev)
    starts[torch.from_numpy(offs[1:-1]).to(dev)] = True
    ok = bool(((k[1:] >= k[:-1]) | starts[1:]).all().item()) and bool((keys_ref[vals.long()] == keys).all().item())
    # saturation: the same requests all at once (one submit_many call)
    keys.copy_(keys_ref)
    kptrs = [kp + 4 * int(offs[i]) for i in range(count)]
    vptrs = [vp + 4 * int(offs[i]) for i in range(count)]
    ns = [int(x) for x in sizes]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ids2 = srv.submit_many(kptrs, vptrs, ns)
    for i in ids2:
        srv.wait(i)
    t_sat = time.perf_counter() - t0
    st2 = srv.stats()
    srv.close()
    return {
        "metric": "serving loop latency p50/p99 at a Poisson request rate, lognormal request sizes",
        "value": round(float(np.percentile(lat, 99)), 1), "unit": "us (p99 latency)", "n_gpus": 1,
        "higher_is_better": False, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"sort_pairs u32/u32 requests, Poisson {a.rate:g} req/s for {a.duration:g} s, sizes "
                               "lognormal median 2048 sigma 0.7356 clipped [32, 2^22]", "requests": count,
                   "keys": n, "batch_keys": 1 << 24, "batch_requests": 8192, "window_us": a.window_us, "depth": 3,
                   "mode": a.serve_mode},
        "latency_us": {"p50": round(float(np.percentile(lat, 50)), 1), "p90": round(float(np.percentile(lat, 90)), 1),
                       "p99": round(float(np.percentile(lat, 99)), 1), "max": round(float(lat.max()), 1),
                       "mean": round(float(lat.mean()), 1), "queue_p50": round(float(np.percentile(q, 50)), 1)},
        "offered_keys_per_s": round(n / a.duration / 1e9, 5),
        "achieved_keys_per_s": round(n / (t_end - t_start) / 1e9, 5),
        "batches": int(st.batches - st0.batches),
        "mean_requests_per_batch": round(count / max(1, int(st.batches - st0.batches)), 2),
        "saturation": {"gkeys_per_s": round(n / t_sat / 1e9, 3), "seconds": round(t_sat, 4),
                       "batches": int(st2.batches - st.batches)},
        "checked": ok,
    }


def _submit_raw(rs, srv, kptr, vptr, n):
    import ctypes

    rid = ctypes.c_uint64(0)
    rs._check(rs.lib().rsort_serve_submit(srv._h, kptr, vptr, n, ctypes.byref(rid)))
    return int(rid.value)


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
        print(json.dumps(bench_segments(rs, a)), flush=True)
        return
    if a.workload == "serve":
        print(json.dumps(bench_serve(rs, a)), flush=True)
        return
    if a.workload in CONFIG_WORKLOADS:
        print(json.dumps(bench_config(rs, a)), flush=True)
        return
    if a.workload == "copy":
        gbs, ms = copy_bandwidth(n, a.steps)
        print(json.dumps({"metric": "device copy bandwidth", "value": round(gbs, 1), "unit": "GB/s",
                          "ms_per_step": round(ms, 4), "bytes_per_step": 8 * n, "n_gpus": 1,
                          "steps": a.steps, "higher_is_better": True}), flush=True)
        return
    kv = a.workload == "kv"
    kb, vb = 4, (4 if kv else 0)
    passes = 4
    ms, stages, ok, keys0, _, dev_steps = run_sort_bench(rs, n, kv, a.descending, a.steps, a.warmup)
    # stage layout (profiled sorts): [clear, histogram, pass0..pass3] (+ copy-back when present)
    pass_ms = [s[2 + p] for s in stages for p in range(passes)]
    hist_ms = [s[1] for s in stages]
    sort_dev_ms = sum(dev_steps) / len(dev_steps)
    avg_pass = sum(pass_ms) / len(pass_ms)
    pass_bytes = 2.0 * n * (kb + vb)  # read + write of keys and values per pass
    achieved = pass_bytes / (avg_pass * 1e-3) / 1e9
    formula_bytes = (1 + passes) * 2.0 * n * (kb + vb)
    gkeys = n / (ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(gkeys, 3),
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
        "roofline": {
            "bound": "hbm", "kernel": "k_onesweep (one LSD pass)",
            "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": measured_traffic("k_onesweep<4,4," if kv else "k_onesweep<4,0,", n),
            "bytes_per_launch": pass_bytes, "avg_launch_ms": round(avg_pass, 5),
        },
        "formula_roofline": {
            "bytes": formula_bytes, "device_ms": round(sort_dev_ms, 4),
            "frac_device": round(formula_bytes / (sort_dev_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
            "frac_wall": round(formula_bytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
        },
        "stages_ms_mean": [round(sum(s[i] for s in stages) / len(stages), 5) for i in range(len(stages[0]))],
        "histogram_ms": round(sum(hist_ms) / len(hist_ms), 5),
        "checked": ok,
    }
    cgbs, _ = copy_bandwidth(n, 5)
    out["stream_copy_gbs"] = round(cgbs, 1)  # measured device copy bandwidth on this box
    if kv and not a.no_extra:
        kms, kst, kok, _, _, kdevs = run_sort_bench(rs, n, False, a.descending, max(3, a.steps // 2), 2, seed=2)
        kdev = sum(kdevs) / len(kdevs)
        kbytes = (1 + passes) * 2.0 * n * kb
        out["key_only"] = {
            "gkeys_per_s": round(n / (kms * 1e-3) / 1e9, 3), "ms_per_step": round(kms, 4),
            "device_ms": round(kdev, 4), "formula_frac_device": round(kbytes / (kdev * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
            "formula_frac_wall": round(kbytes / (kms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4), "checked": kok,
        }
    if not a.no_cpu_baseline:
        from oracle import oracle

        m = 1 << a.cpu_log2n
        sample_k = keys0[:m].cpu().numpy().view("uint32")
        sample_v = None
        if kv:
            import numpy as np

            sample_v = np.arange(m, dtype=np.uint32)
        t_st = oracle.time_stable_sort_u32(sample_k, sample_v, 3)
        t_par = oracle.time_sort_par_u32(sample_k, sample_v, 3)
        par_threads = os.cpu_count() if oracle.pstl_parallel_backend() else 1
        out["cpu_baseline"] = {
            "value": round(m / t_st / 1e9, 5), "unit": "GKeys/s", "cores": 1, "kind": "port",
            "sample": f"std::stable_sort of 2^{a.cpu_log2n} (u32 key{', u32 value' if kv else ''}) records "
                      f"from the same input, median of 3 runs",
            "par_unseq": {"value": round(m / t_par / 1e9, 5), "unit": "GKeys/s", "cores": par_threads,
                          "note": "std::sort(std::execution::par_unseq); libstdc++ without TBB runs it serially"
                          if par_threads == 1 else "std::sort(std::execution::par_unseq)"},
            "host_cores_available": os.cpu_count(),
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
