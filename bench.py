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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["kv", "keys", "copy"], default="kv",
                    help="kv: sort_pairs u32/u32; keys: sort u32; copy: device copy of the keys "
                         "(stream-copy bandwidth, and the known-size kernel that calibrates HBM counters)")
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


def run_sort_bench(rs, n, with_values, descending, steps, warmup, seed=1):
    """Sorts steps+warmup fresh copies; returns (ms_per_step, stage ms list per timed step)."""
    import torch

    dev = torch.device("cuda:0")
    keys0 = gen_u32(n, dev, seed)
    vals0 = torch.arange(n, dtype=torch.int32, device=dev) if with_values else None
    vb = 4 if with_values else 0
    desc = rs.make_desc(rs.KEY_UNSIGNED, 4, vb, descending)
    tb = rs.sort_raw(desc, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device=dev)
    total = steps + warmup
    ks = [keys0.clone() for _ in range(total)]
    vs = [vals0.clone() for _ in range(total)] if with_values else [None] * total
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    for i in range(warmup):
        rs.sort_raw(desc, ks[i].data_ptr(), vs[i].data_ptr() if with_values else None, ws.data_ptr(), tb, n, stream)
    torch.cuda.synchronize()
    stages = []
    t0 = time.perf_counter()
    for i in range(warmup, total):
        st = rs.sort_profiled(desc, ks[i].data_ptr(), vs[i].data_ptr() if with_values else None, ws.data_ptr(), tb,
                              n, stream)
        stages.append(st)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # spot check of the last result (sortedness + pairing) on the device
    k = ks[-1].view(torch.int32) ^ (-(2**31))  # signed order == unsigned order of the bits
    ok = bool((k[1:] >= k[:-1]).all().item())
    if with_values:
        idx = vs[-1].long()
        ok = ok and bool((keys0[idx] == ks[-1]).all().item())
        eq = ks[-1][1:] == ks[-1][:-1]
        ok = ok and bool((vs[-1][1:][eq] > vs[-1][:-1][eq]).all().item())
    del ks, vs
    torch.cuda.empty_cache()
    return (t1 - t0) * 1e3 / steps, stages, ok, keys0, vals0


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
    if a.workload == "copy":
        gbs, ms = copy_bandwidth(n, a.steps)
        print(json.dumps({"metric": "device copy bandwidth", "value": round(gbs, 1), "unit": "GB/s",
                          "ms_per_step": round(ms, 4), "bytes_per_step": 8 * n, "n_gpus": 1,
                          "steps": a.steps, "higher_is_better": True}), flush=True)
        return
    kv = a.workload == "kv"
    kb, vb = 4, (4 if kv else 0)
    passes = 4
    ms, stages, ok, keys0, _ = run_sort_bench(rs, n, kv, a.descending, a.steps, a.warmup)
    # stage layout: [clear, histogram, pass0..pass3] (+ copy-back when present)
    pass_ms = [s[2 + p] for s in stages for p in range(passes)]
    hist_ms = [s[1] for s in stages]
    sort_dev_ms = sum(sum(s) for s in stages) / len(stages)
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
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
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
        kms, kst, kok, _, _ = run_sort_bench(rs, n, False, a.descending, max(3, a.steps // 2), 2, seed=2)
        kdev = sum(sum(s) for s in kst) / len(kst)
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
