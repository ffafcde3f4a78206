"""Per-phase time of the sort kernels, from a -DRSA_PHASE_TIMING build.

    python cuda-fast-sort_amd/build.py --variant timing
    python cuda-fast-sort_amd/phase_timing.py [--log2n 28] [--values] [--reps 3]

Prints one JSON line: microseconds per tile (pass kernel, thread 0's view)
and per workgroup (histogram) for each phase, plus lookback loads and
not-ready polls per tile.  Timing builds serialise a few atomics per tile,
so absolute times are slightly above the release build's.
"""
from __future__ import annotations

import argparse
import json
import os
import sys

os.environ.setdefault("RSORT_AMD_LIB", "librsort_amd_timing.so")
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

PASS_PHASES = ["ticket", "load", "rank", "scan_publish", "lookback", "stage", "scatter_keys", "values"]
SEG_PHASES = {22: "ticket", 23: "load", 24: "zero_rank", 25: "digit_scan", 26: "exchange_keys",
              27: "exchange_values", 28: "store"}
HIST_PHASES = {16: "count_loop", 17: "reduce_global_add", 18: "finalize"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--values", action="store_true")
    ap.add_argument("--key-bytes", type=int, default=4)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--seg-len", type=int, default=0,
                    help="segmented sort of equal segments of this length (one LDS bin) instead")
    ap.add_argument("--segments", type=int, default=30000)
    a = ap.parse_args()
    if a.seg_len:
        return segmented(a)
    import torch

    import rsort_amd as rs

    n = 1 << a.log2n
    kb, vb = a.key_bytes, (4 if a.values else 0)
    dev = torch.device("cuda:0")
    keys0 = torch.randint(0, 256, (n * kb,), dtype=torch.uint8, device=dev)
    vals0 = torch.arange(n, dtype=torch.int32, device=dev) if vb else None
    d = rs.make_desc(rs.KEY_UNSIGNED, kb, vb)
    tb = rs.sort_raw(d, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device=dev)
    plan = rs.plan(d, n)
    k, v = keys0.clone(), (vals0.clone() if vb else None)
    rs.sort_raw(d, k.data_ptr(), v.data_ptr() if vb else None, ws.data_ptr(), tb, n)
    torch.cuda.synchronize()
    rs.debug_phase_slots(reset=True)
    stages = []
    for _ in range(a.reps):
        k, v = keys0.clone(), (vals0.clone() if vb else None)
        torch.cuda.synchronize()
        stages.append(rs.sort_profiled(d, k.data_ptr(), v.data_ptr() if vb else None, ws.data_ptr(), tb, n))
    slots = rs.debug_phase_slots(reset=True)
    if not slots:
        raise SystemExit("the loaded library is not a timing build")
    tiles = max(1, slots[11])
    hist_wgs = max(1, slots[21])
    out = {
        "n": n, "key_bytes": kb, "value_bytes": vb, "tile": plan.tile, "tiles_per_pass": plan.num_tiles,
        "tiles_measured": tiles,
        "pass_us_per_tile": {name: round(slots[i] / tiles / 100.0, 3) for i, name in enumerate(PASS_PHASES)},
        "pass_total_us_per_tile": round(slots[10] / tiles / 100.0, 3),
        "lookback_loads_per_tile": round(slots[8] / tiles, 2),
        "lookback_polls_per_tile": round(slots[9] / tiles, 2),
        "hist_us_per_wg": {name: round(slots[i] / hist_wgs / 100.0, 3) for i, name in HIST_PHASES.items()},
        "hist_total_us_per_wg": round(slots[20] / hist_wgs / 100.0, 3),
        "hist_wgs": hist_wgs,
        "stage_ms": [round(sum(s[i] for s in stages) / len(stages), 4) for i in range(len(stages[0]))],
    }
    print(json.dumps(out), flush=True)


def segmented(a):
    """Phases of the LDS-bin kernel (thread 0 of each workgroup, summed over
    its segments; microseconds per segment) for equal-length segments."""
    import torch

    import rsort_amd as rs

    dev = torch.device("cuda:0")
    n = a.seg_len * a.segments
    vb = 4 if a.values else 0
    keys0 = torch.randint(0, 256, (n * 4,), dtype=torch.uint8, device=dev)
    vals0 = torch.arange(n, dtype=torch.int32, device=dev) if vb else None
    off = torch.arange(0, n + 1, a.seg_len, dtype=torch.int64, device=dev)
    d = rs.make_desc(rs.KEY_UNSIGNED, 4, vb)
    tb = rs.segmented_sort_raw(d, None, None, n, None, 8, a.segments, None, 0)
    ws = torch.empty(tb, dtype=torch.uint8, device=dev)

    def run():
        k, v = keys0.clone(), (vals0.clone() if vb else None)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rs.segmented_sort_raw(d, k.data_ptr(), v.data_ptr() if vb else None, n, off.data_ptr(), 8, a.segments,
                              ws.data_ptr(), tb)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    run()
    rs.debug_phase_slots(reset=True)
    ms = [run() for _ in range(a.reps)]
    slots = rs.debug_phase_slots(reset=True)
    if not slots:
        raise SystemExit("the loaded library is not a timing build")
    segs = max(1, slots[29])
    wgs = max(1, slots[31] if len(slots) > 31 else 1)
    out = {"seg_len": a.seg_len, "segments": a.segments, "value_bytes": vb, "segments_measured": segs,
           "us_per_segment": {name: round(slots[i] / segs / 100.0, 3) for i, name in SEG_PHASES.items()},
           "total_us_per_segment": round(sum(slots[i] for i in SEG_PHASES) / segs / 100.0, 3),
           "workgroups": wgs, "wg_lifetime_us": round(slots[30] / wgs / 100.0, 2),
           "sort_ms": round(sum(ms) / len(ms), 4), "gkeys_per_s": round(n / (sum(ms) / len(ms) * 1e-3) / 1e9, 2)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
