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
HIST_PHASES = {16: "count_loop", 17: "reduce_global_add", 18: "finalize"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--values", action="store_true")
    ap.add_argument("--key-bytes", type=int, default=4)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
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


if __name__ == "__main__":
    main()
