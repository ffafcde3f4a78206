"""Autotuner: measures the library's tunables on this GPU and writes the
persisted tuning table the library reads at first use.

    python cuda-fast-sort_amd/autotune.py [--log2n 28] [--out PATH]

For each tuned (key bytes, value bytes) it times, with the library's own
per-stage device events (rsort_sort_profiled):
  * every pass-kernel tile candidate            -> fastest mean pass time
  * histogram workgroups per CU in 1..4         -> fastest histogram
and, with device events around the whole rsort_sort call at 2^20 .. 2^24
and 2^26 keys:
  * every tile candidate again                  -> up to two small-size tiers
    (tiny_tile_candidate below 2^tiny_log2n, small_tile_candidate below
    2^small_log2n), scored by the summed time relative to the fastest
    candidate per size
and, with device events around rsort_segmented_sort on a batch of
lognormal-sized segments (median 2048, sigma 0.7356, clipped [32, 2^22]):
  * the LDS threshold (each LDS bin capacity)   -> fastest segmented sort.
The table (one line per entry: key_bytes value_bytes tile_candidate
hist_blocks_per_cu lds_threshold small_tile_candidate small_log2n
tiny_tile_candidate tiny_log2n) goes to --out, by default
tuning_gfx950.txt next to librsort_amd.so, and is printed.  Measurements are
on synthetic uniform keys; the reference has no tuner (its geometry is a
fixed sm_86 table, include/rsort/detail/tuning.cuh:138-229).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

# (key bytes, value bytes) tuned by default: the widths with tile candidates
TUNED = [(4, 0), (4, 4), (8, 0), (8, 8), (4, 8), (8, 4), (16, 0), (16, 8)]
# flat sorts are measured at 2^log2n keys, fewer where one array would pass 2 GiB
ARRAY_BYTES_MAX = 1 << 31
# untimed u32/u32 sorts before the first measurement
WARMUP_SORTS = 12
# pass-kernel tile candidates (kTileCands in csrc/rsort_tuning.hpp)
TILE_CANDIDATES = 4


def lognormal_sizes(count, seed=5):
    import numpy as np

    rng = np.random.Generator(np.random.MT19937(seed))
    v = np.floor(2048.0 * np.exp(0.7356 * rng.standard_normal(count)) + 0.5)
    return np.clip(v, 32, 1 << 22).astype(np.int64)


def time_flat(rs, torch, kb, vb, n, reps):
    """Mean pass ms and histogram ms of the current tuning for (kb, vb)."""
    dev = torch.device("cuda:0")
    keys0 = torch.randint(0, 256, (n * kb,), dtype=torch.uint8, device=dev)
    vals0 = torch.randint(0, 256, (n * vb,), dtype=torch.uint8, device=dev) if vb else None
    d = rs.make_desc(rs.KEY_UNSIGNED, kb, vb)
    tb = rs.sort_raw(d, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device=dev)
    passes, hist = [], []
    for r in range(reps + 1):
        k = keys0.clone()
        v = vals0.clone() if vb else None
        torch.cuda.synchronize()
        st = rs.sort_profiled(d, k.data_ptr(), v.data_ptr() if vb else None, ws.data_ptr(), tb, n)
        if r:  # the first run warms up
            hist.append(st[1])
            passes.extend(st[2:2 + kb])
    del ws, keys0, vals0
    return sum(passes) / len(passes), sum(hist) / len(hist)


def time_whole(rs, torch, kb, vb, n, reps):
    """Median ms of whole rsort_sort calls (events around the call) for n keys."""
    dev = torch.device("cuda:0")
    keys0 = torch.randint(0, 256, (n * kb,), dtype=torch.uint8, device=dev)
    vals0 = torch.randint(0, 256, (n * vb,), dtype=torch.uint8, device=dev) if vb else None
    d = rs.make_desc(rs.KEY_UNSIGNED, kb, vb)
    tb = rs.sort_raw(d, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device=dev)
    k = keys0.clone()
    v = vals0.clone() if vb else None
    ms = []
    for r in range(reps + 3):
        k.copy_(keys0)
        if vb:
            v.copy_(vals0)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rs.sort_raw(d, k.data_ptr(), v.data_ptr() if vb else None, ws.data_ptr(), tb, n,
                    torch.cuda.current_stream())
        e1.record()
        torch.cuda.synchronize()
        if r >= 3:
            ms.append(e0.elapsed_time(e1))
    ms.sort()
    return ms[len(ms) // 2]


def small_choice(small_ms, large_cand):
    """Small-size tiers (tiny_cand, tiny_log2n, small_cand, small_log2n):
    sorts below 2^tiny_log2n use tiny_cand, those below 2^small_log2n
    small_cand, larger ones large_cand; a log2n of 0 switches a tier off.
    Every choice of candidates and of thresholds just above measured sizes is
    scored by the sum over the measured sizes of its time relative to the
    fastest candidate there.  The lowest score wins; on ties the choice with
    fewer tiers (none, then the small tier alone)."""
    sizes = sorted(small_ms)
    best = {s: min(small_ms[s].values()) for s in sizes}
    cands = list(small_ms[sizes[0]])
    limits = [s + 1 for s in sizes]

    def score(tc, tl, sc, sl):
        total = 0.0
        for s in sizes:
            c = tc if (tl and s < tl) else sc if (sl and s < sl) else large_cand
            total += small_ms[s][c] / best[s]
        return total

    choice, top = (0, 0, 0, 0), score(0, 0, large_cand, 0)
    for sc in cands:  # one tier
        for sl in limits:
            v = score(0, 0, sc, sl)
            if v < top - 1e-9:
                choice, top = (0, 0, sc, sl), v
    for tc in cands:  # two tiers
        for tl in limits:
            for sc in cands:
                for sl in limits:
                    if sl <= tl or sc == tc:
                        continue
                    v = score(tc, tl, sc, sl)
                    if v < top - 1e-9:
                        choice, top = (tc, tl, sc, sl), v
    return choice


def time_segmented(rs, torch, kb, vb, sizes, reps):
    import numpy as np

    dev = torch.device("cuda:0")
    n = int(sizes.sum())
    off = torch.from_numpy(np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)).to(dev)
    keys0 = torch.randint(0, 256, (n * kb,), dtype=torch.uint8, device=dev)
    vals0 = torch.randint(0, 256, (n * vb,), dtype=torch.uint8, device=dev) if vb else None
    d = rs.make_desc(rs.KEY_UNSIGNED, kb, vb)
    tb = rs.segmented_sort_raw(d, None, None, n, None, 8, sizes.size, None, 0)
    ws = torch.empty(tb, dtype=torch.uint8, device=dev)
    ms = []
    for r in range(reps + 1):
        k = keys0.clone()
        v = vals0.clone() if vb else None
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rs.segmented_sort_raw(d, k.data_ptr(), v.data_ptr() if vb else None, n, off.data_ptr(), 8, sizes.size,
                              ws.data_ptr(), tb)
        e1.record()
        torch.cuda.synchronize()
        if r:
            ms.append(e0.elapsed_time(e1))
    del ws, keys0, vals0
    return sum(ms) / len(ms)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--log2n", type=int, default=28, help="keys per flat sort measurement")
    ap.add_argument("--segments", type=int, default=20000, help="segments per segmented measurement")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning_gfx950.txt"))
    ap.add_argument("--widths", default=None, help="key:value byte pairs to tune, e.g. 4:4,16:8 (default: all)")
    a = ap.parse_args()
    import torch

    import rsort_amd as rs

    assert torch.cuda.is_available(), "the autotuner measures on a GPU"
    rs.tuning_load("")  # start from the built-in defaults
    # The device runs its first sorts of a process slower (up to ~10 % for
    # the first eight or so); warm it up before anything is timed, or the
    # first candidates measured carry that transient.
    for _ in range(WARMUP_SORTS // 2):
        time_flat(rs, torch, 4, 4, 1 << min(a.log2n, 28), 1)  # two sorts each
    sizes = lognormal_sizes(a.segments)
    lines, report = [], []
    widths = TUNED if not a.widths else [tuple(int(x) for x in w.split(":")) for w in a.widths.split(",")]
    for kb, vb in widths:
        log2n = a.log2n
        while (kb + vb) << log2n > ARRAY_BYTES_MAX:
            log2n -= 1
        n = 1 << log2n
        base = rs.tuning_get(kb, vb)
        caps = list(rs.segmented_plan(rs.make_desc(0, kb, vb), 1 << 20, 8, 1).bin_capacity)
        # tile candidates (histogram setting fixed at its default)
        tile_ms = {}
        for c in range(TILE_CANDIDATES):
            rs.tuning_set(kb, vb, c, base.hist_blocks_per_cu, base.lds_threshold)
            tile_ms[c], _ = time_flat(rs, torch, kb, vb, n, a.reps)
        best_tile = min(tile_ms, key=tile_ms.get)
        # histogram workgroups per CU
        hist_ms = {}
        for bpc in (1, 2, 3, 4):
            rs.tuning_set(kb, vb, best_tile, bpc, base.lds_threshold)
            _, hist_ms[bpc] = time_flat(rs, torch, kb, vb, n, a.reps)
        best_bpc = min(hist_ms, key=hist_ms.get)
        # LDS threshold of the segmented sort
        seg_ms = {}
        for cap in caps:
            rs.tuning_set(kb, vb, best_tile, best_bpc, cap)
            seg_ms[cap] = time_segmented(rs, torch, kb, vb, sizes, a.reps)
        best_cap = min(seg_ms, key=seg_ms.get)
        # tile candidate of small sorts
        small_ms = {}
        for l2 in [s for s in (20, 21, 22, 23, 24, 26) if s < log2n]:
            row = {}
            for c in range(TILE_CANDIDATES):
                rs.tuning_set(kb, vb, c, best_bpc, best_cap)
                row[c] = time_whole(rs, torch, kb, vb, 1 << l2, a.reps + 4)
            small_ms[l2] = row
        tiny_cand, tiny_log2n, small_cand, small_log2n = small_choice(small_ms, best_tile)
        rs.tuning_set(kb, vb, best_tile, best_bpc, best_cap, small_cand, small_log2n, tiny_cand, tiny_log2n)
        lines.append(f"{kb} {vb} {best_tile} {best_bpc} {best_cap} {small_cand} {small_log2n} {tiny_cand} {tiny_log2n}")
        report.append({"key_bytes": kb, "value_bytes": vb, "log2n": log2n, "pass_ms_by_tile_candidate": tile_ms,
                       "histogram_ms_by_blocks_per_cu": hist_ms, "segmented_ms_by_lds_threshold": seg_ms,
                       "sort_ms_by_log2n_and_candidate": small_ms,
                       "chosen": {"tile_candidate": best_tile, "hist_blocks_per_cu": best_bpc,
                                  "lds_threshold": best_cap, "small_tile_candidate": small_cand,
                                  "small_log2n": small_log2n, "tiny_tile_candidate": tiny_cand,
                                  "tiny_log2n": tiny_log2n}})
        print(json.dumps(report[-1]), flush=True)
    props = torch.cuda.get_device_properties(0)
    header = [f"# rsort tuning table for {getattr(props, 'gcnArchName', 'gfx950')} ({props.name}), written by autotune.py",
              f"# {time.strftime('%Y-%m-%d')}: flat sorts of 2^{a.log2n} uniform keys (fewer where an array "
              f"would pass {ARRAY_BYTES_MAX >> 30} GiB), segmented sorts of "
              f"{a.segments} lognormal segments, mean of {a.reps} runs each, after {WARMUP_SORTS} untimed sorts",
              "# key_bytes value_bytes tile_candidate hist_blocks_per_cu lds_threshold small_tile_candidate small_log2n "
              "tiny_tile_candidate tiny_log2n"]
    text = "\n".join(header + lines) + "\n"
    with open(a.out, "w") as f:
        f.write(text)
    print(text, end="", flush=True)


if __name__ == "__main__":
    main()
