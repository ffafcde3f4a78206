"""Parity against the oracle of the code paths that environment variables
select: the early tile aggregate on every u32/u32 pass (RSORT_AMD_EARLY_AGG_TILES=1),
the ordinary lookback walk at every size (RSORT_AMD_LB_WIDE_TILES=0), the
two-level histogram fold at every grid size (RSORT_AMD_HIST_DIRECT=0), the
single global tile ticket (RSORT_AMD_XCD_TICKETS=0), the onesweep path
for sorts small enough for one LDS bin (RSORT_AMD_SMALL_LDS=0) and the fill
kernel clearing the control region before every histogram
(RSORT_AMD_HIST_CLEAR=0), and the histogram's clearing claimed only where a
workgroup first needs the counts (RSORT_AMD_HIST_CLAIM=late).

The library reads them once per process, so test_gpu_env_paths.py runs this
file in a process of its own per setting.  Cases: u32 keys and u32/u32 pairs
in both orders, u64/u64 pairs, signed keys, and keys with an odd count of
non-uniform passes (the lowest uniform pass runs as a copy)."""
from __future__ import annotations

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT, os.path.join(ROOT, "cuda-fast-sort_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402
import torch  # noqa: E402

import rsort_amd as rs  # noqa: E402
from helpers import index_values, random_keys  # noqa: E402
from oracle import oracle as orc  # noqa: E402


def check(keys, vals, kind, kb, vb, desc):
    n = keys.size // kb
    print(f"[case] {MODE} kind={kind} kb={kb} vb={vb} desc={int(desc)} n={n}", flush=True)
    dk = torch.from_numpy(keys.copy()).cuda()
    dv = torch.from_numpy(vals.copy()).cuda() if vals is not None else None
    d = rs.make_desc(kind, kb, vb, desc)
    tb = rs.sort_raw(d, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    rs.sort_raw(d, dk.data_ptr(), dv.data_ptr() if dv is not None else None, ws.data_ptr(), tb, n)
    torch.cuda.synchronize()
    ek, ev = orc.sort(keys, vals, kind, kb, desc, vb=vb if vb else None)
    assert np.array_equal(dk.cpu().numpy(), ek), "keys differ"
    if vb:
        assert np.array_equal(dv.cpu().numpy(), ev), "values differ"


MODE = " ".join(f"{k}={v}" for k, v in sorted(os.environ.items()) if k.startswith("RSORT_AMD_") and k != "RSORT_AMD_LIB")


def main():
    assert MODE, "set one of the RSORT_AMD_ path variables"
    assert torch.cuda.is_available()
    rs.lib()
    orc.lib()
    u8 = lambda a: np.ascontiguousarray(a).view(np.uint8).reshape(-1)
    for n, desc in ((1000, False), (300_007, False), (300_007, True)):
        check(u8(random_keys(0, 4, n, seed=n)), None, 0, 4, 0, desc)
        check(u8(random_keys(0, 4, n, seed=n + 1)), u8(index_values(4, n)), 0, 4, 4, desc)
    n = 1_000_003  # 62 histogram workgroups: more than one fold group
    check(u8(random_keys(0, 4, n, seed=4)), u8(index_values(4, n)), 0, 4, 4, True)
    check(u8(random_keys(0, 8, 100_003, seed=5)), u8(index_values(8, 100_003)), 0, 8, 8, False)
    check(u8(random_keys(1, 4, 200_001, seed=6)), u8(index_values(4, 200_001)), 1, 4, 4, True)
    r = orc.mt19937(3, 200_000)
    k_odd = ((r & 0xFF000000) | 0x00123456).astype(np.uint32)
    check(u8(k_odd), u8(index_values(4, 200_000)), 0, 4, 4, False)
    # sorts small enough for one LDS bin (on the onesweep path with
    # RSORT_AMD_SMALL_LDS=0): one histogram workgroup, one or few tiles
    for kind, kb, vb, n in ((0, 2, 0, 9_000), (2, 4, 4, 16_384), (1, 8, 8, 5_003), (0, 16, 8, 4_096), (0, 1, 4, 777)):
        check(u8(random_keys(kind, kb, n, seed=n + kb)), u8(index_values(vb, n)) if vb in (4, 8) else None,
              kind, kb, vb if vb in (4, 8) else 0, False)
    print("env path parity ok", flush=True)


if __name__ == "__main__":
    main()
