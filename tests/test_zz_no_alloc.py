"""The allocation-free entry point (rsort_sort_noalloc) against the workspace
entry point (rsort_sort) through the C ABI: bit-exact keys and values, two
calls in a row on the same device (the second with a smaller input, so state
left by the first must not leak into it), over several widths, both orders,
and the largest u32/u32 size its per-device state holds (2^28)."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import index_values, random_keys

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1).copy()).cuda()


def _both(rs, kind, kb, vb, desc, keys_dev, vals_dev, n):
    """(keys, values) after rsort_sort and after rsort_sort_noalloc of the same input."""
    d = rs.make_desc(kind, kb, vb, desc)
    rk, rv = keys_dev.clone(), vals_dev.clone() if vals_dev is not None else None
    tb = rs.sort_raw(d, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    rs.sort_raw(d, rk.data_ptr(), rv.data_ptr() if rv is not None else None, ws.data_ptr(), tb, n)
    ak, av = keys_dev.clone(), vals_dev.clone() if vals_dev is not None else None
    alt_k = torch.empty(n * kb, dtype=torch.uint8, device="cuda")
    alt_v = torch.empty(max(n * vb, 1), dtype=torch.uint8, device="cuda")
    rs.sort_noalloc_raw(d, ak.data_ptr(), av.data_ptr() if av is not None else None, alt_k.data_ptr(),
                        alt_v.data_ptr() if vb else None, n)
    torch.cuda.synchronize()
    return (rk, rv), (ak, av)


CASES = [
    # kind, kb, vb, desc, first n, second (smaller) n
    (0, 4, 4, False, 3_000_001, 100_003),
    (0, 4, 0, True, (1 << 22) + 5, 5_000),
    (1, 8, 8, False, 1_000_003, 70_001),
    (2, 4, 4, True, 2_500_000, 300),
    (0, 16, 8, False, 600_007, 1_003),
    (1, 2, 1, False, 400_000, 20_000),
]


@pytest.mark.parametrize("kind,kb,vb,desc,n1,n2", CASES)
def test_noalloc_equals_workspace_sort(rs, kind, kb, vb, desc, n1, n2):
    for call, n in enumerate((n1, n2)):
        print(f"[case] noalloc kind={kind} kb={kb} vb={vb} desc={int(desc)} call={call} n={n}", flush=True)
        keys = _dev(random_keys(kind, kb, n, seed=n + kb))
        vals = _dev(index_values(vb, n)) if vb in (4, 8) else (
            _dev(np.random.default_rng(n).integers(0, 256, n * vb, dtype=np.uint8)) if vb else None)
        (rk, rv), (ak, av) = _both(rs, kind, kb, vb, desc, keys, vals, n)
        assert torch.equal(rk, ak), f"call {call}: keys differ from rsort_sort"
        if vb:
            assert torch.equal(rv, av), f"call {call}: values differ from rsort_sort"


def test_noalloc_largest_u32_pairs(rs):
    """u32/u32 at 2^28 (28 MiB of sort state, the largest u32 size the
    per-device array holds), then a 2^20 sort right after it."""
    for call, n in enumerate((1 << 28, 1 << 20)):
        print(f"[case] noalloc u32/u32 call={call} n={n}", flush=True)
        g = torch.Generator(device="cuda")
        g.manual_seed(call + 3)
        keys = torch.randint(-(2**31), 2**31 - 1, (n,), dtype=torch.int32, device="cuda", generator=g)
        vals = torch.arange(n, dtype=torch.int32, device="cuda")
        (rk, rv), (ak, av) = _both(rs, 0, 4, 4, False, keys.view(torch.uint8), vals.view(torch.uint8), n)
        assert torch.equal(rk, ak) and torch.equal(rv, av), f"call {call}: differs from rsort_sort"
        del rk, rv, ak, av
        torch.cuda.empty_cache()
