"""GPU parity of the array-of-structs sort (rsort_sort_records) against the
CPU oracle: the records must come out in the order of the oracle's stable
sort of their key field (its value permutation applied to the records),
byte for byte.  Record sizes cover the direct value path (8, 16 bytes) and
the index-and-gather path (odd and large sizes), aligned and unaligned key
fields."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import random_keys

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.init()


# (record bytes, key offset, key kind, key bytes, descending)
CASES = [(8, 0, 1, 4, False), (16, 8, 0, 8, True), (12, 4, 0, 4, False), (7, 1, 0, 2, False),
         (40, 36, 2, 4, True), (24, 8, 2, 8, False), (32, 16, 0, 16, False), (5, 4, 1, 1, True), (3, 0, 0, 2, False)]


@pytest.mark.parametrize("rb,off,kind,kb,desc", CASES)
@pytest.mark.parametrize("n", [1000, 150_001])
def test_records_match_oracle(rs, orc, rb, off, kind, kb, desc, n):
    rng = np.random.default_rng(rb * 31 + off + n)
    recs = rng.integers(0, 256, size=(n, rb), dtype=np.uint8)
    keys = random_keys(kind, kb, n, seed=rb + kb).reshape(n, kb)
    recs[:, off:off + kb] = keys
    print(f"[case] records rb={rb} off={off} kind={kind} kb={kb} desc={int(desc)} n={n}", flush=True)
    d = torch.from_numpy(recs.reshape(-1).copy()).cuda()
    kt = rs.KeyType(kind, kb)
    tb = rs.sort_records(d, rb, off, kt, None, None, descending=desc)
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    rs.sort_records(d, rb, off, kt, tb, ws, descending=desc)
    torch.cuda.synchronize()
    _, perm = orc.sort(np.ascontiguousarray(keys).reshape(-1), np.arange(n, dtype=np.uint32), kind, kb, desc, vb=4)
    expect = recs[perm.view(np.uint32).astype(np.int64)]
    got = d.cpu().numpy().reshape(n, rb)
    bad = np.where((got != expect).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} records differ, first {bad[:8]}"


def test_records_workspace_reuse_and_stream(rs, orc):
    """Two sorts through one workspace on a side stream, the second smaller."""
    s = torch.cuda.Stream()
    kt = rs.KeyType(0, 4)
    big = 200_000
    tb = rs.sort_records(torch.empty(big * 12, dtype=torch.uint8, device="cuda"), 12, 0, kt, None, None)
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    for n in (big, 777):
        rng = np.random.default_rng(n)
        recs = rng.integers(0, 256, size=(n, 12), dtype=np.uint8)
        d = torch.from_numpy(recs.reshape(-1).copy()).cuda()
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            rs.sort_records(d, 12, 0, kt, tb, ws, stream=s)
        s.synchronize()
        keys = np.ascontiguousarray(recs[:, 0:4]).reshape(-1)
        _, perm = orc.sort(keys, np.arange(n, dtype=np.uint32), 0, 4, False, vb=4)
        assert np.array_equal(d.cpu().numpy().reshape(n, 12), recs[perm.view(np.uint32).astype(np.int64)])
