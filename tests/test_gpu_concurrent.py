"""Sorts that overlap on the device: several sorts launched on separate
streams at once, each with its own workspace, every result bit-exact against
the oracle.  At most one of them holds the device's per-XCD ticket lock at a
time (the others take the global tile ticket), so overlapping passes never
wait on each other's undispatched tiles.  The lock is taken and freed on the
device; a later sort alone finds it free again (its result is checked too)."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import index_values, random_keys

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _sorts_on_streams(rs, orc, cases, rounds):
    """cases: (kind, kb, vb, desc, n, seed); every case sorts on its own stream,
    all launched before any is waited for, `rounds` times over."""
    dev = torch.device("cuda")
    streams = [torch.cuda.Stream() for _ in cases]
    prepared = []
    for kind, kb, vb, desc, n, seed in cases:
        keys = random_keys(kind, kb, n, seed=seed)
        vals = index_values(vb, n) if vb else None
        d = rs.make_desc(kind, kb, vb, desc)
        tb = rs.sort_raw(d, None, None, None, 0, n)
        ek, ev = orc.sort(keys, vals, kind, kb, desc, vb=vb if vb else None)
        prepared.append((d, n, tb, keys, vals, ek, ev, torch.empty(tb, dtype=torch.uint8, device=dev)))
    for r in range(rounds):
        bufs = []
        for (d, n, tb, keys, vals, ek, ev, ws) in prepared:
            dk = torch.from_numpy(keys.copy()).to(dev)
            dv = torch.from_numpy(vals.copy()).to(dev) if vals is not None else None
            bufs.append((dk, dv))
        torch.cuda.synchronize()
        print(f"[round {r}] {len(cases)} sorts on {len(streams)} streams", flush=True)
        for (d, n, tb, *_rest, ws), (dk, dv), st in zip(prepared, bufs, streams):
            rs.sort_raw(d, dk.data_ptr(), dv.data_ptr() if dv is not None else None, ws.data_ptr(), tb, n, st)
        torch.cuda.synchronize()
        for i, ((d, n, tb, keys, vals, ek, ev, ws), (dk, dv)) in enumerate(zip(prepared, bufs)):
            assert np.array_equal(dk.cpu().numpy(), ek), f"round {r} sort {i}: keys differ from the oracle"
            if dv is not None:
                assert np.array_equal(dv.cpu().numpy(), ev), f"round {r} sort {i}: values differ from the oracle"


def test_overlapping_sorts_u32_pairs(rs, orc):
    # 2^21 + 5 keys: 128+ tiles of 16K keys, so every sort's passes fill the device
    n = (1 << 21) + 5
    cases = [(0, 4, 4, False, n, 11), (0, 4, 4, True, n, 12), (0, 4, 0, False, n, 13), (1, 4, 4, False, n, 14)]
    _sorts_on_streams(rs, orc, cases, rounds=3)


def test_overlapping_sorts_mixed_widths(rs, orc):
    cases = [(0, 8, 8, False, 600_001, 21), (2, 4, 0, False, 1_500_007, 22), (0, 4, 4, False, 3_000_001, 23),
             (0, 16, 8, True, 200_003, 24), (1, 8, 0, True, 900_009, 25)]
    _sorts_on_streams(rs, orc, cases, rounds=2)


def test_sort_alone_after_overlap(rs, orc):
    n = (1 << 22) + 77
    _sorts_on_streams(rs, orc, [(0, 4, 4, False, n, 31)], rounds=2)


def test_big_sort_beside_small_sorts(rs, orc):
    """A large sort (its passes on the per-XCD tickets when it takes the lock)
    launched first, then six small sorts whose histograms clear their control
    headers themselves, each on its own stream: every result bit-exact."""
    cases = [(0, 4, 4, False, 1 << 25, 41)] + [(0, 4, 0, False, (1 << 20) + 13 * i, 42 + i) for i in range(6)]
    _sorts_on_streams(rs, orc, cases, rounds=2)
