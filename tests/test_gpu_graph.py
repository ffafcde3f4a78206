"""Graph replay (RSORT_FLAG_GRAPH): the first call with a set of arguments
sorts directly and captures its launch sequence; later identical calls
replay the captured graph.  Every call here gets a fresh copy of the input
written by eager work of the test (a torch copy) with a synchronisation in
between, and every call's result is checked against the oracle -- the first
(direct) call and each replay alike."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import index_values, random_keys, random_values

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

REPLAYS = 4


def _u8(a):
    return np.ascontiguousarray(a).view(np.uint8).reshape(-1)


@pytest.mark.parametrize("kind,kb,vb,desc,n", [
    (0, 4, 4, False, (1 << 22) + 3),   # per-XCD tickets, early aggregate off, lookback groups
    (0, 4, 0, True, 1 << 20),          # segmented lookback walk (few tiles)
    (2, 8, 8, False, 700_001),
    (1, 16, 8, True, 300_007),
    (0, 1, 4, False, 500_000),         # one pass: the copy-back kernel
    (0, 4, 12, False, 200_003),        # values moved by index (iota, gather, copy kernels)
    (0, 2, 0, False, 9_000),           # one histogram workgroup: it clears the state itself
])
def test_sort_graph_replays(rs, orc, kind, kb, vb, desc, n):
    keys = _u8(random_keys(kind, kb, n, seed=n % 1000 + kb))
    vals = _u8(index_values(vb, n)) if vb in (4, 8) else (_u8(random_values(vb, n, seed=vb)) if vb else None)
    ek, ev = orc.sort(keys, vals, kind, kb, desc, vb=vb if vb else None)
    d = rs.make_desc(kind, kb, vb, desc, rs.FLAG_GRAPH)
    tb = rs.sort_raw(d, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    src_k = torch.from_numpy(keys.copy()).cuda()
    src_v = torch.from_numpy(vals.copy()).cuda() if vals is not None else None
    dk = torch.empty_like(src_k)
    dv = torch.empty_like(src_v) if src_v is not None else None
    for r in range(REPLAYS):
        print(f"[case] graph sort kind={kind} kb={kb} vb={vb} desc={int(desc)} n={n} call={r}", flush=True)
        dk.copy_(src_k)
        if dv is not None:
            dv.copy_(src_v)
        torch.cuda.synchronize()
        rs.sort_raw(d, dk.data_ptr(), dv.data_ptr() if dv is not None else None, ws.data_ptr(), tb, n)
        torch.cuda.synchronize()
        assert np.array_equal(dk.cpu().numpy(), ek), f"call {r}: keys differ from the oracle"
        if vb:
            assert np.array_equal(dv.cpu().numpy(), ev), f"call {r}: values differ from the oracle"


def _offsets(lengths):
    return np.concatenate([[0], np.cumsum(lengths)]).astype(np.uint64)


@pytest.mark.parametrize("kind,kb,vb,desc", [(0, 4, 4, False), (1, 8, 8, True), (2, 4, 12, False), (0, 16, 0, False)])
def test_segmented_graph_replays(rs, orc, kind, kb, vb, desc):
    rng = np.random.default_rng(kb * 5 + vb)
    lengths = [int(x) for x in orc.gen_lognormal_sizes(3 + kb, 2000)] + [0, 1, 40_000, 90_001]
    rng.shuffle(lengths)
    n = int(sum(lengths))
    off = _offsets(lengths)
    keys = _u8(random_keys(kind, kb, n, seed=kb + vb))
    vals = _u8(index_values(vb, n)) if vb in (4, 8) else (_u8(random_values(vb, n, seed=vb)) if vb else None)
    ek, ev = orc.segmented_sort(keys, vals, off, kind, kb, desc, vb=vb if vb else None)
    d = rs.make_desc(kind, kb, vb, desc, rs.FLAG_GRAPH)
    doff = torch.from_numpy(off.astype(np.int64)).cuda()
    tb = rs.segmented_sort_raw(d, None, None, n, None, 8, len(lengths), None, 0)
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    src_k = torch.from_numpy(keys.copy()).cuda()
    src_v = torch.from_numpy(vals.copy()).cuda() if vals is not None else None
    dk = torch.empty_like(src_k)
    dv = torch.empty_like(src_v) if src_v is not None else None
    for r in range(REPLAYS):
        print(f"[case] graph segmented kind={kind} kb={kb} vb={vb} n={n} call={r}", flush=True)
        dk.copy_(src_k)
        if dv is not None:
            dv.copy_(src_v)
        torch.cuda.synchronize()
        rs.segmented_sort_raw(d, dk.data_ptr(), dv.data_ptr() if dv is not None else None, n, doff.data_ptr(), 8,
                              len(lengths), ws.data_ptr(), tb)
        torch.cuda.synchronize()
        assert np.array_equal(dk.cpu().numpy(), ek), f"call {r}: keys differ from the oracle"
        if vb:
            assert np.array_equal(dv.cpu().numpy(), ev), f"call {r}: values differ from the oracle"


def test_graph_cache_many_argument_sets(rs, orc):
    """More argument sets than the cache keeps (16): old graphs are dropped
    after their last replay and every call stays correct."""
    n = 50_000
    keys = _u8(orc.mt19937(5, n))
    ek, _ = orc.sort(keys, None, 0, 4, False)
    d = rs.make_desc(0, 4, 0, False, rs.FLAG_GRAPH)
    tb = rs.sort_raw(d, None, None, None, 0, n)
    bufs = [torch.empty(n * 4, dtype=torch.uint8, device="cuda") for _ in range(20)]
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    src = torch.from_numpy(keys.copy()).cuda()
    for rnd in range(3):
        for i, b in enumerate(bufs):
            b.copy_(src)
            rs.sort_raw(d, b.data_ptr(), None, ws.data_ptr(), tb, n)
        torch.cuda.synchronize()
        for i, b in enumerate(bufs):
            assert np.array_equal(b.cpu().numpy(), ek), f"round {rnd} buffer {i}"


def test_segmented_in_callers_capture(rs, orc):
    """A caller captures a segmented sort (no RSORT_FLAG_GRAPH) into its own
    graph: the fork onto the library's side streams and the join back become
    branches of that graph.  Every replay, with eager refills and
    synchronisations between them, equals the oracle."""
    lengths = [int(x) for x in orc.gen_lognormal_sizes(21, 3000)] + [0, 1, 40_000]
    n = int(sum(lengths))
    off = _offsets(lengths)
    keys = _u8(random_keys(0, 4, n, seed=21))
    vals = _u8(index_values(4, n))
    ek, ev = orc.segmented_sort(keys, vals, off, 0, 4, False, vb=4)
    d = rs.make_desc(0, 4, 4, False)
    doff = torch.from_numpy(off.astype(np.int64)).cuda()
    tb = rs.segmented_sort_raw(d, None, None, n, None, 8, len(lengths), None, 0)
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    src_k, src_v = torch.from_numpy(keys.copy()).cuda(), torch.from_numpy(vals.copy()).cuda()
    dk, dv = src_k.clone(), src_v.clone()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    print(f"[case] caller capture segmented n={n} segments={len(lengths)}", flush=True)
    with torch.cuda.graph(g):
        rs.segmented_sort_raw(d, dk.data_ptr(), dv.data_ptr(), n, doff.data_ptr(), 8, len(lengths), ws.data_ptr(), tb,
                              torch.cuda.current_stream())
    for r in range(REPLAYS):
        dk.copy_(src_k)
        dv.copy_(src_v)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(dk.cpu().numpy(), ek), f"replay {r}: keys differ from the oracle"
        assert np.array_equal(dv.cpu().numpy(), ev), f"replay {r}: values differ from the oracle"


def test_segmented_from_two_host_threads(rs, orc):
    """Two host threads sort segmented batches at once, each on its own stream
    and workspace, through the shared side streams: every result equals the
    oracle."""
    import threading

    cases = []
    for t in range(2):
        lengths = [int(x) for x in orc.gen_lognormal_sizes(30 + t, 1500)] + [25_000 + t]
        n = int(sum(lengths))
        off = _offsets(lengths)
        keys = _u8(random_keys(0, 4, n, seed=30 + t))
        vals = _u8(index_values(4, n))
        ek, ev = orc.segmented_sort(keys, vals, off, 0, 4, False, vb=4)
        d = rs.make_desc(0, 4, 4, False)
        tb = rs.segmented_sort_raw(d, None, None, n, None, 8, len(lengths), None, 0)
        cases.append(dict(n=n, segs=len(lengths), d=d, tb=tb, ek=ek, ev=ev,
                          doff=torch.from_numpy(off.astype(np.int64)).cuda(),
                          ws=torch.empty(tb, dtype=torch.uint8, device="cuda"),
                          src_k=torch.from_numpy(keys.copy()).cuda(), src_v=torch.from_numpy(vals.copy()).cuda(),
                          stream=torch.cuda.Stream()))
    errors = []

    def worker(c):
        try:
            for _ in range(6):
                k, v = c["src_k"].clone(), c["src_v"].clone()
                torch.cuda.synchronize()
                rs.segmented_sort_raw(c["d"], k.data_ptr(), v.data_ptr(), c["n"], c["doff"].data_ptr(), 8, c["segs"],
                                      c["ws"].data_ptr(), c["tb"], c["stream"])
                c["stream"].synchronize()
                if not (np.array_equal(k.cpu().numpy(), c["ek"]) and np.array_equal(v.cpu().numpy(), c["ev"])):
                    errors.append(f"n={c['n']}: result differs from the oracle")
        except Exception as e:  # reported below
            errors.append(repr(e))

    print("[case] segmented sorts from two host threads", flush=True)
    threads = [threading.Thread(target=worker, args=(c,)) for c in cases]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
