"""The batched and served workload at its full size: 100 000 segments of u32
keys with u32 values, lognormal sizes (median 2048, sigma 0.7356, clipped to
[32, 2^22]; about 2^28 keys in all), keys uniform.

* Batch: one segmented sort replayed from its captured hipGraph
  (RSORT_FLAG_GRAPH), three calls (the first direct, then two replays) on
  fresh copies of the input.  Every call is checked on the device -- each
  segment in order, each value the index of a key of its own segment
  holding that key, the values of the batch a permutation, equal keys in
  index order -- against the direct call's result bit for bit, and 1200
  sampled segments against the oracle.
* Stream: requests drawn from the same size distribution arrive with
  Poisson gaps at 5000 per second through the serving loop; every result
  equals the oracle's sort of that request, and the requests served again
  alone (one at a time) give the same bytes as in the mix."""
from __future__ import annotations

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SEGMENTS = 100_000


def _device_checks(rs, keys0, k, v, off_d, seg_of_pos, n):
    """Device-side characterisation of a segmented u32/u32 sort (values =
    original index); returns a list of failed checks."""
    bad = []
    ku = k ^ (-(2**31))  # signed order of these words = unsigned order of the keys
    starts = torch.zeros(n, dtype=torch.bool, device="cuda")
    starts[off_d[1:-1]] = True
    if not bool(((ku[1:] >= ku[:-1]) | starts[1:]).all()):
        bad.append("a segment is out of order")
    vl = v.long()
    if not bool(((vl >= 0) & (vl < n)).all()):
        bad.append("value out of range")
        return bad
    if not bool((seg_of_pos[vl] == seg_of_pos).all()):
        bad.append("a value points outside its own segment")
    if not bool((keys0[vl] == k).all()):
        bad.append("a value does not point at its key")
    hits = torch.zeros(n, dtype=torch.int32, device="cuda")
    hits.index_add_(0, vl, torch.ones(n, dtype=torch.int32, device="cuda"))
    if not bool((hits == 1).all()):
        bad.append("values are not a permutation")
    eq = (k[1:] == k[:-1]) & ~starts[1:]
    if not bool((v[1:][eq] > v[:-1][eq]).all()):
        bad.append("equal keys out of index order (stability)")
    return bad


def test_batch_100k_segments_graph(rs, orc):
    sizes = orc.gen_lognormal_sizes(2024, SEGMENTS).astype(np.int64)
    n = int(sizes.sum())
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    print(f"[case] 100k segments n={n} (2^{np.log2(n):.3f})", flush=True)
    assert (1 << 27) < n < (1 << 29)
    g = torch.Generator(device="cuda")
    g.manual_seed(77)
    keys0 = torch.randint(-(2**31), 2**31 - 1, (n,), dtype=torch.int32, device="cuda", generator=g)
    vals0 = torch.arange(n, dtype=torch.int32, device="cuda")
    off_d = torch.from_numpy(off).cuda()
    seg_of_pos = torch.repeat_interleave(torch.arange(SEGMENTS, dtype=torch.int32, device="cuda"),
                                         torch.from_numpy(sizes).cuda())
    d = rs.make_desc(0, 4, 4, False, rs.FLAG_GRAPH)
    tb = rs.segmented_sort_raw(d, None, None, n, None, 8, SEGMENTS, None, 0)
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    k = torch.empty_like(keys0)
    v = torch.empty_like(vals0)
    rng = np.random.default_rng(5)
    sample = np.sort(rng.choice(SEGMENTS, size=1200, replace=False))
    first_k = first_v = None
    for call in range(3):
        print(f"[case] 100k segments graph call={call}", flush=True)
        k.copy_(keys0)
        v.copy_(vals0)
        torch.cuda.synchronize()
        rs.segmented_sort_raw(d, k.data_ptr(), v.data_ptr(), n, off_d.data_ptr(), 8, SEGMENTS, ws.data_ptr(), tb)
        torch.cuda.synchronize()
        bad = _device_checks(rs, keys0, k, v, off_d, seg_of_pos, n)
        assert not bad, f"call {call}: {bad}"
        if first_k is None:
            first_k, first_v = k.clone(), v.clone()
        else:
            assert torch.equal(k, first_k) and torch.equal(v, first_v), f"replay {call} differs from the direct call"
        # sampled segments against the oracle
        hk = keys0.cpu().numpy().view(np.uint32)
        gk, gv = k.cpu().numpy().view(np.uint32), v.cpu().numpy().view(np.uint32)
        pk = np.concatenate([hk[off[s]:off[s + 1]] for s in sample])
        pv = np.concatenate([np.arange(off[s], off[s + 1], dtype=np.uint32) for s in sample])
        poff = np.concatenate([[0], np.cumsum(sizes[sample])]).astype(np.uint64)
        ek, ev = orc.segmented_sort(pk, pv, poff, 0, 4, False, vb=4)
        got_k = np.concatenate([gk[off[s]:off[s + 1]] for s in sample])
        got_v = np.concatenate([gv[off[s]:off[s + 1]] for s in sample])
        assert np.array_equal(got_k, ek), f"call {call}: sampled segments' keys differ from the oracle"
        assert np.array_equal(got_v, ev), f"call {call}: sampled segments' values differ from the oracle"


@pytest.mark.parametrize("mode", ["direct", "graph"])
def test_poisson_stream_5k_per_s(rs, orc, mode):
    count, rate = 2000, 5000.0
    sizes = orc.gen_lognormal_sizes(99, count).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(off[-1])
    host_k = orc.mt19937(1234, n)
    host_v = np.arange(n, dtype=np.uint32)
    ek, ev = orc.segmented_sort(host_k, host_v, off.astype(np.uint64), 0, 4, False, vb=4)
    keys = torch.from_numpy(host_k.view(np.int32).copy()).cuda()
    vals = torch.from_numpy(host_v.view(np.int32).copy()).cuda()
    torch.cuda.synchronize()
    kp, vp = keys.data_ptr(), vals.data_ptr()
    gaps = np.random.default_rng(7).exponential(1.0 / rate, size=count)
    print(f"[case] poisson stream mode={mode} requests={count} keys={n} max={int(sizes.max())}", flush=True)
    with rs.Server(rs.KeyType(0, 4), value_bytes=4, batch_keys=1 << 22, batch_requests=2048, window_us=0, depth=3,
                   mode=mode) as srv:
        ids = []
        t_next = time.perf_counter()
        for i in range(count):
            t_next += gaps[i]
            while time.perf_counter() < t_next:
                pass
            ids.append(srv.submit(_View(kp + 4 * int(off[i])), _View(vp + 4 * int(off[i])), n=int(sizes[i])))
        infos = [srv.wait(i) for i in ids]
        st = srv.stats()
        assert st.completed >= count and st.batches >= 2
        mixed_k, mixed_v = keys.cpu().numpy().view(np.uint32), vals.cpu().numpy().view(np.uint32)
        assert np.array_equal(mixed_k, ek), "a request in the stream differs from the oracle (keys)"
        assert np.array_equal(mixed_v, ev), "a request in the stream differs from the oracle (values)"
        lat = np.array([x.latency_us for x in infos])
        print(f"[stream] p50 {np.percentile(lat, 50):.0f} us p99 {np.percentile(lat, 99):.0f} us "
              f"batches {st.batches}", flush=True)
        # the same requests served alone, one at a time
        alone = list(range(0, count, 40))
        ak = torch.from_numpy(host_k.view(np.int32).copy()).cuda()
        av = torch.from_numpy(host_v.view(np.int32).copy()).cuda()
        torch.cuda.synchronize()
        for i in alone:
            o = int(off[i])
            srv.wait(srv.submit(_View(ak.data_ptr() + 4 * o), _View(av.data_ptr() + 4 * o), n=int(sizes[i])))
        alone_k, alone_v = ak.cpu().numpy().view(np.uint32), av.cpu().numpy().view(np.uint32)
        for i in alone:
            s = slice(int(off[i]), int(off[i + 1]))
            assert np.array_equal(alone_k[s], mixed_k[s]) and np.array_equal(alone_v[s], mixed_v[s]), \
                f"request {i}: alone differs from the mix"


class _View:
    """A device address standing in for a tensor (the server reads data_ptr)."""

    def __init__(self, ptr):
        self.ptr = ptr

    def data_ptr(self):
        return self.ptr
