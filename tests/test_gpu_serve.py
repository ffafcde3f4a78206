"""GPU tests of the serving loop: every request's result is identical alone
(the oracle's sort of that request, which the whole-array GPU sort matches)
and in the batched mix; admission limits; graph and direct launch paths."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.init()


def _requests(orc, count, seed, max_len=5000, long_every=0):
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.integers(0, max_len, size=count)]
    if long_every:
        for i in range(0, count, long_every):
            sizes[i] = int(rng.integers(16385, 90000))
    keys = [orc.mt19937(seed * 1000 + i, m) for i, m in enumerate(sizes)]
    return keys


@pytest.mark.parametrize("mode", ["direct", "graph", "batched"])
@pytest.mark.parametrize("desc", [False, True])
def test_alone_equals_mixed(rs, orc, mode, desc):
    print(f"[case] serve mode={mode} desc={desc}", flush=True)
    host = _requests(orc, 300, seed=3 + int(desc), long_every=37)
    dk = [torch.from_numpy(k.copy()).cuda() for k in host]
    dv = [torch.arange(k.size, dtype=torch.int32, device="cuda") for k in host]
    torch.cuda.synchronize()
    with rs.Server(rs.KeyType(0, 4), value_bytes=4, descending=desc, batch_keys=1 << 20, batch_requests=64,
                   window_us=100, depth=3, mode=mode) as srv:
        ids = [srv.submit(k, v) for k, v in zip(dk, dv)]
        infos = [srv.wait(i) for i in ids]
        st = srv.stats()
        assert st.completed == len(ids) and st.pending == 0
        assert st.graphs == int(mode == "graph") and st.direct == int(mode == "direct")
        assert st.batches >= 2  # 300 requests cannot share one 64-request slot
    for k, v, d_k, d_v, info in zip(host, [np.arange(k.size, dtype=np.int32) for k in host], dk, dv, infos):
        ek, ev = orc.sort(k, v, 0, 4, desc, vb=4)
        assert np.array_equal(d_k.cpu().numpy(), ek)
        assert np.array_equal(d_v.cpu().numpy(), ev)
        assert info.n == k.size and info.latency_us >= info.queue_us


def test_same_request_alone_and_in_mix(rs, orc):
    """One request served alone (a batch of one) and inside a mix of 200."""
    probe = orc.mt19937(99, 7777)
    others = _requests(orc, 200, seed=5)
    with rs.Server(rs.KeyType(0, 4), batch_keys=1 << 21, batch_requests=256, window_us=50, depth=2) as srv:
        a = torch.from_numpy(probe.copy()).cuda()
        srv.wait(srv.submit(a))
        b = torch.from_numpy(probe.copy()).cuda()
        dothers = [torch.from_numpy(k.copy()).cuda() for k in others]
        ids = [srv.submit(x) for x in dothers[:100]] + [srv.submit(b)] + [srv.submit(x) for x in dothers[100:]]
        infos = {i: srv.wait(i) for i in ids}
        assert infos[ids[100]].batch != 0
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    assert np.array_equal(a.cpu().numpy(), np.sort(probe))


@pytest.mark.parametrize("mode", ["direct", "graph"])
def test_key_kinds_and_values(rs, orc, mode):
    from helpers import random_keys, random_values

    for kind, kb, vb in ((2, 4, 8), (1, 8, 4), (0, 2, 0), (2, 8, 12), (0, 16, 16), (1, 1, 2)):
        print(f"[case] serve mode={mode} kind={kind} kb={kb} vb={vb}", flush=True)
        rng = np.random.default_rng(kb + vb)
        sizes = [int(x) for x in rng.integers(1, 20000, size=40)]
        host_k = [random_keys(kind, kb, m, seed=i) for i, m in enumerate(sizes)]
        host_v = [random_values(vb, m, seed=i) if vb else None for i, m in enumerate(sizes)]
        dk = [torch.from_numpy(k.copy()).cuda() for k in host_k]
        dv = [torch.from_numpy(v.copy()).cuda() if v is not None else None for v in host_v]
        with rs.Server(rs.KeyType(kind, kb), value_bytes=vb, batch_keys=1 << 19, batch_requests=32,
                       window_us=20, depth=2, mode=mode) as srv:
            ids = [srv.submit(k, v, n=m) for k, v, m in zip(dk, dv, sizes)]
            for i in ids:
                srv.wait(i)
        for k, v, d_k, d_v in zip(host_k, host_v, dk, dv):
            ek, ev = orc.sort(k, v, kind, kb, False, vb=vb if vb else None)
            assert np.array_equal(d_k.cpu().numpy(), ek)
            if vb:
                assert np.array_equal(d_v.cpu().numpy(), ev)


def test_admission_limits(rs):
    with rs.Server(rs.KeyType(0, 4), batch_keys=4096, batch_requests=8, window_us=1_000_000, depth=1,
                   max_pending=4) as srv:
        big = torch.zeros(5000, dtype=torch.int32, device="cuda")
        with pytest.raises(rs.RsortError) as e:
            srv.submit(big)  # larger than a batch slot
        assert e.value.status == 2
        small = [torch.randint(0, 100, (10,), dtype=torch.int32, device="cuda") for _ in range(5)]
        ids = [srv.submit(x) for x in small[:4]]
        with pytest.raises(rs.RsortError) as e:
            srv.submit(small[4])  # queue full: the window holds the batch open
        assert e.value.status == 5
        srv.flush()
        for i in ids:
            assert srv.wait(i, timeout_us=0).n == 10
        for x in small[:4]:
            v = x.cpu().numpy()
            assert np.all(v[1:] >= v[:-1])


def test_flush_empty_and_zero_length(rs):
    with rs.Server(rs.KeyType(0, 4), batch_keys=1 << 16, batch_requests=16, window_us=10, depth=2) as srv:
        srv.flush()
        z = torch.empty(0, dtype=torch.int32, device="cuda")
        one = torch.tensor([5], dtype=torch.int32, device="cuda")
        i0, i1 = srv.submit(z), srv.submit(one)
        assert srv.wait(i0).n == 0 and srv.wait(i1).n == 1
        assert one.item() == 5


def test_submit_many(rs, orc):
    host = _requests(orc, 120, seed=8, max_len=20000)
    dk = [torch.from_numpy(k.copy()).cuda() for k in host]
    dv = [torch.arange(k.size, dtype=torch.int32, device="cuda") for k in host]
    with rs.Server(rs.KeyType(0, 4), value_bytes=4, batch_keys=1 << 20, batch_requests=64, depth=2) as srv:
        ids = srv.submit_many([k.data_ptr() for k in dk], [v.data_ptr() for v in dv], [k.size for k in host])
        for i in ids:
            srv.wait(i)
    for k, d_k, d_v in zip(host, dk, dv):
        ek, ev = orc.sort(k, np.arange(k.size, dtype=np.int32), 0, 4, False, vb=4)
        assert np.array_equal(d_k.cpu().numpy(), ek) and np.array_equal(d_v.cpu().numpy(), ev)


@pytest.mark.parametrize("kind,kb,vb,desc", [(0, 4, 4, False), (1, 4, 0, True), (0, 8, 8, False), (2, 8, 4, True)])
def test_many_long_requests_in_one_batch(rs, orc, kind, kb, vb, desc):
    """One direct-form batch with a dozen requests longer than every LDS bin
    (copied into the batch buffer and sorted by one segmented sort), short
    requests in every bin beside them, and one batch with a single long
    request (sorted in place); each request equals the oracle's sort."""
    from helpers import random_keys, random_values

    print(f"[case] serve long batch kind={kind} kb={kb} vb={vb} desc={desc}", flush=True)
    rng = np.random.default_rng(40 + kb + vb)
    sizes = [int(x) for x in rng.integers(16385, 120000, size=12)]
    sizes += [int(x) for x in rng.integers(0, 16384, size=60)]
    rng.shuffle(sizes)
    host_k = [random_keys(kind, kb, m, seed=100 + i) for i, m in enumerate(sizes)]
    host_v = [random_values(vb, m, seed=200 + i) if vb else None for i, m in enumerate(sizes)]
    dk = [torch.from_numpy(k.copy()).cuda() for k in host_k]
    dv = [torch.from_numpy(v.copy()).cuda() if vb else None for v in host_v]
    lone_k = random_keys(kind, kb, 50000, seed=7)
    lone_v = random_values(vb, 50000, seed=8) if vb else None
    dlk = torch.from_numpy(lone_k.copy()).cuda()
    dlv = torch.from_numpy(lone_v.copy()).cuda() if vb else None
    torch.cuda.synchronize()
    with rs.Server(rs.KeyType(kind, kb), value_bytes=vb, descending=desc, batch_keys=1 << 22, batch_requests=128,
                   depth=2, mode="direct") as srv:
        ids = srv.submit_many([k.data_ptr() for k in dk], [v.data_ptr() for v in dv] if vb else None, sizes)
        infos = [srv.wait(i) for i in ids]
        lone = srv.wait(srv.submit(dlk, dlv, n=50000))  # the helpers' arrays are bytes
        assert srv.stats().direct == 1
    assert len({x.batch for x in infos}) == 1  # one batch held them all
    assert lone.batch != infos[0].batch
    for i, (k, v, m) in enumerate(zip(host_k + [lone_k], host_v + [lone_v], sizes + [50000])):
        ek, ev = orc.sort(k, v, kind, kb, desc, vb=vb if vb else None)
        got_k = (dk + [dlk])[i].cpu().numpy()
        assert np.array_equal(got_k, ek), f"request {i} (n={m}) keys"
        if vb:
            assert np.array_equal((dv + [dlv])[i].cpu().numpy(), ev), f"request {i} (n={m}) values"


def test_small_slots_with_long_requests(rs, orc):
    """A server whose slots hold fewer requests than its long-request batch
    size (4): created without error, long requests sorted one by one."""
    print("[case] serve small slots", flush=True)
    rng = np.random.default_rng(77)
    sizes = [20000, 30000, 100, 17000, 5000, 40000, 3]
    host = [orc.mt19937(700 + i, m) for i, m in enumerate(sizes)]
    dk = [torch.from_numpy(k.copy()).cuda() for k in host]
    dv = [torch.from_numpy(rng.integers(0, 2**31, size=m, dtype=np.int64).astype(np.int32)).cuda() for m in sizes]
    hv = [v.cpu().numpy() for v in dv]
    with rs.Server(rs.KeyType(0, 4), value_bytes=4, batch_keys=1 << 17, batch_requests=4, depth=2,
                   mode="direct") as srv:
        ids = srv.submit_many([k.data_ptr() for k in dk], [v.data_ptr() for v in dv], sizes)
        for i in ids:
            srv.wait(i)
    for k, v, d_k, d_v in zip(host, hv, dk, dv):
        ek, ev = orc.sort(k, v, 0, 4, False, vb=4)
        assert np.array_equal(d_k.cpu().numpy(), ek) and np.array_equal(d_v.cpu().numpy(), ev)


def test_wait_many(rs, orc):
    """wait_many: BUSY while the batch is held open (nothing released), an
    unknown id refused, then every request's info at once after a flush."""
    print("[case] serve wait_many", flush=True)
    host = [orc.mt19937(900 + i, m) for i, m in enumerate([500, 0, 3000, 20000, 7])]
    dk = [torch.from_numpy(k.copy()).cuda() for k in host]
    with rs.Server(rs.KeyType(0, 4), batch_keys=1 << 20, batch_requests=64, window_us=5_000_000, depth=2) as srv:
        ids = [srv.submit(k) for k in dk]
        with pytest.raises(rs.RsortError) as e:
            srv.wait_many(ids, timeout_us=0)  # the window keeps the batch open
        assert e.value.status == 5
        with pytest.raises(rs.RsortError) as e:
            srv.wait_many(ids + [10**9])
        assert e.value.status == 1
        srv.flush()
        infos = srv.wait_many(ids)
        assert [x.n for x in infos] == [k.size for k in host]
        assert len({x.batch for x in infos}) == 1 and all(x.latency_us >= x.queue_us for x in infos)
        with pytest.raises(rs.RsortError):
            srv.wait(ids[0], timeout_us=0)  # released
    for k, d in zip(host, dk):
        assert np.array_equal(d.cpu().numpy(), np.sort(k))


def test_small_batch_bins_merged(rs, orc):
    """A batch of three requests from three LDS bins (run as one launch of
    the largest of them) and one of two requests: every request exact."""
    print("[case] serve merged bins", flush=True)
    for sizes in ([600, 3000, 12000], [5, 9000]):
        host = [orc.mt19937(1100 + m, m) for m in sizes]
        vals = [np.arange(m, dtype=np.int32) for m in sizes]
        dk = [torch.from_numpy(k.copy()).cuda() for k in host]
        dv = [torch.from_numpy(v.copy()).cuda() for v in vals]
        with rs.Server(rs.KeyType(0, 4), value_bytes=4, batch_keys=1 << 20, batch_requests=64, window_us=5_000_000,
                       depth=2) as srv:
            ids = srv.submit_many([k.data_ptr() for k in dk], [v.data_ptr() for v in dv], sizes)
            srv.flush()
            infos = srv.wait_many(ids)
            assert len({x.batch for x in infos}) == 1
        for k, v, d_k, d_v in zip(host, vals, dk, dv):
            ek, ev = orc.sort(k, v, 0, 4, False, vb=4)
            assert np.array_equal(d_k.cpu().numpy(), ek) and np.array_equal(d_v.cpu().numpy(), ev)
