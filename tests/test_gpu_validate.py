"""GPU tests of the device generator and checks (rsort_init_keys,
rsort_check_sorted, rsort_check_pairings, rsort_digit_counts) against the CPU
oracle, and their use at sizes the oracle cannot check quickly: the 64-bit
(wide) lookback and offset words above 2^29 - 1 keys."""
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


def dev(a: np.ndarray):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1).copy()).cuda()


@pytest.mark.parametrize("kb", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("mode", [1, 2, 3])
def test_init_keys_matches_reference_generator(rs, orc, kb, mode):
    n = 10007
    for i in (0, 3):
        seed = rs.set_seed_radix(i)
        assert seed == orc.set_seed_radix(i)
        out = torch.empty(n * kb, dtype=torch.uint8, device="cuda")
        rs.init_keys(out, kb, mode, seed, n)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), orc.init_keys(kb, mode, seed, n))


@pytest.mark.parametrize("kind,kb", [(0, 4), (1, 4), (2, 4), (2, 8), (0, 16), (2, 16), (1, 2), (2, 2), (0, 1)])
@pytest.mark.parametrize("desc", [False, True])
def test_check_sorted(rs, orc, kind, kb, desc):
    n = 50000
    keys = random_keys(kind, kb, n, seed=kb * 3 + kind)
    sk, _ = orc.sort(keys, None, kind, kb, desc)
    d = rs.make_desc(kind, kb, 0, desc)
    assert rs.check_sorted(dev(sk), d, n) == 0
    assert orc.check_sorted(sk, kind, kb, desc)
    # a planted inversion between two keys with different, non-NaN words
    w = sk.reshape(n, kb).copy()
    i = n // 3
    j = next(k for k in range(i + 1, n) if not np.array_equal(w[k], w[i]))
    w[[i, j]] = w[[j, i]]
    bad = w.reshape(-1)
    assert not orc.check_sorted(bad, kind, kb, desc) or kind == 2  # NaN neighbours may hide it
    got = rs.check_sorted(dev(bad), d, n)
    assert (got > 0) == (not orc.check_sorted(bad, kind, kb, desc))


def test_check_sorted_counts_pairs(rs):
    a = np.array([1, 2, 3, 2, 5, 4, 4, 0], dtype=np.uint32)  # 3 > 2, 5 > 4, 4 > 0
    assert rs.check_sorted(dev(a), rs.make_desc(0, 4, 0), a.size) == 3
    f = np.array([0.0, np.nan, -1.0, 2.0, 1.0], dtype=np.float32)  # NaN never counts; 2 > 1
    assert rs.check_sorted(dev(f), rs.make_desc(2, 4, 0), f.size) == 1


@pytest.mark.parametrize("kb,vb", [(4, 4), (8, 8), (16, 8), (2, 4), (1, 4)])
def test_check_pairings(rs, orc, kb, vb):
    n = 40000
    keys = random_keys(0, kb, n, seed=kb + vb)
    vals = np.arange(n, dtype=np.uint32 if vb == 4 else np.uint64)
    sk, sv = orc.sort(keys, vals, 0, kb, False, vb=vb)
    ko, ks, vs = dev(keys), dev(sk), dev(sv)
    assert rs.check_pairings(ko, ks, vs, kb, vb, n) == 0
    assert orc.check_pairings(keys, sk, sv, kb, vb)
    # a value pointing at another key, and an equal-key pair out of input order
    v2 = sv.view(vals.dtype).copy()
    v2[10], v2[20000] = v2[20000], v2[10]
    assert rs.check_pairings(ko, ks, dev(v2), kb, vb, n) >= 1
    words = sk.reshape(n, kb)
    dup = next(i for i in range(n - 1) if np.array_equal(words[i], words[i + 1]))
    v3 = sv.view(vals.dtype).copy()
    v3[dup], v3[dup + 1] = v3[dup + 1], v3[dup]
    assert rs.check_pairings(ko, ks, dev(v3), kb, vb, n) == 1
    assert not orc.check_pairings(keys, sk, v3, kb, vb)


@pytest.mark.parametrize("kind,kb", [(0, 4), (2, 4), (1, 8), (2, 16), (0, 1)])
@pytest.mark.parametrize("desc", [False, True])
def test_digit_counts(rs, orc, kind, kb, desc):
    n = 70001
    keys = random_keys(kind, kb, n, seed=kb + kind)
    got = rs.digit_counts(dev(keys), rs.make_desc(kind, kb, 0, desc), n).cpu().numpy()
    assert np.array_equal(got.astype(np.uint64), orc.digit_counts(keys, kind, kb, desc))


def test_wide_words_above_2p29(rs):
    """n = 2^29 + 4099 pairs: 64-bit lookback words and offsets.  Checked on
    the device: sorted, every value pairs with its original key, equal keys in
    input order, and the digit histogram of every pass unchanged."""
    n = (1 << 29) + 4099
    d = rs.make_desc(0, 4, 4)
    assert rs.plan(d, n).wide == 1
    print(f"[case] wide sort_pairs n={n}", flush=True)
    keys = torch.empty(n, dtype=torch.int32, device="cuda")
    rs.init_keys(keys, 4, 1, rs.set_seed_radix(5), n)
    keys.bitwise_and_(0x00FFFFFF)  # 2^24 distinct keys: many equal keys (stability)
    orig = keys.clone()
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    before = rs.digit_counts(keys, d, n)
    tb = rs.sort_pairs(keys, vals, None, None, n)
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    rs.sort_pairs(keys, vals, tb, ws, n)
    assert rs.check_sorted(keys, d, n) == 0
    assert rs.check_pairings(orig, keys, vals, 4, 4, n) == 0
    assert torch.equal(rs.digit_counts(keys, d, n), before)


def test_wide_segmented_above_2p29(rs):
    n = (1 << 29) + 777
    d = rs.make_desc(1, 4, 0)  # int32 keys: signed order
    print(f"[case] wide segmented n={n}", flush=True)
    keys = torch.empty(n, dtype=torch.int32, device="cuda")
    rs.init_keys(keys, 4, 1, 11, n)
    cuts = [0, 1000, 5000, (1 << 28) + 3, n]
    off = torch.tensor(cuts, dtype=torch.int64, device="cuda")
    before = rs.digit_counts(keys, d, n)
    tb = rs.segmented_sort(keys, off, None, None)
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    rs.segmented_sort(keys, off, tb, ws)
    for b, e in zip(cuts[:-1], cuts[1:]):
        assert rs.check_sorted(keys[b:e], d, e - b) == 0
    assert torch.equal(rs.digit_counts(keys, d, n), before)
