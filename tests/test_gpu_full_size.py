"""GPU parity at the workload configurations' full sizes, where the CPU oracle
is too slow to compare whole outputs: inputs from the oracle's generators (the
configurations' exact definitions), checks that do not depend on the size:
  * key-only sorts: the output equals, bit for bit, torch.sort of the input
    mapped to its order-preserving unsigned word (a library sort of a
    bijective image; stability is invisible without values);
  * key-value sorts: sorted (device check), every value is the index of an
    input key equal to the output key with equal keys in input order
    (check_pairings), and the values are a permutation of 0..n-1;
  * every sort: per-pass digit histograms unchanged.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SIGN64 = -(2**63)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.init()


def sort_dev(rs, keys, vals, kind, kb, vb, n, desc=False):
    d = rs.make_desc(kind, kb, vb, desc)
    print(f"[case] full-size kind={kind} kb={kb} vb={vb} desc={int(desc)} n={n}", flush=True)
    before = rs.digit_counts(keys, d, n)
    tb = rs.sort_raw(d, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    rs.sort_raw(d, keys.data_ptr(), vals.data_ptr() if vals is not None else None, ws.data_ptr(), tb, n)
    torch.cuda.synchronize()
    del ws
    assert rs.check_sorted(keys, d, n) == 0
    assert torch.equal(rs.digit_counts(keys, d, n), before)
    return d


def u64_order_word(x):
    """int64 view of uint64 keys -> int64 whose signed order is the unsigned order."""
    return x ^ SIGN64


def f32_order_word(x):
    """float32 bits -> int64 in [0, 2^32) ordered as the floats (sign-magnitude twiddle)."""
    b = x.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    neg = (b >> 31) == 1
    return torch.where(neg, b ^ 0xFFFFFFFF, b ^ 0x80000000)


@pytest.mark.parametrize("desc", [False, True])
def test_kv_u32_2p28(rs, orc, desc):
    """2^28 uint32 keys = mt19937 outputs, values = index, both orders."""
    n = 1 << 28
    keys = torch.from_numpy(orc.mt19937(1, n).view(np.int32)).cuda()
    orig = keys.clone()
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    sort_dev(rs, keys, vals, 0, 4, 4, n, desc)
    assert rs.check_pairings(orig, keys, vals, 4, 4, n) == 0
    assert torch.equal(torch.sort(vals)[0], torch.arange(n, dtype=torch.int32, device="cuda"))


def test_u64_low16_2p27(rs, orc):
    """6 of 8 passes have a single digit (skipped on the device)."""
    n = 1 << 27
    keys = torch.from_numpy(orc.gen_u64_low16(3, n).view(np.int64)).cuda()
    expect = torch.sort(keys)[0]
    sort_dev(rs, keys, None, 0, 8, 0, n)
    assert torch.equal(keys, expect)


@pytest.mark.parametrize("desc", [False, True])
def test_zipf_u64_2p27(rs, orc, desc):
    n = 1 << 27
    keys = torch.from_numpy(orc.gen_zipf_u64(4, n).view(np.int64)).cuda()
    expect = u64_order_word(torch.sort(u64_order_word(keys), descending=desc)[0])
    sort_dev(rs, keys, None, 0, 8, 0, n, desc)
    assert torch.equal(keys, expect)


def test_f32_specials_2p27(rs, orc):
    """N(0,1) with 1% +-0.0 and 0.1% +-inf: -inf < negatives < -0.0 < +0.0 < positives < +inf."""
    n = 1 << 27
    keys = torch.from_numpy(orc.gen_f32_normal_specials(5, n)).cuda()
    expect = torch.sort(f32_order_word(keys))[0]
    sort_dev(rs, keys, None, 2, 4, 0, n)
    assert torch.equal(f32_order_word(keys), expect)
    z = keys == 0
    assert bool(z.any())
    first_pos_zero = int(torch.nonzero(z & ~torch.signbit(keys))[0])
    assert not bool(torch.signbit(keys[first_pos_zero:])[z[first_pos_zero:]].any()), "-0.0 after +0.0"


def test_u128_pairs_2p26(rs, orc):
    """{lo, hi} 128-bit keys (hi over 2^20 values), uint64 index values, 16 passes."""
    n = 1 << 26
    keys = torch.from_numpy(orc.gen_u128_pairs(6, n).view(np.int64).reshape(-1)).cuda()
    orig = keys.clone()
    vals = torch.arange(n, dtype=torch.int64, device="cuda")
    sort_dev(rs, keys, vals, 0, 16, 8, n)
    assert rs.check_pairings(orig, keys, vals, 16, 8, n) == 0
    assert torch.equal(torch.sort(vals)[0], torch.arange(n, dtype=torch.int64, device="cuda"))


def test_u32_wide_offsets_2p29(rs, orc):
    """2^29 + 7 keys: past 2^29 - 1 the lookback words and offsets are 64-bit
    ("wide"); key-only and key-value, with the per-XCD tickets' partial last
    round of tiles."""
    n = (1 << 29) + 7
    assert rs.plan(rs.make_desc(0, 4, 0), n).wide == 1
    keys = torch.from_numpy(orc.mt19937(9, n).view(np.int32)).cuda()
    expect = torch.sort(keys.to(torch.int64) & 0xFFFFFFFF)[0]
    k1 = keys.clone()
    sort_dev(rs, k1, None, 0, 4, 0, n)
    assert torch.equal(k1.to(torch.int64) & 0xFFFFFFFF, expect)
    del k1, expect
    orig = keys.clone()
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    sort_dev(rs, keys, vals, 0, 4, 4, n)
    assert rs.check_pairings(orig, keys, vals, 4, 4, n) == 0
    assert torch.equal(torch.sort(vals)[0], torch.arange(n, dtype=torch.int32, device="cuda"))
