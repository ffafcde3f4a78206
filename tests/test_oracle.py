"""Pins the CPU oracle against golden fixtures, exhaustive tables, brute force
and closed-form facts.  CPU only."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import KEY_TYPES, golden_names, load_golden, random_keys, random_values


@pytest.mark.parametrize("name", golden_names())
@pytest.mark.parametrize("desc", [False, True])
def test_oracle_matches_golden(orc, name, desc):
    g = load_golden(name)
    kb = g["key_bytes"]
    perm = orc.permutation(g["key_bytes_array"], g["kind"], kb, desc)
    np.testing.assert_array_equal(perm, g["desc"] if desc else g["asc"])
    # the in-place sort carries values with the keys
    vals = np.arange(g["n"], dtype=np.uint32)
    k, v = orc.sort(g["key_bytes_array"], vals, g["kind"], kb, desc, vb=4)
    np.testing.assert_array_equal(v, perm.astype(np.uint32))
    exp_keys = g["key_bytes_array"].reshape(-1, kb)[perm.astype(np.int64)].reshape(-1)
    np.testing.assert_array_equal(k, exp_keys)


def _numeric(kind, kb, a_bytes):
    dt = {(0, 1): np.uint8, (1, 1): np.int8, (0, 2): np.uint16, (1, 2): np.int16, (2, 2): np.float16}[(kind, kb)]
    return a_bytes.view(dt)


@pytest.mark.parametrize("kind,kb", [(0, 1), (1, 1), (0, 2), (1, 2)])
def test_twiddle_exhaustive_integers(orc, kind, kb):
    """Every 8/16-bit value: the twiddled word orders like the integer, and twiddle_out inverts it."""
    allv = np.arange(1 << (8 * kb), dtype=np.uint64).astype({1: np.uint8, 2: np.uint16}[kb])
    raw = allv.view(np.uint8)
    for desc in (False, True):
        tw = orc.twiddle_in(raw, kind, kb, desc).view(allv.dtype)
        num = _numeric(kind, kb, raw).astype(np.int64)
        order_tw = np.argsort(tw, kind="stable")
        order_num = np.argsort(-num if desc else num, kind="stable")
        np.testing.assert_array_equal(order_tw, order_num)
        back = orc.twiddle_out(tw.view(np.uint8), kind, kb, desc)
        np.testing.assert_array_equal(back, raw)


def test_twiddle_exhaustive_half(orc):
    """Every IEEE half pattern: non-NaN values keep numeric order with -0 < +0;
    negative NaNs sort first, positive NaNs last."""
    allv = np.arange(1 << 16, dtype=np.uint32).astype(np.uint16)
    raw = allv.view(np.uint8)
    f = allv.view(np.float16).astype(np.float64)
    tw = orc.twiddle_in(raw, 2, 2, False).view(np.uint16).astype(np.int64)
    nan = np.isnan(f)
    neg = (allv >> 15) == 1
    finite_idx = np.where(~nan)[0]
    # numeric order, ties (+-0) broken by sign
    key = f[finite_idx] + 0.0
    sgn = np.where(neg[finite_idx], 0, 1)
    by_num = finite_idx[np.lexsort((sgn, key))]
    assert np.all(np.diff(tw[by_num]) > 0)
    assert tw[nan & neg].max() < tw[~nan].min()
    assert tw[nan & ~neg].min() > tw[~nan].max()
    np.testing.assert_array_equal(orc.twiddle_out(tw.astype(np.uint16).view(np.uint8), 2, 2, False), raw)


@pytest.mark.parametrize("kind,kb", KEY_TYPES)
@pytest.mark.parametrize("desc", [False, True])
def test_twiddle_roundtrip_random(orc, kind, kb, desc):
    keys = random_keys(kind, kb, 4000, seed=kb * 10 + kind)
    tw = orc.twiddle_in(keys, kind, kb, desc)
    np.testing.assert_array_equal(orc.twiddle_out(tw, kind, kb, desc), keys)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_float_order_against_numpy(orc, dt):
    """Sorted float keys (no NaN) are numerically nondecreasing with -0 before +0."""
    rng = np.random.default_rng(3)
    x = rng.standard_normal(20000).astype(dt)
    x[::97] = 0.0
    x[::89] = -0.0
    x[::101] = np.inf
    x[::103] = -np.inf
    x[::107] = np.finfo(dt).tiny / 4  # subnormal
    kb = x.itemsize
    k, _ = orc.sort(x.view(np.uint8), None, 2, kb, False)
    s = k.view(dt)
    assert np.all(s[1:] >= s[:-1])
    z = np.where(s == 0)[0]
    signs = np.signbit(s[z])
    assert np.all(signs[: signs.sum()]) and not np.any(signs[signs.sum():])
    np.testing.assert_array_equal(np.sort(x), s)  # numpy's value order agrees for non-NaN


@pytest.mark.parametrize("kind,kb", KEY_TYPES)
@pytest.mark.parametrize("desc", [False, True])
def test_stable_sort_brute_force(orc, kind, kb, desc):
    """Tiny inputs against Python's stable sorted() on the key's value."""
    n = 300
    keys = random_keys(kind, kb, n, seed=100 + kb + kind, dup_frac=0.5)
    kk = keys.reshape(n, kb)
    ints = [int.from_bytes(bytes(r), "little") for r in kk]
    bits = 8 * kb
    if kind == 0:
        val = ints
    elif kind == 1:
        val = [v - (1 << bits) if v >> (bits - 1) else v for v in ints]
    elif kb in (2, 4, 8):
        f = keys.view({2: np.float16, 4: np.float32, 8: np.float64}[kb]).astype(np.float64)
        mag = [v & ((1 << (bits - 1)) - 1) for v in ints]
        sign = [v >> (bits - 1) for v in ints]
        val = []
        for i in range(n):
            if np.isnan(f[i]):
                val.append((0, -mag[i]) if sign[i] else (2, mag[i]))
            else:
                val.append((1, float(f[i]), -1 if sign[i] else 1))
    else:  # 16-byte float: sign-magnitude word
        val = [(0, -(v & ((1 << 127) - 1))) if v >> 127 else (1, v) for v in ints]
    order = sorted(range(n), key=lambda i: val[i], reverse=desc)
    np.testing.assert_array_equal(orc.permutation(keys, kind, kb, desc), np.array(order, dtype=np.uint64))


@pytest.mark.parametrize("kb,vb", [(1, 0), (2, 4), (4, 4), (8, 8), (16, 8), (4, 3)])
def test_lsd_passes_compose_to_stable_sort(orc, kb, vb):
    """LSD invariant: stable counting passes over bytes 0..P-1 == std::stable_sort
    (two independent formulations)."""
    n = 5000
    keys = random_keys(0, kb, n, seed=kb, dup_frac=0.4)
    vals = random_values(vb, n, seed=kb) if vb else None
    w, v = keys.copy(), (vals.copy() if vals is not None else None)
    for p in range(kb):
        w, v = orc.lsd_pass(w, v, kb, vb, p)
    ek, ev = orc.sort(keys, vals, 0, kb, False, vb=vb if vb else None)
    np.testing.assert_array_equal(w, ek)
    if vb:
        np.testing.assert_array_equal(v, ev)


@pytest.mark.parametrize("kind,kb", [(0, 4), (1, 8), (2, 4), (2, 16)])
def test_digit_counts(orc, kind, kb):
    keys = random_keys(kind, kb, 7000, seed=5)
    for desc in (False, True):
        c = orc.digit_counts(keys, kind, kb, desc)
        tw = orc.twiddle_in(keys, kind, kb, desc).reshape(-1, kb)
        for p in range(kb):
            np.testing.assert_array_equal(c[p], np.bincount(tw[:, p], minlength=256))


def test_descending_is_not_reversed_ascending(orc):
    """Stable descending keeps equal keys in input order (not the reverse of ascending)."""
    keys = np.array([5, 1, 5, 3, 5], dtype=np.uint32)
    vals = np.arange(5, dtype=np.uint32)
    _, v = orc.sort(keys.view(np.uint8), vals, 0, 4, True, vb=4)
    np.testing.assert_array_equal(v, [0, 2, 4, 3, 1])


def test_mt19937_known_answer(orc):
    """C++ standard [rand.predef]: the 10000th output of default-seeded mt19937 is 4123659995."""
    assert int(orc.mt19937(5489, 10000)[-1]) == 4123659995


def test_mt19937_matches_numpy_legacy_seeding(orc):
    """numpy's RandomState(seed) uses the same init_genrand seeding and raw outputs."""
    rs = np.random.RandomState(0)
    np.testing.assert_array_equal(orc.mt19937(0, 1000), rs.randint(0, 2**32, size=1000, dtype=np.uint64).astype(np.uint32))


def test_mix32_is_bijective_lowbias32(orc):
    """mix32 (utils.cuh:228-242) is a composition of invertible steps; undo it
    with the modular inverses of its multipliers."""
    inv1, inv2 = pow(0x7FEB352D, -1, 1 << 32), pow(0x846CA68B, -1, 1 << 32)

    def unxorshift(y, s):
        x = y
        for _ in range(32 // s + 1):
            x = y ^ (x >> s)
        return x & 0xFFFFFFFF

    def inverse(y):
        y = unxorshift(y, 16)
        y = (y * inv2) & 0xFFFFFFFF
        y = unxorshift(y, 15)
        y = (y * inv1) & 0xFFFFFFFF
        return unxorshift(y, 16)

    rng = np.random.default_rng(1)
    for x in [0, 1, 0xFFFFFFFF, 0x80000000] + [int(v) for v in rng.integers(0, 2**32, 200)]:
        assert inverse(orc.mix32(x)) == x
    assert orc.mix32(0) == 0  # every step maps 0 to 0


def test_init_keys_modes(orc):
    """init_keys (utils.cuh:245-317): asc = index; random u32 = mix32(i ^ seed);
    byte_skip keeps bytes 1,3 and writes 0x55 in bytes 0,2."""
    n, seed = 1000, orc.set_seed_radix(0)
    assert seed == (0 + 81) * 17
    asc = orc.init_keys(4, 3, seed, n).view(np.uint32)
    np.testing.assert_array_equal(asc, np.arange(n, dtype=np.uint32))
    rnd = orc.init_keys(4, 1, seed, n).view(np.uint32)
    assert all(int(rnd[i]) == orc.mix32(i ^ seed) for i in range(0, n, 37))
    bsk = orc.init_keys(4, 2, seed, n).view(np.uint32)
    np.testing.assert_array_equal(bsk & 0x00FF00FF, np.full(n, 0x00550055, dtype=np.uint32))
    np.testing.assert_array_equal(bsk & 0xFF00FF00, rnd & 0xFF00FF00)
    r64 = orc.init_keys(8, 1, seed, 10).view(np.uint64)
    for i in range(10):
        assert int(r64[i]) == ((orc.mix32(i) ^ seed) << 32) | (orc.mix32((i + 0x9E3779B9) & 0xFFFFFFFF) ^ seed)


def test_checkers_detect_planted_mistakes(orc):
    """check_sorted / check_pairings (utils.cuh:341-432) flag a swapped pair,
    a wrong pairing and a stability break."""
    n = 2000
    keys = random_keys(0, 4, n, seed=9, dup_frac=0.5)
    vals = np.arange(n, dtype=np.uint32)
    ks, vs = orc.sort(keys, vals, 0, 4, False, vb=4)
    assert orc.check_sorted(ks, 0, 4, False)
    assert orc.check_pairings(keys, ks, vs, 4, 4)
    k32 = ks.view(np.uint32)
    bad = k32.copy()
    i = int(np.argmax(np.diff(k32.astype(np.int64)) > 0))
    bad[i], bad[i + 1] = bad[i + 1], bad[i]
    assert not orc.check_sorted(bad.view(np.uint8), 0, 4, False)
    v2 = vs.copy()
    v2[0], v2[-1] = v2[-1], v2[0]
    assert not orc.check_pairings(keys, ks, v2, 4, 4)
    # stability break: swap the values of two equal adjacent keys
    eq = np.where(k32[:-1] == k32[1:])[0]
    assert eq.size > 0
    j = int(eq[0])
    v3 = vs.copy()
    v3[j], v3[j + 1] = v3[j + 1], v3[j]
    # keys still match their originals only if the original keys are equal, which they are
    assert not orc.check_pairings(keys, ks, v3, 4, 4)


def test_generators_shapes_and_statistics(orc):
    low = orc.gen_u64_low16(1, 100000)
    assert low.max() < 65536 and 50500 < len(np.unique(low)) < 52100  # 65536*(1-exp(-100000/65536))
    z = orc.gen_zipf_u64(2, 200000)
    _, counts = np.unique(z, return_counts=True)
    top = counts.max() / z.size
    assert 0.08 < top < 0.14  # Zipf(1.1) over 2^20 ranks: p(rank 1) = 1/H ~ 0.11
    f = orc.gen_f32_normal_specials(3, 1_000_000)
    assert 0.0008 < np.isinf(f).mean() < 0.0012
    assert 0.009 < (f == 0).mean() < 0.011
    fin = f[np.isfinite(f) & (f != 0)]
    assert abs(fin.mean()) < 0.01 and abs(fin.std() - 1) < 0.01
    u = orc.gen_u128_pairs(4, 1000)
    assert u[:, 1].max() < (1 << 20)
    s = orc.gen_lognormal_sizes(5, 100000)
    assert s.min() >= 32 and s.max() <= (1 << 22)
    assert abs(np.median(s) - 2048) < 40
    assert abs(s.sum() / 2**28 - 1.0) < 0.03  # total ~ 2^28 keys


# ---------------------------------------------------------------------------
# Segmented sort oracle: each clamped segment sorted independently.
# ---------------------------------------------------------------------------
def _np_words(orc, keys, kind, kb, desc):
    w = orc.twiddle_in(keys, kind, kb, desc).reshape(-1, kb)
    return w


@pytest.mark.parametrize("kind,kb", [(0, 4), (1, 2), (2, 8), (0, 1)])
@pytest.mark.parametrize("desc", [False, True])
def test_segmented_oracle_matches_numpy_per_segment(orc, kind, kb, desc):
    """Per segment, numpy's stable argsort of the big-endian twiddled words."""
    from helpers import random_keys

    rng = np.random.default_rng(kb * 10 + kind + int(desc))
    n = 3000
    keys = random_keys(kind, kb, n, seed=kb + 3 * kind)
    vals = np.arange(n, dtype=np.uint32)
    cuts = np.sort(rng.integers(0, n, size=40))
    offsets = np.concatenate([[5], cuts[(cuts > 5) & (cuts < n - 7)], [n - 7]]).astype(np.uint64)
    gk, gv = orc.segmented_sort(keys, vals, offsets, kind, kb, desc, vb=4)
    words = _np_words(orc, keys, kind, kb, desc)
    big_endian = words[:, ::-1].copy().view(f"S{kb}").reshape(-1)  # bytes compare as unsigned
    expect = np.arange(n, dtype=np.uint32)
    for b, e in zip(offsets[:-1], offsets[1:]):
        b, e = int(b), int(e)
        if e - b >= 2:
            expect[b:e] = b + np.argsort(big_endian[b:e], kind="stable")
    assert np.array_equal(gv, expect)
    assert np.array_equal(gk.reshape(-1, kb), keys.reshape(-1, kb)[expect])


def test_segmented_oracle_single_segment_is_sort(orc):
    from helpers import random_keys

    keys = random_keys(2, 4, 5000, seed=9)
    a, _ = orc.segmented_sort(keys, None, [0, 5000], 2, 4, True)
    b, _ = orc.sort(keys, None, 2, 4, True)
    assert np.array_equal(a, b)


def test_segmented_oracle_clamps_and_leaves_gaps(orc):
    keys = np.arange(100, 0, -1, dtype=np.uint32)
    # [10,20) sorted, [20,20) empty, [20,30) sorted, [30,1000) clamped to
    # [30,100), [1000,50) clamped to the empty [100,100)
    off = [10, 20, 20, 30, 1000, 50]
    k, _ = orc.segmented_sort(keys, None, off, 0, 4, False)
    exp = keys.copy()
    exp[10:20] = np.sort(keys[10:20])
    exp[20:30] = np.sort(keys[20:30])
    exp[30:100] = np.sort(keys[30:100])
    assert np.array_equal(k, exp)


def test_segmented_oracle_planted_mistake_is_caught(orc):
    """A segment boundary shifted by one changes the result: the per-segment
    reference above would catch an off-by-one in the oracle's clamping."""
    keys = np.array([5, 4, 3, 2, 1, 0], dtype=np.uint32)
    good, _ = orc.segmented_sort(keys, None, [0, 3, 6], 0, 4, False)
    bad, _ = orc.segmented_sort(keys, None, [0, 4, 6], 0, 4, False)
    assert good.tolist() == [3, 4, 5, 0, 1, 2]
    assert not np.array_equal(good, bad)


@pytest.mark.parametrize("threads", [1, 3, 8, 16])
def test_parallel_stable_sort_baseline_is_stable(orc, threads):
    """The CPU baseline's multi-threaded stable sort gives std::stable_sort's
    result (equal keys keep their order across the merged runs)."""
    rng = np.random.default_rng(threads)
    n = 200_003
    keys = rng.integers(0, 5000, size=n, dtype=np.uint32)  # many equal keys
    vals = np.arange(n, dtype=np.uint32)
    ek, ev = orc.sort(keys, vals, 0, 4, False, vb=4)
    pk, pv = orc.stable_sort_par_u32(keys, vals, threads)
    assert np.array_equal(pk, ek.view(np.uint32)) and np.array_equal(pv, ev.view(np.uint32))
    kk, _ = orc.stable_sort_par_u32(keys, None, threads)
    assert np.array_equal(kk, np.sort(keys))
