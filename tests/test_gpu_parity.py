"""GPU parity: the HIP sort through the C ABI against the CPU oracle, bit-exact
(keys and value permutation), over the reference's validation matrix
(benchmark/validate_sort.cuh:36-236: key types x value types x orders x sizes)
plus the golden fixtures and the workload configurations."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import KEY_TYPES, golden_names, index_values, load_golden, random_keys, random_values

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.init()


def dev(a: np.ndarray | None):
    if a is None:
        return None
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1).copy()).cuda()


def gpu_sort(rs, keys: np.ndarray, vals: np.ndarray | None, kind: int, kb: int, vb: int, desc: bool, flags: int = 0):
    n = keys.size // kb
    print(f"[case] kind={kind} kb={kb} vb={vb} desc={int(desc)} n={n} flags={flags}", flush=True)
    dk, dv = dev(keys), dev(vals)
    d = rs.make_desc(kind, kb, vb, desc, flags)
    tb = rs.sort_raw(d, None, None, None, 0, n)
    ws = torch.empty(max(tb, 1), dtype=torch.uint8, device="cuda")
    rs.sort_raw(d, dk.data_ptr() if dk is not None else None, dv.data_ptr() if dv is not None else None,
                ws.data_ptr(), tb, n)
    torch.cuda.synchronize()
    return dk.cpu().numpy(), (dv.cpu().numpy() if dv is not None else None)


def check(rs, orc, keys, vals, kind, kb, vb, desc, flags=0):
    keys = np.ascontiguousarray(keys).view(np.uint8).reshape(-1)
    vals = None if vals is None else np.ascontiguousarray(vals).view(np.uint8).reshape(-1)
    gk, gv = gpu_sort(rs, keys, vals, kind, kb, vb, desc, flags)
    ek, ev = orc.sort(keys, vals, kind, kb, desc, vb=vb if vb else None)
    if not np.array_equal(gk, ek):
        diff = np.where((gk.reshape(-1, kb) != ek.reshape(-1, kb)).any(axis=1))[0]
        raise AssertionError(f"keys differ at {diff.size} positions, first {diff[:8]}")
    if vb:
        diff = np.where((gv.reshape(-1, vb) != ev.reshape(-1, vb)).any(axis=1))[0]
        assert diff.size == 0, f"values differ at {diff.size} positions, first {diff[:8]}"


SIZES = [2, 10, 100, 1000, 8193, 70001]


@pytest.mark.parametrize("kind,kb", KEY_TYPES)
@pytest.mark.parametrize("vb", [0, 4])
@pytest.mark.parametrize("desc", [False, True])
def test_matrix(rs, orc, kind, kb, vb, desc):
    for n in SIZES:
        keys = random_keys(kind, kb, n, seed=n * 31 + kb * 7 + kind)
        vals = index_values(4, n) if vb else None
        check(rs, orc, keys, vals, kind, kb, vb, desc)


@pytest.mark.parametrize("kind,kb", [(0, 4), (2, 4), (1, 8), (0, 16), (0, 1), (1, 2)])
@pytest.mark.parametrize("vb", [1, 2, 8, 16, 3, 12])
def test_value_widths(rs, orc, kind, kb, vb):
    for n, desc in ((777, False), (50001, True)):
        keys = random_keys(kind, kb, n, seed=vb * 13 + kb)
        vals = random_values(vb, n, seed=vb)
        check(rs, orc, keys, vals, kind, kb, vb, desc)


@pytest.mark.parametrize("name", golden_names())
@pytest.mark.parametrize("desc", [False, True])
def test_golden(rs, name, desc):
    g = load_golden(name)
    kb, n = g["key_bytes"], g["n"]
    vals = np.arange(n, dtype=np.uint32)
    gk, gv = gpu_sort(rs, g["key_bytes_array"], vals, g["kind"], kb, 4, desc)
    exp = g["desc"] if desc else g["asc"]
    np.testing.assert_array_equal(gv.view(np.uint32), exp.astype(np.uint32))
    np.testing.assert_array_equal(gk.reshape(-1, kb), g["key_bytes_array"].reshape(-1, kb)[exp.astype(np.int64)])


@pytest.mark.parametrize("n", [1 << 20, (1 << 22) + 12345])
@pytest.mark.parametrize("desc", [False, True])
def test_u32_pairs_large(rs, orc, n, desc):
    keys = orc.mt19937(7, n).view(np.uint8)
    check(rs, orc, keys, index_values(4, n), 0, 4, 4, desc)


@pytest.mark.parametrize("vb", [0, 4])
def test_tile_boundary_sizes(rs, orc, vb):
    """u32 sizes one key either side of whole tiles, for the tile of each
    size tier of the committed tuning (1024 x 8 or 1024 x 10 below 2^22 keys,
    1024 x 16 from there), and of one histogram workgroup's share."""
    sizes = {8191, 8193, 10239, 10241, 16383, 16385, 16384 * 3 + 1, 8192 * 128 - 1, 8192 * 128 + 1,
             10240 * 103 + 1, (1 << 22) - 1, (1 << 22) + 1, (1 << 22) + 16383}
    for i, n in enumerate(sorted(sizes)):
        keys = orc.mt19937(60 + i, n).view(np.uint8)
        check(rs, orc, keys, index_values(4, n) if vb else None, 0, 4, vb, bool(i & 1))


def test_cfg_plain_keys_mt19937_seed0(rs, orc):
    """n = 2^20 uint32 keys = raw mt19937(seed 0), key-only ascending."""
    n = 1 << 20
    keys = orc.mt19937(0, n).view(np.uint8)
    check(rs, orc, keys, None, 0, 4, 0, False)


@pytest.mark.parametrize("flags", [0, 1])
def test_uniform_passes_skipped_and_not(rs, orc, flags):
    """u64 keys with the upper 48 bits zero: 6 of 8 passes have one digit; with
    and without the device-side skip the result is the stable sort."""
    n = 300_000
    keys = orc.gen_u64_low16(11, n).view(np.uint8)
    for desc in (False, True):
        check(rs, orc, keys, index_values(4, n), 0, 8, 4, desc, flags)


@pytest.mark.parametrize("kb", [1, 2, 4, 8, 16])
def test_all_keys_equal(rs, orc, kb):
    """Every pass uniform: every pass is skipped, the data is untouched."""
    n = 100_000
    keys = np.tile(np.arange(1, kb + 1, dtype=np.uint8), n)
    for flags in (0, 1):
        check(rs, orc, keys, index_values(4, n), 1, kb, 4, False, flags)


def test_odd_executed_pass_count(rs, orc):
    """u32 keys whose bytes 1 and 3 vary and bytes 0 and 2 are constant:
    two passes execute (even); bytes 0,1,2 constant: one non-uniform pass
    (odd: pass 0 runs as a copy to make the count even)."""
    n = 200_000
    r = orc.mt19937(3, n)
    k_even = ((r & 0xFF00FF00) | 0x00550055).astype(np.uint32)
    k_odd = ((r & 0xFF000000) | 0x00123456).astype(np.uint32)
    for k in (k_even, k_odd):
        check(rs, orc, k.view(np.uint8), index_values(4, n), 0, 4, 4, False)


@pytest.mark.parametrize("kind,desc", [(1, False), (2, True), (0, True)])
def test_uniform_pass_run_as_copy(rs, orc, kind, desc):
    """Odd count of non-uniform passes: the lowest uniform pass runs as a plain
    copy (twiddled in when it is the first executed pass, out when the last)
    so the result lands in the caller's buffer.  Signed, float-descending and
    unsigned-descending keys; u32 with one varying byte, u64 with three."""
    n = 150_001
    r = orc.mt19937(11, n).astype(np.uint64)
    k32 = ((r & 0x00FF0000) | 0x80123456 if kind else (r & 0x00FF0000) | 0x00123456).astype(np.uint32)
    check(rs, orc, k32.view(np.uint8), index_values(4, n), kind, 4, 4, desc)
    if kind != 2:
        k64 = (r & 0xFF) | ((r & 0xFF00) << 16) | ((r & 0xFF0000) << 32) | np.uint64(0x1200330000440000)
        check(rs, orc, k64.astype(np.uint64).view(np.uint8), index_values(8, n), kind, 8, 8, desc)
        check(rs, orc, k64.astype(np.uint64).view(np.uint8), None, kind, 8, 0, desc)


def test_reference_init_keys_modes(rs, orc):
    """The reference's own generator (utils.cuh:245-317) in its three modes."""
    n = 100_000
    for kb in (4, 8, 16):
        for mode in (1, 2, 3):
            keys = orc.init_keys(kb, mode, orc.set_seed_radix(3), n)
            check(rs, orc, keys, index_values(8, n), 0, kb, 8, False)
            check(rs, orc, keys, None, 2 if kb != 1 else 0, kb, 0, True)


def test_skewed_and_float_configs_small(rs, orc):
    """Zipf(1.1) uint64 and N(0,1) float32 with +-0/+-inf at 2^20."""
    n = 1 << 20
    z = orc.gen_zipf_u64(5, n).view(np.uint8)
    check(rs, orc, z, None, 0, 8, 0, False)
    f = orc.gen_f32_normal_specials(6, n).view(np.uint8)
    check(rs, orc, f, None, 2, 4, 0, False)
    check(rs, orc, f, index_values(4, n), 2, 4, 4, True)


def test_u128_pairs_config_small(rs, orc):
    n = 1 << 18
    k = orc.gen_u128_pairs(8, n).view(np.uint8)
    check(rs, orc, k, index_values(8, n), 0, 16, 8, False)


def test_unaligned_key_pointer(rs, orc):
    """Keys at an offset that is not 16-byte aligned (histogram head/tail paths)."""
    n = 100_003
    keys = random_keys(0, 4, n + 3, seed=77)
    base = dev(keys)
    d = rs.make_desc(0, 4, 0)
    tb = rs.sort_raw(d, None, None, None, 0, n)
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    sub = base[4:4 + 4 * n]  # 4-byte aligned, not 16-byte aligned
    rs.sort_raw(d, sub.data_ptr(), None, ws.data_ptr(), tb, n)
    torch.cuda.synchronize()
    ek, _ = orc.sort(keys[4:4 + 4 * n], None, 0, 4, False)
    np.testing.assert_array_equal(sub.cpu().numpy(), ek)
    np.testing.assert_array_equal(base[:4].cpu().numpy(), keys[:4])  # outside untouched


def test_workspace_reuse_and_stream(rs, orc):
    """One workspace reused by consecutive sorts of different shapes on a side stream."""
    s = torch.cuda.Stream()
    ws = torch.empty(64 << 20, dtype=torch.uint8, device="cuda")
    for i, (kb, vb, n) in enumerate([(4, 4, 300_000), (8, 0, 123_457), (4, 0, 5), (16, 16, 40_000), (2, 8, 99_999)]):
        keys = random_keys(0, kb, n, seed=i)
        vals = random_values(vb, n, seed=i) if vb else None
        dk, dv = dev(keys), dev(vals)
        d = rs.make_desc(0, kb, vb)
        tb = rs.sort_raw(d, None, None, None, 0, n)
        assert tb <= ws.numel()
        with torch.cuda.stream(s):
            rs.sort_raw(d, dk.data_ptr(), dv.data_ptr() if dv is not None else None, ws.data_ptr(), ws.numel(), n,
                        stream=s)
        s.synchronize()
        ek, ev = orc.sort(keys, vals, 0, kb, False, vb=vb if vb else None)
        np.testing.assert_array_equal(dk.cpu().numpy(), ek)
        if vb:
            np.testing.assert_array_equal(dv.cpu().numpy(), ev)


def test_workspace_garbage_and_back_to_back(rs, orc):
    """Sorts whose histogram clears the control region itself (grids of 2 to
    256 workgroups): a workspace first filled with all-ones bytes (every
    lookback flag set, a header and counts of garbage), then three sorts
    queued back to back on it with no synchronisation between them."""
    ws = torch.full((64 << 20,), 0xFF, dtype=torch.uint8, device="cuda")
    cases = [(4, 4, (1 << 20) + 7), (8, 0, 300_001), (4, 0, 1 << 21)]
    dks, dvs, host = [], [], []
    for i, (kb, vb, n) in enumerate(cases):
        print(f"[case] garbage workspace kb={kb} vb={vb} n={n}", flush=True)
        keys = random_keys(0, kb, n, seed=50 + i)
        vals = random_values(vb, n, seed=60 + i) if vb else None
        dk, dv = dev(keys), dev(vals)
        d = rs.make_desc(0, kb, vb)
        assert rs.sort_raw(d, None, None, None, 0, n) <= ws.numel()
        rs.sort_raw(d, dk.data_ptr(), dv.data_ptr() if dv is not None else None, ws.data_ptr(), ws.numel(), n)
        dks.append(dk)
        dvs.append(dv)
        host.append((keys, vals, kb, vb))
    torch.cuda.synchronize()
    for (keys, vals, kb, vb), dk, dv in zip(host, dks, dvs):
        ek, ev = orc.sort(keys, vals, 0, kb, False, vb=vb if vb else None)
        np.testing.assert_array_equal(dk.cpu().numpy(), ek)
        if vb:
            np.testing.assert_array_equal(dv.cpu().numpy(), ev)


def test_python_api_mirror(rs, orc):
    """rsort_amd.sort / sort_pairs_descending with torch dtypes (example_basic flow)."""
    n = 50_000
    x = torch.randint(-2**31, 2**31 - 1, (n,), dtype=torch.int32, device="cuda")
    v = torch.arange(n, dtype=torch.int32, device="cuda")
    ref_k, ref_v = orc.sort(x.cpu().numpy().view(np.uint8), v.cpu().numpy(), 1, 4, True, vb=4)
    tb = rs.sort_pairs_descending(x, v, None, None, n)
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    rs.sort_pairs_descending(x, v, tb, ws, n)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(x.cpu().numpy().view(np.uint8), ref_k)
    np.testing.assert_array_equal(v.cpu().numpy(), ref_v)
    y = torch.randn(n, device="cuda")
    ref = np.sort(y.cpu().numpy())
    tb = rs.sort(y, None, None, n)
    rs.sort(y, tb, torch.empty(tb, dtype=torch.uint8, device="cuda"), n)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), ref)


@pytest.mark.parametrize("kb,vb", [(4, 0), (4, 4), (8, 0), (8, 8), (4, 8), (8, 4), (16, 0), (16, 8)])
@pytest.mark.parametrize("cand", [0, 1, 2, 3])
def test_tile_candidates(rs, orc, kb, vb, cand):
    """Every pass-kernel tile candidate the autotuner may choose, at every
    size (passes of at most 512 tiles use the segmented lookback walk; the
    5M-key u32 case has more tiles and uses the ordinary walk)."""
    try:
        t = rs.tuning_get(kb, vb)
        rs.tuning_set(kb, vb, cand, t.hist_blocks_per_cu, t.lds_threshold)
        cases = [(1000, False), (70001, True), (300007, False)]
        if kb == 4:
            cases.append((5_000_011, False))
        for n, desc in cases:
            keys = random_keys(0, kb, n, seed=cand * 100 + n)
            vals = index_values(vb, n) if vb else None
            check(rs, orc, keys, vals, 0, kb, vb, desc)
    finally:
        rs.tuning_load(None)


@pytest.mark.parametrize("bpc", [1, 3])
def test_histogram_blocks_per_cu(rs, orc, bpc):
    try:
        t = rs.tuning_get(4, 4)
        rs.tuning_set(4, 4, t.tile_candidate, bpc, t.lds_threshold)
        keys = random_keys(0, 4, 1_000_003, seed=bpc)
        check(rs, orc, keys, index_values(4, 1_000_003), 0, 4, 4, False)
    finally:
        rs.tuning_load(None)


@pytest.mark.parametrize("n", [1 << 19, (1 << 19) + 1, 1 << 22, (1 << 22) + 1])
def test_histogram_fold_boundaries(rs, orc, n):
    """u32 keys at the histogram grid sizes where the fold changes (16384 keys
    per 1024-thread workgroup, two per CU): 32 workgroups fold as one group,
    33 to 256 add to the global counts directly, 257 and more fold in two
    levels."""
    keys = random_keys(0, 4, n, seed=n % 1000003)
    check(rs, orc, keys, None, 0, 4, 0, n % 2 == 1)
    if n < (1 << 22):
        check(rs, orc, keys, index_values(4, n), 0, 4, 4, False)
