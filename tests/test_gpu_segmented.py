"""GPU parity of the segmented sort (LDS bins and the long-segment onesweep
path) through the C ABI, bit-exact against the CPU oracle: every segment
sorted as a whole-array sort would sort it, keys outside segments untouched."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import KEY_TYPES, random_keys, random_values

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


def offsets_from_lengths(lengths, start: int = 0, dtype=np.uint64) -> np.ndarray:
    return (start + np.concatenate([[0], np.cumsum(np.asarray(lengths, dtype=np.int64))])).astype(dtype)


def gpu_segmented(rs, keys, vals, n, offsets, kind, kb, vb, desc, ws=None):
    off = np.ascontiguousarray(offsets)
    s = max(off.size - 1, 0)
    print(f"[case] segmented kind={kind} kb={kb} vb={vb} desc={int(desc)} n={n} segments={s} "
          f"offset_bytes={off.itemsize}", flush=True)
    dk, dv, do = dev(keys), dev(vals), dev(off)
    d = rs.make_desc(kind, kb, vb, desc)
    tb = rs.segmented_sort_raw(d, None, None, n, None, off.itemsize, s, None, 0)
    if ws is None or ws.numel() < tb:
        ws = torch.empty(max(tb, 1), dtype=torch.uint8, device="cuda")
    rs.segmented_sort_raw(d, dk.data_ptr(), dv.data_ptr() if dv is not None else None, n, do.data_ptr(),
                          off.itemsize, s, ws.data_ptr(), tb)
    torch.cuda.synchronize()
    return dk.cpu().numpy(), (dv.cpu().numpy() if dv is not None else None)


def check(rs, orc, keys, vals, offsets, kind, kb, vb, desc, ws=None):
    keys = np.ascontiguousarray(keys).view(np.uint8).reshape(-1)
    vals = None if vals is None else np.ascontiguousarray(vals).view(np.uint8).reshape(-1)
    n = keys.size // kb
    gk, gv = gpu_segmented(rs, keys, vals, n, offsets, kind, kb, vb, desc, ws)
    ek, ev = orc.segmented_sort(keys, vals, offsets.astype(np.uint64), kind, kb, desc, vb=vb if vb else None)
    if not np.array_equal(gk, ek):
        diff = np.where((gk.reshape(-1, kb) != ek.reshape(-1, kb)).any(axis=1))[0]
        raise AssertionError(f"keys differ at {diff.size} positions, first {diff[:8]}")
    if vb:
        diff = np.where((gv.reshape(-1, vb) != ev.reshape(-1, vb)).any(axis=1))[0]
        assert diff.size == 0, f"values differ at {diff.size} positions, first {diff[:8]}"


# segment lengths around every LDS bin edge (4-byte keys: 1K, 2K, 4K, 8K, 16K)
# and into the onesweep path
EDGE_LENGTHS = [0, 1, 2, 3, 63, 64, 65, 1023, 1024, 1025, 1535, 1536, 1537, 2047, 2048, 2049, 3071, 3072, 3073,
                4095, 4096, 4097, 6143, 6144, 6145, 8191, 8192, 8193, 12287, 12288, 12289, 16383, 16384, 16385,
                40000, 131072 + 17]


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("offset_dtype", [np.uint32, np.uint64])
def test_bin_edges_u32_pairs(rs, orc, desc, offset_dtype):
    lengths = EDGE_LENGTHS
    n = int(sum(lengths)) + 11
    keys = random_keys(0, 4, n, seed=5)
    vals = np.arange(n, dtype=np.uint32)
    off = offsets_from_lengths(lengths, start=5, dtype=offset_dtype)  # keys [0,5) and the tail stay put
    check(rs, orc, keys, vals, off, 0, 4, 4, desc)


@pytest.mark.parametrize("kind,kb,vb", [(0, 8, 0), (1, 8, 8), (2, 8, 4), (0, 16, 8), (2, 16, 0), (0, 4, 0),
                                         (1, 2, 0), (0, 1, 4), (2, 4, 8)])
def test_bin_edges_every_width(rs, orc, kind, kb, vb):
    """Segments of cap - 1, cap and cap + 1 keys for every LDS bin of the
    width (capacities from rsort_segmented_plan), then one long segment."""
    d = rs.make_desc(kind, kb, vb)
    caps = list(rs.segmented_plan(d, 1 << 20, 8, 1).bin_capacity)
    lengths = [0, 1, 2] + [c + dl for c in caps for dl in (-1, 0, 1)] + [3 * caps[-1] + 7]
    n = int(sum(lengths)) + 3
    keys = random_keys(kind, kb, n, seed=kb * 7 + vb + kind)
    vals = random_values(vb, n, seed=vb + 1) if vb else None
    off = offsets_from_lengths(lengths, start=3)
    check(rs, orc, keys, vals, off, kind, kb, vb, False)


@pytest.mark.parametrize("kind,kb", KEY_TYPES)
@pytest.mark.parametrize("vb", [0, 4])
def test_key_types(rs, orc, kind, kb, vb):
    rng = np.random.default_rng(kb * 3 + kind)
    lengths = list(rng.integers(0, 3000, size=60)) + [1, 0, 70000]
    n = int(sum(lengths))
    keys = random_keys(kind, kb, n, seed=kb + 17 * kind)
    vals = np.arange(n, dtype=np.uint32) if vb else None
    for desc in (False, True):
        check(rs, orc, keys, vals, offsets_from_lengths(lengths), kind, kb, vb, desc)


@pytest.mark.parametrize("kind,kb", [(0, 4), (2, 8), (1, 16), (0, 1), (1, 2)])
@pytest.mark.parametrize("vb", [1, 2, 8, 16, 12])
def test_value_widths(rs, orc, kind, kb, vb):
    rng = np.random.default_rng(vb * 7 + kb)
    lengths = list(rng.integers(0, 9000, size=25)) + [30000]
    n = int(sum(lengths))
    keys = random_keys(kind, kb, n, seed=vb + kb)
    vals = random_values(vb, n, seed=vb)
    check(rs, orc, keys, vals, offsets_from_lengths(lengths), kind, kb, vb, vb % 2 == 0)


def test_lognormal_batch(rs, orc):
    """The serving workload's size distribution (median 2048, sigma 0.7356,
    clipped to [32, 2^22]) at 3000 segments."""
    sizes = orc.gen_lognormal_sizes(7, 3000)
    n = int(sizes.sum())
    keys = orc.mt19937(11, n)
    vals = np.arange(n, dtype=np.uint32)
    check(rs, orc, keys, vals, offsets_from_lengths(sizes), 0, 4, 4, False)


def test_one_long_segment_equals_sort(rs, orc):
    n = 1_500_000
    keys = orc.mt19937(3, n)
    vals = np.arange(n, dtype=np.uint32)
    check(rs, orc, keys, vals, np.array([0, n], dtype=np.uint64), 0, 4, 4, False)


def test_many_long_segments_and_tiny_ones(rs, orc):
    rng = np.random.default_rng(1)
    lengths = []
    for _ in range(40):
        lengths += [int(rng.integers(16385, 60000)), int(rng.integers(0, 40))]
    n = int(sum(lengths))
    keys = random_keys(1, 8, n, seed=2)
    vals = np.arange(n, dtype=np.uint32)
    check(rs, orc, keys, vals, offsets_from_lengths(lengths), 1, 8, 4, True)


def test_uniform_and_sorted_segments(rs, orc):
    """Segments whose every key is equal (each LDS pass keeps the order) and
    already sorted / reverse sorted segments."""
    parts, lengths = [], []
    for m, kind in ((3000, "eq"), (9000, "eq"), (20000, "eq"), (5000, "asc"), (12000, "desc"), (30000, "desc")):
        if kind == "eq":
            parts.append(np.full(m, 0xDEADBEEF, dtype=np.uint32))
        elif kind == "asc":
            parts.append(np.arange(m, dtype=np.uint32) * 3)
        else:
            parts.append(np.arange(m, 0, -1, dtype=np.uint32))
        lengths.append(m)
    keys = np.concatenate(parts)
    vals = np.arange(keys.size, dtype=np.uint32)
    check(rs, orc, keys, vals, offsets_from_lengths(lengths), 0, 4, 4, False)


def test_offsets_clamped_and_gaps(rs, orc):
    n = 50000
    keys = orc.mt19937(5, n)
    vals = np.arange(n, dtype=np.uint32)
    # disjoint segments out of order: [30000, n) (an end past n, a long
    # segment), [100, 900), [900, 5000); decreasing pairs (10^9 -> 100,
    # 5000 -> 4000) and empty segments between them; keys [0, 100) and
    # [5000, 30000) lie in no segment and stay as they are
    off = np.array([30000, 10**9, 100, 900, 900, 5000, 4000, 4000], dtype=np.uint64)
    check(rs, orc, keys, vals, off, 0, 4, 4, False)


def test_overlapping_segments_stay_in_bounds(rs, orc):
    """Overlapping segments sort in an unspecified order (the bins and the
    long segments run at once): the segments that overlap no other are
    exact, and nothing outside the segments' union changes."""
    n = 60000
    keys = orc.mt19937(6, n)
    vals = np.arange(n, dtype=np.uint32)
    # [900, 5000) (an LDS bin) overlaps [4000, 30000) (long); [0, 100),
    # [100, 900) and [30000, 50000) overlap nothing; [50000, 60000) is in no
    # segment
    off = np.array([0, 100, 900, 900, 5000, 4000, 4000, 30000, 50000], dtype=np.uint64)
    kb = np.ascontiguousarray(keys).view(np.uint8).reshape(-1)
    gk, gv = gpu_segmented(rs, kb, vals.view(np.uint8), n, off, 0, 4, 4, False)
    gk, gv = gk.view(np.uint32), gv.view(np.uint32)
    for lo, hi in ((0, 100), (100, 900), (30000, 50000)):
        ek, ev = orc.sort(keys[lo:hi].view(np.uint8), vals[lo:hi], 0, 4, False, vb=4)
        assert np.array_equal(gk[lo:hi], ek.view(np.uint32)), f"segment [{lo}, {hi}) keys"
        assert np.array_equal(gv[lo:hi], ev), f"segment [{lo}, {hi}) values"
    assert np.array_equal(gk[50000:], keys[50000:]) and np.array_equal(gv[50000:], vals[50000:])


def test_workspace_reuse_and_tiny_segments(rs, orc):
    ws = torch.empty(64 << 20, dtype=torch.uint8, device="cuda")
    for seed in range(3):
        rng = np.random.default_rng(seed)
        lengths = list(rng.integers(0, 70, size=5000))
        n = int(sum(lengths))
        keys = random_keys(0, 2, n, seed=seed)
        vals = np.arange(n, dtype=np.uint32)
        check(rs, orc, keys, vals, offsets_from_lengths(lengths), 0, 2, 4, seed == 1, ws=ws)


def test_python_mirror(rs, orc):
    lengths = [5, 3000, 20000, 1]
    n = sum(lengths)
    keys = torch.from_numpy(orc.mt19937(9, n).view(np.int32).copy()).cuda()
    vals = torch.arange(n, dtype=torch.int64, device="cuda")
    off = torch.tensor(offsets_from_lengths(lengths).astype(np.int64), device="cuda")
    tb = rs.segmented_sort(keys, off, None, None, values=vals)
    ws = torch.empty(tb, dtype=torch.uint8, device="cuda")
    ref_k = keys.cpu().numpy()
    rs.segmented_sort(keys, off, tb, ws, values=vals, descending=True)
    torch.cuda.synchronize()
    ek, ev = orc.segmented_sort(ref_k, np.arange(n, dtype=np.int64), off.cpu().numpy(), 1, 4, True, vb=8)
    assert np.array_equal(keys.cpu().numpy(), ek)
    assert np.array_equal(vals.cpu().numpy(), ev)


@pytest.mark.parametrize("threshold", [0, 1024, 4096])
def test_lds_thresholds(rs, orc, threshold):
    """Segments above the tuned LDS threshold take the onesweep path (0: all)."""
    try:
        t = rs.tuning_get(4, 4)
        rs.tuning_set(4, 4, t.tile_candidate, t.hist_blocks_per_cu, threshold)
        lengths = [0, 1, 2, 500, 1024, 1025, 3000, 4096, 4097, 9000, 20000]
        n = int(sum(lengths))
        keys = random_keys(0, 4, n, seed=threshold + 1)
        vals = np.arange(n, dtype=np.uint32)
        check(rs, orc, keys, vals, offsets_from_lengths(lengths), 0, 4, 4, threshold == 1024)
    finally:
        rs.tuning_load(None)
