"""C-ABI checks that need no GPU: exports, size query, plan, argument
validation, and the lookback protocol simulated with host threads."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions() -> list[str]:
    names = []
    inc = os.path.join(ROOT, "include")
    for f in sorted(os.listdir(inc)):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(rsort_[a-z0-9_]+)\s*\(", text, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(rs):
    names = _declared_functions()
    assert "rsort_sort" in names and "rsort_plan" in names
    L = rs.lib()
    for n in names:
        assert hasattr(L, n), f"librsort_amd.so does not export {n}"
    assert L.rsort_abi_version() == 1


def test_status_strings(rs):
    for s in range(6):
        assert rs.lib().rsort_status_string(s)


@pytest.mark.parametrize("kb,vb", [(4, 0), (4, 4), (8, 8), (16, 8), (1, 0), (2, 2), (4, 12)])
def test_size_query_needs_no_device(rs, kb, vb):
    n = (1 << 20) + 3
    d = rs.make_desc(0, kb, vb)
    tb = rs.sort_raw(d, None, None, None, 0, n)
    p = rs.plan(d, n)
    assert tb == p.workspace_bytes
    # ping-pong buffers plus lookback plus the control block
    vcore = 4 if (vb and vb not in (1, 2, 4, 8, 16)) else vb
    assert tb >= n * (kb + vcore) + p.lookback_bytes
    assert p.tile == p.threads * p.items_per_thread
    assert p.num_tiles == (n + p.tile - 1) // p.tile
    assert p.passes == kb
    assert p.index_mode == (1 if vb == 12 else 0)


def test_plan_switches_to_wide_words(rs):
    d = rs.make_desc(0, 4, 0)
    assert rs.plan(d, (1 << 29) - 1).wide == 0
    assert rs.plan(d, 1 << 29).wide == 1
    small, big = rs.plan(d, (1 << 29) - 1), rs.plan(d, 1 << 29)
    # 512 bytes of tile ticket counters (8 per pass, 16 passes at most), tile
    # words, and one group word per 8 tiles for every pass, 8 bytes each
    groups = (big.num_tiles + 7) // 8
    assert big.lookback_bytes == 512 + (big.num_tiles + big.passes * groups) * 256 * 8
    assert small.lookback_bytes == 512 + (small.num_tiles + small.passes * ((small.num_tiles + 7) // 8)) * 256 * 4


@pytest.mark.parametrize("field,value", [("key_kind", 3), ("key_bytes", 3), ("key_bytes", 32), ("descending", 2),
                                         ("flags", 8), ("reserved", 1)])
def test_invalid_descriptor_rejected(rs, field, value):
    d = rs.make_desc(0, 4, 0)
    setattr(d, field, value)
    with pytest.raises(rs.RsortError) as e:
        rs.sort_raw(d, None, None, None, 0, 100)
    assert e.value.status in (1, 2)


def test_float_byte_key_rejected(rs):
    with pytest.raises(rs.RsortError) as e:
        rs.sort_raw(rs.make_desc(2, 1, 0), None, None, None, 0, 100)
    assert e.value.status == 2


def test_null_temp_bytes_rejected(rs):
    d = rs.make_desc(0, 4, 0)
    st = rs.lib().rsort_sort(ctypes.byref(d), None, None, None, None, 10, None)
    assert st == 1


def test_undersized_workspace_rejected_before_device_work(rs):
    d = rs.make_desc(0, 4, 4)
    need = rs.sort_raw(d, None, None, None, 0, 5000)
    with pytest.raises(rs.RsortError) as e:
        rs.sort_raw(d, 0x1000, 0x2000, 0x100000, need - 1, 5000)
    assert e.value.status == 3


def test_trivial_sizes_are_noops(rs):
    """n == 0 and n == 1 need no device work (the reference leaves n == 0 undefined)."""
    d = rs.make_desc(0, 4, 4)
    for n in (0, 1):
        need = rs.sort_raw(d, None, None, None, 0, n)
        rs.sort_raw(d, 0x1000, 0x2000, 0x100000, need, n)


def test_index_mode_limit(rs):
    with pytest.raises(rs.RsortError) as e:
        rs.plan(rs.make_desc(0, 4, 12), 1 << 32)
    assert e.value.status == 2


def test_lookback_protocol_simulation(tmp_path):
    exe = tmp_path / "lookback_sim"
    src = os.path.join(ROOT, "tests", "lookback_sim.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", src, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ": 0 errors" in r.stdout


def test_dispatch_model_per_xcd_lock(tmp_path):
    """Two to four pass kernels sharing the device's XCDs: per-XCD tickets on
    more than one of them can deadlock; with the device lock (one per-XCD
    kernel at a time, the others on the global ticket) no schedule does.  A
    histogram that clears its own header beside a lock-holding pass kernel:
    waiting for workgroup 0 to clear can deadlock; the claim protocol the
    kernel uses (a running claimer that waits on nothing) never does."""
    exe = tmp_path / "dispatch_sim"
    src = os.path.join(ROOT, "tests", "dispatch_sim.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", src, "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "lock   deadlocks: adversarial 0/" in r.stdout and ": 0 errors" in r.stdout
    assert "hist-claim deadlocks: adversarial 0/200, random 0/200" in r.stdout


# ---------------------------------------------------------------------------
# Segmented sort: plan, size query, argument checks (no device work)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kb,vb,cap", [(4, 4, 16384), (4, 0, 16384), (8, 8, 8192), (16, 8, 4096), (1, 0, 16384),
                                       (2, 16, 4096), (4, 12, 16384)])
def test_segmented_plan_bins(rs, kb, vb, cap):
    d = rs.make_desc(0, kb, vb)
    p = rs.segmented_plan(d, 1 << 22, 4, 1000)
    caps = list(p.bin_capacity)
    # nine bins: 1/16, 3/32, 1/8, 3/16, 1/4, 3/8, 1/2, 3/4 and all of the largest
    assert caps[-1] == cap and caps == [cap * q // 32 for q in (2, 3, 4, 6, 8, 12, 16, 24, 32)]
    # at most n / (cap + 1) segments can be too long for the LDS bins
    assert p.max_long_segments == min(1000, (1 << 22) // (cap + 1))
    assert p.max_tiles == ((1 << 22) + p.tile - 1) // p.tile + p.max_long_segments
    assert p.index_mode == (1 if vb == 12 else 0)
    q = rs.segmented_sort_raw(d, None, None, 1 << 22, None, 4, 1000, None, 0)
    assert q == p.workspace_bytes


def test_segmented_plan_small_n_needs_no_long_path(rs):
    d = rs.make_desc(0, 4, 4)
    p = rs.segmented_plan(d, 16384, 4, 10)
    assert p.max_long_segments == 0 and p.max_tiles == 0
    big = rs.segmented_plan(d, 16385, 4, 10)
    assert big.max_long_segments == 1


@pytest.mark.parametrize("args,status", [
    (dict(offset_bytes=2), 1),
    (dict(offset_bytes=4, n=1 << 33), 2),
    (dict(num_segments=1 << 32), 2),
])
def test_segmented_invalid_arguments(rs, args, status):
    a = dict(n=1000, offset_bytes=4, num_segments=10)
    a.update(args)
    with pytest.raises(rs.RsortError) as e:
        rs.segmented_sort_raw(rs.make_desc(0, 4, 0), None, None, a["n"], None, a["offset_bytes"], a["num_segments"],
                              None, 0)
    assert e.value.status == status


def test_segmented_undersized_workspace_and_noops(rs):
    d = rs.make_desc(0, 4, 4)
    need = rs.segmented_sort_raw(d, None, None, 5000, None, 8, 7, None, 0)
    with pytest.raises(rs.RsortError) as e:
        rs.segmented_sort_raw(d, 0x1000, 0x2000, 5000, 0x3000, 8, 7, 0x100000, need - 1)
    assert e.value.status == 3
    # no segments / n <= 1: nothing is launched, no pointer is touched
    rs.segmented_sort_raw(d, 0x1000, 0x2000, 5000, 0x3000, 8, 0, 0x100000, need)
    rs.segmented_sort_raw(d, 0x1000, 0x2000, 1, 0x3000, 8, 7, 0x100000, need)


def test_serve_rejects_bad_config_before_device_work(rs):
    import ctypes

    for field, value in (("depth", 0), ("depth", 9), ("batch_requests", 0), ("batch_keys", 1), ("reserved", 1)):
        cfg = rs.ServeConfig()
        cfg.desc = rs.make_desc(0, 4, 4)
        cfg.batch_keys, cfg.batch_requests, cfg.window_us, cfg.depth = 1 << 20, 64, 100, 2
        setattr(cfg, field, value)
        h = ctypes.c_void_p()
        assert rs.lib().rsort_serve_create(ctypes.byref(cfg), ctypes.byref(h)) == 1, field
        assert not h.value


# ---------------------------------------------------------------------------
# Tuning table
# ---------------------------------------------------------------------------
def test_tuning_defaults_set_and_plan(rs, tmp_path):
    rs.tuning_load("")  # built-in defaults only
    t = rs.tuning_get(4, 4)
    assert (t.tile_candidate, t.lds_threshold) == (1, 16384) and 1 <= t.hist_blocks_per_cu <= 8
    assert (t.small_tile_candidate, t.small_log2n, t.tiny_tile_candidate, t.tiny_log2n) == (3, 22, 0, 0)
    d = rs.make_desc(0, 4, 4)
    # below 2^22 keys the small tier (1024 x 10), from there candidate 1 (1024 x 16)
    small = rs.plan(d, (1 << 22) - 1)
    assert (small.threads, small.items_per_thread) == (1024, 10)
    base = rs.plan(d, 1 << 22)
    assert (base.threads, base.items_per_thread) == (1024, 16)
    base = rs.plan(d, 1 << 28)
    assert (base.threads, base.items_per_thread) == (1024, 16)
    # candidate 0, the built-in geometry
    rs.tuning_set(4, 4, 0, t.hist_blocks_per_cu, 16384)
    base = rs.plan(d, 1 << 27)
    assert (base.threads, base.items_per_thread) == (1024, 20)
    try:
        rs.tuning_set(4, 4, 1, 2, 8192)
        p1 = rs.plan(d, 1 << 20)
        assert (p1.threads, p1.items_per_thread) == (1024, 16)
        assert p1.num_tiles == ((1 << 20) + 16383) // 16384
        rs.tuning_set(4, 4, 2, 2, 8192)
        p2 = rs.plan(d, 1 << 20)
        assert (p2.threads, p2.items_per_thread) == (512, 20)
        rs.tuning_set(4, 4, 3, 2, 8192)
        p3 = rs.plan(d, 1 << 20)
        assert (p3.threads, p3.items_per_thread) == (1024, 10)
        # the LDS threshold moves segments of 8193..16384 keys to the onesweep path
        sp = rs.segmented_plan(d, 1 << 22, 4, 1000)
        assert sp.max_long_segments == (1 << 22) // 8193
        # candidates are only built for the main widths: elsewhere they mean 0
        rs.tuning_set(2, 0, 1, 1, 0)
        assert (rs.plan(rs.make_desc(0, 2, 0), 1000).threads, rs.plan(rs.make_desc(0, 2, 0), 1000).items_per_thread) == (
            1024, 16)
        # a table file round trip
        f = tmp_path / "t.txt"
        f.write_text("# kb vb tile hist lds [small_tile small_log2n [tiny_tile tiny_log2n]]\n4 4 1 3 4096\n"
                     "8 0 2 2 8192 1 24\nbad line\n4 4 9 1 1\n8 8 0 1 1 0\n4 0 0 2 8192 1 24 3 21\n")
        assert rs.tuning_load(str(f)) == 3
        t40 = rs.tuning_get(4, 0)
        assert (t40.small_tile_candidate, t40.small_log2n, t40.tiny_tile_candidate, t40.tiny_log2n) == (1, 24, 3, 21)
        d40 = rs.make_desc(0, 4, 0)
        assert [rs.plan(d40, 1 << l).items_per_thread for l in (20, 21, 23, 24)] == [8, 12, 12, 16]
        t = rs.tuning_get(4, 4)
        assert (t.tile_candidate, t.hist_blocks_per_cu, t.lds_threshold, t.small_log2n) == (1, 3, 4096, 0)
        t8 = rs.tuning_get(8, 0)
        assert (t8.tile_candidate, t8.small_tile_candidate, t8.small_log2n) == (2, 1, 24)
        assert rs.plan(rs.make_desc(0, 8, 0), 1 << 23).threads == rs.plan(rs.make_desc(0, 8, 0), 1 << 25).threads * 2
        assert rs.tuning_load(str(tmp_path / "missing.txt")) == -1
    finally:
        rs.tuning_load(None)
    assert rs.plan(d, 1 << 20).items_per_thread == 10 and rs.plan(d, 1 << 28).items_per_thread == 16


def test_persisted_table_is_read_at_first_use(rs):
    """The autotuner's table next to the library is what the library uses."""
    import os

    path = os.path.join(os.path.dirname(rs.LIB_PATH), "tuning_gfx950.txt")
    rows = [l.split() for l in open(path) if l.strip() and not l.startswith("#")]
    assert rs.tuning_load(None) == len(rows) > 0
    for row in rows:
        kb, vb, tc, bpc, thr = (int(x) for x in row[:5])
        stc, sl = (int(row[5]), int(row[6])) if len(row) >= 7 else (0, 0)
        ttc, tl = (int(row[7]), int(row[8])) if len(row) == 9 else (0, 0)
        t = rs.tuning_get(kb, vb)
        assert (t.tile_candidate, t.hist_blocks_per_cu, t.lds_threshold) == (tc, bpc, thr)
        assert (t.small_tile_candidate, t.small_log2n, t.tiny_tile_candidate, t.tiny_log2n) == (stc, sl, ttc, tl)


@pytest.mark.parametrize("args", [(4, 4, 4, 1, 0), (4, 4, 0, 0, 0), (4, 4, 0, 9, 0), (4, 4, 0, 1, 16385),
                                  (3, 4, 0, 1, 0), (4, 12, 0, 1, 0), (4, 4, 0, 1, 0, 4, 20), (4, 4, 0, 1, 0, 1, 41),
                                  (4, 4, 0, 1, 0, 1, 24, 4, 20), (4, 4, 0, 1, 0, 1, 24, 3, 41)])
def test_tuning_rejects_out_of_range(rs, args):
    with pytest.raises(rs.RsortError):
        rs.tuning_set(*args)


# ---------------------------------------------------------------------------
# Array-of-structs sort: size query and argument checks (no device work)
# ---------------------------------------------------------------------------
def test_sort_records_query_and_arguments(rs):
    import ctypes

    d = rs.make_desc(0, 4, 0)
    tb = rs.sort_records_raw(d, None, 1 << 20, 12, 4, None, 0)
    # extracted keys + a key/value sort of 12-byte values (index mode)
    inner = rs.sort_raw(rs.make_desc(0, 4, 12), None, None, None, 0, 1 << 20)
    assert tb >= inner + (1 << 20) * 4
    for args, why in (((rs.make_desc(0, 4, 4), 12, 4), "value_bytes"), ((d, 12, 9), "key past the record"),
                      ((d, 0, 0), "empty record"), ((rs.make_desc(0, 8, 0), 6, 0), "key wider than the record")):
        desc, rb, off = args
        with pytest.raises(rs.RsortError):
            rs.sort_records_raw(desc, None, 100, rb, off, None, 0)
    # too-small workspace is rejected before any device work
    t = ctypes.c_size_t(16)
    st = rs.lib().rsort_sort_records(ctypes.byref(d), ctypes.c_void_p(0x1000), 1000, 12, 0, ctypes.c_void_p(0x1000),
                                     ctypes.byref(t), None)
    assert st == 3


# ---------------------------------------------------------------------------
# Autotuner: choice of the small-sort tile candidate (host logic only)
# ---------------------------------------------------------------------------
def test_autotune_small_choice():
    import importlib.util
    import os

    spec = importlib.util.spec_from_file_location(
        "autotune", os.path.join(os.path.dirname(__file__), "..", "cuda-fast-sort_amd", "autotune.py"))
    at = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(at)
    # whole-sort microseconds by log2 n and candidate, u32 keys and u32/u32 on MI355X
    keys = {20: {0: 82.4, 1: 77.5, 2: 74.7, 3: 72.1}, 21: {0: 99.0, 1: 95.9, 2: 96.1, 3: 87.8},
            22: {0: 113.2, 1: 119.8, 2: 121.3, 3: 122.9}, 23: {0: 147.9, 1: 161.2, 2: 179.1, 3: 167.3},
            24: {0: 223.3, 1: 233.6, 2: 244.6, 3: 249.3}, 26: {0: 687.5, 1: 743.1, 2: 781.7, 3: 853.8}}
    assert at.small_choice(keys, 0) == (0, 0, 3, 22)  # one tier: candidate 3 below 2^22
    pairs = {20: {0: 108.9, 1: 101.4, 2: 99.7, 3: 87.1}, 21: {0: 118.6, 1: 114.2, 2: 115.7, 3: 97.6},
             22: {0: 136.9, 1: 121.8, 2: 132.7, 3: 133.6}, 23: {0: 201.8, 1: 173.8, 2: 206.5, 3: 188.7},
             24: {0: 342.6, 1: 309.8, 2: 343.1, 3: 330.7}, 26: {0: 1196.7, 1: 1223.0, 2: 1393.3, 3: 1311.3}}
    # two tiers: candidate 3 below 2^22, candidate 1 below 2^25
    assert at.small_choice(pairs, 0) == (3, 22, 1, 25)
    # the large-sort candidate fastest everywhere: no tier
    flat = {20: {0: 1.0, 1: 2.0}, 22: {0: 1.0, 1: 1.5}}
    assert at.small_choice(flat, 0) == (0, 0, 0, 0)


def test_tensor_wrappers_reject_bad_buffers(rs):
    """The tensor wrappers check what the C ABI cannot see: host tensors,
    strided tensors and buffers smaller than n keys / values / the workspace
    raise ValueError before any call (CPU tensors: no device work)."""
    import torch

    keys = torch.zeros(100, dtype=torch.int32)
    ws = torch.zeros(1 << 20, dtype=torch.uint8)
    with pytest.raises(ValueError, match="CUDA"):
        rs.sort(keys, 1 << 20, ws, 100)
    with pytest.raises(ValueError, match="CUDA"):
        rs.sort_pairs(keys, keys.clone(), 1 << 20, ws, 100)
    with pytest.raises(ValueError, match="CUDA"):
        rs.segmented_sort(keys, torch.tensor([0, 50, 100], dtype=torch.int32), 1 << 20, ws)
    # the size checks come after the device check, so build a stand-in that
    # passes as a device tensor to reach them
    class Dev:
        def __init__(self, t):
            self.t, self.is_cuda, self.dtype, self.shape = t, True, t.dtype, t.shape

        def is_contiguous(self):
            return self.t.is_contiguous()

        def numel(self):
            return self.t.numel()

        def element_size(self):
            return self.t.element_size()

        def dim(self):
            return self.t.dim()

        def __getitem__(self, i):
            return self.t[i]

        def data_ptr(self):
            raise AssertionError("no pointer may be taken before the checks")

    with pytest.raises(ValueError, match="keys holds 400 bytes, 404"):
        rs.sort(Dev(keys), 1 << 20, Dev(ws), 101)
    with pytest.raises(ValueError, match="values holds 200 bytes"):
        rs.sort_pairs(Dev(keys), Dev(torch.zeros(50, dtype=torch.int32)), 1 << 20, Dev(ws), 100)
    with pytest.raises(ValueError, match="workspace holds"):
        rs.sort(Dev(keys), (1 << 20) + 1, Dev(ws), 100)
    with pytest.raises(ValueError, match="contiguous"):
        rs.sort(Dev(torch.zeros(200, dtype=torch.int32)[::2]), 1 << 20, Dev(ws), 100)
    with pytest.raises(ValueError, match="offsets must be contiguous"):
        rs.segmented_sort(Dev(keys), Dev(torch.tensor([0, 9, 50, 9, 100], dtype=torch.int32)[::2]), 1 << 20, Dev(ws))


def test_noalloc_rejects_before_device_work(rs):
    """The allocation-free entry point refuses, before any device work, sorts
    whose state exceeds its per-device array and value widths moved by index
    (the pointers here are never dereferenced)."""
    fake = 0x10000
    ok_sizes = [((4, 4), 1 << 28), ((8, 8), 1 << 26), ((16, 8), 1 << 24)]
    too_big = [((4, 4), 1 << 29), ((8, 8), 1 << 27), ((16, 8), 1 << 25)]
    for (kb, vb), n in too_big:
        with pytest.raises(rs.RsortError) as e:
            rs.sort_noalloc_raw(rs.make_desc(0, kb, vb), fake, fake, fake * 2, fake * 2, n)
        assert e.value.status == 2 and "allocation-free" in rs.last_error()
    with pytest.raises(rs.RsortError) as e:
        rs.sort_noalloc_raw(rs.make_desc(0, 4, 12), fake, fake, fake * 2, fake * 2, 1000)
    assert e.value.status == 2
    # argument errors also come before device work
    for args in [(0, fake, fake * 2, fake * 2), (fake, 0, fake * 2, fake * 2), (fake, fake, 0, fake * 2),
                 (fake, fake, fake * 2, 0), (fake, fake, fake, fake * 2)]:
        with pytest.raises(rs.RsortError) as e:
            rs.sort_noalloc_raw(rs.make_desc(0, 4, 4), *args, 1000)
        assert e.value.status == 1
    # trivial sizes are no-ops that touch nothing
    for (kb, vb), n in ok_sizes:
        rs.sort_noalloc_raw(rs.make_desc(0, kb, vb), fake, fake, fake * 2, fake * 2, 1)


@pytest.mark.parametrize("mode,nbytes,bpc,same", [
    (3, 65536 + 16, 4, False),   # scatter: not whole 64 KiB tiles
    (3 | 8, 65536, 4, False),    # scatter has no chunked form
    (3, 65536, 4, True),         # scatter in place
    (16, 4096, 4, False),        # unknown mode bit
    (2, 4096 + 8, 4, False),     # not whole 16-byte units
    (2, 4096, 0, False),         # blocks per CU out of range
    (2, 4096, 33, False),
])
def test_stream_bandwidth_rejects_bad_arguments(rs, mode, nbytes, bpc, same):
    """The streaming kernels check their arguments before any device call."""
    with pytest.raises(rs.RsortError) as e:
        rs.stream_bandwidth(mode, 0x100000, 0x100000 if same else 0x200000, nbytes, bpc)
    assert e.value.status == 1
