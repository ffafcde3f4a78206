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
    assert big.lookback_bytes == 2 * big.num_tiles * 256 * 4


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
