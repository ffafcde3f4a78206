"""numpy front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; the product never imports it.
Each function names the reference code it restates (see rsort_oracle.cpp).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

KIND_UNSIGNED, KIND_SIGNED, KIND_FLOAT = 0, 1, 2

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, i32, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32
        L.oracle_twiddle_in.argtypes = [i32, i32, i32, vp, vp, u64]
        L.oracle_twiddle_out.argtypes = [i32, i32, i32, vp, vp, u64]
        L.oracle_sort.argtypes = [i32, i32, i32, i32, vp, vp, u64]
        L.oracle_sort_permutation.argtypes = [i32, i32, i32, vp, u64, vp]
        L.oracle_segmented_sort.argtypes = [i32, i32, i32, i32, vp, vp, u64, vp, u64]
        L.oracle_lsd_pass.argtypes = [i32, i32, i32, vp, vp, vp, vp, u64]
        L.oracle_digit_counts.argtypes = [i32, i32, i32, vp, u64, vp]
        L.oracle_mix32.argtypes = [u32]
        L.oracle_mix32.restype = u32
        L.oracle_mix64.argtypes = [u64]
        L.oracle_mix64.restype = u64
        L.oracle_set_seed_radix.argtypes = [i32]
        L.oracle_init_keys.argtypes = [i32, i32, u32, vp, u64]
        L.oracle_check_sorted.argtypes = [i32, i32, i32, vp, u64]
        L.oracle_check_pairings.argtypes = [i32, i32, vp, vp, vp, u64]
        L.oracle_mt19937.argtypes = [u32, vp, u64]
        L.oracle_gen_u64_low16.argtypes = [u32, vp, u64]
        L.oracle_gen_zipf_u64.argtypes = [u32, ctypes.c_double, u64, vp, u64]
        L.oracle_gen_f32_normal_specials.argtypes = [u32, vp, u64]
        L.oracle_gen_u128_pairs.argtypes = [u32, vp, u64]
        L.oracle_gen_lognormal_sizes.argtypes = [u32, ctypes.c_double, ctypes.c_double, u64, u64, vp, u64]
        L.oracle_time_stable_sort_u32.argtypes = [vp, vp, u64, i32]
        L.oracle_time_stable_sort_u32.restype = ctypes.c_double
        L.oracle_time_sort_par_u32.argtypes = [vp, vp, u64, i32]
        L.oracle_time_sort_par_u32.restype = ctypes.c_double
        L.oracle_stable_sort_par_u32.argtypes = [vp, vp, u64, i32, vp, vp]
        L.oracle_time_stable_sort_par_u32.argtypes = [vp, vp, u64, i32, i32]
        L.oracle_time_stable_sort_par_u32.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _as_bytes(a: np.ndarray, width: int) -> np.ndarray:
    a = np.ascontiguousarray(a)
    b = a.view(np.uint8).reshape(-1)
    assert b.size % width == 0
    return b


def twiddle_in(keys: np.ndarray, kind: int, kb: int, desc: bool) -> np.ndarray:
    """radix_traits::twiddle_in (tuning.cuh:132-148); returns the kb-byte words as uint8[n*kb]."""
    src = _as_bytes(keys, kb)
    out = np.empty_like(src)
    assert lib().oracle_twiddle_in(kind, kb, int(desc), _p(src), _p(out), src.size // kb) == 0
    return out


def twiddle_out(words: np.ndarray, kind: int, kb: int, desc: bool) -> np.ndarray:
    src = _as_bytes(words, kb)
    out = np.empty_like(src)
    assert lib().oracle_twiddle_out(kind, kb, int(desc), _p(src), _p(out), src.size // kb) == 0
    return out


def sort(keys: np.ndarray, values: np.ndarray | None, kind: int, kb: int, desc: bool, vb: int | None = None):
    """Expected result of rsort::sort / sort_pairs (std::stable_sort on twiddled keys)."""
    k = np.array(keys, copy=True)
    v = None if values is None else np.array(values, copy=True)
    kbytes = _as_bytes(k, kb)
    n = kbytes.size // kb
    if v is not None:
        vb = vb if vb is not None else v.nbytes // max(n, 1)
        vbytes = _as_bytes(v, vb) if n else v.view(np.uint8).reshape(-1)
    else:
        vb, vbytes = 0, None
    assert lib().oracle_sort(kind, kb, vb, int(desc), _p(kbytes), _p(vbytes), n) == 0
    return k, v


def segmented_sort(keys: np.ndarray, values: np.ndarray | None, offsets, kind: int, kb: int, desc: bool,
                   vb: int | None = None):
    """Expected result of rsort_segmented_sort: each clamped segment sorted as sort() does."""
    k = np.array(keys, copy=True)
    v = None if values is None else np.array(values, copy=True)
    kbytes = _as_bytes(k, kb)
    n = kbytes.size // kb
    if v is not None:
        vb = vb if vb is not None else v.nbytes // max(n, 1)
        vbytes = _as_bytes(v, vb) if n else v.view(np.uint8).reshape(-1)
    else:
        vb, vbytes = 0, None
    off = np.ascontiguousarray(np.asarray(offsets, dtype=np.uint64))
    s = max(off.size - 1, 0)
    assert lib().oracle_segmented_sort(kind, kb, vb, int(desc), _p(kbytes), _p(vbytes), n, _p(off), s) == 0
    return k, v


def permutation(keys: np.ndarray, kind: int, kb: int, desc: bool) -> np.ndarray:
    kbytes = _as_bytes(keys, kb)
    n = kbytes.size // kb
    out = np.empty(n, dtype=np.uint64)
    assert lib().oracle_sort_permutation(kind, kb, int(desc), _p(kbytes), n, _p(out)) == 0
    return out


def lsd_pass(words: np.ndarray, vals: np.ndarray | None, kb: int, vb: int, p: int):
    wi = _as_bytes(words, kb)
    n = wi.size // kb
    wo = np.empty_like(wi)
    vi = None if vals is None else _as_bytes(vals, vb)
    vo = None if vals is None else np.empty_like(vi)
    assert lib().oracle_lsd_pass(kb, vb, p, _p(wi), _p(vi), _p(wo), _p(vo), n) == 0
    return wo, vo


def digit_counts(keys: np.ndarray, kind: int, kb: int, desc: bool) -> np.ndarray:
    kbytes = _as_bytes(keys, kb)
    out = np.zeros(kb * 256, dtype=np.uint64)
    assert lib().oracle_digit_counts(kind, kb, int(desc), _p(kbytes), kbytes.size // kb, _p(out)) == 0
    return out.reshape(kb, 256)


def mix32(x: int) -> int:
    return int(lib().oracle_mix32(x))


def mix64(x: int) -> int:
    return int(lib().oracle_mix64(x))


def set_seed_radix(i: int) -> int:
    return int(lib().oracle_set_seed_radix(i))


def init_keys(kb: int, mode: int, seed: int, n: int) -> np.ndarray:
    """init_keys (utils.cuh:245-317); mode 1 random, 2 byte_skip, 3 asc. Returns uint8[n*kb]."""
    out = np.empty(n * kb, dtype=np.uint8)
    assert lib().oracle_init_keys(kb, mode, seed, _p(out), n) == 0
    return out


def check_sorted(keys: np.ndarray, kind: int, kb: int, desc: bool) -> bool:
    kbytes = _as_bytes(keys, kb)
    return lib().oracle_check_sorted(kind, kb, int(desc), _p(kbytes), kbytes.size // kb) == 1


def check_pairings(keys_orig: np.ndarray, keys_sorted: np.ndarray, vals_sorted: np.ndarray, kb: int, vb: int) -> bool:
    ko, ks, vs = _as_bytes(keys_orig, kb), _as_bytes(keys_sorted, kb), _as_bytes(vals_sorted, vb)
    return lib().oracle_check_pairings(kb, vb, _p(ko), _p(ks), _p(vs), ko.size // kb) == 1


def mt19937(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint32)
    lib().oracle_mt19937(seed, _p(out), n)
    return out


def gen_u64_low16(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    lib().oracle_gen_u64_low16(seed, _p(out), n)
    return out


def gen_zipf_u64(seed: int, n: int, s: float = 1.1, ranks: int = 1 << 20) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    lib().oracle_gen_zipf_u64(seed, s, ranks, _p(out), n)
    return out


def gen_f32_normal_specials(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float32)
    lib().oracle_gen_f32_normal_specials(seed, _p(out), n)
    return out


def gen_u128_pairs(seed: int, n: int) -> np.ndarray:
    """uint64[n, 2] = {lo, hi} little-endian 128-bit keys."""
    out = np.empty((n, 2), dtype=np.uint64)
    lib().oracle_gen_u128_pairs(seed, _p(out), n)
    return out


def gen_lognormal_sizes(seed: int, count: int, median: float = 2048.0, sigma: float = 0.7356,
                        lo: int = 32, hi: int = 1 << 22) -> np.ndarray:
    out = np.empty(count, dtype=np.uint64)
    lib().oracle_gen_lognormal_sizes(seed, median, sigma, lo, hi, _p(out), count)
    return out


def time_stable_sort_u32(keys: np.ndarray, vals: np.ndarray | None, reps: int = 3) -> float:
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint32)
    return float(lib().oracle_time_stable_sort_u32(_p(k), _p(v), k.size, reps))


def time_sort_par_u32(keys: np.ndarray, vals: np.ndarray | None, reps: int = 3) -> float:
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint32)
    return float(lib().oracle_time_sort_par_u32(_p(k), _p(v), k.size, reps))


def stable_sort_par_u32(keys: np.ndarray, vals: np.ndarray | None, threads: int):
    """The multi-threaded stable sort of the CPU baseline (result, for its test)."""
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint32)
    ok = np.empty_like(k)
    ov = None if v is None else np.empty_like(v)
    assert lib().oracle_stable_sort_par_u32(_p(k), _p(v), k.size, threads, _p(ok), _p(ov)) == 0
    return ok, ov


def time_stable_sort_par_u32(keys: np.ndarray, vals: np.ndarray | None, threads: int, reps: int = 3) -> float:
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint32)
    return float(lib().oracle_time_stable_sort_par_u32(_p(k), _p(v), k.size, threads, reps))


def pstl_parallel_backend() -> bool:
    return bool(lib().oracle_pstl_parallel_backend())
