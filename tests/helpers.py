"""Shared test helpers: golden fixture loading and key generation."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

KIND_UNSIGNED, KIND_SIGNED, KIND_FLOAT = 0, 1, 2

# (kind, key_bytes) of every key type the reference validates
# (benchmark/validate_sort.cuh:36-40, 59-63) plus IEEE half.
KEY_TYPES = [
    (0, 1), (1, 1), (0, 2), (1, 2), (2, 2), (0, 4), (1, 4), (2, 4),
    (0, 8), (1, 8), (2, 8), (0, 16), (1, 16), (2, 16),
]


def golden_names() -> list[str]:
    return sorted(f[:-5] for f in os.listdir(GOLDEN) if f.endswith(".json"))


def load_golden(name: str) -> dict:
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        g = json.load(f)
    kb = g["key_bytes"]
    ints = [int(k) for k in g["keys"]]
    raw = b"".join(k.to_bytes(kb, "little") for k in ints)
    g["key_bytes_array"] = np.frombuffer(raw, dtype=np.uint8).copy()
    g["n"] = len(ints)
    g["asc"] = np.array(g["asc"], dtype=np.uint64)
    g["desc"] = np.array(g["desc"], dtype=np.uint64)
    return g


def random_keys(kind: int, kb: int, n: int, seed: int, dup_frac: float = 0.25) -> np.ndarray:
    """n keys as uint8[n*kb]: random bit patterns, a share of duplicates, and
    for floating keys +-0, +-inf and NaNs mixed in."""
    rng = np.random.default_rng(seed)
    b = rng.integers(0, 256, size=n * kb, dtype=np.uint8).reshape(n, kb)
    if n >= 4:
        nd = int(n * dup_frac)
        src = rng.integers(0, n, size=nd)
        dst = rng.integers(0, n, size=nd)
        b[dst] = b[src]
        # a run of a single repeated key (equal-digit waves)
        r0 = int(rng.integers(0, n))
        b[r0:r0 + min(n - r0, 300)] = b[r0]
    if kind == KIND_FLOAT and n >= 16 and kb in (2, 4, 8):
        dt = {2: np.float16, 4: np.float32, 8: np.float64}[kb]
        sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 1.0, -1.0], dtype=dt)
        sp_b = sp.view(np.uint8).reshape(-1, kb)
        pos = rng.integers(0, n, size=len(sp))
        b[pos] = sp_b
    return b.reshape(-1).copy()


def random_values(vb: int, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed + 7919)
    return rng.integers(0, 256, size=n * vb, dtype=np.uint8)


def index_values(vb: int, n: int) -> np.ndarray:
    dt = {4: np.uint32, 8: np.uint64}[vb]
    return np.arange(n, dtype=dt).view(np.uint8).copy()
