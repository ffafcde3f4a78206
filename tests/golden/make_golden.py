"""Writes the golden fixtures of tests/golden/ from first principles.

Nothing here uses the oracle or the library: the expected orders come from
Python's stable `sorted` over numeric keys (IEEE totalOrder for floating
keys: -NaN by descending payload, -inf .. -0, +0 .. +inf, +NaN by ascending
payload -- the order the reference's sign-magnitude twiddle produces,
include/rsort/detail/tuning.cuh:132-148).  Descending is `sorted(...,
reverse=True)`, which Python keeps stable, matching the reference's stable
sort on the complemented key (include/rsort/rsort.cuh:33-50).

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import random
import struct

HERE = os.path.dirname(os.path.abspath(__file__))


def float_total_key(bits: int, exp_bits: int, man_bits: int):
    width = 1 + exp_bits + man_bits
    sign = bits >> (width - 1)
    exp = (bits >> man_bits) & ((1 << exp_bits) - 1)
    man = bits & ((1 << man_bits) - 1)
    is_nan = exp == (1 << exp_bits) - 1 and man != 0
    if is_nan:
        return (0, -man) if sign else (2, man)
    # finite or infinite: numeric value with -0 < +0
    if exp == (1 << exp_bits) - 1:
        val = float("-inf") if sign else float("inf")
    else:
        bias = (1 << (exp_bits - 1)) - 1
        if exp == 0:
            mag = man * 2.0 ** (1 - bias - man_bits)
        else:
            mag = (1 + man / 2.0 ** man_bits) * 2.0 ** (exp - bias)
        val = -mag if sign else mag
    return (1, val, -1 if sign else 1)


def specials(exp_bits: int, man_bits: int, rng: random.Random, extra: int):
    width = 1 + exp_bits + man_bits
    e_all = (1 << exp_bits) - 1
    pats = set()
    for sign in (0, 1):
        s = sign << (width - 1)
        pats.update({
            s,                                        # +-0
            s | 1,                                    # smallest subnormal
            s | ((1 << man_bits) - 1),                # largest subnormal
            s | (1 << man_bits),                      # smallest normal
            s | (e_all << man_bits) - (1 << man_bits) | ((1 << man_bits) - 1),  # largest finite
            s | ((e_all - 1) << man_bits),            # 2^(emax)
            s | (e_all << man_bits),                  # inf
            s | (e_all << man_bits) | 1,              # signalling NaN, payload 1
            s | (e_all << man_bits) | (1 << (man_bits - 1)),  # quiet NaN
            s | (e_all << man_bits) | ((1 << man_bits) - 1),  # all-ones payload
        })
        bias = (1 << (exp_bits - 1)) - 1
        pats.add(s | (bias << man_bits))              # +-1.0
    pats = sorted(pats)
    for _ in range(extra):
        pats.append(rng.getrandbits(width))
    # duplicates for stability
    pats += [pats[rng.randrange(len(pats))] for _ in range(extra // 2)]
    rng.shuffle(pats)
    return pats


def stable_orders(keys, keyfn):
    idx = list(range(len(keys)))
    asc = sorted(idx, key=lambda i: keyfn(keys[i]))
    desc = sorted(idx, key=lambda i: keyfn(keys[i]), reverse=True)
    return asc, desc


def write(name, obj):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, separators=(",", ":"))
        f.write("\n")


def main():
    rng = random.Random(20260101)
    fixtures = {}

    for name, eb, mb in (("half", 5, 10), ("float32", 8, 23), ("float64", 11, 52)):
        keys = specials(eb, mb, rng, 40)
        asc, desc = stable_orders(keys, lambda b, eb=eb, mb=mb: float_total_key(b, eb, mb))
        width = 1 + eb + mb
        fixtures[f"{name}_specials"] = {"kind": 2, "key_bytes": width // 8, "keys": keys, "asc": asc, "desc": desc}

    # every int8 and uint8 value, shuffled, plus duplicates
    for kind, name in ((0, "uint8_all"), (1, "int8_all")):
        vals = list(range(256)) * 2
        rng.shuffle(vals)
        if kind == 1:
            keyfn = lambda b: b - 256 if b >= 128 else b
        else:
            keyfn = lambda b: b
        asc, desc = stable_orders(vals, keyfn)
        fixtures[name] = {"kind": kind, "key_bytes": 1, "keys": vals, "asc": asc, "desc": desc}

    # int16 signed with extremes and heavy duplicates
    k16 = [0x8000, 0x7FFF, 0, 1, 0xFFFF, 0x8001] + [rng.getrandbits(16) for _ in range(200)]
    k16 += [k16[rng.randrange(len(k16))] for _ in range(100)]
    rng.shuffle(k16)
    asc, desc = stable_orders(k16, lambda b: b - 65536 if b >= 32768 else b)
    fixtures["int16_mixed"] = {"kind": 1, "key_bytes": 2, "keys": k16, "asc": asc, "desc": desc}

    # u32 stability: few distinct keys, many equal
    ks = [rng.choice([0, 1, 2, 3, 255, 256, 0xFFFFFFFF, 0x80000000, 0x12345678]) for _ in range(3000)]
    asc, desc = stable_orders(ks, lambda b: b)
    fixtures["uint32_stability"] = {"kind": 0, "key_bytes": 4, "keys": ks, "asc": asc, "desc": desc}

    # int64 extremes
    k64 = [0, 1, (1 << 64) - 1, 1 << 63, (1 << 63) - 1] + [rng.getrandbits(64) for _ in range(300)]
    k64 += [k64[rng.randrange(len(k64))] for _ in range(150)]
    rng.shuffle(k64)
    asc, desc = stable_orders(k64, lambda b: b - (1 << 64) if b >= (1 << 63) else b)
    fixtures["int64_mixed"] = {"kind": 1, "key_bytes": 8, "keys": k64, "asc": asc, "desc": desc}

    # 128-bit unsigned and signed keys (hi drawn from few values so low bytes decide)
    k128 = [(rng.randrange(4) << 64 | rng.getrandbits(64)) for _ in range(300)]
    k128 += [(1 << 128) - 1, 1 << 127, 0, (1 << 127) - 1]
    k128 += [k128[rng.randrange(len(k128))] for _ in range(100)]
    rng.shuffle(k128)
    asc, desc = stable_orders(k128, lambda b: b)
    fixtures["uint128_mixed"] = {"kind": 0, "key_bytes": 16, "keys": [str(k) for k in k128], "asc": asc,
                                 "desc": desc}
    asc, desc = stable_orders(k128, lambda b: b - (1 << 128) if b >= (1 << 127) else b)
    fixtures["int128_mixed"] = {"kind": 1, "key_bytes": 16, "keys": [str(k) for k in k128], "asc": asc,
                                "desc": desc}

    # 16-byte floating key = long double reinterpreted as a sign-magnitude 128-bit word
    # (include/rsort/detail/radix_kernel.cuh:351-356): order by sign, then magnitude
    def sm128(b):
        mag = b & ((1 << 127) - 1)
        return (0, -mag) if b >> 127 else (1, mag)

    asc, desc = stable_orders(k128, sm128)
    fixtures["longdouble_bits_mixed"] = {"kind": 2, "key_bytes": 16, "keys": [str(k) for k in k128], "asc": asc,
                                         "desc": desc}

    for name, obj in fixtures.items():
        write(f"{name}.json", obj)
    print(f"wrote {len(fixtures)} fixtures to {HERE}")


if __name__ == "__main__":
    main()
