"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Run in a container that has the reference tree (it needs oracle/_ref/, built by
`make -C oracle ref` from the reference's unmodified SlimeCore.cpp):

    python tests/golden/make_golden.py

Fixtures are data only: inputs and the reference's outputs.
  predicate.json     per-chunk slime bits from the reference's build_chunk_cache
                     (SlimeCore.cpp:234-244): random, edge and constructed
                     rejection-path coordinates
  c1_hits.json       the full ranked hit list of the CPU-runnable config
                     (seed 0x123456789ABCDEF0, radius 2048, threshold 40)
  small_searches.json  a battery of small searches through
                     search_slime_clusters_centered (SlimeCore.cpp:643-651)
"""
import ctypes as C
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as P  # noqa: E402

MASK48 = (1 << 48) - 1
JAVA_MUL = 0x5DEECE66D
JAVA_INV = 0xDFE05BCB1365  # JAVA_MUL^-1 mod 2^48
SCRAMBLE = 0x5DEECE66D ^ 0x3AD8025F
INT32_MAX = (1 << 31) - 1


def s64(v):
    v &= (1 << 64) - 1
    return v - (1 << 64) if v >= (1 << 63) else v


def x_term(x):
    u = x & 0xFFFFFFFF
    a = (u * u * 4987142) & 0xFFFFFFFF
    b = (u * 5947611) & 0xFFFFFFFF
    a = a - (1 << 32) if a >= (1 << 31) else a
    b = b - (1 << 32) if b >= (1 << 31) else b
    return a + b


def z_term(z):
    u = z & 0xFFFFFFFF
    sq = (u * u) & 0xFFFFFFFF
    sq = sq - (1 << 32) if sq >= (1 << 31) else sq
    lin = (u * 389711) & 0xFFFFFFFF
    lin = lin - (1 << 32) if lin >= (1 << 31) else lin
    return sq * 4392871 + lin


def seed_for_state1(state1, x, z, high16):
    """A world seed whose chunk (x, z) has first LCG output `state1`."""
    state0 = ((state1 - 11) * JAVA_INV) & MASK48
    s = state0 ^ SCRAMBLE
    seed = (s - x_term(x) - z_term(z)) & MASK48
    return s64(seed | (high16 << 48))


def load_ref():
    lib = P.ref_lib(avx2=P.cpu_has_avx2())
    if lib is None:
        sys.exit("oracle/_ref not built: run `make -C oracle ref` first")
    fn = getattr(lib, "_Z17build_chunk_cachelllPjRlS0_")
    fn.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.POINTER(C.c_uint32),
                   C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    fn.restype = None
    return lib, fn


def ref_bits(cache_fn, seed, cx, cz):
    """21x21 slime bits around chunk (cx, cz) from the reference."""
    cache = (C.c_uint32 * 21)()
    bx, bz = C.c_int64(), C.c_int64()
    # build_chunk_cache takes block coordinates and starts 10 chunks before.
    cache_fn(seed, cx * 16 + 8, cz * 16 + 8, cache, C.byref(bx), C.byref(bz))
    assert bx.value == cx - 10 and bz.value == cz - 10
    return [(cache[i] >> j) & 1 for i in range(21) for j in range(21)]


def predicate_fixture(cache_fn, rng):
    cases = []

    def add(seed, cx, cz, kind):
        bits = ref_bits(cache_fn, seed, cx, cz)
        # keep the centre chunk and a 21x21 block packed as hex rows
        rows = []
        for i in range(21):
            v = 0
            for j in range(21):
                v |= bits[i * 21 + j] << j
            rows.append(v)
        cases.append({"seed": seed, "cx": cx, "cz": cz, "kind": kind, "rows": rows})

    edges = [-(1 << 31) + 10, -65536 * 3 - 5, -131072, -65536, -17, -1, 0, 1, 17, 65536,
             131072, 1875000, -1875000, (1 << 31) - 11]
    for x in edges:
        for z in edges[::3]:
            add(s64(rng.getrandbits(64)), x, z, "edge")
    for _ in range(24):
        add(s64(rng.getrandbits(64)), rng.randint(-(1 << 31) + 10, (1 << 31) - 11),
            rng.randint(-(1 << 31) + 10, (1 << 31) - 11), "random")
    # Constructed rejection path: force the first draw of the centre chunk
    # into [2^31-8, 2^31) (bits) so Java's nextInt(10) redraws.
    for j in range(8):
        for rep in range(3):
            bits = INT32_MAX - 7 + j
            state1 = (bits << 17) | rng.getrandbits(17)
            cx = rng.randint(-3000000, 3000000)
            cz = rng.randint(-3000000, 3000000)
            seed = seed_for_state1(state1, cx, cz, rng.getrandbits(16))
            add(seed, cx, cz, f"reject{j}")
    return cases


def search_case(lib, params):
    total, rows = P.ref_search(lib, params["seed"], params["rd_max"], params["min_size"],
                               params["max_size"], params["rd_min"], params["cx"],
                               params["cz"], params["max_results"], threads=4)
    out = dict(params)
    out["total"] = total
    out["rows"] = [list(r[:3]) for r in rows]
    # obs/afk columns are fixed for precise_afk == 0: obs 0, afk == centre
    assert all(r[3] == 0 and r[4] == r[1] and r[5] == r[2] for r in rows)
    return out


def small_searches(lib, rng):
    cases = []
    for i in range(64):
        rd = rng.choice([0, 1, 2, 5, 8, 17, 31, 64, 100, 150, 255, 256, 300, 511, 700])
        lim = INT32_MAX - rd - 8
        if i % 8 == 0:
            cx, cz = lim, -lim - 1
        elif i % 8 == 1:
            cx, cz = -lim - 1, lim
        elif i % 8 == 2:
            cx = rng.choice([-1, 1]) * rng.randint(1 << 26, (1 << 27))
            cz = rng.choice([-1, 1]) * rng.randint(1 << 26, (1 << 27))
        else:
            cx, cz = rng.randint(-2000000, 2000000), rng.randint(-2000000, 2000000)
        if rd < 64:
            mn = rng.choice([0, 1, 10, 20, 25, 28, 30, 33, 36])
        else:
            mn = rng.choice([30, 35, 38, 40, 42, 45, 48, 50, 55, 221])
        mx = rng.choice([221, 221, 221, mn, mn + 3, min(221, mn + 10)])
        rmin = rng.choice([0, 0, 0, rd // 3, rd // 2, rd])
        k = rng.choice([1, 2, 7, 100, 1000, 4000])
        cases.append(search_case(lib, {
            "seed": s64(rng.getrandbits(64)), "rd_max": rd, "min_size": mn, "max_size": mx,
            "rd_min": rmin, "cx": cx, "cz": cz, "max_results": k}))
    # searches whose region contains constructed rejection-path chunks
    for j in range(8):
        for rep in range(2):
            bits = INT32_MAX - 7 + j
            state1 = (bits << 17) | rng.getrandbits(17)
            cx, cz = rng.randint(-3000000, 3000000), rng.randint(-3000000, 3000000)
            seed = seed_for_state1(state1, cx + rng.randint(-20, 20), cz + rng.randint(-20, 20),
                                   rng.getrandbits(16))
            cases.append(search_case(lib, {
                "seed": seed, "rd_max": 40, "min_size": rng.choice([0, 20, 25, 30]),
                "max_size": 221, "rd_min": 0, "cx": cx, "cz": cz, "max_results": 1500}))
    # invalid arguments -> -1 (SlimeCore.cpp:378-384)
    bad = [
        dict(rd_max=-1), dict(min_size=-1), dict(min_size=222), dict(max_size=10, min_size=20),
        dict(rd_min=5, rd_max=4), dict(rd_min=-1), dict(max_results=0),
        dict(cx=INT32_MAX - 8, rd_max=1), dict(cz=-(1 << 31) + 8, rd_max=1),
        dict(rd_max=(INT32_MAX - 17) // 2 + 1),
    ]
    for b in bad:
        p = {"seed": 1, "rd_max": 10, "min_size": 40, "max_size": 221, "rd_min": 0,
             "cx": 0, "cz": 0, "max_results": 10}
        p.update(b)
        cases.append(search_case(lib, p))
    return cases


def main():
    rng = random.Random(0x5EED5EED)
    lib, cache_fn = load_ref()
    pred = predicate_fixture(cache_fn, rng)
    with open(os.path.join(HERE, "predicate.json"), "w") as f:
        json.dump(pred, f, separators=(",", ":"))
    seed = 0x123456789ABCDEF0
    total, rows = P.ref_search(lib, seed, 2048, 40, 221, 0, 0, 0, 65536, threads=8)
    with open(os.path.join(HERE, "c1_hits.json"), "w") as f:
        json.dump({"seed": seed, "rd_max": 2048, "min_size": 40, "max_size": 221, "rd_min": 0,
                   "cx": 0, "cz": 0, "max_results": 65536, "total": total,
                   "rows": [list(r[:3]) for r in rows]}, f, separators=(",", ":"))
    cases = small_searches(lib, rng)
    with open(os.path.join(HERE, "small_searches.json"), "w") as f:
        json.dump(cases, f, separators=(",", ":"))
    print(f"predicate cases {len(pred)}, c1 total {total}, small searches {len(cases)}")


if __name__ == "__main__":
    main()
