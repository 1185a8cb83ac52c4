#!/usr/bin/env python3
"""Writes biome_data.inc: the data the biome filter needs to reproduce
Minecraft's multi-noise overworld biomes (versions 1.18 to 26.2).

  * the seeds of the climate and octave noises: MD5 digests of Minecraft's
    noise names ("minecraft:temperature", "octave_-9", ...), computed here;
  * the biome parameter search trees of each version and their parameter
    intervals, and the terrain-offset cubic spline: Minecraft data as encoded
    by cubiomes (MIT, see gpu_noise/LICENSE.cubiomes of the reference), read
    from the reference's gpu_noise/tables/btree*.h and
    gpu_noise/spline_generated.cuh.

Run in a checkout that has the reference next to it:
    python3 slime_amd/csrc/gen_biome_data.py /root/reference
"""
import hashlib
import os
import re
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def md5_pair(name):
    d = hashlib.md5(name.encode()).digest()
    return int.from_bytes(d[:8], "big"), int.from_bytes(d[8:], "big")


def read_tree(path):
    text = open(path).read()
    steps = [int(x) for x in re.search(r"_steps\[\]\s*=\s*\{([^}]*)\}", text).group(1).split(",")
             if x.strip()]
    pblock = re.search(r"_param\[\]\[2\]\s*=\s*\{(.*?)\n\};", text, re.S).group(1)
    pblock = re.sub(r"//[^\n]*", "", pblock)
    params = [(int(a), int(b)) for a, b in re.findall(r"\{\s*(-?\d+)\s*,\s*(-?\d+)\s*\}", pblock)]
    nblock = re.search(r"_nodes\[\]\s*=\s*\{(.*?)\n\};", text, re.S).group(1)
    nblock = re.sub(r"//[^\n]*", "", nblock)
    nodes = [int(x, 16) for x in re.findall(r"0x([0-9A-Fa-f]{16})", nblock)]
    return steps, params, nodes


def f32(s):
    """A float literal as the exact binary32 value (hex bits)."""
    s = s.rstrip("f")
    return struct.unpack("<I", struct.pack("<f", float(s)))[0]


def read_splines(path):
    """Each spline as (coordinate, [(location, derivative, value, child)]),
    child = -1 for a constant value."""
    text = open(path).read()
    bodies = re.findall(r"float spline_(\d+)\(const float\* v\)\n\{(.*?)\n\}\n", text, re.S)
    splines = {}
    num = r"(-?[0-9.e+-]+f)"
    val = r"(spline_(\d+)\(v\)|-?[0-9.e+-]+f)"
    for sid, body in bodies:
        coord = int(re.search(r"const float f = v\[(\d)\];", body).group(1))
        pts = []

        def value(tok, child):
            return (-1, f32("0f")) if child else None

        first = re.search(r"if \(f <= " + num + r"\) \{\s*const float n = " + val +
                          r";\s*return n \+ " + num + r" \* \(f - " + num + r"\);", body)
        loc0, v0, c0, d0 = first.group(1), first.group(2), first.group(3), first.group(4)
        pts.append([f32(loc0), f32(d0), v0, c0])
        for m in re.finditer(
                r"const float g = " + num + r";\s*const float h = " + num + r";.*?"
                r"const float l = " + num + r";\s*const float m = " + num + r";\s*"
                r"const float n = " + val + r";\s*const float o = " + val + r";", body, re.S):
            g, h, l, mm = m.group(1), m.group(2), m.group(3), m.group(4)
            n, nc, o, oc = m.group(5), m.group(6), m.group(7), m.group(8)
            prev = pts[-1]
            assert prev[0] == f32(g) and prev[1] == f32(l) and prev[2] == n, (sid, g)
            pts.append([f32(h), f32(mm), o, oc])
        last = re.search(r"\n    const float n = " + val + r";\s*return n \+ " + num +
                         r" \* \(f - " + num + r"\);\s*$", body)
        assert last and pts[-1][2] == last.group(1) and pts[-1][1] == f32(last.group(3)), sid
        assert pts[-1][0] == f32(last.group(4)), sid
        out = []
        for loc, der, v, child in pts:
            if child is not None:
                out.append((loc, der, 0, int(child)))
            else:
                out.append((loc, der, f32(v), -1))
        splines[int(sid)] = (coord, out)
    assert sorted(splines) == list(range(len(splines)))
    return [splines[i] for i in range(len(splines))]


NOISES = [  # (slot, name, large name, first octave, large first octave, amplitudes)
    ("shift", "minecraft:offset", None, -3, None, [1, 1, 1, 0]),
    ("temperature", "minecraft:temperature", "minecraft:temperature_large", -10, -12,
     [1.5, 0, 1, 0, 0, 0]),
    ("humidity", "minecraft:vegetation", "minecraft:vegetation_large", -8, -10,
     [1, 1, 0, 0, 0, 0]),
    ("continentalness", "minecraft:continentalness", "minecraft:continentalness_large", -9, -11,
     [1, 1, 2, 2, 2, 1, 1, 1, 1]),
    ("erosion", "minecraft:erosion", "minecraft:erosion_large", -9, -11, [1, 1, 0, 1, 1]),
    ("weirdness", "minecraft:ridge", None, -7, None, [1, 2, 1, 0, 0, 0]),
]


# (file suffix, first version code using it, tree shape family): the version
# selection of MinecraftGPUNoise.cu:1785-1820. Family 0 (1.18, 1.19.2): steps
# 1111, 111, 11, 1 and 10 children per node; family 1 (1.19.4 on): steps
# 1555, 259, 43, 7, 1 and 6 children.
TREES = [("262", 2602, 1), ("215", 1215, 1), ("21wd", 1214, 1), ("20", 1206, 1),
         ("19", 1194, 1), ("192", 1192, 0), ("18", 1180, 0)]
FAMILY_STEPS = {0: [1111, 111, 11, 1, 0], 1: [1555, 259, 43, 7, 1, 0]}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    trees = []
    for suffix, version, family in TREES:
        steps, params, nodes = read_tree(
            os.path.join(ref, "gpu_noise", "tables", f"btree{suffix}.h"))
        assert steps == FAMILY_STEPS[family], (suffix, steps)
        trees.append((suffix, version, family, params, nodes))
    splines = read_splines(os.path.join(ref, "gpu_noise", "spline_generated.cuh"))
    L = []
    w = L.append
    w("// Generated by gen_biome_data.py; do not edit. Data of Minecraft's overworld")
    w("// multi-noise biome source (versions 1.18 to 26.2): noise seeds (MD5 of the")
    w("// noise names), the biome parameter search trees and the terrain-offset")
    w("// spline, the latter two as encoded by cubiomes (Copyright (c) 2020 Cubitect,")
    w("// MIT License; the notice is in LICENSE.cubiomes beside this file).")
    w("// BIOME_TABLE marks the tables device code reads (the includer defines it).")
    w("")
    w("// Noise definitions: MD5 seed halves of the name (and of the large-biome")
    w("// name), first octave, number of octaves, amplitudes.")
    w("static const BiomeNoiseDef kBiomeNoises[6] = {")
    for slot, name, large, omin, lomin, amps in NOISES:
        lo, hi = md5_pair(name)
        llo, lhi = md5_pair(large) if large else (lo, hi)
        a = ", ".join(repr(float(x)) for x in amps + [0] * (9 - len(amps)))
        w(f"    {{0x{lo:016x}ull, 0x{hi:016x}ull, 0x{llo:016x}ull, 0x{lhi:016x}ull, {omin}, "
          f"{lomin if lomin is not None else omin}, {len(amps)}, {{{a}}}}},  // {slot}")
    w("};")
    w("")
    w("// Octave seeds: MD5 halves of \"octave_<i>\" for i = -16..0.")
    w("static const uint64_t kOctaveSeeds[17][2] = {")
    for i in range(-16, 1):
        lo, hi = md5_pair(f"octave_{i}")
        w(f"    {{0x{lo:016x}ull, 0x{hi:016x}ull}},  // octave_{i}")
    w("};")
    w("")
    for suffix, version, family, params, nodes in trees:
        w(f"// Parameter tree of version code >= {version} (family {family}).")
        w(f"static const int32_t kTreeParams{suffix}[{len(params)}][2] = {{")
        for i in range(0, len(params), 6):
            w("    " + ", ".join(f"{{{a}, {b}}}" for a, b in params[i:i + 6]) + ",")
        w("};")
        w(f"static const uint64_t kTreeNodes{suffix}[{len(nodes)}] = {{")
        for i in range(0, len(nodes), 4):
            w("    " + ", ".join(f"0x{n:016x}ull" for n in nodes[i:i + 4]) + ",")
        w("};")
    w("")
    w("// The trees, newest first: first version code, shape family, nodes, node")
    w("// count, intervals, interval count.")
    w(f"static const BiomeTreeDef kTreeDefs[{len(trees)}] = {{")
    for suffix, version, family, params, nodes in trees:
        w(f"    {{{version}, {family}, kTreeNodes{suffix}, {len(nodes)}, &kTreeParams{suffix}[0][0], "
          f"{len(params)}}},")
    w("};")
    w(f"#define BIOME_TREE_MAX_PARAMS {max(len(t[3]) for t in trees)}")
    w("")
    npts = sum(len(p) for _, p in splines)
    w("// Splines: coordinate (0 continentalness, 1 erosion, 2 peaks-and-valleys,")
    w("// 3 weirdness), first point, point count. Points: location, derivative and")
    w("// value as binary32 bits, child spline (-1: the value is a constant).")
    w(f"BIOME_TABLE BiomeSpline kSplines[{len(splines)}] = {{")
    first = 0
    for coord, pts in splines:
        w(f"    {{{coord}, {first}, {len(pts)}}},")
        first += len(pts)
    w("};")
    w(f"BIOME_TABLE BiomeSplinePoint kSplinePoints[{npts}] = {{")
    for coord, pts in splines:
        for loc, der, v, child in pts:
            w(f"    {{0x{loc:08x}u, 0x{der:08x}u, 0x{v:08x}u, {child}}},")
    w("};")
    out = os.path.join(HERE, "biome_data.inc")
    open(out, "w").write("\n".join(L) + "\n")
    print(f"wrote {out}: {len(trees)} trees ({sum(len(t[4]) for t in trees)} nodes), "
          f"{len(splines)} splines / {npts} points")


if __name__ == "__main__":
    main()
