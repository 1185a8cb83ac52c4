"""CPU count of the counting-stage A/B's windows (slime_window_ab), from the
same hashed bitplanes the kernels fill their LDS ring with.

Ring of one workgroup: 24 rows x 9 planes x 16 dwords; dword (r, p, d) of
workgroup b has index i = (((b * 24 + r) * 9 + p) * 16 + d) * 3 + seed (mod
2^32); plane 0 (E) is hash(i), a band's hit plane p = band + 1 is
hash(i) & hash(i + 1) & hash(i + 2), planes past the last band are zero.
Chunk column c of a row is bit c % 32 of dword c // 32. Stream row s is ring
row s % 24; strip k is stream rows 4k..4k+3.

Band-phase q = 2 * band + phase counts the cells hit & E when band ^ phase is
odd, hit & ~E when it is even (the scan's E / ~E selection of a phase). The
window of group g (0..123) at strip m (0 .. steps - 5): stream rows
4m..4m+19, chunk columns 4g..4g+19, less rows 0, 1, 18, 19 of the window in
columns 4g..4g+3 and 4g+16..4g+19 (368 cells). A window passes when its count
reaches min_size; the checksum is the wrapping 32-bit sum of
m * 0x9E3779B1 + q * 0x85EBCA77 + g * 0xC2B2AE3D over passing windows.
"""
import numpy as np

RING, PLANES, WORDS, GROUPS = 24, 9, 16, 124
M32 = 0xFFFFFFFF


def ab_hash(x):
    x = np.asarray(x, dtype=np.uint32).copy()
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x7FEB352D)
    x ^= x >> np.uint32(15)
    x *= np.uint32(0x846CA68B)
    x ^= x >> np.uint32(16)
    return x


def ring_bits(block, bands, seed):
    """uint8 [24, 9, 512]: the ring's bits of one workgroup."""
    r = np.arange(RING, dtype=np.uint64)[:, None, None]
    p = np.arange(PLANES, dtype=np.uint64)[None, :, None]
    d = np.arange(WORDS, dtype=np.uint64)[None, None, :]
    idx = ((((block * RING + r) * PLANES + p) * WORDS + d) * 3 + seed) & M32
    idx = idx.astype(np.uint32)
    v = ab_hash(idx)
    hit = v & ab_hash(idx + np.uint32(1)) & ab_hash(idx + np.uint32(2))
    planes = np.where(p.astype(np.int64) == 0, v, hit)
    planes = np.where(p.astype(np.int64) <= bands, planes, np.uint32(0)).astype(np.uint32)
    bits = (planes[..., None] >> np.arange(32, dtype=np.uint32)) & np.uint32(1)
    return bits.reshape(RING, PLANES, WORDS * 32).astype(np.uint8)


def cells(ring, band, phase):
    """[24, 512] cells one band-phase counts."""
    e = ring[:, 0, :]
    h = ring[:, band + 1, :]
    return h & (e if (band ^ phase) & 1 else 1 - e)


def window_counts(cell):
    """[6, 124] counts of the windows at strips m = 0..5 (the stream repeats
    every 6 strips)."""
    ext = np.concatenate([cell, cell], axis=0).astype(np.int64)  # 48 rows
    integ = np.zeros((ext.shape[0] + 1, ext.shape[1] + 1), dtype=np.int64)
    integ[1:, 1:] = ext.cumsum(0).cumsum(1)

    def box(r0, r1, c0, c1):
        return integ[r1, c1] - integ[r0, c1] - integ[r1, c0] + integ[r0, c0]

    g = np.arange(GROUPS)
    out = np.zeros((6, GROUPS), dtype=np.int64)
    for m in range(6):
        r0 = 4 * m
        tot = box(r0, r0 + 20, 4 * g, 4 * g + 20)
        for rr in (0, 1, 18, 19):
            tot -= box(r0 + rr, r0 + rr + 1, 4 * g, 4 * g + 4)
            tot -= box(r0 + rr, r0 + rr + 1, 4 * g + 16, 4 * g + 20)
        out[m] = tot
    return out


def weight(m, q, g):
    return (m * 0x9E3779B1 + q * 0x85EBCA77 + g * 0xC2B2AE3D) & M32


def reference(bands, blocks, steps, min_size, seed):
    """(checksum, passes) over all workgroups, as slime_window_ab reports."""
    cs, n = 0, 0
    nwin = steps - 4
    for b in range(blocks):
        ring = ring_bits(b, bands, seed)
        for q in range(2 * bands):
            cnt = window_counts(cells(ring, q >> 1, q & 1))
            for m6 in range(6):
                ms = np.arange(m6, nwin, 6, dtype=np.int64)
                if ms.size == 0:
                    continue
                gs = np.nonzero(cnt[m6] >= min_size)[0]
                if gs.size == 0:
                    continue
                n += int(ms.size) * int(gs.size)
                msum = int(ms.sum())
                gsum = int(gs.sum())
                cs += (msum * int(gs.size) * 0x9E3779B1 + int(ms.size) * int(gs.size) * q * 0x85EBCA77 +
                       int(ms.size) * gsum * 0xC2B2AE3D)
    return cs & M32, n


def reference_direct(bands, blocks, steps, min_size, seed):
    """The same by a loop over every window and cell (small cases only)."""
    cs, n = 0, 0
    for b in range(blocks):
        ring = ring_bits(b, bands, seed)
        for q in range(2 * bands):
            band, phase = q >> 1, q & 1
            for m in range(steps - 4):
                for g in range(GROUPS):
                    cnt = 0
                    for rr in range(20):
                        row = ring[(4 * m + rr) % RING]
                        for cc in range(20):
                            if rr in (0, 1, 18, 19) and (cc < 4 or cc >= 16):
                                continue
                            c = 4 * g + cc
                            e = int(row[0, c])
                            hit = int(row[band + 1, c])
                            cnt += hit & (e if (band ^ phase) & 1 else 1 - e)
                    if cnt >= min_size:
                        n += 1
                        cs = (cs + weight(m, q, g)) & M32
    return cs, n
