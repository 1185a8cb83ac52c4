"""Counting-stage A/B (slime_window_ab): both methods, VALU word counting and
int8 MFMA window sums, against the CPU count of the same hashed planes.

The CPU count (window_ab_reference.reference) uses integral images and the
stream's 6-strip period; it is pinned against a loop over every cell of
every window on a small case, against hand-checked properties of the hash
planes (bit densities, planes past the last band empty), and against a
planted mistake (windows without the corner cells removed count more and
give another checksum)."""
import numpy as np
import pytest

import window_ab_reference as ref


def test_reference_matches_direct_loop():
    for bands, min_size in ((2, 20), (2, 26), (4, 24)):
        assert ref.reference(bands, 1, 9, min_size, 7) == ref.reference_direct(bands, 1, 9, min_size, 7)


def test_reference_planes():
    ring = ref.ring_bits(3, 4, 0x1234)
    assert ring.shape == (24, 9, 512)
    assert not ring[:, 5:, :].any()  # past the last band
    e = ring[:, 0, :].mean()
    h = ring[:, 1:5, :].mean()
    assert 0.47 < e < 0.53 and 0.11 < h < 0.14  # one hash, three ANDed
    # dword 0 of plane 0, row 0, workgroup 3: bit c of hash(index)
    idx = ((((3 * 24 + 0) * 9 + 0) * 16 + 0) * 3 + 0x1234) & 0xFFFFFFFF
    v = int(ref.ab_hash(np.uint32(idx)))
    assert [int(b) for b in ring[0, 0, :32]] == [(v >> c) & 1 for c in range(32)]


def test_reference_window_size_and_corners():
    """A full ring: every window counts 368 cells; without the corner
    removal it would count 400."""
    full = np.ones((24, 512), dtype=np.uint8)
    cnt = ref.window_counts(full)
    assert (cnt == 368).all()


def test_planted_corner_mistake_changes_result():
    """A reference that forgets the corner cells passes more windows."""
    good = ref.reference(4, 1, 20, 30, 11)
    orig = ref.window_counts

    def no_corners(cell):
        ext = np.concatenate([cell, cell], axis=0).astype(np.int64)
        out = np.zeros((6, ref.GROUPS), dtype=np.int64)
        for m in range(6):
            for g in range(ref.GROUPS):
                out[m, g] = ext[4 * m:4 * m + 20, 4 * g:4 * g + 20].sum()
        return out

    ref.window_counts = no_corners
    try:
        bad = ref.reference(4, 1, 20, 30, 11)
    finally:
        ref.window_counts = orig
    assert bad != good and bad[1] > good[1]


CASES = [(2, 24), (4, 24), (8, 24), (8, 30), (4, 400)]


@pytest.mark.gpu
@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("bands,min_size", CASES)
def test_window_ab_matches_cpu(gpu, method, bands, min_size):
    import slime_amd
    blocks, steps, seed = 3, 40, 0x5EED + bands
    print("window_ab method", method, "bands", bands, "min_size", min_size, flush=True)
    got = slime_amd.window_ab(method, bands, blocks, steps, min_size, seed)
    assert got is not None
    want = ref.reference(bands, blocks, steps, min_size, seed)
    assert got[:2] == want
    if min_size == 400:
        assert want == (0, 0)
    else:
        assert want[1] > 1000


@pytest.mark.gpu
def test_window_ab_methods_agree_many_blocks(gpu):
    import slime_amd
    a = slime_amd.window_ab(0, 8, 600, 64, 30, 99)
    b = slime_amd.window_ab(1, 8, 600, 64, 30, 99)
    assert a is not None and b is not None and a[:2] == b[:2] and a[1] > 0


@pytest.mark.gpu
def test_window_ab_rejects_bad_arguments(gpu):
    import slime_amd
    assert slime_amd.window_ab(2, 8, 1, 10, 30) is None
    assert slime_amd.window_ab(0, 3, 1, 10, 30) is None
    assert slime_amd.window_ab(1, 8, 1, 4, 30) is None
    assert slime_amd.window_ab(1, 8, 1, 10, 0) is None
