"""Known-answer tests of the oracle's scalar helpers and constructor.

Values come from the formulas at the cited reference lines, computed
independently here (bit interleaving by loop, the lowbias32 hash in numpy,
closed forms), and from the survey's Appendix B table for the hash lanes.
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def morton_loop(x, y):
    """Bit interleave by an explicit loop: x in even bits, y in odd bits (src/apic.rs:175-198)."""
    k = 0
    for b in range(16):
        k |= ((x >> b) & 1) << (2 * b)
        k |= ((y >> b) & 1) << (2 * b + 1)
    return k


@pytest.mark.parametrize("xy,key", [((1, 0), 0x1), ((0, 1), 0x2), ((1, 1), 0x3), ((2, 3), 0xE),
                                    ((1023, 1023), 0xFFFFF), ((4095, 2047), 0x7FFFFF),
                                    ((65535, 65535), 0xFFFFFFFF)])
def test_morton_known_answers(xy, key):
    assert po.morton(*xy) == key
    assert morton_loop(*xy) == key


def test_morton_matches_loop_on_random_pairs():
    rng = np.random.default_rng(7)
    for x, y in rng.integers(0, 1 << 16, size=(500, 2)):
        assert po.morton(int(x), int(y)) == morton_loop(int(x), int(y))


def lowbias32_numpy(idx, seed):
    m = (np.asarray(idx, dtype=np.uint32) ^ np.uint32(seed)).astype(np.uint64)
    m ^= m >> 16
    m = (m * 0x7FEB352D) & 0xFFFFFFFF
    m ^= m >> 15
    m = (m * 0x846CA68B) & 0xFFFFFFFF
    m ^= m >> 16
    bits = ((m & 0x007FFFFF) | 0x3F800000).astype(np.uint32)
    return bits.view(np.float32) - np.float32(1.5)


APPENDIX_B = {
    0: [-0.5, -0.4252701, -0.35054, 0.3899494, 0.4872686, 0.0455701, 0.279899, -0.4091218],
    100: [-0.3646634, 0.2679081, -0.3929886, -0.3902403, 0.3893809, -0.1027828, -0.0162574, -0.3644079],
}


@pytest.mark.parametrize("seed", [0, 100])
def test_random_simd_known_answers(seed):
    got = po.random_chunk(0, seed)
    assert np.allclose(got, APPENDIX_B[seed], atol=2e-7)
    assert np.array_equal(got, lowbias32_numpy(np.arange(8), seed))


def test_random_simd_index_xor_timestamp_collision():
    """Hash input is index ^ timestamp (src/apic.rs:960): seed 1 swaps adjacent lanes of seed 0."""
    a, b = po.random_chunk(0, 0), po.random_chunk(0, 1)
    assert np.array_equal(b, a.reshape(4, 2)[:, ::-1].reshape(-1))


def test_random_simd_range_and_chunks():
    for chunk in (0, 1, 1000, 123456):
        v = po.random_chunk(chunk, 7)
        assert np.all(v >= -0.5) and np.all(v < 0.5)
        assert np.array_equal(v, lowbias32_numpy(np.arange(8) + chunk * 8, 7))


def test_temporal_weight():
    """w_t(tau) = max(2.1875 * max(1 - (tau - 1/2)^2, 0)^3, 1e-5) (src/apic.rs:31-46)."""
    assert po.temporal_w(0.0) == 0.9228515625
    assert po.temporal_w(0.5) == 2.1875
    assert po.temporal_w(-0.5) == np.float32(1e-5)
    assert po.temporal_w(1.5) == np.float32(1e-5)
    for t in (0.1, 0.3, 0.45):  # symmetric about 1/2
        assert po.temporal_w(0.5 - t) == po.temporal_w(0.5 + t)


def test_theta_branches():
    """calculate_theta (src/apic.rs:1628-1638) with target density 9 (threshold 4.5)."""
    assert po.theta(5.0, 5.0 + 5e-6) == 1.0                       # |diff| <= 1e-5
    assert po.theta(0.0, 6.0) == np.float32((6.0 - 4.5) / 6.0)   # (rc - 4.5) / (rc - rho)
    assert po.theta(0.0, 4.6) == np.float32(0.1)                  # clamped below
    assert po.theta(7.0, 8.0) == 1.0                              # clamped above
    assert po.theta(20.0, 2.0) == np.float32(np.float32(2.0 - 4.5) / np.float32(2.0 - 20.0))  # air denser


def test_dot_product():
    a = np.arange(1, 20, dtype=np.float32)
    b = np.ones(19, dtype=np.float32)
    assert po.dot(a, b) == 190.0  # exact small integers, with a 3-element tail
    rng = np.random.default_rng(3)
    x = rng.standard_normal(1001).astype(np.float32)
    y = rng.standard_normal(1001).astype(np.float32)
    ref = float(np.dot(x.astype(np.float64), y.astype(np.float64)))
    assert abs(po.dot(x, y) - ref) <= 1e-4 * np.sqrt(1001) * np.abs(x * y).max()


@pytest.mark.parametrize("side,count", [(64, 13365), (128, 63189), (256, 273429)])
def test_reference_scene_particle_counts(side, count):
    """3x3 particles in cells 4<x<W-4, 4<y<H/2 (src/apic.rs:832-849)."""
    assert (side - 9) * (side // 2 - 5) * 9 == count
    o = po.Oracle(side, side, 1.0)
    assert o.status == 0
    d = o.dims()
    assert d["num_particles"] == count
    assert d["num_chunks"] == (count + 7) // 8


def test_reference_scene_positions_stratified():
    o = po.Oracle(64, 64, 1.0)
    x, y = o.read("POS_X"), o.read("POS_Y")
    n = o.dims()["num_particles"]
    # first cell (5,5): offsets (p + 0.5)/3 for p = 0..2, px fastest
    exp_x = [5 + (p + 0.5) / 3 for p in range(3)] * 3
    exp_y = [5 + (p + 0.5) / 3 for p in range(3) for _ in range(3)]
    assert np.allclose(x[:9], exp_x) and np.allclose(y[:9], exp_y)
    # tail lanes of the last chunk repeat the last particle (src/apic.rs:887-898)
    assert np.all(x[n:] == x[n - 1]) and np.all(y[n:] == y[n - 1])


@pytest.mark.parametrize("w,h,levels", [(128, 128, 5), (1024, 1024, 8), (4096, 2048, 9), (256, 128, 5),
                                        (100, 100, 2)])
def test_multigrid_level_counts(w, h, levels):
    """Levels while both sides are even and > 4 (src/apic.rs:900-908)."""
    cw, ch, n = w, h, 0
    while cw % 2 == 0 and ch % 2 == 0 and cw > 4 and ch > 4:
        cw //= 2
        ch //= 2
        n += 1
    assert n == levels
    if w * h <= 256 * 256:
        o = po.Oracle(w, h, 1.0)
        assert o.status == 0 and o.dims()["num_levels"] == levels


def test_odd_side_has_no_level_and_fails():
    """The reference panics in Apic::new when no level can be built (src/apic.rs:910)."""
    o = po.Oracle(97, 97, 1.0)
    assert o.status == 2


def test_open_face_masks():
    """update_world (src/apic.rs:918-944): a face is open when neither adjacent cell is a wall."""
    o = po.Oracle(16, 16, 1.0)
    ou, ov = o.read("OPEN_U").reshape(16, 17), o.read("OPEN_V").reshape(17, 16)
    assert ou[1:15, 2:15].all() and not ou[:, :2].any() and not ou[:, 15:].any()
    assert not ou[0].any() and not ou[15].any()
    assert ov[2:15, 1:15].all() and not ov[:2].any() and not ov[15:].any()
