"""The Morton-code restatement (geometry.py:7-47) against a bit-by-bit
interleave, closed forms at the box corners, and Z-order on a grid."""
import numpy as np

from oracle import oracle as orc


def _interleave(qx, qy, qz):
    code = 0
    for b in range(21):
        code |= ((qx >> b) & 1) << (3 * b + 2) | ((qy >> b) & 1) << (3 * b + 1) | ((qz >> b) & 1) << (3 * b)
    return code


def test_matches_bitwise_interleave():
    rng = np.random.default_rng(0)
    p = rng.uniform(-3, 5, (500, 3)).astype(np.float32)
    codes = orc.morton_codes(p)
    lo, hi = p.min(0), p.max(0)
    q = np.clip((p - lo) / (hi - lo) * np.float32(2097152.0), 0, np.float32(2097151.0)).astype(np.int64)
    for i in range(len(p)):
        assert int(codes[i]) == _interleave(*(int(v) for v in q[i]))


def test_corners_and_flat_axis():
    p = np.array([[0, 0, 0], [1, 1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    c = orc.morton_codes(p)
    top = (1 << 63) - 1  # all 63 bits set at the top corner (clamped to 2^21 - 1 per axis)
    assert list(c) == [0, top, top & 0x4924924924924924, top & 0x2492492492492492, top & 0x1249249249249249]
    # a flat axis (extent 0) maps to 0 on that axis
    flat = orc.morton_codes(np.array([[0, 2, 5], [1, 2, 5]], np.float32))
    assert flat[0] == 0 and flat[1] == 0x4924924924924924 & ((1 << 63) - 1)


def test_z_order_on_a_grid():
    g = np.stack(np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij"), -1).reshape(-1, 3)
    order = np.argsort(orc.morton_codes(g.astype(np.float32)), kind="stable")
    # the first eight cells are the 2x2x2 block at the origin, z fastest
    assert [tuple(g[k]) for k in order[:8]] == [(x, y, z) for x in (0, 1) for y in (0, 1) for z in (0, 1)]
