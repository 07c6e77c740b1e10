"""The blend's conservative Gaussian-vs-tile test, evaluated on the host.

The blend drops a (Gaussian, tile) pair from a tile's batch only when the test
says no pixel of the tile can reach alpha >= 1/255. These checks brute-force
every pixel centre of the tile with the blend's own float32 arithmetic
(power = -0.5 (a dx^2 + c dy^2) - b dx dy, alpha = min(0.99, o exp(power)),
skip when power > 0 or alpha < 1/255) and require that every dropped pair has
no contributing pixel, including pairs placed right at the threshold.
"""
import numpy as np
import pytest

from arapgs_amd import arapgs

TILE = 16


def random_gaussians(rng, n, spread):
    ang = rng.uniform(0, np.pi, n)
    l1 = np.exp(rng.uniform(np.log(0.3), np.log(3000.0), n))
    l2 = np.exp(rng.uniform(np.log(0.3), np.log(50.0), n))
    c, s = np.cos(ang), np.sin(ang)
    A = c * c * l1 + s * s * l2 + 0.3
    B = c * s * (l1 - l2)
    Cc = s * s * l1 + c * c * l2 + 0.3
    det = A * Cc - B * B
    conic = np.stack([Cc / det, -B / det, A / det, rng.uniform(0.0, 1.0, n)], 1).astype(np.float32)
    conic[: n // 10, 3] = rng.uniform(0.0, 0.01, n // 10)  # near / below the 1/255 floor
    xy = rng.uniform(-spread, TILE + spread, size=(n, 2)).astype(np.float32)
    return xy, conic


def contributes(xy, conic, box):
    """Per Gaussian: does any pixel centre of its box pass the blend's tests."""
    out = np.zeros(xy.shape[0], dtype=bool)
    one255 = np.float32(1.0) / np.float32(255.0)
    for i in range(xy.shape[0]):
        x0, y0, x1, y1 = box[i]
        px, py = np.meshgrid(np.arange(x0, x1 + 1, dtype=np.float32), np.arange(y0, y1 + 1, dtype=np.float32))
        dx = xy[i, 0] - px
        dy = xy[i, 1] - py
        a, b, c, o = conic[i]
        power = np.float32(-0.5) * (a * dx * dx + c * dy * dy) - b * dx * dy
        alpha = np.minimum(np.float32(0.99), o * np.exp(power))
        out[i] = bool(np.any((power <= 0) & (alpha >= one255)))
    return out


def test_dropped_pairs_never_contribute():
    rng = np.random.default_rng(3)
    n = 20000
    xy, conic = random_gaussians(rng, n, spread=120.0)
    box = np.tile(np.array([0, 0, TILE - 1, TILE - 1], dtype=np.float32), (n, 1))
    keep = arapgs.tile_test_host(xy, conic, box)
    truth = contributes(xy, conic, box)
    assert not np.any(truth & (keep == 0)), "a dropped pair has a contributing pixel"
    # the test must actually drop something: most far-away pairs do not reach 1/255
    dropped = np.count_nonzero(keep == 0)
    assert dropped > n // 5
    # and it is not far from exact: few kept pairs are invisible in the tile
    kept_invisible = np.count_nonzero((keep == 1) & ~truth)
    assert kept_invisible < 0.05 * n


def test_threshold_boundary():
    """Gaussians whose nearest pixel sits within +-1e-3 of the alpha floor."""
    rng = np.random.default_rng(11)
    n = 4000
    xy, conic = random_gaussians(rng, n, spread=0.0)
    conic[:, 3] = rng.uniform(0.05, 1.0, n).astype(np.float32)
    # put the mean left of the tile on row 7: q(dx, 0) = a dx^2 at pixel column 0
    a = conic[:, 0].astype(np.float64)
    conic[:, 1] = 0.0  # axis-aligned, so the nearest point is (0, 7)
    q_thr = 2.0 * np.log(255.0 * conic[:, 3].astype(np.float64))
    q_target = q_thr * (1.0 + rng.uniform(-1e-3, 1e-3, n))
    dx = np.sqrt(np.maximum(q_target, 0.0) / a)
    xy[:, 0] = (-dx).astype(np.float32)
    xy[:, 1] = 7.0
    box = np.tile(np.array([0, 0, TILE - 1, TILE - 1], dtype=np.float32), (n, 1))
    keep = arapgs.tile_test_host(xy, conic, box)
    truth = contributes(xy, conic, box)
    assert not np.any(truth & (keep == 0))


def test_mean_inside_and_degenerate_are_kept():
    xy = np.array([[5.0, 5.0], [-40.0, 3.0], [3.0, 3.0]], dtype=np.float32)
    conic = np.array([[1e-6, 0.0, 1e-6, 0.001],  # mean inside: always kept
                      [0.0, 0.0, 1.0, 1.0],       # not positive definite: kept
                      [1.0, 2.0, 1.0, 1.0]],      # indefinite: kept
                     dtype=np.float32)
    box = np.tile(np.array([0, 0, 15, 15], dtype=np.float32), (3, 1))
    assert list(arapgs.tile_test_host(xy, conic, box)) == [1, 1, 1]


def test_zero_opacity_and_far_gaussians_dropped():
    xy = np.array([[8.0, 40.0], [-500.0, -500.0]], dtype=np.float32)
    conic = np.array([[1.0, 0.0, 1.0, 0.0], [1.0, 0.0, 1.0, 1.0]], dtype=np.float32)
    box = np.tile(np.array([0, 0, 15, 15], dtype=np.float32), (2, 1))
    assert list(arapgs.tile_test_host(xy, conic, box)) == [0, 0]


def test_bad_shapes_raise():
    with pytest.raises(arapgs.ArapgsError):
        arapgs.tile_test_host(np.zeros((2, 2)), np.zeros((3, 4)), np.zeros((2, 4)))
