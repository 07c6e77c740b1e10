"""GPU parity of the spatial grid (AABB, reorder, cell lists, valid cells,
samples: all bit-exact) and the density sample (FP32 vs the oracle's FP64).
The cell lists are built in chunks of whole x-slabs; a tiny chunk size makes
one build run many chunks (one slab per chunk, and slabs holding more pairs
than the chunk size), which must give the same lists."""
import numpy as np
import pytest

import oracle as O
from arapgs_amd import scenes as S

pytestmark = pytest.mark.gpu


def T(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def H(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("G,chunk", [(64, 0), (32, 0), (64, 20000), (32, 1)])
def test_grid_exact(ctx, G, chunk):
    sc = S.bar_scene(1, n=30000, n_nodes=64)
    g = sc.gaussians
    mn, mx, step = O.scene_aabb(g.pos, G)
    a7 = ctx.scene_aabb(T(g.pos), G)
    assert np.array_equal(a7[:3], mn) and np.array_equal(a7[3:6], mx) and a7[6] == np.float32(step)
    order = H(ctx.grid_reorder(T(g.pos), a7, G)).astype(np.int64)
    assert np.array_equal(order, O.grid_reorder(g.pos, mn, step, G))
    g = g.take(order)
    ab = H(ctx.gaussian_aabbs(T(g.pos), T(g.rot), T(g.scale), T(g.opacity)))
    ab_ref = O.gaussian_aabbs(g.pos, g.rot, g.scale, g.opacity)
    assert np.array_equal(ab, ab_ref)
    ctx.set_grid_chunk(chunk)
    try:
        grid = ctx.grid_build(T(ab), a7, G, padding=1)
    finally:
        ctx.set_grid_chunk(0)
    prefix, lst, valid, samples = grid.arrays()
    assert prefix.dtype == np.int64 and prefix[-1] == grid.n_pairs == lst.shape[0]
    rp, rl, rv, rs = O.grid_lists(ab_ref, mn, step, G, padding=1)
    assert np.array_equal(prefix, rp)
    assert np.array_equal(lst, rl)
    assert np.array_equal(valid, rv)
    assert np.array_equal(samples, rs)
    lpf = float(np.float32(step / 4) ** 2 * np.float32(0.2))
    feat, opac = grid.density(T(g.pos), T(g.rot), T(g.scale), T(g.opacity), T(g.sh), lpf)
    # oracle on a subset of cells (FP64)
    sel = slice(0, min(200, grid.n_valid))
    rf, ro = O.density_sample(rv[sel], rp, rl, rs, g.pos, g.rot, g.scale, g.opacity, g.sh, lpf)
    feat, opac = H(feat)[: rf.shape[0]], H(opac)[: ro.shape[0]]
    scale_o = max(1.0, np.abs(ro).max())
    assert np.abs(opac - ro).max() <= 1e-4 * scale_o
    assert np.abs(feat - rf).max() <= 1e-4 * max(1.0, np.abs(rf).max())
