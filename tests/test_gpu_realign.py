"""GPU parity of the density backward (arapgs_density_backward) and the
density-preserving realignment (arapgs_realign_step).

* the density VJP satisfies the adjoint identity against the oracle's FP64
  forward-mode derivative (oracle.density_jvp) along 32 random directions of
  the Gaussians' positions, rotations, scales, opacities, SH and the sample
  positions: within 1e-3 relative;
* one realignment step from zero Adam moments moves every raw parameter by
  -lr * sign(gradient) (Adam's first step), with the gradient chained through
  the activations (log scale, logit opacity);
* thirty realignment steps from perturbed Gaussians lower the L1 loss against
  the unperturbed Gaussians' densities.
"""
import numpy as np
import pytest

import oracle as O
from arapgs_amd import arapgs
from helpers import chain_scene

pytestmark = pytest.mark.gpu


def T(a):
    import torch
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).cuda()


def H(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def dsc(ctx):
    sc = chain_scene(seed=4, n=2000, n_nodes=16, K=4, G=12, log_scale=-3.2)
    g = sc["g"]
    a7 = ctx.scene_aabb(T(g.pos), 12)
    aabbs = ctx.gaussian_aabbs(T(g.pos), T(g.rot), T(g.scale), T(g.opacity))
    grid = ctx.grid_build(aabbs, a7, 12, padding=1)
    prefix, lst, valid, samples = grid.arrays()
    lpf = float(np.float32(a7[6] / 4) ** 2 * np.float32(0.2))
    return dict(sc=sc, g=g, grid=grid, prefix=prefix, list=lst, valid=valid, samples=samples, lpf=lpf)


def test_density_backward_adjoint_identity(ctx, dsc):
    g, grid = dsc["g"], dsc["grid"]
    rng = np.random.default_rng(7)
    S = grid.n_valid * 64
    gF = rng.normal(size=(S, 48)).astype(np.float32)
    gD = rng.normal(size=S).astype(np.float32)
    dpos, drot, dscale, dop, dsh, dsam = (H(x).astype(np.float64) for x in grid.density_backward(
        T(g.pos), T(g.rot), T(g.scale), T(g.opacity), T(g.sh), dsc["lpf"], dfeat=T(gF), dopac=T(gD),
        samples=T(dsc["samples"]), want_samples=True))
    errs = []
    for k in range(32):
        v = [rng.normal(size=g.pos.shape) * 0.01, rng.normal(size=g.rot.shape), rng.normal(size=g.scale.shape) * g.scale,
             rng.normal(size=g.opacity.shape), rng.normal(size=g.sh.shape), rng.normal(size=dsc["samples"].shape) * 0.01]
        which = k % 7  # one block at a time, then all
        v = [x if (which == 6 or i == which) else None for i, x in enumerate(v)]
        _, _, jf, jd = O.density_jvp(dsc["valid"], dsc["prefix"], dsc["list"], dsc["samples"], g.pos, g.rot, g.scale,
                                     g.opacity, g.sh, dsc["lpf"], *v)
        ref = float(np.sum(gF * jf) + np.sum(gD * jd))
        got = sum(float(np.sum(gr * x)) for gr, x in zip((dpos, drot, dscale, dop, dsh, dsam), v) if x is not None)
        errs.append(abs(got - ref) / abs(ref))
    print(f"density VJP adjoint identity: max rel err {max(errs):.3e}", flush=True)
    assert max(errs) < 1e-3, errs


def _params(g):
    pos, rot, sh = T(g.pos.copy()), T(g.rot.copy()), T(g.sh.copy())
    log_scale = T(np.log(g.scale).astype(np.float32))
    logit = T(np.log(g.opacity / (1.0 - g.opacity)).astype(np.float32))
    return pos, rot, log_scale, logit, sh


def test_first_adam_step_is_signed_learning_rate(ctx, dsc):
    import torch
    g, grid = dsc["g"], dsc["grid"]
    rng = np.random.default_rng(2)
    samples = T(dsc["samples"])
    # aims: the density of slightly different Gaussians
    feat_aim, dens_aim = grid.density(T(g.pos + 0.002 * rng.normal(size=g.pos.shape).astype(np.float32)), T(g.rot),
                                      T(g.scale), T(g.opacity), T(g.sh), dsc["lpf"])
    pos, rot, ls, lo, sh = _params(g)
    lr = dict(lr_pos=1e-4, lr_rot=1e-3, lr_scale=5e-3, lr_opacity=5e-2, lr_sh=2.5e-3)
    r = arapgs.Realign(ctx, grid, g.pos.shape[0], **lr, w_density=1.0)
    before = [H(x).copy() for x in (pos, rot, ls, lo, sh)]
    # the same gradients as the step computes
    scale, opac = torch.exp(ls), torch.sigmoid(lo)
    feat, dens = grid.density(pos, rot, scale, opac, sh, dsc["lpf"])
    dfeat, ddens = torch.sign(feat - feat_aim), torch.sign(dens - dens_aim)
    dpos, drot, dscale, dop, dsh = grid.density_backward(pos, rot, scale, opac, sh, dsc["lpf"], dfeat=dfeat,
                                                         dopac=ddens, samples=samples)
    graw = [H(dpos), H(drot), H(dscale * scale), H(dop * opac * (1 - opac)), H(dsh)]
    r.step(samples, feat_aim, dens_aim, dsc["lpf"], pos, rot, ls, lo, sh)
    torch.cuda.synchronize()
    after = [H(x) for x in (pos, rot, ls, lo, sh)]
    for b, a, gr, l in zip(before, after, graw, lr.values()):
        big = np.abs(gr) > 1e-3 * np.abs(gr).max()  # far from the gradient's rounding level
        assert big.any()
        np.testing.assert_allclose((a - b)[big], -l * np.sign(gr[big]), rtol=1e-3, atol=1e-7)
    r.close()


def test_realignment_lowers_the_l1_loss(ctx, dsc):
    import torch
    g, grid = dsc["g"], dsc["grid"]
    samples = T(dsc["samples"])
    feat_aim, dens_aim = grid.density(T(g.pos), T(g.rot), T(g.scale), T(g.opacity), T(g.sh), dsc["lpf"])
    rng = np.random.default_rng(5)
    pert = g.take(np.arange(g.pos.shape[0]))
    pert.pos = (g.pos + 0.01 * rng.normal(size=g.pos.shape)).astype(np.float32)
    pert.scale = (g.scale * np.exp(0.2 * rng.normal(size=g.scale.shape))).astype(np.float32)
    pos, rot, ls, lo, sh = _params(pert)
    r = arapgs.Realign(ctx, grid, g.pos.shape[0], lr_pos=2e-4, lr_rot=1e-3, lr_scale=5e-3, lr_opacity=1e-2,
                       lr_sh=1e-3)
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    losses = []
    for _ in range(30):
        r.step(samples, feat_aim, dens_aim, dsc["lpf"], pos, rot, ls, lo, sh, loss=loss)
        losses.append(float(loss.item()))
    print(f"realignment L1 loss: {losses[0]:.4g} -> {losses[-1]:.4g}", flush=True)
    assert losses[-1] < 0.8 * losses[0], losses
    r.close()


def test_rebin_equals_a_fresh_build(ctx, dsc):
    """arapgs_grid_rebin (UpdateContainingRelationship, GaussianView.cpp:3634-3744)
    after the Gaussians moved: prefix and lists equal a fresh grid build (and the
    oracle's lists) on the moved Gaussians' boxes and scene box, bit for bit;
    the valid cells and samples stay those of the first build."""
    g = dsc["g"]
    rng = np.random.default_rng(8)
    a7 = ctx.scene_aabb(T(g.pos), 12)
    grid = ctx.grid_build(ctx.gaussian_aabbs(T(g.pos), T(g.rot), T(g.scale), T(g.opacity)), a7, 12, padding=1)
    p0, _, v0, s0 = grid.arrays()
    moved = (g.pos + 0.05 * rng.normal(size=g.pos.shape)).astype(np.float32)
    a7n = grid.rebin(T(moved), T(g.rot), T(g.scale), T(g.opacity))
    prefix, lst, valid, samples = grid.arrays()
    fresh = ctx.grid_build(ctx.gaussian_aabbs(T(moved), T(g.rot), T(g.scale), T(g.opacity)), a7n, 12, padding=1)
    fp, fl, _, _ = fresh.arrays()
    assert np.array_equal(prefix, fp) and np.array_equal(lst, fl)
    mn, mx, step = O.scene_aabb(moved, 12)
    rp, rl, _, _ = O.grid_lists(O.gaussian_aabbs(moved, g.rot, g.scale, g.opacity), mn, step, 12, padding=1)
    assert np.array_equal(prefix, rp) and np.array_equal(lst, rl)
    assert not np.array_equal(prefix, p0)
    assert np.array_equal(valid, v0) and np.array_equal(samples, s0)
    fresh.close()
    grid.close()


def test_moved_mask_static_cells_and_scale_clamp(ctx, dsc):
    """Only Gaussians marked moved change; valid cells whose lists hold no moved
    Gaussian are static (skip flags = a CPU scan of the lists); the activated
    scales never exceed max_scale."""
    import torch
    g, grid = dsc["g"], dsc["grid"]
    rng = np.random.default_rng(6)
    samples = T(dsc["samples"])
    feat_aim, dens_aim = grid.density(T(g.pos), T(g.rot), T(g.scale), T(g.opacity), T(g.sh), dsc["lpf"])
    pert = g.take(np.arange(g.pos.shape[0]))
    pert.pos = (g.pos + 0.01 * rng.normal(size=g.pos.shape)).astype(np.float32)
    pos, rot, ls, lo, sh = _params(pert)
    moved = (rng.random(g.pos.shape[0]) < 0.3).astype(np.uint8)
    max_scale = (pert.scale * 1.001).astype(np.float32)
    r = arapgs.Realign(ctx, grid, g.pos.shape[0], lr_pos=2e-4, lr_rot=1e-3, lr_scale=5e-3, lr_opacity=1e-2,
                       lr_sh=1e-3)
    r.set_options(moved=T(moved), max_scale=T(max_scale))
    before = [H(x).copy() for x in (pos, rot, ls, lo, sh)]
    for _ in range(3):
        r.step(samples, feat_aim, dens_aim, dsc["lpf"], pos, rot, ls, lo, sh)
    torch.cuda.synchronize()
    after = [H(x) for x in (pos, rot, ls, lo, sh)]
    still = moved == 0
    for b, a in zip(before, after):
        assert np.array_equal(a[still], b[still])
    assert any(not np.array_equal(a[~still], b[~still]) for b, a in zip(before, after))
    assert (np.exp(after[2]) <= max_scale * (1 + 1e-6)).all()
    skip, _ = r.cells()
    prefix, lst, valid = dsc["prefix"], dsc["list"], dsc["valid"]
    want = np.zeros(len(valid), np.uint8)
    for v, c in enumerate(valid):
        beg = int(prefix[c - 1]) if c > 0 else 0
        want[v] = 0 if moved[lst[beg:int(prefix[c])]].any() else 1
    assert np.array_equal(skip, want) and want.any() and not want.all()
    r.close()


def test_per_cell_convergence(ctx, dsc):
    """conv_tol: the first step records each cell's loss; at the second step a
    cell whose loss is at most conv_tol times its record is converged (skip) and
    stays so. With conv_tol = 10 every cell with a nonzero loss converges; with
    conv_tol = 1e-9 none does. The per-cell losses add up to the reported loss."""
    import torch
    g, grid = dsc["g"], dsc["grid"]
    rng = np.random.default_rng(12)
    samples = T(dsc["samples"])
    feat_aim, dens_aim = grid.density(T(g.pos), T(g.rot), T(g.scale), T(g.opacity), T(g.sh), dsc["lpf"])
    pert = g.take(np.arange(g.pos.shape[0]))
    pert.pos = (g.pos + 0.01 * rng.normal(size=g.pos.shape)).astype(np.float32)
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    for tol, expect_all in ((10.0, True), (1e-9, False)):
        pos, rot, ls, lo, sh = _params(pert)
        r = arapgs.Realign(ctx, grid, g.pos.shape[0])
        r.set_options(conv_tol=tol)
        r.step(samples, feat_aim, dens_aim, dsc["lpf"], pos, rot, ls, lo, sh, loss=loss)
        skip1, cl = r.cells()
        assert not skip1.any()
        np.testing.assert_allclose(cl.astype(np.float64).sum(), float(loss.item()), rtol=1e-4)
        r.step(samples, feat_aim, dens_aim, dsc["lpf"], pos, rot, ls, lo, sh)
        skip2, _ = r.cells()
        if expect_all:
            assert skip2[cl > 0].all()
        else:
            assert not skip2[cl > 0].any()
        r.close()


def test_realign_loop_with_rebinning(ctx, dsc):
    """The F6 loop (GaussianView.cpp:1695-1733): re-binning before the loop and
    every 5 steps, 15 steps; the L1 loss decreases."""
    import torch
    g = dsc["g"]
    a7 = ctx.scene_aabb(T(g.pos), 12)
    grid = ctx.grid_build(ctx.gaussian_aabbs(T(g.pos), T(g.rot), T(g.scale), T(g.opacity)), a7, 12, padding=1)
    samples = grid.samples_tensor()
    feat_aim, dens_aim = grid.density(T(g.pos), T(g.rot), T(g.scale), T(g.opacity), T(g.sh), dsc["lpf"])
    rng = np.random.default_rng(13)
    pert = g.take(np.arange(g.pos.shape[0]))
    pert.pos = (g.pos + 0.01 * rng.normal(size=g.pos.shape)).astype(np.float32)
    pos, rot, ls, lo, sh = _params(pert)
    r = arapgs.Realign(ctx, grid, g.pos.shape[0], lr_pos=2e-4, lr_rot=1e-3, lr_scale=5e-3, lr_opacity=1e-2,
                       lr_sh=1e-3)
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    losses = []
    arapgs.realign_loop(ctx, grid, r, samples, feat_aim, dens_aim, dsc["lpf"], pos, rot, ls, lo, sh, 15,
                        rebin_interval=5, loss=loss, on_step=lambda i, l: losses.append(float(l.item())))
    print(f"F6 loop L1: {losses[0]:.4g} -> {losses[-1]:.4g}", flush=True)
    assert losses[-1] < 0.9 * losses[0]
    r.close()
    grid.close()


def test_adaptive_filter(ctx, dsc):
    """arapgs_grid_set_ada_lpf (GetAdaLpfRatio, GaussianView.cpp:4670-4749) from
    warped samples equals the oracle's restatement of the reference's 24 x 9
    least squares; with the filters set, the density equals the oracle's with
    the same per-cell filters (1e-4 of the largest value), its VJP satisfies the
    adjoint identity against the FP64 forward-mode oracle (1e-3, 16 directions),
    and switching the filters off gives the scalar-filter density bit for bit."""
    import torch
    g = dsc["g"]
    a7 = ctx.scene_aabb(T(g.pos), 12)
    grid = ctx.grid_build(ctx.gaussian_aabbs(T(g.pos), T(g.rot), T(g.scale), T(g.opacity)), a7, 12, padding=1)
    x = dsc["samples"].astype(np.float64)
    warped = (x + 0.15 * np.sin(3.0 * x[:, [1, 2, 0]]) + 0.05 * x[:, [2, 0, 1]]).astype(np.float32)
    grid.set_ada_lpf(0.2, samples=T(warped))
    mats = grid.ada_lpf()
    ref = O.ada_lpf(warped, 0.2)
    assert mats.shape == ref.shape
    assert np.abs(mats - ref).max() <= 1e-5 * np.abs(ref).max(), np.abs(mats - ref).max()
    feat, dens = grid.density(T(g.pos), T(g.rot), T(g.scale), T(g.opacity), T(g.sh), 123.0)  # lpf unused
    rf, rd = O.density_sample(dsc["valid"], dsc["prefix"], dsc["list"], dsc["samples"], g.pos, g.rot, g.scale,
                              g.opacity, g.sh, 0.0, ada=mats)
    assert np.abs(H(dens) - rd).max() <= 1e-4 * max(1.0, np.abs(rd).max())
    assert np.abs(H(feat) - rf).max() <= 1e-4 * max(1.0, np.abs(rf).max())
    rng = np.random.default_rng(17)
    S = grid.n_valid * 64
    gF = rng.normal(size=(S, 48)).astype(np.float32)
    gD = rng.normal(size=S).astype(np.float32)
    grads = [H(t).astype(np.float64) for t in grid.density_backward(
        T(g.pos), T(g.rot), T(g.scale), T(g.opacity), T(g.sh), 123.0, dfeat=T(gF), dopac=T(gD), samples=T(warped),
        want_samples=True)]
    errs = []
    for k in range(16):
        v = [rng.normal(size=g.pos.shape) * 0.01, rng.normal(size=g.rot.shape), rng.normal(size=g.scale.shape) * g.scale,
             rng.normal(size=g.opacity.shape), rng.normal(size=g.sh.shape), rng.normal(size=warped.shape) * 0.01]
        which = k % 7
        v = [x if (which == 6 or i == which) else None for i, x in enumerate(v)]
        _, _, jf, jd = O.density_jvp(dsc["valid"], dsc["prefix"], dsc["list"], warped, g.pos, g.rot, g.scale,
                                     g.opacity, g.sh, 0.0, *v, ada=mats)
        want = float(np.sum(gF * jf) + np.sum(gD * jd))
        got = sum(float(np.sum(gr * x)) for gr, x in zip(grads, v) if x is not None)
        errs.append(abs(got - want) / abs(want))
    print(f"adaptive-filter density VJP: max rel err {max(errs):.3e}", flush=True)
    assert max(errs) < 1e-3, errs
    grid.set_ada_lpf(0.0)
    assert grid.ada_lpf() is None
    f0, d0 = grid.density(T(g.pos), T(g.rot), T(g.scale), T(g.opacity), T(g.sh), dsc["lpf"])
    f1, d1 = dsc["grid"].density(T(g.pos), T(g.rot), T(g.scale), T(g.opacity), T(g.sh), dsc["lpf"])
    torch.cuda.synchronize()
    assert torch.equal(f0, f1) and torch.equal(d0, d1)
    grid.close()


def test_convergence_survives_rebinning_and_expand_op_range(ctx, dsc):
    """Converged flags stay set across a re-binning (the reference's
    grid_is_converged is cleared only with new features), and
    arapgs_realign_expand (ExpandOpRange, GaussianView.cpp:4320-4360) clears
    exactly the converged valid cells that are, or are face neighbours (clamped
    to the grid) of, a non-converged valid cell (host brute force)."""
    import torch
    g = dsc["g"]
    G = 12
    a7 = ctx.scene_aabb(T(g.pos), G)
    grid = ctx.grid_build(ctx.gaussian_aabbs(T(g.pos), T(g.rot), T(g.scale), T(g.opacity)), a7, G, padding=1)
    samples = grid.samples_tensor()
    feat_aim, dens_aim = grid.density(T(g.pos), T(g.rot), T(g.scale), T(g.opacity), T(g.sh), dsc["lpf"])
    rng = np.random.default_rng(21)
    pert = g.take(np.arange(g.pos.shape[0]))
    pert.pos = (g.pos + 0.01 * rng.normal(size=g.pos.shape)).astype(np.float32)
    # the per-cell loss ratio of the second step to the first, from a dry run
    pos, rot, ls, lo, sh = _params(pert)
    r = arapgs.Realign(ctx, grid, g.pos.shape[0])
    r.step(samples, feat_aim, dens_aim, dsc["lpf"], pos, rot, ls, lo, sh)
    _, cl1 = r.cells()
    r.step(samples, feat_aim, dens_aim, dsc["lpf"], pos, rot, ls, lo, sh)
    _, cl2 = r.cells()
    r.close()
    ratio = cl2[cl1 > 0] / cl1[cl1 > 0]
    tol = float(np.median(ratio))  # about half the cells converge at the second step
    pos, rot, ls, lo, sh = _params(pert)
    r = arapgs.Realign(ctx, grid, g.pos.shape[0])
    r.set_options(conv_tol=tol)
    r.step(samples, feat_aim, dens_aim, dsc["lpf"], pos, rot, ls, lo, sh)  # records
    r.step(samples, feat_aim, dens_aim, dsc["lpf"], pos, rot, ls, lo, sh)
    conv, _ = r.cells()
    assert 0 < conv.sum() < conv.size, (conv.sum(), conv.size)
    grid.rebin(pos, rot, torch.exp(ls), torch.sigmoid(lo))
    r.step(samples, feat_aim, dens_aim, dsc["lpf"], pos, rot, ls, lo, sh)
    after_rebin, _ = r.cells()
    assert np.all(after_rebin[conv == 1] == 1)
    conv = after_rebin.copy()
    r.expand()
    expanded, _ = r.cells()
    _, valid = grid.prefix_valid()
    mark = np.zeros(G ** 3, bool)
    for v, c in enumerate(valid):
        if conv[v]:
            continue
        x, y, z = c // (G * G), (c // G) % G, c % G
        for dx, dy, dz in ((0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
            xx, yy, zz = min(max(x + dx, 0), G - 1), min(max(y + dy, 0), G - 1), min(max(z + dz, 0), G - 1)
            mark[xx * G * G + yy * G + zz] = True
    want = np.where(mark[valid], 0, conv).astype(np.uint8)
    assert np.array_equal(expanded, want)
    assert (expanded < conv).any()  # some converged cell was reopened
    r.close()
    grid.close()
