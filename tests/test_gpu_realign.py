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
