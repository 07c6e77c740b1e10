"""GPU parity of the adjoint chain (arapgs_chain_forward / _backward) against
the oracle's FP64 chain (oracle/chain.cpp, pinned by tests/test_cpu_oracle_chain.py).

* forward: the float images, density features and densities agree with the
  FP64 chain (images: mean |d| < 1e-4, 99.9 % of pixels within 1e-3 — the
  rest are pixels where a contributor sits on the 1/255 alpha cut, which
  float and FP64 place differently; features and densities within 1e-3
  relative to their scale);
* backward: for random output cotangents g (images, features, densities) the
  HIP vector-Jacobian product J^T g agrees with the FP64 directional
  derivatives along 64 random directions v of all inputs (node R and t,
  rest rotations and scales): <J^T g, v> = <g, J v> within 1e-3 relative
  for every direction, with cotangents on the images only, on the density
  outputs only and on both, and per input block along its own directions;
* the backward is reproducible up to float atomics (1e-5).
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
def setup(ctx):
    import torch
    sc = chain_scene(seed=1, n=3000, n_nodes=32, K=8, W=96, H=72, n_views=2, G=16, log_scale=-3.5)
    g = sc["g"]
    # the grid on the GPU (its lists are bit-exact with the oracle's, tests/test_gpu_grid.py)
    a7 = ctx.scene_aabb(T(g.pos), 16)
    aabbs = ctx.gaussian_aabbs(T(g.pos), T(g.rot), T(g.scale), T(g.opacity))
    grid = ctx.grid_build(aabbs, a7, 16, padding=1)
    prefix, lst, valid, samples = grid.arrays()
    sidx, sw = O.knn_weights(sc["nodes"], samples, sc["K"])
    sidx = np.ascontiguousarray(sidx[:, :sc["K"]])
    cams = [dict(view=T(c.view), proj=T(c.proj), campos=T(c.campos), tx=c.tanfovx, ty=c.tanfovy) for c in sc["cams"]]
    sw = sw.astype(np.float32)
    dev = dict(pos0=T(g.pos), sh0=T(g.sh), opac=T(g.opacity), rot0=T(g.rot), scale0=T(g.scale),
               ends_idx=T(sc["ends_idx"]), ends_w=T(sc["ends_w"].astype(np.float32)), sidx=T(sidx), sw=T(sw),
               bg=T(sc["bg"]))
    M = sc["nodes"].shape[0]
    table = np.zeros((M, 16), np.float32)
    table[:, 0:3] = sc["nodes"]
    table[:, 3:12] = sc["R"]
    table[:, 12:15] = sc["t"]
    dev["table"] = T(table)
    chain = arapgs.Chain(ctx, M, dev["pos0"], dev["sh0"], dev["opac"], dev["ends_idx"], dev["ends_w"], cams,
                         sc["W"], sc["H"], dev["bg"], grid=grid, sample_idx=dev["sidx"], sample_w=dev["sw"],
                         lpf=sc["lpf"])
    ocams = [(c.view, c.proj, c.campos, c.tanfovx, c.tanfovy) for c in sc["cams"]]
    ora = O.Chain(g.pos, g.sh, g.opacity, sc["ends_idx"], sc["ends_w"], sc["nodes"], ocams, sc["W"], sc["H"],
                  sc["bg"], grid=(valid, prefix, lst, samples), sample_idx=sidx, sample_w=sw, lpf=sc["lpf"])
    # table R/t as the GPU sees them (float)
    params = (table[:, 3:12].astype(np.float64), table[:, 12:15].astype(np.float64), g.rot.astype(np.float64),
              g.scale.astype(np.float64))
    torch.cuda.synchronize()
    yield dict(sc=sc, dev=dev, chain=chain, ora=ora, params=params, grid=grid, cams=cams)
    chain.close()


def test_chain_forward_matches_fp64_oracle(setup):
    s = setup
    d = s["dev"]
    print(f"chain forward N={s['sc']['g'].pos.shape[0]} S={s['chain'].S}", flush=True)
    images, feat, dens = s["chain"].forward(d["table"], d["rot0"], d["scale0"])
    ref = s["ora"].jvp(*s["params"])
    for v, img in enumerate(images):
        err = np.abs(H(img) - ref["images"][v])
        assert err.mean() < 1e-4, err.mean()
        assert np.quantile(err, 0.999) < 1e-3, np.quantile(err, 0.999)
    fs = max(1.0, np.abs(ref["feat"]).max())
    assert np.abs(H(feat) - ref["feat"]).max() <= 1e-3 * fs
    assert np.abs(H(dens) - ref["dens"]).max() <= 1e-3 * max(1.0, ref["dens"].max())


def _vjp(s, rng, use_images=True, use_density=True):
    import torch
    ch = s["chain"]
    d = s["dev"]
    V, Hh, W, S = ch.V, ch.H, ch.W, ch.S
    gI = rng.normal(size=(V, 3, Hh, W)).astype(np.float32) * (1.0 if use_images else 0.0)
    gF = rng.normal(size=(S, 48)).astype(np.float32) * (1.0 if use_density else 0.0)
    gD = rng.normal(size=S).astype(np.float32) * (1.0 if use_density else 0.0)
    dt, dq, ds = ch.backward(d["table"], d["rot0"], d["scale0"], dimages=[T(x) for x in gI], dfeat=T(gF),
                             ddens=T(gD))
    torch.cuda.synchronize()
    return (gI, gF, gD), (H(dt).astype(np.float64), H(dq).astype(np.float64), H(ds).astype(np.float64))


@pytest.mark.parametrize("outputs", ["images", "density", "both"])
def test_chain_backward_adjoint_identity_64_directions(setup, outputs):
    s = setup
    rng = np.random.default_rng(11)
    (gI, gF, gD), (dtab, dq0, ds0) = _vjp(s, rng, outputs != "density", outputs != "images")
    R, t, q0, s0 = s["params"]
    rows = []
    for k in range(64):
        block = k % 5  # 0: R, 1: t, 2: q0, 3: s0, 4: all
        v = [rng.normal(size=R.shape), rng.normal(size=t.shape), rng.normal(size=q0.shape),
             rng.normal(size=s0.shape) * s0]
        v = [x if (block == 4 or i == block) else None for i, x in enumerate(v)]
        out = s["ora"].jvp(R, t, q0, s0, *v)
        ref = float(np.sum(gI * out["d_images"]) + np.sum(gF * out["d_feat"]) + np.sum(gD * out["d_dens"]))
        got = 0.0
        if v[0] is not None: got += float(np.sum(dtab[:, 3:12] * v[0]))
        if v[1] is not None: got += float(np.sum(dtab[:, 12:15] * v[1]))
        if v[2] is not None: got += float(np.sum(dq0 * v[2]))
        if v[3] is not None: got += float(np.sum(ds0 * v[3]))
        rows.append((block, got, ref))
    errs = np.array([abs(g - r) / abs(r) for _, g, r in rows])
    print(f"adjoint identity ({outputs}): max rel err {errs.max():.3e}, median {np.median(errs):.3e}", flush=True)
    assert all(abs(r) > 0.0 for _, _, r in rows)
    bad = [(b, g, r) for (b, g, r), e in zip(rows, errs) if e > 1e-3]
    assert not bad, bad


def test_chain_backward_reproducible(setup):
    s = setup
    a = _vjp(s, np.random.default_rng(3))[1]
    b = _vjp(s, np.random.default_rng(3))[1]
    for x, y in zip(a, b):
        assert np.abs(x - y).max() <= 1e-5 * max(1.0, np.abs(x).max())


def _chain_function(chain):
    """The chain as a torch.autograd.Function: forward = arapgs_chain_forward,
    backward = arapgs_chain_backward (the vector-Jacobian product), so a torch
    loss over the images and density outputs back-propagates to the node table
    and the rest rotations and scales."""
    import torch

    class ChainFn(torch.autograd.Function):
        @staticmethod
        def forward(fctx, table, rot0, scale0):
            images, feat, dens = chain.forward(table.detach(), rot0.detach(), scale0.detach())
            fctx.save_for_backward(table, rot0, scale0)
            return torch.stack(images), feat, dens

        @staticmethod
        def backward(fctx, g_images, g_feat, g_dens):
            table, rot0, scale0 = fctx.saved_tensors
            dims = [g_images[v].contiguous() for v in range(g_images.shape[0])]
            dt, dq, ds = chain.backward(table.detach(), rot0.detach(), scale0.detach(), dimages=dims,
                                        dfeat=g_feat.contiguous(), ddens=g_dens.contiguous())
            return dt, dq, ds

    return ChainFn.apply


def test_chain_autograd_function(setup):
    """Through torch.autograd: the gradients of a torch loss equal the C-ABI
    backward with that loss's cotangents (up to float atomics), agree with the
    FP64 oracle's directional derivatives, and a few Adam steps on the node
    translations lower an image-fitting loss."""
    import torch
    s = setup
    ch, d = s["chain"], s["dev"]
    fn = _chain_function(ch)
    rng = np.random.default_rng(5)
    gI = T(rng.normal(size=(ch.V, 3, ch.H, ch.W)).astype(np.float32))
    gF = T(rng.normal(size=(ch.S, 48)).astype(np.float32))
    gD = T(rng.normal(size=ch.S).astype(np.float32))
    table = d["table"].clone().requires_grad_(True)
    rot0 = d["rot0"].clone().requires_grad_(True)
    scale0 = d["scale0"].clone().requires_grad_(True)
    img, feat, dens = fn(table, rot0, scale0)
    loss = (img * gI).sum() + (feat * gF).sum() + (dens * gD).sum()
    loss.backward()
    dt, dq, ds = ch.backward(d["table"], d["rot0"], d["scale0"], dimages=[gI[v] for v in range(ch.V)], dfeat=gF,
                             ddens=gD)
    torch.cuda.synchronize()
    for got, ref in ((table.grad, dt), (rot0.grad, dq), (scale0.grad, ds)):
        got, ref = H(got), H(ref)
        assert np.abs(got - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())
    # one direction of every block against the FP64 forward-mode oracle
    R, t, q0, s0 = s["params"]
    v = [rng.normal(size=R.shape), rng.normal(size=t.shape), rng.normal(size=q0.shape), rng.normal(size=s0.shape) * s0]
    out = s["ora"].jvp(R, t, q0, s0, *v)
    ref = float(np.sum(H(gI) * out["d_images"]) + np.sum(H(gF) * out["d_feat"]) + np.sum(H(gD) * out["d_dens"]))
    tg = H(table.grad).astype(np.float64)
    got = float(np.sum(tg[:, 3:12] * v[0]) + np.sum(tg[:, 12:15] * v[1]) +
                np.sum(H(rot0.grad).astype(np.float64) * v[2]) + np.sum(H(scale0.grad).astype(np.float64) * v[3]))
    assert abs(got - ref) <= 1e-3 * abs(ref), (got, ref)
    # fit the images of a translated table with Adam on the translations
    target_tab = d["table"].clone()
    target_tab[:, 12:15] += T(rng.normal(size=(ch.M, 3)).astype(np.float32) * 0.01)
    with torch.no_grad():
        target = fn(target_tab, d["rot0"], d["scale0"])[0].clone()
    trans = d["table"][:, 12:15].clone().requires_grad_(True)
    opt = torch.optim.Adam([trans], lr=2e-3)
    losses = []
    for _ in range(8):
        opt.zero_grad()
        tab = torch.cat([d["table"][:, :12], trans, d["table"][:, 15:]], dim=1)
        img = fn(tab, d["rot0"], d["scale0"])[0]
        lo = ((img - target) ** 2).sum()
        lo.backward()
        opt.step()
        losses.append(float(lo))
    print(f"autograd fit: {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert losses[-1] < 0.8 * losses[0], losses


def test_chain_with_adaptive_filter(ctx):
    """The chain over a grid with per-cell adaptive filters
    (arapgs_grid_set_ada_lpf): forward density = the FP64 chain with the same
    filters, and the backward satisfies the adjoint identity along 16
    directions (density cotangents: the only outputs the filter changes)."""
    import torch
    sc = chain_scene(seed=2, n=2500, n_nodes=24, K=8, W=64, H=48, n_views=1, G=12, log_scale=-3.4)
    g = sc["g"]
    a7 = ctx.scene_aabb(T(g.pos), 12)
    grid = ctx.grid_build(ctx.gaussian_aabbs(T(g.pos), T(g.rot), T(g.scale), T(g.opacity)), a7, 12, padding=1)
    prefix, lst, valid, samples = grid.arrays()
    x = samples.astype(np.float64)
    warped = (x + 0.1 * np.sin(4.0 * x[:, [2, 0, 1]])).astype(np.float32)
    grid.set_ada_lpf(0.3, samples=T(warped))
    mats = grid.ada_lpf()
    sidx, sw = O.knn_weights(sc["nodes"], samples, sc["K"])
    sidx = np.ascontiguousarray(sidx[:, :sc["K"]])
    sw = sw.astype(np.float32)
    cams = [dict(view=T(c.view), proj=T(c.proj), campos=T(c.campos), tx=c.tanfovx, ty=c.tanfovy) for c in sc["cams"]]
    M = sc["nodes"].shape[0]
    table = np.zeros((M, 16), np.float32)
    table[:, 0:3] = sc["nodes"]
    table[:, 3:12] = sc["R"]
    table[:, 12:15] = sc["t"]
    dev = dict(pos0=T(g.pos), sh0=T(g.sh), opac=T(g.opacity), ends_idx=T(sc["ends_idx"]),
               ends_w=T(sc["ends_w"].astype(np.float32)), sidx=T(sidx), sw=T(sw), bg=T(sc["bg"]))
    chain = arapgs.Chain(ctx, M, dev["pos0"], dev["sh0"], dev["opac"], dev["ends_idx"], dev["ends_w"], cams, sc["W"],
                         sc["H"], dev["bg"], grid=grid, sample_idx=dev["sidx"], sample_w=dev["sw"], lpf=sc["lpf"])
    ocams = [(c.view, c.proj, c.campos, c.tanfovx, c.tanfovy) for c in sc["cams"]]
    ora = O.Chain(g.pos, g.sh, g.opacity, sc["ends_idx"], sc["ends_w"], sc["nodes"], ocams, sc["W"], sc["H"], sc["bg"],
                  grid=(valid, prefix, lst, samples), sample_idx=sidx, sample_w=sw, lpf=sc["lpf"], ada=mats)
    params = (table[:, 3:12].astype(np.float64), table[:, 12:15].astype(np.float64), g.rot.astype(np.float64),
              g.scale.astype(np.float64))
    _, feat, dens = chain.forward(T(table), T(g.rot), T(g.scale))
    ref = ora.jvp(*params)
    assert np.abs(H(feat) - ref["feat"]).max() <= 1e-3 * max(1.0, np.abs(ref["feat"]).max())
    assert np.abs(H(dens) - ref["dens"]).max() <= 1e-3 * max(1.0, ref["dens"].max())
    rng = np.random.default_rng(23)
    S = chain.S
    gF = rng.normal(size=(S, 48)).astype(np.float32)
    gD = rng.normal(size=S).astype(np.float32)
    dt, dq, ds = chain.backward(T(table), T(g.rot), T(g.scale), dimages=None, dfeat=T(gF), ddens=T(gD))
    torch.cuda.synchronize()
    dt, dq, ds = (H(a).astype(np.float64) for a in (dt, dq, ds))
    R, t, q0, s0 = params
    errs = []
    for k in range(16):
        v = [rng.normal(size=R.shape), rng.normal(size=t.shape), rng.normal(size=q0.shape),
             rng.normal(size=s0.shape) * s0]
        block = k % 5
        v = [x if (block == 4 or i == block) else None for i, x in enumerate(v)]
        out = ora.jvp(R, t, q0, s0, *v)
        want = float(np.sum(gF * out["d_feat"]) + np.sum(gD * out["d_dens"]))
        got = 0.0
        if v[0] is not None: got += float(np.sum(dt[:, 3:12] * v[0]))
        if v[1] is not None: got += float(np.sum(dt[:, 12:15] * v[1]))
        if v[2] is not None: got += float(np.sum(dq * v[2]))
        if v[3] is not None: got += float(np.sum(ds * v[3]))
        errs.append(abs(got - want) / abs(want))
    print(f"chain with adaptive filters: adjoint identity max rel err {max(errs):.3e}", flush=True)
    assert max(errs) < 1e-3, errs
    chain.close()
    grid.close()
