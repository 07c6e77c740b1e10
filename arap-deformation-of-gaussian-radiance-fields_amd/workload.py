"""Device-resident drag-session state for a synthetic scene, built entirely with
the arapgs kernels (the set-up half of GaussianView: getGridSamples reorder,
farthest-point nodes, setupEdges, computeWeights for end points / samples /
node sources, GetEndPoints, the grid, the samples' initial density features,
and the edit blocks). Shared by bench.py and the tests.
"""
from __future__ import annotations

import numpy as np

from . import arapgs
from . import scenes


def prepare(ctx, sc, device="cuda:0", log=None):
    """Returns dict(state, solver_args, active, cams, bg, opac, grid, K, lpf)."""
    import torch
    g = sc.gaussians
    K = sc.knn_k
    dev = torch.device(device)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    pos, rot, scale, opac, sh = T(g.pos), T(g.rot), T(g.scale), T(g.opacity), T(g.sh)
    a7 = ctx.scene_aabb(pos, sc.grid)
    order = ctx.grid_reorder(pos, a7, sc.grid).long()
    pos, rot, scale, opac, sh = (t.index_select(0, order).contiguous() for t in (pos, rot, scale, opac, sh))
    nidx = ctx.fps(pos, sc.n_nodes).long()
    node_pos = pos.index_select(0, nidx).contiguous()
    nbr = ctx.graph_edges(node_pos, K)
    src_idx_full, src_w = ctx.knn(node_pos, node_pos, K)
    node_src_idx = src_idx_full[:, :K].contiguous()
    ends = ctx.end_points(pos, rot, scale)
    eidx_full, ends_w = ctx.knn(node_pos, ends.view(-1, 3), K)
    ends_idx = eidx_full[:, :K].contiguous()
    del eidx_full
    aabbs = ctx.gaussian_aabbs(pos, rot, scale, opac)
    grid = ctx.grid_build(aabbs, a7, sc.grid, padding=1)
    samples = grid.samples_tensor()
    sidx_full, sample_w = ctx.knn(node_pos, samples, K)
    sample_idx = sidx_full[:, :K].contiguous()
    del sidx_full
    lpf = float(np.float32(a7[6] / 4) ** 2 * np.float32(0.2))
    feat, _ = grid.density(pos, rot, scale, opac, sh, lpf)
    torch.cuda.synchronize(dev)
    if log:
        log(f"grid: {grid.n_valid} valid cells, {samples.shape[0]} samples, {grid.n_pairs} cell-Gaussian pairs")
    npos_h = node_pos.cpu().numpy()
    blocks, types = scenes.blocks_for(sc, npos_h)
    active = [b for b, t in zip(blocks, types) if t == 1]
    active = T(np.concatenate(active).astype(np.int32)) if active else torch.zeros(0, dtype=torch.int32, device=dev)
    solver_args = (nbr.cpu().numpy(), blocks, types, node_src_idx.cpu().numpy(), src_w.cpu().numpy(), npos_h)
    state = dict(pos=pos, rot=rot, scale=scale, sh=sh, scale0=scale.clone(), ends=ends, ends_idx=ends_idx,
                 ends_w=ends_w, gs_static=None, node_pos=node_pos, node_src_idx=node_src_idx, node_src_w=src_w,
                 samples=samples, sample_idx=sample_idx, sample_w=sample_w, sample_static=None, sample_feat=feat)
    cams = [dict(view=T(c.view), proj=T(c.proj), campos=T(c.campos), tx=c.tanfovx, ty=c.tanfovy, W=c.W, H=c.H)
            for c in sc.cameras]
    bg = T(np.array([0.8, 0.8, 0.8], np.float32))
    return dict(state=state, solver_args=solver_args, active=active, cams=cams, bg=bg, opac=opac, grid=grid, K=K,
                lpf=lpf)


def clone_state(state: dict) -> dict:
    return {k: (v.clone() if v is not None else None) for k, v in state.items()}


def drag_delta(sc, drag_frames: int):
    """Per-frame translation of the active block (UpdateAimPosition)."""
    return np.float32(sc.edit["distance"] / drag_frames) * sc.edit["direction"].astype(np.float32)


def make_solver(ctx, solver_args, **kw):
    return arapgs.Solver(ctx, *solver_args, **kw)


def batch_scenes(seed: int, n_scenes: int, n_min: int = 50_000, n_max: int = 2_000_000, W: int = 800,
                 H: int = 800, node_div: int = 500):
    """Scene list of the uneven-batch workload: Gaussian counts log-uniform in
    [n_min, n_max], nodes = ceil(n / node_div) rounded up to a multiple of 64,
    one capsule scene (seed + 1 + i) with one camera each. Yields scenes one at
    a time so that only one scene's host arrays live at once."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_scenes)).astype(np.int64)
    nodes = (np.ceil(counts / node_div).astype(np.int64) + 63) // 64 * 64
    for i, (c, m) in enumerate(zip(counts, nodes)):
        yield scenes.capsule_scene(seed + 1 + i, n=int(c), n_nodes=int(m), W=W, H=H, n_views=1)


def prepare_batch(ctx, scene_iter, device="cuda:0", log=None):
    """Independent scenes stored back to back: every per-Gaussian, per-node and
    per-sample array of prepare() concatenated, node indices offset into the
    concatenated node list (so one node table serves every scene). Returns
    dict(state, seg, node_seg, sample_seg, solver_args (per scene), active (per
    scene), cams (one per scene), bg, opac, K)."""
    import torch
    parts = []
    for sc in scene_iter:
        wl = prepare(ctx, sc, device)
        st = wl["state"]
        keep = {k: st[k] for k in ("pos", "rot", "scale", "sh", "scale0", "ends", "ends_idx", "ends_w", "node_pos",
                                   "node_src_idx", "node_src_w", "samples", "sample_idx", "sample_w", "sample_feat")}
        parts.append(dict(state=keep, opac=wl["opac"], solver_args=wl["solver_args"], active=wl["active"],
                          cam=wl["cams"][0], K=wl["K"]))
        if log:
            _, blocks, types, _, _, _ = wl["solver_args"]
            sizes = {int(t): len(b) for b, t in zip(blocks, types)}
            log(f"batch scene {len(parts) - 1}: {sc.gaussians.n} Gaussians, {sc.n_nodes} nodes, "
                f"{st['samples'].shape[0]} samples, pinned/active nodes {sizes.get(0, 0)}/{sizes.get(1, 0)}")
        del wl
    K = parts[0]["K"]
    seg, node_seg, sample_seg = [0], [0], [0]
    for p in parts:
        seg.append(seg[-1] + p["state"]["pos"].shape[0])
        node_seg.append(node_seg[-1] + p["state"]["node_pos"].shape[0])
        sample_seg.append(sample_seg[-1] + p["state"]["samples"].shape[0])
    state = {}
    for key in parts[0]["state"]:
        chunks = []
        for i, p in enumerate(parts):
            v = p["state"][key]
            if key in ("ends_idx", "node_src_idx", "sample_idx"):
                v = v + node_seg[i]
            chunks.append(v)
        state[key] = torch.cat(chunks).contiguous()
        for p in parts:
            p["state"][key] = None
    state.update(gs_static=None, sample_static=None)
    opac = torch.cat([p["opac"] for p in parts]).contiguous()
    bg = torch.tensor([0.8, 0.8, 0.8], dtype=torch.float32, device=device)
    torch.cuda.synchronize(torch.device(device))
    return dict(state=state, seg=seg, node_seg=node_seg, sample_seg=sample_seg,
                solver_args=[p["solver_args"] for p in parts], active=[p["active"] for p in parts],
                cams=[p["cam"] for p in parts], bg=bg, opac=opac, K=K)
