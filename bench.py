#!/usr/bin/env python3
"""Deform+render frames/s at 1M Gaussians on one MI355X.

Workload (the interactive-scale capsule scene): 1M FP32 Gaussians inside a
union of 9 seeded capsules in the unit sphere, 2048 farthest-point nodes, K=8,
64^3 density grid with 4^3 samples per valid cell, one capsule pinned and another
dragged 0.3 along a seeded unit vector (spread over a 30-frame drag), 10
Gauss-Newton iterations (reference stop rule on), 4 views at 800x800, SH degree 3.

One step = one drag frame: aims -> ARAP GN solve -> node table -> sample blend
-> end-point blend + six-point update (polar, quaternion, scale, SH) -> node
blend -> node quaternions + sample SH rotation -> 4 renders. All inputs are
resident in HBM before the timed region.

Three forms are timed from the same starting state (the drag is reset before each):
  plain      every stage synchronous, three kernels per CG iteration with the
             convergence flag read back to the host after every CG iteration
             (the reference's call pattern: a host-driven solver and a device
             sync after each launch);
  fast       stages queued back to back; the CG phases carry a device-side
             convergence flag (kernels queued past convergence exit at once)
             and the host looks at it only every 16 CG iterations; the views
             are rendered in one pass (arapgs_render_views);
  pipelined  as fast, with frame t's deformation and renders on a second stream
             overlapping frame t+1's solve (pipeline.DragPipeline: two node-table
             slots, events between the halves, one host thread per stream);
             results are bit-identical to the sequential frames.
`value` is the best form's frames/s (normally pipelined); every form's frames/s
and p50/p99 frame intervals are reported under `forms`, with `plain_fps` and
`speedup_vs_plain`. The six-point roofline is timed in the sequential fast form.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "arap-deformation-of-gaussian-radiance-fields_amd")
HBM_PEAK_GBS = 8000.0
# 6 ends 72 + idx 192 + w 192 + s0 12 + static flag 1 + q 16 + SH 192 read;
# mu 12 + q 16 + s 12 + SH 192 + ends 72 written
SIX_POINT_BYTES_PER_GAUSSIAN = 981


def load_package():
    if "arapgs_amd" in sys.modules:
        return sys.modules["arapgs_amd"]
    spec = importlib.util.spec_from_file_location("arapgs_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["arapgs_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def setup(args):
    import torch
    load_package()
    from arapgs_amd import arapgs, autotune, scenes, workload

    # the solve side runs on its own (non-null) stream so that the pipelined form
    # can overlap it with the deform/render stream
    ctx = arapgs.Context(0, stream=torch.cuda.Stream())
    sc = scenes.capsule_scene(args.seed, n=args.n, n_nodes=args.nodes, W=args.width, H=args.height,
                              n_views=args.views)
    log(f"scene: {sc.gaussians.n} Gaussians, {args.nodes} nodes, K={sc.knn_k}")
    with torch.cuda.stream(ctx.stream):
        W = workload.prepare(ctx, sc, log=log)
        solver = workload.make_solver(ctx, W["solver_args"], n_aggregates=args.aggregates)
        log(f"solver: {solver.n_free} free nodes, {solver.n_unknowns} unknowns; "
            f"blocks {[len(b) for b in W['solver_args'][1]]}")
        state = W["state"]
        fb = arapgs.frame_buffers(state, W["K"])
        images = [torch.empty((3, args.height, args.width), dtype=torch.float32, device="cuda:0") for _ in W["cams"]]
        aims = torch.empty_like(state["node_pos"])
    torch.cuda.synchronize()
    return dict(ctx=ctx, sc=sc, state=state, state0=workload.clone_state(state), fb=fb, solver=solver, aims=aims,
                active=W["active"],
                solver_args=W["solver_args"], cfg=dict(autotune.DEFAULTS, aggregates=args.aggregates or 128),
                delta=workload.drag_delta(sc, args.drag_frames), cams=W["cams"], bg=W["bg"], images=images,
                opac=W["opac"], iters=sc.gn_iters, grid=W["grid"], K=W["K"], lpf=W["lpf"])


def reset_state(B):
    """Every form (and every autotuner candidate) starts the drag from the same
    state, so all of them time the same frames. Python's garbage collector runs
    here, outside the timed frames, and is paused until the next reset."""
    import gc
    import torch
    gc.enable()
    gc.collect()
    gc.disable()
    with torch.cuda.stream(B["ctx"].stream):
        for k, v in B["state"].items():
            if v is not None:
                v.copy_(B["state0"][k])
    torch.cuda.synchronize()


def frame(B, plain):
    import torch
    ctx, st = B["ctx"], B["state"]
    with torch.cuda.stream(ctx.stream):
        return _frame(B, plain)


def _frame(B, plain):
    ctx, st = B["ctx"], B["state"]
    sync = ctx.synchronize if plain else (lambda: None)
    ctx.aims_translate(st["node_pos"], B["active"], B["delta"], B["aims"])
    sync()
    _, B["iters_done"] = B["solver"].deform_frame(B["fb"], B["aims"], B["iters"], stop_rule=True)
    sync()
    if not plain:
        # all views in one pass (one sort, one blend launch; bit-identical images)
        img = B["images"][0]
        return ctx.render_views(st["pos"], st["sh"], B["opac"], st["scale"], st["rot"], B["cams"], img.shape[2],
                                img.shape[1], B["bg"], B["images"], tile=B["cfg"]["tile"])
    n_inst = 0
    for cam, img in zip(B["cams"], B["images"]):
        out = ctx.render(st["pos"], st["sh"], B["opac"], st["scale"], st["rot"], cam["view"], cam["proj"],
                         cam["campos"], cam["tx"], cam["ty"], img.shape[2], img.shape[1], B["bg"], out=img,
                         tile=B["cfg"]["tile"])
        n_inst += out["num_rendered"]
        sync()
    return n_inst


def timed_pipelined(B, steps, warmup, solve_cus=None):
    """Frame t's deformation and renders overlap frame t+1's solve (two streams,
    two host threads; pipeline.DragPipeline). solve_cus > 0: the solve stream
    runs on CUs [0, solve_cus) and the deform stream on the others."""
    import torch
    from arapgs_amd import arapgs
    from arapgs_amd.pipeline import DragPipeline
    cus = B["cfg"].get("solve_cus", 0) if solve_cus is None else solve_cus
    reset_state(B)
    ctx = B["ctx"]
    base_stream = ctx.stream
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    streams = []
    if cus > 0:
        streams = [arapgs.CuStream(0, 0, cus), arapgs.CuStream(0, cus, n_cu - cus)]
        ctx.set_stream(streams[0].torch)
    try:
        B["solver"].set_cg_mode(0)
        B["solver"].set_host_checks(B["cfg"]["host_checks"])
        B["solver"].set_warm_start(B["cfg"]["warm"])
        pipe = DragPipeline(B["solver"], B["state"], B["K"], B["active"], B["delta"], B["iters"], B["cams"],
                            B["images"], B["bg"], B["opac"], tile=B["cfg"]["tile"],
                            stream_b=streams[1].torch if streams else None)
        pipe.ctx_b.set_blend_unroll(B["cfg"]["blend_unroll"])
        pipe.run(warmup)
        torch.cuda.synchronize()
        stamps = []
        t0 = time.perf_counter()
        pipe.run(steps, on_frame=lambda t, n: stamps.append(time.perf_counter()))
        torch.cuda.synchronize()
        total = time.perf_counter() - t0
        instances = pipe.instances
        pipe.close()
    finally:
        torch.cuda.synchronize()
        ctx.set_stream(base_stream)
        for s_ in streams:
            s_.close()
    times = [(b - a) * 1e3 for a, b in zip([t0] + stamps[:-1], stamps)]
    return dict(total_s=total, frame_ms=times, kernel_ms=[], instances=instances)


def partition_leg(B, args):
    """Picks the solve stream's CU partition for the pipelined form (0 = both
    streams on all CUs) from a few pipelined frames per candidate."""
    best, best_t = 0, None
    for cus in (0, 16, 32, 48, 64):
        r = timed_pipelined(B, args.tune_steps * 2, 2, solve_cus=cus)
        t = r["total_s"]
        log(f"  partition: solve on {cus} CUs -> {t / (args.tune_steps * 2) * 1e3:.3f} ms/frame")
        if best_t is None or t < best_t:
            best, best_t = cus, t
    return best


def timed(B, plain, steps, warmup):
    import torch
    reset_state(B)
    B["solver"].set_cg_mode(0)
    B["solver"].set_host_checks(1 if plain else B["cfg"]["host_checks"])
    B["solver"].set_warm_start(0 if plain else B["cfg"]["warm"])
    for _ in range(warmup):
        frame(B, plain)
    torch.cuda.synchronize()
    ctx = B["ctx"]
    ctx.profile_enable(True)
    times = []
    n_inst = 0
    t0 = time.perf_counter()
    for i in range(steps):
        a = time.perf_counter()
        n_inst = frame(B, plain)
        torch.cuda.synchronize()
        times.append((time.perf_counter() - a) * 1e3)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    kernel_ms = ctx.profile_read()
    ctx.profile_enable(False)
    gn, cg = B["solver"].last_stats()
    log(f"  last frame: {gn} GN iterations, {cg} CG iterations")
    return dict(total_s=total, frame_ms=times, kernel_ms=kernel_ms, instances=n_inst)


def apply_config(B, cfg):
    """Puts a tuning configuration in place (the solver is rebuilt when the
    aggregate count changes; everything else is a setting)."""
    import torch
    from arapgs_amd import workload
    ctx = B["ctx"]
    if cfg["aggregates"] != B["cfg"]["aggregates"]:
        with torch.cuda.stream(ctx.stream):
            B["solver"].close()
            B["solver"] = workload.make_solver(ctx, B["solver_args"], n_aggregates=cfg["aggregates"])
        torch.cuda.synchronize()
    B["solver"].set_coarse_refresh(cfg["refresh"])
    ctx.set_blend_unroll(cfg["blend_unroll"])
    B["cfg"] = dict(cfg)


def autotune_leg(B, args):
    """Coordinate descent over the tunables on the fast form (a few frames per
    candidate, each from the same starting state)."""
    from arapgs_amd import autotune

    def measure(cfg):
        apply_config(B, cfg)
        r = timed(B, False, args.tune_steps, 2)
        return r["total_s"] / args.tune_steps

    best, trials = autotune.tune(measure, log=log)
    apply_config(B, best)
    return best, trials


def committed_traffic(args):
    """HBM bytes per six-point launch measured with rocprofv3 PMC counters
    (FETCH_SIZE, WRITE_SIZE, scaled by a plain-copy calibration) on the default
    workload, read from the committed summary; None for other workloads."""
    default = (args.n, args.nodes, args.seed) == (1_000_000, 2048, 0)
    path = os.path.join(ROOT, "profiles", "six_point_kernel_capsules-1M_traffic.md")
    if not default or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch six_point_kernel:"):
                return int(line.split(":")[1])
    return None


def pct(v, q):
    s = sorted(v)
    if not s:
        return float("nan")
    k = min(len(s) - 1, max(0, int(math.ceil(q / 100.0 * len(s))) - 1))
    return s[k]


def cpu_baseline(args):
    """The oracle (CPU restatement of the reference path, C++/OpenMP, FP64
    solve) timed on a bounded sample of the same workload."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    load_package()
    from arapgs_amd import scenes
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sc = scenes.capsule_scene(args.seed, n=args.n, n_nodes=args.nodes)
    g = sc.gaussians
    K = sc.knn_k
    rng = np.random.default_rng(0)
    sub = np.sort(rng.choice(g.n, size=min(args.cpu_sample, g.n), replace=False))
    # graph on the full scene (not timed)
    ni = O.fps(g.pos, args.nodes)
    nodes = g.pos[ni]
    nbr = O.graph_edges(nodes, K)
    nidx, nw = O.knn_weights(nodes, nodes, K)
    blocks, types = scenes.blocks_for(sc, nodes)
    aims = scenes.frame_aims(sc, nodes, blocks, types, 1.0 / args.drag_frames)
    prob = O.GNProblem(nodes, nbr, blocks, types, nidx[:, :K], nw, aims)
    log(f"cpu baseline: 1 GN iteration at {prob.n} unknowns ...")
    t = time.perf_counter()
    prob.solve(1, stop_rule=False)
    t_gn1 = time.perf_counter() - t
    # deform of a slice: end-point blend + six-point update (+ SH)
    pos, rot, scale, sh = g.pos[sub], g.rot[sub], g.scale[sub], g.sh[sub]
    ends = O.end_points(pos, rot, scale)
    eidx, ew = O.knn_weights(nodes, ends.reshape(-1, 3), K)
    rot9 = np.tile(np.eye(3).reshape(1, 9), (nodes.shape[0], 1))
    t3 = np.zeros((nodes.shape[0], 3)) + 0.001
    t = time.perf_counter()
    new_ends = O.blend(ends.reshape(-1, 3), eidx[:, :K], ew, nodes, rot9, t3).reshape(-1, 6, 3)
    O.six_point(new_ends, scale, None, pos, rot, scale, sh)
    t_def = time.perf_counter() - t
    frame_s = sc.gn_iters * t_gn1 + t_def * g.n / len(sub)
    return dict(value=1.0 / frame_s, unit="frames/s", cores=threads, kind="port",
                sample=(f"oracle (C++/OpenMP, FP64 envelope-Cholesky GN) on the same scene: 1 of {sc.gn_iters} GN "
                        f"iterations at {prob.n} unknowns ({t_gn1:.1f}s, x{sc.gn_iters}) + end-point blend and "
                        f"six-point update of {len(sub)} of {g.n} Gaussians ({t_def:.2f}s, scaled); no render "
                        f"(the reference renders on the GPU); sample update not included"))


def adjoint_leg(args):
    """The differentiation workload: 500k Gaussians (capsule scene), 1024
    nodes, K=8, 64^3 grid, one 800x800 view. Targets are the chain's outputs
    under a seeded node perturbation (rotations N(0, 5 deg), translations
    N(0, 0.02)); one step = chain forward at the rest table, the L2 loss
    gradient of images + density features + densities, chain backward."""
    import numpy as np
    import torch
    load_package()
    from arapgs_amd import arapgs, scenes, workload
    ctx = arapgs.Context(0, stream=torch.cuda.Stream())
    sc = scenes.capsule_scene(args.seed, n=args.adjoint_n, n_nodes=args.adjoint_nodes, W=args.width, H=args.height,
                              n_views=1)
    with torch.cuda.stream(ctx.stream):
        W = workload.prepare(ctx, sc, log=log)
        st = W["state"]
        M = st["node_pos"].shape[0]
        chain = arapgs.Chain(ctx, M, st["pos"], st["sh"], W["opac"], st["ends_idx"], st["ends_w"], W["cams"],
                             args.width, args.height, W["bg"], grid=W["grid"], sample_idx=st["sample_idx"],
                             sample_w=st["sample_w"], lpf=W["lpf"])
        rng = np.random.default_rng(args.seed + 1)
        table0 = np.zeros((M, 16), np.float32)
        table0[:, 0:3] = st["node_pos"].cpu().numpy()
        table0[:, 3:12] = np.eye(3, dtype=np.float32).reshape(9)
        pert = table0.copy()
        for i in range(M):
            ax = rng.normal(size=3)
            ax /= np.linalg.norm(ax)
            ang = np.radians(5.0) * rng.normal()
            Kx = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
            pert[i, 3:12] = (np.eye(3) + np.sin(ang) * Kx + (1 - np.cos(ang)) * Kx @ Kx).reshape(9)
        pert[:, 12:15] = 0.02 * rng.normal(size=(M, 3))
        table0, pert = torch.from_numpy(table0).cuda(), torch.from_numpy(pert).cuda()
        rot0, scale0 = st["rot"], st["scale0"]
        tgt = chain.forward(pert, rot0, scale0)
        tgt = ([x.clone() for x in tgt[0]], tgt[1].clone(), tgt[2].clone())
    torch.cuda.synchronize()
    log(f"adjoint: {sc.gaussians.n} Gaussians, {M} nodes, {chain.S} samples")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fwd, bwd, tot = [], [], []

    def step():
        with torch.cuda.stream(ctx.stream):
            ev[0].record(ctx.stream)
            imgs, feat, dens = chain.forward(table0, rot0, scale0)
            ev[1].record(ctx.stream)
            dimgs = [2.0 * (a - b) for a, b in zip(imgs, tgt[0])]
            grads = chain.backward(table0, rot0, scale0, dimages=dimgs, dfeat=2.0 * (feat - tgt[1]),
                                   ddens=2.0 * (dens - tgt[2]))
            ev[2].record(ctx.stream)
            return grads

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.adjoint_steps):
        a = time.perf_counter()
        grads = step()
        torch.cuda.synchronize()
        tot.append((time.perf_counter() - a) * 1e3)
        fwd.append(ev[0].elapsed_time(ev[1]))
        bwd.append(ev[1].elapsed_time(ev[2]))
    total = time.perf_counter() - t0
    gnorm = float(torch.linalg.vector_norm(grads[0][:, 3:15]).item())
    chain.close()
    ctx.close()
    return {"workload": f"capsules-{args.adjoint_n // 1000}k differentiation", "gaussians": args.adjoint_n,
            "nodes": M, "samples": int(chain.S), "views": 1, "resolution": [args.width, args.height],
            "steps": args.adjoint_steps, "steps_per_s": round(args.adjoint_steps / total, 3),
            "step_p50_ms": round(pct(tot, 50), 3), "forward_p50_ms": round(pct(fwd, 50), 3),
            "backward_p50_ms": round(pct(bwd, 50), 3), "node_grad_norm": gnorm,
            "check": "vector-Jacobian products vs FP64 forward-mode oracle along 64 directions (tests/test_gpu_chain.py)"}


def realign_leg(B, args):
    """Density-preserving realignment on the main workload after the timed drag:
    the deformed Gaussians are moved by Adam so that their density at the
    deformed samples matches the rotated rest features (the F6 loop). One step =
    activation + density sample + L1 subgradient + density backward + Adam."""
    import torch
    load_package()
    from arapgs_amd import arapgs
    ctx, st, grid = B["ctx"], B["state"], B["grid"]
    with torch.cuda.stream(ctx.stream):
        pos, rot, sh = st["pos"].clone(), st["rot"].clone(), st["sh"].clone()
        log_scale = torch.log(st["scale"])
        o = B["opac"].clamp(1e-6, 1 - 1e-6)
        logit = torch.log(o / (1 - o))
        _, aim_dens = grid.density(B["state0"]["pos"], B["state0"]["rot"], B["state0"]["scale"], B["opac"],
                                   B["state0"]["sh"], B["lpf"])
        r = arapgs.Realign(ctx, grid, pos.shape[0])
        loss = torch.zeros(1, dtype=torch.float64, device=pos.device)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ms, losses = [], []
    for i in range(args.realign_steps + 1):
        with torch.cuda.stream(ctx.stream):
            ev[0].record(ctx.stream)
            r.step(st["samples"], st["sample_feat"], aim_dens, B["lpf"], pos, rot, log_scale, logit, sh, loss=loss)
            ev[1].record(ctx.stream)
        torch.cuda.synchronize()
        losses.append(float(loss.item()))
        if i > 0:
            ms.append(ev[0].elapsed_time(ev[1]))
    r.close()
    return {"gaussians": int(pos.shape[0]), "samples": int(st["samples"].shape[0]),
            "cell_gaussian_pairs": int(grid.n_pairs), "steps": args.realign_steps,
            "step_p50_ms": round(pct(ms, 50), 3), "l1_loss_first_last": [losses[0], losses[-1]]}


def batch_leg(args):
    """The uneven scene batch: args.batch_scenes capsule scenes with Gaussian
    counts log-uniform in [50k, 2M], nodes = ceil(n / 500) rounded up to a
    multiple of 64, K=8, 64^3 grids, one 800x800 view each, 10 GN iterations
    with the stop rule, CG warm start; a seeded random-walk drag (a step of
    0.01 in a new random direction per scene and frame). One frame = every
    scene's aims, solve, node table, node / sample / end-point blends,
    six-point update, sample SH and render. Timed twice: batched (one launch
    per step over all scenes) and scene after scene with single-scene calls."""
    import torch
    load_package()
    from arapgs_amd import arapgs, batch, workload
    ctx = arapgs.Context(0, stream=torch.cuda.Stream())
    n = args.batch_scenes
    with torch.cuda.stream(ctx.stream):
        wl = workload.prepare_batch(ctx, workload.batch_scenes(args.seed, n, W=args.width, H=args.height), log=log)
        sb = batch.SceneBatch(ctx, wl, gn_iters=10)
    seg, ns, ss = wl["seg"], wl["node_seg"], wl["sample_seg"]
    counts = [seg[i + 1] - seg[i] for i in range(n)]
    log(f"batch: {n} scenes, {seg[-1]} Gaussians ({min(counts)}..{max(counts)}), {ns[-1]} nodes, {ss[-1]} samples")
    frames = args.batch_frames
    steps = batch.random_walk_steps(args.seed + 2, n, 2 * (args.warmup + frames), step=0.01)
    res = {}
    k = 0
    for form in ("batched", "looped"):
        run = sb.frame if form == "batched" else sb.frame_looped
        ms = []
        for f in range(args.warmup + frames):
            torch.cuda.synchronize()
            a = time.perf_counter()
            with torch.cuda.stream(ctx.stream):
                inst = run(steps[k])
            torch.cuda.synchronize()
            k += 1
            if f >= args.warmup:
                ms.append((time.perf_counter() - a) * 1e3)
        total = sum(ms) / 1e3
        res[form] = {"frames_per_s": round(frames / total, 3), "scene_frames_per_s": round(n * frames / total, 2),
                     "gaussians_per_s": round(seg[-1] * frames / total, 1), "p50_ms": round(pct(ms, 50), 3),
                     "p99_ms": round(pct(ms, 99), 3), "tile_instances_per_frame": int(inst)}
        log(f"  batch {form}: {res[form]['frames_per_s']} frames/s ({res[form]['scene_frames_per_s']} scene-frames/s), "
            f"p50 {res[form]['p50_ms']} ms, p99 {res[form]['p99_ms']} ms")
    gn = [int(t.item()) for t in sb.iters_done]
    sb.close()
    ctx.close()
    return {"workload": f"{n} uneven capsule scenes", "scenes": n, "gaussians_total": int(seg[-1]),
            "gaussians_min_max": [int(min(counts)), int(max(counts))], "nodes_total": int(ns[-1]),
            "samples_total": int(ss[-1]), "views": n, "resolution": [args.width, args.height], "gn_iters": 10,
            "frames": frames, "last_frame_gn_iters_min_max": [min(gn), max(gn)], **res,
            "speedup_batched_vs_looped": round(res["batched"]["frames_per_s"] / res["looped"]["frames_per_s"], 3),
            "check": "batched frames = scene-by-scene frames bit for bit (tests/test_gpu_scene_batch.py)"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nodes", type=int, default=2048)
    ap.add_argument("--views", type=int, default=4)
    ap.add_argument("--width", type=int, default=800)
    ap.add_argument("--height", type=int, default=800)
    ap.add_argument("--drag-frames", type=int, default=30)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=20_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="plain,fast,pipelined")
    ap.add_argument("--adjoint-steps", type=int, default=10, help="steps of the differentiation leg (0: skip)")
    ap.add_argument("--aggregates", type=int, default=0, help="coarse aggregates of the CG preconditioner (0: 128)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip the autotuner (tile, blend unroll, aggregates, coarse refresh, CG host checks)")
    ap.add_argument("--tune-steps", type=int, default=6, help="timed frames per autotuner candidate")
    ap.add_argument("--partition", action="store_true",
                    help="sweep CU partitions between the solve and deform streams of the pipelined form")
    ap.add_argument("--adjoint-n", type=int, default=500_000)
    ap.add_argument("--realign-steps", type=int, default=5, help="steps of the realignment leg (0: skip)")
    ap.add_argument("--adjoint-nodes", type=int, default=1024)
    ap.add_argument("--batch-scenes", type=int, default=32, help="scenes of the uneven-batch leg (0: skip)")
    ap.add_argument("--batch-frames", type=int, default=8, help="timed frames per form of the uneven-batch leg")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 3 plain device copies of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the per-frame path is single-GPU. "
              "Across GPUs the Gaussians and samples would shard by contiguous ranges with a broadcast of the "
              "solved node table; independent scenes would map one group per GPU.", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.copy_probe > 0:
        n = args.copy_probe * (1 << 20) // 4
        a = torch.rand(n, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n, "copies": 3}), flush=True)
        return
    B = setup(args)
    apply_config(B, B["cfg"])
    tuning = None
    if args.autotune:
        log(f"autotuning on the fast form ({args.tune_steps} frames per candidate)")
        best, trials = autotune_leg(B, args)
        tuning = {"config": best, "candidates_timed": len(trials),
                  "ms_per_frame": {json.dumps(c, sort_keys=True): round(t * 1e3, 3) for c, t in trials}}
        log(f"  autotune: {best}")
    res = {}
    for form in args.forms.split(","):
        log(f"timing form={form}: {args.warmup} warmup + {args.steps} steps")
        if form == "pipelined":
            if args.partition and "solve_cus" not in B["cfg"]:
                B["cfg"]["solve_cus"] = partition_leg(B, args)
                log(f"  partition: solve stream on {B['cfg']['solve_cus']} CUs")
                if tuning is not None:
                    tuning["config"] = dict(B["cfg"])
            res[form] = timed_pipelined(B, args.steps, args.warmup)
        else:
            res[form] = timed(B, form == "plain", args.steps, args.warmup)
        r = res[form]
        log(f"  {form} frame ms: {' '.join(f'{t:.1f}' for t in r['frame_ms'])}")
        log(f"  {form}: {args.steps / r['total_s']:.2f} frames/s, p50 {pct(r['frame_ms'], 50):.2f} ms, "
            f"p99 {pct(r['frame_ms'], 99):.2f} ms, {r['instances']} tile instances/frame")
    best = max(res, key=lambda f: args.steps / res[f]["total_s"])
    r = res[best]
    fps = args.steps / r["total_s"]
    # the six-point kernel is timed (HIP events on its stream) in the sequential forms
    kr = res.get("fast") or res.get("plain") or r
    kms = kr["kernel_ms"]
    avg_k = float(sum(kms) / len(kms)) if len(kms) else float("nan")
    bytes_per_launch = SIX_POINT_BYTES_PER_GAUSSIAN * args.n
    achieved = bytes_per_launch / (avg_k * 1e-3) / 1e9 if len(kms) else None
    line = {
        "metric": "deform+render frames/s @1M Gaussians, adjoint+batched vs plain; p50/p99 ms",
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r["total_s"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "capsules-1M", "gaussians": args.n, "nodes": args.nodes, "knn_k": 8,
                   "grid": 64, "gn_iters": 10, "views": args.views, "resolution": [args.width, args.height],
                   "sh_degree": 3, "form": best},
        "p50_ms": round(pct(r["frame_ms"], 50), 3),
        "p99_ms": round(pct(r["frame_ms"], 99), 3),
    }
    line["forms"] = {f: {"fps": round(args.steps / v["total_s"], 3), "p50_ms": round(pct(v["frame_ms"], 50), 3),
                         "p99_ms": round(pct(v["frame_ms"], 99), 3)} for f, v in res.items()}
    if "plain" in res:
        rp = res["plain"]
        line["plain_fps"] = round(args.steps / rp["total_s"], 3)
        line["plain_p50_ms"] = round(pct(rp["frame_ms"], 50), 3)
        line["plain_p99_ms"] = round(pct(rp["frame_ms"], 99), 3)
        line["speedup_vs_plain"] = round(fps / line["plain_fps"], 3)
    line["roofline"] = {"bound": "hbm", "kernel": "six_point_kernel<8, true> (fused end-point blend + polar + SH)",
                        "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                        "traffic": committed_traffic(args),
                        "bytes_per_launch": bytes_per_launch, "avg_launch_ms": round(avg_k, 4),
                        "launches": int(len(kms))}
    import gc
    gc.enable()
    line["config"]["tuning"] = B["cfg"]
    if tuning:
        line["autotune"] = tuning
    if args.realign_steps > 0:
        log(f"timing the realignment leg: {args.realign_steps} steps")
        line["realign"] = realign_leg(B, args)
        log(f"  realign: {line['realign']['step_p50_ms']} ms/step, L1 {line['realign']['l1_loss_first_last']}")
    if args.adjoint_steps > 0:
        log(f"timing the differentiation leg: {args.warmup} warmup + {args.adjoint_steps} steps")
        line["adjoint"] = adjoint_leg(args)
        log(f"  adjoint: {line['adjoint']['steps_per_s']} steps/s, forward p50 {line['adjoint']['forward_p50_ms']} ms, "
            f"backward p50 {line['adjoint']['backward_p50_ms']} ms")
    if args.batch_scenes > 0:
        log(f"timing the uneven-batch leg: {args.batch_scenes} scenes")
        line["batch"] = batch_leg(args)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
