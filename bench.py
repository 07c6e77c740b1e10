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
             and the host looks at it only every few CG iterations; the views
             are rendered in one pass (arapgs_render_views);
  pipelined  as fast, with frame t's deformation and renders on a second stream
             overlapping frame t+1's solve (pipeline.DragPipeline: two node-table
             slots, events between the halves, one host thread per stream);
             results are bit-identical to the sequential frames.
`value` is the best form's frames/s (normally pipelined); every form's frames/s
and p50/p99 frame intervals are reported under `forms`, with `plain_fps` and
`speedup_vs_plain`. The six-point, sample blend and sample SH kernels are timed
with HIP events in the fast form (`rooflines`; `roofline` is the longest one).

Further legs, each in its own JSON object:
  fixed_iteration  the full edit in one frame, 10 GN iterations with the stop
                   rule off, CG cold-started (plain and fast forms);
  realign          density-preserving realignment steps after the drag;
  cpu_baseline     one complete drag frame of the same workload on the oracle
                   (the CPU restatement of the reference path), measured, plus
                   the GPU's energies of the timed frames against the oracle's
                   FP64 Gauss-Newton solve (the timed frames are re-run with
                   their inputs recorded; the final state must match);
  adjoint          the differentiation workload (500k Gaussians);
  batch            32 uneven scenes, batched vs scene after scene, 64 frames;
  large            3M flat Gaussians, 4096 nodes, 128^3 grid, 8 views at
                   1920x1080: pipelined drag, kernel HBM rates, release step.

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
# Algorithmic bytes per unit of the three deformation kernels timed with HIP events:
#   six-point, per Gaussian: 6 ends 72 + idx 192 + w 192 + s0 12 + static flag 1 + q 16 + SH 192 read;
#                            mu 12 + q 16 + s 12 + SH 192 + ends 72 written
#   sample blend, per sample: idx 32 + w 32 + position 12 read, position 12 written
#   sample SH, per sample: idx 32 + w 32 + feature 192 read, feature 192 written
# (the node tables, M x 64 B and M x 16 B, are read once per launch and left out)
SIX_POINT_BYTES_PER_GAUSSIAN = 981
SAMPLE_BLEND_BYTES_PER_SAMPLE = 88
SAMPLE_SH_BYTES_PER_SAMPLE = 448
KERNELS = (("six_point", "six_point_kernel<8, true, true> (fused end-point blend + polar + SH, LDS node cache)", 0),
           ("sample_blend", "blend_points_kernel<8, true, true> (sample positions, LDS node cache)", 1),
           ("sample_sh", "sample_sh_kernel<8> (slerp of node rotations + SH rotation of the sample features)", 2))


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
    state, Gaussians and solver alike, so all of them time the same frames. Python's garbage collector runs
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
    if B.get("solver") is not None:
        # warm-start steps, coarse operator and refresh count as a new solver's
        B["solver"].reset()


def frame(B, plain, delta=None, stop_rule=True):
    import torch
    ctx, st = B["ctx"], B["state"]
    with torch.cuda.stream(ctx.stream):
        return _frame(B, plain, B["delta"] if delta is None else delta, stop_rule)


def _frame(B, plain, delta, stop_rule):
    ctx, st = B["ctx"], B["state"]
    sync = ctx.synchronize if plain else (lambda: None)
    ctx.aims_translate(st["node_pos"], B["active"], delta, B["aims"])
    sync()
    B["energies"], B["iters_done"] = B["solver"].deform_frame(B["fb"], B["aims"], B["iters"], stop_rule=stop_rule)
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


def partition_streams(ctx, solve_cus, deform_skip, samples_cus=0, side=None):
    """CU-masked streams of the pipelined form. solve_cus > 0: ctx (the solve)
    moves to a stream on CUs [0, solve_cus) and the deform stream runs on the
    other CUs; the samples stream runs beside the deform (samples_cus 0), on
    every CU (1) or beside the solve (2); samples_cus 3 queues the samples on
    the deform stream itself. Else deform_skip > 0: the deform and
    samples streams leave CUs [0, deform_skip) to the solve, which keeps every
    CU. side: a solver whose asynchronous coarse rebuilds move to a stream on
    the deform side's CUs (solve_cus > 0). Returns the streams to close (after
    the caller restores ctx's stream and the solver's side stream) and the
    deform and samples streams (None: the pipeline's own)."""
    import torch
    from arapgs_amd import arapgs
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    if solve_cus > 0:
        streams = [arapgs.CuStream(0, 0, solve_cus), arapgs.CuStream(0, solve_cus, n_cu - solve_cus)]
        if samples_cus == 0:
            streams.append(arapgs.CuStream(0, solve_cus, n_cu - solve_cus))
        elif samples_cus == 2:
            streams.append(arapgs.CuStream(0, 0, solve_cus))
        ctx.set_stream(streams[0].torch)
        stream_c = streams[2].torch if len(streams) > 2 else None
        if side is not None:
            streams.append(arapgs.CuStream(0, solve_cus, n_cu - solve_cus))
            side.set_side_stream(streams[-1].torch)
        streams.append(arapgs.CuStream(0, solve_cus, n_cu - solve_cus))  # the renders (split_render)
        return streams, streams[1].torch, stream_c, streams[-1].torch
    if deform_skip > 0:
        streams = [arapgs.CuStream(0, deform_skip, n_cu - deform_skip) for _ in range(3)]
        return streams, streams[0].torch, streams[1].torch, streams[2].torch
    return [], None, None, None


def timed_pipelined(B, steps, warmup, solve_cus=None, sample_batch=None):
    """Frame t's deformation and renders overlap frame t+1's solve (two streams,
    two host threads; pipeline.DragPipeline). solve_cus > 0: the solve stream
    runs on CUs [0, solve_cus) and the deform stream on the others.
    cfg "deform_skip" > 0 (with solve_cus 0): the solve keeps every CU and the
    deform and samples streams leave CUs [0, deform_skip) to it."""
    import torch
    from arapgs_amd import arapgs
    from arapgs_amd.pipeline import DragPipeline
    cus = B["cfg"].get("solve_cus", 0) if solve_cus is None else solve_cus
    reset_state(B)
    ctx = B["ctx"]
    base_stream = ctx.stream
    side = B["solver"] if B["cfg"].get("side_cus", 0) else None
    streams, stream_b, stream_c, stream_r = partition_streams(ctx, cus, B["cfg"].get("deform_skip", 0),
                                                    B["cfg"].get("samples_cus", 0), side=side)
    try:
        B["solver"].set_cg_mode(B["cfg"].get("cg_mode", 0))
        B["solver"].set_cg_predict(B["cfg"].get("cg_predict", 0))
        B["solver"].set_host_spin(B["cfg"].get("host_spin", 0))
        B["solver"].set_host_checks(B["cfg"]["host_checks"])
        B["solver"].set_warm_start(B["cfg"]["warm"])
        B["solver"].set_cg_graphs(B["cfg"].get("cg_graphs", 0))
        nb = B["cfg"].get("sample_batch", 1) if sample_batch is None else sample_batch
        pipe = DragPipeline(B["solver"], B["state"], B["K"], B["active"], B["delta"], B["iters"], B["cams"],
                            B["images"], B["bg"], B["opac"], tile=B["cfg"]["tile"],
                            stream_b=stream_b, stream_c=stream_c, stream_r=stream_r, sample_batch=nb,
                            split_render=bool(B["cfg"].get("split_render", 0)),
                            samples_on_deform=B["cfg"].get("samples_cus", 0) == 3,
                            node_cache=bool(B["cfg"].get("node_cache", 1)),
                            sh_compose=bool(B["cfg"].get("sh_compose", 0)),
                            samples_late=bool(B["cfg"].get("samples_late", 0)))
        pipe.ctx_render.set_blend_unroll(B["cfg"]["blend_unroll"])
        pipe.ctx_render.set_blend_split(B["cfg"].get("blend_split", 0))
        pipe.ctx_render.set_tile_cull(B["cfg"].get("tile_cull", 0))
        for c_ in {id(pipe.ctx_b): pipe.ctx_b, id(pipe.ctx_render): pipe.ctx_render}.values():
            c_.set_host_spin(B["cfg"].get("host_spin", 0))
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
        if side is not None:
            side.set_side_stream(None)
        for s_ in streams:
            s_.close()
    times = [(b - a) * 1e3 for a, b in zip([t0] + stamps[:-1], stamps)]
    return dict(total_s=total, frame_ms=times, kernel_ms={}, instances=instances)


def sample_batch_leg(B, args):
    """Picks the chunk count of the pipelined form's rolling sample update
    from a few pipelined frames per candidate (results are identical)."""
    best, best_t, trials = 1, None, {}
    for nb in (1, 2, 4, 8):
        n, w = tune_frames(args)
        r = timed_pipelined(B, n, w, sample_batch=nb)
        t = r["total_s"] / n
        trials[nb] = round(t * 1e3, 3)
        log(f"  sample batch {nb}: {t * 1e3:.3f} ms/frame")
        if best_t is None or t < best_t:
            best, best_t = nb, t
    return best, trials


def partition_leg(B, args):
    """Picks the solve stream's CU partition for the pipelined form (0 = both
    streams on all CUs) from a few pipelined frames per candidate."""
    best, best_t = 0, None
    for cus in (0, 16, 32, 48, 64, 128):
        n, w = tune_frames(args)
        r = timed_pipelined(B, n, w, solve_cus=cus)
        t = r["total_s"]
        log(f"  partition: solve on {cus} CUs -> {t / n * 1e3:.3f} ms/frame")
        if best_t is None or t < best_t:
            best, best_t = cus, t
    return best


def timed(B, plain, steps, warmup):
    import torch
    reset_state(B)
    B["solver"].set_cg_mode(0 if plain else B["cfg"].get("cg_mode", 0))
    B["solver"].set_cg_predict(0 if plain else B["cfg"].get("cg_predict", 0))
    B["solver"].set_host_spin(0 if plain else B["cfg"].get("host_spin", 0))
    B["ctx"].set_host_spin(0 if plain else B["cfg"].get("host_spin", 0))
    B["solver"].set_host_checks(1 if plain else B["cfg"]["host_checks"])
    B["solver"].set_warm_start(0 if plain else B["cfg"]["warm"])
    B["solver"].set_cg_graphs(0 if plain else B["cfg"].get("cg_graphs", 0))
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
    kernel_ms = {name: ctx.profile_read(which=which) for name, _, which in KERNELS}
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
    B["solver"].set_coarse_lag(cfg.get("coarse_lag", 1))
    B["solver"].set_cg_block(cfg.get("cg_block", 256))
    B["solver"].set_cg_graphs(cfg.get("cg_graphs", 0))
    B["solver"].set_cg_mode(cfg.get("cg_mode", 0))
    B["solver"].set_cg_predict(cfg.get("cg_predict", 0))
    B["solver"].set_cg_assembled(cfg.get("cg_asm", 0))
    B["solver"].set_cg_tol(cfg.get("cg_tol", 1e-4))
    B["solver"].set_host_spin(cfg.get("host_spin", 0))
    ctx.set_blend_unroll(cfg["blend_unroll"])
    ctx.set_blend_split(cfg.get("blend_split", 0))
    ctx.set_node_cache(cfg.get("node_cache", 1))
    ctx.set_tile_cull(cfg.get("tile_cull", 0))
    ctx.set_host_spin(cfg.get("host_spin", 0))
    B["cfg"] = dict(cfg)


def tune_frames(args):
    """(timed, warmup) frames per tuning candidate: by default the timed drag
    itself, so that settings whose effect grows over the drag (the coarse
    operator's age, the warm start) are judged on the frames that are timed."""
    if args.tune_steps > 0:
        return args.tune_steps, 2
    return args.steps, args.warmup


def autotune_leg(B, args):
    """Coordinate descent over the tunables on the form that is reported, the
    pipelined drag (4 rolling sample chunks while tuning), every candidate
    timed over the same frames from the same starting state."""
    from arapgs_amd import autotune

    def measure(cfg):
        # the best of three runs: frame times vary by a few percent run to run
        apply_config(B, cfg)
        n, w = tune_frames(args)
        return min(timed_pipelined(B, n, w, sample_batch=4)["total_s"] / n for _ in range(3))

    best, trials = autotune.tune(measure, log=log)
    apply_config(B, best)
    return best, trials


def committed_traffic(args, kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 PMC counters
    (FETCH_SIZE, WRITE_SIZE, scaled by a plain-copy calibration) on the default
    workload, read from the committed summary; None for other workloads."""
    default = (args.n, args.nodes, args.seed) == (1_000_000, 2048, 0)
    path = os.path.join(ROOT, "profiles", "deform_kernels_capsules-1M_traffic.md")
    if not default or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                return int(line.split(":")[1])
    return None


def rooflines(args, B, kernel_ms):
    """HBM roofline of each timed deformation kernel: algorithmic bytes per
    launch over its average duration (HIP events on the kernel's stream)."""
    n = args.n
    S = int(B["state"]["samples"].shape[0])
    units = {"six_point": (SIX_POINT_BYTES_PER_GAUSSIAN, n, "Gaussians"),
             "sample_blend": (SAMPLE_BLEND_BYTES_PER_SAMPLE, S, "samples"),
             "sample_sh": (SAMPLE_SH_BYTES_PER_SAMPLE, S, "samples")}
    out = {}
    for name, label, _ in KERNELS:
        kms = kernel_ms.get(name, [])
        per, cnt, what = units[name]
        avg = float(sum(kms) / len(kms)) if len(kms) else float("nan")
        b = per * cnt
        ach = b / (avg * 1e-3) / 1e9 if len(kms) else None
        out[name] = {"bound": "hbm", "kernel": label, "achieved": round(ach, 1) if ach else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4) if ach else None,
                     "traffic": committed_traffic(args, name), "bytes_per_launch": b,
                     "bytes_per_unit": per, "units_per_launch": cnt, "unit_name": what,
                     "avg_launch_ms": round(avg, 4), "launches": int(len(kms))}
    return out


def pct(v, q):
    s = sorted(v)
    if not s:
        return float("nan")
    k = min(len(s) - 1, max(0, int(math.ceil(q / 100.0 * len(s))) - 1))
    return s[k]


def record_frames(B, steps, warmup):
    """Re-runs the timed pipelined drag (same start state, same settings) while
    keeping every frame's solve inputs (node positions, aims) and outputs
    (energies, GN iterations), copied on the solve stream, and the state at the
    start of the first frame. The final state must equal the timed run's bit for
    bit, so the recorded frames are the timed frames."""
    import numpy as np
    import torch
    from arapgs_amd.pipeline import DragPipeline

    class Recording(DragPipeline):
        def _solve_half(self, slot):
            with torch.cuda.stream(self.stream_a):
                g = self.st["node_pos"].clone()
            super()._solve_half(slot)
            with torch.cuda.stream(self.stream_a):
                self.rec.append((g, self.aims.clone(), self.energies[slot].clone(), self.iters_done[slot].clone()))

    final_timed = {k: B["state"][k].clone() for k in ("pos", "rot", "scale", "sh", "node_pos", "samples",
                                                      "sample_feat")}
    reset_state(B)
    start = {k: (v.cpu().numpy() if v is not None else None) for k, v in B["state"].items()}
    B["solver"].set_cg_mode(B["cfg"].get("cg_mode", 0))
    B["solver"].set_cg_predict(B["cfg"].get("cg_predict", 0))
    B["solver"].set_host_spin(B["cfg"].get("host_spin", 0))
    B["solver"].set_host_checks(B["cfg"]["host_checks"])
    B["solver"].set_warm_start(B["cfg"]["warm"])
    B["solver"].set_cg_graphs(B["cfg"].get("cg_graphs", 0))
    pipe = Recording(B["solver"], B["state"], B["K"], B["active"], B["delta"], B["iters"], B["cams"], B["images"],
                     B["bg"], B["opac"], tile=B["cfg"]["tile"], sample_batch=B["cfg"].get("sample_batch", 1),
                     node_cache=bool(B["cfg"].get("node_cache", 1)), sh_compose=bool(B["cfg"].get("sh_compose", 0)),
                     split_render=bool(B["cfg"].get("split_render", 0)),
                     samples_on_deform=B["cfg"].get("samples_cus", 0) == 3,
                     samples_late=bool(B["cfg"].get("samples_late", 0)))
    pipe.rec = []
    pipe.ctx_render.set_blend_unroll(B["cfg"]["blend_unroll"])
    pipe.ctx_render.set_blend_split(B["cfg"].get("blend_split", 0))
    pipe.ctx_render.set_tile_cull(B["cfg"].get("tile_cull", 0))
    pipe.run(warmup + steps)
    torch.cuda.synchronize()
    same = all(torch.equal(final_timed[k], B["state"][k]) for k in final_timed)
    rec = [(g.cpu().numpy(), a.cpu().numpy(), e.cpu().numpy(), int(d.item())) for g, a, e, d in pipe.rec]
    pipe.close()
    return dict(frames=rec, timed_from=warmup, same_as_timed=bool(same), start=start)


def _cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = nproc
    return model, nproc, share


def _set_threads(n):
    """OpenMP threads of the oracle library for the following calls."""
    import ctypes
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))
    except OSError:
        pass


def cpu_frame(B, rec, O):
    """One complete drag frame of the reference's CPU path on the oracle
    (C++/OpenMP, FP64 Gauss-Newton with the reference's per-iteration ordering
    and two factorisations), from the GPU's state at the start of the timed
    drag: solve, node table, sample blend, end-point blend + six-point update,
    node blend, node quaternions + sample SH rotation. The reference renders on
    the GPU, so the frame has no render. Returns per-stage seconds and the
    solve's energy / iterations."""
    import numpy as np
    st = rec["start"]
    g, aims = rec["frames"][0][0], rec["frames"][0][1]
    nbr, blocks, types, src_idx, src_w, _ = B["solver_args"]
    K = B["K"]
    t = {}
    t0 = time.perf_counter()
    prob = O.GNProblem(g, nbr, blocks, types, src_idx, src_w.astype(np.float64), aims)
    x, e, it = prob.solve(B["iters"], stop_rule=True, reference_cost=True)
    t["solve"] = time.perf_counter() - t0
    # putFreeInputs (Deform.hpp:140-151): free nodes in ascending order, x column-major
    M = g.shape[0]
    static = np.zeros(M, bool)
    for blk, ty in zip(blocks, types):
        if ty == -1:
            static[np.asarray(blk)] = True
    free = np.nonzero(~static)[0]
    rot9 = np.tile(np.eye(3).reshape(1, 9), (M, 1))
    t3 = np.zeros((M, 3))
    xr = x.reshape(-1, 12)
    rot9[free] = xr[:, :9]
    t3[free] = xr[:, 9:]
    t0 = time.perf_counter()
    O.blend(st["samples"], st["sample_idx"], st["sample_w"], g, rot9, t3)
    t["sample_blend"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    ends = O.blend(st["ends"].reshape(-1, 3), st["ends_idx"].reshape(-1, K), st["ends_w"].reshape(-1, K), g, rot9,
                   t3).reshape(-1, 6, 3)
    O.six_point(ends, st["scale0"], None, st["pos"], st["rot"], st["scale"], st["sh"])
    t["six_point"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.blend(g, st["node_src_idx"], st["node_src_w"], g, rot9, t3)
    q = O.node_quats(rot9)
    O.sample_sh(st["sample_w"], st["sample_idx"], q, None, st["sample_feat"])
    t["nodes_sample_sh"] = time.perf_counter() - t0
    return dict(seconds=t, total_s=sum(t.values()), energy=float(e[-1]), gn_iters=int(it))


def cpu_baseline(args, B, rec):
    """The oracle (CPU restatement of the reference path, C++/OpenMP, FP64
    solve) running one complete drag frame of the same workload, timed on this
    host's cores at min(16, nproc) threads (the reference's hard-coded 16,
    main.cpp:102-103) and at all the cores of this process; it also checks the
    GPU's energies of the timed frames against the oracle's Gauss-Newton solve."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    model, nproc, share = _cpu_info()
    res = {}
    for label, th in (("ref_threads", min(16, nproc)), ("all_cores", share)):
        if label == "all_cores" and th == res.get("ref_threads", {}).get("threads"):
            res[label] = dict(res["ref_threads"], note="same thread count as ref_threads")
            continue
        _set_threads(th)
        log(f"cpu baseline: one complete frame on the oracle, {th} threads ...")
        f = cpu_frame(B, rec, O)
        f["threads"] = th
        res[label] = f
        log(f"  cpu frame: {f['total_s']:.2f} s ({', '.join(f'{k} {v:.2f}' for k, v in f['seconds'].items())})")
    _set_threads(min(16, nproc))
    # energies of the timed frames against the oracle's solve of the same problems
    frames = rec["frames"][rec["timed_from"]:]
    if args.energy_check_frames >= 0:
        frames = frames[: args.energy_check_frames]
    nbr, blocks, types, src_idx, src_w, _ = B["solver_args"]
    errs, gn = [], []
    t0 = time.perf_counter()
    for g, aims, e, done in frames:
        prob = O.GNProblem(g, nbr, blocks, types, src_idx, src_w.astype(np.float64), aims)
        _, er, it = prob.solve(B["iters"], stop_rule=True)
        errs.append(abs(e[done] - er[-1]) / er[-1])
        gn.append((done, it))
    log(f"  energy check: {len(frames)} timed frames in {time.perf_counter() - t0:.1f} s, "
        f"max rel err {max(errs) if errs else float('nan'):.2e}")
    first = rec["frames"][0]
    main = res["ref_threads"]
    return dict(
        value=round(1.0 / main["total_s"], 5), unit="frames/s", cores=main["threads"], kind="port",
        sample=(f"one complete drag frame of the same workload (1M Gaussians, {B['state']['samples'].shape[0]} "
                f"samples, 2048 nodes, the GPU's first drag frame from the same state) on the oracle (C++/OpenMP, "
                f"FP64 Gauss-Newton with a minimum-degree block Cholesky, re-analysed and factorised twice per "
                f"iteration as Deform.cpp:131-132 does; {main['gn_iters']} GN iterations, the GPU took {first[3]}): "
                f"solve, sample blend, end-point blend + six-point, node blend, sample SH; no render (the reference "
                f"renders on the GPU); measured, not extrapolated"),
        cpu_model=model, nproc=nproc, cpu_share=share, frames=res,
        energy_check={"frames": len(frames), "max_rel_err": max(errs) if errs else None,
                      "gn_iters_gpu_oracle": gn, "bar": 1e-5,
                      "recorded_frames_equal_timed": rec["same_as_timed"]},
        first_frame_energy={"gpu": float(first[2][first[3]]), "oracle": main["energy"]})


def fixed_iteration_leg(B, args):
    """The unit of work with a fixed iteration count: one frame of the whole
    edit (the active capsule translated by the full 0.3 in one frame), 10
    Gauss-Newton iterations with the stop rule off and the CG cold-started,
    then the deformation and the 4 views; every frame starts from the rest
    state (restored outside the timed frame)."""
    import torch
    from arapgs_amd import workload
    full = workload.drag_delta(B["sc"], 1)
    res = {}
    B["solver"].set_warm_start(0)
    for form in ("plain", "fast"):
        plain = form == "plain"
        B["solver"].set_cg_mode(0 if plain else B["cfg"].get("cg_mode", 0))
        B["solver"].set_cg_predict(0 if plain else B["cfg"].get("cg_predict", 0))
        B["solver"].set_host_spin(0 if plain else B["cfg"].get("host_spin", 0))
        B["solver"].set_host_checks(1 if plain else B["cfg"]["host_checks"])
        B["solver"].set_cg_graphs(0 if plain else B["cfg"].get("cg_graphs", 0))
        ms = []
        for i in range(args.warmup + args.fixed_steps):
            reset_state(B)
            a = time.perf_counter()
            frame(B, plain, delta=full, stop_rule=False)
            torch.cuda.synchronize()
            if i >= args.warmup:
                ms.append((time.perf_counter() - a) * 1e3)
        gn, cg = B["solver"].last_stats()
        e = B["energies"].cpu().numpy()
        res[form] = {"fps": round(len(ms) / (sum(ms) / 1e3), 3), "p50_ms": round(pct(ms, 50), 3),
                     "p99_ms": round(pct(ms, 99), 3), "gn_iters": gn, "cg_iters": cg,
                     "energy_first_last": [float(e[0]), float(e[gn])]}
        log(f"  fixed-iteration {form}: {res[form]['fps']} frames/s, p50 {res[form]['p50_ms']} ms "
            f"({gn} GN / {cg} CG iterations)")
    B["solver"].set_warm_start(B["cfg"]["warm"])
    reset_state(B)
    return {"workload": "capsules-1M, full 0.3 edit in one frame", "gn_iters": 10, "stop_rule": False,
            "cg_warm_start": False, "frames": args.fixed_steps, **res,
            "speedup_fast_vs_plain": round(res["fast"]["fps"] / res["plain"]["fps"], 3)}


def large_leg(args, tuning):
    """The large high-frequency workload: 3M flat Gaussians, 4096 nodes, K=8,
    128^3 density grid, the combined bend + twist + 1.3x scale edit spread over a
    30-frame drag (per frame: translate, twist about the active capsule's axis,
    scale about the active centroid), up to 12 Gauss-Newton iterations with the
    stop rule, 8 views at 1920x1080. Pipelined drag frames (p50/p99), the
    deformation kernels' HBM rate from a few sequential frames, and the
    density-preserving step of the release (density sample over the 128^3 cell
    lists + realignment steps)."""
    import math
    import numpy as np
    import torch
    load_package()
    from arapgs_amd import arapgs, scenes, workload
    from arapgs_amd.pipeline import DragPipeline
    sc = scenes.flat_scene(args.seed, n=args.large_n, n_nodes=args.large_nodes)
    ctx = arapgs.Context(0, stream=torch.cuda.Stream())
    t0 = time.perf_counter()
    with torch.cuda.stream(ctx.stream):
        W = workload.prepare(ctx, sc, log=log)
        solver = workload.make_solver(ctx, W["solver_args"], n_aggregates=192)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    st, K = W["state"], W["K"]
    e = sc.edit
    frames_per_edit = 30
    delta = np.float32(e["distance"] / frames_per_edit) * e["direction"].astype(np.float32)
    radian = math.radians(e["angle_deg"]) / frames_per_edit
    kscale = 1.0 + (e["scale"] - 1.0) / frames_per_edit
    active = W["active"]
    with torch.cuda.stream(ctx.stream):
        tmp = [torch.empty_like(st["node_pos"]) for _ in range(2)]

    def aims_fn(c, node_pos, aims):
        # UpdateAimPosition, then UpdateAimPositionTwist, then UpdateAimPositionScale
        c.aims_translate(node_pos, active, delta, tmp[0])
        c.aims_twist(tmp[0], active, e["axis"], radian, tmp[1])
        c.aims_scale(tmp[1], active, kscale, aims)

    images = [torch.empty((3, sc.cameras[0].H, sc.cameras[0].W), dtype=torch.float32, device="cuda:0")
              for _ in W["cams"]]
    solver.set_coarse_refresh(8)
    solver.set_host_checks(4)
    solver.set_warm_start(1)
    solver.set_cg_graphs(tuning.get("cg_graphs", 0))
    solver.set_cg_mode(tuning.get("cg_mode", 0))
    solver.set_cg_predict(tuning.get("cg_predict", 0))
    solver.set_host_spin(tuning.get("host_spin", 0))
    solver.set_coarse_lag(tuning.get("coarse_lag", 1))
    solver.set_cg_assembled(tuning.get("cg_asm", 0))
    s0 = workload.clone_state(st)
    # a few sequential frames with the kernels timed by HIP events
    fb = arapgs.frame_buffers(st, K)
    aims = torch.empty_like(st["node_pos"])
    ctx.profile_enable(True)
    with torch.cuda.stream(ctx.stream):
        for _ in range(3):
            aims_fn(ctx, st["node_pos"], aims)
            solver.deform_frame(fb, aims, sc.gn_iters, stop_rule=True)
    torch.cuda.synchronize()
    kms = {name: ctx.profile_read(which=which) for name, _, which in KERNELS}
    ctx.profile_enable(False)
    S = int(st["samples"].shape[0])
    N = int(st["pos"].shape[0])
    per = {"six_point": (SIX_POINT_BYTES_PER_GAUSSIAN, N), "sample_blend": (SAMPLE_BLEND_BYTES_PER_SAMPLE, S),
           "sample_sh": (SAMPLE_SH_BYTES_PER_SAMPLE, S)}
    kern = {}
    for name, (b, n) in per.items():
        v = kms[name]
        avg = float(sum(v) / len(v)) if len(v) else float("nan")
        kern[name] = {"avg_ms": round(avg, 3), "gbs": round(b * n / (avg * 1e-3) / 1e9, 1) if len(v) else None}
    # pipelined drag
    for k, v in s0.items():
        if v is not None:
            st[k].copy_(v)
    torch.cuda.synchronize()
    # the drag workload's tuned schedule (identical results)
    base_stream = ctx.stream
    masked, stream_b, stream_c, stream_r = partition_streams(ctx, tuning.get("solve_cus", 0),
                                                             tuning.get("deform_skip", 0),
                                                   tuning.get("samples_cus", 0),
                                                   side=solver if tuning.get("side_cus", 0) else None)
    pipe = DragPipeline(solver, st, K, active, delta, sc.gn_iters, W["cams"], images, W["bg"], W["opac"],
                        aims_fn=aims_fn, sample_batch=tuning.get("sample_batch", 1),
                        node_cache=bool(tuning.get("node_cache", 1)), stream_b=stream_b, stream_c=stream_c,
                        stream_r=stream_r, split_render=bool(tuning.get("split_render", 0)),
                        samples_on_deform=tuning.get("samples_cus", 0) == 3,
                        sh_compose=bool(tuning.get("sh_compose", 0)),
                        samples_late=bool(tuning.get("samples_late", 0)))
    pipe.ctx_render.set_blend_unroll(tuning.get("blend_unroll", 4))
    pipe.ctx_render.set_blend_split(tuning.get("blend_split", 0))
    pipe.ctx_render.set_tile_cull(tuning.get("tile_cull", 0))
    pipe.run(args.warmup)
    torch.cuda.synchronize()
    stamps = []
    t0 = time.perf_counter()
    pipe.run(args.large_frames, on_frame=lambda t, n: stamps.append(time.perf_counter()))
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    ms = [(b - a) * 1e3 for a, b in zip([t0] + stamps[:-1], stamps)]
    inst = pipe.instances
    pipe.close()
    torch.cuda.synchronize()
    ctx.set_stream(base_stream)
    solver.set_side_stream(None)
    for s_ in masked:
        s_.close()
    gn, cg = solver.last_stats()
    frame_bytes = sum(b * n for b, n in per.values())
    # release: density sample over the 128^3 lists and realignment steps (high_quality)
    grid = W["grid"]
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    with torch.cuda.stream(ctx.stream):
        ev[0].record(ctx.stream)
        feat, dens = grid.density(st["pos"], st["rot"], st["scale"], W["opac"], st["sh"], W["lpf"])
        ev[1].record(ctx.stream)
    torch.cuda.synchronize()
    dens_ms = ev[0].elapsed_time(ev[1])
    with torch.cuda.stream(ctx.stream):
        pos, rot, shv = st["pos"].clone(), st["rot"].clone(), st["sh"].clone()
        log_scale = torch.log(st["scale"])
        o = W["opac"].clamp(1e-6, 1 - 1e-6)
        logit = torch.log(o / (1 - o))
        _, aim_dens = grid.density(s0["pos"], s0["rot"], s0["scale"], W["opac"], s0["sh"], W["lpf"])
        r = arapgs.Realign(ctx, grid, N)
        loss = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    rms, losses = [], []
    for i in range(args.large_realign_steps):
        with torch.cuda.stream(ctx.stream):
            ev[0].record(ctx.stream)
            r.step(st["samples"], st["sample_feat"], aim_dens, W["lpf"], pos, rot, log_scale, logit, shv, loss=loss)
            ev[1].record(ctx.stream)
        torch.cuda.synchronize()
        rms.append(ev[0].elapsed_time(ev[1]))
        losses.append(float(loss.item()))
    r.close()
    n_pairs, n_valid = int(grid.n_pairs), int(grid.n_valid)
    del feat, dens
    solver.close()
    grid.close()
    ctx.close()
    fps = args.large_frames / total
    out = {"workload": "flat-3M (128^3 grid, 8 views at 1920x1080)", "gaussians": N, "nodes": args.large_nodes,
           "samples": S, "cell_gaussian_pairs": n_pairs, "valid_cells": n_valid, "views": len(images),
           "resolution": [sc.cameras[0].W, sc.cameras[0].H], "gn_iters_max": sc.gn_iters,
           "form": "pipelined", "frames": args.large_frames, "fps": round(fps, 3),
           "p50_ms": round(pct(ms, 50), 3), "p99_ms": round(pct(ms, 99), 3), "last_frame_gn_cg": [gn, cg],
           "tile_instances_per_frame": int(inst), "setup_s": round(setup_s, 2),
           "deform_kernels": kern, "deform_bytes_per_frame": frame_bytes,
           "deform_hbm_gbs_at_frame_rate": round(frame_bytes * fps / 1e9, 1),
           "release": {"density_sample_ms": round(dens_ms, 2),
                       "realign_step_p50_ms": round(pct(rms, 50), 2) if rms else None,
                       "l1_first_last": [losses[0], losses[-1]] if losses else None}}
    log(f"  large: {out['fps']} frames/s, p50 {out['p50_ms']} ms, p99 {out['p99_ms']} ms; density "
        f"{out['release']['density_sample_ms']} ms, realign {out['release']['realign_step_p50_ms']} ms/step")
    return out


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
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg (and the energy check)")
    ap.add_argument("--energy-check-frames", type=int, default=-1,
                    help="timed frames whose energy the CPU leg checks against the oracle (-1: all)")
    ap.add_argument("--fixed-steps", type=int, default=10,
                    help="timed frames per form of the fixed-iteration leg (0: skip)")
    ap.add_argument("--large-frames", type=int, default=20, help="timed frames of the large-scene leg (0: skip)")
    ap.add_argument("--large-n", type=int, default=3_000_000)
    ap.add_argument("--large-nodes", type=int, default=4096)
    ap.add_argument("--large-realign-steps", type=int, default=3)
    ap.add_argument("--forms", default="plain,fast,pipelined")
    ap.add_argument("--adjoint-steps", type=int, default=10, help="steps of the differentiation leg (0: skip)")
    ap.add_argument("--aggregates", type=int, default=0, help="coarse aggregates of the CG preconditioner (0: 128)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip the autotuner (tile, blend unroll, aggregates, coarse refresh, CG host checks)")
    ap.add_argument("--tune-steps", type=int, default=0,
                    help="timed frames per autotuner candidate (0: the timed drag itself, --warmup + --steps)")
    ap.add_argument("--sample-batch", type=int, default=0,
                    help="sample chunks of the rolling sample update in the pipelined form (0: autotuned, or 4)")
    ap.add_argument("--cg-graphs", type=int, default=-1, choices=(-1, 0, 1),
                    help="CG chunks as captured HIP graphs (-1: autotuned, or off without the autotuner)")
    ap.add_argument("--cg-mode", type=int, default=-1, choices=(-1, 0, 2),
                    help="CG iterations as three kernels (0) or two (2) (-1: autotuned, or 0 without the autotuner)")
    ap.add_argument("--cg-block", type=int, default=0, choices=(0, 64, 128, 256, 512),
                    help="CG block size (0: autotuned, or 256 without the autotuner)")
    ap.add_argument("--host-checks", type=int, default=0,
                    help="CG iterations between host checks of the fast and pipelined forms (0: autotuned, or 16)")
    ap.add_argument("--cg-predict", type=int, default=-1, choices=(-1, 0, 1),
                    help="first CG host check at the previous frame's count (-1: autotuned, or off)")
    ap.add_argument("--solve-cus", type=int, default=-1,
                    help="pipelined form: the solve on CUs [0, N), the rest for the deform and samples (-1: autotuned, "
                         "or 0 = all CUs shared)")
    ap.add_argument("--samples-cus", type=int, default=-1, choices=(-1, 0, 1, 2, 3),
                    help="with --solve-cus: the samples stream beside the deform (0), on every CU (1), beside the "
                         "solve (2), or the samples queued on the deform stream itself (3) (-1: autotuned, or 0)")
    ap.add_argument("--sh-compose", type=int, default=-1, choices=(-1, 0, 1),
                    help="rolling sample chunks: one SH rotation by the product of the pending frames' rotations "
                         "(-1: autotuned, or 0)")
    ap.add_argument("--tile-cull", type=int, default=-1, choices=(-1, 0, 1),
                    help="renders emit only the tiles passing the conservative tile test (-1: autotuned, or 0)")
    ap.add_argument("--refresh", type=int, default=0,
                    help="solves between coarse-operator rebuilds (0: autotuned, or 8)")
    ap.add_argument("--side-cus", type=int, default=-1, choices=(-1, 0, 1),
                    help="with --solve-cus: the asynchronous coarse rebuilds on the deform side's CUs (1) or on "
                         "every CU (0) (-1: autotuned, or 0)")
    ap.add_argument("--host-spin", type=int, default=-1, choices=(-1, 0, 1),
                    help="host checks poll the pinned flag (1) or wait for the stream (0) (-1: autotuned, or 0)")
    ap.add_argument("--cg-tol", type=float, default=0.0,
                    help="CG relative residual tolerance of the fast and pipelined forms (0: 1e-4)")
    ap.add_argument("--split-render", type=int, default=-1, choices=(-1, 0, 1),
                    help="renders on their own stream from double-buffered attributes (-1: autotuned, or 0)")
    ap.add_argument("--cg-asm", type=int, default=-1, choices=(-1, 0, 1),
                    help="the CG operator applied from assembled pair blocks (-1: autotuned, or 0)")
    ap.add_argument("--samples-late", type=int, default=-1, choices=(-1, 0, 1),
                    help="the samples stream starts after the six-point update (-1: autotuned, or 0)")
    ap.add_argument("--coarse-lag", type=int, default=0,
                    help="solves between a side-stream coarse rebuild's start and its adoption (0: autotuned, or 1)")
    ap.add_argument("--partition", action="store_true",
                    help="sweep CU partitions between the solve and deform streams of the pipelined form")
    ap.add_argument("--adjoint-n", type=int, default=500_000)
    ap.add_argument("--realign-steps", type=int, default=5, help="steps of the realignment leg (0: skip)")
    ap.add_argument("--adjoint-nodes", type=int, default=1024)
    ap.add_argument("--batch-scenes", type=int, default=32, help="scenes of the uneven-batch leg (0: skip)")
    ap.add_argument("--batch-frames", type=int, default=64, help="timed frames per form of the uneven-batch leg")
    ap.add_argument("--copy-rate", type=int, default=0, metavar="MIB",
                    help="only time 20 plain device copies of MIB MiB (read + write GB/s) and exit")
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
    if args.copy_rate > 0:
        # the practical HBM rate of a plain stream (read + write) for comparison with the roofline kernels
        n = args.copy_rate * (1 << 20) // 4
        a = torch.rand(n, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(20):
            b.copy_(a)
        ev[1].record()
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / 20
        print(json.dumps({"copy_bytes_moved": 8 * n, "copy_ms": round(ms, 4),
                          "copy_gbs": round(8 * n / (ms * 1e-3) / 1e9, 1)}), flush=True)
        return
    B = setup(args)
    apply_config(B, B["cfg"])
    tuning = None
    if args.autotune:
        log(f"autotuning on the pipelined form ({tune_frames(args)[0]} frames per candidate)")
        best, trials = autotune_leg(B, args)
        tuning = {"config": best, "candidates_timed": len(trials),
                  "ms_per_frame": {json.dumps(c, sort_keys=True): round(t * 1e3, 3) for c, t in trials}}
        log(f"  autotune: {best}")
    if args.cg_graphs >= 0:
        apply_config(B, dict(B["cfg"], cg_graphs=args.cg_graphs))
    if args.cg_mode >= 0:
        apply_config(B, dict(B["cfg"], cg_mode=args.cg_mode))
    if args.cg_block > 0:
        apply_config(B, dict(B["cfg"], cg_block=args.cg_block))
    if args.cg_predict >= 0:
        apply_config(B, dict(B["cfg"], cg_predict=args.cg_predict))
    if args.host_checks > 0:
        apply_config(B, dict(B["cfg"], host_checks=args.host_checks))
    if args.solve_cus >= 0:
        apply_config(B, dict(B["cfg"], solve_cus=args.solve_cus))
    if args.samples_cus >= 0:
        apply_config(B, dict(B["cfg"], samples_cus=args.samples_cus))
    if args.sh_compose >= 0:
        apply_config(B, dict(B["cfg"], sh_compose=args.sh_compose))
    if args.tile_cull >= 0:
        apply_config(B, dict(B["cfg"], tile_cull=args.tile_cull))
    if args.coarse_lag > 0:
        apply_config(B, dict(B["cfg"], coarse_lag=args.coarse_lag))
    if args.refresh > 0:
        apply_config(B, dict(B["cfg"], refresh=args.refresh))
    if args.side_cus >= 0:
        apply_config(B, dict(B["cfg"], side_cus=args.side_cus))
    if args.samples_late >= 0:
        apply_config(B, dict(B["cfg"], samples_late=args.samples_late))
    if args.cg_asm >= 0:
        apply_config(B, dict(B["cfg"], cg_asm=args.cg_asm))
    if args.split_render >= 0:
        apply_config(B, dict(B["cfg"], split_render=args.split_render))
    if args.cg_tol > 0:
        apply_config(B, dict(B["cfg"], cg_tol=args.cg_tol))
    if args.host_spin >= 0:
        apply_config(B, dict(B["cfg"], host_spin=args.host_spin))
    res = {}
    for form in args.forms.split(","):
        log(f"timing form={form}: {args.warmup} warmup + {args.steps} steps")
        if form == "pipelined":
            if args.sample_batch > 0:
                B["cfg"]["sample_batch"] = args.sample_batch
            elif args.autotune:
                B["cfg"]["sample_batch"], sb_trials = sample_batch_leg(B, args)
                log(f"  sample batch: {B['cfg']['sample_batch']} rolling sample chunks")
                if tuning is not None:
                    tuning["sample_batch_ms_per_frame"] = sb_trials
                    tuning["config"] = dict(B["cfg"])
            else:
                B["cfg"]["sample_batch"] = 4
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
    # the deformation kernels are timed (HIP events on their stream) in the sequential forms
    kr = res.get("fast") or res.get("plain") or r
    roofs = rooflines(args, B, kr["kernel_ms"])
    dominant = max(roofs, key=lambda k: roofs[k]["avg_launch_ms"] if roofs[k]["launches"] else -1.0)
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
                   "grid": 64, "gn_iters": 10, "stop_rule": True, "views": args.views,
                   "resolution": [args.width, args.height], "sh_degree": 3, "form": best,
                   "drag": f"0.3 over {args.drag_frames} frames"},
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
    line["roofline"] = dict(roofs[dominant], dominant=True)
    line["rooflines"] = roofs
    import gc
    gc.enable()
    line["config"]["tuning"] = B["cfg"]
    if tuning:
        line["autotune"] = tuning
    rec = None
    if "pipelined" in res and not args.no_cpu:
        log("re-running the timed pipelined frames with their solve inputs recorded")
        rec = record_frames(B, args.steps, args.warmup)
        log(f"  recorded frames equal the timed frames: {rec['same_as_timed']}")
    if args.realign_steps > 0:
        log(f"timing the realignment leg: {args.realign_steps} steps")
        line["realign"] = realign_leg(B, args)
        log(f"  realign: {line['realign']['step_p50_ms']} ms/step, L1 {line['realign']['l1_loss_first_last']}")
    if args.fixed_steps > 0:
        log(f"timing the fixed-iteration leg: {args.warmup} warmup + {args.fixed_steps} frames per form")
        line["fixed_iteration"] = fixed_iteration_leg(B, args)
    if rec is not None:
        line["cpu_baseline"] = cpu_baseline(args, B, rec)
    B["ctx"].close()
    del B
    torch.cuda.empty_cache()
    if args.adjoint_steps > 0:
        log(f"timing the differentiation leg: {args.warmup} warmup + {args.adjoint_steps} steps")
        line["adjoint"] = adjoint_leg(args)
        log(f"  adjoint: {line['adjoint']['steps_per_s']} steps/s, forward p50 {line['adjoint']['forward_p50_ms']} ms, "
            f"backward p50 {line['adjoint']['backward_p50_ms']} ms")
        torch.cuda.empty_cache()
    if args.batch_scenes > 0:
        log(f"timing the uneven-batch leg: {args.batch_scenes} scenes")
        line["batch"] = batch_leg(args)
        torch.cuda.empty_cache()
    if args.large_frames > 0:
        log(f"timing the large-scene leg: {args.warmup} warmup + {args.large_frames} frames")
        line["large"] = large_leg(args, line["config"]["tuning"])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
