"""Autotuner of the drag-frame path.

The tunables change only speed, never the images or the solve's result beyond
its tolerance:
  tile          render tile edge: 16 (the 3DGS tile) or 8
  blend_unroll  Gaussians per compositing step of the 16x16 blend: 2, 4, 8
  aggregates    node aggregates of the CG preconditioner's coarse space
  refresh       solves between rebuilds of the coarse operator
  coarse_lag    solves between a side-stream rebuild's start and its adoption
  side_cus      with solve_cus: the side-stream rebuilds on the deform side's
                CUs (1) or on every CU (0)
  samples_late  the samples stream starts a frame after its six-point update (1)
                or beside it (0)
  cg_asm        the CG operator applied from per-frame assembled 4x4 pair blocks
                (1) or matrix-free (0)
  split_render  renders on their own stream from double-buffered attributes, so
                a frame's six-point update overlaps the previous frame's render (1)
  host_checks   CG iterations between host reads of the convergence flag
  warm          CG warm start from the previous frame's GN steps (1) or from 0
  cg_block      workgroup size of the CG operator kernel: 256, 128, 64 or 512 threads
  cg_graphs     CG chunks between host checks as captured HIP graphs, checked
                one chunk behind (1) or launched kernel by kernel (0)
  cg_mode       CG iterations as three kernels (0) or two, with the coarse
                restriction carried by a recurrence (2)
  cg_predict    first CG host check of each GN iteration after as many CG
                iterations as the previous frame needed there, and the GN stop
                flag read only from the previous frame's GN count on (1), or
                after host_checks and after every GN iteration (0)
  host_spin     host checks poll the flag's pinned slot (1) or wait for the
                solve stream (0)
  node_cache    K = 8 blends through the per-workgroup LDS node-row cache (1) or
                with every node row gathered from the global table (0)
  blend_split   16x16 tiles with more instances than this blended in segments by
                separate workgroups and composited (0: one workgroup per tile)
  deform_skip   pipelined frames: CUs [0, deform_skip) left to the solve by the
                deform and samples streams (CU-masked streams; 0: all CUs shared)
  solve_cus     pipelined frames: the solve on CUs [0, solve_cus) and the deform
                stream on the others (0: all CUs shared)
  samples_cus   with solve_cus: the samples stream beside the deform (0), on
                every CU (1), beside the solve (2), or the samples queued on the
                deform stream itself (3)
  sh_compose    rolling sample chunks: features rotated once by the product of
                the pending frames' rotations (1) or frame by frame (0)
  tile_cull     renders emit only the tiles that pass the conservative tile test
                (1; identical images) or every tile of the rectangle (0)

`tune` runs a coordinate descent from the defaults: for each parameter in turn
it times every candidate value (the others held at their current best) with a
caller-supplied `measure(config) -> seconds per frame`, and keeps the fastest.
Results are cached per workload key in a JSON file so a session tunes once.
"""
from __future__ import annotations

import json
import os

# the defaults are the best configuration measured on the capsules-1M drag
DEFAULTS = dict(tile=16, blend_unroll=4, aggregates=128, refresh=4, host_checks=3, warm=1, cg_block=256, cg_graphs=0,
                cg_mode=2, cg_predict=1, node_cache=1, blend_split=0, deform_skip=0,
                solve_cus=32, samples_cus=3, sh_compose=1, tile_cull=0,
                coarse_lag=1, side_cus=1, samples_late=1, cg_asm=0,
                split_render=1, host_spin=1)
CANDIDATES = dict(tile=(16, 8), blend_unroll=(2, 4, 8), aggregates=(64, 128, 192, 256), refresh=(2, 4, 8, 16, 32),
                  host_checks=(2, 3, 4, 6, 8, 16, 32), warm=(1, 0), cg_block=(256, 128, 64, 512), cg_graphs=(0, 1),
                  cg_mode=(0, 2), cg_predict=(0, 1), node_cache=(1, 0), blend_split=(0, 1024, 2048, 4096),
                  deform_skip=(0, 16, 32, 64), solve_cus=(0, 32, 48, 64),
                  samples_cus=(0, 1, 3), sh_compose=(0, 1), tile_cull=(0, 1),
                  coarse_lag=(1, 2, 3), side_cus=(0, 1), samples_late=(0, 1), cg_asm=(0, 1),
                  split_render=(0, 1), host_spin=(0, 1))
# host_checks is revisited after cg_graphs: their best chunk can differ
ORDER = ("warm", "split_render", "solve_cus", "samples_cus", "side_cus", "samples_late", "cg_mode", "cg_asm",
         "cg_predict", "host_spin", "aggregates", "refresh", "coarse_lag", "host_checks", "cg_graphs", "host_checks", "cg_block",
         "tile", "tile_cull", "blend_unroll", "blend_split", "node_cache", "sh_compose", "deform_skip")


def tune(measure, key=None, cache_path=None, candidates=None, order=ORDER, log=None, rounds=1):
    """Coordinate descent over `candidates` (default CANDIDATES). Returns
    (best_config, trials) where trials lists (config, seconds) in the order run."""
    cands = dict(CANDIDATES if candidates is None else candidates)
    if cache_path and key and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache:
            if log:
                log(f"autotune: cached configuration for {key}: {cache[key]}")
            return dict(cache[key]), []
    best = {k: DEFAULTS[k] for k in cands}
    trials = []
    memo = {}

    def run(cfg):
        k = tuple(sorted(cfg.items()))
        if k not in memo:
            memo[k] = float(measure(dict(cfg)))
            trials.append((dict(cfg), memo[k]))
            if log:
                log(f"autotune: {cfg} -> {memo[k] * 1e3:.3f} ms/frame")
        return memo[k]

    best_t = run(best)
    for _ in range(rounds):
        for name in order:
            if name not in cands:
                continue
            for v in cands[name]:
                if v == best[name]:
                    continue
                cfg = dict(best)
                cfg[name] = v
                t = run(cfg)
                if t < best_t:
                    best, best_t = cfg, t
    if cache_path and key:
        cache = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                cache = json.load(f)
        cache[key] = best
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return best, trials
