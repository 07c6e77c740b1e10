"""Drag frames of many independent scenes at once.

The reference edits one scene: per drag frame GaussianView::onUpdate
(GaussianView.cpp:1481-1522) runs aims, the Gauss-Newton solve, putFreeInputs,
the sample / end-point / node blends, the six-point update, the sample SH
rotation and the renders. A batch of scenes of uneven size stores every scene
back to back (workload.prepare_batch: one node list, node indices offset, so
one node table serves all scenes) and runs that frame for all of them with one
launch per step:

  aims            one translate per scene (each scene has its own drag step)
  solve           arapgs_solver_batch_solve: every Gauss-Newton phase one launch
                  over all scenes (scene = blockIdx.z), node tables included
  node blend      one launch over all nodes
  six-point       one launch over all Gaussians
  samples         one launch each for blend, node quaternions, SH rotation
  render          arapgs_render_segments: one view per scene, one pass

frame_looped() runs the same frame one scene after another with the
single-scene calls (what a caller without batching would do); both give the
same state and images bit for bit (GPU test).
"""
from __future__ import annotations

import numpy as np

from . import arapgs
from . import workload


class SceneBatch:
    def __init__(self, ctx, wl: dict, gn_iters: int, warm_start=True, stop_rule=True, host_checks=8,
                 coarse_refresh=8, **solver_kw):
        import torch
        self.torch = torch
        self.ctx = ctx
        self.wl = wl
        self.st = wl["state"]
        self.iters, self.stop_rule = gn_iters, stop_rule
        self.n = len(wl["solver_args"])
        self.solvers = [workload.make_solver(ctx, a, **solver_kw) for a in wl["solver_args"]]
        for s in self.solvers:
            s.set_warm_start(warm_start)
            s.set_host_checks(host_checks)
            s.set_coarse_refresh(coarse_refresh)
        self.batch = arapgs.SolverBatch(ctx, self.solvers)
        self.batch.set_host_checks(host_checks)
        self.batch.set_coarse_refresh(coarse_refresh)
        dev = self.st["pos"].device
        ns, ss, gs = wl["node_seg"], wl["sample_seg"], wl["seg"]
        self.ns, self.ss, self.gs = ns, ss, gs
        M = ns[-1]
        self.table = arapgs.node_table(M, dev)
        self.aims = torch.empty_like(self.st["node_pos"])
        # per-scene views into the concatenated buffers
        self.node_v = [self.st["node_pos"][ns[i]:ns[i + 1]] for i in range(self.n)]
        self.aims_v = [self.aims[ns[i]:ns[i + 1]] for i in range(self.n)]
        self.table_v = [self.table[ns[i]:ns[i + 1]] for i in range(self.n)]
        self.x = [torch.empty(s.n_unknowns, dtype=torch.float32, device=dev) for s in self.solvers]
        self.energies = [torch.zeros(gn_iters + 1, dtype=torch.float64, device=dev) for _ in range(self.n)]
        self.iters_done = [torch.zeros(1, dtype=torch.int32, device=dev) for _ in range(self.n)]
        cam0 = wl["cams"][0]
        self.W, self.H = cam0["W"], cam0["H"]
        self.images = [torch.zeros((3, self.H, self.W), dtype=torch.float32, device=dev) for _ in range(self.n)]
        self.instances = 0

    def close(self):
        self.batch.close()
        for s in self.solvers:
            s.close()

    # ------------------------------------------------------------ batched frame
    def frame(self, deltas, render=True):
        """One drag frame of every scene; deltas (n, 3): each scene's translation
        of its active block this frame (UpdateAimPosition)."""
        st, ctx = self.st, self.ctx
        for i in range(self.n):
            ctx.aims_translate(self.node_v[i], self.wl["active"][i], deltas[i], self.aims_v[i])
        self.batch.solve(self.node_v, self.aims_v, self.iters, self.stop_rule, x_out=self.x, energies=self.energies,
                         iters_done=self.iters_done, tables=self.table_v)
        ctx.blend_points(st["node_pos"], st["node_src_idx"], st["node_src_w"], self.table, out=st["node_pos"])
        self._deform(ctx, self.table, slice(None), slice(None), slice(None))
        if render:
            self.instances = ctx.render_segments(st["pos"], st["sh"], self.wl["opac"], st["scale"], st["rot"],
                                                 self.wl["cams"], self.W, self.H, self.wl["bg"], self.images,
                                                 self.gs)
        return self.instances

    def _deform(self, ctx, table, gsl, esl, ssl):
        """Six-point update of Gaussians gsl (end-point rows esl) and sample
        update of samples ssl with the node table (node ids index the full
        table)."""
        st = self.st
        ctx.six_point_update(table, st["ends_idx"][esl], st["ends_w"][esl], st["ends"][gsl], st["scale0"][gsl], None,
                             st["pos"][gsl], st["rot"][gsl], st["scale"][gsl], st["sh"][gsl])
        samples = st["samples"][ssl]
        ctx.blend_points(samples, st["sample_idx"][ssl], st["sample_w"][ssl], table, out=samples)
        q = ctx.node_quats(table)
        ctx.sample_sh(st["sample_w"][ssl], st["sample_idx"][ssl], q, None, st["sample_feat"][ssl])

    # ------------------------------------------------------------ one scene at a time
    def frame_looped(self, deltas, render=True):
        """The same frame with single-scene calls, scene after scene."""
        st, ctx = self.st, self.ctx
        total = 0
        for i in range(self.n):
            nv, av = self.node_v[i], self.aims_v[i]
            ctx.aims_translate(nv, self.wl["active"][i], deltas[i], av)
            s = self.solvers[i]
            s.solve(nv, av, self.iters, self.stop_rule, x_out=self.x[i], energies=self.energies[i],
                    iters_done=self.iters_done[i])
            s.put_free_inputs(nv, self.x[i], table=self.table_v[i])
            nsl = slice(self.ns[i], self.ns[i + 1])
            ctx.blend_points(st["node_pos"][nsl], st["node_src_idx"][nsl], st["node_src_w"][nsl], self.table,
                             out=st["node_pos"][nsl])
            gsl = slice(self.gs[i], self.gs[i + 1])
            self._deform(ctx, self.table, gsl, slice(6 * self.gs[i], 6 * self.gs[i + 1]),
                         slice(self.ss[i], self.ss[i + 1]))
            if render:
                total += ctx.render_views(st["pos"][gsl], st["sh"][gsl], self.wl["opac"][gsl], st["scale"][gsl],
                                          st["rot"][gsl], [self.wl["cams"][i]], self.W, self.H, self.wl["bg"],
                                          [self.images[i]], fast_culling=False)
        self.instances = total
        return total


def random_walk_steps(seed: int, n_scenes: int, n_frames: int, step: float = 0.01) -> np.ndarray:
    """(n_frames, n_scenes, 3) per-frame translations: a seeded random walk of
    step length `step` per scene (a new uniform direction every frame)."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n_frames, n_scenes, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    return (step * d).astype(np.float32)
