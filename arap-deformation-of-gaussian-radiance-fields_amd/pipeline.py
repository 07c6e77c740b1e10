"""Pipelined drag frames over the arapgs C ABI.

One drag frame of GaussianView::onUpdate (GaussianView.cpp:1481-1522) is, in
order: aims; Gauss-Newton solve; node table (putFreeInputs); sample blend
(UpdatePositionforSamples); end-point blend + six-point update
(UpdatePosition, UpdateAsSixPointsWithdrawBad); node blend; node quaternions +
sample SH rotation (FastUpdateSamplesSH); renders.

The only thing frame t+1 needs from frame t is the blended node positions. So
the frame splits into two halves on two streams:

  solve side   (stream A)  aims(t), solve(t), node table(t) -> slot t%2,
                           node blend(t) (node positions for t+1); event A[t%2]
  deform side  (stream B)  wait A[t%2]; six-point update with table slot t%2;
                           event B[t%2]; the V views in one render pass
               (stream C)  wait A[t%2]; sample blend, node quaternions, sample
                           SH with table slot t%2; event C[t%2]
                           (the slot is free again after B[t%2] and C[t%2])

and frame t's deform side runs while frame t+1's solve runs; the samples (they
read only the node table) run beside the six-point update and the renders.

Rolling sample updates (sample_batch = B > 1): nothing reads the samples during
a drag (they feed the density-preserving step after the release), so the
samples are split into B chunks and frame t updates only chunk t mod B, with
the node tables of every frame since that chunk's last update (at most B) in
one pass (arapgs_samples_frames: frame after frame in order, bit-identical to
per-frame sample updates). Every frame then does the same amount of sample
work as a per-frame update, but reads and writes only 1/B of the sample
streams; B + 2 table slots keep each table alive until the last chunk has
applied it, and run() brings every chunk up to date before it returns. Each
side has its
own host thread, because each side synchronises with the host (the solve's CG
convergence checks, the renderer's instance count) and those waits must not
stall the other side. The two node-table slots bound the lead of the solve side
to one frame. Every kernel sees exactly the inputs of the sequential frame, so
the results are identical bit for bit (GPU test).
"""
from __future__ import annotations

import queue
import threading

from . import arapgs


class DragPipeline:
    def __init__(self, solver, state: dict, k: int, active, delta, gn_iters: int, cameras, images, bg, opacity,
                 stop_rule=True, render_ctx=None, sh_degree=3, tile=16, stream_b=None, split_samples=True,
                 aims_fn=None, sample_batch=1, node_cache=True, stream_c=None, sh_compose=False,
                 samples_late=False, split_render=False, stream_r=None, samples_on_deform=False):
        """The solve side runs on solver.ctx's stream, the deform side on
        stream_b (a new stream when None), e.g. two arapgs.CuStream partitions.
        split_samples: the samples' blend and SH rotation run on a third stream
        beside the six-point update and the renders (they read only the node
        table), so they fill the CUs the renders' deep tiles leave idle.
        aims_fn(ctx, node_pos, aims): the frame's aims on the solve stream
        (default: the active nodes translated by delta, UpdateAimPosition).
        sample_batch: B, the sample chunks of the rolling update (1..8; more
        than one needs split_samples, K = 8 and samples with features).
        node_cache: the deform side's blends through the LDS node-row cache
        (arapgs_set_node_cache; identical results). stream_c: the samples'
        stream (a new stream when None). sh_compose: a chunk's pending frames
        rotate its features once, by the product of their rotations
        (arapgs_set_sh_compose; features within float rounding of per-frame
        rotations, positions bit-identical). samples_late: the samples stream
        starts a frame's work after that frame's six-point update instead of
        beside it, so the two HBM-heavy passes do not overlap (same results).
        split_render: the renders run on a fourth stream (stream_r, or a new
        one) from double-buffered Gaussian attributes, so frame t+1's six-point
        update (reading frame t's set, writing the other) overlaps frame t's
        render; run() leaves the final attributes in `state` (same results).
        ctx_render is the context that renders (set its render options).
        samples_on_deform: the samples' work is queued on the deform stream
        instead of a stream of its own (fewer streams than hardware queues)."""
        import torch
        self.torch = torch
        self.solver = solver
        self.ctx_a = solver.ctx
        dev = state["pos"].device
        self.stream_a = self.ctx_a.stream
        self.stream_b = stream_b if stream_b is not None else torch.cuda.Stream(device=dev)
        self.ctx_b = render_ctx if render_ctx is not None else arapgs.Context(dev.index or 0, stream=self.stream_b)
        if render_ctx is not None:
            self.stream_b = render_ctx.stream
        self.split_render = split_render
        if split_render:
            self.stream_r = stream_r if stream_r is not None else torch.cuda.Stream(device=dev)
            self.ctx_r = arapgs.Context(dev.index or 0, stream=self.stream_r)
            self.ctx_r.set_node_cache(node_cache)
            with torch.cuda.stream(self.stream_b):
                self.alt = {k: state[k].clone() for k in ("pos", "rot", "scale", "sh")}
            self.ev_r = [torch.cuda.Event(), torch.cuda.Event()]
            self.ev_r_recorded = [False, False]
            self.cur = 0  # the set holding the latest attributes: 0 = state, 1 = alt
        else:
            self.stream_r, self.ctx_r = self.stream_b, None
        self.ctx_render = self.ctx_r if split_render else self.ctx_b
        self.split = split_samples and state.get("samples") is not None and state["samples"].shape[0] > 0
        if sample_batch < 1 or sample_batch > 8:
            raise ValueError("sample_batch must be in 1..8")
        self.nb = sample_batch if (self.split and state.get("sample_feat") is not None) else 1
        self.R = 2 if self.nb == 1 else self.nb + 2
        S = state["samples"].shape[0] if self.split else 0
        self.chunks = [(c * S // self.nb, (c + 1) * S // self.nb) for c in range(self.nb)]
        self.applied = [-1] * self.nb  # last frame applied to each chunk
        self.frame = 0                 # frames whose deform side has been queued
        if samples_on_deform:
            stream_c = self.stream_b
        self.stream_c = (stream_c if stream_c is not None else torch.cuda.Stream(device=dev)) if self.split else None
        self.ctx_c = arapgs.Context(dev.index or 0, stream=self.stream_c) if self.split else None
        for c in (self.ctx_b, self.ctx_c):
            if c is not None:
                c.set_node_cache(node_cache)
        if self.ctx_c is not None:
            self.ctx_c.set_sh_compose(sh_compose)
        self.samples_late = samples_late
        self.ev_six = [torch.cuda.Event() for _ in range(self.R)]
        self.ev_c = [torch.cuda.Event() for _ in range(self.R)]
        self.ev_c_recorded = [False] * self.R
        self.st = state
        self.k = k
        self.active, self.delta = active, delta
        self.aims_fn = aims_fn
        self.iters, self.stop_rule = gn_iters, stop_rule
        self.cams, self.images, self.bg, self.opac, self.D = cameras, images, bg, opacity, sh_degree
        self.tile = tile
        M = state["node_pos"].shape[0]
        with torch.cuda.stream(self.stream_a):
            self.aims = torch.empty_like(state["node_pos"])
            self.tables = [arapgs.node_table(M, dev) for _ in range(self.R)]
            self.quats = [torch.empty((M, 4), dtype=torch.float32, device=dev) for _ in range(self.R)] \
                if self.nb > 1 else []
            self.x = [torch.empty(solver.n_unknowns, dtype=torch.float32, device=dev) for _ in range(self.R)]
            self.energies = [torch.zeros(gn_iters + 1, dtype=torch.float64, device=dev) for _ in range(self.R)]
            self.iters_done = [torch.zeros(1, dtype=torch.int32, device=dev) for _ in range(self.R)]
        self.ev_a = [torch.cuda.Event() for _ in range(self.R)]
        self.ev_b = [torch.cuda.Event() for _ in range(self.R)]
        self.ev_b_recorded = [False] * self.R
        self.instances = 0
        torch.cuda.synchronize(dev)

    # ------------------------------------------------------------ the two halves
    def _solve_half(self, slot):
        st, torch = self.st, self.torch
        with torch.cuda.stream(self.stream_a):
            if self.ev_b_recorded[slot]:
                # the slot is free once both deform streams are done with it
                self.stream_a.wait_event(self.ev_b[slot])
            if self.ev_c_recorded[slot]:
                self.stream_a.wait_event(self.ev_c[slot])
            if self.aims_fn is not None:
                self.aims_fn(self.ctx_a, st["node_pos"], self.aims)
            else:
                self.ctx_a.aims_translate(st["node_pos"], self.active, self.delta, self.aims)
            self.solver.solve(st["node_pos"], self.aims, self.iters, self.stop_rule, x_out=self.x[slot],
                              energies=self.energies[slot], iters_done=self.iters_done[slot])
            self.solver.put_free_inputs(st["node_pos"], self.x[slot], table=self.tables[slot])
            self.ctx_a.blend_points(st["node_pos"], st["node_src_idx"], st["node_src_w"], self.tables[slot],
                                    out=st["node_pos"])
            self.ev_a[slot].record(self.stream_a)

    def _samples(self, ctx, table):
        st = self.st
        if st.get("samples") is not None and st["samples"].shape[0] > 0:
            ctx.blend_points(st["samples"], st["sample_idx"], st["sample_w"], table, skip=st.get("sample_static"),
                             out=st["samples"])
            if st.get("sample_feat") is not None:
                q = ctx.node_quats(table)
                ctx.sample_sh(st["sample_w"], st["sample_idx"], q, st.get("sample_static"), st["sample_feat"])

    def _update_chunk(self, c, upto):
        """Chunk c of the samples brought up to frame `upto` (all frames since
        its last update, in order) in one pass on stream C."""
        st = self.st
        frames = list(range(self.applied[c] + 1, upto + 1))
        if not frames:
            return
        b, e = self.chunks[c]
        slots = [t % self.R for t in frames]
        if e > b:
            static = st.get("sample_static")
            self.ctx_c.samples_frames([self.tables[k] for k in slots], [self.quats[k] for k in slots],
                                      st["sample_idx"][b:e], st["sample_w"][b:e],
                                      None if static is None else static[b:e], st["samples"][b:e],
                                      st["sample_feat"][b:e])
        for k in slots:
            self.ev_c[k].record(self.stream_c)
            self.ev_c_recorded[k] = True
        self.applied[c] = upto

    def _rolling_samples(self, slot, t, last):
        """Frame t's node quaternions, then chunk t mod B (every chunk when
        `last`) brought up to frame t."""
        torch = self.torch
        with torch.cuda.stream(self.stream_c):
            self.stream_c.wait_event(self.ev_a[slot])
            if self.samples_late:
                self.stream_c.wait_event(self.ev_six[slot])
            self.ctx_c.node_quats(self.tables[slot], out=self.quats[slot])
            for c in (range(self.nb) if last else [t % self.nb]):
                self._update_chunk(c, t)

    def _deform_half(self, slot, render=True, last=False):
        st, torch = self.st, self.torch
        ctx = self.ctx_b
        table = self.tables[slot]
        t = self.frame
        self.frame += 1

        def samples():
            if self.split and self.nb > 1:
                self._rolling_samples(slot, t, last)
            elif self.split:
                with torch.cuda.stream(self.stream_c):
                    self.stream_c.wait_event(self.ev_a[slot])
                    if self.samples_late:
                        self.stream_c.wait_event(self.ev_six[slot])
                    self._samples(self.ctx_c, table)
                    self.ev_c[slot].record(self.stream_c)
                    self.ev_c_recorded[slot] = True

        if not self.samples_late:
            samples()
        keys = ("pos", "rot", "scale", "sh")
        if self.split_render:
            sets = [{k: st[k] for k in keys}, self.alt]
            src, d = sets[self.cur], 1 - self.cur
            dst = sets[d]
        else:
            dst = {k: st[k] for k in keys}
        with torch.cuda.stream(self.stream_b):
            self.stream_b.wait_event(self.ev_a[slot])
            if not self.split:
                self._samples(ctx, table)
            if self.split_render:
                # the set written here was last read by the render of two frames ago
                if self.ev_r_recorded[d]:
                    self.stream_b.wait_event(self.ev_r[d])
                ctx.six_point_update(table, st["ends_idx"], st["ends_w"], st["ends"], st["scale0"],
                                     st.get("gs_static"), dst["pos"], dst["rot"], dst["scale"], dst["sh"],
                                     src=(src["pos"], src["rot"], src["scale"], src["sh"]))
            else:
                ctx.six_point_update(table, st["ends_idx"], st["ends_w"], st["ends"], st["scale0"],
                                     st.get("gs_static"), st["pos"], st["rot"], st["scale"], st["sh"])
            self.ev_six[slot].record(self.stream_b)
            self.ev_b[slot].record(self.stream_b)
            self.ev_b_recorded[slot] = True
        if self.samples_late:
            samples()
        with torch.cuda.stream(self.stream_r):
            if self.split_render:
                self.stream_r.wait_event(self.ev_six[slot])
            n_inst = 0
            if render and self.images:
                img = self.images[0]
                n_inst = self.ctx_render.render_views(dst["pos"], dst["sh"], self.opac, dst["scale"], dst["rot"],
                                                      self.cams, img.shape[2], img.shape[1], self.bg, self.images,
                                                      D=self.D, tile=self.tile)
            if self.split_render:
                self.ev_r[d].record(self.stream_r)
                self.ev_r_recorded[d] = True
                self.cur = d
            return n_inst

    # ------------------------------------------------------------ driver
    def run(self, n_frames: int, render=True, on_frame=None):
        """Runs n_frames drag frames; on_frame(t, instances) is called on the
        deform thread after frame t's renders were queued (images complete once
        the deform stream reaches that point)."""
        slots = queue.Queue()
        free = threading.Semaphore(2)
        err = []
        base = self.frame  # frames of earlier runs: table slots continue from there

        def solve_side():
            try:
                for t in range(n_frames):
                    free.acquire()
                    if err:
                        return
                    self._solve_half((base + t) % self.R)
                    slots.put((base + t) % self.R)
            except BaseException as e:  # noqa: BLE001 - re-raised on the caller's thread
                err.append(e)
                slots.put(None)

        def deform_side():
            try:
                for t in range(n_frames):
                    slot = slots.get()
                    if slot is None:
                        return
                    n = self._deform_half(slot, render, last=t == n_frames - 1)
                    free.release()
                    self.instances = n
                    if on_frame is not None:
                        on_frame(t, n)
            except BaseException as e:  # noqa: BLE001
                err.append(e)
                for _ in range(n_frames + 2):
                    free.release()

        ta = threading.Thread(target=solve_side, name="arapgs-solve")
        tb = threading.Thread(target=deform_side, name="arapgs-deform")
        ta.start()
        tb.start()
        ta.join()
        tb.join()
        if err:
            raise err[0]
        self.torch.cuda.synchronize(self.st["pos"].device)
        if self.split_render and self.cur == 1:
            # the final attributes back into `state`
            for k, v in self.alt.items():
                self.st[k].copy_(v)
            self.torch.cuda.synchronize(self.st["pos"].device)
            self.cur = 0

    def close(self):
        if self.ctx_c is not None:
            self.ctx_c.close()
            self.ctx_c = None
        if self.ctx_r is not None:
            self.ctx_r.close()
            self.ctx_r = None
        if self.ctx_b is not None:
            self.ctx_b.close()
