"""The benchmarked drag configuration reaches the FP64 oracle's energy.

bench.py times drag frames with the autotuned solver settings: CG relative
tolerance 1e-4 with at most 100 CG iterations (the library defaults), CG warm
start from the previous frame's Gauss-Newton steps, 192 coarse aggregates whose
operator is rebuilt every 8 solves (stale in between), the host reading the CG
flag every 4 iterations, blend unroll 8, pipelined frames, and the drag step of
the capsule workload (0.3 along the edit direction over 30 frames).

Every frame's final ARAP energy must be within 1e-5 relative of the FP64
oracle's Gauss-Newton solve (Deform::optimize, Deform.cpp:95-169) on the same
node positions and aims, with the GN iteration count within one; the last
frame's four 800x800 views must be within 1/255 of the oracle rasteriser; and
the queued sequential form with the same settings gives the pipelined form's
state bit for bit.
"""
import numpy as np
import pytest

import oracle as O
from arapgs_amd import arapgs, scenes, workload
from arapgs_amd.pipeline import DragPipeline

pytestmark = pytest.mark.gpu

FRAMES = 12
TUNED = dict(tile=16, blend_unroll=8, aggregates=192, refresh=8, host_checks=4, warm=1)


class RecordingPipeline(DragPipeline):
    """DragPipeline that keeps every frame's solve inputs (node positions and
    aims) and outputs (energies, GN iterations), copied on the solve stream."""

    def __init__(self, *a, **kw):
        super().__init__(*a, **kw)
        self.rec = []

    def _solve_half(self, slot):
        torch = self.torch
        with torch.cuda.stream(self.stream_a):
            g = self.st["node_pos"].clone()
        super()._solve_half(slot)
        with torch.cuda.stream(self.stream_a):
            self.rec.append((g, self.aims.clone(), self.energies[slot].clone(), self.iters_done[slot].clone()))


def _tuned_solver(ctx, args):
    s = workload.make_solver(ctx, args, n_aggregates=TUNED["aggregates"])
    s.set_coarse_refresh(TUNED["refresh"])
    s.set_host_checks(TUNED["host_checks"])
    s.set_warm_start(TUNED["warm"])
    s.set_cg_mode(0)
    return s


def test_benchmarked_settings_match_oracle(ctx):
    import torch
    print("bench accuracy: capsules 60k Gaussians, 512 nodes, 4 views 800x800, 12 pipelined frames", flush=True)
    sc = scenes.capsule_scene(0, n=60000, n_nodes=512, W=800, H=800, n_views=4)
    ctx_a = arapgs.Context(0, stream=torch.cuda.Stream())
    with torch.cuda.stream(ctx_a.stream):
        W = workload.prepare(ctx_a, sc)
    torch.cuda.synchronize()
    s0 = workload.clone_state(W["state"])
    delta = workload.drag_delta(sc, 30)
    K, iters = W["K"], sc.gn_iters
    nbr, blocks, types, src_idx, src_w, _ = W["solver_args"]

    st = workload.clone_state(s0)
    solver = _tuned_solver(ctx_a, W["solver_args"])
    imgs = [torch.empty((3, c["H"], c["W"]), device="cuda") for c in W["cams"]]
    torch.cuda.synchronize()
    pipe = RecordingPipeline(solver, st, K, W["active"], delta, iters, W["cams"], imgs, W["bg"], W["opac"],
                             tile=TUNED["tile"])
    pipe.ctx_b.set_blend_unroll(TUNED["blend_unroll"])
    pipe.run(FRAMES)
    torch.cuda.synchronize()

    worst = 0.0
    for t, (g, aims, e, done) in enumerate(pipe.rec):
        g, aims, e, done = g.cpu().numpy(), aims.cpu().numpy(), e.cpu().numpy(), int(done.item())
        ref = O.GNProblem(g, nbr, blocks, types, src_idx, src_w.astype(np.float64), aims)
        _, e_ref, it_ref = ref.solve(iters, stop_rule=True)
        rel = abs(e[done] - e_ref[-1]) / e_ref[-1]
        worst = max(worst, rel)
        print(f"frame {t}: GN {done} (oracle {it_ref}), energy {e[done]:.9g} (oracle {e_ref[-1]:.9g}), "
              f"rel {rel:.2e}", flush=True)
        assert rel <= 1e-5, (t, e[done], e_ref[-1])
        assert abs(done - it_ref) <= 1, (t, done, it_ref)
    print(f"max relative energy error over {FRAMES} frames: {worst:.2e}", flush=True)

    # the last frame's views against the oracle rasteriser
    bg = np.array([0.8, 0.8, 0.8], np.float32)
    pos, sh, scale, rot = (st[k].cpu().numpy() for k in ("pos", "sh", "scale", "rot"))
    opac = W["opac"].cpu().numpy()
    for v, cam in enumerate(sc.cameras):
        r = O.raster(pos, sh, opac, scale, rot, cam.view, cam.proj, cam.campos, cam.tanfovx, cam.tanfovy, cam.W,
                     cam.H, bg)
        err = float(np.abs(imgs[v].cpu().numpy() - r["image"]).max())
        print(f"view {v}: max pixel error {err:.2e}", flush=True)
        assert err <= 1.0 / 255.0, (v, err)

    # the queued sequential form with the same settings: the same state bit for bit
    st2 = workload.clone_state(s0)
    solver2 = _tuned_solver(ctx_a, W["solver_args"])
    ctx_a.set_blend_unroll(TUNED["blend_unroll"])
    fb = arapgs.frame_buffers(st2, K)
    aims = torch.empty_like(st2["node_pos"])
    with torch.cuda.stream(ctx_a.stream):
        for _ in range(FRAMES):
            ctx_a.aims_translate(st2["node_pos"], W["active"], delta, aims)
            solver2.deform_frame(fb, aims, iters, stop_rule=True)
    torch.cuda.synchronize()
    for key in ("pos", "rot", "scale", "sh", "ends", "node_pos", "samples", "sample_feat"):
        assert np.array_equal(st[key].cpu().numpy(), st2[key].cpu().numpy()), key
    pipe.close()
    solver.close()
    solver2.close()
    ctx_a.close()
