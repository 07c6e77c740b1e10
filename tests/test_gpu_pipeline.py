"""Pipelined drag frames (solve of frame t+1 overlapping the deformation and
renders of frame t on a second stream) give exactly the sequential frames'
state and images: GaussianView::onUpdate order (GaussianView.cpp:1481-1522)
is preserved per kernel, only the streams differ."""
import numpy as np
import pytest

from arapgs_amd import arapgs, scenes, workload
from arapgs_amd.pipeline import DragPipeline

pytestmark = pytest.mark.gpu

FRAMES = 5


def test_pipelined_frames_bit_identical(ctx):
    import torch
    sc = scenes.capsule_scene(7, n=40000, n_nodes=128, W=160, H=128, n_views=2)
    sc.grid = 32
    W = workload.prepare(ctx, sc)
    s0 = workload.clone_state(W["state"])
    delta = workload.drag_delta(sc, 10)
    K, iters = W["K"], sc.gn_iters

    # sequential frames on the fixture's context
    st = workload.clone_state(s0)
    solver = workload.make_solver(ctx, W["solver_args"])
    solver.set_host_checks(16)
    fb = arapgs.frame_buffers(st, K)
    aims = torch.empty_like(st["node_pos"])
    imgs_seq = [torch.empty((3, c["H"], c["W"]), device="cuda") for c in W["cams"]]
    for _ in range(FRAMES):
        ctx.aims_translate(st["node_pos"], W["active"], delta, aims)
        solver.deform_frame(fb, aims, iters, stop_rule=True)
        for cam, img in zip(W["cams"], imgs_seq):
            ctx.render(st["pos"], st["sh"], W["opac"], st["scale"], st["rot"], cam["view"], cam["proj"],
                       cam["campos"], cam["tx"], cam["ty"], cam["W"], cam["H"], W["bg"], out=img)
    torch.cuda.synchronize()

    # pipelined frames: solve context on its own stream, deform/render context inside the pipeline
    ctx_a = arapgs.Context(0, stream=torch.cuda.Stream())
    st2 = workload.clone_state(s0)
    solver2 = workload.make_solver(ctx_a, W["solver_args"])
    solver2.set_host_checks(16)
    imgs_pipe = [torch.empty((3, c["H"], c["W"]), device="cuda") for c in W["cams"]]
    torch.cuda.synchronize()
    pipe = DragPipeline(solver2, st2, K, W["active"], delta, iters, W["cams"], imgs_pipe, W["bg"], W["opac"])
    pipe.run(FRAMES)
    torch.cuda.synchronize()

    for key in ("pos", "rot", "scale", "sh", "ends", "node_pos", "samples", "sample_feat"):
        a, b = st[key].cpu().numpy(), st2[key].cpu().numpy()
        assert np.array_equal(a, b), key
    for a, b in zip(imgs_seq, imgs_pipe):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    # the drag really moved things
    assert not np.array_equal(st["pos"].cpu().numpy(), s0["pos"].cpu().numpy())
    pipe.close()
    del pipe, solver2
    ctx_a.close()
