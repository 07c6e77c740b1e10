"""Pipelined drag frames (solve of frame t+1 overlapping the deformation and
renders of frame t on a second stream) give exactly the sequential frames'
state and images: GaussianView::onUpdate order (GaussianView.cpp:1481-1522)
is preserved per kernel, only the streams differ. With batched sample frames
(three frames' sample updates per pass, and a final partial batch of two) the
samples and their features are still bit-identical to per-frame updates.
arapgs_samples_frames itself equals its per-frame calls bit for bit on random
node tables, static samples included."""
import numpy as np
import pytest

from arapgs_amd import arapgs, scenes, workload
from arapgs_amd.pipeline import DragPipeline

pytestmark = pytest.mark.gpu

FRAMES = 5


@pytest.mark.parametrize("sample_batch,masked,compose,late,split", [
    (1, False, False, False, False), (3, False, False, False, False), (8, False, False, False, False),
    (3, True, False, False, False), (4, False, True, False, False), (1, False, False, True, False),
    (4, False, False, True, False), (1, False, False, False, True), (4, True, False, False, True)])
def test_pipelined_frames_bit_identical(ctx, sample_batch, masked, compose, late, split):
    """Pipelined frames (and the rolling sample chunks of sample_batch > 1) equal
    sequential frames bit for bit; the drag is run as two run() calls (3 + 2
    frames), so the table slots and chunk records carry over between runs.
    masked: the deform and samples streams are CU-masked streams that leave the
    first 32 CUs to the solve (one from a CU range, one from an explicit mask).
    compose: the chunks' pending frames rotate the features once by the product
    of their rotations; everything but the features stays bit-identical and
    the features agree within 2e-6 of their scale. late: the samples stream
    starts each frame after the six-point update (bit-identical). split: the
    renders run on their own stream from double-buffered attributes (the
    5-frame drag in runs of 3 and 2 ends on each of the two sets once)."""
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
    masks = []
    if masked:
        n_cu = torch.cuda.get_device_properties(0).multi_processor_count
        masks = [arapgs.CuStream(0, 32, n_cu - 32), arapgs.CuStream(0, cus=list(range(32, n_cu))),
                 arapgs.CuStream(0, 32, n_cu - 32)]
    pipe = DragPipeline(solver2, st2, K, W["active"], delta, iters, W["cams"], imgs_pipe, W["bg"], W["opac"],
                        sample_batch=sample_batch, stream_b=masks[0].torch if masks else None,
                        stream_c=masks[1].torch if masks else None, sh_compose=compose, samples_late=late,
                        split_render=split, stream_r=masks[2].torch if (masks and split) else None)
    pipe.run(3)
    pipe.run(FRAMES - 3)
    torch.cuda.synchronize()

    for key in ("pos", "rot", "scale", "sh", "ends", "node_pos", "samples", "sample_feat"):
        a, b = st[key].cpu().numpy(), st2[key].cpu().numpy()
        if compose and key == "sample_feat":
            assert np.abs(a - b).max() <= 2e-6 * np.abs(a).max(), key
            continue
        assert np.array_equal(a, b), key
    for a, b in zip(imgs_seq, imgs_pipe):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    # the drag really moved things
    assert not np.array_equal(st["pos"].cpu().numpy(), s0["pos"].cpu().numpy())
    pipe.close()
    del pipe, solver2
    ctx_a.close()
    torch.cuda.synchronize()
    for m in masks:
        m.close()


def test_samples_frames_equal_per_frame_calls(ctx):
    import torch
    from helpers import random_node_xf, table_from
    rng = np.random.default_rng(9)
    M, S, K = 96, 20000, 8
    nodes = rng.uniform(-1, 1, size=(M, 3)).astype(np.float32)
    pts = rng.uniform(-1, 1, size=(S, 3)).astype(np.float32)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    idx_full, w = ctx.knn(T(nodes), T(pts), K)
    idx = idx_full[:, :K].contiguous()
    feat0 = T(rng.normal(size=(S, 48)).astype(np.float32))
    static = T((rng.random(S) < 0.1).astype(np.uint8))
    tables = []
    for f in range(5):
        rot9, t3 = random_node_xf(M, rng, rot_deg=15.0, trans=0.05)
        tables.append(T(table_from(nodes + 0.01 * f, rot9, t3)))
    quats = [ctx.node_quats(t) for t in tables]
    p_ref, f_ref = T(pts), feat0.clone()
    for t, q in zip(tables, quats):
        ctx.blend_points(p_ref, idx, w, t, skip=static, out=p_ref)
        ctx.sample_sh(w, idx, q, static, f_ref)
    p_bat, f_bat = T(pts), feat0.clone()
    ctx.samples_frames(tables[:3], quats[:3], idx, w, static, p_bat, f_bat)
    ctx.samples_frames(tables[3:], quats[3:], idx, w, static, p_bat, f_bat)
    torch.cuda.synchronize()
    assert torch.equal(p_bat, p_ref)
    assert torch.equal(f_bat, f_ref)
    assert not torch.equal(f_bat, feat0)
    # one SH rotation by the product of the frames' rotations (set_sh_compose):
    # positions bit for bit, features within float rounding of the per-frame
    # rotations (|error| <= 2e-6 of the feature scale after 5 frames), static
    # samples untouched
    c2 = arapgs.Context(0)
    c2.set_sh_compose(True)
    p_cmp, f_cmp = T(pts), feat0.clone()
    c2.samples_frames(tables[:3], quats[:3], idx, w, static, p_cmp, f_cmp)
    c2.samples_frames(tables[3:], quats[3:], idx, w, static, p_cmp, f_cmp)
    torch.cuda.synchronize()
    assert torch.equal(p_cmp, p_ref)
    err = (f_cmp - f_ref).abs().max().item()
    print(f"composed SH rotation: max |feature error| {err:.3e}", flush=True)
    assert err <= 2e-6 * f_ref.abs().max().item()
    st = static.bool()
    assert torch.equal(f_cmp[st], feat0[st])
    c2.close()
