"""Drag frames of a batch of independent scenes (batch.SceneBatch): the
batched frame (one launch per step over all scenes) and the same frame run
scene after scene with the single-scene calls give identical state (Gaussians,
nodes, samples, sample features) and images, bit for bit, over several frames
of a random-walk drag; every scene's Gauss-Newton iteration count and energies
agree too."""
import numpy as np
import pytest

from arapgs_amd import batch as BT
from arapgs_amd import workload

pytestmark = pytest.mark.gpu


def _prepare(ctx):
    scs = workload.batch_scenes(7, 3, n_min=20_000, n_max=60_000, W=256, H=192, node_div=300)
    return workload.prepare_batch(ctx, scs)


def test_batched_frames_equal_looped_frames(ctx):
    import torch
    print("scene batch: preparing 3 scenes", flush=True)
    wa = _prepare(ctx)
    wb = dict(wa)
    wb["state"] = workload.clone_state(wa["state"])
    a = BT.SceneBatch(ctx, wa, gn_iters=6, cg_max_iters=200, cg_tol=1e-5)
    b = BT.SceneBatch(ctx, wb, gn_iters=6, cg_max_iters=200, cg_tol=1e-5)
    steps = BT.random_walk_steps(3, a.n, 3, step=0.02)
    try:
        for f in range(steps.shape[0]):
            print(f"scene batch: frame {f}", flush=True)
            na = a.frame(steps[f])
            nb = b.frame_looped(steps[f])
            torch.cuda.synchronize()
            assert na == nb > 0
            for i in range(a.n):
                ia, ib = int(a.iters_done[i].item()), int(b.iters_done[i].item())
                assert ia == ib >= 1
                assert torch.equal(a.energies[i][: ia + 1], b.energies[i][: ib + 1])
                assert torch.equal(a.images[i], b.images[i]), (f, i)
            for k in ("pos", "rot", "scale", "sh", "node_pos", "samples", "sample_feat", "ends"):
                assert torch.equal(wa["state"][k], wb["state"][k]), (f, k)
        # the drag moved every scene
        e0 = [float(a.energies[i][0].item()) for i in range(a.n)]
        assert all(e > 0 for e in e0)
    finally:
        a.close()
        b.close()
