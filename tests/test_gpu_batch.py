"""Scene batches: the Gauss-Newton solves of independent scenes of uneven size
run as one launch per solver phase (arapgs_solver_batch_solve, scene =
blockIdx.z). Per scene the batch must give exactly what a solve of that scene
alone gives (x, energies, iteration count, putFreeInputs node table, bit for
bit), over consecutive frames, with warm start, centre-mode constraints and
K != 8; and the batched energies match the FP64 oracle's solve.
"""
import numpy as np
import pytest

import oracle as O
from arapgs_amd import arapgs
from arapgs_amd import scenes as S
from helpers import small_scene

pytestmark = pytest.mark.gpu


def T(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def H(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def scenes():
    return [small_scene(n=12000, n_nodes=64, seed=1),
            small_scene(n=30000, n_nodes=256, seed=2, kind="capsules"),
            small_scene(n=20000, n_nodes=128, seed=3)]


def _solver(ctx, sc, center=False, warm=False):
    s = arapgs.Solver(ctx, sc["nbr"], sc["blocks"], sc["types"], sc["src_idx"], sc["src_w"].astype(np.float32),
                      sc["nodes"], cg_max_iters=200, cg_tol=1e-5, constraints_on_center=center)
    if warm:
        s.set_warm_start(True)
    return s


def _aims(sc, fraction):
    return S.frame_aims(sc["scene"], sc["nodes"], sc["blocks"], sc["types"], fraction)


def _run(ctx, scs, centers, warm, batched, fractions, iters=10):
    """Per frame and scene: (x, energies[:it+1], it, table) as numpy."""
    import torch
    solvers = [_solver(ctx, sc, c, warm) for sc, c in zip(scs, centers)]
    batch = arapgs.SolverBatch(ctx, solvers) if batched else None
    npos = [T(sc["nodes"]) for sc in scs]
    out = []
    try:
        for fr in fractions:
            aims = [T(_aims(sc, fr)) for sc in scs]
            tables = [arapgs.node_table(sc["nodes"].shape[0], npos[0].device) for sc in scs]
            if batched:
                xs, es, its = batch.solve(npos, aims, iters, tables=tables)
            else:
                xs, es, its = [], [], []
                for s, p, a, tb in zip(solvers, npos, aims, tables):
                    x, e, it = s.solve(p, a, iters)
                    s.put_free_inputs(p, x, table=tb)
                    xs.append(x), es.append(e), its.append(it)
            torch.cuda.synchronize()
            frame = []
            for x, e, it, tb in zip(xs, es, its, tables):
                n_it = int(H(it)[0])
                frame.append((H(x), H(e)[: n_it + 1], n_it, H(tb)))
            out.append(frame)
    finally:
        if batch is not None:
            batch.close()
        for s in solvers:
            s.close()
    return out


def _assert_identical(a, b):
    for f, (fa, fb) in enumerate(zip(a, b)):
        for s, (ra, rb) in enumerate(zip(fa, fb)):
            assert ra[2] == rb[2], (f, s, ra[2], rb[2])
            assert ra[2] >= 1
            for k in (0, 1, 3):
                assert np.array_equal(ra[k], rb[k]), (f, s, k)


@pytest.mark.parametrize("warm", [False, True])
def test_batch_equals_single_solves(ctx, scenes, warm):
    centers = [False, False, False]
    fr = [0.4, 0.7, 1.0]
    print(f"batch vs single, warm={warm}", flush=True)
    single = _run(ctx, scenes, centers, warm, False, fr)
    batched = _run(ctx, scenes, centers, warm, True, fr)
    _assert_identical(single, batched)


def test_batch_mixed_constraint_modes_and_k(ctx, scenes):
    """A centre-mode scene (multi-member constraint groups: the coarse_multi
    launch) and a K = 6 scene (generic owner walks) beside a plain K = 8 one."""
    k6 = small_scene(n=10000, n_nodes=64, K=6, seed=4)
    scs = [scenes[0], scenes[2], k6]
    centers = [False, True, False]
    print("batch vs single, centre mode + K=6", flush=True)
    single = _run(ctx, scs, centers, False, False, [0.5, 1.0], iters=4)
    batched = _run(ctx, scs, centers, False, True, [0.5, 1.0], iters=4)
    _assert_identical(single, batched)


def test_batch_energies_match_oracle(ctx, scenes):
    out = _run(ctx, scenes, [False] * 3, False, True, [1.0], iters=30)[0]
    for sc, (x, e, it, tb) in zip(scenes, out):
        aims = _aims(sc, 1.0)
        ref = O.GNProblem(sc["nodes"], sc["nbr"], sc["blocks"], sc["types"], sc["src_idx"], sc["src_w"], aims)
        _, e_ref, _ = ref.solve(30)
        assert abs(e[-1] - e_ref[-1]) <= 1e-5 * e_ref[-1] + 1e-9, (e[-1], e_ref[-1])
        # the reported energy is the oracle's energy function at the returned x
        f = ref.residual(x.astype(np.float64))
        assert abs(float(f @ f) - e[-1]) <= 1e-5 * e[-1] + 1e-9


def test_batch_rejects_bad_input(ctx, scenes):
    s = _solver(ctx, scenes[0])
    try:
        with pytest.raises(arapgs.ArapgsError):
            arapgs.SolverBatch(ctx, [s, s])
        b = arapgs.SolverBatch(ctx, [s])
        p = T(scenes[0]["nodes"])
        with pytest.raises(arapgs.ArapgsError):
            b.solve([p], [p], 65)
        b.close()
    finally:
        s.close()


def _cam(c):
    return dict(view=T(c.view), proj=T(c.proj), campos=T(c.campos), tx=c.tanfovx, ty=c.tanfovy)


def test_render_segments_equals_single_renders(ctx, scenes):
    """Scenes back to back, one camera each, one pass: every image equals the
    render of its own scene alone, bit for bit (including an empty segment)."""
    import torch
    gs = [sc["g"] for sc in scenes]
    cams = [_cam(sc["scene"].cameras[0]) for sc in scenes]
    cam0 = scenes[0]["scene"].cameras[0]
    W, H = cam0.W, cam0.H
    # a fourth, empty scene between the first two
    order = [0, None, 1, 2]
    seg = [0]
    parts = []
    vcams = []
    for i in order:
        if i is None:
            seg.append(seg[-1])
            vcams.append(cams[1])
            continue
        parts.append(gs[i])
        seg.append(seg[-1] + gs[i].n)
        vcams.append(cams[i])
    cat = lambda f: T(np.concatenate([getattr(g, f) for g in parts]))  # noqa: E731
    means, rot, scale, opac, sh = cat("pos"), cat("rot"), cat("scale"), cat("opacity"), cat("sh")
    bg = T(np.array([0.8, 0.8, 0.8], np.float32))
    outs = [torch.zeros((3, H, W), dtype=torch.float32, device="cuda") for _ in order]
    print("render_segments: 4 views", flush=True)
    n_all = ctx.render_segments(means, sh, opac, scale, rot, vcams, W, H, bg, outs, seg)
    torch.cuda.synchronize()
    n_sum = 0
    for v, i in enumerate(order):
        ref = torch.zeros((3, H, W), dtype=torch.float32, device="cuda")
        if i is None:
            ref[:] = bg.view(3, 1, 1)
        else:
            g = gs[i]
            n_sum += ctx.render_views(T(g.pos), T(g.sh), T(g.opacity), T(g.scale), T(g.rot), [cams[i]], W, H, bg,
                                      [ref], fast_culling=False)
        torch.cuda.synchronize()
        assert torch.equal(outs[v], ref), (v, (outs[v] - ref).abs().max().item())
    assert n_all == n_sum > 0
