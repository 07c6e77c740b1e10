"""GPU parity of the ARAP Gauss-Newton solve (Deform.cpp:95-581) vs the FP64 oracle.

Bars: the energy function itself within 1e-5 relative at arbitrary x; energy at
every GN iteration non-increasing; the final energy within 1e-5 relative of the
oracle's converged energy, and the oracle re-evaluated at the GPU's x agrees
with the GPU's reported energy.
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
def prob():
    sc = small_scene(n=20000, n_nodes=128)
    return sc


def _translate_aims(nodes, blocks, types, fraction):
    """Aims of a translating drag (UpdateAimPosition, GaussianView.cpp:2920-2931):
    the centre-mode constraint sees only the mean of the block's aims, which a
    twist about the block's centre leaves in place."""
    aims = np.array(nodes, np.float32, copy=True)
    for blk, t in zip(blocks, types):
        if t == 1:
            aims[blk] += np.float32(fraction) * np.array([0.12, 0.05, -0.08], np.float32)
    return aims


def _problem(sc, fraction, wf=1.0, center=False):
    nodes = sc["nodes"]
    if center:
        aims = _translate_aims(nodes, sc["blocks"], sc["types"], fraction)
    else:
        aims = S.frame_aims(sc["scene"], nodes, sc["blocks"], sc["types"], fraction)
    ref = O.GNProblem(nodes, sc["nbr"], sc["blocks"], sc["types"], sc["src_idx"], sc["src_w"], aims, wf=wf,
                      center=center)
    return aims, ref


def _solver(ctx, sc, wf=1.0, blocks=None, types=None, center=False):
    return arapgs.Solver(ctx, sc["nbr"], sc["blocks"] if blocks is None else blocks,
                         sc["types"] if types is None else types, sc["src_idx"], sc["src_w"].astype(np.float32),
                         sc["nodes"], weighting_factor=wf, cg_max_iters=200, cg_tol=1e-5,
                         constraints_on_center=center)


@pytest.mark.parametrize("wf,center", [(1.0, False), (3.0, False), (3.0, True)])
def test_energy_function(ctx, prob, wf, center):
    print(f"solver energy wf={wf} center={center}", flush=True)
    aims, ref = _problem(prob, 0.5, wf, center)
    s = _solver(ctx, prob, wf, center=center)
    rng = np.random.default_rng(0)
    for trial in range(3):
        x = ref.identity_x() + 0.05 * rng.normal(size=ref.n)
        e_ref = ref.energy(x)
        e_gpu = float(H(s.energy(T(prob["nodes"]), T(aims), T(x.astype(np.float32))))[0])
        e_ref32 = ref.energy(x.astype(np.float32).astype(np.float64))
        assert abs(e_gpu - e_ref32) <= 1e-5 * e_ref32, (e_gpu, e_ref32, e_ref)


@pytest.mark.parametrize("fraction", [1.0, 0.1])
def test_gn_solve_matches_oracle(ctx, prob, fraction):
    aims, ref = _problem(prob, fraction)
    x_ref, e_ref, it_ref = ref.solve(8, stop_rule=True)
    s = _solver(ctx, prob)
    x, e, done = s.solve(T(prob["nodes"]), T(aims), 8, stop_rule=True)
    x, e, done = H(x).astype(np.float64), H(e), int(H(done)[0])
    assert 1 <= done <= 8
    e = e[: done + 1]
    assert abs(e[0] - e_ref[0]) <= 1e-6 * e_ref[0]          # same start (identity)
    assert np.all(np.diff(e) <= 1e-9 * e[0])                   # monotone GN with line search
    e_final_ref = e_ref[-1]
    assert abs(e[-1] - e_final_ref) <= 1e-5 * e_final_ref, (e, e_ref)
    e_at_x = ref.energy(x)
    assert abs(e_at_x - e[-1]) <= 1e-5 * e_at_x
    # the solution itself: node transforms close to the oracle's
    assert np.abs(x - x_ref).max() < 1e-3
    # the residual and gradient are FP64, so the stop rule fires where the FP64 oracle's does
    assert abs(done - it_ref) <= 1, (done, it_ref)


def test_static_blocks(ctx, prob):
    """An excluded (-1) block removes its nodes from the unknowns (UpdateIndicies)."""
    blocks = list(prob["blocks"]) + [np.arange(0, 8, dtype=np.int32)]
    types = np.concatenate([prob["types"], [-1]]).astype(np.int32)
    blocks[0] = np.setdiff1d(blocks[0], blocks[-1]).astype(np.int32)
    blocks[1] = np.setdiff1d(blocks[1], blocks[-1]).astype(np.int32)
    aims = S.frame_aims(prob["scene"], prob["nodes"], blocks, types, 0.5)
    ref = O.GNProblem(prob["nodes"], prob["nbr"], blocks, types, prob["src_idx"], prob["src_w"], aims)
    x_ref, e_ref, _ = ref.solve(6, stop_rule=True)
    s = _solver(ctx, prob, blocks=blocks, types=types)
    assert s.n_free == ref.n // 12
    x, e, done = s.solve(T(prob["nodes"]), T(aims), 6, stop_rule=True)
    e = H(e)[: int(H(done)[0]) + 1]
    assert abs(e[-1] - e_ref[-1]) <= 1e-5 * e_ref[-1]


@pytest.mark.parametrize("n_nodes", [128, 1024])
def test_cg_launch_modes_identical(ctx, n_nodes):
    """The persistent cooperative PCG kernel and the three-kernels-per-iteration
    path (with and without a host-driven loop) compute the same iterates bit for
    bit, and repeated persistent solves are deterministic. One solver instance
    is used so that all solves share one coarse operator (assembled with float
    atomics once, on the first solve)."""
    sc = small_scene(n=20000 if n_nodes == 128 else 60000, n_nodes=n_nodes, seed=1)
    aims, _ = _problem(sc, 0.7)
    s = _solver(ctx, sc)
    s.set_coarse_refresh(1000)
    outs = []
    for mode, every in ((1, 0), (0, 1), (0, 0), (1, 0)):
        s.set_cg_mode(mode)
        s.set_host_checks(every)
        x, e, done = s.solve(T(sc["nodes"]), T(aims), 6, stop_rule=True)
        outs.append((H(x), H(e), int(H(done)[0])))
    for x, e, done in outs[1:]:
        assert done == outs[0][2] and done >= 1
        assert np.array_equal(x, outs[0][0])
        assert np.array_equal(e[: done + 1], outs[0][1][: done + 1])


def _center_blocks(prob):
    """The scene's two blocks plus a static block of the 8 nodes farthest from
    them: with only block centres constrained, two non-static blocks leave the
    rotation about the line through their centres free (J^T J singular)."""
    nodes = prob["nodes"]
    used = np.concatenate(prob["blocks"])
    c = nodes[used].mean(0)
    far = [int(i) for i in np.argsort(-np.linalg.norm(nodes - c, axis=1), kind="stable") if i not in set(used)][:8]
    return list(prob["blocks"]) + [np.asarray(far, np.int32)], np.concatenate([prob["types"], [-1]]).astype(np.int32)


@pytest.mark.parametrize("fraction", [1.0, 0.3])
def test_gn_solve_center_mode_matches_oracle(ctx, prob, fraction):
    """Constraints on the centre of each control block (Deform.cpp:34-75,
    483-529): one 3-row group per block over its first min(20, |block|) nodes;
    a static block makes the problem well posed."""
    blocks, types = _center_blocks(prob)
    print(f"solver centre mode fraction={fraction} blocks={[len(b) for b in blocks]}", flush=True)
    nodes = prob["nodes"]
    aims = _translate_aims(nodes, blocks, types, fraction)
    ref = O.GNProblem(nodes, prob["nbr"], blocks, types, prob["src_idx"], prob["src_w"], aims, center=True)
    x_ref, e_ref, it_ref = ref.solve(8, stop_rule=True)
    s = _solver(ctx, prob, blocks=blocks, types=types, center=True)
    assert s.n_free == ref.n // 12
    x, e, done = s.solve(T(nodes), T(aims), 8, stop_rule=True)
    x, e, done = H(x).astype(np.float64), H(e), int(H(done)[0])
    assert 1 <= done <= 8
    e = e[: done + 1]
    assert abs(e[0] - e_ref[0]) <= 1e-6 * e_ref[0]
    assert np.all(np.diff(e) <= 1e-9 * e[0])
    assert abs(e[-1] - e_ref[-1]) <= 1e-5 * e_ref[-1], (e, e_ref)
    e_at_x = ref.energy(x)
    assert abs(e_at_x - e[-1]) <= 1e-5 * e_at_x
    assert abs(done - it_ref) <= 1, (done, it_ref)


def test_center_mode_singular_problem_converges_to_oracle_energy(ctx, prob):
    """Two non-static blocks in centre mode: J^T J is singular (the rotation
    about the line through the two centres costs nothing), so the FP32 CG
    takes more GN iterations than the oracle's FP64 Cholesky; within the
    reference's 30 iterations (Deform.cpp:3) both reach the same energy."""
    aims, ref = _problem(prob, 0.3, center=True)
    x_ref, e_ref, it_ref = ref.solve(30, stop_rule=True)
    s = _solver(ctx, prob, center=True)
    x, e, done = s.solve(T(prob["nodes"]), T(aims), 30, stop_rule=True)
    e, done = H(e), int(H(done)[0])
    assert abs(e[done] - e_ref[-1]) <= 1e-5 * e_ref[-1], (e[: done + 1], e_ref)


def test_center_mode_launch_modes_identical(ctx, prob):
    """Centre mode through the persistent and the per-phase PCG paths: same bits."""
    blocks, types = _center_blocks(prob)
    aims = _translate_aims(prob["nodes"], blocks, types, 0.7)
    s = _solver(ctx, prob, blocks=blocks, types=types, center=True)
    s.set_coarse_refresh(1000)
    outs = []
    for mode, every in ((1, 0), (0, 1)):
        s.set_cg_mode(mode)
        s.set_host_checks(every)
        x, e, done = s.solve(T(prob["nodes"]), T(aims), 6, stop_rule=True)
        outs.append((H(x), H(e), int(H(done)[0])))
    assert outs[0][2] == outs[1][2] >= 1
    assert np.array_equal(outs[0][0], outs[1][0])


def test_center_mode_large_block_keeps_first_20(ctx, prob):
    """A 30-node control block: the centre group uses its first 20 nodes (the
    reference's c_indices[i] selection); energies match the oracle."""
    nodes = prob["nodes"]
    d0 = np.linalg.norm(nodes - nodes[0], axis=1)
    far = int(np.argmax(d0))
    ctrl = np.argsort(d0, kind="stable")[:30].astype(np.int32)
    stat = np.argsort(np.linalg.norm(nodes - nodes[far], axis=1), kind="stable")[:10].astype(np.int32)
    blocks, types = [stat, ctrl], np.array([-1, 1], np.int32)
    aims = _translate_aims(nodes, blocks, types, 0.6)
    ref = O.GNProblem(nodes, prob["nbr"], blocks, types, prob["src_idx"], prob["src_w"], aims, center=True)
    x_ref, e_ref, it_ref = ref.solve(8, stop_rule=True)
    s = _solver(ctx, prob, blocks=blocks, types=types, center=True)
    x, e, done = s.solve(T(nodes), T(aims), 8, stop_rule=True)
    e = H(e)[: int(H(done)[0]) + 1]
    assert abs(e[0] - e_ref[0]) <= 1e-6 * e_ref[0]
    assert abs(e[-1] - e_ref[-1]) <= 1e-5 * e_ref[-1], (e, e_ref)
    # the aims of nodes 21..30 of the block do not enter
    aims2 = aims.copy()
    aims2[ctrl[20:]] += 0.5
    e2 = float(H(s.energy(T(nodes), T(aims2), T(ref.identity_x().astype(np.float32))))[0])
    e1 = float(H(s.energy(T(nodes), T(aims), T(ref.identity_x().astype(np.float32))))[0])
    assert e1 == e2


def test_warm_start_solves_match_oracle(ctx, prob):
    """Warm-started CG (each GN iteration starts from the previous solve's step):
    a sequence of drag frames keeps the oracle's energies (1e-5) and GN
    iteration counts (within 1)."""
    s = _solver(ctx, prob)
    s.set_warm_start(True)
    for fraction in (0.30, 0.35, 0.40, 0.45):
        aims, ref = _problem(prob, fraction)
        x_ref, e_ref, it_ref = ref.solve(8, stop_rule=True)
        x, e, done = s.solve(T(prob["nodes"]), T(aims), 8, stop_rule=True)
        e, done = H(e), int(H(done)[0])
        assert abs(e[done] - e_ref[-1]) <= 1e-5 * e_ref[-1], (fraction, e[: done + 1], e_ref)
        assert abs(done - it_ref) <= 1, (fraction, done, it_ref)
    s.set_warm_start(False)


@pytest.mark.parametrize("center", [False, True])
def test_cg_operator_and_rhs_match_oracle_jacobian(ctx, prob, center):
    """The CG operator q = J^T J v and right-hand side b = -J^T f equal the
    oracle's dense Jacobian products (pinned by central differences)."""
    aims, ref = _problem(prob, 0.5, center=center)
    s = _solver(ctx, prob, center=center)
    rng = np.random.default_rng(4)
    x = (ref.identity_x() + 0.05 * rng.normal(size=ref.n)).astype(np.float32)
    v = rng.normal(size=ref.n).astype(np.float32)
    q, b = s.apply(T(prob["nodes"]), T(aims), T(x), T(v))
    J = ref.jacobian(x.astype(np.float64))
    f = ref.residual(x.astype(np.float64))
    q_ref = J.T @ (J @ v.astype(np.float64))
    b_ref = -J.T @ f
    print(f"center={center}: |q - q_ref| / |q_ref| = {np.linalg.norm(H(q) - q_ref) / np.linalg.norm(q_ref):.3e}, "
          f"|b - b_ref| / |b_ref| = {np.linalg.norm(H(b) - b_ref) / np.linalg.norm(b_ref):.3e}", flush=True)
    assert np.linalg.norm(H(q) - q_ref) <= 1e-5 * np.linalg.norm(q_ref)
    assert np.linalg.norm(H(b) - b_ref) <= 1e-5 * np.linalg.norm(b_ref)


def test_coarse_inverse_is_exact_when_every_node_is_an_aggregate(ctx, prob):
    """With one aggregate per free node the coarse space is the whole space and
    the two-level preconditioner contains (J^T J)^-1 (the blocked Gauss-Jordan
    inverse), so the first GN step's CG needs only a few iterations; a wrong
    inverse shows up as a CG that runs to its cap."""
    aims, ref = _problem(prob, 1.0)
    s = arapgs.Solver(ctx, prob["nbr"], prob["blocks"], prob["types"], prob["src_idx"],
                      prob["src_w"].astype(np.float32), prob["nodes"], cg_max_iters=200, cg_tol=1e-7,
                      n_aggregates=prob["nodes"].shape[0])
    x, e, done = s.solve(T(prob["nodes"]), T(aims), 1, stop_rule=True)
    gn, cg = s.last_stats()
    x_ref, e_ref, _ = ref.solve(1, stop_rule=True)
    print(f"one GN step: {cg} CG iterations, E {float(H(e)[1]):.6e} vs oracle {e_ref[1]:.6e}", flush=True)
    assert gn == 1 and cg <= 20
    assert abs(float(H(e)[1]) - e_ref[1]) <= 1e-5 * e_ref[1]


@pytest.mark.parametrize("block", [64, 128, 256, 512])
def test_cg_block_size_keeps_the_solution(ctx, prob, block):
    """The CG block size (workgroup size of the CG operator kernel, an autotuner
    knob) changes only the order of the p.q partial sums: the final energy stays
    within 1e-5 of the oracle's and the GN count within one, with the asynchronous
    coarse refresh on and off."""
    aims, ref = _problem(prob, 1.0)
    x_ref, e_ref, it_ref = ref.solve(8, stop_rule=True)
    for async_refresh in (True, False):
        s = _solver(ctx, prob)
        s.set_cg_block(block)
        s.set_async_refresh(async_refresh)
        s.set_coarse_refresh(1)
        for _ in range(3):  # consecutive solves: in-line first operator, then side-stream rebuilds
            x, e, it = s.solve(T(prob["nodes"]), T(aims), 8, stop_rule=True)
            n = int(H(it)[0])
            e = H(e)[: n + 1]
            assert abs(e[-1] - e_ref[-1]) <= 1e-5 * e_ref[-1], (block, async_refresh, e[-1], e_ref[-1])
            assert abs(n - it_ref) <= 1
        s.close()


@pytest.mark.parametrize("every,predict", [(4, False), (3, False), (2, True)])
def test_cg_graphs_bit_identical(prob, every, predict):
    """The GN phases and CG chunks replayed as captured HIP graphs (CG flags
    checked one chunk behind) compute the launch loop's iterates bit for bit:
    three consecutive warm-started solves of different aims, on a context with
    its own stream (graphs are captured on a non-null stream), chunk sizes 4
    and 3 (the CG cap of 200 is not a multiple of 3: the last chunk is shorter),
    and predicted first chunks (set_cg_predict) followed by chunks of 2."""
    import torch
    c2 = arapgs.Context(0, stream=torch.cuda.Stream())
    runs = []
    for graphs in (False, True):
        with torch.cuda.stream(c2.stream):
            s = _solver(c2, prob)
            s.set_host_checks(every)
            s.set_warm_start(True)
            s.set_cg_predict(predict)
            s.set_cg_graphs(graphs)
            out = []
            for f in range(3):
                aims, _ = _problem(prob, 0.6 + 0.2 * f)
                x, e, it = s.solve(T(prob["nodes"]), T(aims), 8, stop_rule=True)
                out.append((x.clone(), e.clone(), it.clone()))
        torch.cuda.synchronize()
        runs.append([tuple(H(t) for t in o) for o in out])
        s.close()
    for a, b in zip(*runs):
        assert int(a[2][0]) >= 1
        for u, v in zip(a, b):
            assert np.array_equal(u, v)
    c2.close()


@pytest.mark.parametrize("n_nodes", [128, 1024])
def test_two_kernel_cg_matches_oracle(ctx, n_nodes):
    """CG iterations as two kernels (cg_mode 2: the coarse restriction carried
    by the recurrence rc_{k+1} = rc_k - alpha P^T q_k, update and coarse phases
    in one launch): the final energy within 1e-5 of the oracle's and the GN
    count within one; the iterates do not depend on how the host checks the
    convergence flag (every iteration, every 3, every 16, never, graph chunks of
    4) — the device-side test decides, bit for bit the same."""
    import torch
    sc = small_scene(n=20000 if n_nodes == 128 else 60000, n_nodes=n_nodes, seed=1)
    aims, ref = _problem(sc, 0.7)
    x_ref, e_ref, it_ref = ref.solve(8, stop_rule=True)
    c2 = arapgs.Context(0, stream=torch.cuda.Stream())
    outs = []
    with torch.cuda.stream(c2.stream):
        s = _solver(c2, sc)
        s.set_coarse_refresh(1000)
        s.set_cg_mode(2)
        for every, graphs in ((1, False), (3, False), (16, False), (0, False), (4, True)):
            s.set_host_checks(every)
            s.set_cg_graphs(graphs)
            x, e, done = s.solve(T(sc["nodes"]), T(aims), 8, stop_rule=True)
            gn, cg = s.last_stats()
            outs.append((H(x), H(e), int(H(done)[0]), cg))
    torch.cuda.synchronize()
    x, e, done, cg = outs[0]
    print(f"two-kernel CG, {n_nodes} nodes: GN {done} (oracle {it_ref}), {cg} CG iterations, "
          f"E {e[done]:.9g} vs {e_ref[-1]:.9g}", flush=True)
    assert abs(e[done] - e_ref[-1]) <= 1e-5 * e_ref[-1]
    assert abs(done - it_ref) <= 1
    assert np.all(np.diff(e[: done + 1]) <= 1e-9 * e[0])
    for o in outs[1:]:
        assert o[2] == done and o[3] == cg
        assert np.array_equal(o[0], x)
        assert np.array_equal(o[1][: done + 1], e[: done + 1])
    s.close()
    c2.close()


@pytest.mark.parametrize("block", [64, 256, 512])
def test_two_kernel_cg_warm_drag_and_center_mode(ctx, prob, block):
    """The two-kernel CG form at CG block sizes 64, 256 and 512 (member nodes per
    pass of the aggregate workgroups) over warm-started drag frames (energies
    within 1e-5 of the oracle's, GN counts within one) and in centre mode."""
    s = _solver(ctx, prob)
    s.set_cg_mode(2)
    s.set_cg_block(block)
    s.set_host_checks(4)
    s.set_warm_start(True)
    for fraction in (0.30, 0.35, 0.40, 0.45):
        aims, ref = _problem(prob, fraction)
        x_ref, e_ref, it_ref = ref.solve(8, stop_rule=True)
        x, e, done = s.solve(T(prob["nodes"]), T(aims), 8, stop_rule=True)
        e, done = H(e), int(H(done)[0])
        assert abs(e[done] - e_ref[-1]) <= 1e-5 * e_ref[-1], (fraction, e[: done + 1], e_ref)
        assert abs(done - it_ref) <= 1, (fraction, done, it_ref)
    s.close()
    blocks, types = _center_blocks(prob)
    aims = _translate_aims(prob["nodes"], blocks, types, 0.7)
    ref = O.GNProblem(prob["nodes"], prob["nbr"], blocks, types, prob["src_idx"], prob["src_w"], aims, center=True)
    x_ref, e_ref, it_ref = ref.solve(8, stop_rule=True)
    s = _solver(ctx, prob, blocks=blocks, types=types, center=True)
    s.set_cg_mode(2)
    s.set_cg_block(block)
    x, e, done = s.solve(T(prob["nodes"]), T(aims), 8, stop_rule=True)
    e, done = H(e), int(H(done)[0])
    assert abs(e[done] - e_ref[-1]) <= 1e-5 * e_ref[-1], (e[: done + 1], e_ref)
    assert abs(done - it_ref) <= 1
    s.close()


@pytest.mark.parametrize("mode", [0, 2])
def test_cg_predict_bit_identical(ctx, prob, mode):
    """The first CG host check at the previous solve's count (set_cg_predict),
    with the three-kernel CG also the Gauss-Newton stop flag read only from the
    previous solve's GN count on, and host checks that poll the pinned flag
    (set_host_spin) are host scheduling only: warm-started drag frames give the
    same iterates bit for bit as without them (CG modes 0 and 2, host checks
    every 2). The frames include a fixed-count solve (stop rule off) followed
    by stop-rule solves, so the predicted GN count is both too high and too
    low at least once."""
    runs = []
    frames = ((0.30, True, 8), (0.35, False, 5), (0.40, True, 8), (0.45, True, 8), (0.70, True, 8),
              (0.72, True, 8))
    for predict, spin in ((False, False), (True, False), (True, True)):
        s = _solver(ctx, prob)
        s.set_cg_mode(mode)
        s.set_host_checks(2)
        s.set_warm_start(True)
        s.set_cg_predict(predict)
        s.set_host_spin(spin)
        out = []
        for fraction, stop, iters in frames:
            aims, _ = _problem(prob, fraction)
            x, e, done = s.solve(T(prob["nodes"]), T(aims), iters, stop_rule=stop)
            out.append((H(x), H(e), int(H(done)[0]), s.last_stats()))
        runs.append(out)
        s.close()
    print("GN counts:", [o[2] for o in runs[0]], flush=True)
    for other in runs[2:]:
        for a, b in zip(runs[0], other):
            assert a[2] == b[2] >= 1 and a[3] == b[3]
            assert np.array_equal(a[0], b[0])
            assert np.array_equal(a[1][: a[2] + 1], b[1][: b[2] + 1])
    for a, b in zip(runs[0], runs[1]):
        assert a[2] == b[2] >= 1 and a[3] == b[3]
        assert np.array_equal(a[0], b[0])
        assert np.array_equal(a[1][: a[2] + 1], b[1][: b[2] + 1])


@pytest.mark.parametrize("lag,refresh", [(2, 1), (3, 2)])
def test_coarse_lag_keeps_oracle_energies(ctx, prob, lag, refresh):
    """A side-stream coarse rebuild adopted `lag` solves after it started
    (set_coarse_lag), over warm-started drag frames: every frame's energy within
    1e-5 of the oracle's and the GN count within one; a reset solver repeats
    the same frames bit for bit (the operator sequence depends on the solve
    count only)."""
    runs = []
    for rep in range(2):
        s = _solver(ctx, prob)
        s.set_warm_start(True)
        s.set_async_refresh(True)
        s.set_coarse_refresh(refresh)
        s.set_coarse_lag(lag)
        s.set_host_checks(4)
        out = []
        for fraction in (0.30, 0.33, 0.36, 0.39, 0.42, 0.45):
            aims, ref = _problem(prob, fraction)
            x, e, done = s.solve(T(prob["nodes"]), T(aims), 8, stop_rule=True)
            e, done = H(e), int(H(done)[0])
            if rep == 0:
                _, e_ref, it_ref = ref.solve(8, stop_rule=True)
                assert abs(e[done] - e_ref[-1]) <= 1e-5 * e_ref[-1], (lag, fraction, e[done], e_ref[-1])
                assert abs(done - it_ref) <= 1
            out.append((H(x), e[: done + 1]))
        runs.append(out)
        s.close()
    for a, b in zip(*runs):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_assembled_operator_matches_oracle_and_matrix_free(ctx, prob):
    """The assembled CG operator (set_cg_assembled: per-frame 4x4 pair blocks
    plus the diagonal blocks) equals the oracle's dense J^T J v to 1e-5 and the
    matrix-free walk to 1e-6; centre mode refuses it."""
    aims, ref = _problem(prob, 0.5)
    rng = np.random.default_rng(11)
    x = (ref.identity_x() + 0.05 * rng.normal(size=ref.n)).astype(np.float32)
    v = rng.normal(size=ref.n).astype(np.float32)
    s = _solver(ctx, prob)
    q_mf, b_mf = s.apply(T(prob["nodes"]), T(aims), T(x), T(v))
    q_mf, b_mf = H(q_mf).copy(), H(b_mf).copy()
    s.set_cg_assembled(True)
    q_as, b_as = s.apply(T(prob["nodes"]), T(aims), T(x), T(v))
    q_as, b_as = H(q_as), H(b_as)
    J = ref.jacobian(x.astype(np.float64))
    q_ref = J.T @ (J @ v.astype(np.float64))
    rel_o = np.linalg.norm(q_as - q_ref) / np.linalg.norm(q_ref)
    rel_m = np.linalg.norm(q_as - q_mf) / np.linalg.norm(q_mf)
    print(f"assembled operator: vs oracle {rel_o:.3e}, vs matrix-free {rel_m:.3e}", flush=True)
    assert rel_o <= 1e-5 and rel_m <= 1e-6
    assert np.array_equal(b_as, b_mf)
    s.close()
    blocks, types = _center_blocks(prob)
    sc = _solver(ctx, prob, blocks=blocks, types=types, center=True)
    with pytest.raises(arapgs.ArapgsError):
        sc.set_cg_assembled(True)
    sc.close()


@pytest.mark.parametrize("mode", [0, 2])
def test_assembled_operator_solves_match_oracle(ctx, prob, mode):
    """Warm-started drag frames with the assembled CG operator (CG modes 0 and
    2): energies within 1e-5 of the oracle's, GN counts within one."""
    s = _solver(ctx, prob)
    s.set_cg_assembled(True)
    s.set_cg_mode(mode)
    s.set_host_checks(4)
    s.set_warm_start(True)
    for fraction in (0.30, 0.35, 0.40, 0.45):
        aims, ref = _problem(prob, fraction)
        x_ref, e_ref, it_ref = ref.solve(8, stop_rule=True)
        x, e, done = s.solve(T(prob["nodes"]), T(aims), 8, stop_rule=True)
        e, done = H(e), int(H(done)[0])
        assert abs(e[done] - e_ref[-1]) <= 1e-5 * e_ref[-1], (fraction, e[: done + 1], e_ref)
        assert abs(done - it_ref) <= 1, (fraction, done, it_ref)
    s.close()


def test_cg_tol_setter_matches_creation_tolerance(ctx, prob):
    """set_cg_tol after creation solves exactly like a solver created with that
    tolerance (same bits), the drag energies stay within 1e-5 of the oracle's,
    and tolerances outside (0, 1) are refused."""
    def make(tol):
        return arapgs.Solver(ctx, prob["nbr"], prob["blocks"], prob["types"], prob["src_idx"],
                             prob["src_w"].astype(np.float32), prob["nodes"], cg_max_iters=200, cg_tol=tol)

    a, b = make(1e-4), make(1e-5)
    b.set_cg_tol(1e-4)
    for s in (a, b):
        s.set_warm_start(True)
        s.set_host_checks(4)
    for fraction in (0.30, 0.35, 0.40):
        aims, ref = _problem(prob, fraction)
        xa, ea, da = a.solve(T(prob["nodes"]), T(aims), 8, stop_rule=True)
        xb, eb, db = b.solve(T(prob["nodes"]), T(aims), 8, stop_rule=True)
        da, db = int(H(da)[0]), int(H(db)[0])
        assert da == db and np.array_equal(H(xa), H(xb)) and np.array_equal(H(ea), H(eb))
        _, e_ref, _ = ref.solve(8, stop_rule=True)
        assert abs(H(ea)[da] - e_ref[-1]) <= 1e-5 * e_ref[-1], (fraction, H(ea)[: da + 1], e_ref)
    for bad in (0.0, -1e-3, 1.0, float("nan")):
        with pytest.raises(arapgs.ArapgsError):
            b.set_cg_tol(bad)
    a.close()
    b.close()


def test_host_scheduling_keeps_fixed_count_iterates(ctx, prob):
    """Fixed-count solves (stop rule off, cold CG, a reset solver before each,
    as the fixed-iteration bench leg runs them): host checks after every CG
    iteration or every 3, with or without count prediction and polled flags,
    give the same iterates bit for bit."""
    settings = ((1, False, False), (3, False, False), (3, True, False), (3, False, True), (3, True, True))
    runs = []
    for every, predict, spin in settings:
        s = _solver(ctx, prob)
        s.set_warm_start(False)
        s.set_host_checks(every)
        s.set_cg_predict(predict)
        s.set_host_spin(spin)
        out = []
        for fraction in (0.5, 0.6):
            s.reset()
            aims, _ = _problem(prob, fraction)
            x, e, done = s.solve(T(prob["nodes"]), T(aims), 10, stop_rule=False)
            out.append((H(x), H(e), int(H(done)[0]), s.last_stats()))
        runs.append(out)
        s.close()
    for cfg, other in zip(settings[1:], runs[1:]):
        for a, b in zip(runs[0], other):
            assert a[2] == b[2] and a[3] == b[3], (cfg, a[3], b[3])
            assert np.array_equal(a[0], b[0]), cfg
            assert np.array_equal(a[1][: a[2] + 1], b[1][: b[2] + 1]), cfg
