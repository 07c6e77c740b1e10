"""Oracle Gauss-Newton (Deform.cpp:95-581, Deform.hpp:414-455) pinned on the CPU.

  residual   an independent numpy restatement of CalcEnergyFunc (Deform.cpp:378-581):
             rotation rows, regularisation rows with the same-block weighting factor
             (BelongsToSameBlock, Deform.cpp:20-32), static->free rows, constraint rows
             through each node's source-vertex weights; weights enter as square roots
  Jacobian   FP64 central differences of that residual, except the reference's
             d reg / d t_i entry which carries no weighting factor (Deform.cpp:257 vs :261)
  GN         an independent dense-numpy loop with the reference's step halving
             (alpha = 1 .. 2^-49) and stop rule |h| < (|x| + 1e-6) 1e-6
  invariant  zero energy at identity when the aims are the current node positions
  fixture    the stripes graph (200 nodes) with the node blocks of LoadDeformScript0
  centre     the constraints-on-centre mode (Deform.cpp:34-75, 483-529): the same
             restatement with one 3-row group per control block over its first
             min(20, |block|) nodes; central differences, the dense GN loop, the
             identity "centre rows = sum of the plain rows of the block's nodes"
             (blocks of <= 20 nodes, up to the float rounding of the centre) and
             the reference's first-20 selection (the 21st node's aim is ignored)
"""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class Ref:
    """Independent restatement of the reference energy."""

    def __init__(self, nodes, nbr, blocks, types, src_idx, src_w, aims, w_rot=1.0, w_reg=10.0, w_con=100.0,
                 wf=1.0, center=False):
        self.center = center
        self.aims32 = np.asarray(aims, np.float32)
        self.g = nodes.astype(np.float64)
        self.nbr, self.blocks, self.types = nbr, blocks, list(types)
        self.src_idx, self.src_w, self.aims = src_idx, src_w, aims.astype(np.float64)
        self.sr, self.sg, self.sc, self.wf = np.sqrt(w_rot), np.sqrt(w_reg), np.sqrt(w_con), wf
        M = nodes.shape[0]
        static = sorted({int(n) for b, t in zip(blocks, types) if t < 0 for n in b})
        self.free = [i for i in range(M) if i not in static]
        self.static = static
        self.fid = {n: f for f, n in enumerate(self.free)}
        self.n = 12 * len(self.free)

    def same_block(self, a, b):
        return any(t != -1 and a in blk and b in blk for blk, t in zip(self.blocks, self.types))

    def R(self, x, f):
        return x[12 * f: 12 * f + 9].reshape(3, 3).T   # mat(r, c) = x[3c + r]

    def t(self, x, f):
        return x[12 * f + 9: 12 * f + 12]

    def residual(self, x):
        out = []
        for f in range(len(self.free)):
            c = self.R(x, f).T                          # rows of c = columns of R
            out += [self.sr * c[0] @ c[1], self.sr * c[0] @ c[2], self.sr * c[1] @ c[2],
                    self.sr * (c[0] @ c[0] - 1), self.sr * (c[1] @ c[1] - 1), self.sr * (c[2] @ c[2] - 1)]
        K = self.nbr.shape[1]
        for f, i in enumerate(self.free):
            for tt in range(K):
                j = int(self.nbr[i, tt])
                tk = self.t(x, self.fid[j]) if j in self.fid else np.zeros(3)
                ws = self.wf if self.same_block(i, j) else 1.0
                v = self.R(x, f) @ (self.g[j] - self.g[i]) + self.g[i] + self.t(x, f) - self.g[j] - tk
                out += list(self.sg * v * ws)
        for i in self.static:
            for tt in range(K):
                j = int(self.nbr[i, tt])
                tk = self.t(x, self.fid[j]) if j in self.fid else np.zeros(3)
                out += list(-self.sg * tk)
        for blk, ty in zip(self.blocks, self.types):
            if ty == -1:
                continue
            if self.center:
                cnt = min(20, len(blk))
                ctr = np.zeros(3, np.float32)
                for c in blk[:cnt]:
                    ctr = (ctr + self.aims32[int(c)]).astype(np.float32)
                ctr = (ctr / np.float32(cnt)).astype(np.float64)
                acc = np.zeros(3)
                for c in sorted({int(c) for c in blk[:cnt]}):
                    acc += self.sc * (self.blend(x, c) - ctr)
                out += list(acc)
                continue
            for c in blk:
                out += list(self.sc * (self.blend(x, c) - self.aims[c]))
        return np.array(out)

    def blend(self, x, c):
        K = self.nbr.shape[1]
        ve = self.g[c]
        v = np.zeros(3)
        for jj in range(K):
            nj = int(self.src_idx[c, jj])
            w = float(self.src_w[c, jj])
            if nj in self.fid:
                f = self.fid[nj]
                v += w * (self.R(x, f) @ (ve - self.g[nj]) + self.g[nj] + self.t(x, f))
            else:
                v += w * ve
        return v

    def energy(self, x):
        f = self.residual(x)
        return float(f @ f)

    def quirk_columns(self):
        """(row, col) entries where the reference drops the weighting factor."""
        rows = []
        base = 6 * len(self.free)
        K = self.nbr.shape[1]
        for f, i in enumerate(self.free):
            for tt in range(K):
                j = int(self.nbr[i, tt])
                if self.same_block(i, j):
                    r0 = base + 3 * (f * K + tt)
                    rows += [(r0 + a, 12 * f + 9 + a) for a in range(3)]
        return rows


def tiny(seed=0, M=24, K=4, wf=1.0, aims_at_nodes=False, center=False, blocks=None, types=None):
    rng = np.random.default_rng(seed)
    nodes = rng.uniform(-1, 1, size=(M, 3)).astype(np.float32)
    nbr = O.graph_edges(nodes, K)
    src = nodes + rng.normal(0, 0.05, size=nodes.shape).astype(np.float32)
    sidx, sw = O.knn_weights(nodes, src, K)
    if blocks is None:
        blocks = [np.arange(0, 6), np.arange(6, 12), np.arange(12, 15)]
        types = [0, 1, -1]
    aims = nodes.astype(np.float64).copy()
    if not aims_at_nodes:
        aims[blocks[1]] += np.array([0.2, -0.1, 0.3])
    aims = aims.astype(np.float32)          # aims are float positions in the reference
    args = (nodes, nbr, blocks, types, sidx[:, :K], sw, aims)
    return O.GNProblem(*args, wf=wf, center=center), Ref(*args, wf=wf, center=center)


@pytest.mark.parametrize("wf", [1.0, 3.0])
def test_residual_matches_restatement(wf):
    prob, ref = tiny(wf=wf)
    assert prob.n == ref.n
    rng = np.random.default_rng(1)
    for _ in range(3):
        x = prob.identity_x() + 0.1 * rng.normal(size=prob.n)
        np.testing.assert_allclose(prob.residual(x), ref.residual(x), rtol=0, atol=1e-12)


@pytest.mark.parametrize("wf", [1.0, 3.0])
def test_jacobian_central_differences_and_quirk(wf):
    prob, ref = tiny(seed=2, wf=wf)
    rng = np.random.default_rng(3)
    x = prob.identity_x() + 0.1 * rng.normal(size=prob.n)
    eps = 1e-6
    J_fd = np.zeros((len(ref.residual(x)), prob.n))
    for c in range(prob.n):
        e = np.zeros(prob.n)
        e[c] = eps
        J_fd[:, c] = (ref.residual(x + e) - ref.residual(x - e)) / (2 * eps)
    for r, c in ref.quirk_columns():
        J_fd[r, c] /= wf
    J = prob.jacobian(x)
    np.testing.assert_allclose(J, J_fd, atol=2e-6)
    # the triplet form assembles to the same matrix
    r, c, v = prob.jacobian_triplets(x)
    Jt = np.zeros_like(J)
    np.add.at(Jt, (r, c), v)
    np.testing.assert_allclose(Jt, J, atol=1e-12)


def gn_numpy(prob, ref, iters):
    x = prob.identity_x()
    energies = []
    for it in range(iters):
        f = ref.residual(x)
        J = prob.jacobian(x)
        h = np.linalg.solve(J.T @ J, -J.T @ f)
        e_old = f @ f
        energies.append(e_old)
        normv = np.linalg.norm(x)
        for _ in range(50):
            if ref.energy(x + h) > e_old:
                h = h * 0.5
            else:
                x = x + h
                break
        if np.linalg.norm(h) < (normv + 1e-6) * 1e-6:
            break
    energies.append(ref.energy(x))
    return x, np.array(energies)


def test_gn_matches_dense_restatement():
    prob, ref = tiny(seed=4)
    x_ref, e_ref = gn_numpy(prob, ref, 8)
    x, e, it = prob.solve(8)
    assert it == len(e_ref) - 1
    np.testing.assert_allclose(e, e_ref, rtol=1e-9, atol=1e-14)
    np.testing.assert_allclose(x, x_ref, atol=1e-8)
    assert np.all(np.diff(e) <= 0)


def test_sparse_path_matches_dense_restatement():
    """Above 1536 unknowns the oracle switches to its minimum-degree block
    Cholesky (with and without the reference's per-iteration re-analysis and
    second factorisation); iterates must equal a dense numpy GN."""
    blocks = [np.arange(0, 20), np.arange(20, 40), np.arange(40, 43)]
    prob, ref = tiny(seed=6, M=160, K=6, blocks=blocks, types=[0, 1, -1])
    assert prob.n > 1536
    x_ref, e_ref = gn_numpy(prob, ref, 4)
    for ref_cost in (False, True):
        x, e, it = prob.solve(4, reference_cost=ref_cost)
        assert it == len(e_ref) - 1
        np.testing.assert_allclose(e, e_ref, rtol=1e-9, atol=1e-14)
        np.testing.assert_allclose(x, x_ref, atol=1e-8)


def test_zero_energy_at_identity_when_aims_are_nodes():
    prob, ref = tiny(seed=5, aims_at_nodes=True)
    x0 = prob.identity_x()
    assert prob.energy(x0) < 1e-24
    assert ref.energy(x0) < 1e-24
    # the rotation rows vanish on SO(3) whatever the translation
    rng = np.random.default_rng(6)
    x = x0.copy()
    for f in range(prob.n // 12):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, a, b, c = q
        R = np.array([[1 - 2 * (b * b + c * c), 2 * (a * b - w * c), 2 * (a * c + w * b)],
                      [2 * (a * b + w * c), 1 - 2 * (a * a + c * c), 2 * (b * c - w * a)],
                      [2 * (a * c - w * b), 2 * (b * c + w * a), 1 - 2 * (a * a + b * b)]])
        x[12 * f: 12 * f + 9] = R.T.reshape(9)
        x[12 * f + 9: 12 * f + 12] = rng.normal(size=3)
    nf = prob.n // 12
    assert np.abs(prob.residual(x)[: 6 * nf]).max() < 1e-12


def test_stripes_fixture_gn():
    with open(os.path.join(GOLDEN, "stripes_graph.json")) as f:
        d = json.load(f)
    nodes = np.array(d["nodes"], np.float32)
    K = 8
    nbr = O.graph_edges(nodes, K)
    sidx, sw = O.knn_weights(nodes, nodes, K)
    blocks = [np.array(d["block1"]), np.array(d["block2"])]
    aims = nodes.astype(np.float64).copy()
    aims[blocks[1]] += np.array([0.0, 0.3, 0.1])
    aims = aims.astype(np.float32)
    prob = O.GNProblem(nodes, nbr, blocks, d["block_types"], sidx[:, :K], sw, aims)
    ref = Ref(nodes, nbr, blocks, d["block_types"], sidx[:, :K], sw, aims)
    assert prob.n == 12 * 200
    x, e, it = prob.solve(6)
    x_ref, e_ref = gn_numpy(prob, ref, 6)
    np.testing.assert_allclose(e, e_ref, rtol=1e-8)
    assert np.all(np.diff(e) <= 0) and e[-1] < 0.5 * e[0]


@pytest.mark.parametrize("wf", [1.0, 3.0])
def test_center_mode_residual_and_jacobian(wf):
    prob, ref = tiny(seed=7, wf=wf, center=True)
    rng = np.random.default_rng(8)
    x = prob.identity_x() + 0.1 * rng.normal(size=prob.n)
    f = prob.residual(x)
    np.testing.assert_allclose(f, ref.residual(x), rtol=0, atol=1e-12)
    nf = prob.n // 12
    assert len(f) == 6 * nf + 3 * (nf * 4 + 3 * 4) + 3 * 2       # two control blocks -> 6 centre rows
    eps = 1e-6
    J_fd = np.zeros((len(f), prob.n))
    for c in range(prob.n):
        e = np.zeros(prob.n)
        e[c] = eps
        J_fd[:, c] = (ref.residual(x + e) - ref.residual(x - e)) / (2 * eps)
    for r, c in ref.quirk_columns():
        J_fd[r, c] /= wf
    np.testing.assert_allclose(prob.jacobian(x), J_fd, atol=2e-6)


def test_center_rows_are_sums_of_plain_rows():
    """Blocks of <= 20 distinct nodes: sum_keys (blend - centre) = sum_keys (blend - aim)."""
    pc, _ = tiny(seed=9, center=True)
    pp, _ = tiny(seed=9, center=False)
    rng = np.random.default_rng(10)
    x = pc.identity_x() + 0.1 * rng.normal(size=pc.n)
    fc, fp = pc.residual(x), pp.residual(x)
    nf = pc.n // 12
    head = 6 * nf + 3 * (nf * 4 + 3 * 4)
    np.testing.assert_array_equal(fc[:head], fp[:head])
    plain = fp[head:].reshape(-1, 3)
    np.testing.assert_allclose(fc[head:head + 3], plain[0:6].sum(0), rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(fc[head + 3:head + 6], plain[6:12].sum(0), rtol=1e-5, atol=1e-5)


def test_center_mode_keeps_the_first_20_block_nodes():
    """SelectKeyControls runs FPS but keeps c_indices[i] (the first 20 nodes):
    the aim of a block's 21st node does not enter the centre residual."""
    blocks = [np.arange(0, 4), np.arange(4, 29)]
    prob, ref = tiny(seed=11, M=40, center=True, blocks=blocks, types=[-1, 1])
    x = prob.identity_x()
    f0 = prob.residual(x)
    np.testing.assert_allclose(f0, ref.residual(x), rtol=0, atol=1e-12)
    aims = prob.aims.copy()
    aims[24] += 1.0                      # the 21st node of block 1
    moved = O.GNProblem(prob.node_pos, prob.nbr, blocks, [-1, 1], prob.src_idx, prob.src_w, aims, center=True)
    np.testing.assert_array_equal(moved.residual(x), f0)
    aims[23] += 1.0                      # the 20th node counts
    moved = O.GNProblem(prob.node_pos, prob.nbr, blocks, [-1, 1], prob.src_idx, prob.src_w, aims, center=True)
    assert np.abs(moved.residual(x) - f0).max() > 1.0


def test_center_mode_gn_matches_dense_restatement():
    prob, ref = tiny(seed=12, center=True)
    x_ref, e_ref = gn_numpy(prob, ref, 8)
    x, e, it = prob.solve(8)
    assert it == len(e_ref) - 1
    np.testing.assert_allclose(e, e_ref, rtol=1e-9, atol=1e-14)
    np.testing.assert_allclose(x, x_ref, atol=1e-8)
