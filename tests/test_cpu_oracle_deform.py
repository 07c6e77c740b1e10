"""Oracle per-frame deform pinned on the CPU by closed forms and invariants.

  GetEndPoints      GaussianView.cpp:4643-4668   e = mu +- R(q) ((s + pad) coeff) e_i
  predict_mesh      Deform.hpp:230-268           identity / translation / global rigid motion
  six-point update  GaussianView.cpp:3081-3166   global rigid motion: mu' = R mu + t, s' = s,
                                                 q' = +-quat(R R(q)), SH rotated by R
  polar (JacobiSVD) helper.cpp:429-440           R = M (M^T M)^-1/2, S = (M^T M)^1/2 (numpy eigh);
                                                 RunTestVKVt input matrices (helper.cpp:869-920)
  SH_Rotation       helper.cpp:938-1075 with the sign wrapper of GaussianView.cpp:3134-3156:
                    rotated coefficients evaluate at d to the original at R^T d (3DGS basis);
                    identity, band norms, composition; Test_SH_Rotation input (helper.cpp:925-928)
  node quaternions  helper.cpp:506-517           Newton polar of A = R S gives quat(R)
  sample slerp      cudakdtree.cu:113-148, 201-222  running slerp closed form for two nodes
"""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

C0 = 0.28209479177387814
C1 = 0.4886025119029199
C2 = [1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396]
C3 = [-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154, -0.4570457994644658,
      1.445305721320277, -0.5900435899266435]


def sh_basis(d):
    """16 real SH basis values of the 3DGS colour evaluator at unit directions d (n,3)."""
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    xx, yy, zz = x * x, y * y, z * z
    return np.stack([np.full_like(x, C0), -C1 * y, C1 * z, -C1 * x,
                     C2[0] * x * y, C2[1] * y * z, C2[2] * (2 * zz - xx - yy), C2[3] * x * z, C2[4] * (xx - yy),
                     C3[0] * y * (3 * xx - yy), C3[1] * x * y * z, C3[2] * y * (4 * zz - xx - yy),
                     C3[3] * z * (2 * zz - 3 * xx - 3 * yy), C3[4] * x * (4 * zz - xx - yy), C3[5] * z * (xx - yy),
                     C3[6] * x * (xx - 3 * yy)], 1)


def sh_eval(sh48, d):
    return sh_basis(d) @ np.asarray(sh48, np.float64).reshape(16, 3)


def rand_rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    return quat_mat(q), q


def quat_mat(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def unit_dirs(rng, n):
    d = rng.normal(size=(n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def test_end_points_closed_form():
    rng = np.random.default_rng(0)
    N = 200
    pos = rng.normal(size=(N, 3)).astype(np.float32)
    rot = rng.normal(size=(N, 4)).astype(np.float32)
    scale = np.exp(rng.normal(-4, 0.5, size=(N, 3))).astype(np.float32)
    e = O.end_points(pos, rot, scale)
    for g in range(N):
        q = rot[g].astype(np.float64) / np.linalg.norm(rot[g].astype(np.float64))
        R = quat_mat(q)
        for i in range(3):
            off = R[:, i] * (scale[g, i] + 1e-3) * 2.0
            np.testing.assert_allclose(e[g, 2 * i], pos[g] + off, atol=2e-6)
            np.testing.assert_allclose(e[g, 2 * i + 1], pos[g] - off, atol=2e-6)


def _knn_setup(rng, P=300, M=40, k=8):
    nodes = rng.uniform(-1, 1, size=(M, 3)).astype(np.float32)
    pts = rng.uniform(-1, 1, size=(P, 3)).astype(np.float32)
    idx, w = O.knn_weights(nodes, pts, k)
    return nodes, pts, idx[:, :k], w


def test_blend_identity_translation_rigid():
    rng = np.random.default_rng(1)
    nodes, pts, idx, w = _knn_setup(rng)
    M = nodes.shape[0]
    eye9 = np.tile(np.eye(3).T.reshape(9), (M, 1))
    out = O.blend(pts, idx, w, nodes, eye9, np.zeros((M, 3)))
    np.testing.assert_allclose(out, pts, atol=1e-6)
    t = np.array([0.3, -0.2, 0.05])
    out = O.blend(pts, idx, w, nodes, eye9, np.tile(t, (M, 1)))
    np.testing.assert_allclose(out, pts + t, atol=1e-6)
    # a global rigid motion x -> R x + tau expressed per node: R_j = R, t_j = tau + R g_j - g_j
    R, _ = rand_rot(rng)
    tau = np.array([0.1, 0.2, -0.3])
    g = nodes.astype(np.float64)
    rot9 = np.tile(R.T.reshape(9), (M, 1))        # column-major
    t3 = tau + g @ R.T - g
    out = O.blend(pts, idx, w, nodes, rot9, t3)
    np.testing.assert_allclose(out, pts.astype(np.float64) @ R.T + tau, atol=2e-6)


def test_six_point_global_rigid_motion():
    rng = np.random.default_rng(2)
    N = 100
    pos = rng.normal(size=(N, 3)).astype(np.float32)
    rot = rng.normal(size=(N, 4)).astype(np.float32)
    rot /= np.linalg.norm(rot, axis=1, keepdims=True)
    scale = np.exp(rng.normal(-3, 0.3, size=(N, 3))).astype(np.float32)
    sh = rng.normal(0, 0.2, size=(N, 48)).astype(np.float32)
    ends = O.end_points(pos, rot, scale).astype(np.float64)
    R, qR = rand_rot(rng)
    tau = np.array([0.2, -0.1, 0.4])
    ends2 = (ends @ R.T + tau).astype(np.float32)
    p2, r2, s2, sh2 = O.six_point(ends2, scale, None, pos, rot, scale, sh)
    np.testing.assert_allclose(p2, pos.astype(np.float64) @ R.T + tau, atol=3e-6)
    np.testing.assert_allclose(s2, scale, rtol=2e-4)
    d = unit_dirs(rng, 64)
    for g in range(N):
        Rg = quat_mat(r2[g].astype(np.float64))
        np.testing.assert_allclose(Rg, R @ quat_mat(rot[g].astype(np.float64)), atol=2e-4)
        # SH follows the rigid rotation: f'(d) = f(R^T d)
        np.testing.assert_allclose(sh_eval(sh2[g], d), sh_eval(sh[g], d @ R), atol=2e-4)


def test_six_point_static_untouched():
    rng = np.random.default_rng(3)
    N = 10
    pos = rng.normal(size=(N, 3)).astype(np.float32)
    rot = np.tile(np.array([1, 0, 0, 0], np.float32), (N, 1))
    scale = np.full((N, 3), 0.01, np.float32)
    sh = rng.normal(size=(N, 48)).astype(np.float32)
    ends = O.end_points(pos, rot, scale) + 0.01
    st = np.zeros(N, np.uint8)
    st[::2] = 1
    p2, r2, s2, sh2 = O.six_point(ends, scale, st, pos, rot, scale, sh)
    assert np.array_equal(p2[::2], pos[::2]) and np.array_equal(sh2[::2], sh[::2])
    assert not np.array_equal(p2[1::2], pos[1::2])


def polar_closed_form(M):
    w, V = np.linalg.eigh(M.T @ M)
    S = V @ np.diag(np.sqrt(w)) @ V.T
    R = M @ V @ np.diag(1 / np.sqrt(w)) @ V.T
    return R, S


def test_polar_vs_eigh_and_reference_matrices():
    rng = np.random.default_rng(4)
    mats = [rng.normal(size=(3, 3)) for _ in range(50)]
    mats += [np.diag([1e3, 1.0, 1e-3]) @ rand_rot(rng)[0] for _ in range(10)]   # flat Gaussians
    with open(os.path.join(GOLDEN, "helper_tests.json")) as f:
        fx = json.load(f)
    for t in fx["vkvt"]:
        mats.append(np.array(t["mat"]).reshape(3, 3))
        mats.append(np.array(t["mat"]).reshape(3, 3) @ np.array(t["R0"]).reshape(3, 3))
    for M in mats:
        R, S = O.polar(M)
        Rc, Sc = polar_closed_form(M)
        np.testing.assert_allclose(R, Rc, atol=1e-9 * max(1.0, np.linalg.cond(M)))
        np.testing.assert_allclose(S, Sc, atol=1e-9 * np.abs(M).max() * max(1.0, np.linalg.cond(M)))
        np.testing.assert_allclose(R @ S, M, atol=1e-10 * np.abs(M).max())
        np.testing.assert_allclose(R.T @ R, np.eye(3), atol=1e-12)


def test_sh_rotation_property_and_reference_input():
    rng = np.random.default_rng(5)
    with open(os.path.join(GOLDEN, "helper_tests.json")) as f:
        fx = json.load(f)["sh_rotation"]
    q = np.array(fx["q_wxyz"])
    q /= np.linalg.norm(q)
    sets = [(quat_mat(q), np.array(fx["shs_in"], np.float32))]
    for _ in range(20):
        sets.append((rand_rot(rng)[0], rng.normal(0, 0.3, size=48).astype(np.float32)))
    d = unit_dirs(rng, 200)
    for R, sh in sets:
        out = O.sh_rotate(R.astype(np.float32), sh)[0]
        np.testing.assert_allclose(sh_eval(out, d), sh_eval(sh, d @ R), atol=3e-6 * (1 + np.abs(sh).max()))
        # the DC term is untouched and every band keeps its norm (per channel)
        assert np.array_equal(out[:3], sh[:3])
        for lo, hi in ((1, 4), (4, 9), (9, 16)):
            a = out.reshape(16, 3)[lo:hi]
            b = sh.reshape(16, 3)[lo:hi]
            np.testing.assert_allclose(np.linalg.norm(a, axis=0), np.linalg.norm(b, axis=0), rtol=1e-5, atol=1e-7)


def test_sh_rotation_identity_and_composition():
    rng = np.random.default_rng(6)
    sh = rng.normal(size=48).astype(np.float32)
    np.testing.assert_allclose(O.sh_rotate(np.eye(3, dtype=np.float32), sh)[0], sh, atol=1e-6)
    R1, _ = rand_rot(rng)
    R2, _ = rand_rot(rng)
    a = O.sh_rotate(R2.astype(np.float32), O.sh_rotate(R1.astype(np.float32), sh)[0])[0]
    b = O.sh_rotate((R2 @ R1).astype(np.float32), sh)[0]
    np.testing.assert_allclose(a, b, atol=2e-5)


def test_quaternion_matrix_round_trip():
    rng = np.random.default_rng(7)
    for _ in range(100):
        R, q = rand_rot(rng)
        m = O.quat_to_mat(q.astype(np.float32))
        np.testing.assert_allclose(m, R, atol=1e-6)
        q2 = O.mat_to_quat(R.astype(np.float32)).astype(np.float64)
        assert min(np.abs(q2 - q).max(), np.abs(q2 + q).max()) < 2e-6


def test_node_quats_newton_polar():
    rng = np.random.default_rng(8)
    M = 50
    rot9 = np.zeros((M, 9))
    qs = []
    for i in range(M):
        R, q = rand_rot(rng)
        A = rng.normal(0, 0.05, size=(3, 3))
        S = np.eye(3) + (A + A.T) / 2
        rot9[i] = (R @ S).T.reshape(9)        # column-major
        qs.append(q)
    out = O.node_quats(rot9).astype(np.float64)
    for i in range(M):
        assert min(np.abs(out[i] - qs[i]).max(), np.abs(out[i] + qs[i]).max()) < 1e-5


def slerp(q0, q1, t):
    c = float(q0 @ q1)
    if c < 0:
        q1, c = -q1, -c
    if c > 0.99995:
        r = (1 - t) * q0 + t * q1
    else:
        th = np.arccos(c)
        r = np.sin((1 - t) * th) / np.sin(th) * q0 + np.sin(t * th) / np.sin(th) * q1
    return r / np.linalg.norm(r)


def test_sample_slerp_two_nodes_closed_form():
    rng = np.random.default_rng(9)
    S = 40
    node_q = np.zeros((2, 4), np.float32)
    node_q[0] = rand_rot(rng)[1]
    node_q[1] = rand_rot(rng)[1]
    w = rng.uniform(0.1, 1, size=(S, 2)).astype(np.float32)
    w /= w.sum(1, keepdims=True)
    idx = np.tile(np.array([0, 1], np.uint32), (S, 1))
    feat = rng.normal(0, 0.3, size=(S, 48)).astype(np.float32)
    out = O.sample_sh(w, idx, node_q, None, feat)
    for s in range(S):
        # step 1: t = 1 -> the first node; step 2: t = w1 / (w0 + w1)
        q = slerp(node_q[0].astype(np.float64), node_q[1].astype(np.float64),
                  float(w[s, 1]) / float(w[s, 0] + w[s, 1]))
        ref = O.sh_rotate(quat_mat(q).astype(np.float32), feat[s])[0]
        np.testing.assert_allclose(out[s], ref, atol=5e-5)
