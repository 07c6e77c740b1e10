"""Small seeded scenes shared by the CPU and GPU tests (built with the oracle)."""
import numpy as np

import oracle as O
from arapgs_amd import scenes as S


def small_scene(n=20000, n_nodes=128, K=8, seed=0, kind="bar"):
    sc = S.bar_scene(seed, n=n, n_nodes=n_nodes) if kind == "bar" else S.capsule_scene(seed, n=n, n_nodes=n_nodes)
    g = sc.gaussians
    mn, mx, step = O.scene_aabb(g.pos, 64)
    g = g.take(O.grid_reorder(g.pos, mn, step, 64))
    node_idx = O.fps(g.pos, n_nodes)
    nodes = g.pos[node_idx].copy()
    nbr = O.graph_edges(nodes, K)
    nidx, nw = O.knn_weights(nodes, nodes, K)
    blocks, types = S.blocks_for(sc, nodes)
    ends = O.end_points(g.pos, g.rot, g.scale)
    eidx, ew = O.knn_weights(nodes, ends.reshape(-1, 3), K)
    return dict(scene=sc, g=g, nodes=nodes, node_idx=node_idx, nbr=nbr, src_idx=nidx[:, :K], src_w=nw,
                blocks=blocks, types=types, ends=ends, ends_idx=eidx[:, :K], ends_w=ew, K=K)


def random_node_xf(M, rng, rot_deg=10.0, trans=0.05):
    """Random near-rigid node transforms: rot9 column-major (M,9) and t (M,3)."""
    rot9 = np.zeros((M, 9))
    for i in range(M):
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        a = np.radians(rot_deg) * rng.normal()
        Kx = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        R = np.eye(3) + np.sin(a) * Kx + (1 - np.cos(a)) * Kx @ Kx
        R = R @ (np.eye(3) + 0.02 * rng.normal(size=(3, 3)))  # not exactly orthogonal, like GN output
        rot9[i] = R.T.reshape(9)  # column-major
    return rot9, trans * rng.normal(size=(M, 3))


def table_from(nodes, rot9, t3):
    """arapgs_node_xf rows (g, R row-major, t, pad) as float32 (M,16)."""
    M = nodes.shape[0]
    T = np.zeros((M, 16), np.float32)
    T[:, 0:3] = nodes
    T[:, 3:12] = rot9.reshape(M, 3, 3).transpose(0, 2, 1).reshape(M, 9)  # column-major -> row-major
    T[:, 12:15] = t3
    return T


def chain_scene(seed=0, n=400, n_nodes=16, K=4, W=64, H=48, n_views=2, G=8, log_scale=-3.0):
    """A small scene for the differentiable chain: Gaussians in a cube, FPS
    nodes, end-point and grid-sample KNN, grid cell lists (all from the oracle),
    cameras on a ring, and a perturbed node table (R near I, small t)."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.5, 0.5, size=(n, 3))
    g = S.random_gaussians(rng, pos, log_scale_mean=log_scale, log_scale_std=0.3)
    mn, mx, step = O.scene_aabb(g.pos, G)
    g = g.take(O.grid_reorder(g.pos, mn, step, G))
    nodes = g.pos[O.fps(g.pos, n_nodes)].copy()
    ends = O.end_points(g.pos, g.rot, g.scale)
    eidx, ew = O.knn_weights(nodes, ends.reshape(-1, 3), K)
    aabbs = O.gaussian_aabbs(g.pos, g.rot, g.scale, g.opacity)
    prefix, lst, valid, samples = O.grid_lists(aabbs, mn, step, G, padding=1)
    sidx, sw = O.knn_weights(nodes, samples, K)
    lpf = float(np.float32(step / 4) ** 2 * np.float32(0.2))
    cams = []
    for v in range(n_views):
        az = 2 * np.pi * v / n_views + 0.3
        eye = (2.2 * np.cos(az), 2.2 * np.sin(az), 0.6)
        cams.append(S.look_at_camera(eye, (0.0, 0.0, 0.0), 50.0, W, H))
    R = np.tile(np.eye(3), (n_nodes, 1, 1)) + 0.05 * rng.normal(size=(n_nodes, 3, 3))
    t = 0.03 * rng.normal(size=(n_nodes, 3))
    return dict(g=g, nodes=nodes, K=K, ends_idx=eidx[:, :K].copy(), ends_w=ew, prefix=prefix, list=lst, valid=valid,
                samples=samples, sample_idx=sidx[:, :K].copy(), sample_w=sw, lpf=lpf, cams=cams, W=W, H=H,
                bg=np.array([0.8, 0.8, 0.8], np.float32), R=R.reshape(n_nodes, 9), t=t)


def oracle_chain(sc, with_grid=True):
    g = sc["g"]
    cams = [(c.view, c.proj, c.campos, c.tanfovx, c.tanfovy) for c in sc["cams"]]
    grid = (sc["valid"], sc["prefix"], sc["list"], sc["samples"]) if with_grid else None
    return O.Chain(g.pos, g.sh, g.opacity, sc["ends_idx"], sc["ends_w"], sc["nodes"], cams, sc["W"], sc["H"],
                   sc["bg"], grid=grid, sample_idx=sc["sample_idx"] if with_grid else None,
                   sample_w=sc["sample_w"] if with_grid else None, lpf=sc["lpf"])
