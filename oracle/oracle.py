"""TEST INFRASTRUCTURE — ctypes wrapper of the oracle (oracle/liboracle.so).

The oracle is a CPU restatement of the reference's per-frame path. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module,
and only as the checker / the timed CPU baseline — never as a product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = C.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class GNArgs(C.Structure):
    _fields_ = [("M", C.c_int), ("K", C.c_int), ("node_pos", P), ("nbr", P), ("block_ptr", P),
                ("block_nodes", P), ("block_type", P), ("nblocks", C.c_int), ("src_idx", P), ("src_w", P),
                ("aim", P), ("w_rot", C.c_double), ("w_reg", C.c_double), ("w_con", C.c_double),
                ("wf", C.c_double), ("center", C.c_int)]


class RasterArgs(C.Structure):
    _fields_ = [("P", C.c_int), ("D", C.c_int), ("W", C.c_int), ("H", C.c_int), ("tile", C.c_int),
                ("bg", P), ("means", P), ("shs", P), ("opac", P), ("scales", P), ("scale_mod", C.c_float),
                ("rots", P), ("view", P), ("proj", P), ("campos", P), ("tanfovx", C.c_float),
                ("tanfovy", C.c_float), ("use_rects", C.c_int), ("boxmin", P), ("boxmax", P), ("ortho", C.c_int),
                ("ortho_scale", C.c_float)]


def _declare(L):
    L.orc_knn_weights.argtypes = [P, C.c_int, P, C.c_long, C.c_int, P, P]
    L.orc_graph_edges.argtypes = [P, C.c_int, C.c_int, P]
    L.orc_fps.argtypes = [P, C.c_long, C.c_int, P]
    L.orc_fps.restype = C.c_int
    L.orc_end_points.argtypes = [P, P, P, C.c_long, C.c_float, C.c_float, P]
    L.orc_blend.argtypes = [P, C.c_long, P, P, C.c_int, P, P, P, P]
    L.orc_six_point.argtypes = [C.c_long, P, P, P, C.c_float, C.c_float, P, P, P, P]
    L.orc_node_quats.argtypes = [C.c_int, P, P]
    L.orc_sample_sh.argtypes = [C.c_long, P, P, P, C.c_int, P, P]
    L.orc_polar.argtypes = [P, P, P]
    L.orc_sh_rotate.argtypes = [P, P, C.c_long, C.c_int]
    L.orc_mat_to_quat.argtypes = [P, P]
    L.orc_quat_to_mat.argtypes = [P, P]
    L.orc_gn_sizes.argtypes = [C.POINTER(GNArgs), C.POINTER(C.c_int)]
    L.orc_gn_sizes.restype = C.c_long
    L.orc_gn_residual.argtypes = [C.POINTER(GNArgs), P, P]
    L.orc_gn_jacobian_dense.argtypes = [C.POINTER(GNArgs), P, P]
    L.orc_gn_jacobian_triplets.argtypes = [C.POINTER(GNArgs), P, P, P, P]
    L.orc_gn_jacobian_triplets.restype = C.c_long
    L.orc_gn_solve.argtypes = [C.POINTER(GNArgs), C.c_int, C.c_int, P, P]
    L.orc_gn_solve.restype = C.c_int
    L.orc_gn_reference_cost.argtypes = [C.c_int]
    L.orc_gn_reference_cost.restype = None
    L.orc_scene_aabb.argtypes = [P, C.c_long, C.c_int, P]
    L.orc_grid_reorder.argtypes = [P, C.c_long, P, C.c_float, C.c_int, P]
    L.orc_gaussian_aabbs.argtypes = [P, P, P, P, C.c_long, P]
    L.orc_grid_lists.argtypes = [P, C.c_long, P, C.c_float, C.c_int, C.c_int, P, P, P, P, P]
    L.orc_grid_lists.restype = C.c_long
    L.orc_density_sample.argtypes = [C.c_long, P, P, P, P, P, P, P, P, P, C.c_float, P, P, P]
    L.orc_raster_preprocess.argtypes = [C.POINTER(RasterArgs), P, P, P, P, P, P, P]
    L.orc_chain_jvp.argtypes = [C.c_void_p] + [P] * 19
    L.orc_chain_jvp.restype = None
    L.orc_chain_forward.argtypes = [C.c_void_p] + [P] * 7
    L.orc_chain_forward.restype = None
    L.orc_density_jvp.argtypes = [C.c_long, P, P, P, C.c_long, P, P, P, P, P, P, C.c_float] + [P] * 11
    L.orc_density_jvp.restype = None
    L.orc_raster_bin.argtypes = [C.POINTER(RasterArgs), P, P, P, P, P, P]
    L.orc_raster_bin.restype = C.c_long
    L.orc_raster_blend.argtypes = [C.POINTER(RasterArgs), P, P, P, P, P, P, P]


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# --------------------------------------------------------------------------- graph build
def knn_weights(nodes, queries, k, with_weights=True):
    """Deform.hpp:153-208. Returns idx (Q, k+1) uint32 and w (Q, k) float64."""
    nodes = _f32(nodes).reshape(-1, 3)
    queries = _f32(queries).reshape(-1, 3)
    Q = queries.shape[0]
    idx = np.zeros((Q, k + 1), np.uint32)
    w = np.zeros((Q, k), np.float64) if with_weights else None
    lib().orc_knn_weights(_p(nodes), nodes.shape[0], _p(queries), Q, k, _p(idx), _p(w))
    return idx, w


def graph_edges(nodes, k):
    nodes = _f32(nodes).reshape(-1, 3)
    nbr = np.zeros((nodes.shape[0], k), np.uint32)
    lib().orc_graph_edges(_p(nodes), nodes.shape[0], k, _p(nbr))
    return nbr


def fps(pos, m):
    pos = _f32(pos).reshape(-1, 3)
    # m+1 slots: for m == 1 the reference loop appends a second node
    # (helper.cpp:157-174), see orc_fps.
    out = np.zeros(max(0, min(m, pos.shape[0])) + 1, np.int32)
    n = lib().orc_fps(_p(pos), pos.shape[0], m, _p(out))
    return out[:n]


def end_points(pos, rot, scale, pad=1e-3, coeff=2.0):
    pos, rot, scale = _f32(pos), _f32(rot), _f32(scale)
    N = pos.shape[0]
    ends = np.zeros((N, 6, 3), np.float32)
    lib().orc_end_points(_p(pos), _p(rot), _p(scale), N, pad, coeff, _p(ends))
    return ends


# --------------------------------------------------------------------------- per-frame deform
def blend(pts, idx, w, node_pos, rot9, t3):
    pts = _f32(pts).reshape(-1, 3)
    k = idx.shape[-1]
    idx = _u32(idx).reshape(-1, k)
    w = _f64(w).reshape(-1, k)
    out = np.zeros_like(pts)
    lib().orc_blend(_p(pts), pts.shape[0], _p(idx), _p(w), k, _p(_f32(node_pos)), _p(_f64(rot9)), _p(_f64(t3)),
                    _p(out))
    return out


def six_point(ends, scale0, is_static, pos, rot, scale, sh, pad=1e-3, coeff=2.0):
    """Updates pos/rot/scale/sh copies; returns the new arrays."""
    pos, rot, scale, sh = (np.array(a, dtype=np.float32, copy=True) for a in (pos, rot, scale, sh))
    st = None if is_static is None else np.ascontiguousarray(is_static, dtype=np.uint8)
    lib().orc_six_point(pos.shape[0], _p(_f32(ends)), _p(_f32(scale0)), _p(st), pad, coeff, _p(pos), _p(rot),
                        _p(scale), _p(sh))
    return pos, rot, scale, sh


def node_quats(rot9):
    rot9 = _f64(rot9).reshape(-1, 9)
    q = np.zeros((rot9.shape[0], 4), np.float32)
    lib().orc_node_quats(rot9.shape[0], _p(rot9), _p(q))
    return q


def sample_sh(w_f32, idx, node_q, is_static, feat):
    feat = np.array(feat, dtype=np.float32, copy=True)
    k = idx.shape[-1]
    st = None if is_static is None else np.ascontiguousarray(is_static, dtype=np.uint8)
    lib().orc_sample_sh(feat.shape[0], _p(_f32(w_f32)), _p(_u32(idx)), _p(_f32(node_q)), k, _p(st), _p(feat))
    return feat


def polar(M):
    M = _f64(M).reshape(3, 3)
    R = np.zeros((3, 3)); S = np.zeros((3, 3))
    lib().orc_polar(_p(M), _p(R), _p(S))
    return R, S


def sh_rotate(R, sh, flip_odd=True):
    sh = np.array(sh, dtype=np.float32, copy=True).reshape(-1, 48)
    lib().orc_sh_rotate(_p(_f32(R)), _p(sh), sh.shape[0], int(flip_odd))
    return sh


def mat_to_quat(m):
    q = np.zeros(4, np.float32)
    lib().orc_mat_to_quat(_p(_f32(m)), _p(q))
    return q


def quat_to_mat(q):
    m = np.zeros((3, 3), np.float32)
    lib().orc_quat_to_mat(_p(_f32(q)), _p(m))
    return m


# --------------------------------------------------------------------------- GN solve
class GNProblem:
    """Inputs of Deform::Deform (Deform.hpp:414-455) as flat arrays."""

    def __init__(self, node_pos, nbr, blocks, block_types, src_idx, src_w, aims, w_rot=1.0, w_reg=10.0,
                 w_con=100.0, wf=1.0, center=False):
        self.node_pos = _f32(node_pos).reshape(-1, 3)
        self.M = self.node_pos.shape[0]
        self.nbr = _u32(nbr)
        self.K = self.nbr.shape[1]
        ptr = [0]
        for b in blocks:
            ptr.append(ptr[-1] + len(b))
        self.block_ptr = _i32(ptr)
        self.block_nodes = _i32(np.concatenate([np.asarray(b, np.int32) for b in blocks]) if blocks else [])
        self.block_type = _i32(block_types)
        self.src_idx = _u32(src_idx)
        self.src_w = _f64(src_w)
        self.aims = _f32(aims).reshape(-1, 3)
        self.args = GNArgs(self.M, self.K, _p(self.node_pos), _p(self.nbr), _p(self.block_ptr),
                           _p(self.block_nodes), _p(self.block_type), len(blocks), _p(self.src_idx),
                           _p(self.src_w), _p(self.aims), w_rot, w_reg, w_con, wf, int(center))
        n = C.c_int(0)
        self.rows = lib().orc_gn_sizes(C.byref(self.args), C.byref(n))
        self.n = n.value

    def identity_x(self):
        x = np.zeros(self.n)
        x[0::12] = 1.0
        x[4::12] = 1.0
        x[8::12] = 1.0
        return x

    def residual(self, x):
        f = np.zeros(self.rows)
        lib().orc_gn_residual(C.byref(self.args), _p(_f64(x)), _p(f))
        return f

    def energy(self, x):
        f = self.residual(x)
        return float(f @ f)

    def jacobian(self, x):
        J = np.zeros((self.rows, self.n))
        lib().orc_gn_jacobian_dense(C.byref(self.args), _p(_f64(x)), _p(J))
        return J

    def jacobian_triplets(self, x):
        x = _f64(x)
        n = lib().orc_gn_jacobian_triplets(C.byref(self.args), _p(x), None, None, None)
        r = np.zeros(n, np.int64); c = np.zeros(n, np.int32); v = np.zeros(n)
        lib().orc_gn_jacobian_triplets(C.byref(self.args), _p(x), _p(r), _p(c), _p(v))
        return r, c, v

    def solve(self, iters, stop_rule=True, reference_cost=False):
        """reference_cost: every iteration repeats the reference's ordering +
        two factorisations (Deform.cpp:131-132); same result, for timing."""
        lib().orc_gn_reference_cost(int(bool(reference_cost)))
        x = np.zeros(self.n)
        e = np.zeros(iters + 1)
        it = lib().orc_gn_solve(C.byref(self.args), iters, int(stop_rule), _p(x), _p(e))
        if it < 0:
            raise RuntimeError("oracle GN: factorisation failed")
        return x, e[: it + 1], it


# --------------------------------------------------------------------------- grid + density
def scene_aabb(pos, G):
    out = np.zeros(7, np.float32)
    pos = _f32(pos)
    lib().orc_scene_aabb(_p(pos), pos.shape[0], G, _p(out))
    return out[:3].copy(), out[3:6].copy(), float(out[6])


def grid_reorder(pos, mn, step, G):
    pos = _f32(pos)
    order = np.zeros(pos.shape[0], np.int64)
    lib().orc_grid_reorder(_p(pos), pos.shape[0], _p(_f32(mn)), step, G, _p(order))
    return order


def gaussian_aabbs(pos, rot, scale, opacity):
    N = pos.shape[0]
    out = np.zeros((N, 6), np.float32)
    lib().orc_gaussian_aabbs(_p(_f32(pos)), _p(_f32(rot)), _p(_f32(scale)), _p(_f32(opacity)), N, _p(out))
    return out


def grid_lists(aabbs, mn, step, G, padding=1):
    aabbs = _f32(aabbs)
    N = aabbs.shape[0]
    GC = G ** 3
    prefix = np.zeros(GC, np.int32)
    nv = C.c_long(0)
    total = lib().orc_grid_lists(_p(aabbs), N, _p(_f32(mn)), step, G, padding, _p(prefix), None, None,
                                 C.byref(nv), None)
    lst = np.zeros(total, np.int32)
    valid = np.zeros(nv.value, np.int32)
    samples = np.zeros((nv.value * 64, 3), np.float32)
    lib().orc_grid_lists(_p(aabbs), N, _p(_f32(mn)), step, G, padding, _p(prefix), _p(lst), _p(valid),
                         C.byref(nv), _p(samples))
    return prefix, lst, valid, samples


def density_sample(valid, prefix, lst, samples, pos, rot, scale, opacity, sh, lpf, ada=None):
    """ada: per-cell adaptive filters (n_valid, 6) as (xx, yy, zz, xy, xz, yz), or None for lpf I."""
    nv = valid.shape[0]
    feat = np.zeros((nv * 64, 48), np.float32)
    opac = np.zeros(nv * 64, np.float32)
    ada = None if ada is None else _f32(ada)
    lib().orc_density_sample(nv, _p(_i32(valid)), _p(_i32(prefix)), _p(_i32(lst)), _p(_f32(samples)),
                             _p(_f32(pos)), _p(_f32(rot)), _p(_f32(scale)), _p(_f32(opacity)), _p(_f32(sh)),
                             lpf, _p(feat), _p(opac), _p(ada))
    return feat, opac


def ada_lpf(samples, lpf_param):
    """GetAdaLpfRatio (GaussianView.cpp:4670-4749) restated in FP64 with the
    reference's own construction: per valid cell, the 24 x 9 system A x = b over
    the 8 cube corners (A(3i+j, 3j+k) = Q(k, i), b(3i+j) = P(j, i), P the corner
    samples minus the first, divided in float by the 3 intervals), solved by
    its normal equations; M(i, j) = x(3i + j); returns lpf_param * M M^T
    packed (n_valid, 6) as (xx, yy, zz, xy, xz, yz)."""
    x = np.asarray(samples, np.float32).reshape(-1, 64, 3)
    nv = x.shape[0]
    out = np.zeros((nv, 6))
    for v in range(nv):
        base = x[v, 0]
        A = np.zeros((24, 9))
        b = np.zeros(24)
        row = 0
        for col in range(8):
            qv = np.array([col % 2, (col // 2) % 2, (col // 4) % 2], np.float64)
            idx = 48 * int(qv[0]) + 12 * int(qv[1]) + 3 * int(qv[2])
            pcol = ((x[v, idx] - base) / np.float32(3.0)).astype(np.float64)
            for j in range(3):
                A[row, 3 * j:3 * j + 3] = qv
                b[row] = pcol[j]
                row += 1
        sol = np.linalg.solve(A.T @ A, A.T @ b)
        M = sol.reshape(3, 3)
        L = lpf_param * (M @ M.T)
        out[v] = (L[0, 0], L[1, 1], L[2, 2], L[0, 1], L[0, 2], L[1, 2])
    return out


# --------------------------------------------------------------------------- rasteriser
def raster(means, shs, opac, scales, rots, view, proj, campos, tanfovx, tanfovy, W, H, bg, D=3, tile=16,
           scale_mod=1.0, use_rects=True, boxmin=None, boxmax=None, want_surface=False, ortho_scale=None):
    """ortho_scale: an orthographic camera (proj = orthographic(s, s) @ view) when given."""
    """Returns dict with image (3,H,W), keys, vals, ranges and preprocess outputs."""
    means, shs, opac, scales, rots = _f32(means), _f32(shs), _f32(opac), _f32(scales), _f32(rots)
    view, proj, campos, bg = _f32(view).reshape(16), _f32(proj).reshape(16), _f32(campos), _f32(bg)
    bmin = None if boxmin is None else _f32(boxmin)
    bmax = None if boxmax is None else _f32(boxmax)
    Pn = means.shape[0]
    a = RasterArgs(Pn, D, W, H, tile, _p(bg), _p(means), _p(shs), _p(opac), _p(scales), scale_mod, _p(rots),
                   _p(view), _p(proj), _p(campos), tanfovx, tanfovy, int(use_rects), _p(bmin), _p(bmax),
                   int(ortho_scale is not None), float(ortho_scale or 1.0))
    depth = np.zeros(Pn, np.float32); radius = np.zeros(Pn, np.int32); xy = np.zeros((Pn, 2), np.float32)
    conic = np.zeros((Pn, 4), np.float32); rgb = np.zeros((Pn, 3), np.float32)
    rect = np.zeros((Pn, 4), np.int32); touched = np.zeros(Pn, np.int32)
    lib().orc_raster_preprocess(C.byref(a), _p(depth), _p(radius), _p(xy), _p(conic), _p(rgb), _p(rect),
                                _p(touched))
    n = int(touched.astype(np.int64).sum())
    keys = np.zeros(n, np.uint64); vals = np.zeros(n, np.uint32)
    gx = (W + tile - 1) // tile; gy = (H + tile - 1) // tile
    ranges = np.zeros((gx * gy, 2), np.int32)
    lib().orc_raster_bin(C.byref(a), _p(depth), _p(rect), _p(touched), _p(keys), _p(vals), _p(ranges))
    img = np.zeros((3, H, W), np.float32)
    surface = np.zeros(Pn, np.int32) if want_surface else None
    lib().orc_raster_blend(C.byref(a), _p(vals), _p(ranges), _p(xy), _p(conic), _p(rgb), _p(img), _p(surface))
    return dict(image=img, keys=keys, vals=vals, ranges=ranges, depth=depth, radius=radius, xy=xy, conic=conic,
                rgb=rgb, rect=rect, touched=touched, surface=surface)


# --------------------------------------------------------------------------- adjoint chain
class ChainArgs(C.Structure):
    _fields_ = [("N", C.c_long), ("K", C.c_int), ("M", C.c_int), ("pos0", P), ("sh0", P), ("opac", P),
                ("ends_idx", P), ("ends_w", P), ("node_g", P), ("nv", C.c_long), ("valid", P), ("prefix", P),
                ("list", P), ("samples0", P), ("sample_idx", P), ("sample_w", P), ("lpf", C.c_float),
                ("V", C.c_int), ("W", C.c_int), ("H", C.c_int), ("D", C.c_int), ("bg", P), ("view", P),
                ("proj", P), ("campos", P), ("tanfov", P), ("pad", C.c_float), ("coeff", C.c_float), ("ada", P)]


class Chain:
    """FP64 forward and forward-mode directional derivative of the deform ->
    density-sample -> render chain (oracle/chain.cpp). cams: list of
    (view16, proj16, campos3, tanfovx, tanfovy); grid: (valid, prefix, list,
    samples0) or None; ada: per-cell adaptive filters (n_valid, 6) or None
    for lpf I."""

    def __init__(self, pos0, sh0, opac, ends_idx, ends_w, node_g, cams, W, H, bg, D=3, grid=None,
                 sample_idx=None, sample_w=None, lpf=0.0, pad=1e-3, coeff=2.0, ada=None):
        self.N = pos0.shape[0]
        self.M = node_g.shape[0]
        self.K = np.asarray(ends_idx).size // (6 * self.N)
        self.V, self.W, self.H = len(cams), W, H
        k = dict(pos0=_f32(pos0), sh0=_f32(sh0), opac=_f32(opac), ends_idx=_u32(ends_idx), ends_w=_f32(ends_w),
                 node_g=_f32(node_g), bg=_f32(bg),
                 view=_f32(np.stack([np.asarray(c[0], np.float32).reshape(16) for c in cams])),
                 proj=_f32(np.stack([np.asarray(c[1], np.float32).reshape(16) for c in cams])),
                 campos=_f32(np.stack([np.asarray(c[2], np.float32).reshape(3) for c in cams])),
                 tanfov=_f32(np.array([[c[3], c[4]] for c in cams], np.float32)))
        nv = 0
        if grid is not None:
            valid, prefix, lst, samples0 = grid
            k.update(valid=_i32(valid), prefix=_i32(prefix), list=_i32(lst), samples0=_f32(samples0),
                     sample_idx=_u32(sample_idx), sample_w=_f32(sample_w))
            nv = k["valid"].shape[0]
        else:
            k.update(valid=None, prefix=None, list=None, samples0=None, sample_idx=None, sample_w=None)
        self.nv, self.S = nv, 64 * nv
        k["ada"] = None if ada is None else _f32(ada)
        self._k = k
        self.args = ChainArgs(self.N, self.K, self.M, _p(k["pos0"]), _p(k["sh0"]), _p(k["opac"]),
                              _p(k["ends_idx"]), _p(k["ends_w"]), _p(k["node_g"]), nv, _p(k["valid"]),
                              _p(k["prefix"]), _p(k["list"]), _p(k["samples0"]), _p(k["sample_idx"]),
                              _p(k["sample_w"]), lpf, self.V, W, H, D, _p(k["bg"]), _p(k["view"]), _p(k["proj"]),
                              _p(k["campos"]), _p(k["tanfov"]), pad, coeff, _p(k["ada"]))

    def jvp(self, R, t, q0, s0, dR=None, dt=None, dq0=None, ds0=None):
        L = lib()
        V, H, W, S, N = self.V, self.H, self.W, self.S, self.N
        out = dict(images=np.zeros((V, 3, H, W)), feat=np.zeros((S, 48)), dens=np.zeros(S),
                   d_images=np.zeros((V, 3, H, W)), d_feat=np.zeros((S, 48)), d_dens=np.zeros(S),
                   pos=np.zeros((N, 3)), rot=np.zeros((N, 4)), scale=np.zeros((N, 3)), sh=np.zeros((N, 48)),
                   samples=np.zeros((S, 3)))
        d = [None if x is None else _f64(x) for x in (dR, dt, dq0, ds0)]
        L.orc_chain_jvp(C.byref(self.args), _p(_f64(R)), _p(_f64(t)), _p(_f64(q0)), _p(_f64(s0)),
                        *[_p(x) for x in d], _p(out["images"]), _p(out["feat"]), _p(out["dens"]),
                        _p(out["d_images"]), _p(out["d_feat"]), _p(out["d_dens"]), _p(out["pos"]), _p(out["rot"]),
                        _p(out["scale"]), _p(out["sh"]), _p(out["samples"]))
        return out

    def forward(self, R, t, q0, s0):
        V, H, W, S = self.V, self.H, self.W, self.S
        images, feat, dens = np.zeros((V, 3, H, W)), np.zeros((S, 48)), np.zeros(S)
        lib().orc_chain_forward(C.byref(self.args), _p(_f64(R)), _p(_f64(t)), _p(_f64(q0)), _p(_f64(s0)),
                                _p(images), _p(feat), _p(dens))
        return images, feat, dens


def density_jvp(valid, prefix, lst, samples, pos, rot, scale, opac, sh, lpf, dpos=None, drot=None, dscale=None,
                dopac=None, dsh=None, dsamples=None, ada=None):
    """FP64 density sample and its directional derivative along the given
    directions of the Gaussian attributes and sample positions (oracle/chain.cpp)."""
    nv = np.asarray(valid).shape[0]
    N = np.asarray(pos).shape[0]
    S = 64 * nv
    feat, dens = np.zeros((S, 48)), np.zeros(S)
    d_feat, d_dens = np.zeros((S, 48)), np.zeros(S)
    keep = [_i32(valid), _i32(prefix), _i32(lst), _f32(samples), _f32(pos), _f32(rot), _f32(scale), _f32(opac),
            _f32(sh)]
    dirs = [None if d is None else _f64(d) for d in (dpos, drot, dscale, dopac, dsh, dsamples)]
    lib().orc_density_jvp(nv, _p(keep[0]), _p(keep[1]), _p(keep[2]), N, _p(keep[3]), _p(keep[4]), _p(keep[5]),
                          _p(keep[6]), _p(keep[7]), _p(keep[8]), lpf, *[_p(d) for d in dirs], _p(feat), _p(dens),
                          _p(d_feat), _p(d_dens), _p(None if ada is None else _f32(ada)))
    return feat, dens, d_feat, d_dens


# ---------------------------------------------------------------- PLY codec
PLY_PROPS = (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"] + [f"f_rest_{i}" for i in range(45)] +
             ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])


def ply_write(path, recs):
    """Writes float32 records (n x 62, or n x 63 with a trailing index) as a
    binary little-endian PLY with the 3DGS header (savePly's header,
    GaussianView.cpp:296-307)."""
    recs = np.ascontiguousarray(recs, dtype="<f4")
    props = PLY_PROPS + (["index"] if recs.shape[1] == 63 else [])
    head = f"ply\nformat binary_little_endian 1.0\nelement vertex {recs.shape[0]}\n"
    head += "".join(f"property float {p}\n" for p in props) + "end_header\n"
    with open(path, "wb") as f:
        f.write(head.encode())
        f.write(recs.tobytes())


def ply_read(path):
    """(records float32 (n x stride), stride) of a 3DGS PLY; the header walk of
    loadPly (GaussianView.cpp:53-76) with the property list checked."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    lines = data[:end].decode().splitlines()
    n = int(lines[2].split()[2])
    props = [ln.split()[2] for ln in lines if ln.startswith("property float ")]
    assert props[:62] == PLY_PROPS and len(props) in (62, 63)
    stride = len(props)
    return np.frombuffer(data, dtype="<f4", count=n * stride, offset=end).reshape(n, stride), stride


def morton_codes(pos):
    """63-bit Morton codes of float32 positions over their bounding box, 21 bits
    per axis of trunc((2^21 - 1) * (p - min) / (max - min)), x in bit 3b, y in
    3b + 1, z in 3b + 2 (GaussianView.cpp:205-225). Returns (codes, min, max)."""
    pos = np.asarray(pos, np.float32)
    mn, mx = pos.min(0), pos.max(0)
    with np.errstate(invalid="ignore", divide="ignore"):
        rel = (pos - mn) / (mx - mn)
        scaled = np.float32((1 << 21) - 1) * rel
    xyz = np.where(np.isfinite(scaled), scaled, 0).astype(np.int64)
    code = np.zeros(pos.shape[0], np.uint64)
    for b in range(21):
        for a in range(3):
            code |= ((xyz[:, a] >> b) & 1).astype(np.uint64) << np.uint64(3 * b + a)
    return code, mn, mx


def ply_load(recs, morton=True):
    """Activated SoA of PLY records (loadPly, GaussianView.cpp:226-266): Morton
    order (ties in file order), unit quaternion, exp(scale), sigmoid(opacity),
    SH interleaved as coefficient * 3 + channel; index: the `index` property
    rounded half away from zero, or the file position."""
    recs = np.asarray(recs, np.float32)
    n, stride = recs.shape
    if morton and n:
        code, mn, mx = morton_codes(recs[:, 0:3])
        order = np.argsort(code, kind="stable")
        box = np.concatenate([mn, mx])
    else:
        order = np.arange(n)
        box = np.zeros(6, np.float32)
    r = recs[order]
    q = r[:, 58:62]
    length = np.sqrt((q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1]) + q[:, 2] * q[:, 2] + q[:, 3] * q[:, 3])
    sh = np.zeros((n, 16, 3), np.float32)
    sh[:, 0] = r[:, 6:9]
    for c in range(3):
        sh[:, 1:, c] = r[:, 9 + 15 * c: 24 + 15 * c]
    if stride == 63:
        index = (np.sign(r[:, 62]) * np.floor(np.abs(r[:, 62]) + 0.5)).astype(np.int32)
    else:
        index = order.astype(np.int32)
    return dict(pos=r[:, 0:3].copy(), rot=q / length[:, None], scale=np.exp(r[:, 55:58]),
                opacity=(1.0 / (1.0 + np.exp(-r[:, 54]))).astype(np.float32), sh=sh.reshape(n, 48), index=index,
                box=box)


def ply_save_records(pos, rot, scale, opacity, sh, box=None, too_small=None):
    """62-float records savePly writes (GaussianView.cpp:270-339): the Gaussians
    inside the box (inclusive) and not too small, in order, normals 0,
    log(scale), log(o / (1 - o)), SH back to channel-major."""
    pos = np.asarray(pos, np.float32)
    keep = np.ones(pos.shape[0], bool)
    if box is not None:
        lo, hi = np.asarray(box[0], np.float32), np.asarray(box[1], np.float32)
        keep &= ~((pos < lo) | (pos > hi)).any(1)
    if too_small is not None:
        keep &= ~np.asarray(too_small, bool)
    sh = np.asarray(sh, np.float32).reshape(-1, 16, 3)[keep]
    o = np.asarray(opacity, np.float32)[keep]
    m = int(keep.sum())
    r = np.zeros((m, 62), np.float32)
    r[:, 0:3] = pos[keep]
    r[:, 6:9] = sh[:, 0]
    for c in range(3):
        r[:, 9 + 15 * c: 24 + 15 * c] = sh[:, 1:, c]
    r[:, 54] = np.log(o / (np.float32(1.0) - o))
    r[:, 55:58] = np.log(np.asarray(scale, np.float32)[keep])
    r[:, 58:62] = np.asarray(rot, np.float32)[keep]
    return r
