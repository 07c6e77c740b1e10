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
                ("tanfovy", C.c_float), ("use_rects", C.c_int), ("boxmin", P), ("boxmax", P)]


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
    L.orc_scene_aabb.argtypes = [P, C.c_long, C.c_int, P]
    L.orc_grid_reorder.argtypes = [P, C.c_long, P, C.c_float, C.c_int, P]
    L.orc_gaussian_aabbs.argtypes = [P, P, P, P, C.c_long, P]
    L.orc_grid_lists.argtypes = [P, C.c_long, P, C.c_float, C.c_int, C.c_int, P, P, P, P, P]
    L.orc_grid_lists.restype = C.c_long
    L.orc_density_sample.argtypes = [C.c_long, P, P, P, P, P, P, P, P, P, C.c_float, P, P]
    L.orc_raster_preprocess.argtypes = [C.POINTER(RasterArgs), P, P, P, P, P, P, P]
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

    def solve(self, iters, stop_rule=True):
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


def density_sample(valid, prefix, lst, samples, pos, rot, scale, opacity, sh, lpf):
    nv = valid.shape[0]
    feat = np.zeros((nv * 64, 48), np.float32)
    opac = np.zeros(nv * 64, np.float32)
    lib().orc_density_sample(nv, _p(_i32(valid)), _p(_i32(prefix)), _p(_i32(lst)), _p(_f32(samples)),
                             _p(_f32(pos)), _p(_f32(rot)), _p(_f32(scale)), _p(_f32(opacity)), _p(_f32(sh)),
                             lpf, _p(feat), _p(opac))
    return feat, opac


# --------------------------------------------------------------------------- rasteriser
def raster(means, shs, opac, scales, rots, view, proj, campos, tanfovx, tanfovy, W, H, bg, D=3, tile=16,
           scale_mod=1.0, use_rects=True, boxmin=None, boxmax=None, want_surface=False):
    """Returns dict with image (3,H,W), keys, vals, ranges and preprocess outputs."""
    means, shs, opac, scales, rots = _f32(means), _f32(shs), _f32(opac), _f32(scales), _f32(rots)
    view, proj, campos, bg = _f32(view).reshape(16), _f32(proj).reshape(16), _f32(campos), _f32(bg)
    bmin = None if boxmin is None else _f32(boxmin)
    bmax = None if boxmax is None else _f32(boxmax)
    Pn = means.shape[0]
    a = RasterArgs(Pn, D, W, H, tile, _p(bg), _p(means), _p(shs), _p(opac), _p(scales), scale_mod, _p(rots),
                   _p(view), _p(proj), _p(campos), tanfovx, tanfovy, int(use_rects), _p(bmin), _p(bmax))
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
