"""ctypes binding of libarapgs.so (include/arapgs.h) for Python callers.

Device buffers are passed as raw pointers (``tensor.data_ptr()``); PyTorch is
used only to own device memory and streams. There is no CPU fallback: if the
library is missing or no GPU is present, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libarapgs.so")

P = C.c_void_p
_lib = None


class ArapgsError(RuntimeError):
    pass


class FrameBuffers(C.Structure):
    _fields_ = [("n", C.c_int64), ("k", C.c_int), ("m", C.c_int), ("pos", P), ("rot", P), ("scale", P), ("sh", P),
                ("scale0", P), ("ends", P), ("ends_idx", P), ("ends_w", P), ("gs_static", P), ("node_pos", P),
                ("node_src_idx", P), ("node_src_w", P), ("s", C.c_int64), ("samples", P), ("sample_idx", P),
                ("sample_w", P), ("sample_static", P), ("sample_feat", P), ("axis_padding", C.c_float),
                ("end_coeff", C.c_float)]


class SolverDesc(C.Structure):
    _fields_ = [("m", C.c_int), ("k", C.c_int), ("nbr_h", P), ("block_ptr_h", P), ("block_nodes_h", P),
                ("block_type_h", P), ("nblocks", C.c_int), ("src_idx_h", P), ("src_w_h", P), ("node_pos_h", P),
                ("w_rot", C.c_float), ("w_reg", C.c_float), ("w_con", C.c_float), ("weighting_factor", C.c_float),
                ("constraints_on_center", C.c_int), ("n_aggregates", C.c_int), ("cg_max_iters", C.c_int),
                ("cg_tol", C.c_float)]


class RenderDesc(C.Structure):
    _fields_ = [("P", C.c_int), ("D", C.c_int), ("M", C.c_int), ("background", P), ("width", C.c_int),
                ("height", C.c_int), ("means3D", P), ("shs", P), ("colors_precomp", P), ("opacities", P),
                ("scales", P), ("scale_modifier", C.c_float), ("rotations", P), ("cov3D_precomp", P),
                ("viewmatrix", P), ("projmatrix", P), ("cam_pos", P), ("tan_fovx", C.c_float),
                ("tan_fovy", C.c_float), ("prefiltered", C.c_int), ("out_color", P), ("surface", P), ("radii", P),
                ("rects", P), ("boxmin_h", P), ("boxmax_h", P), ("to_ortho", C.c_int), ("ortho_scale", C.c_float),
                ("tile", C.c_int)]


class ChainDesc(C.Structure):
    _fields_ = [("n", C.c_int64), ("k", C.c_int), ("m", C.c_int), ("pos0", P), ("sh0", P), ("opacity", P),
                ("ends_idx", P), ("ends_w", P), ("grid", P), ("sample_idx", P), ("sample_w", P), ("lpf", C.c_float),
                ("n_views", C.c_int), ("width", C.c_int), ("height", C.c_int), ("sh_degree", C.c_int),
                ("background", P), ("viewmatrix_h", P), ("projmatrix_h", P), ("cam_pos_h", P),
                ("tan_fovx_h", P), ("tan_fovy_h", P), ("axis_padding", C.c_float), ("end_coeff", C.c_float)]


class RealignOptions(C.Structure):
    _fields_ = [("moved", P), ("max_scale", P), ("conv_tol", C.c_float)]


class RealignDesc(C.Structure):
    _fields_ = [("lr_pos", C.c_float), ("lr_rot", C.c_float), ("lr_scale", C.c_float), ("lr_opacity", C.c_float),
                ("lr_sh", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("w_density", C.c_float)]


EXPORTS = {
    "arapgs_create": ([C.c_int, C.POINTER(P)], C.c_int),
    "arapgs_destroy": ([P], None),
    "arapgs_last_error": ([P], C.c_char_p),
    "arapgs_set_stream": ([P, P], C.c_int),
    "arapgs_synchronize": ([P], C.c_int),
    "arapgs_set_grid_chunk": ([P, C.c_int64], C.c_int),
    "arapgs_set_node_cache": ([P, C.c_int], C.c_int),
    "arapgs_set_blend_split": ([P, C.c_int], C.c_int),
    "arapgs_stream_create": ([C.c_int, C.c_int, C.c_int, C.POINTER(P)], C.c_int),
    "arapgs_stream_create_mask": ([C.c_int, P, C.c_int, C.POINTER(P)], C.c_int),
    "arapgs_set_sh_compose": ([P, C.c_int], C.c_int),
    "arapgs_set_tile_cull": ([P, C.c_int], C.c_int),
    "arapgs_stream_destroy": ([P], C.c_int),
    "arapgs_version": ([], C.c_char_p),
    "arapgs_fps": ([P, P, C.c_int64, C.c_int, P, C.POINTER(C.c_int)], C.c_int),
    "arapgs_knn": ([P, P, C.c_int, P, C.c_int64, C.c_int, P, P], C.c_int),
    "arapgs_knn_noalloc": ([P, P, C.c_int, P, C.c_int64, C.c_int, P, P], C.c_int),
    "arapgs_graph_edges": ([P, P, C.c_int, C.c_int, P], C.c_int),
    "arapgs_end_points": ([P, P, P, P, C.c_int64, C.c_float, C.c_float, P], C.c_int),
    "arapgs_solver_create": ([P, C.POINTER(SolverDesc), C.POINTER(P)], C.c_int),
    "arapgs_solver_destroy": ([P], None),
    "arapgs_solver_dims": ([P, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "arapgs_gn_solve": ([P, P, P, P, C.c_int, C.c_int, P, P, P], C.c_int),
    "arapgs_gn_energy": ([P, P, P, P, P, P], C.c_int),
    "arapgs_put_free_inputs": ([P, P, P, P, P], C.c_int),
    "arapgs_solver_batch_create": ([P, P, C.c_int, C.POINTER(P)], C.c_int),
    "arapgs_solver_batch_destroy": ([P], None),
    "arapgs_solver_batch_set_host_checks": ([P, C.c_int], C.c_int),
    "arapgs_solver_batch_set_coarse_refresh": ([P, C.c_int], C.c_int),
    "arapgs_solver_batch_set_async_refresh": ([P, C.c_int], C.c_int),
    "arapgs_solver_set_cg_block": ([P, C.c_int], C.c_int),
    "arapgs_solver_set_async_refresh": ([P, C.c_int], C.c_int),
    "arapgs_solver_set_cg_graphs": ([P, C.c_int], C.c_int),
    "arapgs_solver_batch_set_cg_graphs": ([P, C.c_int], C.c_int),
    "arapgs_solver_batch_solve": ([P, P, P, P, C.c_int, C.c_int, P, P, P, P], C.c_int),
    "arapgs_blend_points": ([P, P, C.c_int64, P, P, C.c_int, P, C.c_int, P, P], C.c_int),
    "arapgs_six_point_update": ([P, C.c_int64, C.c_int, P, C.c_int, P, P, P, P, P, C.c_float, C.c_float, P, P, P, P],
                                C.c_int),
    "arapgs_six_point_update_from": ([P, C.c_int64, C.c_int, P, C.c_int, P, P, P, P, P, C.c_float, C.c_float,
                                      P, P, P, P, P, P, P, P], C.c_int),
    "arapgs_node_quats": ([P, P, C.c_int, P], C.c_int),
    "arapgs_sample_sh": ([P, C.c_int64, C.c_int, P, P, P, P, P], C.c_int),
    "arapgs_samples_frames": ([P, C.c_int64, C.c_int, C.c_int, P, P, P, P, P, P, P], C.c_int),
    "arapgs_render": ([P, C.POINTER(RenderDesc), C.POINTER(C.c_int)], C.c_int),
    "arapgs_render_debug": ([P, P, P, C.c_int64, P, C.c_int64], C.c_int),
    "arapgs_scene_aabb": ([P, P, C.c_int64, C.c_int, P], C.c_int),
    "arapgs_grid_reorder": ([P, P, C.c_int64, P, C.c_int, P], C.c_int),
    "arapgs_gaussian_aabbs": ([P, P, P, P, P, C.c_int64, P], C.c_int),
    "arapgs_grid_build": ([P, P, C.c_int64, P, C.c_int, C.c_int, C.POINTER(P)], C.c_int),
    "arapgs_grid_destroy": ([P], None),
    "arapgs_grid_rebin": ([P, P, P, C.c_int64, P], C.c_int),
    "arapgs_grid_info": ([P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(P), C.POINTER(P), C.POINTER(P),
                          C.POINTER(P)], C.c_int),
    "arapgs_density_sample": ([P, P, P, P, P, P, P, C.c_float, P, P], C.c_int),
    "arapgs_grid_set_ada_lpf": ([P, P, P, C.c_float], C.c_int),
    "arapgs_grid_ada_lpf": ([P, C.POINTER(P)], C.c_int),
    "arapgs_solver_set_host_checks": ([P, C.c_int], C.c_int),
    "arapgs_solver_set_coarse_refresh": ([P, C.c_int], C.c_int),
    "arapgs_solver_set_cg_mode": ([P, C.c_int], C.c_int),
    "arapgs_solver_set_cg_predict": ([P, C.c_int], C.c_int),
    "arapgs_solver_set_coarse_lag": ([P, C.c_int], C.c_int),
    "arapgs_solver_set_cg_assembled": ([P, C.c_int], C.c_int),
    "arapgs_solver_set_cg_tol": ([P, C.c_float], C.c_int),
    "arapgs_solver_set_host_spin": ([P, C.c_int], C.c_int),
    "arapgs_set_host_spin": ([P, C.c_int], C.c_int),
    "arapgs_solver_set_side_stream": ([P, P], C.c_int),
    "arapgs_solver_set_warm_start": ([P, C.c_int], C.c_int),
    "arapgs_solver_reset": ([P], C.c_int),
    "arapgs_solver_last_stats": ([P, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "arapgs_solver_apply": ([P, P, P, P, P, P, P, P], C.c_int),
    "arapgs_deform_frame": ([P, P, C.POINTER(FrameBuffers), P, C.c_int, C.c_int, P, P], C.c_int),
    "arapgs_aims_translate": ([P, P, C.c_int, P, C.c_int, P, P], C.c_int),
    "arapgs_aims_twist": ([P, P, C.c_int, P, C.c_int, P, C.c_float, P], C.c_int),
    "arapgs_aims_scale": ([P, P, C.c_int, P, C.c_int, C.c_float, P], C.c_int),
    "arapgs_static_rows": ([P, P, P, C.c_int64, C.c_int, P], C.c_int),
    "arapgs_aims_set": ([P, P, C.c_int, P, C.c_int, P, P], C.c_int),
    "arapgs_profile_enable": ([P, C.c_int], C.c_int),
    "arapgs_profile_read": ([P, P, C.c_int], C.c_int),
    "arapgs_profile_read_kernel": ([P, C.c_int, P, C.c_int], C.c_int),
    "arapgs_sh_rotate": ([P, C.c_int64, P, P], C.c_int),
    "arapgs_polar": ([P, C.c_int64, P, P, P], C.c_int),
    "arapgs_tile_test_host": ([C.c_int64, P, P, P, P], C.c_int),
    "arapgs_render_views": ([P, P, C.c_int, P], C.c_int),
    "arapgs_render_segments": ([P, P, C.c_int, P, P], C.c_int),
    "arapgs_set_blend_unroll": ([P, C.c_int], C.c_int),
    "arapgs_density_backward": ([P, P, P, P, P, P, P, P, C.c_float, P, P, P, P, P, P, P, P], C.c_int),
    "arapgs_realign_create": ([P, P, C.c_int64, C.POINTER(RealignDesc), C.POINTER(P)], C.c_int),
    "arapgs_realign_destroy": ([P], None),
    "arapgs_realign_step": ([P, P, P, P, P, C.c_float, P, P, P, P, P, P], C.c_int),
    "arapgs_realign_set_options": ([P, P, P], C.c_int),
    "arapgs_realign_expand": ([P, P], C.c_int),
    "arapgs_realign_cells": ([P, C.POINTER(C.c_int64), C.POINTER(P), C.POINTER(P)], C.c_int),
    "arapgs_chain_create": ([P, C.POINTER(ChainDesc), C.POINTER(P)], C.c_int),
    "arapgs_chain_destroy": ([P], None),
    "arapgs_chain_forward": ([P, P, P, P, P, P, P, P], C.c_int),
    "arapgs_chain_backward": ([P, P, P, P, P, P, P, P, P, P, P], C.c_int),
    "arapgs_chain_state": ([P, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P)], C.c_int),
    "arapgs_ply_info": ([C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int)], C.c_int),
    "arapgs_ply_load": ([P, C.c_char_p, C.c_int, P, P, P, P, P, P, P], C.c_int),
    "arapgs_ply_save": ([P, C.c_char_p, C.c_int64, P, P, P, P, P, P, P, P, C.POINTER(C.c_int64)], C.c_int),
}


def build(force: bool = False) -> None:
    """Compile libarapgs.so for gfx950 with hipcc (make, in place)."""
    cmd = ["make", "-s", "-C", HERE, "-j8"]
    if force:
        subprocess.check_call(["make", "-s", "-C", HERE, "clean"])
    subprocess.check_call(cmd)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ArapgsError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C {HERE}")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in EXPORTS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def ptr(t):
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ArapgsError("arapgs calls take device tensors")
    if not t.is_contiguous():
        raise ArapgsError("arapgs calls take contiguous tensors")
    return C.c_void_p(t.data_ptr())


class CuStream:
    """A HIP stream limited to compute units [cu_begin, cu_begin + cu_count)
    (all of them when cu_count is 0), usable as a torch stream (.torch)."""

    def __init__(self, device=0, cu_begin=0, cu_count=0, cus=None):
        """cus: an explicit list of compute-unit indices (instead of a range)."""
        import torch
        h = P()
        if cus is not None:
            words = (max(cus) + 32) // 32 if cus else 1
            mask = (C.c_uint32 * words)()
            for cu in cus:
                mask[cu // 32] |= 1 << (cu % 32)
            rc = lib().arapgs_stream_create_mask(device, C.cast(mask, P), words, C.byref(h))
        else:
            rc = lib().arapgs_stream_create(device, cu_begin, cu_count, C.byref(h))
        if rc != 0:
            raise ArapgsError(f"arapgs_stream_create({cu_begin}, {cu_count}, {cus}) failed ({rc})")
        self.h = h
        self.torch = torch.cuda.ExternalStream(h.value, device=torch.device("cuda", device))

    def close(self):
        if self.h:
            lib().arapgs_stream_destroy(self.h)
            self.h = None


def ply_info(path):
    """(vertex count, has_index) of a 3DGS PLY (header check only, no GPU)."""
    n = C.c_int64(0)
    has = C.c_int(0)
    if lib().arapgs_ply_info(os.fsencode(path), C.byref(n), C.byref(has)) != 0:
        raise ArapgsError(f"{path}: not a 3DGS PLY of 62 (or 63 with index) float properties")
    return n.value, bool(has.value)


class Context:
    """One arapgs context on a device, bound to torch's current stream."""

    def __init__(self, device: int = 0, stream=None):
        import torch
        self._torch = torch
        h = P()
        rc = lib().arapgs_create(device, C.byref(h))
        if rc != 0:
            raise ArapgsError(f"arapgs_create failed ({rc})")
        self.h = h
        self.device = device
        self.set_stream(stream if stream is not None else torch.cuda.current_stream(device))

    def set_stream(self, stream):
        self.stream = stream
        self._check(lib().arapgs_set_stream(self.h, C.c_void_p(stream.cuda_stream)))

    def _check(self, rc):
        if rc != 0:
            msg = lib().arapgs_last_error(self.h)
            raise ArapgsError(f"arapgs error {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            lib().arapgs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        self._check(lib().arapgs_synchronize(self.h))

    # ------------------------------------------------------------ graph build
    def fps(self, pos, m):
        t = self._torch
        out = t.zeros(m + 1, dtype=t.int32, device=pos.device)
        cnt = C.c_int(0)
        self._check(lib().arapgs_fps(self.h, ptr(pos), pos.shape[0], m, ptr(out), C.byref(cnt)))
        return out[: cnt.value]

    def knn(self, nodes, queries, k, with_weights=True):
        t = self._torch
        Q = queries.shape[0]
        idx = t.empty((Q, k + 1), dtype=t.int32, device=queries.device)
        w = t.empty((Q, k), dtype=t.float32, device=queries.device) if with_weights else None
        self._check(lib().arapgs_knn(self.h, ptr(nodes), nodes.shape[0], ptr(queries), Q, k, ptr(idx), ptr(w)))
        return idx, w

    def knn_noalloc(self, nodes, queries, k, idx, w=None):
        """arapgs_knn_noalloc into caller-owned idx (Q, k+1) int32 and w (Q, k)
        float32 (or None): the library allocates nothing."""
        self._check(lib().arapgs_knn_noalloc(self.h, ptr(nodes), nodes.shape[0], ptr(queries), queries.shape[0], k,
                                             ptr(idx), ptr(w)))
        return idx, w

    def graph_edges(self, nodes, k):
        t = self._torch
        nbr = t.empty((nodes.shape[0], k), dtype=t.int32, device=nodes.device)
        self._check(lib().arapgs_graph_edges(self.h, ptr(nodes), nodes.shape[0], k, ptr(nbr)))
        return nbr

    def end_points(self, pos, rot, scale, pad=1e-3, coeff=2.0):
        t = self._torch
        ends = t.empty((pos.shape[0], 6, 3), dtype=t.float32, device=pos.device)
        self._check(lib().arapgs_end_points(self.h, ptr(pos), ptr(rot), ptr(scale), pos.shape[0], pad, coeff,
                                            ptr(ends)))
        return ends

    # ------------------------------------------------------------ deformation
    def blend_points(self, pts, idx, w, table, skip=None, out=None):
        t = self._torch
        out = t.empty_like(pts) if out is None else out
        self._check(lib().arapgs_blend_points(self.h, ptr(pts), pts.shape[0], ptr(idx), ptr(w), idx.shape[-1],
                                              ptr(table), table.shape[0], ptr(skip), ptr(out)))
        return out

    def six_point_update(self, table, ends_idx, ends_w, ends, scale0, is_static, pos, rot, scale, sh, pad=1e-3,
                         coeff=2.0, src=None):
        """src: (pos, rot, scale, sh) to read instead of the written set
        (arapgs_six_point_update_from)."""
        k = ends_idx.shape[-1]
        if src is not None:
            self._check(lib().arapgs_six_point_update_from(
                self.h, pos.shape[0], k, ptr(table), table.shape[0], ptr(ends_idx), ptr(ends_w), ptr(ends),
                ptr(scale0), ptr(is_static), pad, coeff, ptr(src[0]), ptr(src[1]), ptr(src[2]), ptr(src[3]),
                ptr(pos), ptr(rot), ptr(scale), ptr(sh)))
            return
        self._check(lib().arapgs_six_point_update(self.h, pos.shape[0], k, ptr(table), table.shape[0], ptr(ends_idx),
                                                  ptr(ends_w), ptr(ends), ptr(scale0), ptr(is_static), pad, coeff,
                                                  ptr(pos), ptr(rot), ptr(scale), ptr(sh)))

    def node_quats(self, table, out=None):
        t = self._torch
        q = t.empty((table.shape[0], 4), dtype=t.float32, device=table.device) if out is None else out
        self._check(lib().arapgs_node_quats(self.h, ptr(table), table.shape[0], ptr(q)))
        return q

    def samples_frames(self, tables, node_qs, idx, w, is_static, samples, feat):
        """len(tables) consecutive frames of the sample blend + sample SH in one
        pass (arapgs_samples_frames; bit-identical to the per-frame calls, or with
        set_sh_compose one SH rotation by the product of the frames' rotations)."""
        nf = len(tables)
        tp = _ptr_array(tables)
        qp = _ptr_array(node_qs)
        self._check(lib().arapgs_samples_frames(self.h, samples.shape[0], idx.shape[1], nf, tp, qp, ptr(idx), ptr(w),
                                                ptr(is_static), ptr(samples), ptr(feat)))

    def sample_sh(self, w, idx, node_q, is_static, feat):
        self._check(lib().arapgs_sample_sh(self.h, feat.shape[0], idx.shape[-1], ptr(w), ptr(idx), ptr(node_q),
                                           ptr(is_static), ptr(feat)))

    def render(self, means, shs, opac, scales, rots, view, proj, campos, tanfovx, tanfovy, W, H, bg, D=3,
               scale_mod=1.0, fast_culling=True, boxmin=None, boxmax=None, tile=16, out=None, want_surface=False,
               want_radii=False, ortho_scale=None):
        """Rasterizer::forward (GaussianView.cpp:1099-1128). view/proj/campos/bg are device tensors.
        ortho_scale: orthographic camera (to_ortho with that ortho_scale; proj must be its matrix)."""
        t = self._torch
        dev = means.device
        P_ = means.shape[0]
        out = t.empty((3, H, W), dtype=t.float32, device=dev) if out is None else out
        surface = t.zeros(P_, dtype=t.int32, device=dev) if want_surface else None
        radii = t.zeros(P_, dtype=t.int32, device=dev) if want_radii else None
        rects = t.zeros((P_, 2), dtype=t.int32, device=dev) if fast_culling else None
        import numpy as np
        self._box = None
        bmin = bmax = None
        if boxmin is not None:
            self._box = (np.ascontiguousarray(boxmin, np.float32), np.ascontiguousarray(boxmax, np.float32))
            bmin, bmax = self._box[0].ctypes.data_as(P), self._box[1].ctypes.data_as(P)
        d = RenderDesc(P_, D, 16, ptr(bg), W, H, ptr(means), ptr(shs), None, ptr(opac), ptr(scales), scale_mod,
                       ptr(rots), None, ptr(view), ptr(proj), ptr(campos), tanfovx, tanfovy, 0, ptr(out), ptr(surface),
                       ptr(radii), ptr(rects), bmin, bmax, int(ortho_scale is not None), float(ortho_scale or 1.0),
                       tile)
        n = C.c_int(0)
        self._check(lib().arapgs_render(self.h, C.byref(d), C.byref(n)))
        return dict(image=out, num_rendered=n.value, surface=surface, radii=radii, rects=rects)

    def set_blend_unroll(self, unroll):
        """Gaussians per compositing step of the 16x16 blend (2, 4, 8); images do not change."""
        self._check(lib().arapgs_set_blend_unroll(self.h, unroll))

    def render_views(self, means, shs, opac, scales, rots, cams, W, H, bg, outs, D=3, scale_mod=1.0,
                     fast_culling=True, tile=16):
        """All views in one pass (arapgs_render_views); cams: dicts with view, proj, campos (device
        tensors), tx, ty; outs: (3, H, W) device tensors. Returns the total instance count."""
        t = self._torch
        dev = means.device
        P_ = means.shape[0]
        V = len(cams)
        if not hasattr(self, "_rects") or self._rects is None or self._rects.shape[0] != V * P_:
            self._rects = t.zeros((V * P_, 2), dtype=t.int32, device=dev) if fast_culling else None
        arr = (RenderDesc * V)()
        for v, (cam, out) in enumerate(zip(cams, outs)):
            rects = None
            if fast_culling:
                rects = C.c_void_p(self._rects.data_ptr() + v * P_ * 8)
            arr[v] = RenderDesc(P_, D, 16, ptr(bg), W, H, ptr(means), ptr(shs), None, ptr(opac), ptr(scales),
                                scale_mod, ptr(rots), None, ptr(cam["view"]), ptr(cam["proj"]), ptr(cam["campos"]),
                                cam["tx"], cam["ty"], 0, ptr(out), None, None, rects, None, None, 0, 1.0, tile)
        n = C.c_int(0)
        self._check(lib().arapgs_render_views(self.h, arr, V, C.byref(n)))
        return n.value

    def render_segments(self, means, shs, opac, scales, rots, cams, W, H, bg, outs, seg, D=3, scale_mod=1.0,
                        tile=16):
        """Scenes stored back to back, one camera each (arapgs_render_segments):
        cams[v] draws Gaussians seg[v]..seg[v+1]-1 into outs[v]. Returns the total
        instance count."""
        import numpy as np
        P_ = means.shape[0]
        V = len(cams)
        seg = np.ascontiguousarray(seg, np.int64)
        if seg.shape != (V + 1,):
            raise ArapgsError("render_segments: need len(cams) + 1 segment offsets")
        arr = (RenderDesc * V)()
        for v, (cam, out) in enumerate(zip(cams, outs)):
            arr[v] = RenderDesc(P_, D, 16, ptr(bg), W, H, ptr(means), ptr(shs), None, ptr(opac), ptr(scales),
                                scale_mod, ptr(rots), None, ptr(cam["view"]), ptr(cam["proj"]), ptr(cam["campos"]),
                                cam["tx"], cam["ty"], 0, ptr(out), None, None, None, None, None, 0, 1.0, tile)
        n = C.c_int(0)
        self._check(lib().arapgs_render_segments(self.h, arr, V, seg.ctypes.data_as(P), C.byref(n)))
        return n.value

    # ------------------------------------------------------------ PLY codec
    def ply_load(self, path, morton=True):
        """Loads a 3DGS PLY (loadPly, GaussianView.cpp:43-268) into device
        tensors with the activations applied; Morton-ordered unless morton is
        False. Returns dict(pos, rot, scale, opacity, sh (N, 48), index, box)."""
        import numpy as np
        t = self._torch
        n, _ = ply_info(path)
        dev = t.device("cuda", self.device)
        out = dict(pos=t.empty((n, 3), dtype=t.float32, device=dev), rot=t.empty((n, 4), dtype=t.float32, device=dev),
                   scale=t.empty((n, 3), dtype=t.float32, device=dev),
                   opacity=t.empty(n, dtype=t.float32, device=dev), sh=t.empty((n, 48), dtype=t.float32, device=dev),
                   index=t.empty(n, dtype=t.int32, device=dev))
        box = np.zeros(6, np.float32)
        self._check(lib().arapgs_ply_load(self.h, os.fsencode(path), 1 if morton else 0, ptr(out["pos"]),
                                          ptr(out["rot"]), ptr(out["scale"]), ptr(out["opacity"]), ptr(out["sh"]),
                                          ptr(out["index"]), box.ctypes.data_as(P)))
        out["box"] = box
        return out

    def ply_save(self, path, pos, rot, scale, opacity, sh, box=None, too_small=None):
        """Writes the Gaussians inside box ((min xyz), (max xyz)) and not flagged
        in too_small (device bool/uint8) as a 3DGS PLY (savePly,
        GaussianView.cpp:270-339). Returns the number written."""
        import numpy as np
        t = self._torch
        lo = hi = None
        if box is not None:
            lo = np.ascontiguousarray(box[0], np.float32)
            hi = np.ascontiguousarray(box[1], np.float32)
        ts = None
        if too_small is not None:
            ts = too_small.to(t.uint8).contiguous()
        written = C.c_int64(0)
        self._check(lib().arapgs_ply_save(self.h, os.fsencode(path), pos.shape[0], ptr(pos), ptr(rot), ptr(scale),
                                          ptr(opacity), ptr(sh), None if lo is None else lo.ctypes.data_as(P),
                                          None if hi is None else hi.ctypes.data_as(P), ptr(ts), C.byref(written)))
        return written.value

    # ------------------------------------------------------------ grid
    def scene_aabb(self, pos, G):
        import numpy as np
        out = np.zeros(7, np.float32)
        self._check(lib().arapgs_scene_aabb(self.h, ptr(pos), pos.shape[0], G, out.ctypes.data_as(P)))
        return out

    def grid_reorder(self, pos, aabb7, G):
        import numpy as np
        t = self._torch
        order = t.empty(pos.shape[0], dtype=t.int32, device=pos.device)
        a = np.ascontiguousarray(aabb7, np.float32)
        self._check(lib().arapgs_grid_reorder(self.h, ptr(pos), pos.shape[0], a.ctypes.data_as(P), G, ptr(order)))
        return order

    def gaussian_aabbs(self, pos, rot, scale, opacity):
        t = self._torch
        out = t.empty((pos.shape[0], 6), dtype=t.float32, device=pos.device)
        self._check(lib().arapgs_gaussian_aabbs(self.h, ptr(pos), ptr(rot), ptr(scale), ptr(opacity), pos.shape[0],
                                                ptr(out)))
        return out

    def set_blend_split(self, entries):
        """Deep 16x16 tiles blended in segments of `entries` instances (0: off)."""
        self._check(lib().arapgs_set_blend_split(self.h, int(entries)))

    def set_node_cache(self, on=True):
        """Six-point end-point blend through the LDS node-row cache (bit-identical)."""
        self._check(lib().arapgs_set_node_cache(self.h, int(bool(on))))

    def set_host_spin(self, on=True):
        """The render's instance-count read polls its pinned slot (images unchanged)."""
        self._check(lib().arapgs_set_host_spin(self.h, int(bool(on))))

    def set_tile_cull(self, on=True):
        """Renders emit only the tiles that pass the conservative tile test (identical images)."""
        self._check(lib().arapgs_set_tile_cull(self.h, int(bool(on))))

    def set_sh_compose(self, on=True):
        """samples_frames: one SH rotation by the product of the frames' rotations (within float rounding)."""
        self._check(lib().arapgs_set_sh_compose(self.h, int(bool(on))))

    def set_grid_chunk(self, pairs):
        """Cell-Gaussian pairs per chunk of later grid builds (0: default); test hook."""
        self._check(lib().arapgs_set_grid_chunk(self.h, int(pairs)))

    def grid_build(self, aabbs, aabb7, G, padding=1):
        import numpy as np
        a = np.ascontiguousarray(aabb7, np.float32)
        h = P()
        self._check(lib().arapgs_grid_build(self.h, ptr(aabbs), aabbs.shape[0], a.ctypes.data_as(P), G, padding,
                                            C.byref(h)))
        return Grid(self, h, G)

    def render_debug(self, n, ntiles):
        import numpy as np
        keys = np.zeros(n, np.uint64)
        vals = np.zeros(n, np.uint32)
        ranges = np.zeros((ntiles, 2), np.uint32)
        self._check(lib().arapgs_render_debug(self.h, keys.ctypes.data_as(P), vals.ctypes.data_as(P), n,
                                              ranges.ctypes.data_as(P), ntiles))
        return keys, vals, ranges

    def aims_translate(self, node_pos, active, delta, aims):
        import numpy as np
        d = np.ascontiguousarray(delta, np.float32)
        self._check(lib().arapgs_aims_translate(self.h, ptr(node_pos), node_pos.shape[0], ptr(active),
                                                active.shape[0], d.ctypes.data_as(P), ptr(aims)))

    def aims_twist(self, node_pos, active, axis, radian, aims):
        import numpy as np
        a = np.ascontiguousarray(axis, np.float32)
        self._check(lib().arapgs_aims_twist(self.h, ptr(node_pos), node_pos.shape[0], ptr(active), active.shape[0],
                                            a.ctypes.data_as(P), radian, ptr(aims)))

    def aims_scale(self, node_pos, active, s, aims):
        self._check(lib().arapgs_aims_scale(self.h, ptr(node_pos), node_pos.shape[0], ptr(active), active.shape[0],
                                            float(s), ptr(aims)))

    def aims_set(self, node_pos, idx, vals, aims):
        self._check(lib().arapgs_aims_set(self.h, ptr(node_pos), node_pos.shape[0], ptr(idx), idx.shape[0],
                                          ptr(vals), ptr(aims)))

    def static_rows(self, node_static, idx, out=None):
        """uint8 per row of idx (n, len): 1 when all its node ids are static."""
        t = self._torch
        n = idx.shape[0]
        length = idx[0].numel() if n else 1
        if out is None:
            out = t.empty(n, dtype=t.uint8, device=idx.device)
        self._check(lib().arapgs_static_rows(self.h, ptr(node_static), ptr(idx), n, length, ptr(out)))
        return out

    def profile_enable(self, on=True):
        self._check(lib().arapgs_profile_enable(self.h, int(on)))

    def profile_read(self, cap=4096, which=0):
        """Per-launch durations (ms) of the six-point (0), sample blend (1) or
        sample SH (2) launches of deform_frame since profiling was enabled."""
        import numpy as np
        buf = np.zeros(cap, np.float32)
        n = lib().arapgs_profile_read_kernel(self.h, which, buf.ctypes.data_as(P), cap)
        if n < 0:
            self._check(n)
        return buf[:n]

    def sh_rotate(self, R9, sh):
        self._check(lib().arapgs_sh_rotate(self.h, sh.shape[0], ptr(R9), ptr(sh)))

    def polar(self, M9):
        t = self._torch
        R = t.empty_like(M9)
        S = t.empty((M9.shape[0], 3), dtype=t.float64, device=M9.device)
        self._check(lib().arapgs_polar(self.h, M9.shape[0], ptr(M9), ptr(R), ptr(S)))
        return R, S


def tile_test_host(xy, conic, box):
    """Blend's conservative Gaussian-vs-tile test evaluated on the host (numpy
    float32 arrays xy (n, 2), conic (n, 4), box (n, 4)); returns uint8 keep."""
    import numpy as np
    xy = np.ascontiguousarray(xy, dtype=np.float32)
    conic = np.ascontiguousarray(conic, dtype=np.float32)
    box = np.ascontiguousarray(box, dtype=np.float32)
    n = xy.shape[0]
    if conic.shape != (n, 4) or box.shape != (n, 4) or xy.shape != (n, 2):
        raise ArapgsError("tile_test_host: shapes must be (n,2), (n,4), (n,4)")
    keep = np.zeros(n, dtype=np.uint8)
    rc = lib().arapgs_tile_test_host(n, xy.ctypes.data, conic.ctypes.data, box.ctypes.data, keep.ctypes.data)
    if rc != 0:
        raise ArapgsError(f"tile_test_host failed ({rc})")
    return keep


def node_table(n, device):
    """Empty arapgs_node_xf table (n x 16 float32)."""
    import torch
    return torch.zeros((n, 16), dtype=torch.float32, device=device)


class Solver:
    """arapgs_solver for one block layout (Deform::Deform, Deform.hpp:414-455)."""

    def __init__(self, ctx: Context, nbr, blocks, block_types, src_idx, src_w, node_pos, w_rot=1.0, w_reg=10.0,
                 w_con=100.0, weighting_factor=1.0, n_aggregates=0, cg_max_iters=0, cg_tol=0.0,
                 constraints_on_center=False):
        import numpy as np
        self.ctx = ctx
        nbr = np.ascontiguousarray(nbr, np.uint32)
        M, K = nbr.shape
        ptrs = [0]
        for b in blocks:
            ptrs.append(ptrs[-1] + len(b))
        self._keep = dict(
            nbr=nbr,
            bptr=np.ascontiguousarray(ptrs, np.int32),
            bnodes=np.ascontiguousarray(np.concatenate([np.asarray(b, np.int32) for b in blocks]) if blocks
                                        else np.zeros(1, np.int32), np.int32),
            btype=np.ascontiguousarray(block_types if len(blocks) else np.zeros(1), np.int32),
            src_idx=np.ascontiguousarray(src_idx, np.uint32),
            src_w=np.ascontiguousarray(src_w, np.float32),
            pos=np.ascontiguousarray(node_pos, np.float32),
        )
        kp = self._keep
        d = SolverDesc(M, K, kp["nbr"].ctypes.data_as(P), kp["bptr"].ctypes.data_as(P),
                       kp["bnodes"].ctypes.data_as(P), kp["btype"].ctypes.data_as(P), len(blocks),
                       kp["src_idx"].ctypes.data_as(P), kp["src_w"].ctypes.data_as(P), kp["pos"].ctypes.data_as(P),
                       w_rot, w_reg, w_con, weighting_factor, int(bool(constraints_on_center)), n_aggregates,
                       cg_max_iters, cg_tol)
        h = P()
        ctx._check(lib().arapgs_solver_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        nf, nu = C.c_int(0), C.c_int(0)
        lib().arapgs_solver_dims(h, C.byref(nf), C.byref(nu))
        self.n_free, self.n_unknowns = nf.value, nu.value
        self.M, self.K = M, K

    def close(self):
        if self.h:
            lib().arapgs_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, node_pos, aims, iters, stop_rule=True, x_out=None, energies=None, iters_done=None):
        t = self.ctx._torch
        dev = node_pos.device
        x_out = t.empty(self.n_unknowns, dtype=t.float32, device=dev) if x_out is None else x_out
        energies = t.zeros(iters + 1, dtype=t.float64, device=dev) if energies is None else energies
        iters_done = t.zeros(1, dtype=t.int32, device=dev) if iters_done is None else iters_done
        self.ctx._check(lib().arapgs_gn_solve(self.ctx.h, self.h, ptr(node_pos), ptr(aims), iters, int(stop_rule),
                                              ptr(x_out), ptr(energies), ptr(iters_done)))
        return x_out, energies, iters_done

    def energy(self, node_pos, aims, x):
        t = self.ctx._torch
        e = t.zeros(1, dtype=t.float64, device=node_pos.device)
        self.ctx._check(lib().arapgs_gn_energy(self.ctx.h, self.h, ptr(node_pos), ptr(aims), ptr(x), ptr(e)))
        return e

    def set_host_checks(self, every):
        self.ctx._check(lib().arapgs_solver_set_host_checks(self.h, every))

    def set_coarse_refresh(self, frames):
        self.ctx._check(lib().arapgs_solver_set_coarse_refresh(self.h, frames))

    def set_cg_block(self, threads):
        """Workgroup size of the CG operator kernel (64, 128, 256, 512; both CG kernels in cg_mode 2)."""
        if lib().arapgs_solver_set_cg_block(self.h, int(threads)) != 0:
            raise ArapgsError(f"set_cg_block({threads}) failed")

    def set_async_refresh(self, on=True):
        """Coarse rebuilds on the solver's side stream, adopted by the next solve (default on)."""
        self.ctx._check(lib().arapgs_solver_set_async_refresh(self.h, int(bool(on))))

    def set_cg_graphs(self, on=True):
        """GN phases and CG chunks as captured HIP graphs (CG flags checked one chunk behind)."""
        self.ctx._check(lib().arapgs_solver_set_cg_graphs(self.h, int(bool(on))))

    def apply(self, node_pos, aims, x, v):
        """(J^T J v, -J^T f) at unknowns x (test hook)."""
        t = self.ctx._torch
        q = t.empty_like(v)
        b = t.empty_like(v)
        self.ctx._check(lib().arapgs_solver_apply(self.ctx.h, self.h, ptr(node_pos), ptr(aims), ptr(x), ptr(v), ptr(q),
                                                  ptr(b)))
        return q, b

    def last_stats(self):
        """(GN iterations, total CG iterations) of the last solve (synchronous)."""
        gn, cg = C.c_int(0), C.c_int(0)
        if lib().arapgs_solver_last_stats(self.h, C.byref(gn), C.byref(cg)) != 0:
            raise ArapgsError("last_stats failed")
        return gn.value, cg.value

    def set_warm_start(self, on=True):
        if lib().arapgs_solver_set_warm_start(self.h, int(bool(on))) != 0:
            raise ArapgsError("set_warm_start failed")

    def reset(self):
        """Forget the frame-to-frame state (warm-start steps, coarse operator,
        refresh count): the next solve computes what a new solver's would."""
        if lib().arapgs_solver_reset(self.h) != 0:
            raise ArapgsError("solver reset failed")

    def set_side_stream(self, stream=None):
        """The stream of the asynchronous coarse rebuilds (a torch stream, or None for the solver's own)."""
        h = stream.cuda_stream if stream is not None else 0
        if lib().arapgs_solver_set_side_stream(self.h, P(h) if h else None) != 0:
            raise ArapgsError("set_side_stream failed")

    def set_host_spin(self, on=True):
        """Host checks poll the pinned flag instead of waiting for the stream."""
        if lib().arapgs_solver_set_host_spin(self.h, int(bool(on))) != 0:
            raise ArapgsError("set_host_spin failed")

    def set_cg_tol(self, tol):
        """The CG's relative residual tolerance (inexact Newton; default 1e-4)."""
        if lib().arapgs_solver_set_cg_tol(self.h, float(tol)) != 0:
            raise ArapgsError(f"set_cg_tol({tol}) failed")

    def set_cg_assembled(self, on=True):
        """The CG operator applied from per-frame assembled 4x4 pair blocks (plain mode)."""
        rc = lib().arapgs_solver_set_cg_assembled(self.h, int(bool(on)))
        if rc != 0:
            raise ArapgsError(f"set_cg_assembled({on}) failed ({rc})")

    def set_coarse_lag(self, solves):
        """A side-stream coarse rebuild is adopted by the solves-th solve after the one that started it."""
        if lib().arapgs_solver_set_coarse_lag(self.h, int(solves)) != 0:
            raise ArapgsError(f"set_coarse_lag({solves}) failed")

    def set_cg_predict(self, on=True):
        """First CG host check of each GN iteration at the previous solve's count (same results)."""
        self.ctx._check(lib().arapgs_solver_set_cg_predict(self.h, int(bool(on))))

    def set_cg_mode(self, mode):
        """0 = three kernels per CG iteration (default), 1 = persistent cooperative PCG kernel,
        2 = two kernels per CG iteration (coarse restriction by recurrence)."""
        self.ctx._check(lib().arapgs_solver_set_cg_mode(self.h, mode))

    def deform_frame(self, fb, aims, iters, stop_rule=True, energies=None, iters_done=None):
        t = self.ctx._torch
        dev = aims.device
        energies = t.zeros(iters + 1, dtype=t.float64, device=dev) if energies is None else energies
        iters_done = t.zeros(1, dtype=t.int32, device=dev) if iters_done is None else iters_done
        self.ctx._check(lib().arapgs_deform_frame(self.ctx.h, self.h, C.byref(fb), ptr(aims), iters, int(stop_rule),
                                                  ptr(energies), ptr(iters_done)))
        return energies, iters_done

    def put_free_inputs(self, node_pos, x, table=None):
        table = node_table(self.M, node_pos.device) if table is None else table
        self.ctx._check(lib().arapgs_put_free_inputs(self.ctx.h, self.h, ptr(node_pos), ptr(x), ptr(table)))
        return table


def _ptr_array(tensors):
    """Host array of device pointers (None -> null entry)."""
    arr = (P * len(tensors))()
    for i, t in enumerate(tensors):
        arr[i] = None if t is None else ptr(t)
    return arr


class SolverBatch:
    """arapgs_solver_batch: independent scenes whose Gauss-Newton solves run as
    one launch per solver phase (scene = blockIdx.z). Per scene, the result is
    bit-identical to Solver.solve + Solver.put_free_inputs."""

    def __init__(self, ctx: Context, solvers):
        self.ctx = ctx
        self.solvers = list(solvers)
        arr = (P * len(self.solvers))(*[s.h for s in self.solvers])
        h = P()
        ctx._check(lib().arapgs_solver_batch_create(ctx.h, arr, len(self.solvers), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().arapgs_solver_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_host_checks(self, every):
        self.ctx._check(lib().arapgs_solver_batch_set_host_checks(self.h, every))

    def set_coarse_refresh(self, frames):
        self.ctx._check(lib().arapgs_solver_batch_set_coarse_refresh(self.h, frames))

    def set_cg_graphs(self, on=True):
        self.ctx._check(lib().arapgs_solver_batch_set_cg_graphs(self.h, int(bool(on))))

    def set_async_refresh(self, on=True):
        self.ctx._check(lib().arapgs_solver_batch_set_async_refresh(self.h, int(bool(on))))

    def solve(self, node_pos, aims, iters, stop_rule=True, x_out=None, energies=None, iters_done=None, tables=None):
        """Lists (one entry per scene) of device tensors; outputs default to new
        tensors (tables: None = no node tables)."""
        t = self.ctx._torch
        n = len(self.solvers)
        dev = node_pos[0].device
        if x_out is None:
            x_out = [t.empty(s.n_unknowns, dtype=t.float32, device=dev) for s in self.solvers]
        if energies is None:
            energies = [t.zeros(iters + 1, dtype=t.float64, device=dev) for _ in range(n)]
        if iters_done is None:
            iters_done = [t.zeros(1, dtype=t.int32, device=dev) for _ in range(n)]
        tb = _ptr_array(tables) if tables is not None else None
        self.ctx._check(lib().arapgs_solver_batch_solve(self.ctx.h, self.h, _ptr_array(node_pos), _ptr_array(aims),
                                                        iters, int(stop_rule), _ptr_array(x_out),
                                                        _ptr_array(energies), _ptr_array(iters_done), tb))
        return x_out, energies, iters_done


class Grid:
    """arapgs_grid: per-cell Gaussian lists, valid cells and samples (device views)."""

    def __init__(self, ctx, h, G):
        self.ctx, self.h, self.G = ctx, h, G
        self._info()

    def _info(self):
        n_pairs, n_valid = C.c_int64(0), C.c_int64(0)
        pr, li, va, sa = P(), P(), P(), P()
        lib().arapgs_grid_info(self.h, C.byref(n_pairs), C.byref(n_valid), C.byref(pr), C.byref(li), C.byref(va),
                               C.byref(sa))
        self.n_pairs, self.n_valid = n_pairs.value, n_valid.value
        self._ptrs = (pr.value, li.value, va.value, sa.value)

    def rebin(self, pos, rot, scale, opacity):
        """UpdateContainingRelationship (GaussianView.cpp:3634-3744): cell lists of
        the Gaussians' current (activated) state on the geometry of their current
        scene box; valid cells and samples are kept. Returns the geometry (aabb7)."""
        import numpy as np
        ctx = self.ctx
        a7 = ctx.scene_aabb(pos, self.G)
        aabbs = ctx.gaussian_aabbs(pos, rot, scale, opacity)
        a = np.ascontiguousarray(a7, np.float32)
        ctx._check(lib().arapgs_grid_rebin(ctx.h, self.h, ptr(aabbs), aabbs.shape[0], a.ctypes.data_as(P)))
        self._info()
        return a7

    def set_ada_lpf(self, lpf_param, samples=None):
        """GetAdaLpfRatio (GaussianView.cpp:4670-4749): per-cell adaptive low-pass
        filters lpf_param * M M^T from the cells' corner samples (the grid's own
        samples when None); the density calls on this grid use them until
        set_ada_lpf(0)."""
        self.ctx._check(lib().arapgs_grid_set_ada_lpf(self.ctx.h, self.h, ptr(samples), float(lpf_param)))

    def ada_lpf(self):
        """Host copy of the adaptive filters (n_valid, 6) as (xx, yy, zz, xy, xz, yz),
        or None while the grid uses lpf I."""
        import numpy as np
        import torch
        p = P()
        lib().arapgs_grid_ada_lpf(self.h, C.byref(p))
        if not p.value:
            return None
        a = np.empty((self.n_valid, 6), np.float32)
        torch.cuda.synchronize()
        if a.size and _hip_memcpy_d2h(a.ctypes.data, p.value, a.nbytes) != 0:
            raise ArapgsError("hipMemcpy failed")
        return a

    def arrays(self):
        """Copies of (prefix, list, valid, samples) on the host, via torch."""
        import numpy as np
        import torch
        G3 = self.G ** 3
        out = []
        for p, n, dt, shp in ((self._ptrs[0], G3, np.int64, (G3,)), (self._ptrs[1], self.n_pairs, np.int32,
                                                                      (self.n_pairs,)),
                              (self._ptrs[2], self.n_valid, np.int32, (self.n_valid,)),
                              (self._ptrs[3], self.n_valid * 64 * 3, np.float32, (self.n_valid * 64, 3))):
            a = np.empty(shp, dt)
            if n:
                torch.cuda.synchronize()
                rc = _hip_memcpy_d2h(a.ctypes.data, p, a.nbytes)
                if rc != 0:
                    raise ArapgsError(f"hipMemcpy failed ({rc})")
            out.append(a)
        return out

    def prefix_valid(self):
        """Host copies of the 64-bit per-cell prefix and the valid cells only
        (the lists of a large grid stay on the device)."""
        import numpy as np
        import torch
        G3 = self.G ** 3
        prefix = np.empty(G3, np.int64)
        valid = np.empty(self.n_valid, np.int32)
        torch.cuda.synchronize()
        for a, p in ((prefix, self._ptrs[0]), (valid, self._ptrs[2])):
            if a.size and _hip_memcpy_d2h(a.ctypes.data, p, a.nbytes) != 0:
                raise ArapgsError("hipMemcpy failed")
        return prefix, valid

    def list_range(self, beg, end):
        """Host copy of list entries [beg, end)."""
        import numpy as np
        a = np.empty(max(0, end - beg), np.int32)
        if a.size and _hip_memcpy_d2h(a.ctypes.data, self._ptrs[1] + 4 * beg, a.nbytes) != 0:
            raise ArapgsError("hipMemcpy failed")
        return a

    def samples_tensor(self):
        """Device copy of the samples as a torch tensor (S, 3)."""
        import torch
        out = torch.empty((self.n_valid * 64, 3), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        if out.numel() and _hip_memcpy(out.data_ptr(), self._ptrs[3], out.numel() * 4, 3) != 0:
            raise ArapgsError("hipMemcpy failed")
        return out

    def density_backward(self, pos, rot, scale, opacity, sh, lpf, dfeat=None, dopac=None, samples=None,
                         want_samples=False):
        """VJP of the density sample (arapgs_density_backward): returns
        (dpos, drot, dscale, dopacity, dsh[, dsamples])."""
        t = self.ctx._torch
        dev = pos.device
        N = pos.shape[0]
        out = [t.empty((N, 3), device=dev), t.empty((N, 4), device=dev), t.empty((N, 3), device=dev),
               t.empty(N, device=dev), t.empty((N, 48), device=dev)]
        ds = t.empty((self.n_valid * 64, 3), device=dev) if want_samples else None
        self.ctx._check(lib().arapgs_density_backward(self.ctx.h, self.h, ptr(samples), ptr(pos), ptr(rot), ptr(scale),
                                                      ptr(opacity), ptr(sh), lpf, ptr(dfeat), ptr(dopac),
                                                      *[ptr(o) for o in out], ptr(ds)))
        return tuple(out) + ((ds,) if want_samples else ())

    def density(self, pos, rot, scale, opacity, sh, lpf):
        t = self.ctx._torch
        S = self.n_valid * 64
        feat = t.empty((S, 48), dtype=t.float32, device=pos.device)
        opac = t.empty(S, dtype=t.float32, device=pos.device)
        self.ctx._check(lib().arapgs_density_sample(self.ctx.h, self.h, ptr(pos), ptr(rot), ptr(scale), ptr(opacity),
                                                    ptr(sh), lpf, ptr(feat), ptr(opac)))
        return feat, opac

    def close(self):
        if self.h:
            lib().arapgs_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_hip = None


def _hiplib():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    return _hip


def _hip_memcpy(dst, src, nbytes, kind):
    """hipMemcpy with an explicit kind (2 device-to-host, 3 device-to-device)."""
    return _hiplib().hipMemcpy(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), kind)


def _hip_memcpy_d2h(dst, src, nbytes):
    return _hip_memcpy(dst, src, nbytes, 2)


def frame_buffers(state: dict, k: int, axis_padding=1e-3, end_coeff=2.0) -> FrameBuffers:
    """FrameBuffers over a dict of device tensors (keys as in arapgs_frame_buffers)."""
    g = state.get
    return FrameBuffers(int(g("pos").shape[0]), k, int(g("node_pos").shape[0]), ptr(g("pos")), ptr(g("rot")),
                        ptr(g("scale")), ptr(g("sh")), ptr(g("scale0")), ptr(g("ends")), ptr(g("ends_idx")),
                        ptr(g("ends_w")), ptr(g("gs_static")), ptr(g("node_pos")), ptr(g("node_src_idx")),
                        ptr(g("node_src_w")), int(g("samples").shape[0]) if g("samples") is not None else 0,
                        ptr(g("samples")), ptr(g("sample_idx")), ptr(g("sample_w")), ptr(g("sample_static")),
                        ptr(g("sample_feat")), axis_padding, end_coeff)


class Chain:
    """arapgs_chain: the differentiable deform -> density-sample -> render chain.

    forward(table, rot0, scale0) -> (images [V x (3,H,W)], feat (S,48), dens (S,))
    backward(table, rot0, scale0, dimages, dfeat, ddens) -> (dtable (M,16), drot0 (N,4), dscale0 (N,3))
    cams: dicts with device tensors view, proj, campos and floats tx, ty (tan of the half FoVs).
    """

    def __init__(self, ctx: Context, m, pos0, sh0, opacity, ends_idx, ends_w, cams, W, H, bg, grid=None,
                 sample_idx=None, sample_w=None, lpf=0.0, sh_degree=3, axis_padding=1e-3, end_coeff=2.0):
        t = ctx._torch
        for name, x, dts in (("pos0", pos0, (t.float32,)), ("sh0", sh0, (t.float32,)),
                             ("opacity", opacity, (t.float32,)), ("ends_w", ends_w, (t.float32,)),
                             ("ends_idx", ends_idx, (t.int32, t.uint32)), ("bg", bg, (t.float32,)),
                             ("sample_w", sample_w, (t.float32,)), ("sample_idx", sample_idx, (t.int32, t.uint32))):
            if x is not None and x.dtype not in dts:
                raise ArapgsError(f"Chain: {name} must be {dts}, got {x.dtype}")
        self.ctx = ctx
        self.N, self.M = pos0.shape[0], m
        self.K = ends_w.numel() // (6 * self.N)
        self.V, self.W, self.H = len(cams), W, H
        self.S = 0 if grid is None else 64 * grid.n_valid
        self.grid = grid
        self._keep = [pos0, sh0, opacity, ends_idx, ends_w, bg, sample_idx, sample_w, cams]
        V = len(cams)
        self._vh = (P * V)(*[ptr(cm["view"]) for cm in cams])
        self._ph = (P * V)(*[ptr(cm["proj"]) for cm in cams])
        self._ch = (P * V)(*[ptr(cm["campos"]) for cm in cams])
        self._tx = (C.c_float * V)(*[cm["tx"] for cm in cams])
        self._ty = (C.c_float * V)(*[cm["ty"] for cm in cams])
        d = ChainDesc(self.N, self.K, m, ptr(pos0), ptr(sh0), ptr(opacity), ptr(ends_idx), ptr(ends_w),
                      grid.h if grid is not None else None, ptr(sample_idx), ptr(sample_w), lpf, V, W, H,
                      sh_degree, ptr(bg), C.cast(self._vh, P), C.cast(self._ph, P), C.cast(self._ch, P),
                      C.cast(self._tx, P), C.cast(self._ty, P), axis_padding, end_coeff)
        h = P()
        ctx._check(lib().arapgs_chain_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().arapgs_chain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, table, rot0, scale0, images=None, feat=None, dens=None):
        t = self.ctx._torch
        dev = rot0.device
        images = images or [t.empty((3, self.H, self.W), dtype=t.float32, device=dev) for _ in range(self.V)]
        if self.S > 0:
            feat = t.empty((self.S, 48), dtype=t.float32, device=dev) if feat is None else feat
            dens = t.empty(self.S, dtype=t.float32, device=dev) if dens is None else dens
        ih = (P * self.V)(*[ptr(i) for i in images])
        self.ctx._check(lib().arapgs_chain_forward(self.ctx.h, self.h, ptr(table), ptr(rot0), ptr(scale0),
                                                   C.cast(ih, P), ptr(feat), ptr(dens)))
        return images, feat, dens

    def backward(self, table, rot0, scale0, dimages=None, dfeat=None, ddens=None):
        t = self.ctx._torch
        dev = rot0.device
        dtable = t.empty((self.M, 16), dtype=t.float32, device=dev)
        drot0 = t.empty((self.N, 4), dtype=t.float32, device=dev)
        dscale0 = t.empty((self.N, 3), dtype=t.float32, device=dev)
        dih = None
        if dimages is not None:
            dih = (P * self.V)(*[ptr(i.contiguous()) for i in dimages])
            self._dkeep = dimages
        self.ctx._check(lib().arapgs_chain_backward(self.ctx.h, self.h, ptr(table), ptr(rot0), ptr(scale0),
                                                    C.cast(dih, P) if dih is not None else None, ptr(dfeat),
                                                    ptr(ddens), ptr(dtable), ptr(drot0), ptr(dscale0)))
        return dtable, drot0, dscale0

    def state(self):
        """Device pointers of the deformed (pos, rot, scale, sh, samples) as torch views."""
        t = self.ctx._torch
        ps = [P() for _ in range(5)]
        lib().arapgs_chain_state(self.h, *[C.byref(p) for p in ps])
        return ps


class Realign:
    """arapgs_realign: Adam steps of the density-preserving realignment over the
    raw Gaussian parameters (pos, rot, log scales, logit opacities, SH)."""

    def __init__(self, ctx: Context, grid, n, lr_pos=1e-5, lr_rot=1e-4, lr_scale=5e-4, lr_opacity=5e-3, lr_sh=2.5e-4,
                 beta1=0.9, beta2=0.999, eps=1e-15, w_density=1.0):
        # Adam moves every parameter by about lr per step whatever the loss scale;
        # where ~1000 Gaussians overlap a sample their steps add up, so the step
        # sizes here are a tenth of the 3DGS training rates.
        self.ctx, self.grid = ctx, grid
        d = RealignDesc(lr_pos, lr_rot, lr_scale, lr_opacity, lr_sh, beta1, beta2, eps, w_density)
        h = P()
        ctx._check(lib().arapgs_realign_create(ctx.h, grid.h, n, C.byref(d), C.byref(h)))
        self.h = h

    def set_options(self, moved=None, max_scale=None, conv_tol=0.0):
        """moved (u8 device tensor, N): only these Gaussians move and cells without
        them are static; max_scale (f32, N x 3): clamp of the activated scales;
        conv_tol: per-cell convergence (0: off). See arapgs_realign_set_options."""
        self._opts = (moved, max_scale)
        o = RealignOptions(ptr(moved), ptr(max_scale), float(conv_tol))
        self.ctx._check(lib().arapgs_realign_set_options(self.ctx.h, self.h, C.byref(o)))

    def cells(self):
        """(skip u8[n_valid], cell L1 loss f32[n_valid]) after the last step, on the host."""
        import numpy as np
        import torch
        nv, sk, cl = C.c_int64(0), P(), P()
        lib().arapgs_realign_cells(self.h, C.byref(nv), C.byref(sk), C.byref(cl))
        skip = np.empty(nv.value, np.uint8)
        loss = np.empty(nv.value, np.float32)
        torch.cuda.synchronize()
        if nv.value:
            _hip_memcpy_d2h(skip.ctypes.data, sk.value, skip.nbytes)
            _hip_memcpy_d2h(loss.ctypes.data, cl.value, loss.nbytes)
        return skip, loss

    def expand(self):
        """ExpandOpRange (GaussianView.cpp:4320-4360): converged cells next to a
        non-converged one (or that are one) are optimised again."""
        self.ctx._check(lib().arapgs_realign_expand(self.ctx.h, self.h))

    def step(self, samples, aim_feat, aim_dens, lpf, pos, rot, log_scale, logit_opacity, sh, loss=None):
        self.ctx._check(lib().arapgs_realign_step(self.ctx.h, self.h, ptr(samples), ptr(aim_feat), ptr(aim_dens), lpf,
                                                  ptr(pos), ptr(rot), ptr(log_scale), ptr(logit_opacity), ptr(sh),
                                                  ptr(loss)))

    def close(self):
        if self.h:
            lib().arapgs_realign_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def realign_loop(ctx, grid, realign, samples, aim_feat, aim_dens, lpf, pos, rot, log_scale, logit_opacity, sh, steps,
                 rebin_interval=10, loss=None, on_step=None, high_quality=False, ada_lpf_param=None):
    """The F6 handler of GaussianView (GaussianView.cpp:1695-1733) over the
    library: ExpandOpRange first when high_quality (:1699-1701); the adaptive
    low-pass filters from the current samples when ada_lpf_param is given
    (GetAdaLpfRatio, :1703-1705; enable_adaptive_lpf, GaussianView.hpp:180);
    the cell lists rebuilt from the Gaussians' current state before the loop
    (UpdateContainingRelationshipWithStaticGrids) and every rebin_interval
    steps after the first (UpdateContainingRelationship,
    _containing_relation_interval = 10 by default, GaussianView.hpp:427); and
    every step one arapgs_realign_step (GPUOptimize). on_step(i, loss) is
    called after each step when given (it reads the loss: one sync per step)."""
    import torch
    if high_quality:
        realign.expand()
    if ada_lpf_param is not None:
        grid.set_ada_lpf(ada_lpf_param, samples=samples)
    for i in range(steps):
        if i % max(1, rebin_interval) == 0:
            grid.rebin(pos, rot, torch.exp(log_scale), torch.sigmoid(logit_opacity))
        realign.step(samples, aim_feat, aim_dens, lpf, pos, rot, log_scale, logit_opacity, sh, loss=loss)
        if on_step is not None:
            on_step(i, loss)
