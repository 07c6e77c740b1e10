// Launchers of the chain's reverse-mode kernels (adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "common.hpp"
#include "grid.hpp"

namespace arapgs {

int launch_zero_f32(float* p, int64_t n, hipStream_t st);

// Density backward over the grid's valid cells at the given (deformed) sample
// positions: gsamp (S*3, may be null) written, gdens (58 x N component-major:
// d mu 3, d Q 6, d sh 48, d opacity) accumulated (zeroed by the caller).
// dfeat / ddens may be null.
int launch_density_backward(int64_t nv, const int* valid, const int64_t* prefix, const int* list, const float* samples,
                            const float* pos, const float* rot, const float* scale, const float* opacity,
                            const float* sh, DensityFilter flt, const float* dfeat, const float* ddens, float* gsamp,
                            float* gdens, int64_t n_gauss, hipStream_t st, const uint8_t* skip = nullptr);

// Gradients of each Gaussian's own attributes from the density accumulators
// (58 x N): dpos N*3, drot N*4 (w.r.t. the stored rotation), dscale N*3,
// dopac N, dsh N*48.
int launch_density_attr_grad(int64_t N, const float* rot, const float* scale, const float* gd, DensityFilter flt, float* dpos,
                             float* drot, float* dscale, float* dopac, float* dsh, hipStream_t st);

// Per-Gaussian adjoint from (mu', q', s', sh') to the blended end points (gends,
// N*18), q0 (drot0, N*4) and s0 (dscale0, N*3). gr: N*58 renderer gradients;
// gd: density gradients (57 x N component-major) or null.
int launch_six_point_backward(int64_t N, int K, const NodeXf* table, const uint32_t* ends_idx, const float* ends_w,
                              const float* ends, const float* rot0, const float* scale0, const float* sh0,
                              const float* rot1, const float* scale1, const float* gr, const float* gd, DensityFilter flt,
                              float pad, float coeff, const ShRotTables* Tsh, float* gends, float* drot0,
                              float* dscale0, hipStream_t st);

// Node -> (end point | sample) x slot inverse index: entries (n_entries u32,
// sorted by node, stable) and per-node [begin, end) ranges. ws holds
// inverse_index_words(n_entries) words.
size_t inverse_index_words(int64_t n_entries);
int build_inverse_index(int64_t n_end_entries, int64_t n_entries, int M, const uint32_t* ends_idx,
                        const uint32_t* sample_idx, uint32_t* ws, uint32_t* entries, uint2* rng, hipStream_t st);

// d R_j, d t_j of every node (dtable, NodeXf layout, g and pad zero).
int launch_node_gather(int M, int K, int64_t n_end_entries, const uint2* rng, const uint32_t* entries,
                       const NodeXf* table, const float* ends_w, const float* ends0, const float* gends,
                       const float* sample_w, const float* samples0, const float* gsamp, NodeXf* dtable,
                       hipStream_t st);

// Realignment (density-preserving step): activation of the raw parameters,
// the L1 loss subgradient at the samples, one Adam step over the 59 raw
// parameters per Gaussian (lr: pos, rot, log-scale, logit-opacity, SH).
int launch_activate(int64_t N, const float* log_scale, const float* logit_opac, float* scale, float* opac,
                    hipStream_t st);
// cell_loss (nullable, one float per valid cell = 64 samples): that cell's L1 loss
int launch_l1_grad(int64_t S, const float* feat, const float* dens, const float* aim_feat, const float* aim_dens,
                   float w, float* dfeat, float* ddens, double* loss, float* cell_loss, hipStream_t st);
// stat[v] = 1 when valid cell v's list holds no Gaussian with moved[g] != 0
int launch_static_cells(int64_t nv, const int* valid, const int64_t* prefix, const int* list, const uint8_t* moved,
                        uint8_t* stat, hipStream_t st);
// first: cell_ref = cell_loss, conv = 0, skip = stat; else conv |= (tol > 0 &&
// cell_loss <= tol * cell_ref), skip = stat | conv
int launch_converge(int64_t nv, const float* cell_loss, float* cell_ref, const uint8_t* stat, float tol, int first,
                    uint8_t* conv, uint8_t* skip, hipStream_t st);
// ExpandOpRange: valid cells that are, or are face neighbours of, a
// non-converged valid cell lose their converged flag (map: G^3 bytes scratch).
int launch_expand_op_range(int64_t nv, const int* valid, int G, const uint8_t* stat, uint8_t* conv, uint8_t* skip,
                           uint8_t* map, hipStream_t st);
// moved (nullable): only Gaussians with moved[g] != 0 are updated; max_scale
// (nullable, N*3): the activated scales are clamped to it after the update
int launch_adam(int64_t N, const float* dpos, const float* drot, const float* dscale, const float* dopac,
                const float* dsh, const float* scale, const float* opac, float* pos, float* rot, float* log_scale,
                float* logit_opac, float* sh, float* m, float* v, const float lr[5], float b1, float b2, float eps,
                float bc1, float bc2, const uint8_t* moved, const float* max_scale, hipStream_t st);

}  // namespace arapgs
