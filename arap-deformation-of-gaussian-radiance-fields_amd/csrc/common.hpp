// Shared host/device declarations of the arapgs HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_math.hpp"

#define KNN_MAX_T 12  // KNN_MAX of helper.hpp:48

#define HIP_RET(expr)                              \
    do {                                           \
        hipError_t _e = (expr);                    \
        if (_e != hipSuccess) return -(int)_e - 1000; \
    } while (0)

namespace arapgs {

// ----------------------------------------------------------------- knn.hip
int launch_knn(const float* nodes, int M, const float* queries, int64_t Q, int k, uint32_t* out_idx, float* out_w,
               uint32_t* scratch_list, uint32_t* scratch_count, hipStream_t st);
// Largest query count of arapgs_knn_noalloc (its static scratch: 32 MB per GPU)
// and the devices it serves.
constexpr int64_t KNN_NOALLOC_MAX_QUERIES = 1 << 23;
constexpr int KNN_NOALLOC_MAX_DEVICES = 64;
// Device address of this GPU's static KNN scratch (hipGetSymbolAddress on the
// first call per device; `device` must be current). Allocates nothing.
int knn_noalloc_scratch(int device, uint32_t** out);
int launch_fps(const float* pos, int64_t N, int m_steps, int* out, float* dist, unsigned long long* partial,
               int n_partial, unsigned int* counter, unsigned long long* best, hipStream_t st);

// ----------------------------------------------------------------- deform.hip
// Per-node deformation table: g (current position), R (row-major), t.
struct NodeXf {
    float g[3];
    float R[9];
    float t[3];
    float pad;
};
static_assert(sizeof(NodeXf) == 64, "NodeXf is one 64-byte record");


int launch_end_points(const float* pos, const float* rot, const float* scale, int64_t N, float pad, float coeff,
                      float* ends, hipStream_t st);
int launch_build_node_table(int M, const float* node_pos, const float* x_free, const int* free_id, NodeXf* table,
                            hipStream_t st);
int launch_blend_points(const float* pts, int64_t P, const uint32_t* idx, const float* w, int k, const NodeXf* table,
                        int M, const uint8_t* skip, float* out, hipStream_t st,
                        bool node_cache = true);
// The attributes a six-point update starts from when they are not the ones it
// writes (double-buffered Gaussians: read one set, write the other).
struct SixPointIn {
    const float* pos;
    const float* rot;
    const float* scale;
    const float* sh;
};
int launch_six_point(int64_t N, int k, const NodeXf* table, int M, const uint32_t* ends_idx, const float* ends_w,
                     float* ends, const float* scale0, const uint8_t* is_static, float pad, float coeff, float* pos,
                     float* rot, float* scale, float* sh, const ShRotTables* Td, hipStream_t st,
                     bool node_cache = true, const SixPointIn* in = nullptr);
int launch_node_quats(int M, const NodeXf* table, float* q, hipStream_t st);
// Node tables and node quaternions of up to SAMPLE_FRAMES_MAX consecutive
// frames, applied to the samples in order by one launch (K = 8 only; idx and w
// rows 16-byte aligned).
constexpr int SAMPLE_FRAMES_MAX = 8;
struct SampleFrames {
    int n;
    const NodeXf* table[SAMPLE_FRAMES_MAX];
    const float* q[SAMPLE_FRAMES_MAX];
};
int launch_samples_frames(int64_t S, int k, const SampleFrames& F, const uint32_t* idx, const float* w,
                          const uint8_t* is_static, float* pts, float* feat, const ShRotTables* Td, hipStream_t st,
                          bool compose = false);
int launch_sample_sh(int64_t S, int k, const float* w, const uint32_t* idx, const float* node_q,
                     const uint8_t* is_static, float* feat, const ShRotTables* Td, hipStream_t st);
int launch_sh_rotate(int64_t n, const float* R9, float* sh, const ShRotTables* Td, hipStream_t st);
int launch_polar(int64_t n, const double* M9, double* R9, double* Sdiag, hipStream_t st);

void build_shrot_tables(ShRotTables* T);

// ----------------------------------------------------------------- frame.hip
int launch_aims_translate(const float* node_pos, int m, const int* active, int n_active, const float d[3], float* aims,
                          hipStream_t st);
int launch_aims_twist(const float* node_pos, int m, const int* active, int n_active, const float axis[3], float radian,
                      float* aims, hipStream_t st);
int launch_aims_scale(const float* node_pos, int m, const int* active, int n_active, float sc, float* aims,
                      hipStream_t st);
int launch_aims_set(const float* node_pos, int m, const int* idx, int n, const float* vals, float* aims,
                    hipStream_t st);
int launch_static_rows(const uint8_t* node_static, const uint32_t* idx, int64_t n, int len, uint8_t* out,
                       hipStream_t st);
int launch_copy_f32(float* dst, const float* src, int64_t n, hipStream_t st);

}  // namespace arapgs
