// Per-frame glue kernels: drag aims (UpdateAimPosition / UpdateAimPositionTwist,
// GaussianView.cpp:2920-2983 with ReloadAimPositions, :5240-5247) and buffer
// copies, so a whole frame stays on the device.
#include "common.hpp"

namespace arapgs {

namespace {

// aims = node positions; active nodes += delta (float)
__global__ void aims_translate_kernel(const float* __restrict__ node_pos, int m, const int* __restrict__ active,
                                      int n_active, float dx, float dy, float dz, float* __restrict__ aims) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        aims[3 * i] = node_pos[3 * i];
        aims[3 * i + 1] = node_pos[3 * i + 1];
        aims[3 * i + 2] = node_pos[3 * i + 2];
    }
}
__global__ void aims_translate_active_kernel(const float* __restrict__ node_pos, const int* __restrict__ active,
                                             int n_active, float dx, float dy, float dz, float* __restrict__ aims) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_active) return;
    const int i = active[j];
    aims[3 * i] = node_pos[3 * i] + dx;
    aims[3 * i + 1] = node_pos[3 * i + 1] + dy;
    aims[3 * i + 2] = node_pos[3 * i + 2] + dz;
}

// PointRotateByAxis (helper.cpp:1077-1100) about the centroid of the active
// nodes (UpdateAimPositionTwist). One workgroup: the float centroid is summed
// in node order like the reference's loop.
__global__ void __launch_bounds__(256) aims_twist_kernel(const float* __restrict__ node_pos, int m,
                                                         const int* __restrict__ active, int n_active, float ax,
                                                         float ay, float az, float radian, float* __restrict__ aims) {
#pragma clang fp contract(off)
    __shared__ float cen[3];
    if (threadIdx.x == 0) {
        float cx = 0.0f, cy = 0.0f, cz = 0.0f;
        for (int j = 0; j < n_active; j++) {
            const int i = active[j];
            cx += node_pos[3 * i]; cy += node_pos[3 * i + 1]; cz += node_pos[3 * i + 2];
        }
        const float cnt = (float)n_active;
        cen[0] = cx / cnt; cen[1] = cy / cnt; cen[2] = cz / cnt;
    }
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        aims[3 * i] = node_pos[3 * i];
        aims[3 * i + 1] = node_pos[3 * i + 1];
        aims[3 * i + 2] = node_pos[3 * i + 2];
    }
    __syncthreads();
    const float cost = cosf(radian), sint = sinf(radian);
    const float norm = __builtin_sqrtf(ax * ax + ay * ay + az * az);
    const float x = ax / norm, y = ay / norm, z = az / norm;
    const float a = cen[0], b = cen[1], c = cen[2];
    for (int j = threadIdx.x; j < n_active; j += blockDim.x) {
        const int i = active[j];
        const float p0 = node_pos[3 * i], p1 = node_pos[3 * i + 1], p2 = node_pos[3 * i + 2];
        float o0 = (x * x * (1 - cost) + cost) * p0 + (x * y * (1 - cost) - z * sint) * p1 + (x * z * (1 - cost) + y * sint) * p2;
        float o1 = (y * x * (1 - cost) + z * sint) * p0 + (y * y * (1 - cost) + cost) * p1 + (y * z * (1 - cost) - x * sint) * p2;
        float o2 = (z * x * (1 - cost) - y * sint) * p0 + (z * y * (1 - cost) + x * sint) * p1 + (z * z * (1 - cost) + cost) * p2;
        o0 += (a * (y * y + z * z) - x * (b * y + c * z)) * (1 - cost) + (b * z - c * y) * sint;
        o1 += (b * (x * x + z * z) - y * (a * x + c * z)) * (1 - cost) + (c * x - a * z) * sint;
        o2 += (c * (x * x + y * y) - z * (a * x + b * y)) * (1 - cost) + (a * y - b * x) * sint;
        aims[3 * i] = o0; aims[3 * i + 1] = o1; aims[3 * i + 2] = o2;
    }
}

// UpdateAimPositionScale (GaussianView.cpp:2961-2983): aims of the active nodes
// scaled by s about their float centroid (summed in node order, one thread).
__global__ void __launch_bounds__(256) aims_scale_kernel(const float* __restrict__ node_pos, int m,
                                                         const int* __restrict__ active, int n_active, float sc,
                                                         float* __restrict__ aims) {
#pragma clang fp contract(off)
    __shared__ float cen[3];
    if (threadIdx.x == 0) {
        float cx = 0.0f, cy = 0.0f, cz = 0.0f;
        for (int j = 0; j < n_active; j++) {
            const int i = active[j];
            cx += node_pos[3 * i]; cy += node_pos[3 * i + 1]; cz += node_pos[3 * i + 2];
        }
        const float cnt = (float)n_active;
        cen[0] = cx / cnt; cen[1] = cy / cnt; cen[2] = cz / cnt;
    }
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
        aims[3 * i] = node_pos[3 * i];
        aims[3 * i + 1] = node_pos[3 * i + 1];
        aims[3 * i + 2] = node_pos[3 * i + 2];
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n_active; j += blockDim.x) {
        const int i = active[j];
#pragma unroll
        for (int c = 0; c < 3; c++) aims[3 * i + c] = cen[c] + sc * (node_pos[3 * i + c] - cen[c]);
    }
}

// aims = node_pos, then aims[idx[j]] = vals[j] for j in order (one thread, so
// a node listed twice keeps its last value, like the reference's loop).
__global__ void __launch_bounds__(256) aims_set_kernel(const float* __restrict__ node_pos, int m,
                                                       const int* __restrict__ idx, int n,
                                                       const float* __restrict__ vals, float* __restrict__ aims) {
    for (int i = threadIdx.x; i < 3 * m; i += blockDim.x) aims[i] = node_pos[i];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int j = 0; j < n; j++) {
            const int i = idx[j];
            if (i < 0 || i >= m) continue;
            aims[3 * i] = vals[3 * j];
            aims[3 * i + 1] = vals[3 * j + 1];
            aims[3 * i + 2] = vals[3 * j + 2];
        }
}

// CheckStaticSamples / CheckMovedGaussians (GaussianView.cpp:2023-2109): row r
// of n rows of len node ids is static when every id is a static node.
__global__ void __launch_bounds__(256) static_rows_kernel(const uint8_t* __restrict__ node_static,
                                                          const uint32_t* __restrict__ idx, int64_t n, int len,
                                                          uint8_t* __restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint8_t all = 1;
    for (int j = 0; j < len; j++)
        if (!node_static[idx[r * len + j]]) all = 0;
    out[r] = all;
}

__global__ void copy_f32_n_kernel(float* __restrict__ dst, const float* __restrict__ src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

int launch_aims_translate(const float* node_pos, int m, const int* active, int n_active, const float d[3], float* aims,
                          hipStream_t st) {
    hipLaunchKernelGGL(aims_translate_kernel, dim3((m + 255) / 256), dim3(256), 0, st, node_pos, m, active, n_active,
                       d[0], d[1], d[2], aims);
    if (n_active > 0)
        hipLaunchKernelGGL(aims_translate_active_kernel, dim3((n_active + 255) / 256), dim3(256), 0, st, node_pos,
                           active, n_active, d[0], d[1], d[2], aims);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_aims_twist(const float* node_pos, int m, const int* active, int n_active, const float axis[3], float radian,
                      float* aims, hipStream_t st) {
    hipLaunchKernelGGL(aims_twist_kernel, dim3(1), dim3(256), 0, st, node_pos, m, active, n_active, axis[0], axis[1],
                       axis[2], radian, aims);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_aims_scale(const float* node_pos, int m, const int* active, int n_active, float sc, float* aims,
                      hipStream_t st) {
    hipLaunchKernelGGL(aims_scale_kernel, dim3(1), dim3(256), 0, st, node_pos, m, active, n_active, sc, aims);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_aims_set(const float* node_pos, int m, const int* idx, int n, const float* vals, float* aims,
                    hipStream_t st) {
    hipLaunchKernelGGL(aims_set_kernel, dim3(1), dim3(256), 0, st, node_pos, m, idx, n, vals, aims);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_static_rows(const uint8_t* node_static, const uint32_t* idx, int64_t n, int len, uint8_t* out,
                       hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(static_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, node_static, idx, n,
                       len, out);
    HIP_RET(hipGetLastError());
    return 0;
}

int launch_copy_f32(float* dst, const float* src, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(copy_f32_n_kernel, dim3(std::min<int64_t>((n + 255) / 256, 1024)), dim3(256), 0, st, dst, src, n);
    HIP_RET(hipGetLastError());
    return 0;
}

}  // namespace arapgs
