// 3DGS PLY codec (the on-disk format of GaussianView.cpp:43-339): binary
// little-endian vertices of 62 floats (x y z, nx ny nz, f_dc_0..2,
// f_rest_0..44 channel-major, opacity (logit), scale_0..2 (log), rot_0..3), or
// 63 with a trailing `index` float (RichPoint_Soup). The host reads or writes
// the file with one blocking transfer of the raw records; the per-Gaussian work
// runs on the GPU:
//   load  positions' bounding box, 3D Morton code of each position (21 bits
//         per axis of (p - min) / (max - min) * (2^21 - 1), as the reference),
//         stable radix sort by code, gather of the records in that order with
//         activations: unit quaternion, exp(scale), sigmoid(opacity), SH
//         re-interleaved to (coefficient * 3 + channel)
//   save  mask of the Gaussians inside the crop box and not flagged too small,
//         compaction (exclusive scan), inverse activations (log(scale),
//         logit(opacity)), SH back to channel-major
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "grid.hpp"
#include "ply.hpp"
#include "sort.hpp"

namespace arapgs {

namespace {

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

__global__ void __launch_bounds__(256) morton_kernel(int64_t n, int stride, const float* __restrict__ rec,
                                                     const float* __restrict__ mm, uint64_t* __restrict__ keys,
                                                     uint32_t* __restrict__ vals) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rec + (int64_t)stride * i;
    uint64_t code = 0;
    int xyz[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float rel = (r[c] - mm[c]) / (mm[3 + c] - mm[c]);
        xyz[c] = (int)((float)((1 << 21) - 1) * rel);
    }
    for (int b = 0; b < 21; b++) {
        code |= (uint64_t)(xyz[0] & (1 << b)) << (2 * b + 0);
        code |= (uint64_t)(xyz[1] & (1 << b)) << (2 * b + 1);
        code |= (uint64_t)(xyz[2] & (1 << b)) << (2 * b + 2);
    }
    keys[i] = code;
    vals[i] = (uint32_t)i;
}

__global__ void __launch_bounds__(256) positions_kernel(int64_t n, int stride, const float* __restrict__ rec,
                                                        float* __restrict__ pos) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int c = 0; c < 3; c++) pos[3 * i + c] = rec[(int64_t)stride * i + c];
}

__global__ void __launch_bounds__(256) iota_kernel(int64_t n, uint32_t* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) vals[i] = (uint32_t)i;
}

// record layout: 0-2 pos, 3-5 normal, 6-8 f_dc, 9-53 f_rest (channel-major: 15
// per channel), 54 opacity, 55-57 scale, 58-61 rot, 62 index (stride 63)
__global__ void __launch_bounds__(256) unpack_kernel(int64_t n, int stride, const float* __restrict__ rec,
                                                     const uint32_t* __restrict__ order, float* __restrict__ pos,
                                                     float* __restrict__ rot, float* __restrict__ scale,
                                                     float* __restrict__ opacity, float* __restrict__ sh,
                                                     int* __restrict__ index) {
#pragma clang fp contract(off)
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float* r = rec + (int64_t)stride * order[k];
#pragma unroll
    for (int c = 0; c < 3; c++) pos[3 * k + c] = r[c];
    float len2 = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; j++) len2 += r[58 + j] * r[58 + j];
    const float len = __builtin_sqrtf(len2);
#pragma unroll
    for (int j = 0; j < 4; j++) rot[4 * k + j] = r[58 + j] / len;
#pragma unroll
    for (int j = 0; j < 3; j++) scale[3 * k + j] = expf(r[55 + j]);
    opacity[k] = 1.0f / (1.0f + expf(-r[54]));
    float* s = sh + 48 * k;
    s[0] = r[6]; s[1] = r[7]; s[2] = r[8];
#pragma unroll
    for (int j = 1; j < 16; j++) {
        s[3 * j + 0] = r[9 + (j - 1)];
        s[3 * j + 1] = r[9 + (j - 1) + 15];
        s[3 * j + 2] = r[9 + (j - 1) + 30];
    }
    if (index) index[k] = stride == 63 ? (int)roundf(r[62]) : (int)order[k];
}

__global__ void __launch_bounds__(256) keep_kernel(int64_t n, const float* __restrict__ pos, float3 bmin, float3 bmax,
                                                   int use_box, const uint8_t* __restrict__ too_small,
                                                   uint32_t* __restrict__ keep) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    bool k = true;
    if (use_box && (x < bmin.x || y < bmin.y || z < bmin.z || x > bmax.x || y > bmax.y || z > bmax.z)) k = false;
    if (too_small && too_small[i]) k = false;
    keep[i] = k ? 1u : 0u;
}

__global__ void __launch_bounds__(256) pack_kernel(int64_t n, const uint32_t* __restrict__ keep,
                                                   const uint32_t* __restrict__ slot, const float* __restrict__ pos,
                                                   const float* __restrict__ rot, const float* __restrict__ scale,
                                                   const float* __restrict__ opacity, const float* __restrict__ sh,
                                                   float* __restrict__ rec) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    float* r = rec + 62 * (int64_t)slot[i];
#pragma unroll
    for (int c = 0; c < 3; c++) { r[c] = pos[3 * i + c]; r[3 + c] = 0.0f; }
    const float* s = sh + 48 * i;
    r[6] = s[0]; r[7] = s[1]; r[8] = s[2];
#pragma unroll
    for (int j = 1; j < 16; j++) {
        r[9 + (j - 1)] = s[3 * j + 0];
        r[9 + (j - 1) + 15] = s[3 * j + 1];
        r[9 + (j - 1) + 30] = s[3 * j + 2];
    }
    const float o = opacity[i];
    r[54] = logf(o / (1.0f - o));
#pragma unroll
    for (int j = 0; j < 3; j++) r[55 + j] = logf(scale[3 * i + j]);
#pragma unroll
    for (int j = 0; j < 4; j++) r[58 + j] = rot[4 * i + j];
}

}  // namespace

// ---------------------------------------------------------------- host side
int ply_parse_header(const char* path, PlyHeader* h, std::string* err) {
    FILE* f = std::fopen(path, "rb");
    if (!f) { *err = std::string("cannot open ") + path; return -1; }
    char line[512];
    int nline = 0, nprops = 0;
    bool binary_le = false, vertex = false, has_index = false;
    int64_t count = -1;
    const char* expect[] = {"x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"};
    std::vector<std::string> props;
    while (std::fgets(line, sizeof(line), f)) {
        nline++;
        std::string s(line);
        while (!s.empty() && (s.back() == '\n' || s.back() == '\r')) s.pop_back();
        if (nline == 1 && s != "ply") { std::fclose(f); *err = "not a PLY file"; return -1; }
        if (s == "format binary_little_endian 1.0") binary_le = true;
        else if (s.rfind("element vertex ", 0) == 0) { count = std::atoll(s.c_str() + 15); vertex = true; }
        else if (s.rfind("element ", 0) == 0) vertex = false;
        else if (s.rfind("property float ", 0) == 0 && vertex) props.push_back(s.substr(15));
        else if (s.rfind("property ", 0) == 0 && vertex) { std::fclose(f); *err = "only float vertex properties"; return -1; }
        else if (s == "end_header") break;
    }
    h->data_offset = std::ftell(f);
    std::fclose(f);
    if (!binary_le || count < 0) { *err = "need binary_little_endian vertices"; return -1; }
    nprops = (int)props.size();
    if (nprops != 62 && nprops != 63) { *err = "need the 3DGS layout of 62 (or 63 with index) float properties"; return -1; }
    for (int i = 0; i < 9; i++)
        if (props[i] != expect[i]) { *err = "unexpected property " + props[i]; return -1; }
    for (int i = 0; i < 45; i++)
        if (props[9 + i] != "f_rest_" + std::to_string(i)) { *err = "unexpected property " + props[9 + i]; return -1; }
    const char* tail[] = {"opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"};
    for (int i = 0; i < 8; i++)
        if (props[54 + i] != tail[i]) { *err = "unexpected property " + props[54 + i]; return -1; }
    if (nprops == 63) {
        if (props[62] != "index") { *err = "unexpected property " + props[62]; return -1; }
        has_index = true;
    }
    h->count = count;
    h->stride = nprops;
    h->has_index = has_index ? 1 : 0;
    return 0;
}

int launch_ply_unpack(int64_t n, int stride, const float* rec, int morton, uint32_t* ws, float* mm6, float* pos,
                      float* rot, float* scale, float* opacity, float* sh, int* index, hipStream_t st) {
    if (n <= 0) return 0;
    uint64_t* keys = reinterpret_cast<uint64_t*>(ws);
    uint64_t* keys_alt = keys + n;
    uint32_t* vals = reinterpret_cast<uint32_t*>(keys_alt + n);
    uint32_t* vals_alt = vals + n;
    uint32_t* mm_ws = vals_alt + n;           // 64 words for the bounding-box reduction
    uint32_t* sort_ws = mm_ws + 64;
    if (morton) {
        // bounding box on the positions (gathered into `pos`, overwritten by the unpack below)
        hipLaunchKernelGGL(positions_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, stride, rec, pos);
        int rc = launch_minmax(pos, n, mm_ws, mm6, st);
        if (rc) return rc;
        hipLaunchKernelGGL(morton_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, stride, rec, mm6, keys, vals);
        rc = launch_radix_sort(keys, vals, keys_alt, vals_alt, n, 63, sort_ws, st);
        if (rc) return rc;
    } else {
        hipLaunchKernelGGL(iota_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, vals);
    }
    hipLaunchKernelGGL(unpack_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, stride, rec, vals, pos, rot, scale,
                       opacity, sh, index);
    HIP_RET(hipGetLastError());
    return 0;
}

size_t ply_workspace_words(int64_t n) { return (size_t)n * 6 + 64 + radix_workspace_words(n) + 64; }

int launch_ply_pack(int64_t n, const float* pos, const float* rot, const float* scale, const float* opacity,
                    const float* sh, const float* boxmin_h, const float* boxmax_h, const uint8_t* too_small,
                    uint32_t* ws, float* rec, uint32_t* d_count, hipStream_t st) {
    if (n <= 0) return 0;
    uint32_t* keep = ws;
    uint32_t* slot = keep + n;
    uint32_t* scan_ws = slot + n;
    const int use_box = boxmin_h && boxmax_h;
    const float3 bmin = use_box ? make_float3(boxmin_h[0], boxmin_h[1], boxmin_h[2]) : make_float3(0, 0, 0);
    const float3 bmax = use_box ? make_float3(boxmax_h[0], boxmax_h[1], boxmax_h[2]) : make_float3(0, 0, 0);
    hipLaunchKernelGGL(keep_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, pos, bmin, bmax, use_box, too_small,
                       keep);
    int rc = launch_exclusive_scan(keep, n, slot, d_count, scan_ws, st);
    if (rc) return rc;
    hipLaunchKernelGGL(pack_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, keep, slot, pos, rot, scale, opacity, sh,
                       rec);
    HIP_RET(hipGetLastError());
    return 0;
}

size_t ply_pack_workspace_words(int64_t n) { return (size_t)n * 2 + scan_workspace_words(n) + 64; }

}  // namespace arapgs
