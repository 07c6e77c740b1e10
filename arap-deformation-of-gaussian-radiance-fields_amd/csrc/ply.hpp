// 3DGS PLY codec: host header parsing and the GPU unpack / pack (ply.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace arapgs {

struct PlyHeader {
    int64_t count = 0;
    int stride = 62;       // floats per vertex: 62, or 63 with the trailing index
    int has_index = 0;
    long data_offset = 0;  // byte offset of the first vertex
};

// Parses and checks the header of a binary little-endian 3DGS PLY.
int ply_parse_header(const char* path, PlyHeader* h, std::string* err);

// Records (n x stride floats, device) -> activated SoA in Morton order (morton
// != 0) or file order. mm6 (device f32[6]) receives min.xyz, max.xyz of the
// positions when morton != 0. index (optional): the records' `index` field
// (stride 63) or their file position. ws holds ply_workspace_words(n) words.
size_t ply_workspace_words(int64_t n);
int launch_ply_unpack(int64_t n, int stride, const float* rec, int morton, uint32_t* ws, float* mm6, float* pos,
                      float* rot, float* scale, float* opacity, float* sh, int* index, hipStream_t st);

// Activated SoA -> 62-float records of the Gaussians inside the crop box (host
// floats, optional) and not flagged in too_small (device, optional), in their
// order; *d_count (device) receives the number written. ws holds
// ply_pack_workspace_words(n) words.
size_t ply_pack_workspace_words(int64_t n);
int launch_ply_pack(int64_t n, const float* pos, const float* rot, const float* scale, const float* opacity,
                    const float* sh, const float* boxmin_h, const float* boxmax_h, const uint8_t* too_small,
                    uint32_t* ws, float* rec, uint32_t* d_count, hipStream_t st);

}  // namespace arapgs
