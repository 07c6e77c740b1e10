// Scan and radix sort launchers (sort.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <algorithm>

namespace arapgs {

// Exclusive scan of n u32 values; *total receives the sum (device). ws holds
// scan_workspace_words(n) words.
size_t scan_workspace_words(int64_t n);
int launch_exclusive_scan(const uint32_t* in, int64_t n, uint32_t* out, uint32_t* total, uint32_t* ws,
                          hipStream_t st);

// Stable LSD radix sort of (key, value) pairs by key bits [0, bits). ws holds
// radix_workspace_words(n) words; keys_alt/vals_alt are n-element scratch.
size_t radix_workspace_words(int64_t n);
int launch_radix_sort(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, int64_t n, int bits,
                      uint32_t* ws, hipStream_t st);
// The same for u32 keys (same workspace).
int launch_radix_sort32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, int64_t n, int bits,
                        uint32_t* ws, hipStream_t st);

}  // namespace arapgs
