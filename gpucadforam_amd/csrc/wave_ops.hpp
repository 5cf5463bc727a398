// Device helpers shared by the drop-in kernels (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gc {

// Inclusive prefix sum over the 64 lanes of a wave (lane order), uint32
// wrapping; every lane of the wave must call it.  DPP steps: shifts by 1, 2,
// 4 and 8 lanes within each row of 16 (lanes shifted in from outside the row
// add 0), then row 0's last lane added to row 1 and row 2's to row 3
// (row_bcast:15), then lane 31 added to rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    v += __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

// Point occupancy bits (uint32 words, bit p % 32 of word p / 32): bits p
// and p + 1 (word p / 32 + 1 must exist).
__device__ __forceinline__ uint32_t bit_pair(const uint32_t *w, uint32_t p) {
    const uint32_t k = p >> 5;
    return __builtin_amdgcn_alignbit(w[k + 1], w[k], p & 31u) & 3u;
}

__device__ __forceinline__ uint32_t swap2(uint32_t b) { return ((b & 1u) << 1) | (b >> 1); }

// The cube index of the cell whose corner 0 is point lin (corners in the
// reference's order: 0 (x,y,z), 1 (x+1,y,z), 2 (x+1,y+1,z), 3 (x,y+1,z), then
// the same at z+1), gx and gxy the point strides along y and z.
__device__ __forceinline__ uint32_t cube_from_bits(const uint32_t *bits, uint32_t lin, uint32_t gx,
                                                   uint32_t gxy) {
    const uint32_t b01 = bit_pair(bits, lin);
    const uint32_t b32 = bit_pair(bits, lin + gx);
    const uint32_t b45 = bit_pair(bits, lin + gxy);
    const uint32_t b76 = bit_pair(bits, lin + gxy + gx);
    return b01 | swap2(b32) << 2 | b45 << 4 | swap2(b76) << 6;
}

// The 8 corner values of a cell whose corners lie inside the volume (every
// cell of a layout the drop-in grid check admits: the reference's clamp to
// gridSize never applies), corner 0 at point lin, in the reference's corner
// order: each x-adjacent pair in one 8-byte load (4-byte aligned)
typedef float F2a4 __attribute__((ext_vector_type(2), aligned(4)));
__device__ __forceinline__ void corners_inside(const float *v, uint32_t lin, uint32_t gx,
                                               uint32_t gxy, float f[8]) {
    const F2a4 a = *(const F2a4 *)(v + lin), b = *(const F2a4 *)(v + lin + gx);
    const F2a4 c = *(const F2a4 *)(v + lin + gxy), d = *(const F2a4 *)(v + lin + gxy + gx);
    f[0] = a.x;
    f[1] = a.y;
    f[3] = b.x;
    f[2] = b.y;
    f[4] = c.x;
    f[5] = c.y;
    f[7] = d.x;
    f[6] = d.y;
}

}  // namespace gc
