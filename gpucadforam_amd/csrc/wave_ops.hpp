// Wave-level helpers shared by the drop-in kernels (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gc {

// Inclusive prefix sum over the 64 lanes of a wave (lane order), uint32
// wrapping; every lane of the wave must call it.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

}  // namespace gc
