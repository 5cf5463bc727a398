// Shared host-side plumbing of libgpucad: status reporting, HIP error checks,
// per-device state (marching-cubes tables in device memory).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/gpucad.h"

static_assert(sizeof(gc_uint3) == sizeof(uint3), "gc_uint3 must mirror uint3");
static_assert(sizeof(gc_float3) == sizeof(float3), "gc_float3 must mirror float3");
static_assert(sizeof(gc_float4) == sizeof(float4), "gc_float4 must mirror float4");

namespace gc {

// Records `msg` as the calling thread's last error and returns `status`.
int fail(int status, const char *fmt, ...);
void clear_error();

#define GC_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t gc_err_ = (call);                                                  \
        if (gc_err_ != hipSuccess)                                                    \
            return ::gc::fail(GPUCAD_E_HIP, "%s failed: %s (%s:%d)", #call,          \
                              hipGetErrorString(gc_err_), __FILE__, __LINE__);        \
    } while (0)

#define GC_CHECK_LAUNCH() GC_HIP(hipGetLastError())

#define GC_ARG(cond, ...)                                                             \
    do {                                                                              \
        if (!(cond)) return ::gc::fail(GPUCAD_E_ARG, __VA_ARGS__);                    \
    } while (0)

// Device-resident copies of the marching-cubes tables (one set per device).
struct DeviceTables {
    uint64_t *tri_packed = nullptr;  // 256 x uint64: 15 edge nibbles + count in nibble 15
    uint8_t *num_verts = nullptr;    // 256
    uint8_t *tri_table = nullptr;    // 256*16 (255 terminated)
};

// Returns the tables of the current device, initialising them on first use.
int device_tables(const DeviceTables **out);

// Host tables (decoded once from the compact encoding in mc_tables.cpp).
const uint8_t *host_tri_table();     // 256*16
const uint8_t *host_num_verts();     // 256
const uint16_t *host_edge_mask();    // 256
const uint64_t *host_tri_packed();   // 256

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Shared pieces of the drop-in marching cubes (dropin_mc.hip): the grid
// layout check, the exclusive uint32 scan (enqueued on st) and the read of
// the last elements.
int check_dropin_grid(gc_uint3 s, gc_uint3 sh, gc_uint3 m, uint32_t n);
// With `comp`, the scan of a 0/1 occupancy also writes the compaction
// comp[out[i]] = i where in[i] != 0.
int dropin_scan_u32(uint32_t *out, const uint32_t *in, uint32_t n, hipStream_t st,
                    uint32_t *comp = nullptr);
// h[0] = a[n-1], h[1] = b[n-1] through pinned host memory, after the work on st
int dropin_read_last_pair(const uint32_t *a, const uint32_t *b, uint32_t n, uint32_t *h,
                          hipStream_t st);

// The orchestrations' classification and scans.  Bit p of `bits` (uint32
// words, bit p % 32 of word p / 32) is point p's side for the cube index;
// dropin_classify_scans classifies the n cells from it, scans occupancy and
// vertex counts, compacts the active cells (with each one's vertex offset in
// vbase) and leaves the totals (active voxels, vertices) in tot[0..1] and in
// host words 0-1.  With no active cell, verts_scan, comp and vbase are not
// written (and may be NULL).  With every later buffer present and at most
// kDropinPrefixMaxTiles scan tiles (dropin_prefix_path), each tile of the
// last pass sums the earlier tiles' sums itself, and the leading verts_scan
// entries (zero) of tiles before the first active cell are left to the
// generate launch: their count is tot[2] (see GenArgs::zfill).
constexpr int kDropinBitsWaveWords = 16;  // 64-bit words of point bits per wave
constexpr uint32_t kDropinPrefixMaxTiles = 4096;
struct DropinScratch {
    uint64_t *bits;          // the point bits (written by bits_blocks x 256 threads,
                             // 64 x kDropinBitsWaveWords points per wave), then a padding word
    uint32_t *bsum_o, *bsum_v, *tot, *vbase;  // tot: active, vertices, zero-filled entries
    uint32_t bits_blocks, nb;
};
inline bool dropin_prefix_path(const DropinScratch &w, bool later_null) {
    return !later_null && w.nb <= kDropinPrefixMaxTiles;
}
int dropin_scratch(gc_uint3 s, uint32_t n, DropinScratch *out);  // from the arena
int dropin_classify_scans(const DropinScratch &w, uint32_t *verts, uint32_t *occ,
                          uint32_t *occ_scan, uint32_t *verts_scan, uint32_t *comp, gc_uint3 s,
                          gc_uint3 sh, gc_uint3 m, uint32_t n, uint32_t *hw, hipStream_t st);
// workgroups of a persistent generate launch over at most n active cells
// (a multiple of the 8 XCDs)
uint32_t dropin_generate_blocks(uint32_t n);

// Scratch memory owned by the library for drop-in calls (grown on demand,
// one arena per device, never shrunk).  Callers hold lock_device_calls().
int scratch_arena(size_t bytes, void **out);
// 16 words of pinned host memory per device, written by kernels and read by
// the host after a stream synchronisation (allocated on first use or by
// gpucad_init).  Words 0-1 serve the drop-in orchestrations.
int host_words(uint32_t **out);

// The same tables and host words, without allocating: GPUCAD_E_ARG when
// gpucad_init has not prepared them on the current device.
int tables_if_ready(const DeviceTables **out);
int host_words_if_ready(uint32_t **out);

// Per-device lock held by the synchronous entry points that use the scratch
// arena or the host words for the whole call (recursive: entry points call
// each other's pieces).
std::unique_lock<std::recursive_mutex> lock_device_calls();

// Called by gpucad_init with the device current: resolves the address of the
// allocation-free drop-in's fixed scratch array on that device.
int noalloc_setup();

}  // namespace gc
