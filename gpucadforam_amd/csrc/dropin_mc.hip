// Drop-in marching cubes: the reference's per-stage entry points and their
// orchestration (Isosurface::computeIsosurface_lattice / _latticeone), filling
// the caller's uint arrays and non-indexed float4 pos/norm buffers exactly as
// the reference does (src/Isosurface.cu:401-572, src/MarchingCubes_kernel.cu).
//
// Differences by design (documented deviations from reference launch bugs):
//  * every cell is classified (the reference's 2-D grid fold truncates beyond
//    65535 blocks, src/Isosurface.cu:411-415);
//  * an edge whose band interpolation takes none of the branches gets t = 0
//    (the reference leaves t uninitialised, src/MarchingCubes_kernel.cu:3274, 3598);
//  * lerp(a, b, t) is evaluated as fma(t, b - a, a) and cross products as
//    fma(a.y, b.z, -(a.z*b.y)), ... (fused forms chosen for this build).
#include "gc_common.hpp"
#include "wave_ops.hpp"

// Fixed scratch of gpucad_compute_isosurface_lattice_noalloc: the tile sums
// of its two scans (occupancy, then vertex counts), kNoAllocMaxTiles words
// each, 2 MiB per device, held until the process ends.  Host code never uses
// the array's name as an address: noalloc_setup resolves the array's address
// on each device and the entry point passes that pointer to the kernels.
constexpr uint32_t kNoAllocMaxTiles = 1u << 18;
constexpr uint32_t kNoAllocMaxVoxels = 1u << 30;
static __device__ __align__(256) uint32_t scratch[2 * kNoAllocMaxTiles];

namespace gc {
namespace {

struct GridDesc {
    uint3 size, shift, mask;
    float3 vs, center;
};

__device__ __forceinline__ uint3 calc_grid_pos(uint32_t i, uint3 shift, uint3 mask) {
    // src/MarchingCubes_kernel.cu:120-136
    uint32_t zq = i / shift.z;
    uint32_t zr = i % shift.z;
    return make_uint3(zr % mask.x, zr / mask.x, zq);
}

__device__ __forceinline__ float sample(const float *v, uint3 p, uint3 gs) {
    // src/MarchingCubes_kernel.cu:88-96 (clamped to gridSize)
    p.x = min(p.x, gs.x);
    p.y = min(p.y, gs.y);
    p.z = min(p.z, gs.z);
    uint32_t i = p.z * gs.x * gs.y + p.y * gs.x + p.x;
    return v[i];
}

__device__ __forceinline__ void corners(const float *v, uint3 g, uint3 gs, float f[8]) {
    f[0] = sample(v, g, gs);
    f[1] = sample(v, make_uint3(g.x + 1, g.y, g.z), gs);
    f[2] = sample(v, make_uint3(g.x + 1, g.y + 1, g.z), gs);
    f[3] = sample(v, make_uint3(g.x, g.y + 1, g.z), gs);
    f[4] = sample(v, make_uint3(g.x, g.y, g.z + 1), gs);
    f[5] = sample(v, make_uint3(g.x + 1, g.y, g.z + 1), gs);
    f[6] = sample(v, make_uint3(g.x + 1, g.y + 1, g.z + 1), gs);
    f[7] = sample(v, make_uint3(g.x, g.y + 1, g.z + 1), gs);
}

__device__ __forceinline__ uint32_t cube_of(const float f[8], float iso) {
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) c |= (uint32_t)(f[k] < iso) << k;
    return c;
}

__global__ void classify_kernel(uint32_t *verts, uint32_t *occ, const float *vol, GridDesc g,
                                uint32_t n, float iso, const uint8_t *nverts) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint3 p = calc_grid_pos(i, g.shift, g.mask);
    float f[8];
    corners(vol, p, g.size, f);
    uint32_t nv = nverts[cube_of(f, iso)];
    verts[i] = nv;
    occ[i] = nv > 0;
}

// ---- exclusive uint32 scan (wrapping), reduce-then-scan
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;
static_assert((uint64_t)kNoAllocMaxTiles * kScanTile == kNoAllocMaxVoxels,
              "the fixed scratch holds the tile sums of the largest allocation-free grid");

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t s[kScanThreads];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        uint32_t a = threadIdx.x >= (uint32_t)off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += a;
        __syncthreads();
    }
    uint32_t incl = s[threadIdx.x];
    *total = s[kScanThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ void scan_reduce_kernel(const uint32_t *in, uint32_t n, uint32_t *bsum) {
    uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t acc = 0;
    for (int k = 0; k < kScanItems; ++k) {
        uint64_t i = base + (uint64_t)k * kScanThreads + threadIdx.x;
        if (i < n) acc += in[i];
    }
    uint32_t total;
    block_exclusive_scan(acc, &total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ void scan_bsum_kernel(uint32_t *bsum, uint32_t nb) {
    // one block: exclusive scan of nb block sums
    uint32_t per = (nb + kScanThreads - 1) / kScanThreads;
    uint32_t beg = threadIdx.x * per, end = min(nb, beg + per);
    uint32_t acc = 0;
    for (uint32_t i = beg; i < end; ++i) acc += bsum[i];
    uint32_t total;
    uint32_t run = block_exclusive_scan(acc, &total);
    for (uint32_t i = beg; i < end; ++i) {
        uint32_t v = bsum[i];
        bsum[i] = run;
        run += v;
    }
}

// Tile elements are read and written as kScanItems chunks of kScanThreads
// consecutive elements (every access coalesced); element k * kScanThreads +
// threadIdx.x is scanned in wave order through shuffles, and the 4 waves x
// kScanItems chunk totals by the first wave.  With `comp`, also the
// compaction of the scanned 0/1 occupancy: comp[out[i]] = i where in[i] != 0.
__global__ void scan_apply_kernel(uint32_t *out, const uint32_t *in, uint32_t n,
                                  const uint32_t *bsum, uint32_t *comp) {
    constexpr int kWaves = kScanThreads / 64;
    __shared__ uint32_t wtot[kScanItems * kWaves];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x;
    const int wave = threadIdx.x >> 6;
    uint32_t vals[kScanItems], incl[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + k * kScanThreads;
        vals[k] = i < n ? in[i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        incl[k] = wave_inclusive_scan(vals[k]);
        if ((threadIdx.x & 63) == 63) wtot[k * kWaves + wave] = incl[k];
    }
    __syncthreads();
    if (threadIdx.x < kScanItems * kWaves) {
        const uint32_t v = wtot[threadIdx.x];
        const uint32_t e = wave_inclusive_scan(v) - v;
        wtot[threadIdx.x] = e + bsum[blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + k * kScanThreads;
        if (i < n) {
            const uint32_t e = wtot[k * kWaves + wave] + incl[k] - vals[k];
            out[i] = e;
            if (comp && vals[k]) comp[e] = (uint32_t)i;
        }
    }
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// classify_kernel over the scan's tiles, for the orchestrations: the cells
// of tile b are read and written as kScanItems chunks of kScanThreads
// consecutive cells, and the tile's sums of verts and occ go to bsum_v[b]
// and bsum_o[b] (the scan's first pass, without reading the arrays again)
__global__ void classify_tile_kernel(uint32_t *verts, uint32_t *occ, const float *vol, GridDesc g,
                                     uint32_t n, float iso, const uint8_t *nverts,
                                     uint32_t *bsum_v, uint32_t *bsum_o) {
    constexpr int kWaves = kScanThreads / 64;
    __shared__ uint32_t part[2][kWaves];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x;
    uint32_t sv = 0, so = 0;
#pragma unroll 4
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + k * kScanThreads;
        if (i < n) {
            const uint3 p = calc_grid_pos((uint32_t)i, g.shift, g.mask);
            float f[8];
            corners(vol, p, g.size, f);
            const uint32_t nv = nverts[cube_of(f, iso)];
            verts[i] = nv;
            occ[i] = nv > 0;
            sv += nv;
            so += nv > 0;
        }
    }
    sv = wave_sum(sv);
    so = wave_sum(so);
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = sv;
        part[1][threadIdx.x >> 6] = so;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tv = 0, to = 0;
        for (int w = 0; w < kWaves; ++w) {
            tv += part[0][w];
            to += part[1][w];
        }
        bsum_v[blockIdx.x] = tv;
        bsum_o[blockIdx.x] = to;
    }
}

__global__ void compact_kernel(uint32_t *comp, const uint32_t *occ, const uint32_t *occ_scan,
                               uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && occ[i]) comp[occ_scan[i]] = i;
}

// ---- generate
__device__ __forceinline__ float3 lerp3(float3 a, float3 b, float t) {
    return make_float3(__builtin_fmaf(t, b.x - a.x, a.x), __builtin_fmaf(t, b.y - a.y, a.y),
                       __builtin_fmaf(t, b.z - a.z, a.z));
}

__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(__builtin_fmaf(a.y, b.z, -(a.z * b.y)), __builtin_fmaf(a.z, b.x, -(a.x * b.z)),
                       __builtin_fmaf(a.x, b.y, -(a.y * b.x)));
}

// One band (src/MarchingCubes_kernel.cu:3593-3674 and the two halves of
// 3269-3416): with f0 <= f1 after the swap, lo then hi level, endpoint snapping
// at 5e-4.  Returns false when no branch applies (t stays as it was).
__device__ __forceinline__ bool band_t(float lo, float hi, float3 &p0, float3 &p1, float f0,
                                       float f1, float &t, float3 &snap, bool &snapped) {
    if (f1 < f0) {
        float3 tp = p1; p1 = p0; p0 = tp;
        float tf = f1; f1 = f0; f0 = tf;
    }
    if ((f1 >= lo) && (f0 <= lo)) {
        if (fabsf(lo - f0) < 0.0005f) { snap = p0; snapped = true; return true; }
        if (fabsf(lo - f1) < 0.0005f) { snap = p1; snapped = true; return true; }
        if (fabsf(f1 - f0) < 0.0005f) { snap = p0; snapped = true; return true; }
        t = (lo - f0) / (f1 - f0);
        return true;
    }
    if ((f1 >= hi) && (f0 <= hi)) {
        if (fabsf(hi - f0) < 0.0005f) { snap = p0; snapped = true; return true; }
        if (fabsf(hi - f1) < 0.0005f) { snap = p1; snapped = true; return true; }
        if (fabsf(f1 - f0) < 0.0005f) { snap = p0; snapped = true; return true; }
        t = (hi - f0) / (f1 - f0);
        return true;
    }
    if (f1 == f0) {
        t = (p0.z == 0.0f) ? 1.0f : 0.0f;
        return true;
    }
    return false;
}

// vertexInterp3_new (two fields; indicator values 1/0 use field one, 2/0 field two)
__device__ float3 interp3(float lo1, float hi1, float3 p0, float3 p1, float f0, float f1, float g0,
                          float g1, float lo2, float hi2, float id0, float id1) {
    float t = 0.0f;
    float3 snap;
    bool snapped = false;
    if ((id0 == 1.f && id1 == 0.f) || (id0 == 0.f && id1 == 1.f)) {
        band_t(lo1, hi1, p0, p1, f0, f1, t, snap, snapped);
        if (snapped) return snap;
    }
    if ((id0 == 2.f && id1 == 0.f) || (id1 == 2.f && id0 == 0.f)) {
        band_t(lo2, hi2, p0, p1, g0, g1, t, snap, snapped);
        if (snapped) return snap;
    }
    return lerp3(p0, p1, t);
}

// vertexInterp2_new (one field)
__device__ float3 interp2(float lo, float hi, float3 p0, float3 p1, float f0, float f1, float id0,
                          float id1) {
    float t = 0.0f;
    float3 snap;
    bool snapped = false;
    if ((id0 == 1.f && id1 == 0.f) || (id0 == 0.f && id1 == 1.f)) {
        band_t(lo, hi, p0, p1, f0, f1, t, snap, snapped);
        if (snapped) return snap;
    }
    return lerp3(p0, p1, t);
}

// vertexInterp4 (textbook)
__device__ __forceinline__ float3 interp4(float iso, float3 p0, float3 p1, float f0, float f1) {
    float t = (iso - f0) / (f1 - f0);
    return lerp3(p0, p1, t);
}

// edge e joins corners c_edge_a[e] and c_edge_b[e]; the loops over edges
// are unrolled, so corner arrays stay in registers
constexpr int c_edge_a[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
constexpr int c_edge_b[12] = {1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7};

enum { GEN_BAND2 = 0, GEN_BAND1 = 1, GEN_STD = 2 };

struct GenArgs {
    float4 *pos, *norm;
    const uint32_t *comp, *scan;
    const float *vol, *vol1, *vol2;
    GridDesc g;
    float iso, lo1, hi1, lo2, hi2;
    uint32_t active, max_verts;
    const uint8_t *tri;  // 256*16
    const uint8_t *nverts;
};

// One wave of 64 active voxels (one per lane): each lane interpolates its
// 12 edge vertices into LDS as the reference's thread does; then the wave
// writes its voxels' vertices one vertex per lane.  Vertex j of a voxel goes
// to scan[voxel] + j with its triangle's normal, and a triangle is written
// only if its first index is below maxVerts - 3, as in the reference; with
// the scan of the voxels' vertex counts these are one contiguous range
// (active voxels are compacted in voxel order), so every store of pos and
// norm is coalesced.
template <int MODE>
__global__ void __launch_bounds__(64) generate_kernel(GenArgs a) {
    __shared__ float3 vl[12][64];
    __shared__ uint8_t owner[64 * 15];  // output vertex -> lane
    __shared__ uint32_t s_first[64], s_base[64];
    __shared__ uint4 s_tri[64];  // each lane's triTable row
    const uint32_t lane = threadIdx.x;
    const uint32_t i = blockIdx.x * 64u + lane;
    uint32_t cube = 0, nv = 0, vbase = 0;
    if (i < a.active) {
        const uint32_t voxel = a.comp[i];
        vbase = a.scan[voxel];
        const uint3 gp = calc_grid_pos(voxel, a.g.shift, a.g.mask);
        float3 p;
        p.x = ((float)gp.x - a.g.center.x) * a.g.vs.x;
        p.y = ((float)gp.y - a.g.center.y) * a.g.vs.y;
        p.z = ((float)gp.z - a.g.center.z) * a.g.vs.z;
        const float3 vs = a.g.vs;
        float3 v[8];
        v[0] = p;
        v[1] = make_float3(p.x + vs.x, p.y + 0.f, p.z + 0.f);
        v[2] = make_float3(p.x + vs.x, p.y + vs.y, p.z + 0.f);
        v[3] = make_float3(p.x + 0.f, p.y + vs.y, p.z + 0.f);
        v[4] = make_float3(p.x + 0.f, p.y + 0.f, p.z + vs.z);
        v[5] = make_float3(p.x + vs.x, p.y + 0.f, p.z + vs.z);
        v[6] = make_float3(p.x + vs.x, p.y + vs.y, p.z + vs.z);
        v[7] = make_float3(p.x + 0.f, p.y + vs.y, p.z + vs.z);
        float fid[8];
        corners(a.vol, gp, a.g.size, fid);
        cube = cube_of(fid, a.iso);
        nv = a.nverts[cube];
        if (MODE == GEN_STD) {
            // an active voxel has nv > 0 (src/MarchingCubes_kernel.cu:3103)
#pragma unroll
            for (int e = 0; e < 12; ++e)
                vl[e][lane] = interp4(a.iso, v[c_edge_a[e]], v[c_edge_b[e]], fid[c_edge_a[e]],
                                      fid[c_edge_b[e]]);
        } else {
            float f1[8];
            corners(a.vol1, gp, a.g.size, f1);
            if (MODE == GEN_BAND2) {
                float f2[8];
                corners(a.vol2, gp, a.g.size, f2);
#pragma unroll
                for (int e = 0; e < 12; ++e) {
                    int ca = c_edge_a[e], cb = c_edge_b[e];
                    vl[e][lane] = interp3(a.lo1, a.hi1, v[ca], v[cb], f1[ca], f1[cb], f2[ca],
                                          f2[cb], a.lo2, a.hi2, fid[ca], fid[cb]);
                }
            } else {
#pragma unroll
                for (int e = 0; e < 12; ++e) {
                    int ca = c_edge_a[e], cb = c_edge_b[e];
                    vl[e][lane] =
                        interp2(a.lo1, a.hi1, v[ca], v[cb], f1[ca], f1[cb], fid[ca], fid[cb]);
                }
            }
        }
    }
    const uint32_t incl = wave_inclusive_scan(nv);
    const uint32_t first = incl - nv;
    const uint32_t total = __shfl(incl, 63, 64);
    s_first[lane] = first;
    s_base[lane] = vbase;
    s_tri[lane] = *(const uint4 *)(a.tri + cube * 16);
    for (uint32_t j = 0; j < nv; ++j) owner[first + j] = (uint8_t)lane;
    __syncthreads();
    for (uint32_t k = lane; k < total; k += 64) {
        const uint32_t o = owner[k];
        const uint32_t j = k - s_first[o];
        const uint32_t t0 = j - j % 3u;  // the triangle's first vertex
        const uint8_t *row = (const uint8_t *)&s_tri[o];
        const float3 a0 = vl[row[t0]][o];
        const float3 a1 = vl[row[t0 + 1]][o];
        const float3 a2 = vl[row[t0 + 2]][o];
        const float3 e0 = make_float3(a1.x - a0.x, a1.y - a0.y, a1.z - a0.z);
        const float3 e1 = make_float3(a2.x - a0.x, a2.y - a0.y, a2.z - a0.z);
        const float3 n = cross3(e0, e1);
        const uint32_t index = s_base[o] + t0;  // scan[voxel] + t0, uint arithmetic
        if (index < a.max_verts - 3u) {
            const uint32_t r = j - t0;
            const float3 q = r == 0 ? a0 : (r == 1 ? a1 : a2);
            a.pos[index + r] = make_float4(q.x, q.y, q.z, 1.0f);
            a.norm[index + r] = make_float4(n.x, n.y, n.z, 0.0f);
        }
    }
}

GridDesc grid_desc(gc_uint3 s, gc_uint3 sh, gc_uint3 m, gc_float3 vs, gc_float3 c) {
    GridDesc g;
    g.size = make_uint3(s.x, s.y, s.z);
    g.shift = make_uint3(sh.x, sh.y, sh.z);
    g.mask = make_uint3(m.x, m.y, m.z);
    g.vs = make_float3(vs.x, vs.y, vs.z);
    g.center = make_float3(c.x, c.y, c.z);
    return g;
}

// The reference derives cell positions from gridSizeShift/gridSizeMask
// (calcGridPos) and clamps corner reads to gridSize, not gridSize-1; only the
// layout initMC_two sets up (src/main.cu:2124-2138) keeps every corner inside
// the volume, so that layout is required here.
int check_grid(gc_uint3 s, gc_uint3 sh, gc_uint3 m, uint32_t n) {
    GC_ARG(s.x >= 2 && s.y >= 2 && s.z >= 2, "gridSize must be >= 2 per axis");
    GC_ARG((uint64_t)s.x * s.y * s.z <= 0xffffffffull, "grid exceeds 32-bit point indices");
    GC_ARG(m.x >= 1 && m.y >= 1 && m.z >= 1 && m.x <= s.x - 1 && m.y <= s.y - 1 && m.z <= s.z - 1,
           "gridSizeMask must be in [1, gridSize-1]");
    GC_ARG(sh.z == m.x * m.y, "gridSizeShift.z must equal gridSizeMask.x * gridSizeMask.y");
    GC_ARG((uint64_t)n <= (uint64_t)m.x * m.y * m.z, "numVoxels exceeds the cell count");
    return GPUCAD_OK;
}

int classify(uint32_t *verts, uint32_t *occ, const float *vol, gc_uint3 s, gc_uint3 sh, gc_uint3 m,
             uint32_t n, float iso, hipStream_t st) {
    GC_ARG(verts && occ && vol, "NULL buffer");
    int r = check_grid(s, sh, m, n);
    if (r) return r;
    const DeviceTables *t = nullptr;
    r = device_tables(&t);
    if (r) return r;
    if (n == 0) return GPUCAD_OK;
    GridDesc g = grid_desc(s, sh, m, gc_float3{1, 1, 1}, gc_float3{0, 0, 0});
    hipLaunchKernelGGL(classify_kernel, dim3((n + 255) / 256), dim3(256), 0, st, verts, occ, vol, g,
                       n, iso, t->num_verts);
    GC_CHECK_LAUNCH();
    return GPUCAD_OK;
}

// The scan's second and third passes, given the tile sums bsum[nb].
int scan_finish(uint32_t *out, const uint32_t *in, uint32_t n, uint32_t *bsum, uint32_t nb,
                hipStream_t st, uint32_t *comp) {
    hipLaunchKernelGGL(scan_bsum_kernel, dim3(1), dim3(kScanThreads), 0, st, bsum, nb);
    GC_CHECK_LAUNCH();
    hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(kScanThreads), 0, st, out, in, n, bsum,
                       comp);
    GC_CHECK_LAUNCH();
    return GPUCAD_OK;
}

int scan_u32(uint32_t *out, const uint32_t *in, uint32_t n, hipStream_t st,
             uint32_t *comp = nullptr) {
    GC_ARG(out && in, "NULL buffer");
    if (n == 0) return GPUCAD_OK;
    uint32_t nb = (uint32_t)(((uint64_t)n + kScanTile - 1) / kScanTile);
    void *ws = nullptr;
    int r = scratch_arena((size_t)nb * sizeof(uint32_t), &ws);
    if (r) return r;
    uint32_t *bsum = (uint32_t *)ws;
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(kScanThreads), 0, st, in, n, bsum);
    GC_CHECK_LAUNCH();
    return scan_finish(out, in, n, bsum, nb, st, comp);
}

int compact(uint32_t *comp, const uint32_t *occ, const uint32_t *occ_scan, uint32_t n,
            hipStream_t st) {
    GC_ARG(comp && occ && occ_scan, "NULL buffer");
    if (n == 0) return GPUCAD_OK;
    hipLaunchKernelGGL(compact_kernel, dim3((n + 255) / 256), dim3(256), 0, st, comp, occ,
                       occ_scan, n);
    GC_CHECK_LAUNCH();
    return GPUCAD_OK;
}

template <int MODE>
int generate(GenArgs &a, hipStream_t st, const DeviceTables *t = nullptr) {
    GC_ARG(a.pos && a.norm && a.comp && a.scan && a.vol, "NULL buffer");
    if (MODE != GEN_STD) GC_ARG(a.vol1, "volume_one is NULL");
    if (MODE == GEN_BAND2) GC_ARG(a.vol2, "volume_two is NULL");
    if (!t) {
        int r = device_tables(&t);
        if (r) return r;
    }
    a.tri = t->tri_table;
    a.nverts = t->num_verts;
    if (a.active == 0) return GPUCAD_OK;
    hipLaunchKernelGGL(generate_kernel<MODE>, dim3((a.active + 63) / 64), dim3(64), 0, st, a);
    GC_CHECK_LAUNCH();
    return GPUCAD_OK;
}

__global__ void last_pair_kernel(const uint32_t *a, const uint32_t *b, uint32_t n,
                                 uint32_t *out) {
    if (threadIdx.x == 0) {
        out[0] = a[n - 1];
        out[1] = b[n - 1];
    }
}

// h[0] = a[n-1], h[1] = b[n-1]: a one-thread kernel writes them into pinned
// host memory, read after the stream synchronisation (the reference copies
// each with its own cudaMemcpy, Isosurface.cu:424-436, 453-472)
int read_last_pair(const uint32_t *a, const uint32_t *b, uint32_t n, uint32_t *h,
                   hipStream_t st, uint32_t *hw = nullptr) {
    if (!hw) {
        int r = host_words(&hw);
        if (r) return r;
    }
    hipLaunchKernelGGL(last_pair_kernel, dim3(1), dim3(64), 0, st, a, b, n, hw);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(st));
    h[0] = hw[0];
    h[1] = hw[1];
    return GPUCAD_OK;
}

// Isosurface.cu:401-486 / 488-572 orchestration; MODE selects the generate.
// With `fixed` (2 * kNoAllocMaxTiles words of device memory) the scans'
// tile sums go there and the tables and host words must already exist:
// nothing is allocated.
template <int MODE>
int compute_isosurface(const float *vol, gc_float4 *pos, gc_float4 *norm, float iso, uint32_t n,
                       uint32_t *d_verts, uint32_t *d_verts_scan, uint32_t *d_occ,
                       uint32_t *d_occ_scan, gc_uint3 s, gc_uint3 sh, gc_uint3 m, gc_float3 vs,
                       gc_float3 c, uint32_t *active, uint32_t *total, uint32_t *d_comp,
                       uint32_t max_verts, const float *vol1, const float *vol2, float lo1,
                       float hi1, float lo2, float hi2, uint32_t *fixed = nullptr) {
    GC_ARG(active && total, "host outputs are NULL");
    GC_ARG(n > 0, "numVoxels must be > 0");
    GC_ARG(d_verts && d_occ && vol && d_occ_scan, "NULL buffer");
    hipStream_t st = 0;
    int r = check_grid(s, sh, m, n);
    if (r) return r;
    const DeviceTables *t = nullptr;
    uint32_t *hw = nullptr;
    if ((r = fixed ? tables_if_ready(&t) : device_tables(&t))) return r;
    if ((r = fixed ? host_words_if_ready(&hw) : host_words(&hw))) return r;
    // scratch: the tile sums of occ and verts
    const uint32_t nb = (uint32_t)(((uint64_t)n + kScanTile - 1) / kScanTile);
    uint32_t *bsum_o = fixed, *bsum_v = fixed + kNoAllocMaxTiles;
    if (!fixed) {
        void *ws = nullptr;
        if ((r = scratch_arena(2 * (size_t)nb * sizeof(uint32_t), &ws))) return r;
        bsum_o = (uint32_t *)ws;
        bsum_v = bsum_o + nb;
    }
    // classify (src/Isosurface.cu:408-420) with the first pass of both scans
    const GridDesc g = grid_desc(s, sh, m, gc_float3{1, 1, 1}, gc_float3{0, 0, 0});
    hipLaunchKernelGGL(classify_tile_kernel, dim3(nb), dim3(kScanThreads), 0, st, d_verts, d_occ,
                       vol, g, n, iso, t->num_verts, bsum_v, bsum_o);
    GC_CHECK_LAUNCH();
    // the occupancy scan also writes the compacted voxel list (nothing when
    // no voxel is active, as the reference's early return leaves it)
    if ((r = scan_finish(d_occ_scan, d_occ, n, bsum_o, nb, st, d_comp))) return r;
    uint32_t last[2] = {0, 0};
    if ((r = read_last_pair(d_occ, d_occ_scan, n, last, st, hw))) return r;
    *active = last[0] + last[1];
    if (*active == 0) {
        *total = 0;
        return GPUCAD_OK;
    }
    GC_ARG(d_verts_scan && d_comp, "NULL buffer");
    if ((r = scan_finish(d_verts_scan, d_verts, n, bsum_v, nb, st, nullptr))) return r;
    if ((r = read_last_pair(d_verts, d_verts_scan, n, last, st, hw))) return r;
    *total = last[0] + last[1];
    GenArgs a;
    a.pos = (float4 *)pos;
    a.norm = (float4 *)norm;
    a.comp = d_comp;
    a.scan = d_verts_scan;
    a.vol = vol;
    a.vol1 = vol1;
    a.vol2 = vol2;
    a.g = grid_desc(s, sh, m, vs, c);
    a.iso = iso;
    a.lo1 = lo1;
    a.hi1 = hi1;
    a.lo2 = lo2;
    a.hi2 = hi2;
    a.active = *active;
    a.max_verts = max_verts;
    r = generate<MODE>(a, st, t);
    if (r) return r;
    GC_HIP(hipStreamSynchronize(st));
    return GPUCAD_OK;
}

constexpr int kNoAllocDevices = 64;
uint32_t *g_fixed_scratch[kNoAllocDevices];  // the array's address on each device
std::mutex g_fixed_mu;

int fixed_scratch(uint32_t **out) {
    int dev = 0;
    GC_HIP(hipGetDevice(&dev));
    GC_ARG(dev >= 0 && dev < kNoAllocDevices, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_fixed_mu);
    GC_ARG(g_fixed_scratch[dev], "gpucad_init was not called on device %d", dev);
    *out = g_fixed_scratch[dev];
    return GPUCAD_OK;
}

}  // namespace

int noalloc_setup() {
    int dev = 0;
    GC_HIP(hipGetDevice(&dev));
    GC_ARG(dev >= 0 && dev < kNoAllocDevices, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_fixed_mu);
    if (!g_fixed_scratch[dev]) {
        void *p = nullptr;
        GC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
        g_fixed_scratch[dev] = (uint32_t *)p;
    }
    return GPUCAD_OK;
}

int check_dropin_grid(gc_uint3 s, gc_uint3 sh, gc_uint3 m, uint32_t n) {
    return check_grid(s, sh, m, n);
}
int dropin_scan_u32(uint32_t *out, const uint32_t *in, uint32_t n, hipStream_t st,
                    uint32_t *comp) {
    return scan_u32(out, in, n, st, comp);
}
int dropin_read_last_pair(const uint32_t *a, const uint32_t *b, uint32_t n, uint32_t *h,
                          hipStream_t st) {
    return read_last_pair(a, b, n, h, st);
}

}  // namespace gc

using namespace gc;

extern "C" {

int gpucad_classify_voxel_lattice_new(uint32_t *voxel_verts, uint32_t *voxel_occupied,
                                      const float *volume, gc_uint3 grid_size,
                                      gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                      uint32_t num_voxels, gc_float3 voxel_size, float iso) {
    clear_error();
    (void)voxel_size;
    int r = classify(voxel_verts, voxel_occupied, volume, grid_size, grid_size_shift,
                     grid_size_mask, num_voxels, iso, 0);
    if (r) return r;
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_exclusive_scan_u32(uint32_t *output, const uint32_t *input, uint32_t n) {
    clear_error();
    auto calls = lock_device_calls();  // shared scratch arena and host words
    int r = scan_u32(output, input, n, 0);
    if (r) return r;
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_compact_voxels(uint32_t *compacted, const uint32_t *voxel_occupied,
                          const uint32_t *voxel_occupied_scan, uint32_t num_voxels) {
    clear_error();
    int r = compact(compacted, voxel_occupied, voxel_occupied_scan, num_voxels, 0);
    if (r) return r;
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

static GenArgs gen_args(gc_float4 *pos, gc_float4 *norm, const uint32_t *comp,
                        const uint32_t *scan, const float *vol, gc_uint3 s, gc_uint3 sh,
                        gc_uint3 m, gc_float3 vs, gc_float3 c, float iso, uint32_t active,
                        uint32_t max_verts) {
    GenArgs a;
    a.pos = (float4 *)pos;
    a.norm = (float4 *)norm;
    a.comp = comp;
    a.scan = scan;
    a.vol = vol;
    a.vol1 = nullptr;
    a.vol2 = nullptr;
    a.g = grid_desc(s, sh, m, vs, c);
    a.iso = iso;
    a.lo1 = a.hi1 = a.lo2 = a.hi2 = 0.f;
    a.active = active;
    a.max_verts = max_verts;
    return a;
}

int gpucad_generate_triangles_lattice_new(gc_float4 *pos, gc_float4 *norm,
                                          const uint32_t *compacted,
                                          const uint32_t *num_verts_scanned, const float *volume,
                                          gc_uint3 grid_size, gc_uint3 grid_size_shift,
                                          gc_uint3 grid_size_mask, gc_float3 voxel_size,
                                          gc_float3 grid_center, float iso,
                                          uint32_t active_voxels, uint32_t max_verts,
                                          uint32_t total_verts, const float *volume_one,
                                          const float *volume_two, float isovalue1,
                                          float isovalue2, float iso1, float iso2) {
    clear_error();
    (void)total_verts;
    GenArgs a = gen_args(pos, norm, compacted, num_verts_scanned, volume, grid_size,
                         grid_size_shift, grid_size_mask, voxel_size, grid_center, iso,
                         active_voxels, max_verts);
    a.vol1 = volume_one;
    a.vol2 = volume_two;
    a.lo1 = isovalue1;
    a.hi1 = isovalue2;
    a.lo2 = iso1;
    a.hi2 = iso2;
    int r = generate<GEN_BAND2>(a, 0);
    if (r) return r;
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_generate_triangles_lattice_newone(gc_float4 *pos, gc_float4 *norm,
                                             const uint32_t *compacted,
                                             const uint32_t *num_verts_scanned,
                                             const float *volume, gc_uint3 grid_size,
                                             gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                             gc_float3 voxel_size, gc_float3 grid_center,
                                             float iso, uint32_t active_voxels,
                                             uint32_t max_verts, uint32_t total_verts,
                                             const float *volume_one, float isovalue1,
                                             float isovalue2) {
    clear_error();
    (void)total_verts;
    GenArgs a = gen_args(pos, norm, compacted, num_verts_scanned, volume, grid_size,
                         grid_size_shift, grid_size_mask, voxel_size, grid_center, iso,
                         active_voxels, max_verts);
    a.vol1 = volume_one;
    a.lo1 = isovalue1;
    a.hi1 = isovalue2;
    int r = generate<GEN_BAND1>(a, 0);
    if (r) return r;
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_generate_triangles_standard(gc_float4 *pos, gc_float4 *norm, const uint32_t *compacted,
                                       const uint32_t *num_verts_scanned, const float *volume,
                                       gc_uint3 grid_size, gc_uint3 grid_size_shift,
                                       gc_uint3 grid_size_mask, gc_float3 voxel_size,
                                       gc_float3 grid_center, float iso, uint32_t active_voxels,
                                       uint32_t max_verts) {
    clear_error();
    GenArgs a = gen_args(pos, norm, compacted, num_verts_scanned, volume, grid_size,
                         grid_size_shift, grid_size_mask, voxel_size, grid_center, iso,
                         active_voxels, max_verts);
    int r = generate<GEN_STD>(a, 0);
    if (r) return r;
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_compute_isosurface_lattice(const float *vol, gc_float4 *pos, gc_float4 *norm, float iso,
                                      uint32_t num_voxels, uint32_t *d_voxel_verts,
                                      uint32_t *d_voxel_verts_scan, uint32_t *d_voxel_occupied,
                                      uint32_t *d_voxel_occupied_scan, gc_uint3 grid_size,
                                      gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                      gc_float3 voxel_size, gc_float3 grid_center,
                                      uint32_t *active_voxels, uint32_t *total_verts,
                                      uint32_t *d_comp_voxel_array, uint32_t max_verts,
                                      const float *vol_one, const float *vol_two, float isovalue1,
                                      float isovalue2, float iso1, float iso2) {
    clear_error();
    auto calls = lock_device_calls();  // shared scratch arena and host words
    return compute_isosurface<GEN_BAND2>(
        vol, pos, norm, iso, num_voxels, d_voxel_verts, d_voxel_verts_scan, d_voxel_occupied,
        d_voxel_occupied_scan, grid_size, grid_size_shift, grid_size_mask, voxel_size, grid_center,
        active_voxels, total_verts, d_comp_voxel_array, max_verts, vol_one, vol_two, isovalue1,
        isovalue2, iso1, iso2);
}

int gpucad_compute_isosurface_lattice_noalloc(
    const float *vol, gc_float4 *pos, gc_float4 *norm, float iso, uint32_t num_voxels,
    uint32_t *d_voxel_verts, uint32_t *d_voxel_verts_scan, uint32_t *d_voxel_occupied,
    uint32_t *d_voxel_occupied_scan, gc_uint3 grid_size, gc_uint3 grid_size_shift,
    gc_uint3 grid_size_mask, gc_float3 voxel_size, gc_float3 grid_center, uint32_t *active_voxels,
    uint32_t *total_verts, uint32_t *d_comp_voxel_array, uint32_t max_verts, const float *vol_one,
    const float *vol_two, float isovalue1, float isovalue2, float iso1, float iso2) {
    clear_error();
    auto calls = lock_device_calls();  // one fixed scratch per device
    GC_ARG(num_voxels <= kNoAllocMaxVoxels,
           "numVoxels %u exceeds the allocation-free limit of %u voxels", num_voxels,
           kNoAllocMaxVoxels);
    uint32_t *fixed = nullptr;
    int r = fixed_scratch(&fixed);
    if (r) return r;
    return compute_isosurface<GEN_BAND2>(
        vol, pos, norm, iso, num_voxels, d_voxel_verts, d_voxel_verts_scan, d_voxel_occupied,
        d_voxel_occupied_scan, grid_size, grid_size_shift, grid_size_mask, voxel_size, grid_center,
        active_voxels, total_verts, d_comp_voxel_array, max_verts, vol_one, vol_two, isovalue1,
        isovalue2, iso1, iso2, fixed);
}

int gpucad_compute_isosurface_latticeone(const float *vol, gc_float4 *pos, gc_float4 *norm,
                                         float iso, uint32_t num_voxels, uint32_t *d_voxel_verts,
                                         uint32_t *d_voxel_verts_scan, uint32_t *d_voxel_occupied,
                                         uint32_t *d_voxel_occupied_scan, gc_uint3 grid_size,
                                         gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                         gc_float3 voxel_size, gc_float3 grid_center,
                                         uint32_t *active_voxels, uint32_t *total_verts,
                                         uint32_t *d_comp_voxel_array, uint32_t max_verts,
                                         const float *vol_one, float isovalue1, float isovalue2) {
    clear_error();
    auto calls = lock_device_calls();  // shared scratch arena and host words
    return compute_isosurface<GEN_BAND1>(
        vol, pos, norm, iso, num_voxels, d_voxel_verts, d_voxel_verts_scan, d_voxel_occupied,
        d_voxel_occupied_scan, grid_size, grid_size_shift, grid_size_mask, voxel_size, grid_center,
        active_voxels, total_verts, d_comp_voxel_array, max_verts, vol_one, nullptr, isovalue1,
        isovalue2, 0.f, 0.f);
}

int gpucad_compute_isosurface_standard(const float *vol, gc_float4 *pos, gc_float4 *norm, float iso,
                                       uint32_t num_voxels, uint32_t *d_voxel_verts,
                                       uint32_t *d_voxel_verts_scan, uint32_t *d_voxel_occupied,
                                       uint32_t *d_voxel_occupied_scan, gc_uint3 grid_size,
                                       gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                       gc_float3 voxel_size, gc_float3 grid_center,
                                       uint32_t *active_voxels, uint32_t *total_verts,
                                       uint32_t *d_comp_voxel_array, uint32_t max_verts) {
    clear_error();
    auto calls = lock_device_calls();  // shared scratch arena and host words
    return compute_isosurface<GEN_STD>(
        vol, pos, norm, iso, num_voxels, d_voxel_verts, d_voxel_verts_scan, d_voxel_occupied,
        d_voxel_occupied_scan, grid_size, grid_size_shift, grid_size_mask, voxel_size, grid_center,
        active_voxels, total_verts, d_comp_voxel_array, max_verts, nullptr, nullptr, 0.f, 0.f, 0.f,
        0.f);
}

}  // extern "C"
