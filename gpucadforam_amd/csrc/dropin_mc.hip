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
// each, then the two totals: 2 MiB + 256 B per device, held until the
// process ends.  Host code never uses the array's name as an address:
// noalloc_setup resolves the array's address on each device and the entry
// point passes that pointer to the kernels.
constexpr uint32_t kNoAllocMaxTiles = 1u << 18;
constexpr uint32_t kNoAllocMaxVoxels = 1u << 30;
constexpr uint32_t kNoAllocFixedWords = 2 * kNoAllocMaxTiles + 64;
static __device__ __align__(256) uint32_t scratch[kNoAllocFixedWords];

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
// point bits: 64-bit words per wave, points per wave, per 256-thread block
constexpr int kBitsWaveWords = kDropinBitsWaveWords;
constexpr uint32_t kBitsWavePoints = 64 * kBitsWaveWords;
constexpr uint32_t kBitsBlockPoints = 4 * kBitsWavePoints;
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

// ---- the orchestrations' classify and scans from point occupancy bits
//
// Bit p of `bits` (uint32 words, bit p % 32 of word p / 32) is corner p's
// side of the surface.  Corner reads in the reference clamp to gridSize, and
// check_grid only admits layouts whose cells keep all eight corners inside
// the volume, so a cell's corners are the points lin, lin + 1, lin + gx,
// lin + gx + 1 and the same in the next plane (lin = z*gx*gy + y*gx + x).

// bit p of bits = vol[p] < iso (the cube-index test of classifyVoxel_lattice_new);
// each wave sets kBitsWavePoints consecutive bits
__global__ void __launch_bounds__(256) point_bits_kernel(uint64_t *bits, const float *vol,
                                                         uint32_t np, float iso) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    uint64_t mine = 0;
#pragma unroll
    for (int j = 0; j < kBitsWaveWords; ++j) {
        const uint64_t i = wave * kBitsWavePoints + j * 64 + lane;
        const uint64_t m = __ballot(i < np && vol[i] < iso);
        if (lane == (uint32_t)j) mine = m;
    }
    if (lane < kBitsWaveWords) bits[wave * kBitsWaveWords + lane] = mine;
}

// The cells of a scan tile in groups of 4 consecutive cells: thread t's
// groups start at cells tile + k * kGroupStep + 4t (k < kGroups), each
// kGroupStep cells on from the last, walked without a division per group.
constexpr int kGroups = kScanItems / 4;
constexpr uint32_t kGroupStep = kScanThreads * 4;

struct CellWalk {
    uint3 shift, mask;       // calcGridPos of the first group
    uint32_t gx, gxy;        // point strides: gridSize.x, gridSize.x * gridSize.y
    uint32_t dx, dy, dz;     // kGroupStep cells as steps along x, y, z
};

CellWalk cell_walk(gc_uint3 s, gc_uint3 sh, gc_uint3 m) {
    CellWalk w;
    w.shift = make_uint3(sh.x, sh.y, sh.z);
    w.mask = make_uint3(m.x, m.y, m.z);
    w.gx = s.x;
    w.gxy = s.x * s.y;
    const uint32_t plane = m.x * m.y;  // == gridSizeShift.z (check_grid)
    w.dz = kGroupStep / plane;
    w.dy = (kGroupStep % plane) / m.x;
    w.dx = (kGroupStep % plane) % m.x;
    return w;
}

__device__ __forceinline__ void walk(uint3 &p, uint32_t dx, uint32_t dy, uint32_t dz,
                                     const uint3 &mask) {
    // dx < mask.x and dy < mask.y, so one carry per axis suffices
    p.x += dx;
    p.y += dy;
    p.z += dz;
    if (p.x >= mask.x) {
        p.x -= mask.x;
        ++p.y;
    }
    if (p.y >= mask.y) {
        p.y -= mask.y;
        ++p.z;
    }
}

// 32 bits of the point bits from point p on
__device__ __forceinline__ uint32_t bit_window(const uint32_t *w, uint32_t p) {
    const uint32_t k = p >> 5;
    return __builtin_amdgcn_alignbit(w[k + 1], w[k], p & 31u);
}

// The count passes index their LDS copy of numVertsTable by a cell's corner
// bits in row order: bits 0-1 the (x, x+1) pair of row (y, z), 2-3 of row
// (y+1, z), 4-5 of (y, z+1), 6-7 of (y+1, z+1), each pair as it lies in the
// point bits.  The reference's cube index takes corners 3 and 2 (and 7 and
// 6) the other way round, so the copy holds nverts[row_order_cube(i)] at i
// (the swap is its own inverse).
__device__ __forceinline__ uint32_t row_order_cube(uint32_t i) {
    return (i & 0x33u) | ((i & 0x44u) << 1) | ((i & 0x88u) >> 1);
}

// corner bits in row order of the cell whose corner 0 is point lin
__device__ __forceinline__ uint32_t row_order_bits(const uint32_t *bits, uint32_t lin,
                                                   uint32_t gx, uint32_t gxy) {
    return bit_pair(bits, lin) | bit_pair(bits, lin + gx) << 2 | bit_pair(bits, lin + gxy) << 4 |
           bit_pair(bits, lin + gxy + gx) << 6;
}

// vertex counts of a thread's kGroups groups of 4 cells, group k starting at
// cell g0 + k * kGroupStep (g0 < n; cells at or beyond n count 0), from the
// row-order table s_nv: the four 32-bit bit windows of every group (one per
// corner row, from corner 0's point of its first cell) are loaded first, so
// all groups share one round trip, and each cell takes its 2-bit field of
// each; a group that crosses a plane (or whose rows lie far apart) reads its
// cells' corners one by one.  Cells past the end of a row are gridSize.x -
// gridSizeMask.x points further on (rows of at least 4 cells, so a group
// crosses at most one row end).
__device__ __forceinline__ void tile_counts(uint32_t (&nv)[kGroups][4], const uint32_t *bits,
                                            uint64_t g0, uint32_t n, const CellWalk &cw,
                                            const uint8_t *s_nv) {
    const uint32_t skip = cw.gx - cw.mask.x;
    const bool rows_near = skip <= 27 && cw.mask.x >= 4;
    uint3 p = calc_grid_pos((uint32_t)g0, cw.shift, cw.mask);
    uint3 pk[kGroups];
    uint32_t w[kGroups][4];
    bool fast[kGroups];
#pragma unroll
    for (int k = 0; k < kGroups; ++k) {
        const uint64_t c0 = g0 + (uint64_t)k * kGroupStep;
        pk[k] = p;
        fast[k] = p.x + 3 < cw.mask.x || (p.y + 1 < cw.mask.y && rows_near);
        // (a group past the cells reads the first windows and counts 0)
        const uint32_t lin = c0 < n ? p.z * cw.gxy + p.y * cw.gx + p.x : 0u;
        w[k][0] = bit_window(bits, lin);
        w[k][1] = bit_window(bits, lin + cw.gx);
        w[k][2] = bit_window(bits, lin + cw.gxy);
        w[k][3] = bit_window(bits, lin + cw.gxy + cw.gx);
        walk(p, cw.dx, cw.dy, cw.dz, cw.mask);
    }
#pragma unroll
    for (int k = 0; k < kGroups; ++k) {
        const uint64_t c0 = g0 + (uint64_t)k * kGroupStep;
        if (fast[k]) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const uint32_t sh = c + (pk[k].x + c >= cw.mask.x ? skip : 0u);
                const uint32_t idx = __builtin_amdgcn_ubfe(w[k][0], sh, 2) |
                                     __builtin_amdgcn_ubfe(w[k][1], sh, 2) << 2 |
                                     __builtin_amdgcn_ubfe(w[k][2], sh, 2) << 4 |
                                     __builtin_amdgcn_ubfe(w[k][3], sh, 2) << 6;
                nv[k][c] = c0 + c < n ? s_nv[idx] : 0u;
            }
        } else {
            uint3 q = pk[k];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                nv[k][c] = 0;
                if (c0 + c < n) {
                    const uint32_t lin = q.z * cw.gxy + q.y * cw.gx + q.x;
                    nv[k][c] = s_nv[row_order_bits(bits, lin, cw.gx, cw.gxy)];
                    walk(q, 1, 0, 0, cw.mask);
                }
            }
        }
    }
}

// the scan's first pass from the point bits: the tile's sums of vertex
// counts and occupancy
__global__ void __launch_bounds__(kScanThreads) tile_sums_bits_kernel(
    const uint32_t *bits, CellWalk cw, uint32_t n, const uint8_t *nverts, uint32_t *bsum_v,
    uint32_t *bsum_o) {
    constexpr int kWaves = kScanThreads / 64;
    static_assert(kScanThreads == 256, "one table entry per thread");
    __shared__ uint8_t s_nv[256];
    __shared__ uint32_t part[2][kWaves];
    s_nv[threadIdx.x] = nverts[row_order_cube(threadIdx.x)];  // the row-order copy
    __syncthreads();
    const uint64_t g0 = (uint64_t)blockIdx.x * kScanTile + 4u * threadIdx.x;
    uint32_t sv = 0, so = 0;
    if (g0 < n) {
        uint32_t nv[kGroups][4];
        tile_counts(nv, bits, g0, n, cw, s_nv);
#pragma unroll
        for (int k = 0; k < kGroups; ++k)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                sv += nv[k][c];
                so += nv[k][c] > 0;
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

typedef uint32_t U4nt __attribute__((ext_vector_type(4)));

// the outputs of the orchestrations' classify, scans and compaction
struct ScanOut {
    uint32_t *verts, *occ, *occ_scan, *verts_scan, *comp, *vbase;
    int vec;  // verts, occ, occ_scan and verts_scan are 16-byte aligned
};

__device__ __forceinline__ void store4(uint32_t *p, uint64_t c0, const uint32_t (&v)[4], bool vec,
                                       uint32_t n) {
    if (vec && c0 + 3 < n) {
        __builtin_nontemporal_store(U4nt{v[0], v[1], v[2], v[3]}, (U4nt *)(p + c0));
    } else {
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (c0 + c < n) p[c0 + c] = v[c];
    }
}

// the scans' third pass from the point bits: every array the reference's
// classify, scans and compaction write (verts_scan only when some voxel is
// active), each tile's cells classified again from the bits (in L2) rather
// than read back; vbase[j] = verts_scan of the j-th active cell.  The tile's
// offsets are bsum_o/bsum_v[tile], and tot[0] says whether any voxel is
// active; or, PREFIX, bsum_o/bsum_v hold each tile's own sums and the tile
// adds up those of the tiles before it (its loads issued with the bit
// windows'), writes its verts_scan entries once its own prefix shows an
// active voxel, and the first tile with an active voxel records in tot[2]
// how many leading entries (all zero) it left to the generate launch; the
// last tile writes the totals to tot[0..1] and host words 0-1.
template <bool PREFIX>
__global__ void __launch_bounds__(kScanThreads) classify_scan_bits_kernel(
    ScanOut o, const uint32_t *bits, CellWalk cw, uint32_t n, const uint8_t *nverts,
    const uint32_t *bsum_o, const uint32_t *bsum_v, uint32_t *tot, uint32_t *hw) {
    constexpr int kWaves = kScanThreads / 64;
    constexpr int kChunks = kGroups * kWaves;
    static_assert(kScanThreads == 256 && kChunks <= 64, "tile layout");
    __shared__ uint8_t s_nv[256];
    __shared__ uint32_t wtot[2][kChunks];
    __shared__ uint32_t s_pre[2][kWaves], s_any;
    // the earlier tiles' sums, four per thread and step (loads in flight
    // together with the table load and the bit windows below)
    uint32_t po = 0, pv = 0;
    if (PREFIX) {
        for (uint32_t i = threadIdx.x; i < blockIdx.x; i += 4 * kScanThreads) {
            uint32_t qo[4], qv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t j = i + u * kScanThreads;
                const bool in = j < blockIdx.x;
                qo[u] = in ? bsum_o[j] : 0u;
                qv[u] = in ? bsum_v[j] : 0u;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                po += qo[u];
                pv += qv[u];
            }
        }
    }
    s_nv[threadIdx.x] = nverts[row_order_cube(threadIdx.x)];  // the row-order copy
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63;
    if (PREFIX) {
        po = wave_sum(po);
        pv = wave_sum(pv);
        if (lane == 0) {
            s_pre[0][wave] = po;
            s_pre[1][wave] = pv;
        }
    }
    const uint64_t g0 = (uint64_t)blockIdx.x * kScanTile + 4u * threadIdx.x;
    uint32_t nv[kGroups][4];
#pragma unroll
    for (int k = 0; k < kGroups; ++k)
#pragma unroll
        for (int c = 0; c < 4; ++c) nv[k][c] = 0;
    if (g0 < n) tile_counts(nv, bits, g0, n, cw, s_nv);
    // chunk (k, wave) holds 256 consecutive cells, 4 per lane
    uint32_t ex_o[kGroups], ex_v[kGroups];
#pragma unroll
    for (int k = 0; k < kGroups; ++k) {
        uint32_t so = 0, sv = 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            so += nv[k][c] > 0;
            sv += nv[k][c];
        }
        const uint32_t io = wave_inclusive_scan(so), iv = wave_inclusive_scan(sv);
        ex_o[k] = io - so;
        ex_v[k] = iv - sv;
        if (lane == 63) {
            wtot[0][k * kWaves + wave] = io;
            wtot[1][k * kWaves + wave] = iv;
        }
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t vo = lane < kChunks ? wtot[0][lane] : 0u;
        const uint32_t vv = lane < kChunks ? wtot[1][lane] : 0u;
        const uint32_t io = wave_inclusive_scan(vo), iv = wave_inclusive_scan(vv);
        uint32_t bo, bv;  // the tile's offsets
        if (PREFIX) {
            bo = 0;
            bv = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                bo += s_pre[0][w];
                bv += s_pre[1][w];
            }
            const uint32_t ao = __shfl(io, 63, 64), av = __shfl(iv, 63, 64);
            if (lane == 0) {
                s_any = bo + ao != 0;
                if (bo == 0 && ao != 0) tot[2] = blockIdx.x * (uint32_t)kScanTile;
                if (blockIdx.x == gridDim.x - 1) {
                    tot[0] = bo + ao;
                    tot[1] = bv + av;
                    hw[0] = bo + ao;
                    hw[1] = bv + av;
                }
            }
        } else {
            bo = bsum_o[blockIdx.x];
            bv = bsum_v[blockIdx.x];
            if (lane == 0) s_any = tot[0] != 0;
        }
        if (lane < kChunks) {
            wtot[0][lane] = io - vo + bo;
            wtot[1][lane] = iv - vv + bv;
        }
    }
    __syncthreads();
    const bool any_active = s_any != 0;
    const bool vec = o.vec != 0;
#pragma unroll
    for (int k = 0; k < kGroups; ++k) {
        const uint64_t c0 = g0 + (uint64_t)k * kGroupStep;
        if (c0 >= n) continue;
        uint32_t eo = wtot[0][k * kWaves + wave] + ex_o[k];
        uint32_t ev = wtot[1][k * kWaves + wave] + ex_v[k];
        uint32_t occ[4], so[4], sv[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            occ[c] = nv[k][c] > 0;
            so[c] = eo;
            sv[c] = ev;
            if (occ[c] && c0 + c < n) {
                o.comp[eo] = (uint32_t)(c0 + c);
                if (o.vbase) o.vbase[eo] = ev;
            }
            eo += occ[c];
            ev += nv[k][c];
        }
        store4(o.verts, c0, nv[k], vec, n);
        store4(o.occ, c0, occ, vec, n);
        store4(o.occ_scan, c0, so, vec, n);
        if (any_active) store4(o.verts_scan, c0, sv, vec, n);
    }
}

// both scans' second pass (one block of kSumsThreads, each scanning a run
// of consecutive tile sums of both arrays): the totals (active voxels,
// vertices, wrapping as the reference's last element + last scan element)
// go to tot[0..1] and to the host words
constexpr int kSumsThreads = 1024;
__global__ void __launch_bounds__(kSumsThreads) scan_sums2_kernel(uint32_t *bsum_o,
                                                                  uint32_t *bsum_v, uint32_t nb,
                                                                  uint32_t *tot, uint32_t *hw) {
    constexpr int kWaves = kSumsThreads / 64;
    __shared__ uint32_t s_o[kWaves + 1], s_v[kWaves + 1];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t per = (nb + kSumsThreads - 1) / kSumsThreads;
    const uint32_t beg = min(nb, threadIdx.x * per), end = min(nb, beg + per);
    uint32_t ao = 0, av = 0;
    for (uint32_t i = beg; i < end; ++i) {
        ao += bsum_o[i];
        av += bsum_v[i];
    }
    const uint32_t io = wave_inclusive_scan(ao), iv = wave_inclusive_scan(av);
    if (lane == 63) {
        s_o[threadIdx.x >> 6] = io;
        s_v[threadIdx.x >> 6] = iv;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t vo = lane < kWaves ? s_o[lane] : 0u;
        const uint32_t vv = lane < kWaves ? s_v[lane] : 0u;
        const uint32_t jo = wave_inclusive_scan(vo), jv = wave_inclusive_scan(vv);
        if (lane < kWaves) {
            s_o[lane] = jo - vo;
            s_v[lane] = jv - vv;
        }
        if (lane == kWaves - 1) {
            s_o[kWaves] = jo;
            s_v[kWaves] = jv;
        }
    }
    __syncthreads();
    uint32_t ro = s_o[threadIdx.x >> 6] + io - ao, rv = s_v[threadIdx.x >> 6] + iv - av;
    for (uint32_t i = beg; i < end; ++i) {
        const uint32_t vo = bsum_o[i], vv = bsum_v[i];
        bsum_o[i] = ro;
        bsum_v[i] = rv;
        ro += vo;
        rv += vv;
    }
    if (threadIdx.x == 0) {
        tot[0] = s_o[kWaves];
        tot[1] = s_v[kWaves];
        hw[0] = s_o[kWaves];
        hw[1] = s_v[kWaves];
    }
}

// exclusive scan of a tile's kScanItems x kScanThreads values (element
// k * kScanThreads + threadIdx.x is vals[k]) given the tile's offset
__device__ __forceinline__ void tile_scan(const uint32_t (&vals)[kScanItems],
                                          uint32_t (&excl)[kScanItems], uint32_t offset,
                                          uint32_t *wtot) {
    constexpr int kWaves = kScanThreads / 64;
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        excl[k] = wave_inclusive_scan(vals[k]);
        if ((threadIdx.x & 63) == 63) wtot[k * kWaves + wave] = excl[k];
    }
    __syncthreads();
    if (threadIdx.x < kScanItems * kWaves) {
        const uint32_t v = wtot[threadIdx.x];
        wtot[threadIdx.x] = wave_inclusive_scan(v) - v + offset;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) excl[k] += wtot[k * kWaves + wave] - vals[k];
}

// both scans' third pass from verts alone (occ[i] = verts[i] > 0, as the
// classify wrote it): occ_scan, the compacted list (with, when vbase is set,
// each active voxel's verts_scan value beside it) and verts_scan, except that
// with no active voxel verts_scan is not written (the reference returns
// before that scan)
__global__ void __launch_bounds__(kScanThreads) scan_apply2_kernel(
    uint32_t *occ_scan, uint32_t *comp, uint32_t *verts_scan, uint32_t *vbase,
    const uint32_t *verts, uint32_t n, const uint32_t *bsum_o, const uint32_t *bsum_v,
    const uint32_t *tot) {
    constexpr int kWaves = kScanThreads / 64;
    __shared__ uint32_t wtot[2][kScanItems * kWaves];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x;
    uint32_t vals[kScanItems], occ[kScanItems], eo[kScanItems], ev[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + k * kScanThreads;
        vals[k] = i < n ? verts[i] : 0u;
        occ[k] = vals[k] != 0;
    }
    tile_scan(occ, eo, bsum_o[blockIdx.x], wtot[0]);
    tile_scan(vals, ev, bsum_v[blockIdx.x], wtot[1]);
    const bool any_active = tot[0] != 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + k * kScanThreads;
        if (i < n) {
            occ_scan[i] = eo[k];
            if (any_active) verts_scan[i] = ev[k];
            if (occ[k]) {
                comp[eo[k]] = (uint32_t)i;
                if (vbase) vbase[eo[k]] = ev[k];
            }
        }
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
// 3269-3416), on the edge from pa to pb with field values fa, fb, when `on`
// (the indicator pair selects the band): the endpoints are ordered so that
// f0 <= f1 (swapped when fb < fa); the lo level applies if f0 <= lo <= f1,
// else the hi level if f0 <= hi <= f1; at the level, the vertex snaps to the
// first endpoint when the level is within 5e-4 of f0, else to the second
// when within 5e-4 of f1, else to the first when f1 - f0 < 5e-4, and lies at
// t = (level - f0) / (f1 - f0) otherwise; with neither level, f1 == f0 gives
// t = (first endpoint's z == 0 ? 1 : 0) and anything else t = 0.  Without
// the band, t = 0 on the unordered endpoints.  Evaluated without branches
// (one division per edge, every lane busy), the same values the reference's
// branches give.
__device__ __forceinline__ float3 band_vertex(float3 pa, float3 pb, float fa, float fb, float lo,
                                              float hi, bool on) {
    const bool sw = fb < fa;
    const float f0 = sw ? fb : fa, f1 = sw ? fa : fb;
    const float3 P0 = sw ? pb : pa, P1 = sw ? pa : pb;
    const bool cl = (f1 >= lo) && (f0 <= lo);
    const bool hit = cl || ((f1 >= hi) && (f0 <= hi));
    const float lev = cl ? lo : hi;
    const bool s0 = fabsf(lev - f0) < 0.0005f;
    const bool s1 = fabsf(lev - f1) < 0.0005f;
    const bool s2 = fabsf(f1 - f0) < 0.0005f;
    const float td = (lev - f0) / (f1 - f0);
    const float tz = (f1 == f0) ? (P0.z == 0.0f ? 1.0f : 0.0f) : 0.0f;
    const float t = hit ? td : tz;
    const float3 q0 = on ? P0 : pa, q1 = on ? P1 : pb;
    const float3 r = lerp3(q0, q1, on ? t : 0.0f);
    if (on && hit && !s0 && s1) return P1;
    if (on && hit && (s0 || s2)) return P0;
    return r;
}

// vertexInterp3_new (two fields; indicator values 1/0 use field one, 2/0 field two)
__device__ __forceinline__ float3 interp3(float lo1, float hi1, float3 p0, float3 p1, float f0,
                                          float f1, float g0, float g1, float lo2, float hi2,
                                          float id0, float id1) {
    const bool on1 = (id0 == 1.f && id1 == 0.f) || (id0 == 0.f && id1 == 1.f);
    const bool on2 = (id0 == 2.f && id1 == 0.f) || (id1 == 2.f && id0 == 0.f);
    // (the two selections exclude each other)
    return band_vertex(p0, p1, on2 ? g0 : f0, on2 ? g1 : f1, on2 ? lo2 : lo1, on2 ? hi2 : hi1,
                       on1 || on2);
}

// vertexInterp2_new (one field)
__device__ __forceinline__ float3 interp2(float lo, float hi, float3 p0, float3 p1, float f0,
                                          float f1, float id0, float id1) {
    const bool on = (id0 == 1.f && id1 == 0.f) || (id0 == 0.f && id1 == 1.f);
    return band_vertex(p0, p1, f0, f1, lo, hi, on);
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
    const uint32_t *active_dev = nullptr;  // when set, the active voxel count is read here
    const uint32_t *vbase = nullptr;       // when set, scan[comp[i]] is read as vbase[i]
    // when set (and some voxel is active): scan[0 .. *zfill) are written with
    // 0 first (the leading verts_scan entries the prefix path leaves)
    const uint32_t *zfill = nullptr;
    const uint8_t *tri;  // 256*16
    const uint8_t *nverts;
};

typedef float F4nt __attribute__((ext_vector_type(4)));

// One wave of 64 active voxels (one per lane): each lane interpolates its
// 12 edge vertices into LDS as the reference's thread does; then the wave
// writes its voxels' vertices one vertex per lane.  Vertex j of a voxel goes
// to scan[voxel] + j with its triangle's normal, and a triangle is written
// only if its first index is below maxVerts - 3, as in the reference; with
// the scan of the voxels' vertex counts these are one contiguous range
// (active voxels are compacted in voxel order), so every store of pos and
// norm is coalesced (and non-temporal: nothing reads them back here).
//
// Batches of 64 active voxels: a launch of one workgroup per batch takes
// batch blockIdx.x; the orchestrations' persistent launch (active count on
// the device, gridDim.x a multiple of 8) gives each XCD (workgroup b runs on
// XCD b % 8) a contiguous eighth of the batches, so the corner gathers of
// its waves share its L2.  Field two is read only by a wave with a corner
// whose indicator is 2 (no other corner takes vertexInterp3_new's second
// branch).
template <int MODE>
__global__ void __launch_bounds__(64) generate_kernel(GenArgs a) {
    __shared__ float3 vl[12][65];  // (stride 65: the edges of one voxel in different banks)
    __shared__ uint8_t owner[64 * 15];  // output vertex -> lane
    __shared__ uint32_t s_first[64], s_base[64];
    __shared__ uint4 s_tri[64];  // each lane's triTable row
    const uint32_t lane = threadIdx.x;
    const uint32_t active = a.active_dev ? *a.active_dev : a.active;
    if (a.zfill && active) {
        const uint32_t z = *a.zfill;
        for (uint32_t i = blockIdx.x * 64 + lane; i < z; i += gridDim.x * 64)
            ((uint32_t *)a.scan)[i] = 0;  // (the caller's verts_scan)
    }
    const uint32_t nbat = (uint32_t)(((uint64_t)active + 63) / 64);
    uint32_t bat = blockIdx.x, bat_end = nbat, bat_step = gridDim.x;
    if (a.active_dev) {
        const uint32_t per = (nbat + 7) / 8, x = blockIdx.x & 7u;
        bat = x * per + (blockIdx.x >> 3);
        bat_end = min(nbat, x * per + per);
        bat_step = gridDim.x >> 3;
    }
    // each batch's voxel ids and vertex offsets are read one batch ahead
    uint32_t next_voxel = 0, next_vbase = 0;
    auto fetch = [&](uint32_t b) {
        const uint64_t j = (uint64_t)b * 64u + lane;
        if (b < bat_end && j < active) {
            next_voxel = a.comp[j];
            next_vbase = a.vbase ? a.vbase[j] : a.scan[next_voxel];
        }
    };
    fetch(bat);
    for (; bat < bat_end; bat += bat_step) {
        const uint64_t i = (uint64_t)bat * 64u + lane;
        const uint32_t voxel = next_voxel;
        uint32_t cube = 0, nv = 0, vbase = next_vbase;
        fetch(bat + bat_step);
        if (i < active) {
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
            const uint32_t gx = a.g.size.x, gxy = a.g.size.x * a.g.size.y;
            const uint32_t lin = gp.z * gxy + gp.y * gx + gp.x;
            float fid[8];
            corners_inside(a.vol, lin, gx, gxy, fid);
            cube = cube_of(fid, a.iso);
            nv = a.nverts[cube];
            if (MODE == GEN_STD) {
                // an active voxel has nv > 0 (src/MarchingCubes_kernel.cu:3103)
#pragma unroll
                for (int e = 0; e < 12; ++e)
                    vl[e][lane] = interp4(a.iso, v[c_edge_a[e]], v[c_edge_b[e]],
                                          fid[c_edge_a[e]], fid[c_edge_b[e]]);
            } else {
                float f1[8];
                corners_inside(a.vol1, lin, gx, gxy, f1);
                if (MODE == GEN_BAND2) {
                    bool two = false;
#pragma unroll
                    for (int k = 0; k < 8; ++k) two |= fid[k] == 2.f;
                    float f2[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                    if (__any(two)) corners_inside(a.vol2, lin, gx, gxy, f2);
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
                __builtin_nontemporal_store(F4nt{q.x, q.y, q.z, 1.0f},
                                            (F4nt *)(a.pos + index + r));
                __builtin_nontemporal_store(F4nt{n.x, n.y, n.z, 0.0f},
                                            (F4nt *)(a.norm + index + r));
            }
        }
        __syncthreads();  // the next batch reuses the LDS
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

// Both scans' second and third passes after a classify that left the tile
// sums in w.bsum_o / w.bsum_v.  When a buffer reached only with active
// voxels is NULL (later_null), the active count is read first and must be 0.
int scans(const DropinScratch &w, const uint32_t *verts, uint32_t *occ_scan, uint32_t *verts_scan,
          uint32_t *comp, uint32_t n, uint32_t *hw, bool later_null, hipStream_t st) {
    hipLaunchKernelGGL(scan_sums2_kernel, dim3(1), dim3(kSumsThreads), 0, st, w.bsum_o, w.bsum_v,
                       w.nb, w.tot, hw);
    GC_CHECK_LAUNCH();
    if (later_null) {
        GC_HIP(hipStreamSynchronize(st));
        GC_ARG(hw[0] == 0, "NULL buffer");
    }
    hipLaunchKernelGGL(scan_apply2_kernel, dim3(w.nb), dim3(kScanThreads), 0, st, occ_scan, comp,
                       verts_scan, w.vbase, verts, n, w.bsum_o, w.bsum_v, w.tot);
    GC_CHECK_LAUNCH();
    return GPUCAD_OK;
}

// On the prefix path (dropin_prefix_path) without the pass that scans the
// tile sums; otherwise with it and, where a later buffer is NULL, the active
// count read (it must be 0) before the third pass.
int classify_scans(const DropinScratch &w, uint32_t *verts, uint32_t *occ, uint32_t *occ_scan,
                   uint32_t *verts_scan, uint32_t *comp, gc_uint3 s, gc_uint3 sh, gc_uint3 m,
                   uint32_t n, const DeviceTables *t, uint32_t *hw, bool later_null,
                   hipStream_t st) {
    const CellWalk cw = cell_walk(s, sh, m);
    const uint32_t *bits = (const uint32_t *)w.bits;
    const bool prefix = dropin_prefix_path(w, later_null);
    hipLaunchKernelGGL(tile_sums_bits_kernel, dim3(w.nb), dim3(kScanThreads), 0, st, bits, cw, n,
                       t->num_verts, w.bsum_v, w.bsum_o);
    GC_CHECK_LAUNCH();
    if (!prefix) {
        hipLaunchKernelGGL(scan_sums2_kernel, dim3(1), dim3(kSumsThreads), 0, st, w.bsum_o,
                           w.bsum_v, w.nb, w.tot, hw);
        GC_CHECK_LAUNCH();
    }
    if (later_null) {
        GC_HIP(hipStreamSynchronize(st));
        GC_ARG(hw[0] == 0, "NULL buffer");
    }
    ScanOut o;
    o.verts = verts;
    o.occ = occ;
    o.occ_scan = occ_scan;
    o.verts_scan = verts_scan;
    o.comp = comp;
    o.vbase = w.vbase;
    auto a16 = [](const void *q) { return ((uintptr_t)q & 15u) == 0; };
    o.vec = a16(verts) && a16(occ) && a16(occ_scan) && a16(verts_scan);
    if (prefix)
        hipLaunchKernelGGL(classify_scan_bits_kernel<true>, dim3(w.nb), dim3(kScanThreads), 0, st,
                           o, bits, cw, n, t->num_verts, (const uint32_t *)w.bsum_o,
                           (const uint32_t *)w.bsum_v, w.tot, hw);
    else
        hipLaunchKernelGGL(classify_scan_bits_kernel<false>, dim3(w.nb), dim3(kScanThreads), 0,
                           st, o, bits, cw, n, t->num_verts, (const uint32_t *)w.bsum_o,
                           (const uint32_t *)w.bsum_v, w.tot, hw);
    GC_CHECK_LAUNCH();
    return GPUCAD_OK;
}

// Isosurface.cu:401-486 / 488-572 orchestration; MODE selects the generate.
// Classification goes through the point occupancy bits; both scans share
// one pass over the tile sums and one over verts; the totals stay on the
// device for the generate launch (a persistent grid that reads the active
// count), so the call synchronises once.  Caller-visible buffers end as the
// reference leaves them: with no active voxel, verts_scan, the compacted
// list, pos and norm are not written.
// With `fixed` (kNoAllocFixedWords words of device memory) nothing is
// allocated: the tables and host words must already exist, cells are
// classified from their corners directly and the generate gathers scan[].
constexpr uint32_t kGenerateBlocks = 256 * 12;  // a multiple of the 8 XCDs

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
    // buffers the reference reaches only when some voxel is active may be
    // NULL; the call then learns the active count before going on
    const bool later_null = !d_verts_scan || !d_comp || !pos || !norm ||
                            (MODE != GEN_STD && !vol1) || (MODE == GEN_BAND2 && !vol2);
    hipStream_t st = 0;
    int r = check_grid(s, sh, m, n);
    if (r) return r;
    const DeviceTables *t = nullptr;
    uint32_t *hw = nullptr;
    if ((r = fixed ? tables_if_ready(&t) : device_tables(&t))) return r;
    if ((r = fixed ? host_words_if_ready(&hw) : host_words(&hw))) return r;
    DropinScratch w = {};
    if (fixed) {
        w.nb = (uint32_t)(((uint64_t)n + kScanTile - 1) / kScanTile);
        w.bsum_o = fixed;
        w.bsum_v = fixed + kNoAllocMaxTiles;
        w.tot = fixed + 2 * kNoAllocMaxTiles;
        // classify (src/Isosurface.cu:408-420) with the first pass of both scans
        const GridDesc g = grid_desc(s, sh, m, gc_float3{1, 1, 1}, gc_float3{0, 0, 0});
        hipLaunchKernelGGL(classify_tile_kernel, dim3(w.nb), dim3(kScanThreads), 0, st, d_verts,
                           d_occ, vol, g, n, iso, t->num_verts, w.bsum_v, w.bsum_o);
        GC_CHECK_LAUNCH();
        if ((r = scans(w, d_verts, d_occ_scan, d_verts_scan, d_comp, n, hw, later_null, st)))
            return r;
    } else {
        if ((r = dropin_scratch(s, n, &w))) return r;
        hipLaunchKernelGGL(point_bits_kernel, dim3(w.bits_blocks), dim3(256), 0, st, w.bits, vol,
                           s.x * s.y * s.z, iso);
        GC_CHECK_LAUNCH();
        if ((r = classify_scans(w, d_verts, d_occ, d_occ_scan, d_verts_scan, d_comp, s, sh, m, n,
                                t, hw, later_null, st)))
            return r;
    }
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
    a.active = 0;
    a.active_dev = w.tot;
    a.vbase = w.vbase;
    a.max_verts = max_verts;
    a.tri = t->tri_table;
    a.nverts = t->num_verts;
    if (!fixed && dropin_prefix_path(w, later_null)) a.zfill = w.tot + 2;
    if (!later_null) {  // (otherwise no voxel is active)
        hipLaunchKernelGGL(generate_kernel<MODE>, dim3(dropin_generate_blocks(n)), dim3(64), 0,
                           st, a);
        GC_CHECK_LAUNCH();
    }
    GC_HIP(hipStreamSynchronize(st));
    *active = hw[0];
    *total = *active ? hw[1] : 0u;
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
int dropin_scratch(gc_uint3 s, uint32_t n, DropinScratch *out) {
    DropinScratch w = {};
    const uint64_t np = (uint64_t)s.x * s.y * s.z;
    w.nb = (uint32_t)(((uint64_t)n + kScanTile - 1) / kScanTile);
    w.bits_blocks = (uint32_t)((np + kBitsBlockPoints - 1) / kBitsBlockPoints);
    const size_t bits_u64 = (size_t)w.bits_blocks * (kBitsBlockPoints / 64) + 1;
    void *ws = nullptr;
    int r = scratch_arena(bits_u64 * 8 + (2 * (size_t)w.nb + 4 + (size_t)n) * 4, &ws);
    if (r) return r;
    w.bits = (uint64_t *)ws;
    w.bsum_o = (uint32_t *)(w.bits + bits_u64);
    w.bsum_v = w.bsum_o + w.nb;
    w.tot = w.bsum_v + w.nb;
    w.vbase = w.tot + 4;
    *out = w;
    return GPUCAD_OK;
}
int dropin_classify_scans(const DropinScratch &w, uint32_t *verts, uint32_t *occ,
                          uint32_t *occ_scan, uint32_t *verts_scan, uint32_t *comp, gc_uint3 s,
                          gc_uint3 sh, gc_uint3 m, uint32_t n, uint32_t *hw, hipStream_t st) {
    const DeviceTables *t = nullptr;
    int r = device_tables(&t);
    if (r) return r;
    return classify_scans(w, verts, occ, occ_scan, verts_scan, comp, s, sh, m, n, t, hw, false,
                          st);
}
uint32_t dropin_generate_blocks(uint32_t n) {
    const uint64_t max_batches = ((uint64_t)n + 63) / 64;
    return (uint32_t)std::min<uint64_t>(kGenerateBlocks, (max_batches + 7) / 8 * 8);
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
