// Native marching-cubes extraction with shared-edge vertex dedup.
//
// Classification follows the reference (cube bit k = [f(corner k) < iso],
// corner/edge numbering of src/MarchingCubes_kernel.cu:3449-3456, 3506-3517;
// triangle lists of src/tables.h).  The output is canonical:
//   vertex k   = k-th crossing grid edge in edge-id order 3*point + axis,
//                point = x + nx*(y + ny*z); each edge interpolated once from its
//                low end: t = (iso - f0)/(f1 - f0), v = fma(t, p1 - p0, p0);
//   triangle j = j-th triangle in cell order, triTable order inside a cell.
//
// The grid is seen through its sign words: bit b of word w of row r = y+ny*z
// is [f(64w+b, y, z) < iso].  A word (k = r*W + w) is the unit of
// bookkeeping: the lane that owns it derives the crossing edges of its 64
// points and the active cells of its 64 cells with 64-bit bit arithmetic.
// Work on active cells is balanced by compacting them, per wave, into an LDS
// list in canonical order and running one lane per active cell; a lane reads
// the word-level data of its cell from the owning lane (ds_bpermute).
//
// Launches (one stream, no host round trip, no inter-block waits):
//   signs : ballot of f < iso per 64 points; skipped when the field kernel
//           already produced the words
//   count : blocks of 256 words: triangles per wave (case of each active
//           cell), vertex and triangle sums per block, and each half word's
//           first vertex number within its block
//   scan  : one workgroup scans the block sums (64-bit), writes the totals
//           and the vertex/triangle range of each grid
//   emit  : the vertex and triangle stages in one launch, their workgroups
//           interleaved (both depend only on count and scan):
//     vert: each wave compacts its 64 words' crossing edges into an LDS list
//           in canonical order and emits them one lane per edge (loads of f0
//           and f1, interpolation); a wave's vertices form one contiguous
//           range, stored coalesced
//     tri : one lane per active cell computes the 12 edge vertex numbers from
//           popcounts of the crossing words of the four rows touching the cell
//           (plus the rows' first vertex numbers) and writes its triangles; a
//           wave's triangles form one contiguous range starting at the wave's
//           scanned base
//
// The kernels see grids as segments: a grid's words occupy a contiguous range
// of global blocks.  Every launch runs one workgroup per block (three per
// block for emit: its vertices and its two triangle halves); the many-grid
// form (batched and served) finds each block's grid in a block -> grid table.
// Both run the same block code, so a grid's mesh does not depend on what it
// is launched with.
#include "extract_types.hpp"

namespace gc {
namespace {

constexpr int kWaveVertCap = 1024; // per-wave LDS edge-list entries per round
constexpr int kEdgeStep = 4;   // edges per thread and step of the vertex emission

struct ExtractLayout {
    uint64_t signs_off, base_off, twave_off, bsum_off, ctl_off, res_off, bytes;
    uint64_t words, nb;
};

inline uint64_t align256(uint64_t v) { return (v + 255) & ~uint64_t(255); }

ExtractLayout layout(int nx, int ny, int nz) {
    ExtractLayout L;
    L.words = gpucad_sign_words(nx, ny, nz);
    L.nb = (L.words + kBlock - 1) / kBlock;
    uint64_t off = 0;
    L.signs_off = off;
    off = align256(off + L.words * 8);
    L.base_off = off;
    off = align256(off + L.nb * kBaseBytesPerBlock);
    L.twave_off = off;
    off = align256(off + L.nb * kTwaveBytesPerBlock);
    L.bsum_off = off;
    off = align256(off + L.nb * 16);
    L.ctl_off = off;  // [0] n_vertices [1] n_triangles [2] vertex numbers overflow 32 bits
    L.res_off = off + 64;
    off = align256(off + 64 + sizeof(SegResult));
    L.bytes = off;
    return L;
}

enum { CTL_NV = 0, CTL_NT = 1, CTL_OVF = 2 };

__device__ __forceinline__ uint64_t shift1(uint64_t cur, uint64_t next) {
    return (cur >> 1) | (next << 63);
}
__device__ __forceinline__ uint32_t popc64(uint64_t w) { return (uint32_t)__popcll(w); }
__device__ __forceinline__ uint64_t mask_below(int64_t n) {
    if (n >= 64) return ~0ull;
    if (n <= 0) return 0ull;
    return (1ull << n) - 1ull;
}
__device__ __forceinline__ uint32_t bit_at(uint64_t w, uint32_t b) {
    return (uint32_t)(w >> b) & 1u;
}
__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint32_t gather32(uint32_t v, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
}
__device__ __forceinline__ uint32_t readlane32(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}

// inclusive scan over the 64 lanes of a wave with DPP moves (no LDS): shifts
// by 1, 2, 4, 8 inside each row of 16 lanes, then the last lane of row 0 / 2
// is broadcast into rows 1 / 3 (row_bcast:15) and the last lane of row 1 into
// rows 2 and 3 (row_bcast:31).  Lanes without a source add the `old` operand, 0.
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v) {
    int x = (int)v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return (uint32_t)x;
}

// 12-byte records with 4-byte alignment (stored as one dwordx3 where possible)
struct V3 {
    float x, y, z;
};
struct T3 {
    uint32_t a, b, c;
};

// The mesh is written once and not read back by this launch: its 12-byte
// records are stored with the non-temporal hint (one dwordx3 store each).
typedef float F3nt __attribute__((ext_vector_type(3), aligned(4)));
typedef uint32_t U3nt __attribute__((ext_vector_type(3), aligned(4)));
__device__ __forceinline__ void store_nt(float *p, V3 v) {
    __builtin_nontemporal_store(F3nt{v.x, v.y, v.z}, (F3nt *)p);
}
__device__ __forceinline__ void store_nt(uint32_t *p, T3 t) {
    __builtin_nontemporal_store(U3nt{t.a, t.b, t.c}, (U3nt *)p);
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f) {
    uint32_t q = (uint32_t)((double)n * f.inv);
    const int64_t r = (int64_t)n - (int64_t)q * f.d;
    if (r < 0) --q;
    else if (r >= (int64_t)f.d) ++q;
    return q;
}

// sign word `idx` of a grid (32-bit word index: a grid has fewer than 2^32
// words) when `ok`, else 0.  The load is unconditional (word 0 stands in),
// so a run of these has no branches and its loads issue back to back.
__device__ __forceinline__ uint64_t ldw(const uint64_t *s, uint32_t idx, bool ok) {
    const uint64_t v = s[ok ? idx : 0u];
    return ok ? v : 0ull;
}

struct WordPos {
    int y, z, w;
};
__device__ __forceinline__ WordPos word_pos(uint64_t k, const Grid &g) {
    const uint32_t kk = (uint32_t)k;
    const uint32_t r = fdiv(kk, g.divW);
    const uint32_t z = fdiv(r, g.divNy);
    WordPos p;
    p.w = (int)(kk - r * (uint32_t)g.W);
    p.z = (int)z;
    p.y = (int)(r - z * (uint32_t)g.ny);
    return p;
}

struct Cross {
    uint64_t ex, ey, ez, sx, sy;
};

// Crossing words of the row with words (cur, nxt), +y neighbour (py, pyn) and
// +z neighbour pz; sx/sy: the same for point b+1.
__device__ __forceinline__ Cross row_cross(uint64_t cur, uint64_t nxt, uint64_t py, uint64_t pyn,
                                           uint64_t pz, bool has_y, bool has_z, uint64_t ved,
                                           uint64_t vpt, bool ved_n, bool vpt_n) {
    Cross c;
    c.ex = (cur ^ shift1(cur, nxt)) & ved;
    c.ey = has_y ? (cur ^ py) & vpt : 0ull;
    c.ez = has_z ? (cur ^ pz) & vpt : 0ull;
    const uint64_t exn = ved_n ? ((nxt ^ (nxt >> 1)) & 1ull) : 0ull;
    const uint64_t eyn = (has_y && vpt_n) ? ((nxt ^ pyn) & 1ull) : 0ull;
    c.sx = shift1(c.ex, exn);
    c.sy = shift1(c.ey, eyn);
    return c;
}

// ---------------------------------------------------------------------------
// grids and block scheduling
// ---------------------------------------------------------------------------
struct OneGrid {
    Seg s;
    __device__ __forceinline__ uint32_t seg_of(uint32_t) const { return 0; }
    __device__ __forceinline__ Seg get(uint32_t) const { return s; }
};

struct ManyGrids {
    const Seg *__restrict__ segs;
    const uint32_t *__restrict__ blk2seg;
    __device__ __forceinline__ uint32_t seg_of(uint32_t b) const {
        return (uint32_t)__builtin_amdgcn_readfirstlane((int)blk2seg[b]);
    }
    __device__ __forceinline__ Seg get(uint32_t i) const { return segs[i]; }
};

// ---------------------------------------------------------------------------
// signs
// ---------------------------------------------------------------------------
constexpr int kSignZ = 8;

__global__ void __launch_bounds__(256)
signs_kernel(const float *__restrict__ field, uint64_t *__restrict__ signs, int nx, int ny, int nz,
             int W, float iso) {
    const uint32_t lane = threadIdx.x & 63;
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int x = blockIdx.x * 64 + lane;
    const int z0 = blockIdx.z * kSignZ;
    if (y >= ny) return;
    const bool xin = x < nx;
    const uint64_t plane = (uint64_t)nx * ny;
    for (int z = z0; z < min(z0 + kSignZ, nz); ++z) {
        float f = xin ? field[(uint64_t)z * plane + (uint64_t)y * nx + x] : 0.f;
        uint64_t m = __ballot(xin && f < iso);
        if (lane == 0) signs[((uint64_t)z * ny + y) * W + blockIdx.x] = m;
    }
}

// ---------------------------------------------------------------------------
// count
// ---------------------------------------------------------------------------
// Triangles the case table gives the active cells of one word, from the
// corner words k0..k7 (bit x of kc: corner c of cell x is inside, f < iso;
// corners 0 (x,y,z) 1 (x+1,y,z) 2 (x+1,y+1,z) 3 (x,y+1,z) 4 (x,y,z+1)
// 5 (x+1,y,z+1) 6 (x+1,y+1,z+1) 7 (x,y+1,z+1), the case table's numbering,
// bit c of the case index).  For every one of the table's 255 cases
// with an outside corner, a cell with inside-corner set S has
//   |S| - 2 (faces of the cube with all four corners in S)
//       - 4 [S = the cube minus both ends of a body diagonal]
// triangles: each crossing edge is a vertex of exactly one polygon of the
// case and a polygon of k vertices is cut into k - 2 triangles, so the count
// is (3|S| - 2 edges in S) - 2 polygons; the polygons are as many as the
// Euler characteristic |S| - edges + faces of S, except in the four
// body-diagonal cases, which have two (characteristic 0).  The identity is
// checked case by case against the table by the tests.  So a word's count is
// a weighted sum of popcounts over its 64 cells, with no per-cell lookup.
// The sums run over all cells of the word, not only the active ones: a cell
// with every corner outside adds 0 to each of them, and one with every corner
// inside adds 8 - 2 * 6 = -4, which the last term takes back.  Corner bits of
// cells past the end of the row must be 0 (the caller masks them).
__device__ __forceinline__ uint32_t word_triangles(uint64_t k0, uint64_t k1, uint64_t k2,
                                                   uint64_t k3, uint64_t k4, uint64_t k5,
                                                   uint64_t k6, uint64_t k7) {
    const uint32_t v = popc64(k0) + popc64(k1) + popc64(k2) + popc64(k3) + popc64(k4) +
                       popc64(k5) + popc64(k6) + popc64(k7);
    const uint64_t p01 = k0 & k1, p23 = k2 & k3, p45 = k4 & k5, p67 = k6 & k7;
    const uint64_t p12 = k1 & k2, p56 = k5 & k6, p03 = k0 & k3, p47 = k4 & k7;
    const uint64_t bottom = p01 & p23, top = p45 & p67;
    const uint32_t f = popc64(bottom) + popc64(top) + popc64(p01 & p45) + popc64(p23 & p67) +
                       popc64(p12 & p56) + popc64(p03 & p47);
    // body diagonals (0,6) (1,7) (2,4) (3,5): both ends outside, the rest inside
    const uint32_t d = popc64(~(k0 | k6) & p12 & p45 & k3 & k7) +
                       popc64(~(k1 | k7) & p03 & p56 & k2 & k4) +
                       popc64(~(k2 | k4) & p01 & p67 & k3 & k5) +
                       popc64(~(k3 | k5) & p12 & p67 & k0 & k4);
    const uint32_t inside = popc64(bottom & top);
    return v - 2 * f - 4 * d + 4 * inside;
}

struct CountArgs {
    const uint64_t *signs;  // sign buffer base
    uint32_t *twave;
    uint64_t *bsum;
    uint32_t *base;  // block-local first vertex of each half word
};

template <class Src>
__global__ void __launch_bounds__(kBlock) count_kernel(Src src, CountArgs a) {
    __shared__ uint32_t s_sum[4][2];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    {
        const uint32_t gb = blockIdx.x;
        const Seg s = src.get(src.seg_of(gb));
        const Grid &g = s.g;
        const uint64_t *signs = a.signs + s.sign_off;
        const uint64_t k = (uint64_t)(gb - s.block_off) * kBlock + tid;
        uint32_t vc = 0, tc = 0, vlo = 0;
        if (k < g.words) {
            const WordPos p = word_pos(k, g);
            const bool has_y = p.y + 1 < g.ny, has_z = p.z + 1 < g.nz;
            const uint32_t kk = (uint32_t)k, W = (uint32_t)g.W, RW = (uint32_t)g.ny * W;
            const bool hw = p.w + 1 < g.W;
            const uint64_t wa = signs[kk];
            const uint64_t an = ldw(signs, kk + 1, hw);
            const uint64_t wb = ldw(signs, kk + W, has_y);
            const uint64_t bn = ldw(signs, kk + W + 1, has_y && hw);
            const uint64_t wc = ldw(signs, kk + RW, has_z);
            const uint64_t cn = ldw(signs, kk + RW + 1, has_z && hw);
            const uint64_t wd = ldw(signs, kk + RW + W, has_y && has_z);
            const uint64_t dn = ldw(signs, kk + RW + W + 1, has_y && has_z && hw);
            // all eight loads issue before any of the arithmetic on them
            __builtin_amdgcn_sched_barrier(0);
            const uint64_t sa = shift1(wa, an);
            const uint64_t sb = shift1(wb, bn);
            const uint64_t sc = shift1(wc, cn);
            const uint64_t vpt = mask_below((int64_t)g.nx - 64 * p.w);
            const uint64_t ved = mask_below((int64_t)g.nx - 1 - 64 * p.w);
            const uint64_t ex = (wa ^ sa) & ved;
            const uint64_t ey = has_y ? (wa ^ wb) & vpt : 0ull;
            const uint64_t ez = has_z ? (wa ^ wc) & vpt : 0ull;
            vc = popc64(ex) + popc64(ey) + popc64(ez);
            vlo = popc64(ex & 0xffffffffull) + popc64(ey & 0xffffffffull) +
                  popc64(ez & 0xffffffffull);
            if (has_y && has_z) {
                const uint64_t sd = shift1(wd, dn);
                tc = word_triangles(wa & ved, sa & ved, sb & ved, wb & ved, wc & ved, sc & ved,
                                    sd & ved, wd & ved);
            }
        }
        const uint32_t vincl = wave_scan_incl(vc);
        const uint32_t vw = readlane32(vincl, 63);
        const uint32_t tincl = wave_scan_incl(tc);
        const uint32_t tlo = readlane32(tincl, 31), tw = readlane32(tincl, 63);
        if (lane == 0) {
            s_sum[wave][0] = vw;
            s_sum[wave][1] = tw;
            // triangle counts of the wave's two groups of 32 words
            a.twave[kGroups * (uint64_t)gb + 2 * wave] = tlo;
            a.twave[kGroups * (uint64_t)gb + 2 * wave + 1] = tw - tlo;
        }
        __syncthreads();
        if (k < g.words) {
            // the word's first vertex within the block, for each 32-point half
            uint32_t v0 = vincl - vc;
            for (uint32_t i = 0; i < 4; ++i) v0 += i < wave ? s_sum[i][0] : 0u;
            *(uint2 *)(a.base + 2 * ((uint64_t)gb * kBlock + tid)) = make_uint2(v0, v0 + vlo);
        }
        if (tid == 0) {
            a.bsum[2 * (uint64_t)gb] =
                (uint64_t)s_sum[0][0] + s_sum[1][0] + s_sum[2][0] + s_sum[3][0];
            a.bsum[2 * (uint64_t)gb + 1] =
                (uint64_t)s_sum[0][1] + s_sum[1][1] + s_sum[2][1] + s_sum[3][1];
        }
    }
}

// ---------------------------------------------------------------------------
// scan of the block sums (one workgroup); ctl = totals and the overflow flag;
// res = each grid's vertex and triangle range
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024)
scan_blocks_kernel(uint64_t *bsum, uint64_t nb, uint64_t *ctl, uint64_t *d_counts,
                   const Seg *segs, uint32_t nseg, SegResult *res) {
    // Wave w scans the contiguous entries [w*per, (w+1)*per) in steps of 64
    // (lane l takes entry 64*step + l: coalesced 16-byte loads and stores).
    // A block's sums (at most 256 words x 192 crossing edges, or 256 x 64
    // cells x 5 triangles) and a step's (64 of them) fit 32 bits, so each
    // step is a DPP wave scan of 32-bit values with 64-bit carries.  The
    // first kKeep steps stay in registers for the write-back (8192 blocks:
    // a 512^3 grid); steps past them are read again and scanned again.
    constexpr int kWaves = 16, kKeep = 8;
    static_assert((uint64_t)64 * kBlock * 64 * 5 < (1ull << 32), "a step's sums fit 32 bits");
    __shared__ uint64_t sw_v[kWaves], sw_t[kWaves];
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    ulonglong2 *io = (ulonglong2 *)bsum;
    const uint64_t per = (nb + kWaves * 64 - 1) / (kWaves * 64) * 64;
    const uint64_t w0 = min(nb, (uint64_t)wave * per), w1 = min(nb, w0 + per);
    const uint32_t nsteps = (uint32_t)((w1 - w0 + 63) / 64);
    uint32_t qv[kKeep], qt[kKeep];
#pragma unroll
    for (int u = 0; u < kKeep; ++u) {
        const uint64_t i = w0 + 64 * (uint64_t)u + lane;
        const ulonglong2 q = i < w1 ? io[i] : make_ulonglong2(0ull, 0ull);
        qv[u] = (uint32_t)q.x;
        qt[u] = (uint32_t)q.y;
    }
    uint32_t ev[kKeep], et[kKeep], tv[kKeep], tt[kKeep];  // exclusive scans, step totals
    uint64_t sv = 0, st = 0;                              // the wave's totals
#pragma unroll
    for (int u = 0; u < kKeep; ++u) {
        const uint32_t iv = wave_scan_incl(qv[u]), it = wave_scan_incl(qt[u]);
        ev[u] = iv - qv[u];
        et[u] = it - qt[u];
        tv[u] = readlane32(iv, 63);
        tt[u] = readlane32(it, 63);
        sv += tv[u];
        st += tt[u];
    }
    for (uint32_t u = kKeep; u < nsteps; ++u) {
        const uint64_t i = w0 + 64 * (uint64_t)u + lane;
        const ulonglong2 q = i < w1 ? io[i] : make_ulonglong2(0ull, 0ull);
        sv += readlane32(wave_scan_incl((uint32_t)q.x), 63);
        st += readlane32(wave_scan_incl((uint32_t)q.y), 63);
    }
    if (lane == 0) {
        sw_v[wave] = sv;
        sw_t[wave] = st;
    }
    __syncthreads();
    uint64_t rv = 0, rt = 0, carry_v = 0, carry_t = 0;
    for (uint32_t w = 0; w < kWaves; ++w) {
        if (w < wave) {
            rv += sw_v[w];
            rt += sw_t[w];
        }
        carry_v += sw_v[w];
        carry_t += sw_t[w];
    }
#pragma unroll
    for (int u = 0; u < kKeep; ++u) {
        const uint64_t i = w0 + 64 * (uint64_t)u + lane;
        if (i < w1) io[i] = make_ulonglong2(rv + ev[u], rt + et[u]);
        rv += tv[u];
        rt += tt[u];
    }
    for (uint32_t u = kKeep; u < nsteps; ++u) {
        const uint64_t i = w0 + 64 * (uint64_t)u + lane;
        const ulonglong2 q = i < w1 ? io[i] : make_ulonglong2(0ull, 0ull);
        const uint32_t iv = wave_scan_incl((uint32_t)q.x), it = wave_scan_incl((uint32_t)q.y);
        if (i < w1) io[i] = make_ulonglong2(rv + (iv - (uint32_t)q.x), rt + (it - (uint32_t)q.y));
        rv += readlane32(iv, 63);
        rt += readlane32(it, 63);
    }
    __syncthreads();
    // every thread holds the totals; the scanned sums are visible to the workgroup
    int any_ovf = 0;
    for (uint32_t i = t; i < nseg; i += 1024) {
        const uint64_t b0 = segs ? segs[i].block_off : 0;
        const uint64_t b1 = (segs && i + 1 < nseg) ? segs[i + 1].block_off : nb;
        // one grid: its first block's prefix is 0 (no read back)
        const uint64_t v0 = segs ? bsum[2 * b0] : 0ull, t0 = segs ? bsum[2 * b0 + 1] : 0ull;
        const uint64_t v1 = b1 < nb ? bsum[2 * b1] : carry_v;
        const uint64_t t1 = b1 < nb ? bsum[2 * b1 + 1] : carry_t;
        SegResult r;
        r.v_off = v0;
        r.nv = v1 - v0;
        r.t_off = t0;
        r.nt = t1 - t0;
        r.ovf = r.nv > 0xffffffffull ? 1u : 0u;
        r.pad = 0;
        res[i] = r;
        any_ovf |= (int)r.ovf;
    }
    any_ovf = __syncthreads_or(any_ovf);
    if (t == 0) {
        ctl[CTL_NV] = carry_v;
        ctl[CTL_NT] = carry_t;
        ctl[CTL_OVF] = any_ovf ? 1 : 0;
        if (d_counts) {
            d_counts[0] = carry_v;
            d_counts[1] = carry_t;
        }
    }
}

// ---------------------------------------------------------------------------
// vertices (+ per-word first vertex numbers)
// ---------------------------------------------------------------------------
// vertex on the edge from point (xf, yf, zf) along axis ax with end values f0, f1:
// t = (iso - f0)/(f1 - f0), p = fma(t, p1 - p0, p0) on that axis
__device__ __forceinline__ V3 edge_vertex(const Seg &s, float xf, float yf, float zf, uint32_t ax,
                                          float f0, float f1) {
    const float t = (s.iso - f0) / (f1 - f0);
    const float px = __builtin_fmaf(xf, s.sx, s.ox);
    const float py = __builtin_fmaf(yf, s.sy, s.oy);
    const float pz = __builtin_fmaf(zf, s.sz, s.oz);
    // the edge's axis by selects (lanes of a wave hold edges of all three
    // axes): its low end p0, its high end p1, and p0 + t (p1 - p0) on it
    const float c = ax == 0 ? xf : (ax == 1 ? yf : zf);
    const float sc = ax == 0 ? s.sx : (ax == 1 ? s.sy : s.sz);
    const float oc = ax == 0 ? s.ox : (ax == 1 ? s.oy : s.oz);
    const float p0 = ax == 0 ? px : (ax == 1 ? py : pz);
    const float p1 = __builtin_fmaf(c + 1.0f, sc, oc);
    const float pe = __builtin_fmaf(t, p1 - p0, p0);
    return V3{ax == 0 ? pe : px, ax == 1 ? pe : py, ax == 2 ? pe : pz};
}

struct VertArgs {
    const float *field;     // field buffer base
    const uint64_t *signs;  // sign buffer base
    const uint64_t *bsum;   // scanned block bases
    const SegResult *res;
    float *vertices;
    uint64_t vcap;
};

// The vertex stage's LDS: one edge list per wave (a wave fills and emits its
// own 64 words' edges, with no wait on the block's other waves).
struct VertShared {
    uint16_t list[kBlock / 64][kWaveVertCap + 64];  // + one discard slot per lane
    uint64_t p0[kBlock];   // linear index of the word's first point
    uint32_t xyz[kBlock];  // first x | y << 16
    uint32_t z[kBlock];
    uint32_t wsum[4];
};

// Vertices of count block gb.
template <class Src>
__device__ __forceinline__ void vert_block(const Src &src, const VertArgs &a, uint32_t gb,
                                           VertShared &sh) {
    uint64_t *s_p0 = sh.p0;
    uint32_t *s_xyz = sh.xyz, *s_z = sh.z, *s_wsum = sh.wsum;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    {
        const uint32_t si = src.seg_of(gb);
        const Seg s = src.get(si);
        const Grid &g = s.g;
        const uint64_t *signs = a.signs + s.sign_off;
        const uint64_t k = (uint64_t)(gb - s.block_off) * kBlock + tid;
        // block base and grid range first: their loads overlap the sign loads
        const uint64_t vblock = a.bsum[2 * (uint64_t)gb];
        const uint32_t ovf = a.res[si].ovf;
        uint64_t ex = 0, ey = 0, ez = 0;
        if (k < g.words) {
            const WordPos p = word_pos(k, g);
            const bool has_y = p.y + 1 < g.ny, has_z = p.z + 1 < g.nz;
            const uint32_t kk = (uint32_t)k, W = (uint32_t)g.W, RW = (uint32_t)g.ny * W;
            // every load before any use (one round trip)
            const uint64_t w0 = signs[kk], w0n = ldw(signs, kk + 1, p.w + 1 < g.W);
            const uint64_t wy = ldw(signs, kk + W, has_y), wz = ldw(signs, kk + RW, has_z);
            const uint64_t vpt = mask_below((int64_t)g.nx - 64 * p.w);
            const uint64_t ved = mask_below((int64_t)g.nx - 1 - 64 * p.w);
            ex = (w0 ^ shift1(w0, w0n)) & ved;
            ey = has_y ? (w0 ^ wy) & vpt : 0ull;
            ez = has_z ? (w0 ^ wz) & vpt : 0ull;
            s_p0[tid] = ((uint64_t)p.z * g.ny + p.y) * g.nx + 64 * (uint64_t)p.w;
            s_xyz[tid] = (uint32_t)(64 * p.w) | ((uint32_t)p.y << 16);
            s_z[tid] = (uint32_t)p.z;
        }
        const uint32_t vc = popc64(ex) + popc64(ey) + popc64(ez);
        const uint32_t incl = wave_scan_incl(vc);
        const uint32_t wtot = readlane32(incl, 63);
        if (lane == 63) s_wsum[wave] = incl;
        __syncthreads();
        uint32_t wbase = 0;
        for (uint32_t i = 0; i < 4; ++i)
            if (i < wave) wbase += s_wsum[i];
        const uint32_t vexw = incl - vc;  // first vertex of this word within the wave
        if (ovf) return;  // vertex numbers do not fit 32 bits: nothing is written
        const float *field = a.field + s.field_off;
        const uint64_t plane = (uint64_t)g.nx * g.ny;
        uint16_t *list = sh.list[wave];
        const uint64_t vwave = vblock + wbase;  // the wave's first vertex
        for (uint32_t r0 = 0; r0 < wtot; r0 += kWaveVertCap) {
            // this lane's edges, canonical order, that fall into the round's window
            if (wtot <= kWaveVertCap) {
                // one round holds the wave: no window tests, 32-bit halves;
                // an edge a point lacks goes to the lane's discard slot
                uint16_t *dump = list + kWaveVertCap + lane;
                uint32_t i = vexw;
#pragma unroll
                for (uint32_t h = 0; h < 2; ++h) {
                    const uint32_t hx = (uint32_t)(ex >> (32 * h));
                    const uint32_t hy = (uint32_t)(ey >> (32 * h));
                    const uint32_t hz = (uint32_t)(ez >> (32 * h));
                    uint32_t pts = hx | hy | hz;
                    while (pts) {
                        const uint32_t b = (uint32_t)__builtin_ctz(pts);
                        pts &= pts - 1;
                        const uint32_t ent = (tid << 8) | ((b + 32 * h) << 2);
                        const uint32_t bx = (hx >> b) & 1u, by = (hy >> b) & 1u, bz = (hz >> b) & 1u;
                        *(bx ? list + i : dump) = (uint16_t)ent;
                        i += bx;
                        *(by ? list + i : dump) = (uint16_t)(ent | 1u);
                        i += by;
                        *(bz ? list + i : dump) = (uint16_t)(ent | 2u);
                        i += bz;
                    }
                }
            } else if (vc && vexw < r0 + kWaveVertCap && vexw + vc > r0) {
                uint64_t pts = ex | ey | ez;
                uint32_t i = vexw;
                while (pts) {
                    const uint32_t b = (uint32_t)__builtin_ctzll(pts);
                    pts &= pts - 1;
                    const uint32_t m =
                        bit_at(ex, b) | (bit_at(ey, b) << 1) | (bit_at(ez, b) << 2);
#pragma unroll
                    for (uint32_t ax = 0; ax < 3; ++ax) {
                        if (!((m >> ax) & 1)) continue;
                        if (i >= r0 && i < r0 + kWaveVertCap)
                            list[i - r0] = (uint16_t)((tid << 8) | (b << 2) | ax);
                        ++i;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            const uint32_t n = min((uint32_t)kWaveVertCap, wtot - r0);
            // kEdgeStep edges per lane and step, so that all their field
            // loads are in flight before the first interpolation
            for (uint32_t e0 = lane; e0 < n; e0 += kEdgeStep * 64) {
                uint32_t ent[kEdgeStep];
                float f0[kEdgeStep], f1[kEdgeStep];
#pragma unroll
                for (int q = 0; q < kEdgeStep; ++q) {
                    const uint32_t e = e0 + q * 64;
                    ent[q] = list[e < n ? e : e0];
                    const uint32_t j = ent[q] >> 8, b = (ent[q] >> 2) & 63, ax = ent[q] & 3;
                    const uint64_t pt = s_p0[j] + b;
                    const uint64_t st = ax == 0 ? 1ull : (ax == 1 ? (uint64_t)g.nx : plane);
                    f0[q] = field[pt];
                    f1[q] = field[pt + st];
                }
#pragma unroll
                for (int q = 0; q < kEdgeStep; ++q) {
                    const uint32_t e = e0 + q * 64;
                    const uint32_t j = ent[q] >> 8, b = (ent[q] >> 2) & 63, ax = ent[q] & 3;
                    const uint32_t xy = s_xyz[j];
                    const V3 v = edge_vertex(s, (float)((xy & 0xffff) + b), (float)(xy >> 16),
                                             (float)s_z[j], ax, f0[q], f1[q]);
                    const uint64_t id = vwave + r0 + e;
                    if (e < n && id < a.vcap) store_nt(a.vertices + 3 * id, v);
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// ---------------------------------------------------------------------------
// triangles
// ---------------------------------------------------------------------------
// One thread per half word: 32 points of a row, bit b of half h of word w is
// point 64w + 32h + b, and the sign buffer is read as 32-bit halves.  A
// triangle block is half of a count block: 128 words, 4 waves of 32 words,
// each wave one of the count stage's triangle groups.  32-bit halves halve
// the gathers and the bit work per cell and the per-thread state.
constexpr int kTriThreads = kBlock;
constexpr int kTriWords = kTriThreads / 2;  // words per triangle block
constexpr int kCellCap = 256;  // per-wave LDS cell-list entries per round

__device__ __forceinline__ uint32_t sh1(uint32_t cur, uint32_t next) {
    return (cur >> 1) | (next << 31);
}
// bits [0, n) of a 32-bit half
__device__ __forceinline__ uint32_t below32(int64_t n) {
    return n <= 0 ? 0u : (n >= 32 ? ~0u : (1u << n) - 1u);
}
// half `idx` when `ok`, else 0; unconditional load as ldw
__device__ __forceinline__ uint32_t ldh(const uint32_t *s, uint32_t idx, bool ok) {
    const uint32_t v = s[ok ? idx : 0u];
    return ok ? v : 0u;
}
__device__ __forceinline__ uint32_t bit32(uint32_t w, uint32_t b) { return (w >> b) & 1u; }

struct Cross32 {
    uint32_t ex, ey, ez, sx, sy;
};

// Crossing bits of the row half with signs (cur, nxt = next half), +y
// neighbour (py, pyn) and +z neighbour pz; sx/sy: the same for point b+1.
__device__ __forceinline__ Cross32 row_cross32(uint32_t cur, uint32_t nxt, uint32_t py,
                                               uint32_t pyn, uint32_t pz, bool has_y, bool has_z,
                                               uint32_t ved, uint32_t vpt, bool ved_n,
                                               bool vpt_n) {
    Cross32 c;
    c.ex = (cur ^ sh1(cur, nxt)) & ved;
    c.ey = has_y ? (cur ^ py) & vpt : 0u;
    c.ez = has_z ? (cur ^ pz) & vpt : 0u;
    const uint32_t exn = ved_n ? ((nxt ^ (nxt >> 1)) & 1u) : 0u;
    const uint32_t eyn = (has_y && vpt_n) ? ((nxt ^ pyn) & 1u) : 0u;
    c.sx = sh1(c.ex, exn);
    c.sy = sh1(c.ey, eyn);
    return c;
}

// Fills the calling wave's LDS list with its active cells of entries
// [r0, r0 + kCellCap) of the canonical order (lane | x << 6).
__device__ __forceinline__ void fill_cells(uint16_t *list, uint32_t act, uint32_t nact,
                                           uint32_t aex, uint32_t atot, uint32_t r0,
                                           uint32_t lane) {
    if (atot <= kCellCap) {
        uint32_t i = aex;
        uint32_t m = act;
        while (m) {
            const uint32_t x = (uint32_t)__builtin_ctz(m);
            m &= m - 1;
            list[i++] = (uint16_t)(lane | (x << 6));
        }
    } else if (nact && aex < r0 + kCellCap && aex + nact > r0) {
        uint32_t m = act;
        uint32_t i = aex;
        while (m) {
            const uint32_t x = (uint32_t)__builtin_ctz(m);
            m &= m - 1;
            if (i >= r0 && i < r0 + kCellCap) list[i - r0] = (uint16_t)(lane | (x << 6));
            ++i;
        }
    }
    __builtin_amdgcn_wave_barrier();
}

struct TriArgs {
    const uint64_t *signs;  // sign buffer base
    const uint32_t *base;   // block-local first vertex number of each half word
    const uint32_t *twave;  // triangles of each group of 32 words
    const uint64_t *bsum;
    const uint64_t *tri_packed;
    const SegResult *res;
    uint32_t *triangles;
    uint64_t tcap;
};

// The triangle stage's LDS.
struct TriShared {
    static constexpr int kWaves = kTriThreads / 64;
    uint64_t tri[256];
    uint16_t list[kWaves][kCellCap];
    uint32_t vid[kWaves][12 * 64];  // id k of lane l at k*64 + (l ^ k)
    // triangle -> lane | its three edges (4 bits each) << 8
    uint32_t own[kWaves][5 * 64];
};

// Triangles of triangle block tb (half tb & 1 of count block tb >> 1).
template <class Src>
__device__ __forceinline__ void tri_block(const Src &src, const TriArgs &a, uint32_t tb,
                                          TriShared &sh) {
    static_assert(kTriThreads == 256, "one case-table entry per thread");
    uint64_t *s_tri = sh.tri;
    // the case table's entry is loaded first and stored to LDS after the
    // sign loads below have returned: both share one round trip
    const uint64_t tri_entry = a.tri_packed[threadIdx.x];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, h = tid & 1;
    {
        const uint32_t gb = tb >> 1, part = tb & 1;  // count block and its half
        const uint32_t si = src.seg_of(gb);
        // the block's bookkeeping, loaded with no branch on it so that its
        // round trip is shared with the sign loads': the grid's overflow
        // flag, the block's first triangle and its 8 group counts
        const uint32_t ovf = a.res[si].ovf;
        const uint64_t v_off = a.res[si].v_off;
        const uint64_t tblock = a.bsum[2 * (uint64_t)gb + 1];
        const uint4 ga = *(const uint4 *)(a.twave + kGroups * (uint64_t)gb);
        const uint4 gc4 = *(const uint4 *)(a.twave + kGroups * (uint64_t)gb + 4);
        const Seg s = src.get(si);
        const Grid &g = s.g;
        const uint32_t *signs = (const uint32_t *)(a.signs + s.sign_off);
        const uint64_t k = (uint64_t)(gb - s.block_off) * kBlock + part * kTriWords + (tid >> 1);
        const uint32_t *base = a.base + (uint64_t)s.block_off * kBlock * 2;  // by grid half word
        const uint64_t *gbsum = a.bsum + 2 * (uint64_t)s.block_off;       // by grid block
        // ---- per half word (owner lane): the sign half of row (y, z) and the
        // crossing halves of the four rows (y..y+1, z..z+1) touching its
        // cells; a cell's cube index and edge vertex numbers follow from these
        // bits (sign of a neighbour corner = own sign ^ crossing of the edge
        // between)
        uint32_t wa = 0, act = 0;
        uint32_t ex0 = 0, ey0 = 0, ez0 = 0, ex1 = 0, ey1 = 0, ez1 = 0;
        uint32_t ex2 = 0, ey2 = 0, ez2 = 0, ex3 = 0, ey3 = 0, ez3 = 0;
        uint32_t carry = 0;  // crossing bits of point 32 (next half): x0 y0 x1 y1 x2
        uint32_t vb0 = 0, vb1 = 0, vb2 = 0, vb3 = 0;
        if (k < g.words) {
            const WordPos p = word_pos(k, g);
            const int xh = 64 * p.w + 32 * (int)h;  // first point of the half
            const uint32_t ved = below32((int64_t)g.nx - 1 - xh);
            if (p.y + 1 < g.ny && p.z + 1 < g.nz && ved) {
                const bool has_y2 = p.y + 2 < g.ny, has_z2 = p.z + 2 < g.nz;
                const uint32_t W2 = 2u * (uint32_t)g.W, RW2 = (uint32_t)g.ny * W2;
                const uint32_t i0 = 2u * (uint32_t)k + h;
                const bool hn = h == 0 || p.w + 1 < g.W;  // the row has a next half
                // every load of the half at once (one round trip); a half
                // without active cells discards them
                wa = signs[i0];
                const uint32_t na = ldh(signs, i0 + 1, hn);
                const uint32_t wb = signs[i0 + W2];
                const uint32_t nb = ldh(signs, i0 + W2 + 1, hn);
                const uint32_t wc = signs[i0 + RW2];
                const uint32_t nc = ldh(signs, i0 + RW2 + 1, hn);
                const uint32_t wd = signs[i0 + RW2 + W2];
                const uint32_t nd = ldh(signs, i0 + RW2 + W2 + 1, hn);
                const uint32_t we = ldh(signs, i0 + 2 * W2, has_y2),
                               ne = ldh(signs, i0 + 2 * W2 + 1, has_y2 && hn);
                const uint32_t wf = ldh(signs, i0 + 2 * W2 + RW2, has_y2);
                const uint32_t wg = ldh(signs, i0 + 2 * RW2, has_z2);
                const uint32_t wh = ldh(signs, i0 + 2 * RW2 + W2, has_z2);
                // the first vertex ids (block-local, plus the first vertex
                // of the block holding the row's word) are needed only by
                // halves with active cells, but are loaded (and kept)
                // unconditionally so that they share the sign loads' round trip
                static_assert(kBlock == 256, "a word's block is its index >> 8");
                const uint32_t kw = (uint32_t)k, Wd = (uint32_t)g.W, RWd = (uint32_t)g.ny * Wd;
                vb0 = base[i0] + (uint32_t)(gbsum[2 * (kw >> 8)] - v_off);
                vb1 = base[i0 + W2] + (uint32_t)(gbsum[2 * ((kw + Wd) >> 8)] - v_off);
                vb2 = base[i0 + RW2] + (uint32_t)(gbsum[2 * ((kw + RWd) >> 8)] - v_off);
                vb3 = base[i0 + RW2 + W2] + (uint32_t)(gbsum[2 * ((kw + RWd + Wd) >> 8)] - v_off);
                const uint32_t sa = sh1(wa, na), sb = sh1(wb, nb), sc = sh1(wc, nc),
                               sd = sh1(wd, nd);
                act = ((wa ^ sa) | (wa ^ wb) | (wa ^ sb) | (wa ^ wc) | (wa ^ sc) | (wa ^ wd) |
                       (wa ^ sd)) & ved;
                if (act) {
                    const uint32_t vpt = below32((int64_t)g.nx - xh);
                    const bool vpt_n = xh + 32 < g.nx, ved_n = xh + 32 < g.nx - 1;
                    const Cross32 c0 = row_cross32(wa, na, wb, nb, wc, true, true, ved, vpt, ved_n, vpt_n);
                    const Cross32 c1 = row_cross32(wb, nb, we, ne, wd, has_y2, true, ved, vpt, ved_n, vpt_n);
                    const Cross32 c2 = row_cross32(wc, nc, wd, nd, wg, true, has_z2, ved, vpt, ved_n, vpt_n);
                    const Cross32 c3 = row_cross32(wd, nd, wf, 0u, wh, has_y2, has_z2, ved, vpt, ved_n, vpt_n);
                    ex0 = c0.ex; ey0 = c0.ey; ez0 = c0.ez;
                    ex1 = c1.ex; ey1 = c1.ey; ez1 = c1.ez;
                    ex2 = c2.ex; ey2 = c2.ey; ez2 = c2.ez;
                    ex3 = c3.ex; ey3 = c3.ey; ez3 = c3.ez;
                    carry = (c0.sx >> 31) | ((c0.sy >> 31) << 1) | ((c1.sx >> 31) << 2) |
                            ((c1.sy >> 31) << 3) | ((c2.sx >> 31) << 4);
                }
            }
        }
        s_tri[tid] = tri_entry;
        __syncthreads();
        // vertex numbers that do not fit 32 bits: the grid's triangles are
        // not written
        const uint64_t tcap = ovf ? 0ull : a.tcap;
        // the wave's first triangle follows the part * 4 + wave groups before it
        static_assert(kGroups == 8, "two uint4 loads carry a block's group counts");
        const uint32_t ng = part * 4 + wave;
        const uint32_t gcount[kGroups] = {ga.x, ga.y, ga.z, ga.w, gc4.x, gc4.y, gc4.z, gc4.w};
        uint64_t tnext = tblock;
#pragma unroll
        for (uint32_t i = 0; i < kGroups; ++i) tnext += i < ng ? gcount[i] : 0u;
        const uint32_t nact = (uint32_t)__popc(act);
        const uint32_t aincl = wave_scan_incl(nact);
        const uint32_t atot = readlane32(aincl, 63);
        const uint32_t aex = aincl - nact;
        uint16_t *list = sh.list[wave];
        uint32_t *vid = sh.vid[wave];
        uint32_t *own = sh.own[wave];
        for (uint32_t r0 = 0; r0 < atot; r0 += kCellCap) {
            fill_cells(list, act, nact, aex, atot, r0, lane);
            const uint32_t n = min((uint32_t)kCellCap, atot - r0);
            for (uint32_t e0 = 0; e0 < n; e0 += 64) {
                const uint32_t e = e0 + lane;
                const bool on = e < n;
                const uint32_t ent = on ? list[e] : 0u;
                const uint32_t j = ent & 63, x = ent >> 6;
                // row by row: each row's edge ids go to LDS as soon as they
                // are known, so only one row's gathered halves are live
                const bool last = x == 31;
                const uint32_t xn = last ? 0u : x + 1;
                const uint32_t below = (1u << x) - 1u;
                const uint32_t cy = gather32(carry, j);
                const uint32_t a0 = bit32(gather32(wa, j), x);
                uint32_t x0, y0, z0, x1, z1, x2, x3;
                {  // row (y, z): x- and y-edges at x and x+1, z-edges at x and x+1
                    const uint32_t gx = gather32(ex0, j), gy = gather32(ey0, j), gz = gather32(ez0, j);
                    x0 = bit32(gx, x);
                    y0 = bit32(gy, x);
                    z0 = bit32(gz, x);
                    const uint32_t sx = last ? (cy & 1u) : bit32(gx, xn);
                    const uint32_t sy = last ? ((cy >> 1) & 1u) : bit32(gy, xn);
                    const uint32_t p = gather32(vb0, j) + __popc(gx & below) + __popc(gy & below) +
                                       __popc(gz & below);
                    const uint32_t q = p + x0 + y0 + z0;  // first edge of point x+1
                    vid[0 * 64 + (lane ^ 0)] = p;             // x-edge of (x,   y,   z)
                    vid[3 * 64 + (lane ^ 3)] = p + x0;        // y-edge of (x,   y,   z)
                    vid[8 * 64 + (lane ^ 8)] = p + x0 + y0;   // z-edge of (x,   y,   z)
                    vid[1 * 64 + (lane ^ 1)] = q + sx;        // y-edge of (x+1, y,   z)
                    vid[9 * 64 + (lane ^ 9)] = q + sx + sy;   // z-edge of (x+1, y,   z)
                }
                {  // row (y+1, z)
                    const uint32_t gx = gather32(ex1, j), gy = gather32(ey1, j), gz = gather32(ez1, j);
                    x1 = bit32(gx, x);
                    const uint32_t y1 = bit32(gy, x);
                    z1 = bit32(gz, x);
                    const uint32_t sx = last ? ((cy >> 2) & 1u) : bit32(gx, xn);
                    const uint32_t sy = last ? ((cy >> 3) & 1u) : bit32(gy, xn);
                    const uint32_t p = gather32(vb1, j) + __popc(gx & below) + __popc(gy & below) +
                                       __popc(gz & below);
                    vid[2 * 64 + (lane ^ 2)] = p;                         // x-edge of (x,   y+1, z)
                    vid[11 * 64 + (lane ^ 11)] = p + x1 + y1;             // z-edge of (x,   y+1, z)
                    vid[10 * 64 + (lane ^ 10)] = p + x1 + y1 + z1 + sx + sy;  // z-edge of (x+1, y+1, z)
                }
                {  // row (y, z+1)
                    const uint32_t gx = gather32(ex2, j), gy = gather32(ey2, j), gz = gather32(ez2, j);
                    x2 = bit32(gx, x);
                    const uint32_t y2 = bit32(gy, x), z2 = bit32(gz, x);
                    const uint32_t sx = last ? ((cy >> 4) & 1u) : bit32(gx, xn);
                    const uint32_t p = gather32(vb2, j) + __popc(gx & below) + __popc(gy & below) +
                                       __popc(gz & below);
                    vid[4 * 64 + (lane ^ 4)] = p;                    // x-edge of (x,   y,   z+1)
                    vid[7 * 64 + (lane ^ 7)] = p + x2;               // y-edge of (x,   y,   z+1)
                    vid[5 * 64 + (lane ^ 5)] = p + x2 + y2 + z2 + sx;  // y-edge of (x+1, y,   z+1)
                }
                {  // row (y+1, z+1)
                    const uint32_t gx = gather32(ex3, j), gy = gather32(ey3, j), gz = gather32(ez3, j);
                    x3 = bit32(gx, x);
                    vid[6 * 64 + (lane ^ 6)] = gather32(vb3, j) + __popc(gx & below) +
                                               __popc(gy & below) + __popc(gz & below);
                }
                // corners: 0 (x,y,z) 1 (x+1,y,z) 2 (x+1,y+1,z) 3 (x,y+1,z)
                //          4 (x,y,z+1) 5 (x+1,y,z+1) 6 (x+1,y+1,z+1) 7 (x,y+1,z+1)
                const uint32_t k0 = a0, k1 = a0 ^ x0, k3 = a0 ^ y0, k2 = k3 ^ x1;
                const uint32_t k4 = a0 ^ z0, k5 = k4 ^ x2, k7 = k3 ^ z1, k6 = k7 ^ x3;
                const uint32_t cube = k0 | (k1 << 1) | (k2 << 2) | (k3 << 3) | (k4 << 4) |
                                      (k5 << 5) | (k6 << 6) | (k7 << 7);
                const uint64_t packed = s_tri[cube];
                const uint32_t nt = on ? (uint32_t)(packed >> 60) : 0u;
                const uint32_t tincl = wave_scan_incl(nt);
                const uint32_t tex = tincl - nt, ttot = readlane32(tincl, 63);
                if (on) {
                    uint64_t rest = packed;
                    for (uint32_t jt = 0; jt < nt; ++jt, rest >>= 12)
                        own[tex + jt] = lane | (((uint32_t)rest & 0xfffu) << 8);
                }
                __builtin_amdgcn_wave_barrier();
                // the wave's triangles, one per lane: consecutive lanes store
                // consecutive 12-byte records
                for (uint32_t te = lane; te < ttot; te += 64) {
                    const uint32_t o = own[te], l = o & 63u;
                    const uint64_t t = tnext + te;
                    if (t < tcap) {
                        const uint32_t e0v = (o >> 8) & 15u;
                        const uint32_t e1v = (o >> 12) & 15u;
                        const uint32_t e2v = (o >> 16) & 15u;
                        const uint32_t v0 = vid[e0v * 64 + (l ^ e0v)];
                        const uint32_t v1 = vid[e1v * 64 + (l ^ e1v)];
                        const uint32_t v2 = vid[e2v * 64 + (l ^ e2v)];
                        store_nt(a.triangles + 3 * t, T3{v0, v1, v2});
                    }
                }
                tnext += ttot;
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
}

// The vertex and triangle stages in one launch: they depend only on the count
// stage and the block scan, so their workgroups are interleaved (workgroup w:
// w % 3 == 0 runs count block w / 3's vertices, the other two its two
// triangle halves) and run side by side on every CU; the stages wait on
// different things (field loads, LDS gathers), so each hides the other's
// latency.
union EmitShared {
    VertShared v;
    TriShared t;
};

template <class Src>
__global__ void __launch_bounds__(kBlock) emit_kernel(Src src, VertArgs va, TriArgs ta) {
    static_assert(kTriThreads == kBlock, "both stages run 256-thread workgroups");
    __shared__ EmitShared sh;
    const uint32_t w = blockIdx.x, gb = w / 3, r = w - 3 * gb;
    if (r == 0)
        vert_block(src, va, gb, sh.v);
    else
        tri_block(src, ta, 2 * gb + (r - 1), sh.t);
}

}  // namespace

// Enqueues the whole one-grid extraction on `stream` given a workspace.
int enqueue_extract(const gc_extract_args *a, hipStream_t stream) {
    GC_ARG(a, "args is NULL");
    GC_ARG(a->field, "field is NULL");
    GC_ARG(a->nx >= 2 && a->ny >= 2 && a->nz >= 2, "grid must have >= 2 points per axis");
    GC_ARG(a->ny < 65536 && a->nz < 65536, "ny and nz must be < 65536");
    GC_ARG(a->normals == nullptr, "per-vertex normals are not supported by this build");
    GC_ARG(a->vertices || a->vert_capacity == 0, "vertices is NULL");
    GC_ARG(a->triangles || a->tri_capacity == 0, "triangles is NULL");
    const ExtractLayout L = layout(a->nx, a->ny, a->nz);
    GC_ARG(a->workspace && a->workspace_bytes >= L.bytes,
           "workspace too small (%llu < %llu bytes)", (unsigned long long)a->workspace_bytes,
           (unsigned long long)L.bytes);
    GC_ARG(L.words < 0xffffffffull && a->nx <= 65536, "grid too large");
    const DeviceTables *tabs = nullptr;
    int st = device_tables(&tabs);
    if (st) return st;
    char *ws = (char *)a->workspace;
    OneGrid one;
    Seg &s = one.s;
    std::memset(&s, 0, sizeof(s));
    s.g = make_grid(a->nx, a->ny, a->nz);
    s.iso = a->iso;
    s.ox = a->origin.x;
    s.oy = a->origin.y;
    s.oz = a->origin.z;
    s.sx = a->spacing.x;
    s.sy = a->spacing.y;
    s.sz = a->spacing.z;
    const uint64_t *signs = a->signs;
    if (!signs) {
        GC_ARG((a->nz + kSignZ - 1) / kSignZ <= 65535 && (a->ny + 3) / 4 <= 65535,
               "grid too large");
        uint64_t *sw = (uint64_t *)(ws + L.signs_off);
        dim3 grid(s.g.W, (a->ny + 3) / 4, (a->nz + kSignZ - 1) / kSignZ);
        hipLaunchKernelGGL(signs_kernel, grid, dim3(256), 0, stream, a->field, sw, a->nx, a->ny,
                           a->nz, s.g.W, a->iso);
        GC_CHECK_LAUNCH();
        signs = sw;
    }
    uint32_t *base = (uint32_t *)(ws + L.base_off);
    uint32_t *twave = (uint32_t *)(ws + L.twave_off);
    uint64_t *bsum = (uint64_t *)(ws + L.bsum_off);
    uint64_t *ctl = (uint64_t *)(ws + L.ctl_off);
    SegResult *res = (SegResult *)(ws + L.res_off);
    hipLaunchKernelGGL(count_kernel<OneGrid>, dim3((unsigned)L.nb), dim3(kBlock), 0, stream, one,
                       CountArgs{signs, twave, bsum, base});
    GC_CHECK_LAUNCH();
    hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(1024), 0, stream, bsum, L.nb, ctl,
                       a->d_counts, (const Seg *)nullptr, 1u, res);
    GC_CHECK_LAUNCH();
    const VertArgs va = {a->field, signs, bsum, res, a->vertices,
                         a->vertices ? a->vert_capacity : 0};
    const TriArgs ta = {signs, base, twave, bsum, tabs->tri_packed, res, a->triangles,
                        a->triangles ? a->tri_capacity : 0};
    GC_ARG(L.nb <= 0xffffffffull / 3, "grid too large");
    if (a->emit_event) GC_HIP(hipEventRecord((hipEvent_t)a->emit_event, stream));
    hipLaunchKernelGGL(emit_kernel<OneGrid>, dim3(3 * (unsigned)L.nb), dim3(kBlock), 0, stream,
                       one, va, ta);
    GC_CHECK_LAUNCH();
    return GPUCAD_OK;
}

int enqueue_extract_batch(const BatchExtract &b, int phases, hipStream_t stream) {
    GC_ARG(b.nseg > 0 && b.nb > 0 && b.segs && b.blk2seg, "empty batch");
    GC_ARG(b.nb <= 0xffffffffu / 3, "batch too large");
    const DeviceTables *tabs = nullptr;
    int st = device_tables(&tabs);
    if (st) return st;
    const ManyGrids src = {b.segs, b.blk2seg};
    // one workgroup per block (two per block for the triangle stage), as for
    // one grid
    if (phases & EXTRACT_COUNT) {
        hipLaunchKernelGGL(count_kernel<ManyGrids>, dim3(b.nb), dim3(kBlock), 0, stream, src,
                           CountArgs{b.signs, b.twave, b.bsum, b.base});
        GC_CHECK_LAUNCH();
        hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(1024), 0, stream, b.bsum,
                           (uint64_t)b.nb, b.ctl, (uint64_t *)nullptr, b.segs, b.nseg, b.res);
        GC_CHECK_LAUNCH();
    }
    if (phases & EXTRACT_EMIT) {
        const VertArgs va = {b.fields, b.signs, b.bsum, b.res, b.vertices,
                             b.vertices ? b.vcap : 0};
        const TriArgs ta = {b.signs, b.base, b.twave, b.bsum, tabs->tri_packed, b.res,
                            b.triangles, b.triangles ? b.tcap : 0};
        hipLaunchKernelGGL(emit_kernel<ManyGrids>, dim3(3 * b.nb), dim3(kBlock), 0, stream, src,
                           va, ta);
        GC_CHECK_LAUNCH();
    }
    return GPUCAD_OK;
}

}  // namespace gc

extern "C" {

uint64_t gpucad_extract_workspace_bytes(int32_t nx, int32_t ny, int32_t nz) {
    if (nx < 2 || ny < 2 || nz < 2) return 0;
    return gc::layout(nx, ny, nz).bytes;
}

int gpucad_extract(const gc_extract_args *a, gc_mesh_counts *counts, void *stream) {
    gc::clear_error();
    hipStream_t s = gc::as_stream(stream);
    int st = gc::enqueue_extract(a, s);
    if (st) return st;
    if (!counts) return GPUCAD_OK;
    GC_HIP(hipStreamSynchronize(s));
    uint64_t ctl[3];
    const char *ws = (const char *)a->workspace;
    GC_HIP(hipMemcpy(ctl, ws + gc::layout(a->nx, a->ny, a->nz).ctl_off, sizeof(ctl),
                     hipMemcpyDeviceToHost));
    counts->n_vertices = ctl[gc::CTL_NV];
    counts->n_triangles = ctl[gc::CTL_NT];
    if (ctl[gc::CTL_OVF])
        return gc::fail(GPUCAD_E_OVERFLOW, "%llu vertices exceed 32-bit indices",
                        (unsigned long long)ctl[gc::CTL_NV]);
    if (ctl[gc::CTL_NV] > a->vert_capacity || ctl[gc::CTL_NT] > a->tri_capacity)
        return gc::fail(GPUCAD_E_OVERFLOW,
                        "mesh needs %llu vertices / %llu triangles, capacity %llu / %llu",
                        (unsigned long long)ctl[gc::CTL_NV], (unsigned long long)ctl[gc::CTL_NT],
                        (unsigned long long)a->vert_capacity, (unsigned long long)a->tri_capacity);
    return GPUCAD_OK;
}

}  // extern "C"
