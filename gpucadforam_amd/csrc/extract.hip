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
// Data: the grid is seen through its sign words (bit b of word w of row
// r = y + ny*z is [f(64w+b, y, z) < iso]).  Word k = r*W + w is the unit of
// work for bookkeeping: a lane owns one word and derives the crossing edges
// of its 64 points and the active cells of its 64 cells with 64-bit bit
// arithmetic.  Pipeline (one stream, no host round trip):
//   signs : ballot of f < iso per 64 points (skipped when the field kernel
//           already wrote them)
//   count : per word, crossing edges owned (vertices) and triangles of its
//           cells; per block of 256 words, their sums
//   scan  : one workgroup scans the block sums (64-bit) and writes totals
//   vert  : per block, exclusive scan of the word counts -> per-word vertex
//           and triangle bases (stored for `tri`); the crossing edges of the
//           block are compacted into an LDS list in canonical order, then one
//           lane per edge loads f0/f1, interpolates and stores: the block's
//           vertices are one contiguous range, written coalesced
//   tri   : per block, the active cells are compacted into an LDS list (word,
//           bit, case, triangle offset); one lane per active cell computes the
//           12 edge vertex ids from popcounts of the crossing words of the four
//           rows touching the cell and writes its triangles (contiguous range)
#include "gc_common.hpp"

namespace gc {
namespace {

constexpr int kBlockWords = 256;  // words per count / vert block
constexpr int kTriWords = 128;    // words per tri block
constexpr int kVertCap = 4096;    // LDS edge-list entries per round
constexpr int kTriCap = 2048;     // LDS active-cell entries per round

struct ExtractLayout {
    uint64_t signs_off, cnt_off, base_off, bsum_off, totals_off, bytes;
    uint64_t words, nb;
};

inline uint64_t align256(uint64_t v) { return (v + 255) & ~uint64_t(255); }

ExtractLayout layout(int nx, int ny, int nz) {
    ExtractLayout L;
    L.words = gpucad_sign_words(nx, ny, nz);
    L.nb = (L.words + kBlockWords - 1) / kBlockWords;
    uint64_t off = 0;
    L.signs_off = off;
    off = align256(off + L.words * 8);
    L.cnt_off = off;
    off = align256(off + L.words * 8);
    L.base_off = off;
    off = align256(off + L.words * 8);
    L.bsum_off = off;
    off = align256(off + L.nb * 16);
    L.totals_off = off;
    off = align256(off + 64);
    L.bytes = off;
    return L;
}

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

struct Grid {
    int nx, ny, nz, W;
    uint64_t words;
};

// word w of row (y, z); zero outside the grid
__device__ __forceinline__ uint64_t sword(const uint64_t *s, const Grid &g, int y, int z, int w) {
    return (y < g.ny && z < g.nz && w < g.W) ? s[((uint64_t)z * g.ny + y) * g.W + w] : 0ull;
}

struct WordPos {
    int y, z, w;
};

__device__ __forceinline__ WordPos word_pos(uint64_t k, const Grid &g) {
    uint64_t r = k / (uint64_t)g.W;
    WordPos p;
    p.w = (int)(k - r * g.W);
    p.y = (int)(r % (uint64_t)g.ny);
    p.z = (int)(r / (uint64_t)g.ny);
    return p;
}

// Crossing words of the row whose words are (cur, nxt) with +y neighbour
// words (py, pyn) and +z neighbour word pz: bit b of ex/ey/ez marks a crossing
// owned edge of point 64w+b; sx/sy hold the same for point 64w+b+1.
struct Cross {
    uint64_t ex, ey, ez, sx, sy;
};

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

// Cube index of cell bit b from the sign words of rows a=(y,z) b=(y+1,z)
// c=(y,z+1) d=(y+1,z+1) and their one-point shifts (corner order of
// src/MarchingCubes_kernel.cu:3449-3456).
__device__ __forceinline__ uint32_t cube_at(uint64_t a, uint64_t sa, uint64_t b, uint64_t sb,
                                            uint64_t c, uint64_t sc, uint64_t d, uint64_t sd,
                                            uint32_t x) {
    return bit_at(a, x) | (bit_at(sa, x) << 1) | (bit_at(sb, x) << 2) | (bit_at(b, x) << 3) |
           (bit_at(c, x) << 4) | (bit_at(sc, x) << 5) | (bit_at(sd, x) << 6) |
           (bit_at(d, x) << 7);
}

// Active cells of the word (a cell is active iff one of its 12 edges crosses).
__device__ __forceinline__ uint64_t active_cells(uint64_t a, uint64_t an, uint64_t b, uint64_t bn,
                                                 uint64_t c, uint64_t cn, uint64_t d, uint64_t dn,
                                                 uint64_t ved) {
    const uint64_t sa = shift1(a, an), sb = shift1(b, bn), sc = shift1(c, cn),
                   sd = shift1(d, dn);
    // a cell's corners all agree iff a == sa == b == sb == c == sc == d == sd at its bit
    const uint64_t diff = (a ^ sa) | (a ^ b) | (a ^ sb) | (a ^ c) | (a ^ sc) | (a ^ d) | (a ^ sd);
    return diff & ved;
}

// Block-wide exclusive scan of two u32 values over blockDim.x threads (<= 256).
__device__ __forceinline__ void block_scan2(uint32_t v, uint32_t t, uint32_t &vex, uint32_t &tex,
                                            uint32_t &vtot, uint32_t &ttot) {
    __shared__ uint32_t sv[256], st[256];
    const uint32_t tid = threadIdx.x;
    sv[tid] = v;
    st[tid] = t;
    __syncthreads();
    for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
        uint32_t av = tid >= off ? sv[tid - off] : 0u, at = tid >= off ? st[tid - off] : 0u;
        __syncthreads();
        sv[tid] += av;
        st[tid] += at;
        __syncthreads();
    }
    vex = sv[tid] - v;
    tex = st[tid] - t;
    vtot = sv[blockDim.x - 1];
    ttot = st[blockDim.x - 1];
    __syncthreads();
}

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
__global__ void __launch_bounds__(kBlockWords)
count_kernel(const uint64_t *__restrict__ signs, uint2 *__restrict__ cnt,
             uint64_t *__restrict__ bsum, const uint8_t *__restrict__ nverts, Grid g) {
    __shared__ uint8_t s_ntri[256];
    s_ntri[threadIdx.x] = nverts[threadIdx.x] / 3;
    __syncthreads();
    const uint64_t k = (uint64_t)blockIdx.x * kBlockWords + threadIdx.x;
    uint32_t vc = 0, tc = 0;
    if (k < g.words) {
        const WordPos p = word_pos(k, g);
        const bool has_y = p.y + 1 < g.ny, has_z = p.z + 1 < g.nz;
        const uint64_t a = sword(signs, g, p.y, p.z, p.w), an = sword(signs, g, p.y, p.z, p.w + 1);
        const uint64_t b = sword(signs, g, p.y + 1, p.z, p.w),
                       bn = sword(signs, g, p.y + 1, p.z, p.w + 1);
        const uint64_t c = sword(signs, g, p.y, p.z + 1, p.w),
                       cn = sword(signs, g, p.y, p.z + 1, p.w + 1);
        const uint64_t vpt = mask_below((int64_t)g.nx - 64 * p.w);
        const uint64_t ved = mask_below((int64_t)g.nx - 1 - 64 * p.w);
        const uint64_t ex = (a ^ shift1(a, an)) & ved;
        const uint64_t ey = has_y ? (a ^ b) & vpt : 0ull;
        const uint64_t ez = has_z ? (a ^ c) & vpt : 0ull;
        vc = popc64(ex) + popc64(ey) + popc64(ez);
        if (has_y && has_z) {
            const uint64_t d = sword(signs, g, p.y + 1, p.z + 1, p.w),
                           dn = sword(signs, g, p.y + 1, p.z + 1, p.w + 1);
            uint64_t act = active_cells(a, an, b, bn, c, cn, d, dn, ved);
            const uint64_t sa = shift1(a, an), sb = shift1(b, bn), sc = shift1(c, cn),
                           sd = shift1(d, dn);
            while (act) {
                const uint32_t x = (uint32_t)__builtin_ctzll(act);
                act &= act - 1;
                tc += s_ntri[cube_at(a, sa, b, sb, c, sc, d, sd, x)];
            }
        }
        cnt[k] = make_uint2(vc, tc);
    }
    uint32_t vex, tex, vtot, ttot;
    block_scan2(vc, tc, vex, tex, vtot, ttot);
    if (threadIdx.x == 0) {
        bsum[2 * blockIdx.x] = vtot;
        bsum[2 * blockIdx.x + 1] = ttot;
    }
}

// ---------------------------------------------------------------------------
// scan of the block sums (one workgroup); totals[0..1] = sums, totals[2] =
// 1 when a sum does not fit 32-bit ids
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(1024)
scan_blocks_kernel(uint64_t *bsum, uint64_t nb, uint64_t *totals, uint64_t *d_counts) {
    __shared__ uint64_t sv[1024], st[1024];
    const uint32_t t = threadIdx.x;
    const uint64_t per = (nb + 1023) / 1024;
    const uint64_t beg = t * per, end = min(nb, beg + per);
    uint64_t v = 0, tr = 0;
    for (uint64_t i = beg; i < end; ++i) {
        v += bsum[2 * i];
        tr += bsum[2 * i + 1];
    }
    sv[t] = v;
    st[t] = tr;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint64_t av = 0, at = 0;
        if (t >= off) {
            av = sv[t - off];
            at = st[t - off];
        }
        __syncthreads();
        sv[t] += av;
        st[t] += at;
        __syncthreads();
    }
    uint64_t rv = sv[t] - v, rt = st[t] - tr;
    for (uint64_t i = beg; i < end; ++i) {
        uint64_t cv = bsum[2 * i], ct = bsum[2 * i + 1];
        bsum[2 * i] = rv;
        bsum[2 * i + 1] = rt;
        rv += cv;
        rt += ct;
    }
    if (t == 1023) {
        totals[0] = sv[1023];
        totals[1] = st[1023];
        totals[2] = (sv[1023] > 0xffffffffull || st[1023] > 0xffffffffull) ? 1 : 0;
        if (d_counts) {
            d_counts[0] = sv[1023];
            d_counts[1] = st[1023];
        }
    }
}

// ---------------------------------------------------------------------------
// vertices (+ per-word bases)
// ---------------------------------------------------------------------------
struct VertArgs {
    const float *field;
    const uint64_t *signs;
    const uint2 *cnt;
    const uint64_t *bsum;  // scanned: block bases
    const uint64_t *totals;
    uint2 *base;
    float *vertices;
    uint64_t vcap;
    Grid g;
    float iso, ox, oy, oz, sx, sy, sz;
};

__global__ void __launch_bounds__(kBlockWords) vert_kernel(VertArgs a) {
    __shared__ uint16_t s_list[kVertCap];
    __shared__ uint64_t s_p0[kBlockWords];  // linear index of the word's first point
    __shared__ uint32_t s_yz[kBlockWords];  // y | z << 16 ... (y, z < 65536)
    const uint32_t tid = threadIdx.x;
    const uint64_t k = (uint64_t)blockIdx.x * kBlockWords + tid;
    const Grid &g = a.g;
    uint64_t ex = 0, ey = 0, ez = 0;
    uint32_t vc = 0, tc = 0;
    if (k < g.words) {
        const WordPos p = word_pos(k, g);
        const bool has_y = p.y + 1 < g.ny, has_z = p.z + 1 < g.nz;
        const uint64_t w0 = sword(a.signs, g, p.y, p.z, p.w),
                       w0n = sword(a.signs, g, p.y, p.z, p.w + 1);
        const uint64_t vpt = mask_below((int64_t)g.nx - 64 * p.w);
        const uint64_t ved = mask_below((int64_t)g.nx - 1 - 64 * p.w);
        ex = (w0 ^ shift1(w0, w0n)) & ved;
        ey = has_y ? (w0 ^ sword(a.signs, g, p.y + 1, p.z, p.w)) & vpt : 0ull;
        ez = has_z ? (w0 ^ sword(a.signs, g, p.y, p.z + 1, p.w)) & vpt : 0ull;
        vc = popc64(ex) + popc64(ey) + popc64(ez);
        tc = a.cnt[k].y;
        s_p0[tid] = ((uint64_t)p.z * g.ny + p.y) * g.nx + 64 * (uint64_t)p.w;
        s_yz[tid] = (uint32_t)p.y | ((uint32_t)p.z << 16);
    }
    uint32_t vex, tex, vtot, ttot;
    block_scan2(vc, tc, vex, tex, vtot, ttot);
    if (a.totals[2]) return;  // ids would not fit 32 bits: nothing is written
    const uint64_t vblock = a.bsum[2 * blockIdx.x], tblock = a.bsum[2 * blockIdx.x + 1];
    if (k < g.words) a.base[k] = make_uint2((uint32_t)(vblock + vex), (uint32_t)(tblock + tex));
    const uint64_t plane = (uint64_t)g.nx * g.ny;
    for (uint32_t r0 = 0; r0 < vtot; r0 += kVertCap) {
        // compact this lane's edges that fall into the round's window
        if (vc && vex < r0 + kVertCap && vex + vc > r0) {
            uint64_t pts = ex | ey | ez;
            uint32_t i = vex;
            while (pts) {
                const uint32_t b = (uint32_t)__builtin_ctzll(pts);
                pts &= pts - 1;
                const uint32_t m = bit_at(ex, b) | (bit_at(ey, b) << 1) | (bit_at(ez, b) << 2);
                for (uint32_t ax = 0; ax < 3; ++ax) {
                    if (!((m >> ax) & 1)) continue;
                    if (i >= r0 && i < r0 + kVertCap) s_list[i - r0] = (uint16_t)((tid << 8) | (b << 2) | ax);
                    ++i;
                }
            }
        }
        __syncthreads();
        const uint32_t n = min((uint32_t)kVertCap, vtot - r0);
        for (uint32_t e = tid; e < n; e += kBlockWords) {
            const uint32_t ent = s_list[e];
            const uint32_t j = ent >> 8, b = (ent >> 2) & 63, ax = ent & 3;
            const uint64_t pidx = s_p0[j] + b;
            const uint32_t yz = s_yz[j];
            const uint64_t kj = (uint64_t)blockIdx.x * kBlockWords + j;
            const float xf = (float)(64 * (int)(kj % (uint64_t)g.W) + (int)b);
            const float yf = (float)(yz & 0xffff), zf = (float)(yz >> 16);
            const float f0 = a.field[pidx];
            const uint64_t step = ax == 0 ? 1ull : (ax == 1 ? (uint64_t)g.nx : plane);
            const float f1 = a.field[pidx + step];
            const float t = (a.iso - f0) / (f1 - f0);
            float px = __builtin_fmaf(xf, a.sx, a.ox);
            float py = __builtin_fmaf(yf, a.sy, a.oy);
            float pz = __builtin_fmaf(zf, a.sz, a.oz);
            if (ax == 0) {
                const float p1 = __builtin_fmaf(xf + 1.0f, a.sx, a.ox);
                px = __builtin_fmaf(t, p1 - px, px);
            } else if (ax == 1) {
                const float p1 = __builtin_fmaf(yf + 1.0f, a.sy, a.oy);
                py = __builtin_fmaf(t, p1 - py, py);
            } else {
                const float p1 = __builtin_fmaf(zf + 1.0f, a.sz, a.oz);
                pz = __builtin_fmaf(t, p1 - pz, pz);
            }
            const uint64_t id = vblock + r0 + e;
            if (id < a.vcap) {
                float *v = a.vertices + 3 * id;
                v[0] = px;
                v[1] = py;
                v[2] = pz;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// triangles
// ---------------------------------------------------------------------------
struct TriArgs {
    const uint64_t *signs;
    const uint2 *base;
    const uint64_t *tri_packed;  // 256: 15 edge nibbles + count in nibble 15
    const uint64_t *totals;
    uint32_t *triangles;
    uint64_t tcap;
    Grid g;
};

// per-word data kept in LDS for the per-cell phase
struct WordData {
    uint64_t ex[4], ey[4], ez[4];  // crossing words of rows R0..R3
    uint64_t sx0, sy0, sx1, sy1, sx2;
    uint32_t vb[4];
    uint32_t tb;
    uint32_t pad;
};

__global__ void __launch_bounds__(kTriWords) tri_kernel(TriArgs a) {
    __shared__ uint64_t s_tri[256];
    __shared__ WordData s_w[kTriWords];
    __shared__ uint32_t s_list[kTriCap];
    __shared__ uint32_t s_vid[12][kTriWords];
    __shared__ uint32_t s_scan[kTriWords];
    __shared__ uint32_t s_total;
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < 256; i += kTriWords) s_tri[i] = a.tri_packed[i];
    if (a.totals[2]) return;  // uniform: ids would not fit 32 bits
    const Grid &g = a.g;
    const uint64_t k = (uint64_t)blockIdx.x * kTriWords + tid;
    uint64_t act = 0;
    uint64_t wa = 0, sa = 0, wb = 0, sb = 0, wc = 0, sc = 0, wd = 0, sd = 0;
    uint32_t nact = 0;
    if (k < g.words) {
        const WordPos p = word_pos(k, g);
        const bool has_y = p.y + 1 < g.ny, has_z = p.z + 1 < g.nz;
        if (has_y && has_z) {
            const bool has_y2 = p.y + 2 < g.ny, has_z2 = p.z + 2 < g.nz;
            const int y = p.y, z = p.z, w = p.w;
            wa = sword(a.signs, g, y, z, w);
            const uint64_t na = sword(a.signs, g, y, z, w + 1);
            wb = sword(a.signs, g, y + 1, z, w);
            const uint64_t nb = sword(a.signs, g, y + 1, z, w + 1);
            wc = sword(a.signs, g, y, z + 1, w);
            const uint64_t nc = sword(a.signs, g, y, z + 1, w + 1);
            wd = sword(a.signs, g, y + 1, z + 1, w);
            const uint64_t nd = sword(a.signs, g, y + 1, z + 1, w + 1);
            const uint64_t we = sword(a.signs, g, y + 2, z, w), ne = sword(a.signs, g, y + 2, z, w + 1);
            const uint64_t wf = sword(a.signs, g, y + 2, z + 1, w);
            const uint64_t wg = sword(a.signs, g, y, z + 2, w);
            const uint64_t wh = sword(a.signs, g, y + 1, z + 2, w);
            const uint64_t vpt = mask_below((int64_t)g.nx - 64 * w);
            const uint64_t ved = mask_below((int64_t)g.nx - 1 - 64 * w);
            const bool vpt_n = 64 * (w + 1) < g.nx, ved_n = 64 * (w + 1) < g.nx - 1;
            act = active_cells(wa, na, wb, nb, wc, nc, wd, nd, ved);
            if (act) {
                const Cross c0 = row_cross(wa, na, wb, nb, wc, true, true, ved, vpt, ved_n, vpt_n);
                const Cross c1 = row_cross(wb, nb, we, ne, wd, has_y2, true, ved, vpt, ved_n, vpt_n);
                const Cross c2 = row_cross(wc, nc, wd, nd, wg, true, has_z2, ved, vpt, ved_n, vpt_n);
                const Cross c3 = row_cross(wd, nd, wf, 0ull, wh, has_y2, has_z2, ved, vpt, ved_n, vpt_n);
                WordData &d = s_w[tid];
                d.ex[0] = c0.ex; d.ey[0] = c0.ey; d.ez[0] = c0.ez;
                d.ex[1] = c1.ex; d.ey[1] = c1.ey; d.ez[1] = c1.ez;
                d.ex[2] = c2.ex; d.ey[2] = c2.ey; d.ez[2] = c2.ez;
                d.ex[3] = c3.ex; d.ey[3] = c3.ey; d.ez[3] = c3.ez;
                d.sx0 = c0.sx; d.sy0 = c0.sy; d.sx1 = c1.sx; d.sy1 = c1.sy; d.sx2 = c2.sx;
                const uint64_t rW = (uint64_t)g.W, rowW = (uint64_t)g.ny * g.W;
                const uint2 b0 = a.base[k], b1 = a.base[k + rW], b2 = a.base[k + rowW],
                            b3 = a.base[k + rowW + rW];
                d.vb[0] = b0.x; d.vb[1] = b1.x; d.vb[2] = b2.x; d.vb[3] = b3.x;
                d.tb = b0.y;
                sa = shift1(wa, na); sb = shift1(wb, nb); sc = shift1(wc, nc); sd = shift1(wd, nd);
                nact = popc64(act);
            }
        }
    }
    // exclusive scan of the active-cell counts over the block
    s_scan[tid] = nact;
    __syncthreads();
    for (uint32_t off = 1; off < kTriWords; off <<= 1) {
        uint32_t v = tid >= off ? s_scan[tid - off] : 0u;
        __syncthreads();
        s_scan[tid] += v;
        __syncthreads();
    }
    const uint32_t aex = s_scan[tid] - nact;
    if (tid == kTriWords - 1) s_total = s_scan[tid];
    __syncthreads();
    const uint32_t total = s_total;
    for (uint32_t r0 = 0; r0 < total; r0 += kTriCap) {
        if (nact && aex < r0 + kTriCap && aex + nact > r0) {
            uint64_t m = act;
            uint32_t i = aex, toff = 0;
            while (m) {
                const uint32_t x = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                const uint32_t cube = cube_at(wa, sa, wb, sb, wc, sc, wd, sd, x);
                if (i >= r0 && i < r0 + kTriCap)
                    s_list[i - r0] = tid | (x << 7) | (cube << 13) | (toff << 21);
                toff += (uint32_t)(s_tri[cube] >> 60);
                ++i;
            }
        }
        __syncthreads();
        const uint32_t n = min((uint32_t)kTriCap, total - r0);
        for (uint32_t e = tid; e < n; e += kTriWords) {
            const uint32_t ent = s_list[e];
            const uint32_t j = ent & 127, x = (ent >> 7) & 63, cube = (ent >> 13) & 255,
                           toff = ent >> 21;
            const WordData &d = s_w[j];
            const uint64_t below = x ? (~0ull >> (64 - x)) : 0ull;
            uint32_t p[4], cx[4], cy[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                p[r] = d.vb[r] + popc64(d.ex[r] & below) + popc64(d.ey[r] & below) +
                       popc64(d.ez[r] & below);
                cx[r] = bit_at(d.ex[r], x);
                cy[r] = bit_at(d.ey[r], x);
            }
            const uint32_t cz0 = bit_at(d.ez[0], x), cz1 = bit_at(d.ez[1], x),
                           cz2 = bit_at(d.ez[2], x);
            const uint32_t q0 = p[0] + cx[0] + cy[0] + cz0;  // first edge of point x+1
            const uint32_t q1 = p[1] + cx[1] + cy[1] + cz1;
            const uint32_t q2 = p[2] + cx[2] + cy[2] + cz2;
            const uint32_t sx0 = bit_at(d.sx0, x), sy0 = bit_at(d.sy0, x);
            const uint32_t sx1 = bit_at(d.sx1, x), sy1 = bit_at(d.sy1, x);
            const uint32_t sx2 = bit_at(d.sx2, x);
            s_vid[0][tid] = p[0];                 // x-edge of (x,   y,   z)
            s_vid[1][tid] = q0 + sx0;             // y-edge of (x+1, y,   z)
            s_vid[2][tid] = p[1];                 // x-edge of (x,   y+1, z)
            s_vid[3][tid] = p[0] + cx[0];         // y-edge of (x,   y,   z)
            s_vid[4][tid] = p[2];                 // x-edge of (x,   y,   z+1)
            s_vid[5][tid] = q2 + sx2;             // y-edge of (x+1, y,   z+1)
            s_vid[6][tid] = p[3];                 // x-edge of (x,   y+1, z+1)
            s_vid[7][tid] = p[2] + cx[2];         // y-edge of (x,   y,   z+1)
            s_vid[8][tid] = p[0] + cx[0] + cy[0]; // z-edge of (x,   y,   z)
            s_vid[9][tid] = q0 + sx0 + sy0;       // z-edge of (x+1, y,   z)
            s_vid[10][tid] = q1 + sx1 + sy1;      // z-edge of (x+1, y+1, z)
            s_vid[11][tid] = p[1] + cx[1] + cy[1];// z-edge of (x,   y+1, z)
            const uint64_t packed = s_tri[cube];
            const uint32_t nt = (uint32_t)(packed >> 60);
            const uint64_t t0 = (uint64_t)d.tb + toff;
            for (uint32_t jt = 0; jt < nt; ++jt) {
                const uint64_t tid_out = t0 + jt;
                if (tid_out < a.tcap) {
                    uint32_t *tri = a.triangles + 3 * tid_out;
                    tri[0] = s_vid[(packed >> (12 * jt)) & 15][tid];
                    tri[1] = s_vid[(packed >> (12 * jt + 4)) & 15][tid];
                    tri[2] = s_vid[(packed >> (12 * jt + 8)) & 15][tid];
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace

// Enqueues the whole extraction on `stream` given a workspace (also used by
// the batched and served forms).
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
    GC_ARG(L.nb < 0x7fffffffull, "grid too large");
    const DeviceTables *tabs = nullptr;
    int st = device_tables(&tabs);
    if (st) return st;
    char *ws = (char *)a->workspace;
    Grid g;
    g.nx = a->nx;
    g.ny = a->ny;
    g.nz = a->nz;
    g.W = (int)gpucad_sign_words(a->nx, 1, 1);
    g.words = L.words;
    const uint64_t *signs = a->signs;
    if (!signs) {
        GC_ARG((a->nz + kSignZ - 1) / kSignZ <= 65535 && (a->ny + 3) / 4 <= 65535,
               "grid too large");
        uint64_t *s = (uint64_t *)(ws + L.signs_off);
        dim3 grid(g.W, (a->ny + 3) / 4, (a->nz + kSignZ - 1) / kSignZ);
        hipLaunchKernelGGL(signs_kernel, grid, dim3(256), 0, stream, a->field, s, a->nx, a->ny,
                           a->nz, g.W, a->iso);
        GC_CHECK_LAUNCH();
        signs = s;
    }
    uint2 *cnt = (uint2 *)(ws + L.cnt_off);
    uint2 *base = (uint2 *)(ws + L.base_off);
    uint64_t *bsum = (uint64_t *)(ws + L.bsum_off);
    uint64_t *totals = (uint64_t *)(ws + L.totals_off);

    hipLaunchKernelGGL(count_kernel, dim3((unsigned)L.nb), dim3(kBlockWords), 0, stream, signs,
                       cnt, bsum, tabs->num_verts, g);
    GC_CHECK_LAUNCH();
    hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(1024), 0, stream, bsum, L.nb, totals,
                       a->d_counts);
    GC_CHECK_LAUNCH();

    VertArgs va;
    va.field = a->field;
    va.signs = signs;
    va.cnt = cnt;
    va.bsum = bsum;
    va.totals = totals;
    va.base = base;
    va.vertices = a->vertices;
    va.vcap = a->vertices ? a->vert_capacity : 0;
    va.g = g;
    va.iso = a->iso;
    va.ox = a->origin.x;
    va.oy = a->origin.y;
    va.oz = a->origin.z;
    va.sx = a->spacing.x;
    va.sy = a->spacing.y;
    va.sz = a->spacing.z;
    hipLaunchKernelGGL(vert_kernel, dim3((unsigned)L.nb), dim3(kBlockWords), 0, stream, va);
    GC_CHECK_LAUNCH();

    TriArgs ta;
    ta.signs = signs;
    ta.base = base;
    ta.tri_packed = tabs->tri_packed;
    ta.totals = totals;
    ta.triangles = a->triangles;
    ta.tcap = a->triangles ? a->tri_capacity : 0;
    ta.g = g;
    const uint64_t ntb = (L.words + kTriWords - 1) / kTriWords;
    hipLaunchKernelGGL(tri_kernel, dim3((unsigned)ntb), dim3(kTriWords), 0, stream, ta);
    GC_CHECK_LAUNCH();
    return GPUCAD_OK;
}

uint64_t extract_totals_offset(int nx, int ny, int nz) { return layout(nx, ny, nz).totals_off; }

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
    uint64_t tot[3];
    const char *ws = (const char *)a->workspace;
    GC_HIP(hipMemcpy(tot, ws + gc::layout(a->nx, a->ny, a->nz).totals_off, sizeof(tot),
                     hipMemcpyDeviceToHost));
    counts->n_vertices = tot[0];
    counts->n_triangles = tot[1];
    if (tot[2])
        return gc::fail(GPUCAD_E_OVERFLOW, "%llu vertices / %llu triangles exceed 32-bit ids",
                        (unsigned long long)tot[0], (unsigned long long)tot[1]);
    if (tot[0] > a->vert_capacity || tot[1] > a->tri_capacity)
        return gc::fail(GPUCAD_E_OVERFLOW,
                        "mesh needs %llu vertices / %llu triangles, capacity %llu / %llu",
                        (unsigned long long)tot[0], (unsigned long long)tot[1],
                        (unsigned long long)a->vert_capacity, (unsigned long long)a->tri_capacity);
    return GPUCAD_OK;
}

}  // extern "C"
