// Drop-in boolean meshing: the reference's computeIsosurface path for
// primitives combined with each other or with a lattice band
// (src/Isosurface.cu:44-134, src/MarchingCubes_kernel.cu:158-463 classify_copy_Voxel,
// 869-1079 classifyVoxel, 1640-1784 vertexInterp_primitive/_primitive_one/_new,
// 1865-2218 generateTriangles_lattice_kernel), writing the caller's uint arrays
// and non-indexed float4 pos/norm exactly as the reference does.
//
// Fields: F = primitive_fixed (grid_points: val + per-point edge parameters
// t_x, t_y, t_z of its +x/+y/+z edges), D = primitive_dynamic (float),
// L = lattice field (float), band(L) = iso1 < L < iso2.  The corner
// occupancy of each mode (flags, GC_BOOL_*):
//   make_region            F < iso
//   union     fixed        D < iso  |  band(L)
//             dynamic      F < iso  |  band(L)
//             (neither)    F < iso  |  D < iso
//   diff      fixed        D >= iso &  band(L)
//             dynamic      F < iso  &  !(iso1 <= L <= iso2)   (L < iso1 | L > iso2)
//             (neither)    D >= iso &  F < iso
//   intersect fixed        D < iso  &  band(L)
//             dynamic      F < iso  &  band(L)
//             (neither)    F < iso  &  D < iso
// Vertices: make_region lerps with F's edge parameter; fixed uses
// vertexInterp_new on (D, L); dynamic vertexInterp_primitive_one on L with F's
// parameter; otherwise vertexInterp_primitive on D with F's parameter.  Face
// normal cross(v2 - v0, v1 - v0), w = 0.5 (src/MarchingCubes_kernel.cu:2178-2191).
//
// Decided differences (the reference leaves these undefined): with none of
// union/diff/intersect/make_region set the cube index is 0 (uninitialised in
// the reference); every cell is classified (no 2-D grid fold); a volume the
// mode does not read may be NULL.
#include "gc_common.hpp"
#include "wave_ops.hpp"

namespace gc {
namespace {

struct GridPoint {  // grid_points, src/MarchingCubes_kernel.h:12-18
    int val;
    float t_x, t_y, t_z;
};
static_assert(sizeof(GridPoint) == 16, "grid_points is 16 bytes");

struct BGrid {
    uint3 size, shift, mask;
    float3 vs, center;
};

__device__ __forceinline__ uint3 calc_grid_pos(uint32_t i, uint3 shift, uint3 mask) {
    // src/MarchingCubes_kernel.cu:120-136
    uint32_t zq = i / shift.z;
    uint32_t zr = i % shift.z;
    return make_uint3(zr % mask.x, zr / mask.x, zq);
}

__device__ __forceinline__ uint32_t sample_index(uint3 p, uint3 gs) {
    // src/MarchingCubes_kernel.cu:88-96, 99-118: clamped to gridSize
    p.x = min(p.x, gs.x);
    p.y = min(p.y, gs.y);
    p.z = min(p.z, gs.z);
    return p.z * gs.x * gs.y + p.y * gs.x + p.x;
}

__device__ __forceinline__ void corner_index(uint3 g, uint3 gs, uint32_t idx[8]) {
    idx[0] = sample_index(g, gs);
    idx[1] = sample_index(make_uint3(g.x + 1, g.y, g.z), gs);
    idx[2] = sample_index(make_uint3(g.x + 1, g.y + 1, g.z), gs);
    idx[3] = sample_index(make_uint3(g.x, g.y + 1, g.z), gs);
    idx[4] = sample_index(make_uint3(g.x, g.y, g.z + 1), gs);
    idx[5] = sample_index(make_uint3(g.x + 1, g.y, g.z + 1), gs);
    idx[6] = sample_index(make_uint3(g.x + 1, g.y + 1, g.z + 1), gs);
    idx[7] = sample_index(make_uint3(g.x, g.y + 1, g.z + 1), gs);
}

__device__ __forceinline__ bool in_band(float l, float lo, float hi) { return (l > lo) && (l < hi); }

// occupancy of one corner (the table in the header)
__device__ __forceinline__ uint32_t occupied(uint32_t flags, float fval, float d, float l,
                                             float iso, float lo, float hi) {
    const bool fixed = flags & GC_BOOL_FIXED, dynamic = flags & GC_BOOL_DYNAMIC;
    if (flags & GC_BOOL_MAKE_REGION) return fval < iso;
    if (flags & GC_BOOL_UNION) {
        if (fixed) return (d < iso) | in_band(l, lo, hi);
        if (dynamic) return (fval < iso) | in_band(l, lo, hi);
        return (fval < iso) | (d < iso);
    }
    if (flags & GC_BOOL_DIFF) {
        if (fixed) return (d >= iso) & in_band(l, lo, hi);
        if (dynamic) return (fval < iso) & ((l < lo) | (l > hi));
        return (d >= iso) & (fval < iso);
    }
    if (flags & GC_BOOL_INTERSECT) {
        if (fixed) return (d < iso) & in_band(l, lo, hi);
        if (dynamic) return (fval < iso) & in_band(l, lo, hi);
        return (fval < iso) & (d < iso);
    }
    return 0;
}

struct Vols {
    const GridPoint *fixed;
    const float *dyn, *lat;
};

// which volumes a mode reads (fixed takes precedence over dynamic, as in the
// reference's if / else-if chains)
// (returned by value: flags written through pointers were kept in private
// memory by the kernels)
struct Needs {
    bool f, d, l;
};
__host__ __device__ __forceinline__ Needs needs(uint32_t flags) {
    Needs n = {false, false, false};
    if (flags & GC_BOOL_MAKE_REGION) {
        n.f = true;
    } else if (flags & (GC_BOOL_UNION | GC_BOOL_DIFF | GC_BOOL_INTERSECT)) {
        if (flags & GC_BOOL_FIXED) {
            n.d = n.l = true;
        } else if (flags & GC_BOOL_DYNAMIC) {
            n.f = n.l = true;
        } else {
            n.f = n.d = true;
        }
    }
    return n;
}

__device__ __forceinline__ uint32_t cube_index(uint32_t flags, const Vols &v, const uint32_t idx[8],
                                               float iso, float lo, float hi, bool nf, bool nd,
                                               bool nl) {
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float fv = nf ? (float)v.fixed[idx[k]].val : 0.f;
        const float d = nd ? v.dyn[idx[k]] : 0.f;
        const float l = nl ? v.lat[idx[k]] : 0.f;
        c |= occupied(flags, fv, d, l, iso, lo, hi) << k;
    }
    return c;
}

__global__ void bool_classify_kernel(uint32_t *verts, uint32_t *occ, Vols v, BGrid g, uint32_t n,
                                     float iso, float lo, float hi, uint32_t flags,
                                     const uint8_t *nverts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Needs nn = needs(flags);
    const bool nf = nn.f, nd = nn.d, nl = nn.l;
    const uint3 p = calc_grid_pos(i, g.shift, g.mask);
    uint32_t idx[8];
    corner_index(p, g.size, idx);
    const uint32_t nv = nverts[cube_index(flags, v, idx, iso, lo, hi, nf, nd, nl)];
    verts[i] = nv;
    occ[i] = nv > 0;
}

// bit p of bits = the point's occupancy (the header's table), for the
// orchestration's classify from bits; each wave sets 64 x kWords
// consecutive bits (the layout dropin_scratch sizes)
__global__ void __launch_bounds__(256) bool_point_bits_kernel(uint64_t *bits, Vols v, uint32_t np,
                                                              float iso, float lo, float hi,
                                                              uint32_t flags) {
    constexpr int kWords = kDropinBitsWaveWords;
    const Needs nn = needs(flags);
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    uint64_t mine = 0;
#pragma unroll
    for (int j = 0; j < kWords; ++j) {
        const uint64_t i = wave * (64 * kWords) + j * 64 + lane;
        uint32_t b = 0;
        if (i < np) {
            const float fv = nn.f ? (float)v.fixed[i].val : 0.f;
            const float d = nn.d ? v.dyn[i] : 0.f;
            const float l = nn.l ? v.lat[i] : 0.f;
            b = occupied(flags, fv, d, l, iso, lo, hi);
        }
        const uint64_t m = __ballot(b);
        if (lane == (uint32_t)j) mine = m;
    }
    if (lane < kWords) bits[wave * kWords + lane] = mine;
}

// src/Isosurface.cu:120-121: bytes bytes of pos and norm cleared once some
// voxel is active (the count is on the device)
__global__ void clear_if_active_kernel(uint8_t *pos, uint8_t *norm, uint64_t bytes,
                                       const uint32_t *tot) {
    if (tot[0] == 0) return;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n16 = bytes / 16;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) {
        ((uint4 *)pos)[i] = make_uint4(0, 0, 0, 0);
        ((uint4 *)norm)[i] = make_uint4(0, 0, 0, 0);
    }
    if (blockIdx.x == 0 && threadIdx.x < bytes % 16) {
        pos[n16 * 16 + threadIdx.x] = 0;
        norm[n16 * 16 + threadIdx.x] = 0;
    }
}

// ---- classify_copy_Voxel: occupancy as val = -1 / 1 and averaged edge parameters
__device__ __forceinline__ bool crosses(float a, float b, float level) {
    // (a < level && b >= level) || (a >= level && b < level), a = neighbour, b = own
    return ((a < level) && (b >= level)) || ((a >= level) && (b < level));
}

__device__ __forceinline__ void blend(float &tp, float t) {
    if (tp > 0) tp = (tp + t) * 0.5f;
    else tp = t;
}

__global__ void bool_copy_kernel(GridPoint *vol_one, const float *vol_two, const float *vol_lat,
                                 uint3 gs, float iso, float lo, float hi, uint32_t flags,
                                 uint32_t n_points) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1 >= n_points) return;  // the reference stops before the last point
    const uint32_t plane = gs.x * gs.y;
    const uint3 p = make_uint3((i % plane) % gs.x, (i % plane) / gs.x, i / plane);
    const bool dynamic = flags & GC_BOOL_DYNAMIC;
    GridPoint g = vol_one[i];
    const float v = dynamic ? 0.f : vol_two[sample_index(p, gs)];
    const float vl = dynamic ? vol_lat[sample_index(p, gs)] : 0.f;
    const float fv = (float)g.val;
    bool inside;
    bool set = true;
    if (flags & GC_BOOL_UNION) {
        inside = dynamic ? (in_band(vl, lo, hi) | (fv < iso)) : ((v < iso) | (fv < iso));
    } else if (flags & GC_BOOL_DIFF) {
        inside = dynamic ? (in_band(vl, lo, hi) & (fv >= iso)) : ((v >= iso) & (fv < iso));
    } else if (flags & GC_BOOL_INTERSECT) {
        inside = dynamic ? (in_band(vl, lo, hi) & (fv < iso)) : ((v < iso) & (fv < iso));
    } else {
        inside = false;
        set = false;
    }
    if (set) g.val = inside ? -1 : 1;
    const float own = dynamic ? vl : v;
    const float *src = dynamic ? vol_lat : vol_two;
    for (int ax = 0; ax < 3; ++ax) {
        const uint32_t c = ax == 0 ? p.x : (ax == 1 ? p.y : p.z);
        const uint32_t lim = ax == 0 ? gs.x : (ax == 1 ? gs.y : gs.z);
        if (c >= lim - 1) continue;
        const uint3 q = make_uint3(p.x + (ax == 0), p.y + (ax == 1), p.z + (ax == 2));
        const float nb = src[sample_index(q, gs)];
        float &tp = ax == 0 ? g.t_x : (ax == 1 ? g.t_y : g.t_z);
        if (dynamic) {
            if (crosses(nb, own, lo)) {
                blend(tp, (lo - own) / (nb - own));
            } else if (crosses(nb, own, hi)) {
                blend(tp, (hi - own) / (nb - own));
            }
        } else if (crosses(nb, own, iso)) {
            blend(tp, (iso - own) / (nb - own));
        }
    }
    vol_one[i] = g;
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

__device__ __forceinline__ bool straddles(float f0, float f1, float level) {
    return ((f1 >= level) && (f0 <= level)) || ((f0 >= level) && (f1 <= level));
}

// merge of an edge parameter t1 with an interpolated t2 (vertexInterp_primitive)
__device__ __forceinline__ float merge_t(float t1, float t2, float t) {
    if ((t1 > 0) && (t2 > 0)) return (t1 + t2) * 0.5f;
    if ((t1 > 0) && (t2 == 0)) return t1;
    if ((t2 > 0) && (t1 == 0)) return t2;
    return t;
}

// The interpolations below compute each quotient unconditionally and
// select it where the reference's branch would take it (the same values,
// without divergent branches).

// src/MarchingCubes_kernel.cu:1640-1671
__device__ __forceinline__ float3 interp_primitive(float iso, float3 p0, float3 p1, float f0,
                                                   float f1, float te) {
    const float q = (iso - f0) / (f1 - f0);
    const float t2 = straddles(f0, f1, iso) ? q : 0.0f;
    return lerp3(p0, p1, merge_t(te, t2, 0.0f));
}

// src/MarchingCubes_kernel.cu:1674-1726
__device__ __forceinline__ float3 interp_primitive_one(float lo, float hi, float3 p0, float3 p1,
                                                       float f0, float f1, float te) {
    const float d = f1 - f0;
    const float ql = (lo - f0) / d, qh = (hi - f0) / d;
    const float t2 = straddles(f0, f1, lo) ? ql : 0.0f;
    const float t3 = straddles(f0, f1, hi) ? qh : 0.0f;
    return lerp3(p0, p1, merge_t(te, t3, merge_t(te, t2, 0.0f)));
}

// src/MarchingCubes_kernel.cu:1729-1784
__device__ __forceinline__ float3 interp_new(float iso, float lo, float hi, float3 p0, float3 p1,
                                             float f0, float f1, float f2, float f3) {
    const float q1 = (iso - f0) / (f1 - f0);
    const float q2 = (lo - f2) / (f3 - f2), q3 = (hi - f2) / (f3 - f2);
    const float t1 = (((f0 < iso) && (f1 >= iso)) || ((f1 < iso) && (f0 >= iso))) ? q1 : 0.0f;
    const float t2 = (((f2 < lo) && (f3 >= lo)) || ((f3 < lo) && (f2 >= lo))) ? q2 : 0.0f;
    const float t3 = (((f2 < hi) && (f3 >= hi)) || ((f3 < hi) && (f2 >= hi))) ? q3 : 0.0f;
    float t = 0.0f;
    if ((t1 > 0.0f) && (t2 > 0.0f) && (t3 == 0.0f)) t = (t1 + t2) * 0.5f;
    else if ((t1 > 0.0f) && (t3 > 0.0f) && (t2 == 0.0f)) t = (t1 + t3) * 0.5f;
    else if ((t1 > 0.0f) && (t2 == 0.0f) && (t3 == 0.0f)) t = t1;
    else if ((t2 > 0.0f) && (t1 == 0.0f) && (t3 == 0.0f)) t = t2;
    else if ((t3 > 0.0f) && (t1 == 0.0f) && (t2 == 0.0f)) t = t3;
    return lerp3(p0, p1, t);
}

// edges as the primitive modes read them: from corner a to corner b with the
// edge parameter of corner c along axis ax (edge 2: corner 3 -> 2, edge 6: 7 -> 6)
// (the loops over edges are unrolled, so corner arrays stay in registers)
constexpr int c_pa[12] = {0, 1, 3, 0, 4, 5, 7, 4, 0, 1, 2, 3};
constexpr int c_pb[12] = {1, 2, 2, 3, 5, 6, 6, 7, 4, 5, 6, 7};
constexpr int c_pax[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
// edges of vertexInterp_new (fixed mode): corner pairs in the cube's edge order
constexpr int c_na[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
constexpr int c_nb[12] = {1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7};

struct BGenArgs {
    float4 *pos, *norm;
    const uint32_t *comp, *scan;
    Vols v;
    BGrid g;
    float iso, lo, hi;
    uint32_t active, max_verts, flags;
    const uint8_t *tri;
    const uint8_t *nverts;
    const uint32_t *active_dev = nullptr;  // when set, the active voxel count is read here
    const uint32_t *vbase = nullptr;       // when set, scan[comp[i]] is read as vbase[i]
    const uint32_t *bits = nullptr;        // when set, cube indices come from the point bits
    // when set (and some voxel is active): scan[0 .. *zfill) are written with
    // 0 first (the leading verts_scan entries the prefix path leaves)
    const uint32_t *zfill = nullptr;
};

__device__ __forceinline__ float tparam(const GridPoint &p, int ax) {
    return ax == 0 ? p.t_x : (ax == 1 ? p.t_y : p.t_z);
}

typedef float F4nt __attribute__((ext_vector_type(4)));

// One lane per active voxel computes its 12 edge vertices into LDS as the
// reference's thread does; the wave then writes its voxels' vertices one
// output vertex per lane (contiguous with the scan of the voxels' counts, so
// the stores are coalesced; non-temporal, nothing reads them back here),
// vertex j of a voxel at scan[voxel] + j, each triangle only if its first
// index is below maxVerts - 3.  Batches of 64 active voxels as in the
// lattice generate: one workgroup per batch, or the orchestration's
// persistent launch with a contiguous eighth of the batches per XCD.
__global__ void __launch_bounds__(64) bool_generate_kernel(BGenArgs a) {
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
    const Needs nn = needs(a.flags);
    const bool nf = nn.f, nd = nn.d, nl = nn.l;
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
            if (a.bits) {
                cube = cube_from_bits(a.bits, lin, gx, gxy);
            } else {
                uint32_t idx[8];
                corner_index(gp, a.g.size, idx);
                cube = cube_index(a.flags, a.v, idx, a.iso, a.lo, a.hi, nf, nd, nl);
            }
            nv = a.nverts[cube];
            // the corner values each mode's interpolation reads (corner order of
            // the reference; grid_points whole, for their edge parameters)
            const uint32_t f = a.flags;
            const bool fixed = f & GC_BOOL_FIXED, make = f & GC_BOOL_MAKE_REGION;
            float D[8], L[8];
            float4 G[8];
            const bool load = nv != 0;  // (a mode's volumes exist when it yields vertices)
            if (load && !make && (fixed || !(f & GC_BOOL_DYNAMIC)))
                corners_inside(a.v.dyn, lin, gx, gxy, D);
            if (load && !make && (fixed || (f & GC_BOOL_DYNAMIC)))
                corners_inside(a.v.lat, lin, gx, gxy, L);
            if (load && (make || !fixed)) {
                const uint32_t cl[8] = {lin, lin + 1, lin + gx + 1, lin + gx,
                                        lin + gxy, lin + gxy + 1, lin + gxy + gx + 1, lin + gxy + gx};
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (k != 6) G[k] = ((const float4 *)a.v.fixed)[cl[k]];  // (corner 6: no edge)
            }
            // edge parameter of corner c along axis ax (t_x, t_y, t_z)
            auto te = [&](int c, int ax) { return ax == 0 ? G[c].y : (ax == 1 ? G[c].z : G[c].w); };
            if (nv == 0) {
                // nothing to write (src/MarchingCubes_kernel.cu:2078)
            } else if (make) {
#pragma unroll
                for (int e = 0; e < 12; ++e)
                    vl[e][lane] = lerp3(v[c_pa[e]], v[c_pb[e]], te(c_pa[e], c_pax[e]));
            } else if (fixed) {
#pragma unroll
                for (int e = 0; e < 12; ++e) {
                    const int ca = c_na[e], cb = c_nb[e];
                    vl[e][lane] = interp_new(a.iso, a.lo, a.hi, v[ca], v[cb], D[ca], D[cb], L[ca],
                                             L[cb]);
                }
            } else if (f & GC_BOOL_DYNAMIC) {
#pragma unroll
                for (int e = 0; e < 12; ++e) {
                    const int ca = c_pa[e], cb = c_pb[e];
                    vl[e][lane] = interp_primitive_one(a.lo, a.hi, v[ca], v[cb], L[ca], L[cb],
                                                       te(ca, c_pax[e]));
                }
            } else {
#pragma unroll
                for (int e = 0; e < 12; ++e) {
                    const int ca = c_pa[e], cb = c_pb[e];
                    vl[e][lane] = interp_primitive(a.iso, v[ca], v[cb], D[ca], D[cb],
                                                   te(ca, c_pax[e]));
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
            // calcNormal(ver[0], ver[2], ver[1]) = cross(v2 - v0, v1 - v0)
            const float3 e0 = make_float3(a2.x - a0.x, a2.y - a0.y, a2.z - a0.z);
            const float3 e1 = make_float3(a1.x - a0.x, a1.y - a0.y, a1.z - a0.z);
            const float3 n = cross3(e0, e1);
            const uint32_t index = s_base[o] + t0;  // scan[voxel] + t0, uint arithmetic
            if (index < a.max_verts - 3u) {
                const uint32_t r = j - t0;
                const float3 q = r == 0 ? a0 : (r == 1 ? a1 : a2);
                __builtin_nontemporal_store(F4nt{q.x, q.y, q.z, 1.0f},
                                            (F4nt *)(a.pos + index + r));
                __builtin_nontemporal_store(F4nt{n.x, n.y, n.z, 0.5f},
                                            (F4nt *)(a.norm + index + r));
            }
        }
        __syncthreads();  // the next batch reuses the LDS
    }
}

BGrid bgrid(gc_uint3 s, gc_uint3 sh, gc_uint3 m, gc_float3 vs, gc_float3 c) {
    BGrid g;
    g.size = make_uint3(s.x, s.y, s.z);
    g.shift = make_uint3(sh.x, sh.y, sh.z);
    g.mask = make_uint3(m.x, m.y, m.z);
    g.vs = make_float3(vs.x, vs.y, vs.z);
    g.center = make_float3(c.x, c.y, c.z);
    return g;
}

int check_vols(uint32_t flags, const void *fixed, const float *dyn, const float *lat) {
    const Needs nn = needs(flags);
    const bool nf = nn.f, nd = nn.d, nl = nn.l;
    GC_ARG(!nf || fixed, "primitive_fixed is NULL but the mode reads it");
    GC_ARG(!nd || dyn, "primitive_dynamic is NULL but the mode reads it");
    GC_ARG(!nl || lat, "lattice_field is NULL but the mode reads it");
    return GPUCAD_OK;
}

int bool_classify(uint32_t *verts, uint32_t *occ, const Vols &v, gc_uint3 s, gc_uint3 sh,
                  gc_uint3 m, uint32_t n, float iso, float lo, float hi, uint32_t flags,
                  hipStream_t st) {
    GC_ARG(verts && occ, "NULL buffer");
    int r = check_dropin_grid(s, sh, m, n);
    if (r) return r;
    if ((r = check_vols(flags, v.fixed, v.dyn, v.lat))) return r;
    const DeviceTables *t = nullptr;
    if ((r = device_tables(&t))) return r;
    if (n == 0) return GPUCAD_OK;
    const BGrid g = bgrid(s, sh, m, gc_float3{1, 1, 1}, gc_float3{0, 0, 0});
    hipLaunchKernelGGL(bool_classify_kernel, dim3((n + 255) / 256), dim3(256), 0, st, verts, occ, v,
                       g, n, iso, lo, hi, flags, t->num_verts);
    GC_CHECK_LAUNCH();
    return GPUCAD_OK;
}

int bool_generate(BGenArgs &a, hipStream_t st) {
    GC_ARG(a.pos && a.norm && a.comp && a.scan, "NULL buffer");
    int r = check_vols(a.flags, a.v.fixed, a.v.dyn, a.v.lat);
    if (r) return r;
    const DeviceTables *t = nullptr;
    if ((r = device_tables(&t))) return r;
    a.tri = t->tri_table;
    a.nverts = t->num_verts;
    if (a.active == 0) return GPUCAD_OK;
    hipLaunchKernelGGL(bool_generate_kernel, dim3((a.active + 63) / 64), dim3(64), 0, st, a);
    GC_CHECK_LAUNCH();
    return GPUCAD_OK;
}

}  // namespace
}  // namespace gc

using namespace gc;

extern "C" {

int gpucad_classify_voxel_boolean(uint32_t *voxel_verts, uint32_t *voxel_occupied,
                                  const void *primitive_fixed, const float *primitive_dynamic,
                                  const float *lattice_field, gc_uint3 grid_size,
                                  gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                  uint32_t num_voxels, float iso1, float iso2, float iso,
                                  uint32_t flags) {
    clear_error();
    const Vols v = {(const GridPoint *)primitive_fixed, primitive_dynamic, lattice_field};
    int r = bool_classify(voxel_verts, voxel_occupied, v, grid_size, grid_size_shift,
                          grid_size_mask, num_voxels, iso, iso1, iso2, flags, 0);
    if (r) return r;
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_classify_copy_voxel(void *vol_one, const float *volume_two, const float *vol_lattice,
                               gc_uint3 grid_size, float iso1, float iso2, float iso,
                               uint32_t flags) {
    clear_error();
    GC_ARG(vol_one, "vol_one is NULL");
    const bool dynamic = flags & GC_BOOL_DYNAMIC;
    GC_ARG(dynamic ? vol_lattice != nullptr : volume_two != nullptr,
           "the volume this mode reads is NULL");
    const uint64_t n = (uint64_t)grid_size.x * grid_size.y * grid_size.z;
    GC_ARG(grid_size.x >= 1 && grid_size.y >= 1 && grid_size.z >= 1 && n <= 0xffffffffull,
           "bad grid size");
    if (n < 2) return GPUCAD_OK;
    hipLaunchKernelGGL(bool_copy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0,
                       (GridPoint *)vol_one, volume_two, vol_lattice,
                       make_uint3(grid_size.x, grid_size.y, grid_size.z), iso, iso1, iso2, flags,
                       (uint32_t)n);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_generate_triangles_boolean(gc_float4 *pos, gc_float4 *norm, const uint32_t *compacted,
                                      const uint32_t *num_verts_scanned, gc_uint3 grid_size,
                                      gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                      gc_float3 voxel_size, gc_float3 grid_center, float iso,
                                      uint32_t active_voxels, uint32_t max_verts,
                                      const void *primitive_fixed,
                                      const float *primitive_dynamic, const float *lattice_field,
                                      float iso1, float iso2, uint32_t flags) {
    clear_error();
    int r = check_dropin_grid(grid_size, grid_size_shift, grid_size_mask, 0);
    if (r) return r;
    BGenArgs a;
    a.pos = (float4 *)pos;
    a.norm = (float4 *)norm;
    a.comp = compacted;
    a.scan = num_verts_scanned;
    a.v = Vols{(const GridPoint *)primitive_fixed, primitive_dynamic, lattice_field};
    a.g = bgrid(grid_size, grid_size_shift, grid_size_mask, voxel_size, grid_center);
    a.iso = iso;
    a.lo = iso1;
    a.hi = iso2;
    a.active = active_voxels;
    a.max_verts = max_verts;
    a.flags = flags;
    if ((r = bool_generate(a, 0))) return r;
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_compute_isosurface(gc_float4 *pos, gc_float4 *norm, float iso, uint32_t num_voxels,
                              uint32_t *d_voxel_verts, uint32_t *d_voxel_verts_scan,
                              uint32_t *d_voxel_occupied, uint32_t *d_voxel_occupied_scan,
                              gc_uint3 grid_size, gc_uint3 grid_size_shift,
                              gc_uint3 grid_size_mask, gc_float3 voxel_size,
                              gc_float3 grid_center, uint32_t *active_voxels,
                              uint32_t *total_verts, uint32_t *d_comp_voxel_array,
                              uint32_t max_verts, const void *primitive_fixed,
                              const float *primitive_dynamic, const float *lattice_field,
                              float iso1, float iso2, uint32_t flags, uint64_t *nfacets) {
    clear_error();
    auto calls = lock_device_calls();  // shared scratch arena and host words
    GC_ARG(active_voxels && total_verts && nfacets, "host outputs are NULL");
    GC_ARG(num_voxels > 0, "numVoxels must be > 0");
    GC_ARG(pos && norm && d_comp_voxel_array && d_voxel_verts_scan && d_voxel_occupied_scan,
           "NULL buffer");
    const Vols v = {(const GridPoint *)primitive_fixed, primitive_dynamic, lattice_field};
    const uint32_t n = num_voxels;
    GC_ARG(d_voxel_verts && d_voxel_occupied, "NULL buffer");
    int r = check_dropin_grid(grid_size, grid_size_shift, grid_size_mask, n);
    if (r) return r;
    if ((r = check_vols(flags, v.fixed, v.dyn, v.lat))) return r;
    const DeviceTables *t = nullptr;
    uint32_t *hw = nullptr;
    if ((r = device_tables(&t)) || (r = host_words(&hw))) return r;
    // classify (src/Isosurface.cu:63-67) from the point occupancy bits, both
    // scans and the compaction (69-116); the totals stay on the device, so
    // the call synchronises once
    DropinScratch w = {};
    if ((r = dropin_scratch(grid_size, n, &w))) return r;
    const uint32_t np = grid_size.x * grid_size.y * grid_size.z;
    hipLaunchKernelGGL(bool_point_bits_kernel, dim3(w.bits_blocks), dim3(256), 0, 0, w.bits, v, np,
                       iso, iso1, iso2, flags);
    GC_CHECK_LAUNCH();
    if ((r = dropin_classify_scans(w, d_voxel_verts, d_voxel_occupied, d_voxel_occupied_scan,
                                   d_voxel_verts_scan, d_comp_voxel_array, grid_size,
                                   grid_size_shift, grid_size_mask, n, hw, 0)))
        return r;
    // src/Isosurface.cu:120-121 clears max_verts bytes (not elements) of each,
    // when some voxel is active
    hipLaunchKernelGGL(clear_if_active_kernel, dim3(1024), dim3(256), 0, 0, (uint8_t *)pos,
                       (uint8_t *)norm, (uint64_t)max_verts, w.tot);
    GC_CHECK_LAUNCH();
    BGenArgs a;
    a.pos = (float4 *)pos;
    a.norm = (float4 *)norm;
    a.comp = d_comp_voxel_array;
    a.scan = d_voxel_verts_scan;
    a.v = v;
    a.g = bgrid(grid_size, grid_size_shift, grid_size_mask, voxel_size, grid_center);
    a.iso = iso;
    a.lo = iso1;
    a.hi = iso2;
    a.active = 0;
    a.max_verts = max_verts;
    a.flags = flags;
    a.tri = t->tri_table;
    a.nverts = t->num_verts;
    a.active_dev = w.tot;
    a.vbase = w.vbase;
    a.bits = (const uint32_t *)w.bits;
    if (dropin_prefix_path(w, false)) a.zfill = w.tot + 2;
    hipLaunchKernelGGL(bool_generate_kernel, dim3(dropin_generate_blocks(n)), dim3(64), 0, 0, a);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    *active_voxels = hw[0];
    if (*active_voxels == 0) {  // src/Isosurface.cu:83-87 (nfacets left as it was)
        *total_verts = 0;
        return GPUCAD_OK;
    }
    *total_verts = hw[1];
    *nfacets = *total_verts / 3;
    return GPUCAD_OK;
}

}  // extern "C"
