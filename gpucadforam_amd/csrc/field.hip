// Field evaluation kernels.
//
//  * lattice_field_kernel: the native spatially varying TPMS lattice (see
//    gc_lattice_params in gpucad.h for the definition), one 64 x 4 x lz brick
//    per workgroup (lattice_body.hpp): each lane owns one (x, y) column and
//    walks `lz` z planes; each z step of a wave stores 64 consecutive floats
//    (256 B) and, if requested, the ballot of G < iso as one 64-bit sign word.
//
//  * drop-in kernels for the reference's field entry points:
//    Fft_lattice::create_lattice (src/lattice_files/Fft_lattice.cu:12-74),
//    the Modelling primitives (src/Modelling.cu:244-761),
//    Gratings::GPU_buffer_normalise_four (src/lattice_files/Gratings.cu:1089-1134,
//    1394-1495, 1579-1617) and Gratings::primitive_field
//    (src/lattice_files/Gratings.cu:1695-1734).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "gc_common.hpp"
#include "gc_math.hpp"
#include "lattice_body.hpp"

namespace gc {
namespace {

// Rows per lane of the one-grid launch: a rotated lattice evaluates its z
// phase (and its cos) per lane and z, so four rows share it; without
// rotation the z terms come from the brick's table and one row per lane
// keeps small grids spread over more workgroups.
constexpr int rows_per_lane(bool rot) { return rot ? 4 : 1; }

template <bool ROT, int TPMS, bool THICK, int PRIM, int OP>
__global__ void __launch_bounds__(64 * kFieldWaves)
lattice_field_kernel(FieldConsts c, float *__restrict__ field, uint64_t *__restrict__ signs) {
    __shared__ float ztab[kMaxLZ * kZTab];
    constexpr int RY = rows_per_lane(ROT);
    lattice_brick<ROT, TPMS, THICK, 64, PRIM, OP, RY>(c, field, signs, blockIdx.x * 64,
                                                      blockIdx.y * kFieldWaves * RY,
                                                      blockIdx.z * c.lz, ztab);
}

typedef void (*FieldKernel)(FieldConsts, float *, uint64_t *);

// the primitive and the boolean are template parameters of the one-grid
// kernel: the per-point loop then carries no branch on them
template <bool ROT, int TPMS, bool THICK, int PRIM>
FieldKernel pick_op(int op) {
    switch (op) {
        case GC_OP_UNION: return lattice_field_kernel<ROT, TPMS, THICK, PRIM, GC_OP_UNION>;
        case GC_OP_SUBTRACT: return lattice_field_kernel<ROT, TPMS, THICK, PRIM, GC_OP_SUBTRACT>;
        case GC_OP_INTERSECT: return lattice_field_kernel<ROT, TPMS, THICK, PRIM, GC_OP_INTERSECT>;
        default: return lattice_field_kernel<ROT, TPMS, THICK, PRIM, GC_OP_NONE>;
    }
}

template <bool ROT, int TPMS, bool THICK>
FieldKernel pick_prim(int prim, int op) {
    switch (prim) {
        case GC_PRIM_SPHERE: return pick_op<ROT, TPMS, THICK, GC_PRIM_SPHERE>(op);
        case GC_PRIM_BOX: return pick_op<ROT, TPMS, THICK, GC_PRIM_BOX>(op);
        default: return lattice_field_kernel<ROT, TPMS, THICK, GC_PRIM_NONE, GC_OP_NONE>;
    }
}

template <bool ROT, int TPMS>
FieldKernel pick_thick(bool thick, int prim, int op) {
    return thick ? pick_prim<ROT, TPMS, true>(prim, op) : pick_prim<ROT, TPMS, false>(prim, op);
}

template <bool ROT>
FieldKernel pick_tpms(int tpms, bool thick, int prim, int op) {
    switch (tpms) {
        case GC_TPMS_GYROID: return pick_thick<ROT, GC_TPMS_GYROID>(thick, prim, op);
        case GC_TPMS_SCHWARZ_P: return pick_thick<ROT, GC_TPMS_SCHWARZ_P>(thick, prim, op);
        default: return pick_thick<ROT, GC_TPMS_DIAMOND>(thick, prim, op);
    }
}

}  // namespace

int validate_lattice(const gc_lattice_params *p) {
    GC_ARG(p, "params is NULL");
    GC_ARG(p->nx >= 2 && p->ny >= 2 && p->nz >= 2, "grid must have >= 2 points per axis");
    GC_ARG((uint64_t)p->nx * p->ny * p->nz < (1ull << 40), "grid too large");
    // the field kernels address a z plane with 32-bit byte offsets
    GC_ARG((uint64_t)p->nx * p->ny < (1ull << 30), "a z plane must hold fewer than 2^30 points");
    GC_ARG(p->tpms >= 0 && p->tpms <= 2, "unknown tpms %d", p->tpms);
    GC_ARG(p->period0 > 0.f && p->period1 > 0.f, "periods must be positive");
    GC_ARG(p->bend_radius >= 0.f, "bend radius must be >= 0");
    GC_ARG(!(p->bend_radius > 0.f && (p->rot0 != 0.f || p->rot1 != 0.f)),
           "rotation and bend are exclusive");
    GC_ARG(p->n_modes >= 0 && p->n_modes <= GC_MAX_MODES, "n_modes out of range");
    for (int m = 0; m < p->n_modes; ++m)
        for (int a = 0; a < 3; ++a)
            GC_ARG(p->mode_k[m][a] >= 0 && p->mode_k[m][a] <= 3, "mode wave numbers must be 0..3");
    GC_ARG(p->prim >= 0 && p->prim <= 2, "unknown primitive %d", p->prim);
    GC_ARG(p->op >= 0 && p->op <= 3, "unknown op %d", p->op);
    return GPUCAD_OK;
}

int launch_lattice_field(const gc_lattice_params *p, float *d_field, uint64_t *d_signs, float iso,
                         int lz, hipStream_t stream) {
    int st = validate_lattice(p);
    if (st) return st;
    GC_ARG(d_field, "field pointer is NULL");
    const bool rot = p->rot0 != 0.f || p->rot1 != 0.f;
    const int rows = rows_per_lane(rot);
    // default z run: long enough to amortise the per-column work (with four
    // rows per lane that work is four rows' thickness modes, so the run is
    // about a fifth of the grid, rounded up to 16 planes: 112 at 512^3,
    // measured best among 32..128), short enough to leave several workgroups
    // per z column
    if (lz <= 0)
        lz = rows > 1 ? std::max(16, std::min(kMaxLZ, ((p->nz + 4) / 5 + 15) / 16 * 16))
                      : std::max(16, std::min(kMaxLZ, p->nz / 8));
    if (lz > kMaxLZ) lz = kMaxLZ;
    const FieldConsts c = make_field_consts(p, iso, lz);
    const bool thick = p->n_modes > 0;
    FieldKernel k = rot ? pick_tpms<true>(p->tpms, thick, c.prim, c.op)
                        : pick_tpms<false>(p->tpms, thick, c.prim, c.op);
    const int brick_y = kFieldWaves * rows;
    dim3 grid((p->nx + 63) / 64, (p->ny + brick_y - 1) / brick_y, (p->nz + lz - 1) / lz);
    GC_ARG(grid.y <= 65535 && grid.z <= 65535, "grid too large for one launch");
    hipLaunchKernelGGL(k, grid, dim3(64 * kFieldWaves), 0, stream, c, d_field, d_signs);
    GC_CHECK_LAUNCH();
    return GPUCAD_OK;
}

// ===========================================================================
// Drop-in field kernels
// ===========================================================================
namespace {

// Euler rotation rows exactly as the reference forms them in each primitive
// kernel (src/Modelling.cu:401-407), evaluated once on the host.
struct Rows {
    float3 x, y, z;
};

Rows euler_rows(gc_float3 a) {
    float cx = cosf(a.x), sx = sinf(a.x);
    float cy = cosf(a.y), sy = sinf(a.y);
    float cz = cosf(a.z), sz = sinf(a.z);
    Rows r;
    r.x = make_float3(cz * cy, cz * sy * sx - sz * cx, cz * sy * cx + sz * sx);
    r.y = make_float3(sz * cy, sz * sy * sx + cz * cx, sz * sy * cx - cz * sx);
    r.z = make_float3(-1.0f * sy, cy * sx, cy * cx);
    return r;
}

struct Grid3 {
    int nx, ny, nz;
    float dx, dy, dz;
    float mx, my, mz;  // (N-1)/2.0
};

Grid3 make_grid(int nx, int ny, int nz, float dx, float dy, float dz) {
    Grid3 g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.dx = dx; g.dy = dy; g.dz = dz;
    g.mx = (float)((nx - 1) / 2.0);
    g.my = (float)((ny - 1) / 2.0);
    g.mz = (float)((nz - 1) / 2.0);
    return g;
}

__device__ __forceinline__ bool grid_point(const Grid3 &g, uint64_t &i, int &xx, int &yy, int &zz) {
    i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t n = (uint64_t)g.nx * g.ny * g.nz;
    if (i >= n) return false;
    uint64_t plane = (uint64_t)g.nx * g.ny;
    zz = (int)(i / plane);
    uint64_t rem = i - (uint64_t)zz * plane;
    yy = (int)(rem / g.nx);
    xx = (int)(rem - (uint64_t)yy * g.nx);
    return true;
}

__device__ __forceinline__ float3 centred(const Grid3 &g, int xx, int yy, int zz, float3 c) {
    return make_float3(((float)xx - g.mx) * g.dx - c.x, ((float)yy - g.my) * g.dy - c.y,
                       ((float)zz - g.mz) * g.dz - c.z);
}

__device__ __forceinline__ float dot_rows(float3 v, float3 r) {
    return v.x * r.x + v.y * r.y + v.z * r.z;
}

__global__ void sphere_kernel(Grid3 g, float *out, float3 c, float radius, float thick, int shell) {
    uint64_t i; int xx, yy, zz;
    if (!grid_point(g, i, xx, yy, zz)) return;
    float3 v = centred(g, xx, yy, zz, c);
    float t_diff = (float)(thick / 2.0);
    float r2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if (shell) {
        float ri = radius - t_diff, ro = radius + t_diff;
        float f1 = r2 - ri * ri;
        float f2 = r2 - ro * ro;
        out[i] = (float)fmax((double)f1 * -1.0, (double)f2);
    } else {
        out[i] = r2 - radius * radius;
    }
}

__global__ void cuboid_kernel(Grid3 g, float *out, float3 c, Rows R, float xw, float yw, float zw) {
    uint64_t i; int xx, yy, zz;
    if (!grid_point(g, i, xx, yy, zz)) return;
    float3 v = centred(g, xx, yy, zz, c);
    float f1 = fabsf(dot_rows(v, R.x)) - xw;
    float f2 = fabsf(dot_rows(v, R.y)) - yw;
    float f3 = fabsf(dot_rows(v, R.z)) - zw;
    out[i] = fmaxf(fmaxf(f1, f2), f3);
}

__global__ void cuboid_shell_kernel(Grid3 g, float *out, float3 c, Rows R, float xw, float yw,
                                    float zw, float thick) {
    uint64_t i; int xx, yy, zz;
    if (!grid_point(g, i, xx, yy, zz)) return;
    float3 v = centred(g, xx, yy, zz, c);
    float f1 = dot_rows(v, R.x);
    float f11 = fabsf(f1) - xw;
    float f12 = fabsf(f1) - (xw - thick);
    float f2 = dot_rows(v, R.y);
    float f21 = fabsf(f2) - yw;
    float f22 = fabsf(f2) - (yw - thick);
    float f3 = fabsf(dot_rows(v, R.z)) - zw;
    double inner = (double)fmaxf(f12, f22) * -1.0;
    out[i] = (float)fmax(fmax((double)fmaxf(f11, f21), inner), (double)f3);
}

__global__ void pyramid_kernel(Grid3 g, float *out, float3 c, Rows R, float xwb, float xwt,
                               float yh, float zwb, float zwt) {
    uint64_t i; int xx, yy, zz;
    if (!grid_point(g, i, xx, yy, zz)) return;
    float3 v = centred(g, xx, yy, zz, c);
    float f1 = dot_rows(v, R.x);
    float f2 = dot_rows(v, R.y);
    float f3 = dot_rows(v, R.z);
    float ratio = (yh - f2) / yh;
    float xw = ratio * (xwb - xwt) + xwt;
    f1 = fabsf(f1) - xw;
    f2 = fabsf(f2 - yh / 2) - yh / 2;
    float zw = ratio * (zwb - zwt) + zwt;
    f3 = fabsf(f3) - zw;
    out[i] = fmaxf(fmaxf(f1, f2), f3);
}

__global__ void torus_kernel(Grid3 g, float *out, float3 c, Rows R, float tr, float cr) {
    uint64_t i; int xx, yy, zz;
    if (!grid_point(g, i, xx, yy, zz)) return;
    float3 v = centred(g, xx, yy, zz, c);
    float vx = dot_rows(v, R.x), vy = dot_rows(v, R.y), vz = dot_rows(v, R.z);
    float side = tr - __builtin_sqrtf(vx * vx + vy * vy);
    out[i] = vz * vz - cr * cr + side * side;
}

__global__ void cone_kernel(Grid3 g, float *out, float3 c, Rows R, float br, float h) {
    uint64_t i; int xx, yy, zz;
    if (!grid_point(g, i, xx, yy, zz)) return;
    float3 v = centred(g, xx, yy, zz, c);
    float q = h / br;
    float f2 = q * q;
    float vx = dot_rows(v, R.x), vy = dot_rows(v, R.y), vz = dot_rows(v, R.z);
    float gg = (float)((double)vy - (double)h / 2.0);
    float hh = (float)fmax(((double)gg - (double)h / 2.0) * 100.0,
                           ((double)gg + (double)h / 2.0) * 100.0 * -1.0);
    float d = vy - h;
    float f1 = ((vx * vx + vz * vz) * f2) - d * d;
    out[i] = fmaxf(f1, hh);
}

__global__ void cone_frustum_kernel(Grid3 g, float *out, float3 c, Rows R, float tr, float br,
                                    float h) {
    uint64_t i; int xx, yy, zz;
    if (!grid_point(g, i, xx, yy, zz)) return;
    float3 v = centred(g, xx, yy, zz, c);
    float vx = dot_rows(v, R.x), vy = dot_rows(v, R.y), vz = dot_rows(v, R.z);
    float r_diff = ((h - vy) / h) * (br - tr);
    float gg = (float)((double)vy - (double)h / 2.0);
    float hh = (float)fmax(((double)gg - (double)h / 2.0) * 100.0,
                           ((double)gg + (double)h / 2.0) * 100.0 * -1.0);
    double rr = (double)(r_diff + tr);
    float f1 = (float)((double)(vx * vx + vz * vz) - rr * rr);
    out[i] = fmaxf(f1, hh);
}

__global__ void line_kernel(Grid3 g, float *out, float3 c, float3 axis, float radius, float t_diff,
                            float t_diff_ax, int cyl) {
    uint64_t i; int xx, yy, zz;
    if (!grid_point(g, i, xx, yy, zz)) return;
    float x1 = ((float)xx - g.mx) * g.dx;
    float y1 = ((float)yy - g.my) * g.dy;
    float z1 = ((float)zz - g.mz) * g.dz;
    float3 end = make_float3(axis.x + c.x, axis.y + c.y, axis.z + c.z);
    float3 w1 = make_float3(x1 - c.x, y1 - c.y, z1 - c.z);
    float3 w2 = make_float3(x1 - end.x, y1 - end.y, z1 - end.z);
    float3 w3 = make_float3(end.x - c.x, end.y - c.y, end.z - c.z);
    float3 d = make_float3(w1.y * w2.z - w1.z * w2.y, w1.z * w2.x - w1.x * w2.z,
                           w1.x * w2.y - w1.y * w2.x);
    float e = __builtin_sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    float dis = __builtin_sqrtf(w3.x * w3.x + w3.y * w3.y + w3.z * w3.z);
    float f = e / dis;
    float gg = (x1 - c.x) * axis.x + (y1 - c.y) * axis.y + (z1 - c.z) * axis.z;
    float f1 = fmaxf(gg - t_diff_ax, (gg + t_diff_ax) * -1.0f);
    float f2;
    if (cyl)
        f2 = fmaxf(f - (radius + t_diff), (f - (radius - t_diff)) * -1.0f);
    else
        f2 = f - radius;
    out[i] = fmaxf(f1, f2);
}

// Fft_lattice::create_lattice unit-cell forms (src/lattice_files/Fft_lattice.cu:12-66).
__global__ void unit_lattice_kernel(float *out, uint32_t nx, uint32_t ny, uint32_t nz,
                                    uint32_t type) {
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    uint32_t z = blockIdx.z * blockDim.z + threadIdx.z;
    if (x >= nx || y >= ny || z >= nz) return;
    float xx = (float)((((double)x * 1.0) / (double)(nx - 1) - 0.5) / 0.5);
    float yy = (float)((((double)y * 1.0) / (double)(ny - 1) - 0.5) / 0.5);
    float zz = (float)((((double)z * 1.0) / (double)(nz - 1) - 0.5) / 0.5);
    float ax = (float)(3.14 * (double)xx), ay = (float)(3.14 * (double)yy),
          az = (float)(3.14 * (double)zz);
    float aa = 0.f;
    if (type == 0) {
        aa = cosf(ax) * sinf(ay) + cosf(ay) * sinf(az) + cosf(az) * sinf(ax);
    } else if (type == 1) {
        aa = cosf(ax) + cosf(ay) + cosf(az);
    } else if (type == 2) {
        aa = 4 * (cosf(xx) * cosf(yy) * cosf(zz)) -
             (cosf(2 * xx) * cosf(2 * yy) + cosf(2 * yy) * cosf(2 * zz) + cosf(2 * zz) * cosf(2 * xx));
    } else if (type == 3) {
        float bx = (float)(2 * 3.14 * (double)xx), by = (float)(2 * 3.14 * (double)yy),
              bz = (float)(2 * 3.14 * (double)zz);
        aa = 2 * (cosf(ax) * cosf(ay) + cosf(ay) * cosf(az) + cosf(az) * cosf(ax)) -
             (cosf(bx) + cosf(by) + cosf(bz));
    } else if (type == 4) {
        double X2 = (double)xx * (double)xx, Y2 = (double)yy * (double)yy,
               Z2 = (double)zz * (double)zz;
        double a = (double)(xx * xx) + Y2;
        aa = (float)fmin(fmin(a, Y2 + Z2), Z2 + X2);
    } else if (type == 5) {
        double t1 = cos(3.14 * (double)xx) * (double)cosf(ay) * (double)cosf(az);
        aa = (float)(t1 - (double)(sinf(ax) * sinf(ay) * sinf(az)));
    }
    out[(uint64_t)x + (uint64_t)y * nx + (uint64_t)z * nx * ny] = aa;
}

// ---- normalise_four: min/max (seeded with 0) then band indicator
__global__ void minmax_partial_kernel(const float *in, uint64_t n, float2 *partial) {
    __shared__ float smin[256], smax[256];
    float lo = 0.f, hi = 0.f;  // the reference's second pass seeds with 0
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        float v = in[i];
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = make_float2(smin[0], smax[0]);
}

__global__ void minmax_final_kernel(float2 *partial, int nparts) {
    __shared__ float smin[256], smax[256];
    float lo = 0.f, hi = 0.f;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        lo = fminf(lo, partial[i].x);
        hi = fmaxf(hi, partial[i].y);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[0] = make_float2(smin[0], smax[0]);
}

__global__ void band_kernel(const float *in, float *indicator, float *normalised, const float2 *ab,
                            int nx, int ny, int nz, float iso1, float iso2) {
    uint64_t tx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t size = (uint64_t)nx * ny * nz;
    if (tx >= size) return;
    float a = ab->x, b = ab->y;
    int xx = (int)(tx % (uint64_t)nx);
    int yy = (int)((tx % ((uint64_t)nx * ny)) / (uint64_t)nx);
    int zz = (int)(tx / ((uint64_t)nx * ny));
    float k = in[tx];
    k = (k - a) / (b - a);
    float m;
    if (xx == 0 || xx == nx - 1 || yy == 0 || yy == ny - 1 || zz == 0 || zz == nz - 1) {
        m = 0.f;
        k = 0.f;
    } else {
        m = (k >= iso1 && k <= iso2) ? 1.f : 0.f;
    }
    indicator[tx] = m;
    normalised[tx] = k;
}

__global__ void primitive_dynamic_kernel(const float *active, float *iso, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && active[i] >= 0.f) iso[i] = FLT_MAX;
}

// grid_points {int val; float t_x, t_y, t_z;}: only val is read
__global__ void primitive_fixed_kernel(const int4 *fixed, float *iso, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && fixed[i].x > -1) iso[i] = FLT_MAX;
}

int point_launch(int nx, int ny, int nz, dim3 *grid) {
    GC_ARG(nx >= 1 && ny >= 1 && nz >= 1, "grid sizes must be positive");
    uint64_t n = (uint64_t)nx * ny * nz;
    uint64_t blocks = (n + 255) / 256;
    GC_ARG(blocks <= 0x7fffffffull, "grid too large");
    *grid = dim3((unsigned)blocks);
    return GPUCAD_OK;
}

}  // namespace
}  // namespace gc

using namespace gc;

namespace gc {
// min and max of in[0, n) seeded with 0 (the reference's reductions seed their
// second pass with 0) into part[0]; part holds 1024 float2
int enqueue_minmax0(const float *in, uint64_t n, float2 *part, hipStream_t st) {
    hipLaunchKernelGGL(minmax_partial_kernel, dim3(1024), dim3(256), 0, st, in, n, part);
    GC_CHECK_LAUNCH();
    hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(256), 0, st, part, 1024);
    GC_CHECK_LAUNCH();
    return GPUCAD_OK;
}
}  // namespace gc

extern "C" {

uint64_t gpucad_sign_words(int32_t nx, int32_t ny, int32_t nz) {
    if (nx <= 0 || ny <= 0 || nz <= 0) return 0;
    return (uint64_t)((nx + 63) / 64) * (uint64_t)ny * (uint64_t)nz;
}

int gpucad_lattice_field(const gc_lattice_params *p, float *d_field, uint64_t *d_signs, float iso,
                         void *stream) {
    clear_error();
    return launch_lattice_field(p, d_field, d_signs, iso, 0, as_stream(stream));
}

// Same launch with an explicit z-run length (autotuning hook).
int gpucad_lattice_field_lz(const gc_lattice_params *p, float *d_field, uint64_t *d_signs,
                            float iso, int32_t lz, void *stream) {
    clear_error();
    return launch_lattice_field(p, d_field, d_signs, iso, lz, as_stream(stream));
}

int gpucad_create_lattice(float *d, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t size,
                          uint32_t type) {
    clear_error();
    GC_ARG(d, "output is NULL");
    GC_ARG(nx >= 2 && ny >= 2 && nz >= 2, "grid must have >= 2 points per axis");
    GC_ARG(type <= 5, "lattice type must be 0..5");
    (void)size;
    dim3 tids(16, 8, 8);
    dim3 grids((nx + 15) / 16, (ny + 7) / 8, (nz + 7) / 8);
    hipLaunchKernelGGL(unit_lattice_kernel, grids, tids, 0, 0, d, nx, ny, nz, type);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

#define GC_PRIM_PROLOGUE(ptr)                                   \
    clear_error();                                              \
    GC_ARG(ptr, "output is NULL");                              \
    dim3 grid;                                                  \
    {                                                           \
        int st_ = point_launch(nx, ny, nz, &grid);              \
        if (st_) return st_;                                    \
    }                                                           \
    Grid3 g = make_grid(nx, ny, nz, dx, dy, dz);

#define GC_PRIM_EPILOGUE()            \
    GC_CHECK_LAUNCH();                \
    GC_HIP(hipStreamSynchronize(0)); \
    return GPUCAD_OK;

int gpucad_sphere_with_center(float *data, gc_float3 center, float radius, float thickness,
                              int32_t nx, int32_t ny, int32_t nz, float dx, float dy, float dz,
                              int32_t shell) {
    GC_PRIM_PROLOGUE(data)
    hipLaunchKernelGGL(sphere_kernel, grid, dim3(256), 0, 0, g, data,
                       make_float3(center.x, center.y, center.z), radius, thickness, shell);
    GC_PRIM_EPILOGUE()
}

int gpucad_cuboid(float *data, gc_float3 center, gc_float3 angles, float x_width, float y_width,
                  float z_width, int32_t nx, int32_t ny, int32_t nz, float dx, float dy,
                  float dz) {
    GC_PRIM_PROLOGUE(data)
    Rows R = euler_rows(angles);
    hipLaunchKernelGGL(cuboid_kernel, grid, dim3(256), 0, 0, g, data,
                       make_float3(center.x, center.y, center.z), R, (float)(x_width / 2.0),
                       (float)(y_width / 2.0), (float)(z_width / 2.0));
    GC_PRIM_EPILOGUE()
}

int gpucad_cuboid_shell(float *data, gc_float3 center, gc_float3 angles, float x_width,
                        float y_width, float z_width, float thickness, int32_t nx, int32_t ny,
                        int32_t nz, float dx, float dy, float dz) {
    GC_PRIM_PROLOGUE(data)
    Rows R = euler_rows(angles);
    hipLaunchKernelGGL(cuboid_shell_kernel, grid, dim3(256), 0, 0, g, data,
                       make_float3(center.x, center.y, center.z), R, (float)(x_width / 2.0),
                       (float)(y_width / 2.0), (float)(z_width / 2.0), thickness);
    GC_PRIM_EPILOGUE()
}

int gpucad_pyramid_frustum(float *data, gc_float3 center, gc_float3 angles, float x_width_base,
                           float x_width_top, float y_height, float z_width_base,
                           float z_width_top, int32_t nx, int32_t ny, int32_t nz, float dx,
                           float dy, float dz) {
    GC_PRIM_PROLOGUE(data)
    Rows R = euler_rows(angles);
    hipLaunchKernelGGL(pyramid_kernel, grid, dim3(256), 0, 0, g, data,
                       make_float3(center.x, center.y, center.z), R, (float)(x_width_base / 2.0),
                       (float)(x_width_top / 2.0), y_height, (float)(z_width_base / 2.0),
                       (float)(z_width_top / 2.0));
    GC_PRIM_EPILOGUE()
}

int gpucad_torus_with_center(float *data, gc_float3 center, gc_float3 angles, float torus_radius,
                             float circle_radius, int32_t nx, int32_t ny, int32_t nz, float dx,
                             float dy, float dz) {
    GC_PRIM_PROLOGUE(data)
    Rows R = euler_rows(angles);
    hipLaunchKernelGGL(torus_kernel, grid, dim3(256), 0, 0, g, data,
                       make_float3(center.x, center.y, center.z), R, torus_radius, circle_radius);
    GC_PRIM_EPILOGUE()
}

int gpucad_cone_with_base_radius_height(float *data, gc_float3 center, gc_float3 angles,
                                        float base_radius, float cone_height, int32_t nx,
                                        int32_t ny, int32_t nz, float dx, float dy, float dz) {
    GC_PRIM_PROLOGUE(data)
    Rows R = euler_rows(angles);
    hipLaunchKernelGGL(cone_kernel, grid, dim3(256), 0, 0, g, data,
                       make_float3(center.x, center.y, center.z), R, base_radius, cone_height);
    GC_PRIM_EPILOGUE()
}

int gpucad_cone_frustum(float *data, gc_float3 center, gc_float3 angles, float top_radius,
                        float bottom_radius, float height, int32_t nx, int32_t ny, int32_t nz,
                        float dx, float dy, float dz) {
    GC_PRIM_PROLOGUE(data)
    Rows R = euler_rows(angles);
    hipLaunchKernelGGL(cone_frustum_kernel, grid, dim3(256), 0, 0, g, data,
                       make_float3(center.x, center.y, center.z), R, top_radius, bottom_radius,
                       height);
    GC_PRIM_EPILOGUE()
}

int gpucad_distance_from_line(float *data, gc_float3 center, gc_float3 axis, float radius,
                              float thickness_radial, float thickness_axial, int32_t nx,
                              int32_t ny, int32_t nz, float dx, float dy, float dz,
                              int32_t cylinder_disc) {
    GC_PRIM_PROLOGUE(data)
    float mag = sqrtf(axis.x * axis.x + axis.y * axis.y + axis.z * axis.z);
    GC_ARG(mag > 0.f, "axis must be non-zero");
    float3 ax = make_float3(axis.x / mag, axis.y / mag, axis.z / mag);
    hipLaunchKernelGGL(line_kernel, grid, dim3(256), 0, 0, g, data,
                       make_float3(center.x, center.y, center.z), ax, radius,
                       (float)(thickness_radial / 2.0), (float)(thickness_axial / 2.0),
                       cylinder_disc);
    GC_PRIM_EPILOGUE()
}

int gpucad_normalise_four(const float *in, float *indicator, float *normalised, uint64_t size,
                          int32_t nx, int32_t ny, int32_t nz, float iso1, float iso2) {
    clear_error();
    auto calls = lock_device_calls();  // shared scratch arena
    GC_ARG(in && indicator && normalised, "NULL buffer");
    GC_ARG(nx >= 1 && ny >= 1 && nz >= 1, "grid sizes must be positive");
    GC_ARG(size == (uint64_t)nx * ny * nz, "size must equal nx*ny*nz");
    const int nparts = 1024;
    void *ws = nullptr;
    int st = scratch_arena(nparts * sizeof(float2), &ws);
    if (st) return st;
    float2 *part = (float2 *)ws;
    hipLaunchKernelGGL(minmax_partial_kernel, dim3(nparts), dim3(256), 0, 0, in, size, part);
    GC_CHECK_LAUNCH();
    hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(256), 0, 0, part, nparts);
    GC_CHECK_LAUNCH();
    dim3 grid((unsigned)((size + 255) / 256));
    hipLaunchKernelGGL(band_kernel, grid, dim3(256), 0, 0, in, indicator, normalised, part, nx, ny,
                       nz, iso1, iso2);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_primitive_field_dynamic(const float *primitive_active, float *isosurf, int32_t nx,
                                   int32_t ny, int32_t nz) {
    clear_error();
    GC_ARG(primitive_active && isosurf, "NULL buffer");
    dim3 grid;
    int st = point_launch(nx, ny, nz, &grid);
    if (st) return st;
    hipLaunchKernelGGL(primitive_dynamic_kernel, grid, dim3(256), 0, 0, primitive_active, isosurf,
                       (uint64_t)nx * ny * nz);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_primitive_field_fixed(const void *primitive_fixed, float *isosurf, int32_t nx,
                                 int32_t ny, int32_t nz) {
    clear_error();
    GC_ARG(primitive_fixed && isosurf, "NULL buffer");
    dim3 grid;
    int st = point_launch(nx, ny, nz, &grid);
    if (st) return st;
    hipLaunchKernelGGL(primitive_fixed_kernel, grid, dim3(256), 0, 0,
                       (const int4 *)primitive_fixed, isosurf, (uint64_t)nx * ny * nz);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

}  // extern "C"
