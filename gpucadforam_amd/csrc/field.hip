// Field evaluation kernels.
//
//  * lattice_field_kernel: the native spatially varying TPMS lattice (see
//    gc_lattice_params in gpucad.h for the definition).  A workgroup is 64 x 4
//    lanes: each lane owns one (x, y) column and walks `lz` consecutive z
//    planes.  Everything that depends on (x, y) only (period, phase offsets,
//    per-axis Fourier factors of the thickness, primitive partial distances)
//    is evaluated once per lane; everything that depends on z only (rotation
//    angle, z Fourier factors, primitive z term) is evaluated once per
//    workgroup into an LDS table and read back as broadcasts.  The per-point
//    work is the TPMS itself, the thickness combination, the sheet and the
//    primitive boolean.  Each z step of a wave stores 64 consecutive floats
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

namespace gc {
namespace {

constexpr int kFieldWX = 64;   // lanes along x per wave
constexpr int kFieldWY = 4;    // waves (y rows) per workgroup
constexpr int kZTab = 16;      // floats per z entry of the LDS table
constexpr int kMaxLZ = 128;    // largest z run per workgroup

// Host-prepared constants of one lattice field launch.
struct FieldConsts {
    int nx, ny, nz, lz;
    uint64_t sign_words_per_row;
    float iso;
    float hx, hy, hz;
    float cx, cy;                 // grid centre (voxels)
    float p0, dp;                 // period(x) = p0 + dp * x
    float inv_p_uniform;          // 1/p0 when the period is uniform
    int uniform_period;
    float th_rev0, dth_rev;       // rotation angle in revolutions: th0 + dth * z
    int bend;
    float bend_r, bend_cx, bend_cy;
    int sheet;
    float t0;
    int n_modes;
    int mode_k[GC_MAX_MODES][3];
    float phi_re[GC_MAX_MODES], phi_im[GC_MAX_MODES];  // a_m * exp(2 pi i phi_m)
    int group_mask;               // bit k set when some mode has kz == k
    int prim, op;
    float pc[3], pr, ph[3];
};

struct Cplx {
    float re, im;
};

__device__ __forceinline__ Cplx cmul(Cplx a, Cplx b) {
    Cplx r;
    r.re = __builtin_fmaf(a.re, b.re, -(a.im * b.im));
    r.im = __builtin_fmaf(a.re, b.im, a.im * b.re);
    return r;
}

__device__ __forceinline__ Cplx expi_rev(float r) {
    Cplx c;
    c.re = cos_rev(r);
    c.im = sin_rev(r);
    return c;
}

// z-table layout
enum {
    ZT_ZF = 0,
    ZT_CTH = 1,
    ZT_STH = 2,
    ZT_E1 = 3,  // re, im for k = 1..3 at 3..8
    ZT_CZ = 9,
    ZT_SZ = 10,
    ZT_PRIMZ = 11,
};

template <bool ROT, int TPMS, bool THICK>
__global__ void __launch_bounds__(kFieldWX *kFieldWY)
lattice_field_kernel(FieldConsts c, float *__restrict__ field, uint64_t *__restrict__ signs) {
    __shared__ float ztab[kMaxLZ * kZTab];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int x = blockIdx.x * kFieldWX + lane;
    const int y = blockIdx.y * kFieldWY + wave;
    const int z0 = blockIdx.z * c.lz;
    const int zn = min(c.lz, c.nz - z0);

    // ---- per-z table (one z per thread of the first waves)
    for (int i = threadIdx.x; i < zn; i += blockDim.x) {
        float *e = ztab + i * kZTab;
        float zf = (float)(z0 + i);
        float pz = zf * c.hz;
        e[ZT_ZF] = zf;
        if (ROT) {
            float th = __builtin_fmaf(c.dth_rev, zf, c.th_rev0);
            e[ZT_CTH] = cos_rev(th);
            e[ZT_STH] = sin_rev(th);
        }
        if (THICK) {
            for (int k = 1; k <= 3; ++k) {
                Cplx ez = expi_rev((float)k * pz);
                e[ZT_E1 + 2 * (k - 1)] = ez.re;
                e[ZT_E1 + 2 * (k - 1) + 1] = ez.im;
            }
        }
        if (!ROT && c.uniform_period) {
            float zr = zf * c.inv_p_uniform;
            e[ZT_CZ] = cos_rev(zr);
            e[ZT_SZ] = sin_rev(zr);
        }
        if (c.prim == GC_PRIM_SPHERE) {
            e[ZT_PRIMZ] = pz - c.pc[2];
        } else if (c.prim == GC_PRIM_BOX) {
            e[ZT_PRIMZ] = __builtin_fabsf(pz - c.pc[2]) - c.ph[2];
        }
    }
    __syncthreads();
    if (y >= c.ny) return;
    const bool xin = x < c.nx;

    // ---- per-lane (x, y) quantities
    const float xf = (float)x, yf = (float)y;
    const float px = xf * c.hx, py = yf * c.hy;
    const float period = __builtin_fmaf(c.dp, xf, c.p0);
    const float inv_p = 1.0f / period;
    float dxP = 0.f, dyP = 0.f, cxP = 0.f, cyP = 0.f;
    float hA = 0.f, hB = 0.f, cX = 0.f, sX = 0.f;  // hoisted TPMS partials (non-rotated)
    if (ROT) {
        dxP = (xf - c.cx) * inv_p;
        dyP = (yf - c.cy) * inv_p;
        cxP = c.cx * inv_p;
        cyP = c.cy * inv_p;
    } else {
        float X, Y;
        if (c.bend) {
            float ax = px - c.bend_cx, ay = py - c.bend_cy;
            float r = __builtin_sqrtf(__builtin_fmaf(ay, ay, ax * ax));
            float phi = atan2_poly(ax, ay);
            float ud = __builtin_fmaf(c.bend_r, phi, c.bend_cx);
            float vd = c.bend_cy + r;
            X = (ud / c.hx) * inv_p;
            Y = (vd / c.hy) * inv_p;
        } else {
            X = xf * inv_p;
            Y = yf * inv_p;
        }
        sX = sin_rev(X);
        cX = cos_rev(X);
        float sY = sin_rev(Y), cY = cos_rev(Y);
        if (TPMS == GC_TPMS_GYROID) {
            hA = sX * cY;  // sX cY
            hB = sY;       // sY (times cZ)
        } else if (TPMS == GC_TPMS_SCHWARZ_P) {
            hA = cX + cY;
        } else {
            hA = __builtin_fmaf(sX, sY, cX * cY);  // multiplies sZ
            hB = __builtin_fmaf(sX, cY, cX * sY);  // multiplies cZ
        }
    }
    // thickness: S_k = sum over modes with kz == k of a_m e^{2 pi i phi_m} Ex Ey
    float T0 = c.t0;
    Cplx S1 = {0.f, 0.f}, S2 = {0.f, 0.f}, S3 = {0.f, 0.f};
    if (THICK) {
        // e^{2 pi i k px} and e^{2 pi i k py} for k = 0..3, selected without
        // dynamic register indexing (which would spill to scratch)
        const Cplx ex1 = expi_rev(px), ex2 = expi_rev(2.0f * px), ex3 = expi_rev(3.0f * px);
        const Cplx ey1 = expi_rev(py), ey2 = expi_rev(2.0f * py), ey3 = expi_rev(3.0f * py);
        auto pick = [](int k, Cplx e1, Cplx e2, Cplx e3) {
            Cplx r = {1.f, 0.f};
            r = k == 1 ? e1 : r;
            r = k == 2 ? e2 : r;
            r = k == 3 ? e3 : r;
            return r;
        };
        Cplx S0 = {0.f, 0.f};
        for (int m = 0; m < c.n_modes; ++m) {
            Cplx w = cmul(pick(c.mode_k[m][0], ex1, ex2, ex3), pick(c.mode_k[m][1], ey1, ey2, ey3));
            Cplx ph = {c.phi_re[m], c.phi_im[m]};
            Cplx a = cmul(w, ph);
            int kz = c.mode_k[m][2];
            if (kz == 0) { S0.re += a.re; S0.im += a.im; }
            if (kz == 1) { S1.re += a.re; S1.im += a.im; }
            if (kz == 2) { S2.re += a.re; S2.im += a.im; }
            if (kz == 3) { S3.re += a.re; S3.im += a.im; }
        }
        T0 = c.t0 + S0.re;
    }
    // primitive partials
    float pxy = 0.f, pbxy = 0.f;
    if (c.prim == GC_PRIM_SPHERE) {
        float dx = px - c.pc[0], dy = py - c.pc[1];
        pxy = __builtin_fmaf(dy, dy, dx * dx);
    } else if (c.prim == GC_PRIM_BOX) {
        float bx = __builtin_fabsf(px - c.pc[0]) - c.ph[0];
        float by = __builtin_fabsf(py - c.pc[1]) - c.ph[1];
        float mx = __builtin_fmaxf(bx, 0.f), my = __builtin_fmaxf(by, 0.f);
        pxy = __builtin_fmaf(my, my, mx * mx);
        pbxy = __builtin_fmaxf(bx, by);
    }

    const uint64_t plane = (uint64_t)c.nx * (uint64_t)c.ny;
    float *out = field + (uint64_t)z0 * plane + (uint64_t)y * c.nx + x;
    uint64_t *srow = signs ? signs + ((uint64_t)z0 * c.ny + y) * c.sign_words_per_row + blockIdx.x
                           : nullptr;
    const uint64_t sstride = (uint64_t)c.ny * c.sign_words_per_row;

    for (int i = 0; i < zn; ++i) {
        const float *e = ztab + i * kZTab;
        float f;
        if (ROT) {
            const float zf = e[ZT_ZF];
            const float cth = e[ZT_CTH], sth = e[ZT_STH];
            float X = __builtin_fmaf(cth, dxP, __builtin_fmaf(-sth, dyP, cxP));
            float Y = __builtin_fmaf(sth, dxP, __builtin_fmaf(cth, dyP, cyP));
            float Z = zf * inv_p;
            if (TPMS == GC_TPMS_SCHWARZ_P) {
                f = (cos_rev(X) + cos_rev(Y)) + cos_rev(Z);
            } else {
                float sx = sin_rev(X), cx = cos_rev(X), sy = sin_rev(Y), cy = cos_rev(Y);
                float sz = sin_rev(Z), cz = cos_rev(Z);
                if (TPMS == GC_TPMS_GYROID)
                    f = __builtin_fmaf(sz, cx, __builtin_fmaf(sy, cz, sx * cy));
                else
                    f = __builtin_fmaf(sz, __builtin_fmaf(sx, sy, cx * cy),
                                       cz * __builtin_fmaf(sx, cy, cx * sy));
            }
        } else {
            float cz, sz;
            if (c.uniform_period) {
                cz = e[ZT_CZ];
                sz = e[ZT_SZ];
            } else {
                float Z = e[ZT_ZF] * inv_p;
                cz = cos_rev(Z);
                sz = sin_rev(Z);
            }
            if (TPMS == GC_TPMS_GYROID)
                f = __builtin_fmaf(sz, cX, __builtin_fmaf(hB, cz, hA));
            else if (TPMS == GC_TPMS_SCHWARZ_P)
                f = hA + cz;
            else
                f = __builtin_fmaf(sz, hA, cz * hB);
        }
        float t = T0;
        if (THICK) {
            if (c.group_mask & 2) {
                t = __builtin_fmaf(S1.re, e[ZT_E1 + 0], t);
                t = __builtin_fmaf(-S1.im, e[ZT_E1 + 1], t);
            }
            if (c.group_mask & 4) {
                t = __builtin_fmaf(S2.re, e[ZT_E1 + 2], t);
                t = __builtin_fmaf(-S2.im, e[ZT_E1 + 3], t);
            }
            if (c.group_mask & 8) {
                t = __builtin_fmaf(S3.re, e[ZT_E1 + 4], t);
                t = __builtin_fmaf(-S3.im, e[ZT_E1 + 5], t);
            }
        }
        float g = c.sheet ? __builtin_fabsf(f) - t : f;
        if (c.prim != GC_PRIM_NONE) {
            float s;
            const float pz = e[ZT_PRIMZ];
            if (c.prim == GC_PRIM_SPHERE) {
                s = __builtin_sqrtf(__builtin_fmaf(pz, pz, pxy)) - c.pr;
            } else {
                float mz = __builtin_fmaxf(pz, 0.f);
                s = __builtin_sqrtf(__builtin_fmaf(mz, mz, pxy)) +
                    __builtin_fminf(__builtin_fmaxf(pbxy, pz), 0.f);
            }
            if (c.op == GC_OP_INTERSECT) g = __builtin_fmaxf(g, s);
            else if (c.op == GC_OP_UNION) g = __builtin_fminf(g, s);
            else if (c.op == GC_OP_SUBTRACT) g = __builtin_fmaxf(g, -s);
        }
        if (xin) out[(uint64_t)i * plane] = g;
        if (srow) {
            uint64_t m = __ballot(xin && g < c.iso);
            if (lane == 0) srow[(uint64_t)i * sstride] = m;
        }
    }
}

typedef void (*FieldKernel)(FieldConsts, float *, uint64_t *);

template <bool ROT, int TPMS>
FieldKernel pick_thick(bool thick) {
    return thick ? lattice_field_kernel<ROT, TPMS, true> : lattice_field_kernel<ROT, TPMS, false>;
}

template <bool ROT>
FieldKernel pick_tpms(int tpms, bool thick) {
    switch (tpms) {
        case GC_TPMS_GYROID: return pick_thick<ROT, GC_TPMS_GYROID>(thick);
        case GC_TPMS_SCHWARZ_P: return pick_thick<ROT, GC_TPMS_SCHWARZ_P>(thick);
        default: return pick_thick<ROT, GC_TPMS_DIAMOND>(thick);
    }
}

}  // namespace

int validate_lattice(const gc_lattice_params *p) {
    GC_ARG(p, "params is NULL");
    GC_ARG(p->nx >= 2 && p->ny >= 2 && p->nz >= 2, "grid must have >= 2 points per axis");
    GC_ARG((uint64_t)p->nx * p->ny * p->nz < (1ull << 40), "grid too large");
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
    // default z run: long enough to amortise the per-column work, short enough
    // to leave about 8 workgroups per z column of the grid
    if (lz <= 0) lz = std::max(16, std::min(kMaxLZ, p->nz / 8));
    if (lz > kMaxLZ) lz = kMaxLZ;
    FieldConsts c;
    std::memset(&c, 0, sizeof(c));
    c.nx = p->nx;
    c.ny = p->ny;
    c.nz = p->nz;
    c.lz = lz;
    c.sign_words_per_row = gpucad_sign_words(p->nx, 1, 1);
    c.iso = iso;
    c.hx = (float)(1.0 / (p->nx - 1));
    c.hy = (float)(1.0 / (p->ny - 1));
    c.hz = (float)(1.0 / (p->nz - 1));
    c.cx = 0.5f * (float)(p->nx - 1);
    c.cy = 0.5f * (float)(p->ny - 1);
    c.p0 = p->period0;
    c.dp = (float)(((double)p->period1 - (double)p->period0) / (double)(p->nx - 1));
    c.uniform_period = p->period0 == p->period1;
    c.inv_p_uniform = 1.0f / p->period0;
    const bool rot = p->rot0 != 0.f || p->rot1 != 0.f;
    const double two_pi = 6.283185307179586476925286766559;
    c.th_rev0 = (float)((double)p->rot0 / two_pi);
    c.dth_rev = (float)(((double)p->rot1 - (double)p->rot0) / (two_pi * (double)(p->nz - 1)));
    c.bend = p->bend_radius > 0.f;
    c.bend_r = p->bend_radius;
    c.bend_cx = 0.5f;
    c.bend_cy = (float)(0.5 - (double)p->bend_radius);
    c.sheet = p->sheet != 0;
    c.t0 = p->t0;
    c.n_modes = p->n_modes;
    c.group_mask = 0;
    for (int m = 0; m < p->n_modes; ++m) {
        for (int a = 0; a < 3; ++a) c.mode_k[m][a] = p->mode_k[m][a];
        c.phi_re[m] = (float)((double)p->mode_a[m] * std::cos(two_pi * (double)p->mode_phi[m]));
        c.phi_im[m] = (float)((double)p->mode_a[m] * std::sin(two_pi * (double)p->mode_phi[m]));
        c.group_mask |= 1 << p->mode_k[m][2];
    }
    c.prim = p->prim;
    c.op = p->prim == GC_PRIM_NONE ? GC_OP_NONE : p->op;
    for (int a = 0; a < 3; ++a) {
        c.pc[a] = p->prim_center[a];
        c.ph[a] = p->prim_half[a];
    }
    c.pr = p->prim_radius;
    const bool thick = p->n_modes > 0;
    FieldKernel k = rot ? pick_tpms<true>(p->tpms, thick) : pick_tpms<false>(p->tpms, thick);
    dim3 grid((p->nx + kFieldWX - 1) / kFieldWX, (p->ny + kFieldWY - 1) / kFieldWY,
              (p->nz + lz - 1) / lz);
    GC_ARG(grid.y <= 65535 && grid.z <= 65535, "grid too large for one launch");
    hipLaunchKernelGGL(k, grid, dim3(kFieldWX * kFieldWY), 0, stream, c, d_field, d_signs);
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
    GC_ARG(in && indicator && normalised, "NULL buffer");
    GC_ARG(nx >= 1 && ny >= 1 && nz >= 1, "grid sizes must be positive");
    GC_ARG(size == (uint64_t)nx * ny * nz, "size must equal nx*ny*nz");
    const int nparts = 1024;
    void *ws = nullptr;
    int st = scratch(nparts * sizeof(float2), &ws);
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
