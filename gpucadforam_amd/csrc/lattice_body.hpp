// The native spatially varying TPMS lattice field, evaluated one brick at a
// time (shared by the one-grid launch in field.hip and the batched launch in
// batch.hip, so both produce the same bits).
//
// A brick is what one workgroup of 4 waves evaluates: each wave owns BX x BY
// (x, y) columns (BX * BY = 64 lanes, the 4 waves stacked along y, so the
// brick is BX x 4*BY columns) and walks `lz` consecutive z planes.  Work that
// depends on (x, y) only (period, phase offsets, per-axis Fourier factors of
// the thickness, primitive partial distances) is done once per lane; work that
// depends on z only (rotation angle, z Fourier factors, primitive z term) is
// done once per brick into an LDS table read back as broadcasts.  Each z step
// of a row segment stores BX consecutive floats and the BX sign bits
// [G < iso] of the segment (bit b of sign word w of a row = point 64w + b).
#pragma once

#include <cmath>
#include <cstring>

#include "gc_common.hpp"
#include "gc_math.hpp"

namespace gc {

constexpr int kFieldWaves = 4;  // waves per brick workgroup
constexpr int kZTab = 16;       // floats per z entry of the LDS table
constexpr int kMaxLZ = 128;     // largest z run of a brick

// Host-prepared constants of one lattice grid.
struct FieldConsts {
    int nx, ny, nz, lz;
    uint64_t sign_words_per_row;
    uint64_t field_off;           // first point of the grid in the field buffer (floats)
    uint64_t sign_off;            // first sign word of the grid in the sign buffer
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

// Field variant: which template instance evaluates a grid.
struct FieldVariant {
    bool rot;
    int tpms;
    bool thick;
    int index() const { return (rot ? 6 : 0) + tpms * 2 + (thick ? 1 : 0); }
};

inline FieldVariant field_variant(const gc_lattice_params *p) {
    FieldVariant v;
    v.rot = p->rot0 != 0.f || p->rot1 != 0.f;
    v.tpms = p->tpms;
    v.thick = p->n_modes > 0;
    return v;
}

int validate_lattice(const gc_lattice_params *p);

// Fills the constants of grid `p` (validated) for isovalue `iso` and z run `lz`.
inline FieldConsts make_field_consts(const gc_lattice_params *p, float iso, int lz) {
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
    return c;
}

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

// Stores the `bits` sign bits of a row segment of BX points starting at x0
// (x0 a multiple of BX, BX dividing 64) into the row's sign words.
template <int BX>
__device__ __forceinline__ void store_sign_bits(uint64_t *row, int x0, uint64_t bits) {
    if (BX == 64) {
        row[x0 >> 6] = bits;
    } else {
        char *p = (char *)(row + (x0 >> 6)) + ((x0 & 63) >> 3);
        if (BX == 32) *(uint32_t *)p = (uint32_t)bits;
        else if (BX == 16) *(uint16_t *)p = (uint16_t)bits;
        else *(uint8_t *)p = (uint8_t)bits;
    }
}

// Buffer stores of the brick: word 3 of a gfx9 buffer resource for untyped
// access, the offset given to lanes that must not store (past any plane:
// planes are below 4 GiB, validated by the launches), and the non-temporal
// cache-policy bit of the store.
constexpr int kBufferWord3 = 0x00020000;
constexpr uint32_t kPastPlane = 0xfffffff0u;
constexpr int kStoreNonTemporal = 2;
typedef uint32_t U32x2 __attribute__((ext_vector_type(2)));

// Primitive of a lattice field known at compile time (PRIM_ANY: read from
// the constants at run time).
constexpr int PRIM_ANY = -1;

// The lattice value g combined with the primitive whose distance root is
// `root` (sphere: root - r; box: root + min(max(partial, z term), 0)).
template <int PRIM, int OP>
__device__ __forceinline__ float prim_op(const FieldConsts &c, float g, float root, float pz,
                                         float pbxy) {
    const int prim = PRIM == PRIM_ANY ? c.prim : PRIM;
    const float s = prim == GC_PRIM_SPHERE
                        ? root - c.pr
                        : root + __builtin_fminf(__builtin_fmaxf(pbxy, pz), 0.f);
    const int op = OP == PRIM_ANY ? c.op : OP;
    const float gi = __builtin_fmaxf(g, s), gu = __builtin_fminf(g, s), gs = __builtin_fmaxf(g, -s);
    return op == GC_OP_INTERSECT ? gi
                                 : (op == GC_OP_UNION ? gu : (op == GC_OP_SUBTRACT ? gs : g));
}

// Evaluates the brick with first point (x0, y0, z0) of the grid `c`.  Must be
// called by all 256 threads of the workgroup (it synchronises through LDS);
// `ztab` is kMaxLZ * kZTab floats of LDS.  `field` and `signs` point at the
// grid's own first point and sign word; `signs` may be NULL.  PRIM and OP are
// c.prim and c.op, or PRIM_ANY (read at run time).  The per-point loop has no
// data-independent branches: absent thickness groups have zero coefficients
// (their fmas add exactly 0) and the boolean is a select among max(g, s),
// min(g, s), max(g, -s).
//
// RY > 1 (one-grid launches, BX = 64): each lane owns RY consecutive rows of
// its x column, so the brick is 64 x 4 RY columns and the work that depends
// on (x, z) only (the z table reads, the z phase and its cos / sin) is done
// once per RY points.  Every point's arithmetic is the same for any RY.
template <bool ROT, int TPMS, bool THICK, int BX, int PRIM = PRIM_ANY, int OP = PRIM_ANY,
          int RY = 1>
__device__ __forceinline__ void lattice_brick(const FieldConsts &c, float *__restrict__ field,
                                              uint64_t *__restrict__ signs, int x0, int y0,
                                              int z0, float *ztab) {
    static_assert(RY == 1 || BX == 64, "rows per lane need one row per wave");
    constexpr int BY = 64 / BX;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int lx = lane % BX, ly = lane / BX;
    const int x = x0 + lx;
    const int yb = y0 + (wave * BY + ly) * RY;  // the lane's first row
    const int zn = min(c.lz, c.nz - z0);

    // ---- per-z table (one z per thread)
    for (int i = threadIdx.x; i < zn; i += 64 * kFieldWaves) {
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
    const bool xin = x < c.nx;

    // ---- per-lane x quantities
    const float xf = (float)x;
    const float px = xf * c.hx;
    const float period = __builtin_fmaf(c.dp, xf, c.p0);
    const float inv_p = 1.0f / period;
    float dxP = 0.f, cxP = 0.f, cyP = 0.f;
    if (ROT) {
        dxP = (xf - c.cx) * inv_p;
        cxP = c.cx * inv_p;
        cyP = c.cy * inv_p;
    }
    Cplx ex1 = {1.f, 0.f}, ex2 = ex1, ex3 = ex1;
    if (THICK) {
        ex1 = expi_rev(px);
        ex2 = expi_rev(2.0f * px);
        ex3 = expi_rev(3.0f * px);
    }
    // ---- per-row (x, y) quantities
    bool yin[RY];
    float dyP[RY], hA[RY], hB[RY], cX[RY];  // hoisted TPMS partials (non-rotated)
    float T0[RY];
    Cplx S1[RY], S2[RY], S3[RY];
    float pxy[RY], pbxy[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int y = yb + r;
        yin[r] = y < c.ny;
        const float yf = (float)y;
        const float py = yf * c.hy;
        dyP[r] = hA[r] = hB[r] = cX[r] = 0.f;
        if (ROT) {
            dyP[r] = (yf - c.cy) * inv_p;
        } else {
            float X, Y;
            if (c.bend) {
                float ax = px - c.bend_cx, ay = py - c.bend_cy;
                float rr = __builtin_sqrtf(__builtin_fmaf(ay, ay, ax * ax));
                float phi = atan2_poly(ax, ay);
                float ud = __builtin_fmaf(c.bend_r, phi, c.bend_cx);
                float vd = c.bend_cy + rr;
                X = (ud / c.hx) * inv_p;
                Y = (vd / c.hy) * inv_p;
            } else {
                X = xf * inv_p;
                Y = yf * inv_p;
            }
            const float sX = sin_rev(X);
            cX[r] = cos_rev(X);
            float sY = sin_rev(Y), cY = cos_rev(Y);
            if (TPMS == GC_TPMS_GYROID) {
                hA[r] = sX * cY;  // sX cY
                hB[r] = sY;       // sY (times cZ)
            } else if (TPMS == GC_TPMS_SCHWARZ_P) {
                hA[r] = cX[r] + cY;
            } else {
                hA[r] = __builtin_fmaf(sX, sY, cX[r] * cY);  // multiplies sZ
                hB[r] = __builtin_fmaf(sX, cY, cX[r] * sY);  // multiplies cZ
            }
        }
        // thickness: S_k = sum over modes with kz == k of a_m e^{2 pi i phi_m} Ex Ey
        T0[r] = c.t0;
        S1[r] = S2[r] = S3[r] = Cplx{0.f, 0.f};
        if (THICK) {
            // e^{2 pi i k py} for k = 0..3, selected without dynamic register
            // indexing (which would spill to scratch)
            const Cplx ey1 = expi_rev(py), ey2 = expi_rev(2.0f * py), ey3 = expi_rev(3.0f * py);
            auto pick = [](int k, Cplx e1, Cplx e2, Cplx e3) {
                Cplx q = {1.f, 0.f};
                q = k == 1 ? e1 : q;
                q = k == 2 ? e2 : q;
                q = k == 3 ? e3 : q;
                return q;
            };
            Cplx S0 = {0.f, 0.f};
            for (int m = 0; m < c.n_modes; ++m) {
                Cplx w = cmul(pick(c.mode_k[m][0], ex1, ex2, ex3),
                              pick(c.mode_k[m][1], ey1, ey2, ey3));
                Cplx ph = {c.phi_re[m], c.phi_im[m]};
                Cplx a = cmul(w, ph);
                int kz = c.mode_k[m][2];
                if (kz == 0) { S0.re += a.re; S0.im += a.im; }
                if (kz == 1) { S1[r].re += a.re; S1[r].im += a.im; }
                if (kz == 2) { S2[r].re += a.re; S2[r].im += a.im; }
                if (kz == 3) { S3[r].re += a.re; S3[r].im += a.im; }
            }
            T0[r] = c.t0 + S0.re;
        }
        // primitive partials
        pxy[r] = pbxy[r] = 0.f;
        if (c.prim == GC_PRIM_SPHERE) {
            float dx = px - c.pc[0], dy = py - c.pc[1];
            pxy[r] = __builtin_fmaf(dy, dy, dx * dx);
        } else if (c.prim == GC_PRIM_BOX) {
            float bx = __builtin_fabsf(px - c.pc[0]) - c.ph[0];
            float by = __builtin_fabsf(py - c.pc[1]) - c.ph[1];
            float mx = __builtin_fmaxf(bx, 0.f), my = __builtin_fmaxf(by, 0.f);
            pxy[r] = __builtin_fmaf(my, my, mx * mx);
            pbxy[r] = __builtin_fmaxf(bx, by);
        }
    }

    const uint64_t plane = (uint64_t)c.nx * (uint64_t)c.ny;
    // stores address a z plane through a base shared by the workgroup plus a
    // 32-bit byte offset per lane (a plane of the field and of the sign words
    // is well below 4 GiB); rows past the grid compute but never store
    char *fplane = (char *)(field + (uint64_t)z0 * plane);
    const uint64_t fstep = 4 * plane;
    const uint64_t sstep = 8 * (uint64_t)c.ny * c.sign_words_per_row;
    char *splane = signs ? (char *)(signs + (uint64_t)z0 * c.ny * c.sign_words_per_row) : nullptr;
    uint32_t foff[RY], soff[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const uint32_t yrow = yin[r] ? (uint32_t)(yb + r) : 0u;
        foff[r] = 4u * (yrow * (uint32_t)c.nx + (uint32_t)x);
        soff[r] = 8u * (yrow * (uint32_t)c.sign_words_per_row);
    }
    const bool sign_lane = signs != nullptr && lx == 0;
    // lanes whose point lies in the grid (the sign mask keeps only these)
    const uint64_t in_mask = __ballot(xin);

    for (int i = 0; i < zn; ++i) {
        const float *e = ztab + i * kZTab;
        // (x, z) work, shared by the lane's rows
        float cth = 0.f, sth = 0.f, cZ = 0.f, sZ = 0.f, Zr = 0.f;
        if (ROT) {
            cth = e[ZT_CTH];
            sth = e[ZT_STH];
            Zr = e[ZT_ZF] * inv_p;
            if (TPMS == GC_TPMS_SCHWARZ_P) {
                cZ = cos_rev(Zr);
            } else {
                sZ = sin_rev(Zr);
                cZ = cos_rev(Zr);
            }
        } else {
            if (c.uniform_period) {
                cZ = e[ZT_CZ];
                sZ = e[ZT_SZ];
            } else {
                const float Z = e[ZT_ZF] * inv_p;
                cZ = cos_rev(Z);
                sZ = sin_rev(Z);
            }
        }
        float E[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (THICK) {
#pragma unroll
            for (int q = 0; q < 6; ++q) E[q] = e[ZT_E1 + q];
        }
        const int prim = PRIM == PRIM_ANY ? c.prim : PRIM;
        const float pz = prim != GC_PRIM_NONE ? e[ZT_PRIMZ] : 0.f;
        // every row's value first (straight-line code, so the rows' independent
        // chains interleave), then the stores, which are guarded per row
        float gv[RY];
        float gl[RY], px2[RY];  // the lattice value and the primitive's squared distance
        bool small = false;     // a squared distance in (0, 2^-96) (or NaN)
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            float f;
            if (ROT) {
                float X = __builtin_fmaf(cth, dxP, __builtin_fmaf(-sth, dyP[r], cxP));
                float Y = __builtin_fmaf(sth, dxP, __builtin_fmaf(cth, dyP[r], cyP));
                if (TPMS == GC_TPMS_SCHWARZ_P) {
                    f = (cos_rev(X) + cos_rev(Y)) + cZ;
                } else {
                    float sx = sin_rev(X), cx = cos_rev(X), sy = sin_rev(Y), cy = cos_rev(Y);
                    if (TPMS == GC_TPMS_GYROID)
                        f = __builtin_fmaf(sZ, cx, __builtin_fmaf(sy, cZ, sx * cy));
                    else
                        f = __builtin_fmaf(sZ, __builtin_fmaf(sx, sy, cx * cy),
                                           cZ * __builtin_fmaf(sx, cy, cx * sy));
                }
            } else {
                if (TPMS == GC_TPMS_GYROID)
                    f = __builtin_fmaf(sZ, cX[r], __builtin_fmaf(hB[r], cZ, hA[r]));
                else if (TPMS == GC_TPMS_SCHWARZ_P)
                    f = hA[r] + cZ;
                else
                    f = __builtin_fmaf(sZ, hA[r], cZ * hB[r]);
            }
            float t = T0[r];
            if (THICK) {
                // a group without modes has S = 0: its fmas add exactly 0
                t = __builtin_fmaf(S1[r].re, E[0], t);
                t = __builtin_fmaf(-S1[r].im, E[1], t);
                t = __builtin_fmaf(S2[r].re, E[2], t);
                t = __builtin_fmaf(-S2[r].im, E[3], t);
                t = __builtin_fmaf(S3[r].re, E[4], t);
                t = __builtin_fmaf(-S3[r].im, E[5], t);
            }
            float g = c.sheet ? __builtin_fabsf(f) - t : f;
            gl[r] = g;
            px2[r] = 0.f;
            if (prim != GC_PRIM_NONE) {
                const float mz = __builtin_fmaxf(pz, 0.f);
                px2[r] = prim == GC_PRIM_SPHERE ? __builtin_fmaf(pz, pz, pxy[r])
                                                : __builtin_fmaf(mz, mz, pxy[r]);
                // exact zeros (a box's inside) are exact on the fast path too
                small |= !(px2[r] >= 0x1p-96f) && px2[r] != 0.0f;
                g = prim_op<PRIM, OP>(c, g, sqrt_rn_normal(px2[r]), pz, pbxy[r]);
            }
            gv[r] = g;
        }
        // a squared distance in (0, 2^-96), a point within 2^-48 of the
        // primitive's centre: the wave takes the rescaled root for every row
        // (one uniform test per z step, after the rows' arithmetic)
        if (prim != GC_PRIM_NONE && __ballot(small) != 0) {
#pragma unroll
            for (int r = 0; r < RY; ++r)
                gv[r] = prim_op<PRIM, OP>(c, gl[r], sqrt_rn(px2[r]), pz, pbxy[r]);
        }
        if (BX == 64) {
            // buffer stores over this z plane: a lane whose point lies outside
            // the grid (and every lane but the first for the sign word) gets
            // an offset past the plane, which the buffer's range check
            // discards, so the stores need no branch and the rows' arithmetic
            // above stays one block
            const __amdgpu_buffer_rsrc_t fr =
                __builtin_amdgcn_make_buffer_rsrc(fplane + (uint64_t)i * fstep, (short)0,
                                                  (int)(uint32_t)fstep, kBufferWord3);
#pragma unroll
            for (int r = 0; r < RY; ++r)
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(gv[r]), fr,
                                                      (int)((xin && yin[r]) ? foff[r] : kPastPlane),
                                                      0, kStoreNonTemporal);
            if (splane) {
                // the rows' sign words in one store: lane r holds row r's
                // word (one compare per row gives the wave's mask of lanes with
                // g < iso; 4 is the ordered less-than predicate, NaN compares
                // false), every other lane an offset past the plane
                const __amdgpu_buffer_rsrc_t sr =
                    __builtin_amdgcn_make_buffer_rsrc(splane + (uint64_t)i * sstep, (short)0,
                                                      (int)(uint32_t)sstep, kBufferWord3);
                uint64_t mine = 0;
                uint32_t moff = kPastPlane;
#pragma unroll
                for (int r = 0; r < RY; ++r) {
                    const uint64_t m = __builtin_amdgcn_fcmpf(gv[r], c.iso, 4) & in_mask;
                    const bool me = signs != nullptr && lane == r && yin[r];
                    mine = me ? m : mine;
                    moff = me ? soff[r] + 8u * (uint32_t)(x0 >> 6) : moff;
                }
                const U32x2 w = {(uint32_t)mine, (uint32_t)(mine >> 32)};
                __builtin_amdgcn_raw_buffer_store_b64(w, sr, (int)moff, 0, 0);
            }
            continue;
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const float g = gv[r];
            // the field is written once per launch (the extraction reads it
            // back from HBM): non-temporal stores
            if (xin && yin[r])
                __builtin_nontemporal_store(g, (float *)(fplane + (uint64_t)i * fstep + foff[r]));
            if (splane) {
                // one compare gives the wave's mask of lanes with g < iso (4 is
                // the ordered less-than predicate: NaN compares false)
                const uint64_t m = __builtin_amdgcn_fcmpf(g, c.iso, 4) & in_mask;
                if (sign_lane && yin[r]) {
                    const uint64_t bits =
                        BX == 64 ? m : (m >> (ly * BX)) & ((1ull << (BX & 63)) - 1ull);
                    store_sign_bits<BX>((uint64_t *)(splane + (uint64_t)i * sstep + soff[r]), x0,
                                        bits);
                }
            }
        }
    }
}

}  // namespace gc
