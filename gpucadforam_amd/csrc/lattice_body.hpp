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

// Primitive of a lattice field known at compile time (PRIM_ANY: read from
// the constants at run time).
constexpr int PRIM_ANY = -1;

// Evaluates the brick with first point (x0, y0, z0) of the grid `c`.  Must be
// called by all 256 threads of the workgroup (it synchronises through LDS);
// `ztab` is kMaxLZ * kZTab floats of LDS.  `field` and `signs` point at the
// grid's own first point and sign word; `signs` may be NULL.  PRIM and OP are
// c.prim and c.op, or PRIM_ANY (read at run time).  The per-point loop has no data-independent branches: absent
// thickness groups have zero coefficients (their fmas add exactly 0) and the
// boolean is a select among max(g, s), min(g, s), max(g, -s).
template <bool ROT, int TPMS, bool THICK, int BX, int PRIM = PRIM_ANY, int OP = PRIM_ANY>
__device__ __forceinline__ void lattice_brick(const FieldConsts &c, float *__restrict__ field,
                                              uint64_t *__restrict__ signs, int x0, int y0,
                                              int z0, float *ztab) {
    constexpr int BY = 64 / BX;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int lx = lane % BX, ly = lane / BX;
    const int x = x0 + lx;
    const int y = y0 + wave * BY + ly;
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
    const bool yin = y < c.ny;
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
    const int yrow = yin ? y : 0;  // rows past the grid compute but never store
    // stores address a z plane through a base shared by the workgroup plus a
    // 32-bit byte offset per lane (a plane of the field and of the sign words
    // is well below 4 GiB), so the per-plane step is scalar work
    char *fplane = (char *)(field + (uint64_t)z0 * plane);
    const uint32_t foff = 4u * ((uint32_t)yrow * (uint32_t)c.nx + (uint32_t)x);
    const uint64_t fstep = 4 * plane;
    const uint64_t sstep = 8 * (uint64_t)c.ny * c.sign_words_per_row;
    char *splane = signs ? (char *)(signs + (uint64_t)z0 * c.ny * c.sign_words_per_row) : nullptr;
    const uint32_t soff = 8u * ((uint32_t)yrow * (uint32_t)c.sign_words_per_row);
    const bool sign_lane = signs != nullptr && lx == 0 && yin;
    // lanes whose point lies in the grid (the sign mask keeps only these)
    const uint64_t in_mask = __ballot(xin);

#pragma unroll 2
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
            // a group without modes has S = 0: its fmas add exactly 0
            t = __builtin_fmaf(S1.re, e[ZT_E1 + 0], t);
            t = __builtin_fmaf(-S1.im, e[ZT_E1 + 1], t);
            t = __builtin_fmaf(S2.re, e[ZT_E1 + 2], t);
            t = __builtin_fmaf(-S2.im, e[ZT_E1 + 3], t);
            t = __builtin_fmaf(S3.re, e[ZT_E1 + 4], t);
            t = __builtin_fmaf(-S3.im, e[ZT_E1 + 5], t);
        }
        float g = c.sheet ? __builtin_fabsf(f) - t : f;
        const int prim = PRIM == PRIM_ANY ? c.prim : PRIM;
        if (prim != GC_PRIM_NONE) {
            float s;
            const float pz = e[ZT_PRIMZ];
            if (prim == GC_PRIM_SPHERE) {
                s = sqrt_rn(__builtin_fmaf(pz, pz, pxy)) - c.pr;
            } else {
                float mz = __builtin_fmaxf(pz, 0.f);
                s = sqrt_rn(__builtin_fmaf(mz, mz, pxy)) +
                    __builtin_fminf(__builtin_fmaxf(pbxy, pz), 0.f);
            }
            const int op = OP == PRIM_ANY ? c.op : OP;
            const float gi = __builtin_fmaxf(g, s), gu = __builtin_fminf(g, s),
                        gs = __builtin_fmaxf(g, -s);
            g = op == GC_OP_INTERSECT ? gi
                                      : (op == GC_OP_UNION ? gu : (op == GC_OP_SUBTRACT ? gs : g));
        }
        if (xin && yin) *(float *)(fplane + (uint64_t)i * fstep + foff) = g;
        if (splane) {
            // one compare gives the wave's mask of lanes with g < iso (4 is
            // the ordered less-than predicate: NaN compares false)
            const uint64_t m = __builtin_amdgcn_fcmpf(g, c.iso, 4) & in_mask;
            if (sign_lane) {
                const uint64_t bits =
                    BX == 64 ? m : (m >> (ly * BX)) & ((1ull << (BX & 63)) - 1ull);
                store_sign_bits<BX>((uint64_t *)(splane + (uint64_t)i * sstep + soff), x0, bits);
            }
        }
    }
}

}  // namespace gc
