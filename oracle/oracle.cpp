/*
 * ORACLE — test infrastructure only.
 *
 * A host C++ restatement of the reference's hot-path arithmetic, used as the
 * checker of the HIP library (tests, smoke check) and as the CPU baseline of
 * the benchmark.  It is never linked into, loaded by or called from the
 * library's product path.
 *
 * Each function cites the reference lines it restates.  Parity pins (which
 * independent fact each part is checked against) are listed in the tests:
 *   - case tables     : src/tables.h when present; 256-case brute force
 *   - classify/scan   : definitions (src/MarchingCubes_kernel.cu:3207-3249,
 *                       3198-3203, 1594-1608) + invariants
 *   - extraction      : linear fields (exact t), sphere (O(h^2)), closed-surface
 *                       manifold/Euler checks, FP64 twin
 *   - lattice field   : FP64 closed form of the formula (independent code path)
 *   - sin/cos/atan    : libm in double precision, error bound 2e-7
 *   - primitives      : float64 geometry of each shape, scipy Euler rotations
 *                       (tests/test_oracle_primitives.py)
 *   - band interp     : linear fields, snapping, mirror symmetry
 *                       (tests/test_oracle_band.py); parity unpinned for the
 *                       equal-values branch (t = 1 when p0.z == 0) and the
 *                       uninitialised-t case (defined as 0)
 *   - SVL             : operator = sum of Dt Dt^T, linear phases, closed-form
 *                       upsampling (tests/test_oracle_svl.py)
 *   - unit-cell lattice, cuFFT spectra, CG reduction order: tolerances only
 *                       (parity unpinned at bit level)
 *   - .obj writer     : std::map restatement of src/File_output.cu:5-81
 *
 * Built with g++ -O3 -fopenmp -ffp-contract=off: every fused multiply-add is
 * an explicit fmaf, as in the kernels, so the binary32 paths reproduce the
 * device arithmetic operation by operation.
 */
#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <vector>

#include "../include/gpucad.h"

namespace {

// ---------------------------------------------------------------------------
// Tables (src/tables.h:11-45, 49-307, 311-569)
// ---------------------------------------------------------------------------
const char *const kCases[256] = {
#include "mc_cases.inc"
};

const int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                           {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
const int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                          {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

struct Tables {
    uint8_t tri[256][16];
    uint8_t nverts[256];
    Tables() {
        for (int c = 0; c < 256; ++c) {
            const char *s = kCases[c];
            int n = (int)strlen(s);
            for (int j = 0; j < 16; ++j) tri[c][j] = 255;
            for (int j = 0; j < n; ++j)
                tri[c][j] = (uint8_t)(s[j] <= '9' ? s[j] - '0' : s[j] - 'a' + 10);
            nverts[c] = (uint8_t)n;
        }
    }
};
const Tables T;

// ---------------------------------------------------------------------------
// binary32 sin/cos in revolutions and atan2 (same constants and operation
// order as the device; see derive_poly.py for the Remez derivation)
// ---------------------------------------------------------------------------
const float S1 = 0x1.921fb4p+2f, S3 = -0x1.4abbb6p+5f, S5 = 0x1.46676ep+6f,
            S7 = -0x1.3232fap+6f, S9 = 0x1.3c4b2cp+5f;

inline float sin2pi_red(float d) {
    float d2 = d * d;
    float p = fmaf(S9, d2, S7);
    p = fmaf(p, d2, S5);
    p = fmaf(p, d2, S3);
    p = fmaf(p, d2, S1);
    return d * p;
}
inline float cos_rev(float r) {
    float a = r - rintf(r);
    float d = 0.25f - fabsf(a);
    return sin2pi_red(d);
}
inline float sin_rev(float r) {
    float a = r - rintf(r);
    float d = (fabsf(a) > 0.25f) ? copysignf(0.5f, a) - a : a;
    return sin2pi_red(d);
}

const float A[11] = {0x1p+0f,       -0x1.55554p-2f, 0x1.99950ap-3f, -0x1.245722p-3f,
                     0x1.c3bbccp-4f, -0x1.64eb78p-4f, 0x1.0bef6p-4f, -0x1.58c94ep-5f,
                     0x1.4f589ap-6f, -0x1.a316f6p-8f, 0x1.eb8dep-11f};
const float HALF_PI = 0x1.921fb6p+0f, PI_F = 0x1.921fb6p+1f;

inline float atan2_poly(float y, float x) {
    float ax = fabsf(x), ay = fabsf(y);
    float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    float a = (mx == 0.0f) ? 0.0f : mn / mx;
    float a2 = a * a;
    float q = fmaf(A[10], a2, A[9]);
    for (int k = 8; k >= 0; --k) q = fmaf(q, a2, A[k]);
    float t = a * q;
    if (ay > ax) t = HALF_PI - t;
    if (x < 0.0f) t = PI_F - t;
    if (y < 0.0f) t = -t;
    return t;
}

struct Cx {
    float re, im;
};
inline Cx cmul(Cx a, Cx b) {
    return Cx{fmaf(a.re, b.re, -(a.im * b.im)), fmaf(a.re, b.im, a.im * b.re)};
}
inline Cx expi(float r) { return Cx{cos_rev(r), sin_rev(r)}; }

const double TWO_PI = 6.283185307179586476925286766559;

// ---------------------------------------------------------------------------
// Native lattice field, binary32 (the operation sequence of the definition in
// gpucad.h, evaluated per point)
// ---------------------------------------------------------------------------
struct LatticeHost {
    float hx, hy, hz, cx, cy, p0, dp, inv_p0, th0, dth, bcx, bcy;
    bool rot, uniform;
    float phre[8], phim[8];
    int gmask;
};

LatticeHost lattice_host(const gc_lattice_params &p) {
    LatticeHost h;
    h.hx = (float)(1.0 / (p.nx - 1));
    h.hy = (float)(1.0 / (p.ny - 1));
    h.hz = (float)(1.0 / (p.nz - 1));
    h.cx = 0.5f * (float)(p.nx - 1);
    h.cy = 0.5f * (float)(p.ny - 1);
    h.p0 = p.period0;
    h.dp = (float)(((double)p.period1 - (double)p.period0) / (double)(p.nx - 1));
    h.inv_p0 = 1.0f / p.period0;
    h.uniform = p.period0 == p.period1;
    h.rot = p.rot0 != 0.f || p.rot1 != 0.f;
    h.th0 = (float)((double)p.rot0 / TWO_PI);
    h.dth = (float)(((double)p.rot1 - (double)p.rot0) / (TWO_PI * (double)(p.nz - 1)));
    h.bcx = 0.5f;
    h.bcy = (float)(0.5 - (double)p.bend_radius);
    h.gmask = 0;
    for (int m = 0; m < p.n_modes; ++m) {
        h.phre[m] = (float)((double)p.mode_a[m] * std::cos(TWO_PI * (double)p.mode_phi[m]));
        h.phim[m] = (float)((double)p.mode_a[m] * std::sin(TWO_PI * (double)p.mode_phi[m]));
        h.gmask |= 1 << p.mode_k[m][2];
    }
    return h;
}

float lattice_point_f32(const gc_lattice_params &p, const LatticeHost &h, int x, int y, int z) {
    const float xf = (float)x, yf = (float)y, zf = (float)z;
    const float px = xf * h.hx, py = yf * h.hy, pz = zf * h.hz;
    const float period = fmaf(h.dp, xf, h.p0);
    const float inv_p = 1.0f / period;
    float f;
    if (h.rot) {
        float th = fmaf(h.dth, zf, h.th0);
        float cth = cos_rev(th), sth = sin_rev(th);
        float dxP = (xf - h.cx) * inv_p, dyP = (yf - h.cy) * inv_p;
        float cxP = h.cx * inv_p, cyP = h.cy * inv_p;
        float X = fmaf(cth, dxP, fmaf(-sth, dyP, cxP));
        float Y = fmaf(sth, dxP, fmaf(cth, dyP, cyP));
        float Z = zf * inv_p;
        if (p.tpms == GC_TPMS_SCHWARZ_P) {
            f = (cos_rev(X) + cos_rev(Y)) + cos_rev(Z);
        } else {
            float sx = sin_rev(X), cx = cos_rev(X), sy = sin_rev(Y), cy = cos_rev(Y);
            float sz = sin_rev(Z), cz = cos_rev(Z);
            if (p.tpms == GC_TPMS_GYROID)
                f = fmaf(sz, cx, fmaf(sy, cz, sx * cy));
            else
                f = fmaf(sz, fmaf(sx, sy, cx * cy), cz * fmaf(sx, cy, cx * sy));
        }
    } else {
        float X, Y;
        if (p.bend_radius > 0.f) {
            float ax = px - h.bcx, ay = py - h.bcy;
            float r = sqrtf(fmaf(ay, ay, ax * ax));
            float phi = atan2_poly(ax, ay);
            float ud = fmaf(p.bend_radius, phi, h.bcx);
            float vd = h.bcy + r;
            X = (ud / h.hx) * inv_p;
            Y = (vd / h.hy) * inv_p;
        } else {
            X = xf * inv_p;
            Y = yf * inv_p;
        }
        float sX = sin_rev(X), cX = cos_rev(X), sY = sin_rev(Y), cY = cos_rev(Y);
        float cz, sz;
        if (h.uniform) {
            float zr = zf * h.inv_p0;
            cz = cos_rev(zr);
            sz = sin_rev(zr);
        } else {
            float Z = zf * inv_p;
            cz = cos_rev(Z);
            sz = sin_rev(Z);
        }
        if (p.tpms == GC_TPMS_GYROID)
            f = fmaf(sz, cX, fmaf(sY, cz, sX * cY));
        else if (p.tpms == GC_TPMS_SCHWARZ_P)
            f = (cX + cY) + cz;
        else
            f = fmaf(sz, fmaf(sX, sY, cX * cY), cz * fmaf(sX, cY, cX * sY));
    }
    float t = p.t0;
    if (p.n_modes > 0) {
        Cx ex[4], ey[4];
        ex[0] = ey[0] = Cx{1.f, 0.f};
        for (int k = 1; k <= 3; ++k) {
            ex[k] = expi((float)k * px);
            ey[k] = expi((float)k * py);
        }
        Cx S[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
        for (int m = 0; m < p.n_modes; ++m) {
            Cx w = cmul(ex[p.mode_k[m][0]], ey[p.mode_k[m][1]]);
            Cx a = cmul(w, Cx{h.phre[m], h.phim[m]});
            int kz = p.mode_k[m][2];
            S[kz].re += a.re;
            S[kz].im += a.im;
        }
        t = p.t0 + S[0].re;
        for (int k = 1; k <= 3; ++k) {
            if (!(h.gmask & (1 << k))) continue;
            Cx ez = expi((float)k * pz);
            t = fmaf(S[k].re, ez.re, t);
            t = fmaf(-S[k].im, ez.im, t);
        }
    }
    float g = p.sheet ? fabsf(f) - t : f;
    if (p.prim != GC_PRIM_NONE) {
        float s;
        if (p.prim == GC_PRIM_SPHERE) {
            float dx = px - p.prim_center[0], dy = py - p.prim_center[1];
            float dz = pz - p.prim_center[2];
            s = sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx))) - p.prim_radius;
        } else {
            float bx = fabsf(px - p.prim_center[0]) - p.prim_half[0];
            float by = fabsf(py - p.prim_center[1]) - p.prim_half[1];
            float bz = fabsf(pz - p.prim_center[2]) - p.prim_half[2];
            float mx = fmaxf(bx, 0.f), my = fmaxf(by, 0.f), mz = fmaxf(bz, 0.f);
            s = sqrtf(fmaf(mz, mz, fmaf(my, my, mx * mx))) + fminf(fmaxf(fmaxf(bx, by), bz), 0.f);
        }
        if (p.op == GC_OP_INTERSECT) g = fmaxf(g, s);
        else if (p.op == GC_OP_UNION) g = fminf(g, s);
        else if (p.op == GC_OP_SUBTRACT) g = fmaxf(g, -s);
    }
    return g;
}

// FP64 closed form of the same definition (independent formulation: direct
// trigonometric evaluation, the four-term diamond, the direct Fourier sum).
double lattice_point_f64(const gc_lattice_params &p, int x, int y, int z) {
    const double hx = 1.0 / (p.nx - 1), hy = 1.0 / (p.ny - 1), hz = 1.0 / (p.nz - 1);
    const double px = x * hx, py = y * hy, pz = z * hz;
    const double cx = 0.5 * (p.nx - 1), cy = 0.5 * (p.ny - 1);
    double u, v;
    const double w = z;
    if (p.bend_radius > 0.f) {
        const double R = p.bend_radius, Cx_ = 0.5, Cy_ = 0.5 - R;
        const double ax = px - Cx_, ay = py - Cy_;
        const double r = std::sqrt(ax * ax + ay * ay), phi = std::atan2(ax, ay);
        u = (Cx_ + R * phi) / hx;
        v = (Cy_ + r) / hy;
    } else {
        const double th = (double)p.rot0 + ((double)p.rot1 - (double)p.rot0) * z / (p.nz - 1);
        u = cx + std::cos(th) * (x - cx) - std::sin(th) * (y - cy);
        v = cy + std::sin(th) * (x - cx) + std::cos(th) * (y - cy);
    }
    const double P = (double)p.period0 + ((double)p.period1 - (double)p.period0) * x / (p.nx - 1);
    const double X = TWO_PI * u / P, Y = TWO_PI * v / P, Z = TWO_PI * w / P;
    double f;
    if (p.tpms == GC_TPMS_GYROID)
        f = std::sin(X) * std::cos(Y) + std::sin(Y) * std::cos(Z) + std::sin(Z) * std::cos(X);
    else if (p.tpms == GC_TPMS_SCHWARZ_P)
        f = std::cos(X) + std::cos(Y) + std::cos(Z);
    else
        f = std::sin(X) * std::sin(Y) * std::sin(Z) + std::sin(X) * std::cos(Y) * std::cos(Z) +
            std::cos(X) * std::sin(Y) * std::cos(Z) + std::cos(X) * std::cos(Y) * std::sin(Z);
    double t = p.t0;
    for (int m = 0; m < p.n_modes; ++m)
        t += (double)p.mode_a[m] *
             std::cos(TWO_PI * (p.mode_k[m][0] * px + p.mode_k[m][1] * py + p.mode_k[m][2] * pz +
                                (double)p.mode_phi[m]));
    double g = p.sheet ? std::fabs(f) - t : f;
    if (p.prim != GC_PRIM_NONE) {
        const double dx = px - p.prim_center[0], dy = py - p.prim_center[1],
                     dz = pz - p.prim_center[2];
        double s;
        if (p.prim == GC_PRIM_SPHERE) {
            s = std::sqrt(dx * dx + dy * dy + dz * dz) - p.prim_radius;
        } else {
            const double bx = std::fabs(dx) - p.prim_half[0], by = std::fabs(dy) - p.prim_half[1],
                         bz = std::fabs(dz) - p.prim_half[2];
            const double mx = std::max(bx, 0.0), my = std::max(by, 0.0), mz = std::max(bz, 0.0);
            s = std::sqrt(mx * mx + my * my + mz * mz) + std::min(std::max(bx, std::max(by, bz)), 0.0);
        }
        if (p.op == GC_OP_INTERSECT) g = std::max(g, s);
        else if (p.op == GC_OP_UNION) g = std::min(g, s);
        else if (p.op == GC_OP_SUBTRACT) g = std::max(g, -s);
    }
    return g;
}

// ---------------------------------------------------------------------------
// Marching cubes
// ---------------------------------------------------------------------------
inline uint64_t pidx(int x, int y, int z, int nx, int ny) {
    return (uint64_t)x + (uint64_t)nx * ((uint64_t)y + (uint64_t)ny * z);
}

inline int cube_at(const float *f, int x, int y, int z, int nx, int ny, float iso) {
    int c = 0;
    for (int k = 0; k < 8; ++k)
        if (f[pidx(x + kCorner[k][0], y + kCorner[k][1], z + kCorner[k][2], nx, ny)] < iso)
            c |= 1 << k;
    return c;
}

// local edge -> owning point offset (dx, dy, dz) and axis of the canonical edge
void edge_owner(int e, int &dx, int &dy, int &dz, int &axis) {
    const int *a = kCorner[kEdge[e][0]], *b = kCorner[kEdge[e][1]];
    dx = std::min(a[0], b[0]);
    dy = std::min(a[1], b[1]);
    dz = std::min(a[2], b[2]);
    axis = a[0] != b[0] ? 0 : (a[1] != b[1] ? 1 : 2);
}

struct RowEdges {
    // per point x: first vertex id (row-relative) and crossing flags per axis
    std::vector<uint32_t> first;
    std::vector<uint8_t> cross;  // bit a set: edge along axis a crosses
};

void row_edges(const uint8_t *s, int y, int z, int nx, int ny, int nz, RowEdges &re) {
    re.first.assign(nx + 1, 0);
    re.cross.assign(nx, 0);
    uint32_t run = 0;
    for (int x = 0; x < nx; ++x) {
        re.first[x] = run;
        uint8_t c = 0;
        const uint8_t s0 = s[pidx(x, y, z, nx, ny)];
        if (x + 1 < nx && s0 != s[pidx(x + 1, y, z, nx, ny)]) c |= 1;
        if (y + 1 < ny && s0 != s[pidx(x, y + 1, z, nx, ny)]) c |= 2;
        if (z + 1 < nz && s0 != s[pidx(x, y, z + 1, nx, ny)]) c |= 4;
        re.cross[x] = c;
        run += (c & 1) + ((c >> 1) & 1) + ((c >> 2) & 1);
    }
    re.first[nx] = run;
}

// Canonical indexed extraction (vertex k = k-th crossing edge in id order
// 3*point+axis; triangle order = cell order, triTable order per cell).
// z_index0: global z index of the field's first plane (positions use global
// indices, so a z-slab reproduces the whole grid's vertices bit for bit)
int extract_indexed(const float *f, int nx, int ny, int nz, float iso, const float *origin,
                    const float *spacing, float *verts, uint64_t vcap, uint32_t *tris,
                    uint64_t tcap, uint64_t *counts, uint64_t *edge_ids, int threads,
                    int z_index0) {
    const uint64_t npts = (uint64_t)nx * ny * nz;
    std::vector<uint8_t> s(npts);
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t i = 0; i < (int64_t)npts; ++i) s[i] = f[i] < iso;
    const int64_t rows = (int64_t)ny * nz;
    std::vector<uint64_t> vb(rows + 1, 0), tb(rows + 1, 0);
#pragma omp parallel num_threads(threads)
    {
        RowEdges re;
#pragma omp for schedule(dynamic, 16)
        for (int64_t r = 0; r < rows; ++r) {
            int y = (int)(r % ny), z = (int)(r / ny);
            row_edges(s.data(), y, z, nx, ny, nz, re);
            vb[r + 1] = re.first[nx];
            uint64_t tcnt = 0;
            if (y + 1 < ny && z + 1 < nz)
                for (int x = 0; x + 1 < nx; ++x) tcnt += T.nverts[cube_at(f, x, y, z, nx, ny, iso)] / 3;
            tb[r + 1] = tcnt;
        }
    }
    for (int64_t r = 0; r < rows; ++r) {
        vb[r + 1] += vb[r];
        tb[r + 1] += tb[r];
    }
    counts[0] = vb[rows];
    counts[1] = tb[rows];
    const float ox = origin[0], oy = origin[1], oz = origin[2];
    const float sx = spacing[0], sy = spacing[1], sz = spacing[2];
#pragma omp parallel num_threads(threads)
    {
        RowEdges R[4];
#pragma omp for schedule(dynamic, 16)
        for (int64_t r = 0; r < rows; ++r) {
            const int y = (int)(r % ny), z = (int)(r / ny);
            row_edges(s.data(), y, z, nx, ny, nz, R[0]);
            // vertices owned by this row
            const float yf = (float)y, zf = (float)(z + z_index0);
            const float py = fmaf(yf, sy, oy), pz = fmaf(zf, sz, oz);
            for (int x = 0; x < nx; ++x) {
                const uint8_t c = R[0].cross[x];
                if (!c) continue;
                uint64_t id = vb[r] + R[0].first[x];
                const float f0 = f[pidx(x, y, z, nx, ny)];
                const float xf = (float)x;
                const float px = fmaf(xf, sx, ox);
                for (int a = 0; a < 3; ++a) {
                    if (!(c & (1 << a))) continue;
                    float f1, v[3] = {px, py, pz};
                    if (a == 0) {
                        f1 = f[pidx(x + 1, y, z, nx, ny)];
                        float p1 = fmaf(xf + 1.0f, sx, ox);
                        float t = (iso - f0) / (f1 - f0);
                        v[0] = fmaf(t, p1 - px, px);
                    } else if (a == 1) {
                        f1 = f[pidx(x, y + 1, z, nx, ny)];
                        float p1 = fmaf(yf + 1.0f, sy, oy);
                        float t = (iso - f0) / (f1 - f0);
                        v[1] = fmaf(t, p1 - py, py);
                    } else {
                        f1 = f[pidx(x, y, z + 1, nx, ny)];
                        float p1 = fmaf(zf + 1.0f, sz, oz);
                        float t = (iso - f0) / (f1 - f0);
                        v[2] = fmaf(t, p1 - pz, pz);
                    }
                    if (id < vcap) {
                        verts[3 * id] = v[0];
                        verts[3 * id + 1] = v[1];
                        verts[3 * id + 2] = v[2];
                        if (edge_ids) edge_ids[id] = 3 * pidx(x, y, z, nx, ny) + a;
                    }
                    ++id;
                }
            }
            if (!(y + 1 < ny && z + 1 < nz)) continue;
            row_edges(s.data(), y + 1, z, nx, ny, nz, R[1]);
            row_edges(s.data(), y, z + 1, nx, ny, nz, R[2]);
            row_edges(s.data(), y + 1, z + 1, nx, ny, nz, R[3]);
            const uint64_t base[4] = {vb[r], vb[r + 1], vb[r + ny], vb[r + ny + 1]};
            uint64_t tid = tb[r];
            for (int x = 0; x + 1 < nx; ++x) {
                const int cube = cube_at(f, x, y, z, nx, ny, iso);
                const int n = T.nverts[cube];
                for (int j = 0; j < n; j += 3) {
                    uint32_t ids[3];
                    for (int k = 0; k < 3; ++k) {
                        int dx, dy, dz, axis;
                        edge_owner(T.tri[cube][j + k], dx, dy, dz, axis);
                        const int row = dy + 2 * dz;
                        const RowEdges &q = R[row];
                        uint64_t id = base[row] + q.first[x + dx];
                        const uint8_t c = q.cross[x + dx];
                        for (int b = 0; b < axis; ++b) id += (c >> b) & 1;
                        ids[k] = (uint32_t)id;
                    }
                    if (tid < tcap) {
                        tris[3 * tid] = ids[0];
                        tris[3 * tid + 1] = ids[1];
                        tris[3 * tid + 2] = ids[2];
                    }
                    ++tid;
                }
            }
        }
    }
    return (counts[0] <= vcap && counts[1] <= tcap) ? GPUCAD_OK : GPUCAD_E_OVERFLOW;
}

// --- reference-shaped drop-in MC (src/MarchingCubes_kernel.cu) --------------
struct Grid {
    uint32_t sx, sy, sz;       // gridSize
    uint32_t shz, mx;          // gridSizeShift.z, gridSizeMask.x
    float vsx, vsy, vsz, cx, cy, cz;
};

inline void grid_pos(uint32_t i, const Grid &g, uint32_t &x, uint32_t &y, uint32_t &z) {
    // src/MarchingCubes_kernel.cu:120-136
    z = i / g.shz;
    uint32_t rem = i % g.shz;
    y = rem / g.mx;
    x = rem % g.mx;
}

inline float sample(const float *v, uint32_t x, uint32_t y, uint32_t z, const Grid &g) {
    // src/MarchingCubes_kernel.cu:88-96
    x = std::min(x, g.sx);
    y = std::min(y, g.sy);
    z = std::min(z, g.sz);
    return v[z * g.sx * g.sy + y * g.sx + x];
}

void corners8(const float *v, uint32_t x, uint32_t y, uint32_t z, const Grid &g, float f[8]) {
    for (int k = 0; k < 8; ++k)
        f[k] = sample(v, x + kCorner[k][0], y + kCorner[k][1], z + kCorner[k][2], g);
}

struct F3 {
    float x, y, z;
};

inline F3 lerp3(F3 a, F3 b, float t) {
    return F3{fmaf(t, b.x - a.x, a.x), fmaf(t, b.y - a.y, a.y), fmaf(t, b.z - a.z, a.z)};
}

inline F3 cross3(F3 a, F3 b) {
    return F3{fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)),
              fmaf(a.x, b.y, -(a.y * b.x))};
}

// one band of vertexInterp2_new / vertexInterp3_new
// (src/MarchingCubes_kernel.cu:3593-3674, 3269-3416); t unchanged when no
// branch applies (the reference leaves it uninitialised; the oracle starts at 0)
bool band(float lo, float hi, F3 &p0, F3 &p1, float f0, float f1, float &t, F3 &snap,
          bool &snapped) {
    if (f1 < f0) {
        std::swap(p0, p1);
        std::swap(f0, f1);
    }
    const float lv[2] = {lo, hi};
    for (int k = 0; k < 2; ++k) {
        const float L = lv[k];
        if (f1 >= L && f0 <= L) {
            if (std::fabs(L - f0) < 0.0005) { snap = p0; snapped = true; return true; }
            if (std::fabs(L - f1) < 0.0005) { snap = p1; snapped = true; return true; }
            if (std::fabs(f1 - f0) < 0.0005) { snap = p0; snapped = true; return true; }
            t = (L - f0) / (f1 - f0);
            return true;
        }
    }
    if (f1 == f0) {
        t = (p0.z == 0.0f) ? 1.0f : 0.0f;
        return true;
    }
    return false;
}

enum { MODE_BAND2 = 0, MODE_BAND1 = 1, MODE_STD = 2 };

F3 edge_vertex(int mode, F3 p0, F3 p1, float id0, float id1, float a0, float a1, float b0,
               float b1, float iso, float lo1, float hi1, float lo2, float hi2) {
    if (mode == MODE_STD) {
        float t = (iso - id0) / (id1 - id0);  // vertexInterp4 on the classified field
        return lerp3(p0, p1, t);
    }
    float t = 0.0f;
    F3 snap;
    bool snapped = false;
    if ((id0 == 1.f && id1 == 0.f) || (id0 == 0.f && id1 == 1.f)) {
        band(lo1, hi1, p0, p1, a0, a1, t, snap, snapped);
        if (snapped) return snap;
    }
    if (mode == MODE_BAND2 && ((id0 == 2.f && id1 == 0.f) || (id1 == 2.f && id0 == 0.f))) {
        band(lo2, hi2, p0, p1, b0, b1, t, snap, snapped);
        if (snapped) return snap;
    }
    return lerp3(p0, p1, t);
}

void generate(int mode, const float *vol, const float *vol1, const float *vol2, const Grid &g,
              float iso, float lo1, float hi1, float lo2, float hi2, const uint32_t *comp,
              const uint32_t *scan, uint32_t active, uint32_t max_verts, float *pos, float *norm) {
    // every active voxel writes only its own output range
#pragma omp parallel for schedule(dynamic, 4096)
    for (uint32_t i = 0; i < active; ++i) {
        const uint32_t voxel = comp[i];
        uint32_t gx, gy, gz;
        grid_pos(voxel, g, gx, gy, gz);
        F3 p{((float)gx - g.cx) * g.vsx, ((float)gy - g.cy) * g.vsy, ((float)gz - g.cz) * g.vsz};
        F3 v[8];
        for (int k = 0; k < 8; ++k)
            v[k] = F3{p.x + (kCorner[k][0] ? g.vsx : 0.f), p.y + (kCorner[k][1] ? g.vsy : 0.f),
                      p.z + (kCorner[k][2] ? g.vsz : 0.f)};
        float fid[8], f1[8] = {0}, f2[8] = {0};
        corners8(vol, gx, gy, gz, g, fid);
        if (mode != MODE_STD) corners8(vol1, gx, gy, gz, g, f1);
        if (mode == MODE_BAND2) corners8(vol2, gx, gy, gz, g, f2);
        int cube = 0;
        for (int k = 0; k < 8; ++k)
            if (fid[k] < iso) cube |= 1 << k;
        const int nv = T.nverts[cube];
        if (mode == MODE_STD && nv == 0) continue;
        F3 vl[12];
        for (int e = 0; e < 12; ++e) {
            int a = kEdge[e][0], b = kEdge[e][1];
            vl[e] = edge_vertex(mode, v[a], v[b], fid[a], fid[b], f1[a], f1[b], f2[a], f2[b], iso,
                                lo1, hi1, lo2, hi2);
        }
        for (int j = 0; j < nv; j += 3) {
            const uint32_t index = scan[voxel] + j;
            F3 a0 = vl[T.tri[cube][j]], a1 = vl[T.tri[cube][j + 1]], a2 = vl[T.tri[cube][j + 2]];
            F3 n = cross3(F3{a1.x - a0.x, a1.y - a0.y, a1.z - a0.z},
                          F3{a2.x - a0.x, a2.y - a0.y, a2.z - a0.z});
            if (index < max_verts - 3u) {
                const F3 vv[3] = {a0, a1, a2};
                for (int k = 0; k < 3; ++k) {
                    float *P = pos + 4 * (uint64_t)(index + k), *N = norm + 4 * (uint64_t)(index + k);
                    P[0] = vv[k].x; P[1] = vv[k].y; P[2] = vv[k].z; P[3] = 1.0f;
                    N[0] = n.x; N[1] = n.y; N[2] = n.z; N[3] = 0.0f;
                }
            }
        }
    }
}

// --- boolean meshing (Isosurface::computeIsosurface) ------------------------
// grid_points (src/MarchingCubes_kernel.h:12-18)
struct GPt {
    int val;
    float t_x, t_y, t_z;
};
enum { B_UNION = 1, B_DIFF = 2, B_INTERSECT = 4, B_FIXED = 8, B_DYNAMIC = 16, B_REGION = 32 };

// corner occupancy of classifyVoxel (src/MarchingCubes_kernel.cu:869-1047):
// F = fixed grid_points .val (as float), D = dynamic float, L = lattice float
int bool_corner(unsigned flags, float F, float D, float L, float iso, float iso1, float iso2) {
    const int band = (int)(L > iso1) & (int)(L < iso2);
    if (flags & B_REGION) return F < iso;
    if (flags & B_UNION) {
        if (flags & B_FIXED) return (int)(D < iso) | band;
        if (flags & B_DYNAMIC) return (int)(F < iso) | band;
        return (int)(F < iso) | (int)(D < iso);
    }
    if (flags & B_DIFF) {
        if (flags & B_FIXED) return (int)(D >= iso) & band;
        if (flags & B_DYNAMIC) return (int)(F < iso) & ((int)(L < iso1) | (int)(L > iso2));
        return (int)(D >= iso) & (int)(F < iso);
    }
    if (flags & B_INTERSECT) {
        if (flags & B_FIXED) return (int)(D < iso) & band;
        if (flags & B_DYNAMIC) return (int)(F < iso) & band;
        return (int)(F < iso) & (int)(D < iso);
    }
    return 0;  // the reference leaves the cube index uninitialised here
}

int bool_cube(unsigned flags, const GPt *fx, const float *dy, const float *lt, uint32_t x,
              uint32_t y, uint32_t z, const Grid &g, float iso, float iso1, float iso2) {
    int cube = 0;
    for (int k = 0; k < 8; ++k) {
        uint32_t px = std::min(x + kCorner[k][0], g.sx), py = std::min(y + kCorner[k][1], g.sy),
                 pz = std::min(z + kCorner[k][2], g.sz);
        uint32_t i = pz * g.sx * g.sy + py * g.sx + px;
        float F = fx ? (float)fx[i].val : 0.f, D = dy ? dy[i] : 0.f, L = lt ? lt[i] : 0.f;
        cube |= bool_corner(flags, F, D, L, iso, iso1, iso2) << k;
    }
    return cube;
}

// vertexInterp_primitive (src/MarchingCubes_kernel.cu:1640-1671)
F3 vi_primitive(float iso, F3 p0, F3 p1, float f0, float f1, float edge_id) {
    float t1 = edge_id, t2 = 0.0f, t = 0.0f;
    if (((f1 >= iso) && (f0 <= iso)) || ((f0 >= iso) && (f1 <= iso))) t2 = (iso - f0) / (f1 - f0);
    if ((t1 > 0) && (t2 > 0)) t = (t1 + t2) * 0.5f;
    else if ((t1 > 0) && (t2 == 0)) t = t1;
    else if ((t2 > 0) && (t1 == 0)) t = t2;
    return lerp3(p0, p1, t);
}

// vertexInterp_primitive_one (src/MarchingCubes_kernel.cu:1674-1726)
F3 vi_primitive_one(float lo, float hi, F3 p0, F3 p1, float f0, float f1, float edge_id) {
    float t1 = edge_id, t2 = 0.0f, t3 = 0.0f, t = 0.0f;
    if (((f1 >= lo) && (f0 <= lo)) || ((f0 >= lo) && (f1 <= lo))) t2 = (lo - f0) / (f1 - f0);
    if ((t1 > 0) && (t2 > 0)) t = (t1 + t2) * 0.5f;
    else if ((t1 > 0) && (t2 == 0)) t = t1;
    else if ((t2 > 0) && (t1 == 0)) t = t2;
    if (((f1 >= hi) && (f0 <= hi)) || ((f0 >= hi) && (f1 <= hi))) t3 = (hi - f0) / (f1 - f0);
    if ((t1 > 0) && (t3 > 0)) t = (t1 + t3) * 0.5f;
    else if ((t1 > 0) && (t3 == 0)) t = t1;
    else if ((t3 > 0) && (t1 == 0)) t = t3;
    return lerp3(p0, p1, t);
}

// vertexInterp_new (src/MarchingCubes_kernel.cu:1729-1784)
F3 vi_new(float iso, float lo, float hi, F3 p0, F3 p1, float f0, float f1, float f2, float f3) {
    float t1 = 0.0f, t2 = 0.0f, t3 = 0.0f, t = 0.0f;
    if (((f0 < iso) && (f1 >= iso)) || ((f1 < iso) && (f0 >= iso))) t1 = (iso - f0) / (f1 - f0);
    if (((f2 < lo) && (f3 >= lo)) || ((f3 < lo) && (f2 >= lo))) t2 = (lo - f2) / (f3 - f2);
    if (((f2 < hi) && (f3 >= hi)) || ((f3 < hi) && (f2 >= hi))) t3 = (hi - f2) / (f3 - f2);
    if ((t1 > 0.0f) && (t2 > 0.0f) && (t3 == 0.0f)) t = (t1 + t2) * 0.5f;
    else if ((t1 > 0.0f) && (t3 > 0.0f) && (t2 == 0.0f)) t = (t1 + t3) * 0.5f;
    else if ((t1 > 0.0f) && (t2 == 0.0f) && (t3 == 0.0f)) t = t1;
    else if ((t2 > 0.0f) && (t1 == 0.0f) && (t3 == 0.0f)) t = t2;
    else if ((t3 > 0.0f) && (t1 == 0.0f) && (t2 == 0.0f)) t = t3;
    return lerp3(p0, p1, t);
}

// generateTriangles_lattice_kernel (src/MarchingCubes_kernel.cu:1865-2199)
void bool_generate(unsigned flags, const GPt *fx, const float *dy, const float *lt, const Grid &g,
                   float iso, float iso1, float iso2, const uint32_t *comp, const uint32_t *scan,
                   uint32_t active, uint32_t max_verts, float *pos, float *norm) {
    // the primitive modes' edges: (start corner, end corner, axis of the start
    // corner's edge parameter); vertexInterp_new walks the plain edge list
    static const int pe[12][3] = {{0, 1, 0}, {1, 2, 1}, {3, 2, 0}, {0, 3, 1}, {4, 5, 0}, {5, 6, 1},
                                  {7, 6, 0}, {4, 7, 1}, {0, 4, 2}, {1, 5, 2}, {2, 6, 2}, {3, 7, 2}};
#pragma omp parallel for schedule(dynamic, 4096)
    for (uint32_t i = 0; i < active; ++i) {
        const uint32_t voxel = comp[i];
        uint32_t gx, gy, gz;
        grid_pos(voxel, g, gx, gy, gz);
        F3 p{((float)gx - g.cx) * g.vsx, ((float)gy - g.cy) * g.vsy, ((float)gz - g.cz) * g.vsz};
        F3 v[8];
        uint32_t ci[8];
        for (int k = 0; k < 8; ++k) {
            v[k] = F3{p.x + (kCorner[k][0] ? g.vsx : 0.f), p.y + (kCorner[k][1] ? g.vsy : 0.f),
                      p.z + (kCorner[k][2] ? g.vsz : 0.f)};
            uint32_t px = std::min(gx + kCorner[k][0], g.sx), py = std::min(gy + kCorner[k][1], g.sy),
                     pz = std::min(gz + kCorner[k][2], g.sz);
            ci[k] = pz * g.sx * g.sy + py * g.sx + px;
        }
        const int cube = bool_cube(flags, fx, dy, lt, gx, gy, gz, g, iso, iso1, iso2);
        const int nv = T.nverts[cube];
        if (nv == 0) continue;
        F3 vl[12];
        for (int e = 0; e < 12; ++e) {
            const int a = pe[e][0], b = pe[e][1], ax = pe[e][2];
            float te = 0.f;
            if (fx) te = ax == 0 ? fx[ci[a]].t_x : (ax == 1 ? fx[ci[a]].t_y : fx[ci[a]].t_z);
            if (flags & B_REGION) {
                vl[e] = lerp3(v[a], v[b], te);
            } else if (flags & B_FIXED) {
                const int na = kEdge[e][0], nb = kEdge[e][1];
                vl[e] = vi_new(iso, iso1, iso2, v[na], v[nb], dy[ci[na]], dy[ci[nb]], lt[ci[na]],
                               lt[ci[nb]]);
            } else if (flags & B_DYNAMIC) {
                vl[e] = vi_primitive_one(iso1, iso2, v[a], v[b], lt[ci[a]], lt[ci[b]], te);
            } else {
                vl[e] = vi_primitive(iso, v[a], v[b], dy[ci[a]], dy[ci[b]], te);
            }
        }
        for (int j = 0; j < nv; j += 3) {
            const uint32_t index = scan[voxel] + j;
            F3 a0 = vl[T.tri[cube][j]], a1 = vl[T.tri[cube][j + 1]], a2 = vl[T.tri[cube][j + 2]];
            // calcNormal(ver[0], ver[2], ver[1])
            F3 n = cross3(F3{a2.x - a0.x, a2.y - a0.y, a2.z - a0.z},
                          F3{a1.x - a0.x, a1.y - a0.y, a1.z - a0.z});
            if (index < max_verts - 3u) {
                const F3 vv[3] = {a0, a1, a2};
                for (int k = 0; k < 3; ++k) {
                    float *P = pos + 4 * (uint64_t)(index + k), *N = norm + 4 * (uint64_t)(index + k);
                    P[0] = vv[k].x; P[1] = vv[k].y; P[2] = vv[k].z; P[3] = 1.0f;
                    N[0] = n.x; N[1] = n.y; N[2] = n.z; N[3] = 0.5f;
                }
            }
        }
    }
}

// --- reference primitives (src/Modelling.cu:244-761), FP32 as restated -----
struct G3 {
    int nx, ny, nz;
    float dx, dy, dz, mx, my, mz;
};
G3 g3(int nx, int ny, int nz, float dx, float dy, float dz) {
    return G3{nx, ny, nz, dx, dy, dz, (float)((nx - 1) / 2.0), (float)((ny - 1) / 2.0),
              (float)((nz - 1) / 2.0)};
}
struct Rows {
    F3 x, y, z;
};
Rows euler_rows(F3 a) {
    float cx = cosf(a.x), sx = sinf(a.x), cy = cosf(a.y), sy = sinf(a.y), cz = cosf(a.z),
          sz = sinf(a.z);
    return Rows{F3{cz * cy, cz * sy * sx - sz * cx, cz * sy * cx + sz * sx},
                F3{sz * cy, sz * sy * sx + cz * cx, sz * sy * cx - cz * sx},
                F3{-1.0f * sy, cy * sx, cy * cx}};
}
inline float dotr(F3 v, F3 r) { return v.x * r.x + v.y * r.y + v.z * r.z; }

template <class Fn>
void for_points(const G3 &g, F3 c, float *out, int threads, Fn fn) {
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int z = 0; z < g.nz; ++z)
        for (int y = 0; y < g.ny; ++y)
            for (int x = 0; x < g.nx; ++x) {
                F3 v{((float)x - g.mx) * g.dx - c.x, ((float)y - g.my) * g.dy - c.y,
                     ((float)z - g.mz) * g.dz - c.z};
                out[pidx(x, y, z, g.nx, g.ny)] = fn(v, x, y, z);
            }
}

}  // namespace

extern "C" {

// ---- math ------------------------------------------------------------------
void orc_cos_rev(const float *r, float *out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = cos_rev(r[i]);
}
void orc_sin_rev(const float *r, float *out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = sin_rev(r[i]);
}
void orc_atan2(const float *y, const float *x, float *out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = atan2_poly(y[i], x[i]);
}

// ---- tables ----------------------------------------------------------------
void orc_tables(uint8_t *tri, uint8_t *nverts) {
    std::memcpy(tri, T.tri, sizeof(T.tri));
    std::memcpy(nverts, T.nverts, sizeof(T.nverts));
}

// ---- lattice field ---------------------------------------------------------
void orc_lattice_field_f32(const gc_lattice_params *p, float *out, int z0, int z1, int threads) {
    const LatticeHost h = lattice_host(*p);
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int z = z0; z < z1; ++z)
        for (int y = 0; y < p->ny; ++y)
            for (int x = 0; x < p->nx; ++x)
                out[(uint64_t)(z - z0) * p->nx * p->ny + (uint64_t)y * p->nx + x] =
                    lattice_point_f32(*p, h, x, y, z);
}

void orc_lattice_field_f64(const gc_lattice_params *p, double *out, int z0, int z1, int threads) {
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int z = z0; z < z1; ++z)
        for (int y = 0; y < p->ny; ++y)
            for (int x = 0; x < p->nx; ++x)
                out[(uint64_t)(z - z0) * p->nx * p->ny + (uint64_t)y * p->nx + x] =
                    lattice_point_f64(*p, x, y, z);
}

// ---- sign words (bit = f < iso), rows padded to 64 points -----------------
void orc_signs(const float *f, int nx, int ny, int nz, float iso, uint64_t *out) {
    const int W = (nx + 63) / 64;
    for (int64_t r = 0; r < (int64_t)ny * nz; ++r)
        for (int w = 0; w < W; ++w) {
            uint64_t m = 0;
            for (int b = 0; b < 64; ++b) {
                int x = 64 * w + b;
                if (x < nx && f[(uint64_t)r * nx + x] < iso) m |= 1ull << b;
            }
            out[r * W + w] = m;
        }
}

// ---- canonical indexed extraction ------------------------------------------
int orc_extract_indexed(const float *f, int nx, int ny, int nz, float iso, const float *origin,
                        const float *spacing, float *verts, uint64_t vcap, uint32_t *tris,
                        uint64_t tcap, uint64_t *counts, uint64_t *edge_ids, int threads,
                        int z_index0) {
    return extract_indexed(f, nx, ny, nz, iso, origin, spacing, verts, vcap, tris, tcap, counts,
                           edge_ids, threads, z_index0);
}

// ---- reference-shaped classify / scan / compact / generate -----------------
// classifyVoxel_new (src/MarchingCubes_kernel.cu:3207-3249) over the layout of
// initMC_two: cells in linear order x + (nx-1)(y + (ny-1) z).
void orc_classify(const float *vol, int nx, int ny, int nz, float iso, uint8_t *cube,
                  uint32_t *verts, uint32_t *occ) {
    const int cx = nx - 1, cy = ny - 1, cz = nz - 1;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < cz; ++z)
        for (int y = 0; y < cy; ++y)
            for (int x = 0; x < cx; ++x) {
                uint64_t i = (uint64_t)x + (uint64_t)cx * (y + (uint64_t)cy * z);
                int c = cube_at(vol, x, y, z, nx, ny, iso);
                if (cube) cube[i] = (uint8_t)c;
                if (verts) verts[i] = T.nverts[c];
                if (occ) occ[i] = T.nverts[c] > 0;
            }
}

// thrust::exclusive_scan over uint (src/MarchingCubes_kernel.cu:3198-3203): wraps mod 2^32
void orc_exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n) {
    uint32_t run = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t v = in[i];
        out[i] = run;
        run += v;
    }
}

// compactVoxels (src/MarchingCubes_kernel.cu:1594-1608)
void orc_compact(const uint32_t *occ, const uint32_t *occ_scan, uint64_t n, uint32_t *comp) {
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; ++i)
        if (occ[i]) comp[occ_scan[i]] = (uint32_t)i;
}

// generateTriangles_lattice_kernel_new / _newone / textbook (mode 0/1/2)
void orc_generate(int mode, const float *vol, const float *vol1, const float *vol2, int nx, int ny,
                  int nz, const float *voxel_size, const float *center, float iso, float lo1,
                  float hi1, float lo2, float hi2, const uint32_t *comp, const uint32_t *scan,
                  uint32_t active, uint32_t max_verts, float *pos, float *norm) {
    Grid g{(uint32_t)nx, (uint32_t)ny, (uint32_t)nz, (uint32_t)((nx - 1) * (ny - 1)),
           (uint32_t)(nx - 1), voxel_size[0], voxel_size[1], voxel_size[2], center[0], center[1],
           center[2]};
    generate(mode, vol, vol1, vol2, g, iso, lo1, hi1, lo2, hi2, comp, scan, active, max_verts, pos,
             norm);
}

// Isosurface::computeIsosurface_lattice / _latticeone orchestration
// (src/Isosurface.cu:401-572) with the textbook variant as mode 2.
int orc_compute_isosurface(int mode, const float *vol, const float *vol1, const float *vol2,
                           int nx, int ny, int nz, const float *voxel_size, const float *center,
                           float iso, float lo1, float hi1, float lo2, float hi2,
                           uint32_t max_verts, uint32_t *verts, uint32_t *verts_scan,
                           uint32_t *occ, uint32_t *occ_scan, uint32_t *comp, float *pos,
                           float *norm, uint32_t *active, uint32_t *total) {
    const uint64_t n = (uint64_t)(nx - 1) * (ny - 1) * (nz - 1);
    orc_classify(vol, nx, ny, nz, iso, nullptr, verts, occ);
    orc_exclusive_scan_u32(occ, occ_scan, n);
    *active = occ[n - 1] + occ_scan[n - 1];
    if (*active == 0) {
        *total = 0;
        return 0;
    }
    orc_compact(occ, occ_scan, n, comp);
    orc_exclusive_scan_u32(verts, verts_scan, n);
    *total = verts[n - 1] + verts_scan[n - 1];
    orc_generate(mode, vol, vol1, vol2, nx, ny, nz, voxel_size, center, iso, lo1, hi1, lo2, hi2,
                 comp, verts_scan, *active, max_verts, pos, norm);
    return 0;
}

// ---- boolean meshing entry points (grid layout as orc_generate: gridSize =
// points, cells = points - 1 per axis)
void orc_bool_classify(unsigned flags, const void *fixed, const float *dyn, const float *lat,
                       int nx, int ny, int nz, float iso, float iso1, float iso2, uint32_t *verts,
                       uint32_t *occ) {
    Grid g{(uint32_t)nx, (uint32_t)ny, (uint32_t)nz, (uint32_t)((nx - 1) * (ny - 1)),
           (uint32_t)(nx - 1), 1, 1, 1, 0, 0, 0};
    const uint64_t n = (uint64_t)(nx - 1) * (ny - 1) * (nz - 1);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t x, y, z;
        grid_pos((uint32_t)i, g, x, y, z);
        const int c = bool_cube(flags, (const GPt *)fixed, dyn, lat, x, y, z, g, iso, iso1, iso2);
        verts[i] = T.nverts[c];
        occ[i] = T.nverts[c] > 0;
    }
}

// classify_copy_Voxel (src/MarchingCubes_kernel.cu:158-446)
void orc_bool_copy(unsigned flags, void *vol_one, const float *vol_two, const float *vol_lat,
                   int nx, int ny, int nz, float iso, float iso1, float iso2) {
    GPt *vo = (GPt *)vol_one;
    const uint64_t n = (uint64_t)nx * ny * nz;
    const bool dynamic = flags & B_DYNAMIC;
    // each point rewrites only its own record and reads neighbours' fields
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n - 1; ++i) {  // n >= 1
        const uint32_t plane = (uint32_t)(nx * ny);
        const uint32_t z = (uint32_t)(i / plane), rem = (uint32_t)(i % plane);
        const uint32_t y = rem / nx, x = rem % nx;
        GPt p = vo[i];
        const float v = dynamic ? 0.f : vol_two[i], vl = dynamic ? vol_lat[i] : 0.f;
        const int band = (int)(vl > iso1) & (int)(vl < iso2);
        if (flags & B_UNION) {
            p.val = (dynamic ? (band | (int)(p.val < iso)) : ((int)(v < iso) | (int)(p.val < iso)))
                        ? -1 : 1;
        } else if (flags & B_DIFF) {
            p.val = (dynamic ? (band & (int)(p.val >= iso)) : ((int)(v >= iso) & (int)(p.val < iso)))
                        ? -1 : 1;
        } else if (flags & B_INTERSECT) {
            p.val = (dynamic ? (band & (int)(p.val < iso)) : ((int)(v < iso) & (int)(p.val < iso)))
                        ? -1 : 1;
        }
        const uint32_t coord[3] = {x, y, z};
        const uint32_t lim[3] = {(uint32_t)nx, (uint32_t)ny, (uint32_t)nz};
        const uint64_t step[3] = {1, (uint64_t)nx, (uint64_t)nx * ny};
        float *tp[3] = {&p.t_x, &p.t_y, &p.t_z};
        for (int ax = 0; ax < 3; ++ax) {
            if (coord[ax] >= lim[ax] - 1) continue;
            if (dynamic) {
                const float q = vol_lat[i + step[ax]];
                if ((q < iso1 && vl >= iso1) || (q >= iso1 && vl < iso1)) {
                    const float t = (iso1 - vl) / (q - vl);
                    *tp[ax] = *tp[ax] > 0 ? (*tp[ax] + t) * 0.5f : t;
                } else if ((q < iso2 && vl >= iso2) || (q >= iso2 && vl < iso2)) {
                    const float t = (iso2 - vl) / (q - vl);
                    *tp[ax] = *tp[ax] > 0 ? (*tp[ax] + t) * 0.5f : t;
                }
            } else {
                const float q = vol_two[i + step[ax]];
                if ((q < iso && v >= iso) || (q >= iso && v < iso)) {
                    const float t = (iso - v) / (q - v);
                    *tp[ax] = *tp[ax] > 0 ? (*tp[ax] + t) * 0.5f : t;
                }
            }
        }
        vo[i] = p;
    }
}

void orc_bool_generate(unsigned flags, const void *fixed, const float *dyn, const float *lat,
                       int nx, int ny, int nz, const float *voxel_size, const float *center,
                       float iso, float iso1, float iso2, const uint32_t *comp,
                       const uint32_t *scan, uint32_t active, uint32_t max_verts, float *pos,
                       float *norm) {
    Grid g{(uint32_t)nx, (uint32_t)ny, (uint32_t)nz, (uint32_t)((nx - 1) * (ny - 1)),
           (uint32_t)(nx - 1), voxel_size[0], voxel_size[1], voxel_size[2], center[0], center[1],
           center[2]};
    bool_generate(flags, (const GPt *)fixed, dyn, lat, g, iso, iso1, iso2, comp, scan, active,
                  max_verts, pos, norm);
}

// ---- Gratings::GPU_buffer_normalise_four (src/lattice_files/Gratings.cu:1089-1134,
// 1394-1495, 1579-1617): a = min(0, min f), b = max(0, max f)
void orc_normalise_four(const float *in, float *ind, float *nrm, int nx, int ny, int nz, float iso1,
                        float iso2) {
    const uint64_t n = (uint64_t)nx * ny * nz;
    float a = 0.f, b = 0.f;
    for (uint64_t i = 0; i < n; ++i) {
        a = std::min(a, in[i]);
        b = std::max(b, in[i]);
    }
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                uint64_t i = pidx(x, y, z, nx, ny);
                float k = (in[i] - a) / (b - a), m;
                if (x == 0 || x == nx - 1 || y == 0 || y == ny - 1 || z == 0 || z == nz - 1) {
                    m = 0.f;
                    k = 0.f;
                } else {
                    m = (k >= iso1 && k <= iso2) ? 1.f : 0.f;
                }
                ind[i] = m;
                nrm[i] = k;
            }
}

// ---- Fft_lattice::create_lattice unit cells (src/lattice_files/Fft_lattice.cu:12-66)
void orc_create_lattice(float *out, int nx, int ny, int nz, int type, int f64) {
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                float xx = (float)((((double)x * 1.0) / (double)(nx - 1) - 0.5) / 0.5);
                float yy = (float)((((double)y * 1.0) / (double)(ny - 1) - 0.5) / 0.5);
                float zz = (float)((((double)z * 1.0) / (double)(nz - 1) - 0.5) / 0.5);
                double aa = 0;
                if (f64) {  // the same formulas in double precision
                    const double X = xx, Y = yy, Z = zz, k = 3.14;
                    if (type == 0)
                        aa = cos(k * X) * sin(k * Y) + cos(k * Y) * sin(k * Z) + cos(k * Z) * sin(k * X);
                    else if (type == 1)
                        aa = cos(k * X) + cos(k * Y) + cos(k * Z);
                    else if (type == 2)
                        aa = 4 * cos(X) * cos(Y) * cos(Z) -
                             (cos(2 * X) * cos(2 * Y) + cos(2 * Y) * cos(2 * Z) + cos(2 * Z) * cos(2 * X));
                    else if (type == 3)
                        aa = 2 * (cos(k * X) * cos(k * Y) + cos(k * Y) * cos(k * Z) + cos(k * Z) * cos(k * X)) -
                             (cos(2 * k * X) + cos(2 * k * Y) + cos(2 * k * Z));
                    else if (type == 4)
                        aa = std::min(std::min(X * X + Y * Y, Y * Y + Z * Z), Z * Z + X * X);
                    else
                        aa = cos(k * X) * cos(k * Y) * cos(k * Z) - sin(k * X) * sin(k * Y) * sin(k * Z);
                } else {
                    float ax = (float)(3.14 * (double)xx), ay = (float)(3.14 * (double)yy),
                          az = (float)(3.14 * (double)zz);
                    if (type == 0)
                        aa = cosf(ax) * sinf(ay) + cosf(ay) * sinf(az) + cosf(az) * sinf(ax);
                    else if (type == 1)
                        aa = cosf(ax) + cosf(ay) + cosf(az);
                    else if (type == 2)
                        aa = 4 * (cosf(xx) * cosf(yy) * cosf(zz)) -
                             (cosf(2 * xx) * cosf(2 * yy) + cosf(2 * yy) * cosf(2 * zz) +
                              cosf(2 * zz) * cosf(2 * xx));
                    else if (type == 3) {
                        float bx = (float)(2 * 3.14 * (double)xx), by = (float)(2 * 3.14 * (double)yy),
                              bz = (float)(2 * 3.14 * (double)zz);
                        aa = 2 * (cosf(ax) * cosf(ay) + cosf(ay) * cosf(az) + cosf(az) * cosf(ax)) -
                             (cosf(bx) + cosf(by) + cosf(bz));
                    } else if (type == 4) {
                        double X2 = (double)xx * xx, Y2 = (double)yy * yy, Z2 = (double)zz * zz;
                        aa = std::min(std::min((double)(xx * xx) + Y2, Y2 + Z2), Z2 + X2);
                    } else {
                        double t1 = cos(3.14 * (double)xx) * (double)cosf(ay) * (double)cosf(az);
                        aa = t1 - (double)(sinf(ax) * sinf(ay) * sinf(az));
                    }
                }
                out[pidx(x, y, z, nx, ny)] = (float)aa;
            }
}

// ---- Modelling primitives (src/Modelling.cu:244-761) -----------------------
// prim: 0 sphere, 1 cuboid, 2 cuboid shell, 3 pyramid frustum, 4 torus,
// 5 cone, 6 cone frustum, 7 distance from line.  prm holds the scalar
// arguments in declaration order of the corresponding gpucad_* entry point
// after (center, angles|axis): see tests for the packing.
void orc_primitive(int prim, float *out, int nx, int ny, int nz, float dx, float dy, float dz,
                   const float *center, const float *angles, const float *prm, int flag,
                   int threads) {
    const G3 g = g3(nx, ny, nz, dx, dy, dz);
    const F3 c{center[0], center[1], center[2]};
    const Rows R = euler_rows(F3{angles[0], angles[1], angles[2]});
    switch (prim) {
        case 0: {
            const float radius = prm[0], t_diff = (float)(prm[1] / 2.0);
            for_points(g, c, out, threads, [&](F3 v, int, int, int) {
                float r2 = v.x * v.x + v.y * v.y + v.z * v.z;
                if (flag) {
                    float ri = radius - t_diff, ro = radius + t_diff;
                    return (float)std::max((double)(r2 - ri * ri) * -1.0, (double)(r2 - ro * ro));
                }
                return r2 - radius * radius;
            });
            break;
        }
        case 1: {
            const float xw = (float)(prm[0] / 2.0), yw = (float)(prm[1] / 2.0), zw = (float)(prm[2] / 2.0);
            for_points(g, c, out, threads, [&](F3 v, int, int, int) {
                float f1 = fabsf(dotr(v, R.x)) - xw, f2 = fabsf(dotr(v, R.y)) - yw,
                      f3 = fabsf(dotr(v, R.z)) - zw;
                return fmaxf(fmaxf(f1, f2), f3);
            });
            break;
        }
        case 2: {
            const float xw = (float)(prm[0] / 2.0), yw = (float)(prm[1] / 2.0), zw = (float)(prm[2] / 2.0),
                        th = prm[3];
            for_points(g, c, out, threads, [&](F3 v, int, int, int) {
                float f1 = dotr(v, R.x), f2 = dotr(v, R.y);
                float f11 = fabsf(f1) - xw, f12 = fabsf(f1) - (xw - th);
                float f21 = fabsf(f2) - yw, f22 = fabsf(f2) - (yw - th);
                float f3 = fabsf(dotr(v, R.z)) - zw;
                double inner = (double)fmaxf(f12, f22) * -1.0;
                return (float)std::max(std::max((double)fmaxf(f11, f21), inner), (double)f3);
            });
            break;
        }
        case 3: {
            const float xwb = (float)(prm[0] / 2.0), xwt = (float)(prm[1] / 2.0), yh = prm[2],
                        zwb = (float)(prm[3] / 2.0), zwt = (float)(prm[4] / 2.0);
            for_points(g, c, out, threads, [&](F3 v, int, int, int) {
                float f1 = dotr(v, R.x), f2 = dotr(v, R.y), f3 = dotr(v, R.z);
                float ratio = (yh - f2) / yh;
                float xw = ratio * (xwb - xwt) + xwt;
                f1 = fabsf(f1) - xw;
                f2 = fabsf(f2 - yh / 2) - yh / 2;
                float zw = ratio * (zwb - zwt) + zwt;
                f3 = fabsf(f3) - zw;
                return fmaxf(fmaxf(f1, f2), f3);
            });
            break;
        }
        case 4: {
            const float tr = prm[0], cr = prm[1];
            for_points(g, c, out, threads, [&](F3 v, int, int, int) {
                float vx = dotr(v, R.x), vy = dotr(v, R.y), vz = dotr(v, R.z);
                float side = tr - sqrtf(vx * vx + vy * vy);
                return vz * vz - cr * cr + side * side;
            });
            break;
        }
        case 5: {
            const float br = prm[0], h = prm[1];
            for_points(g, c, out, threads, [&](F3 v, int, int, int) {
                float q = h / br, f2 = q * q;
                float vx = dotr(v, R.x), vy = dotr(v, R.y), vz = dotr(v, R.z);
                float gg = (float)((double)vy - (double)h / 2.0);
                float hh = (float)std::max(((double)gg - (double)h / 2.0) * 100.0,
                                           ((double)gg + (double)h / 2.0) * 100.0 * -1.0);
                float d = vy - h;
                float f1 = ((vx * vx + vz * vz) * f2) - d * d;
                return fmaxf(f1, hh);
            });
            break;
        }
        case 6: {
            const float tr = prm[0], br = prm[1], h = prm[2];
            for_points(g, c, out, threads, [&](F3 v, int, int, int) {
                float vx = dotr(v, R.x), vy = dotr(v, R.y), vz = dotr(v, R.z);
                float r_diff = ((h - vy) / h) * (br - tr);
                float gg = (float)((double)vy - (double)h / 2.0);
                float hh = (float)std::max(((double)gg - (double)h / 2.0) * 100.0,
                                           ((double)gg + (double)h / 2.0) * 100.0 * -1.0);
                double rr = (double)(r_diff + tr);
                float f1 = (float)((double)(vx * vx + vz * vz) - rr * rr);
                return fmaxf(f1, hh);
            });
            break;
        }
        case 7: {
            // angles holds the axis for the line primitive
            const float radius = prm[0], t_diff = (float)(prm[1] / 2.0), t_ax = (float)(prm[2] / 2.0);
            float mag = sqrtf(angles[0] * angles[0] + angles[1] * angles[1] + angles[2] * angles[2]);
            const F3 ax{angles[0] / mag, angles[1] / mag, angles[2] / mag};
#pragma omp parallel for num_threads(threads) schedule(static)
            for (int z = 0; z < nz; ++z)
                for (int y = 0; y < ny; ++y)
                    for (int x = 0; x < nx; ++x) {
                        float x1 = ((float)x - g.mx) * dx, y1 = ((float)y - g.my) * dy,
                              z1 = ((float)z - g.mz) * dz;
                        F3 end{ax.x + c.x, ax.y + c.y, ax.z + c.z};
                        F3 w1{x1 - c.x, y1 - c.y, z1 - c.z}, w2{x1 - end.x, y1 - end.y, z1 - end.z},
                            w3{end.x - c.x, end.y - c.y, end.z - c.z};
                        F3 d{w1.y * w2.z - w1.z * w2.y, w1.z * w2.x - w1.x * w2.z,
                             w1.x * w2.y - w1.y * w2.x};
                        float e = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
                        float dis = sqrtf(w3.x * w3.x + w3.y * w3.y + w3.z * w3.z);
                        float fv = e / dis;
                        float gg = (x1 - c.x) * ax.x + (y1 - c.y) * ax.y + (z1 - c.z) * ax.z;
                        float f1 = fmaxf(gg - t_ax, (gg + t_ax) * -1.0f);
                        float f2 = flag ? fmaxf(fv - (radius + t_diff), (fv - (radius - t_diff)) * -1.0f)
                                        : fv - radius;
                        out[pidx(x, y, z, nx, ny)] = fmaxf(f1, f2);
                    }
            break;
        }
        default:
            break;
    }
}

// ---- File_output::file_write_obj (src/File_output.cu:5-81), std::map form --
int orc_file_write_obj(const float *pos4, uint32_t total, const char *filename) {
    std::ofstream m(filename, std::ios::out);
    if (!m) return 1;
    m << "##Sample latttice new Obj \n";
    m << "o Solid \n";
    std::map<std::vector<float>, int> vmap;
    std::vector<uint32_t> faces;
    int index = 0;
    for (uint32_t i = 0; i < total; ++i) {
        float vx = int(pos4[4 * i] * 1000) * 0.001;
        float vy = int(pos4[4 * i + 1] * 1000) * 0.001;
        float vz = int(pos4[4 * i + 2] * 1000) * 0.001;
        std::vector<float> key{vx, vy, vz};
        if (vmap.count(key) == 0) {
            ++index;
            vmap[key] = index;
            faces.push_back(index);
            m << "v " << vx << " " << vy << " " << vz << "\n";
        } else {
            faces.push_back(vmap[key]);
        }
    }
    m << "\n";
    std::map<std::vector<uint32_t>, int> fmap;
    m << "\n";
    for (size_t i = 0; i + 2 < faces.size(); i += 3) {
        if (faces[i] != faces[i + 1] && faces[i] != faces[i + 2] && faces[i + 1] != faces[i + 2]) {
            std::vector<uint32_t> key{faces[i], faces[i + 1], faces[i + 2]};
            if (fmap.count(key) == 0) {
                fmap[key] = 1;
                m << " f  " << faces[i] << " " << faces[i + 2] << " " << faces[i + 1] << "\n";
            }
        }
    }
    return 0;
}

int orc_max_threads(void) { return omp_get_max_threads(); }

// ===========================================================================
// Spatially variant lattice synthesis (src/lattice_files/Gratings.cu,
// src/main.cu:3578-3711, 3904-4037).  The binary32 sequences match the
// kernels' (same radians-through-revolutions sin/cos, same atan2, the
// reference's double subexpressions in double); CG sums its dot products in
// double in index order.
// ===========================================================================
static const float kInv2PiF = (float)0.15915494309189533577;
static const double kTwoPiD = 6.283185307179586476925286766559;
static inline float cos_rad(float t) { return cos_rev(t * kInv2PiF); }
static inline float sin_rad(float t) { return sin_rev(t * kInv2PiF); }

// ---- set_period_kernel / set_theta_kernel (Gratings.cu:775-860)
void orc_svl_period_angle(float *out, int nx, int ny, int nz, float dx, float dy, float dz,
                          float mx, float my, float mz, int axis, int angle) {
#pragma omp parallel for collapse(2)
    for (int zz = 0; zz < nz; ++zz)
        for (int yy = 0; yy < ny; ++yy)
            for (int xx = 0; xx < nx; ++xx) {
                const float x = ((float)xx - mx) * dx, y = ((float)yy - my) * dy,
                            z = ((float)zz - mz) * dz;
                float v;
                if (angle) {
                    v = axis == 'y' ? atan2_poly(z + 1.0f, x + 1.0f) : atan2_poly(y + 1.0f, x + 1.0f);
                } else {
                    float u, w;
                    if (axis == 'z') { u = x + 1.0f; w = y + 1.0f; }
                    else if (axis == 'y') { u = x + 1.0f; w = z + 1.0f; }
                    else { u = z + 1.0f; w = y + 1.0f; }
                    v = sqrtf(u * u + w * w);
                }
                out[pidx(xx, yy, zz, nx, ny)] = v;
            }
}

// ---- GPU_buffer_normalise_three / _buffer (Gratings.cu:1500-1575, 1439-1495):
// min/max seeded with 0, then a1 + b1 k or k
void orc_svl_rescale(const float *in, float *out, uint64_t n, float a1, float b1, int affine) {
    float a = 0.f, b = 0.f;
    for (uint64_t i = 0; i < n; ++i) {
        a = std::min(a, in[i]);
        b = std::max(b, in[i]);
    }
    for (uint64_t i = 0; i < n; ++i) {
        const float k = (in[i] - a) / (b - a);
        out[i] = affine ? a1 + b1 * k : k;
    }
}

// ---- finding_phi_kernel (Gratings.cu:100-417)
static void axis_weights(int c, int n, float *w1, float *w2, int *c1, int *c2) {
    if (c == 0) { *w1 = -1.0f; *w2 = -0.5f; *c1 = c; *c2 = c + 1; }
    else if (c == 1) { *w1 = 1.0f; *w2 = -0.5f; *c1 = c - 1; *c2 = c + 1; }
    else if (c == n - 2) { *w1 = 0.5f; *w2 = -1.0f; *c1 = c - 1; *c2 = c + 1; }
    else if (c == n - 1) { *w1 = 0.5f; *w2 = 1.0f; *c1 = c - 1; *c2 = c; }
    else { *w1 = 0.5f; *w2 = -0.5f; *c1 = c - 1; *c2 = c + 1; }
}
static inline float wave(float per, float w) { return (float)((kTwoPiD / (double)per) * (double)w); }
static inline float sine_angle(float lcon, float lcon_1, float u) {
    const float arg = (float)((6.28 * (double)lcon_1) * (double)u);
    return lcon * sin_rad(arg);
}

void orc_svl_finding_phi(float *phi, const float *period, int nx, int ny, int nz, int i, int j,
                         int k, float dx, float dy, float dz, int type, int uniform_type,
                         float const_period, float x_period, float y_period, float z_period,
                         float lcon, float lcon_1, int sinewave_zaxis) {
#pragma omp parallel for collapse(2)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                float a1, a2, b1, b2, c1, c2;
                int x1, x2, y1, y2, z1, z2;
                axis_weights(x, nx, &a1, &a2, &x1, &x2);
                axis_weights(y, ny, &b1, &b2, &y1, &y2);
                axis_weights(z, nz, &c1, &c2, &z1, &z2);
                float th1 = 0, th2 = 0, th3 = 0, th4 = 0, th5 = 0, th6 = 0, th7 = 0, th8 = 0;
                if (type == 'r' || type == 'b') {
                    const float mx = type == 'r' ? (float)(nx + 1) / 2.0f : 0.f;
                    const float my = type == 'r' ? (float)(ny + 1) / 2.0f : 0.f;
                    const float yy = ((float)(y + 1) - my) * dx;
                    th1 = atan2_poly(yy, ((float)(x1 + 1) - mx) * dx);
                    th2 = atan2_poly(yy, ((float)(x2 + 1) - mx) * dx);
                    const float xx = ((float)(x + 1) - mx) * dx;
                    th3 = atan2_poly(((float)(y1 + 1) - my) * dx, xx);
                    th4 = atan2_poly(((float)(y2 + 1) - my) * dx, xx);
                } else if (type == 's') {
                    th1 = sine_angle(lcon, lcon_1, (float)(x1 + 1) * dx);
                    th5 = sine_angle(lcon, lcon_1, (float)(z1 + 1) * dz);
                    th2 = sine_angle(lcon, lcon_1, (float)(x2 + 1) * dx);
                    th6 = sine_angle(lcon, lcon_1, (float)(z2 + 1) * dz);
                    th3 = sine_angle(lcon, lcon_1, (float)(x + 1) * dx);
                    th7 = sine_angle(lcon, lcon_1, (float)(z + 1) * dz);
                    th4 = th3;
                    th8 = th7;
                }
                float p1, p2, p3, p4, p5, p6;
                if (uniform_type == 0) {
                    p1 = p2 = p3 = p4 = p5 = p6 = const_period;
                } else if (uniform_type == 1) {
                    p1 = p2 = x_period;
                    p3 = p4 = y_period;
                    p5 = p6 = z_period;
                } else {
                    p1 = period[pidx(x1, y, z, nx, ny)];
                    p2 = period[pidx(x2, y, z, nx, ny)];
                    p3 = period[pidx(x, y1, z, nx, ny)];
                    p4 = period[pidx(x, y2, z, nx, ny)];
                    p5 = period[pidx(x, y, z1, nx, ny)];
                    p6 = period[pidx(x, y, z2, nx, ny)];
                }
                const float fi = (float)i, fj = (float)j, fk = (float)k;
                float kx1 = wave(p1, fi * cos_rad(th1) - fj * sin_rad(th1));
                float kx2 = wave(p2, fi * cos_rad(th2) - fj * sin_rad(th2));
                float ky1 = wave(p3, fi * sin_rad(th3) + fj * cos_rad(th3));
                float ky2 = wave(p4, fi * sin_rad(th4) + fj * cos_rad(th4));
                float kz1 = wave(p5, fk), kz2 = wave(p6, fk);
                float ph = (fmaf(a1, kx1, a2 * kx2) + fmaf(b1, ky1, b2 * ky2)) +
                           fmaf(c1, kz1, c2 * kz2);
                if (type == 's' && sinewave_zaxis) {
                    kz1 = wave(p5, fj * sin_rad(th5) + fk * cos_rad(th5));
                    kz2 = wave(p6, fj * sin_rad(th6) + fk * cos_rad(th6));
                    ky1 = wave(p3, fj * cos_rad(th7) - fk * sin_rad(th7));
                    ky2 = wave(p4, fj * cos_rad(th8) - fk * sin_rad(th8));
                    kx1 = wave(p1, fi);
                    kx2 = wave(p2, fi);
                    ph += (fmaf(a1, kx1, a2 * kx2) + fmaf(b1, ky1, b2 * ky2)) +
                          fmaf(c1, kz1, c2 * kz2);
                }
                phi[pidx(x, y, z, nx, ny)] = ph;
            }
}

// ---- GPUMatvec_lattice_kernel (Gratings.cu:420-597)
static inline void axis_terms(const float *d, int c, int n, int64_t t, int64_t s, float *w,
                              float *t2, float *t3) {
    if (c == 0) { *w = 1.25f; *t2 = -d[t + s]; *t3 = -0.25f * d[t + 2 * s]; }
    else if (c == 1) { *w = 1.25f; *t2 = -d[t - s]; *t3 = -0.25f * d[t + 2 * s]; }
    else if (c == n - 2) { *w = 1.25f; *t2 = -0.25f * d[t - 2 * s]; *t3 = -d[t + s]; }
    else if (c == n - 1) { *w = 1.25f; *t2 = -0.25f * d[t - 2 * s]; *t3 = -d[t - s]; }
    else { *w = 0.5f; *t2 = -0.25f * d[t - 2 * s]; *t3 = -0.25f * d[t + 2 * s]; }
}

void orc_svl_matvec(const float *d, float *q, int nx, int ny, int nz) {
#pragma omp parallel for collapse(2)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const int64_t t = (int64_t)pidx(x, y, z, nx, ny);
                float wx, x2, x3, wy, y2, y3, wz, z2, z3;
                axis_terms(d, x, nx, t, 1, &wx, &x2, &x3);
                axis_terms(d, y, ny, t, nx, &wy, &y2, &y3);
                axis_terms(d, z, nz, t, (int64_t)nx * ny, &wz, &z2, &z3);
                const float w = (wx + wy) + wz;
                float a = fmaf(w, d[t], x2);
                a = a + x3;
                a = a + y2;
                a = a + y3;
                a = a + z2;
                a = a + z3;
                q[t] = a;
            }
}

static double dot_d(const float *a, const float *b, uint64_t n) {
    double s = 0.0;
    for (uint64_t i = 0; i < n; ++i) s += (double)a[i] * (double)b[i];
    return s;
}

// ---- GPUCG_lattice (Gratings.cu:875-974): x = 0, r = d = b; while
// (iCounter < iter && r.r > end_res^2) { q = A d; alpha = r.r / d.q;
// x += alpha d; r -= alpha q; beta = r.r_new / r.r_old; d = r + beta d }
void orc_svl_cg(float *b, int nx, int ny, int nz, int iter, float end_res, int *final_iter,
                float *final_res) {
    const uint64_t n = (uint64_t)nx * ny * nz;
    std::vector<float> x(n, 0.f), r(b, b + n), d(b, b + n), q(n);
    float delta = (float)dot_d(r.data(), d.data(), n);
    const float term = end_res * end_res;
    int it = 1;
    while (it < iter && delta > term) {
        orc_svl_matvec(d.data(), q.data(), nx, ny, nz);
        const float alpha = delta / (float)dot_d(d.data(), q.data(), n);
        for (uint64_t i = 0; i < n; ++i) {
            x[i] = fmaf(alpha, d[i], x[i]);
            r[i] = fmaf(-alpha, q[i], r[i]);
        }
        const float old = delta;
        delta = (float)dot_d(r.data(), r.data(), n);
        const float beta = delta / old;
        for (uint64_t i = 0; i < n; ++i) d[i] = fmaf(beta, d[i], r[i]);
        ++it;
    }
    std::copy(x.begin(), x.end(), b);
    *final_iter = it;
    *final_res = sqrtf(delta);
}

// ---- refine / grating texture sampling (Gratings.cu:653-722,
// Interpolations.cu:57-107): texel coordinate f*h, weights quantised to
// 1/256, clamped, lerps x then y then z
struct OTap {
    int64_t i0, i1;
    float w;
};
static OTap otap(int f, float h, int n) {
    const float u = (float)f * h;
    const float fl = floorf(u);
    float w = u - fl;
    w = rintf(w * 256.0f) * (1.0f / 256.0f);
    int64_t i0 = (int64_t)fl;
    if (w >= 1.0f) {
        i0 += 1;
        w = 0.0f;
    }
    OTap t;
    t.i0 = std::min<int64_t>(std::max<int64_t>(i0, 0), n - 1);
    t.i1 = std::min<int64_t>(std::max<int64_t>(i0 + 1, 0), n - 1);
    t.w = w;
    return t;
}
static inline float olerp(float a, float b, float w) { return fmaf(w, b, (1.0f - w) * a); }
static float osample(const float *c, const OTap &tx, const OTap &ty, const OTap &tz, int64_t row,
                     int64_t plane) {
    const int64_t z0 = tz.i0 * plane, z1 = tz.i1 * plane, y0 = ty.i0 * row, y1 = ty.i1 * row;
    const float c00 = olerp(c[z0 + y0 + tx.i0], c[z0 + y0 + tx.i1], tx.w);
    const float c10 = olerp(c[z0 + y1 + tx.i0], c[z0 + y1 + tx.i1], tx.w);
    const float c01 = olerp(c[z1 + y0 + tx.i0], c[z1 + y0 + tx.i1], tx.w);
    const float c11 = olerp(c[z1 + y1 + tx.i0], c[z1 + y1 + tx.i1], tx.w);
    return olerp(olerp(c00, c10, ty.w), olerp(c01, c11, ty.w), tz.w);
}

// mode 0: out[f] = sample; mode 1: out[2f], out[2f+1] = cos, sin of the sample
void orc_svl_upsample(const float *c, int nx, int ny, int nz, float *out, int fx, int fy, int fz,
                      float hx, float hy, float hz, int mode) {
#pragma omp parallel for collapse(2)
    for (int z = 0; z < fz; ++z)
        for (int y = 0; y < fy; ++y)
            for (int x = 0; x < fx; ++x) {
                const float v = osample(c, otap(x, hx, nx), otap(y, hy, ny), otap(z, hz, nz), nx,
                                        (int64_t)nx * ny);
                const uint64_t t = pidx(x, y, z, fx, fy);
                if (mode == 0) out[t] = v;
                else {
                    out[2 * t] = cos_rad(v);
                    out[2 * t + 1] = sin_rad(v);
                }
            }
}

// ---- svl_kernel (Gratings.cu:724-752)
void orc_svl_accumulate(float *svl, const float *grating, uint64_t n, int index,
                        const float *coeffs) {
    const float cr = coeffs[2 * index], ci = coeffs[2 * index + 1];
    for (uint64_t t = 0; t < n; ++t)
        svl[t] = svl[t] + fmaf(grating[2 * t], cr, -(grating[2 * t + 1] * ci));
}

// ---- the synthesis loop of spatial_lattice_run as one field: the three unit
// phases (each solved alone, to end_res / (3 range)), then per fine point the
// sum over the first `harmonics` coefficients of Re(c e^{i phi}) with phi =
// i Phi_x + j Phi_y + k Phi_z evaluated directly
void orc_svl_field(int nx, int ny, int nz, float dx, float dy, float dz, int type,
                   int uniform_type, float const_period, float x_period, float y_period,
                   float z_period, float lcon, float lcon_1, int sinewave_zaxis,
                   const float *period, const float *coeffs, int range, int harmonics,
                   int cg_max_iter, float cg_end_res, int fx, int fy, int fz, float hx, float hy,
                   float hz, float *svl, int *iters, float *res) {
    const uint64_t n = (uint64_t)nx * ny * nz;
    std::vector<float> ph(3 * n);
    const int unit[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const float tol = cg_end_res / (float)(3 * std::max(1, range));
    for (int a = 0; a < 3; ++a) {
        orc_svl_finding_phi(ph.data() + a * n, period, nx, ny, nz, unit[a][0], unit[a][1],
                            unit[a][2], dx, dy, dz, type, uniform_type, const_period, x_period,
                            y_period, z_period, lcon, lcon_1, sinewave_zaxis);
        orc_svl_cg(ph.data() + a * n, nx, ny, nz, cg_max_iter, tol, iters + a, res + a);
    }
    const int side = 2 * range + 1;
#pragma omp parallel for collapse(2)
    for (int z = 0; z < fz; ++z)
        for (int y = 0; y < fy; ++y)
            for (int x = 0; x < fx; ++x) {
                const OTap tx = otap(x, hx, nx), ty = otap(y, hy, ny), tz = otap(z, hz, nz);
                const int64_t row = nx, plane = (int64_t)nx * ny;
                const float px = osample(ph.data(), tx, ty, tz, row, plane);
                const float py = osample(ph.data() + n, tx, ty, tz, row, plane);
                const float pz = osample(ph.data() + 2 * n, tx, ty, tz, row, plane);
                float acc = 0.f;
                for (int h = 0; h < harmonics; ++h) {
                    const int i = h % side - range, j = (h / side) % side - range,
                              k = h / (side * side) - range;
                    const float phi = ((float)i * px + (float)j * py) + (float)k * pz;
                    acc = acc + fmaf(cos_rad(phi), coeffs[2 * h], -(sin_rad(phi) * coeffs[2 * h + 1]));
                }
                svl[pidx(x, y, z, fx, fy)] = acc;
            }
}

// ---- unit_lattice's coefficients (src/main.cu:3578-3711 with Fft_lattice.cu
// fft_scalar): (2r+1)^3 DFT coefficients / N, loop order k, j, i, in double
void orc_unit_spectrum(const float *cell, int nx, int ny, int nz, int range, double *out) {
    const int side = 2 * range + 1;
#pragma omp parallel for
    for (int h = 0; h < side * side * side; ++h) {
        const int i = h % side - range, j = (h / side) % side - range, k = h / (side * side) - range;
        double re = 0.0, im = 0.0;
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const double ang = -kTwoPiD * ((double)i * x / nx + (double)j * y / ny +
                                                   (double)k * z / nz);
                    const double f = cell[pidx(x, y, z, nx, ny)];
                    re += f * cos(ang);
                    im += f * sin(ang);
                }
        const double N = (double)nx * ny * nz;
        out[2 * h] = re / N;
        out[2 * h + 1] = im / N;
    }
}

// ---- fft_fill_kernel (Fft_lattice.cu:150-225)
void orc_fft_fill(const float *half, float *full, int nx, int ny, int nz) {
    const int nx2 = nx / 2 + 1;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx2; ++x) {
                const uint64_t t = (uint64_t)x + (uint64_t)y * nx2 + (uint64_t)z * nx2 * ny;
                if (t == 0) {
                    full[0] = half[0];
                    full[1] = half[1];
                    continue;
                }
                const int e = x == 0 ? 0 : nx - x, f = y == 0 ? 0 : nx - y, g = z == 0 ? 0 : nx - z;
                const uint64_t d1 = pidx(x, y, z, nx, ny), d2 = pidx(e, f, g, nx, ny);
                full[2 * d1] = half[2 * t];
                full[2 * d1 + 1] = half[2 * t + 1];
                full[2 * d2] = half[2 * t];
                full[2 * d2 + 1] = -half[2 * t + 1];
            }
}

}  // extern "C"
