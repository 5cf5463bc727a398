// Spatially variant lattice (SVL) synthesis: the step before the meshing path
// in the reference application (src/main.cu:3904-4037, spatial_lattice_run).
//
// The reference builds the lattice field as a sum of gratings: for each
// harmonic (i, j, k) of the unit cell's Fourier series it derives a phase
// field phi on a coarse grid by solving D^T D phi = D^T K with conjugate
// gradients (K: the spatially varying grating vector), samples e^{i phi} on a
// grid twice as fine and accumulates Re(c_ijk e^{i phi}).  This file holds
//   * the reference-shaped steps (period/angle data, normalise, finding_phi,
//     the operator, CG, refine/grating, svl accumulation, FFT fill/scale), each
//     with the reference's arithmetic, and
//   * the native form, gpucad_svl_field: the right-hand side of finding_phi
//     is linear in (i, j, k), so three solves (unit wave numbers) replace one
//     solve per harmonic; the three are solved together (one CG over three
//     right-hand sides, all scalars on the device, no per-iteration host
//     round trip) and one fused pass upsamples the three phase fields and
//     sums every harmonic per fine point as products of e^{i Phi_x},
//     e^{i Phi_y}, e^{i Phi_z}.
//   * the unit-cell Fourier coefficients by direct DFT (FP64 accumulation)
//     for exactly the (2r+1)^3 coefficients the reference keeps.
#include <cmath>
#include <vector>

#include "gc_common.hpp"
#include "gc_math.hpp"

namespace gc {

int enqueue_minmax0(const float *in, uint64_t n, float2 *part, hipStream_t st);

namespace {

constexpr float kInv2Pi = (float)0.15915494309189533577;  // 1/(2 pi), binary32
constexpr double kTwoPi = 6.283185307179586476925286766559;

// cos/sin of an angle in radians through the revolution polynomials
__device__ __forceinline__ float cos_rad(float t) { return cos_rev(t * kInv2Pi); }
__device__ __forceinline__ float sin_rad(float t) { return sin_rev(t * kInv2Pi); }

// (x, y, z) of linear point t; grids have fewer than 2^32 points (checked
// at the entry points), so the divisions are 32-bit
__device__ __forceinline__ void point_xyz(uint64_t t, int nx, int ny, int *x, int *y, int *z) {
    const uint32_t tt = (uint32_t)t, plane = (uint32_t)nx * (uint32_t)ny;
    const uint32_t zz = tt / plane;
    const uint32_t r = tt - zz * plane;
    const uint32_t yy = r / (uint32_t)nx;
    *z = (int)zz;
    *y = (int)yy;
    *x = (int)(r - yy * (uint32_t)nx);
}

// n / d for n < 2^32 through a double reciprocal and one correction step
struct Div32 {
    uint32_t d;
    double inv;
};
inline Div32 make_div32(uint32_t d) { return Div32{d, 1.0 / (double)d}; }
__device__ __forceinline__ uint32_t div32(uint32_t n, const Div32 &f) {
    uint32_t q = (uint32_t)((double)n * f.inv);
    const int64_t r = (int64_t)n - (int64_t)q * f.d;
    if (r < 0) --q;
    else if (r >= (int64_t)f.d) ++q;
    return q;
}
__device__ __forceinline__ void point_xyz(uint32_t t, const Div32 &dnx, const Div32 &dplane,
                                          int *x, int *y, int *z) {
    const uint32_t zz = div32(t, dplane);
    const uint32_t r = t - zz * dplane.d;
    const uint32_t yy = div32(r, dnx);
    *z = (int)zz;
    *y = (int)yy;
    *x = (int)(r - yy * dnx.d);
}

// ---------------------------------------------------------------------------
// period and angle data (src/lattice_files/Gratings.cu:775-860)
// ---------------------------------------------------------------------------
__global__ void period_angle_kernel(float *out, int nx, int ny, int nz, float dx, float dy,
                                    float dz, float mx, float my, float mz, int axis, int angle) {
    const uint64_t n = (uint64_t)nx * ny * nz;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        int xx, yy, zz;
        point_xyz(t, nx, ny, &xx, &yy, &zz);
        const float x = ((float)xx - mx) * dx;
        const float y = ((float)yy - my) * dy;
        const float z = ((float)zz - mz) * dz;
        float v;
        if (angle) {
            v = axis == 'y' ? atan2_poly(z + 1.0f, x + 1.0f) : atan2_poly(y + 1.0f, x + 1.0f);
        } else {
            float u, w;
            if (axis == 'z') {
                u = x + 1.0f;
                w = y + 1.0f;
            } else if (axis == 'y') {
                u = x + 1.0f;
                w = z + 1.0f;
            } else {
                u = z + 1.0f;
                w = y + 1.0f;
            }
            v = __builtin_sqrtf(u * u + w * w);
        }
        out[t] = v;
    }
}

// k = (f - a)/(b - a); out = a1 + b1 k (three) or k (buffer); (a, b) = part[0]
__global__ void rescale_kernel(const float *in, float *out, uint64_t n, const float2 *ab,
                               float a1, float b1, int affine) {
    const float a = ab->x, b = ab->y;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const float k = (in[t] - a) / (b - a);
        out[t] = affine ? a1 + b1 * k : k;
    }
}

// ---------------------------------------------------------------------------
// finding_phi: right-hand side D^T K of one harmonic (Gratings.cu:100-417)
// ---------------------------------------------------------------------------
struct PhiArgs {
    int nx, ny, nz;
    int i, j, k;
    float dx, dy, dz;
    int type;          // 'n' 'b' 'r' 's'
    int uniform_type;  // 0 const, 1 per axis, 2 period field
    float const_period, x_period, y_period, z_period;
    float lcon, lcon_1;
    int sinewave_zaxis;
};

// stencil weights and the two neighbours of coordinate c in [0, n)
__device__ __forceinline__ void axis_weights(int c, int n, float *w1, float *w2, int *c1, int *c2) {
    if (c == 0) {
        *w1 = -1.0f; *w2 = -0.5f; *c1 = c; *c2 = c + 1;
    } else if (c == 1) {
        *w1 = 1.0f; *w2 = -0.5f; *c1 = c - 1; *c2 = c + 1;
    } else if (c == n - 2) {
        *w1 = 0.5f; *w2 = -1.0f; *c1 = c - 1; *c2 = c + 1;
    } else if (c == n - 1) {
        *w1 = 0.5f; *w2 = 1.0f; *c1 = c - 1; *c2 = c;
    } else {
        *w1 = 0.5f; *w2 = -0.5f; *c1 = c - 1; *c2 = c + 1;
    }
}

// (2 pi / per) * w, the reference's double expression rounded to binary32
__device__ __forceinline__ float wave(float per, float w) {
    return (float)((kTwoPi / (double)per) * (double)w);
}

// lcon * sin(6.28 * lcon_1 * u): the product in double as the reference's
// double literal makes it, the sine of the binary32 argument
__device__ __forceinline__ float sine_angle(float lcon, float lcon_1, float u) {
    const float arg = (float)((6.28 * (double)lcon_1) * (double)u);
    return lcon * sin_rad(arg);
}

__global__ void finding_phi_kernel(float *phi, const float *period, PhiArgs p) {
    const uint64_t n = (uint64_t)p.nx * p.ny * p.nz;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        int x, y, z;
        point_xyz(t, p.nx, p.ny, &x, &y, &z);
        float a1, a2, b1, b2, c1, c2;
        int x1, x2, y1, y2, z1, z2;
        axis_weights(x, p.nx, &a1, &a2, &x1, &x2);
        axis_weights(y, p.ny, &b1, &b2, &y1, &y2);
        axis_weights(z, p.nz, &c1, &c2, &z1, &z2);
        float th1 = 0.f, th2 = 0.f, th3 = 0.f, th4 = 0.f, th5 = 0.f, th6 = 0.f, th7 = 0.f,
              th8 = 0.f;
        if (p.type == 'r' || p.type == 'b') {
            const float mx = p.type == 'r' ? (float)(p.nx + 1) / 2.0f : 0.f;
            const float my = p.type == 'r' ? (float)(p.ny + 1) / 2.0f : 0.f;
            const float yy = ((float)(y + 1) - my) * p.dx;
            th1 = atan2_poly(yy, ((float)(x1 + 1) - mx) * p.dx);
            th2 = atan2_poly(yy, ((float)(x2 + 1) - mx) * p.dx);
            const float xx = ((float)(x + 1) - mx) * p.dx;
            th3 = atan2_poly(((float)(y1 + 1) - my) * p.dx, xx);
            th4 = atan2_poly(((float)(y2 + 1) - my) * p.dx, xx);
        } else if (p.type == 's') {
            th1 = sine_angle(p.lcon, p.lcon_1, (float)(x1 + 1) * p.dx);
            th5 = sine_angle(p.lcon, p.lcon_1, (float)(z1 + 1) * p.dz);
            th2 = sine_angle(p.lcon, p.lcon_1, (float)(x2 + 1) * p.dx);
            th6 = sine_angle(p.lcon, p.lcon_1, (float)(z2 + 1) * p.dz);
            th3 = sine_angle(p.lcon, p.lcon_1, (float)(x + 1) * p.dx);
            th7 = sine_angle(p.lcon, p.lcon_1, (float)(z + 1) * p.dz);
            th4 = th3;
            th8 = th7;
        }
        float per1, per2, per3, per4, per5, per6;
        if (p.uniform_type == 0) {
            per1 = per2 = per3 = per4 = per5 = per6 = p.const_period;
        } else if (p.uniform_type == 1) {
            per1 = per2 = p.x_period;
            per3 = per4 = p.y_period;
            per5 = per6 = p.z_period;
        } else {
            const uint64_t row = (uint64_t)p.nx, plane = (uint64_t)p.nx * p.ny;
            const uint64_t c = (uint64_t)y * row + (uint64_t)z * plane;
            per1 = period[(uint64_t)x1 + c];
            per2 = period[(uint64_t)x2 + c];
            per3 = period[(uint64_t)x + (uint64_t)y1 * row + (uint64_t)z * plane];
            per4 = period[(uint64_t)x + (uint64_t)y2 * row + (uint64_t)z * plane];
            per5 = period[(uint64_t)x + (uint64_t)y * row + (uint64_t)z1 * plane];
            per6 = period[(uint64_t)x + (uint64_t)y * row + (uint64_t)z2 * plane];
        }
        const float fi = (float)p.i, fj = (float)p.j, fk = (float)p.k;
        float kx1 = wave(per1, fi * cos_rad(th1) - fj * sin_rad(th1));
        float kx2 = wave(per2, fi * cos_rad(th2) - fj * sin_rad(th2));
        float ky1 = wave(per3, fi * sin_rad(th3) + fj * cos_rad(th3));
        float ky2 = wave(per4, fi * sin_rad(th4) + fj * cos_rad(th4));
        float kz1 = wave(per5, fk);
        float kz2 = wave(per6, fk);
        float ph = (__builtin_fmaf(a1, kx1, a2 * kx2) + __builtin_fmaf(b1, ky1, b2 * ky2)) +
                   __builtin_fmaf(c1, kz1, c2 * kz2);
        if (p.type == 's' && p.sinewave_zaxis) {
            kz1 = wave(per5, fj * sin_rad(th5) + fk * cos_rad(th5));
            kz2 = wave(per6, fj * sin_rad(th6) + fk * cos_rad(th6));
            ky1 = wave(per3, fj * cos_rad(th7) - fk * sin_rad(th7));
            ky2 = wave(per4, fj * cos_rad(th8) - fk * sin_rad(th8));
            kx1 = wave(per1, fi);
            kx2 = wave(per2, fi);
            ph += (__builtin_fmaf(a1, kx1, a2 * kx2) + __builtin_fmaf(b1, ky1, b2 * ky2)) +
                  __builtin_fmaf(c1, kz1, c2 * kz2);
        }
        phi[t] = ph;
    }
}

// ---------------------------------------------------------------------------
// the operator D^T D of the phase solve (GPUMatvec_lattice_kernel,
// Gratings.cu:420-597): per axis a diagonal weight and two neighbour terms
// ---------------------------------------------------------------------------
// Coordinate c of n (n >= 4) along stride s:
//   c = 0:     w 1.25, t2 = -d(t+s),         t3 = -0.25 d(t+2s)
//   c = 1:     w 1.25, t2 = -d(t-s),         t3 = -0.25 d(t+2s)
//   c = n-2:   w 1.25, t2 = -0.25 d(t-2s),   t3 = -d(t+s)
//   c = n-1:   w 1.25, t2 = -0.25 d(t-2s),   t3 = -d(t-s)
//   otherwise: w 0.5,  t2 = -0.25 d(t-2s),   t3 = -0.25 d(t+2s)
// chosen by selects, so both loads are unconditional (every point's loads
// can be in flight together); -1 * v and -0.25 * v are the exact negation
// and quarter the cases name.
template <class G>
__device__ __forceinline__ void axis_terms(const G &get, int c, int n, int64_t t, int64_t s,
                                           float *w, float *t2, float *t3) {
    const bool lo = c <= 1, hi1 = c == n - 2, hi0 = c == n - 1;
    *w = (lo || hi1 || hi0) ? 1.25f : 0.5f;
    const int64_t o2 = c == 0 ? s : (c == 1 ? -s : -2 * s);
    const float k2 = lo ? -1.0f : -0.25f;
    const int64_t o3 = lo ? 2 * s : (hi1 ? s : (hi0 ? -s : 2 * s));
    const float k3 = lo ? -0.25f : ((hi1 || hi0) ? -1.0f : -0.25f);
    *t2 = k2 * get(t + o2);
    *t3 = k3 * get(t + o3);
}

// (wx + wy + wz) * d + x2 + x3 + y2 + y3 + z2 + z3 in this order, the first
// product and sum fused (the reference's contracted form)
template <class G>
__device__ __forceinline__ float apply_op(const G &get, int x, int y, int z, int nx, int ny,
                                          int nz, int64_t t) {
    float wx, x2, x3, wy, y2, y3, wz, z2, z3;
    axis_terms(get, x, nx, t, 1, &wx, &x2, &x3);
    axis_terms(get, y, ny, t, (int64_t)nx, &wy, &y2, &y3);
    axis_terms(get, z, nz, t, (int64_t)nx * ny, &wz, &z2, &z3);
    const float w = (wx + wy) + wz;
    float a = __builtin_fmaf(w, get(t), x2);
    a = a + x3;
    a = a + y2;
    a = a + y3;
    a = a + z2;
    a = a + z3;
    return a;
}

__global__ void matvec_kernel(const float *d, float *q, int nx, int ny, int nz) {
    const uint64_t n = (uint64_t)nx * ny * nz;
    auto get = [d](int64_t i) { return d[i]; };
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        int x, y, z;
        point_xyz(t, nx, ny, &x, &y, &z);
        q[t] = apply_op(get, x, y, z, nx, ny, nz, (int64_t)t);
    }
}

// ---------------------------------------------------------------------------
// conjugate gradients over m right-hand sides (GPUCG_lattice,
// Gratings.cu:875-974, scalars kept on the device)
// ---------------------------------------------------------------------------
constexpr int kCgThreads = 256;
constexpr int kCgBlocks = 512;  // blocks of the stand-alone dot product
constexpr int kCgResident = 1280;  // CG workgroups over all right-hand sides (5 per CU)
constexpr int kMaxRhs = 4;

struct CgState {
    float delta[2];  // r.r of the last two iterations (by iteration parity)
    int done;
    int iters;       // the reference's FinalIter
    float res;       // the reference's FinalRes = sqrt(r.r)
    int pad[3];
};

constexpr int kCgUnroll = 2;  // points per thread and step, all their loads in flight
constexpr int kCgStep = kCgThreads * kCgUnroll;

struct CgArgs {
    int nx, ny, nz, G;             // G: workgroups per right-hand side, a multiple of 8
    uint64_t n, C;                 // C: points per workgroup range, a multiple of kCgThreads
    Div32 dnx, dplane;
    float *phi, *r, *q, *d0, *d1;  // m vectors of n each (right-hand side s at s*n)
    double *pq, *prr;              // partial sums: pq[m][G], prr[2][m][G]
    CgState *st;
    int max_iter;
    float term;                    // end_res^2
};

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double s[kCgThreads];
    __syncthreads();
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = kCgThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    return s[0];
}

// the same fixed-order sum of G partials in every block (so every block
// takes the same decisions without communicating)
__device__ __forceinline__ double sum_partials(const double *p, int G) {
    double v = 0.0;
    for (int i = threadIdx.x; i < G; i += kCgThreads) v += p[i];
    return block_sum(v);
}

// Points [lo, hi) of workgroup blockIdx.x.  Workgroups are dispatched to
// the 8 XCDs round robin, so range (b % 8) * (G / 8) + b / 8 gives each XCD
// one contiguous run of z planes: the operator's neighbours (two planes up
// and down) then sit in the L2 of the XCD that reads them.
__device__ __forceinline__ void cg_range(const CgArgs &a, uint64_t *lo, uint64_t *hi) {
    const uint32_t b = blockIdx.x;
    const uint64_t r = (uint64_t)((b & 7u) * ((uint32_t)a.G >> 3) + (b >> 3));
    *lo = r * a.C < a.n ? r * a.C : a.n;
    *hi = *lo + a.C < a.n ? *lo + a.C : a.n;
}

__global__ void cg_init_kernel(const float *b, CgArgs a) {
    const int s = blockIdx.y;
    const uint64_t off = (uint64_t)s * a.n;
    uint64_t lo, hi;
    cg_range(a, &lo, &hi);
    double acc = 0.0;
    for (uint64_t t = lo + threadIdx.x; t < hi; t += kCgThreads) {
        const float v = b[off + t];
        a.phi[off + t] = 0.0f;
        a.r[off + t] = v;
        a.d0[off + t] = v;
        acc += (double)v * (double)v;
    }
    const double tot = block_sum(acc);
    if (threadIdx.x == 0) a.prr[(uint64_t)s * a.G + blockIdx.x] = tot;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        CgState z = {};
        a.st[s] = z;
    }
}

// iteration it, first half: convergence test, d_it = r + beta d_{it-1}
// (recomputed for the stencil's neighbours), q = A d_it, partial d.q
__global__ void cg_a_kernel(CgArgs a, int it) {
    const int s = blockIdx.y;
    if (a.st[s].done) return;
    const uint64_t m_off = (uint64_t)s * a.n;
    const int m = gridDim.y;
    const double *prr = a.prr + ((uint64_t)(it & 1) * m + s) * a.G;
    const float delta = (float)sum_partials(prr, a.G);
    // the reference loops while (iCounter < iter && delta_new > term),
    // iCounter = it + 1
    if (!(it + 1 < a.max_iter && delta > a.term)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.st[s].iters = it + 1;
            a.st[s].res = sqrtf(delta);
            a.st[s].done = 1;
        }
        return;
    }
    const float beta = it == 0 ? 0.0f : delta / a.st[s].delta[(it + 1) & 1];
    if (blockIdx.x == 0 && threadIdx.x == 0) a.st[s].delta[it & 1] = delta;
    float *dcur = (it & 1 ? a.d1 : a.d0) + m_off;
    const float *r = a.r + m_off;
    const bool first = it == 0;
    // d = r + beta d_prev; the first direction is r itself (fma(0, r, r) = r
    // exactly), so one load pattern serves every iteration
    const float *dprev = first ? r : ((it - 1) & 1 ? a.d1 : a.d0) + m_off;
    const float bt = first ? 0.0f : beta;
    auto get = [=](int64_t i) { return __builtin_fmaf(bt, dprev[i], r[i]); };
    // points t0 + u * kCgThreads (u < kCgUnroll) per step: every load of the
    // step is issued before its stores; the partial sum adds a thread's points
    // in increasing order
    uint64_t lo, hi;
    cg_range(a, &lo, &hi);
    double acc = 0.0;
    for (uint64_t t0 = lo + threadIdx.x; t0 < hi; t0 += kCgStep) {
        float dv[kCgUnroll], qv[kCgUnroll];
#pragma unroll
        for (int u = 0; u < kCgUnroll; ++u) {
            const uint32_t t = (uint32_t)(t0 + u * kCgThreads < hi ? t0 + u * kCgThreads : t0);
            int x, y, z;
            point_xyz(t, a.dnx, a.dplane, &x, &y, &z);
            dv[u] = get((int64_t)t);
            qv[u] = apply_op(get, x, y, z, a.nx, a.ny, a.nz, (int64_t)t);
        }
#pragma unroll
        for (int u = 0; u < kCgUnroll; ++u) {
            const uint64_t t = t0 + u * kCgThreads;
            if (t < hi) {
                if (!first) dcur[t] = dv[u];
                a.q[m_off + t] = qv[u];
                acc += (double)dv[u] * (double)qv[u];
            }
        }
    }
    const double tot = block_sum(acc);
    if (threadIdx.x == 0) a.pq[(uint64_t)s * a.G + blockIdx.x] = tot;
}

// iteration it, second half: alpha = delta / d.q, phi += alpha d,
// r -= alpha q, partial r.r for the next test
__global__ void cg_b_kernel(CgArgs a, int it) {
    const int s = blockIdx.y;
    if (a.st[s].done) return;
    const uint64_t m_off = (uint64_t)s * a.n;
    const int m = gridDim.y;
    const float dq = (float)sum_partials(a.pq + (uint64_t)s * a.G, a.G);
    const float alpha = a.st[s].delta[it & 1] / dq;
    const float *d = (it & 1 ? a.d1 : a.d0) + m_off;
    float *phi = a.phi + m_off, *r = a.r + m_off;
    const float *q = a.q + m_off;
    uint64_t lo, hi;
    cg_range(a, &lo, &hi);
    double acc = 0.0;
    for (uint64_t t0 = lo + threadIdx.x; t0 < hi; t0 += kCgStep) {
        float dv[kCgUnroll], pv[kCgUnroll], qv[kCgUnroll], rv[kCgUnroll];
#pragma unroll
        for (int u = 0; u < kCgUnroll; ++u) {
            const uint32_t t = (uint32_t)(t0 + u * kCgThreads < hi ? t0 + u * kCgThreads : t0);
            dv[u] = d[t];
            pv[u] = phi[t];
            qv[u] = q[t];
            rv[u] = r[t];
        }
#pragma unroll
        for (int u = 0; u < kCgUnroll; ++u) {
            const uint64_t t = t0 + u * kCgThreads;
            if (t < hi) {
                phi[t] = __builtin_fmaf(alpha, dv[u], pv[u]);
                const float rn = __builtin_fmaf(-alpha, qv[u], rv[u]);
                r[t] = rn;
                acc += (double)rn * (double)rn;
            }
        }
    }
    const double tot = block_sum(acc);
    if (threadIdx.x == 0) a.prr[((uint64_t)((it + 1) & 1) * m + s) * a.G + blockIdx.x] = tot;
}

// Solves A x = b for m right-hand sides (b: m vectors of n, x written over
// b when x == b).  Iterations run in chunks between host checks of the done
// flags; a finished right-hand side leaves the kernels at their first test.
int cg_solve(const float *b, float *x, int m, int nx, int ny, int nz, int max_iter, float end_res,
             int *iters, float *res, hipStream_t st) {
    const uint64_t n = (uint64_t)nx * ny * nz;
    CgArgs a;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.n = n;
    a.dnx = make_div32((uint32_t)nx);
    a.dplane = make_div32((uint32_t)nx * (uint32_t)ny);
    // about kCgResident workgroups in all (one wave of them, none waiting),
    // fewer when ranges of one step per thread cover the grid
    const uint64_t steps = (n + kCgStep - 1) / kCgStep;
    a.G = (int)std::min<uint64_t>((steps + 7) & ~uint64_t(7), std::max(8, (kCgResident / m) & ~7));
    a.C = ((n + a.G - 1) / a.G + kCgThreads - 1) / kCgThreads * kCgThreads;
    a.max_iter = max_iter;
    a.term = end_res * end_res;
    const uint64_t vec = m * n * sizeof(float);
    const uint64_t slot = (vec + 255) & ~uint64_t(255);  // 256-byte aligned vectors
    const uint64_t bytes = 5 * slot + (uint64_t)m * a.G * 8 * 3 + kMaxRhs * sizeof(CgState) + 1024;
    void *ws = nullptr;
    int rc = scratch_arena(bytes, &ws);
    if (rc) return rc;
    char *w = (char *)ws;
    a.phi = (float *)w;
    a.r = (float *)(w + slot);
    a.q = (float *)(w + 2 * slot);
    a.d0 = (float *)(w + 3 * slot);
    a.d1 = (float *)(w + 4 * slot);
    a.pq = (double *)(w + 5 * slot);
    a.prr = a.pq + (uint64_t)m * a.G;
    a.st = (CgState *)(a.prr + 2 * (uint64_t)m * a.G);
    const dim3 grid((unsigned)a.G, (unsigned)m);
    hipLaunchKernelGGL(cg_init_kernel, grid, dim3(kCgThreads), 0, st, b, a);
    GC_CHECK_LAUNCH();
    constexpr int kChunk = 16;
    CgState h[kMaxRhs];
    for (int it0 = 0; it0 < max_iter; it0 += kChunk) {
        for (int it = it0; it < std::min(max_iter, it0 + kChunk); ++it) {
            hipLaunchKernelGGL(cg_a_kernel, grid, dim3(kCgThreads), 0, st, a, it);
            hipLaunchKernelGGL(cg_b_kernel, grid, dim3(kCgThreads), 0, st, a, it);
        }
        GC_CHECK_LAUNCH();
        GC_HIP(hipMemcpyAsync(h, a.st, m * sizeof(CgState), hipMemcpyDeviceToHost, st));
        GC_HIP(hipStreamSynchronize(st));
        bool all = true;
        for (int s = 0; s < m; ++s) all = all && h[s].done;
        if (all) break;
    }
    for (int s = 0; s < m; ++s) {
        if (!h[s].done) return fail(GPUCAD_E_HIP, "internal: CG did not stop by iteration %d",
                                    max_iter);
        if (iters) iters[s] = h[s].iters;
        if (res) res[s] = h[s].res;
    }
    GC_HIP(hipMemcpyAsync(x, a.phi, vec, hipMemcpyDeviceToDevice, st));
    GC_HIP(hipStreamSynchronize(st));
    return GPUCAD_OK;
}

// ---------------------------------------------------------------------------
// 2x upsampling of a coarse field as the reference's texture fetch does it
// (tex3D, unnormalised coordinates, texel centres at +0.5, linear filter,
// clamp at the far end: Gratings.cu:653-722, Interpolations.cu:57-107).
// Fine point f samples texel coordinate u = f * h; weights are the fraction
// of u quantised to 1/256 (the filter's 8 fraction bits); lerps run along
// x, then y, then z.
// ---------------------------------------------------------------------------
struct Tap {
    int64_t i0, i1;
    float w;
};

__device__ __forceinline__ Tap tap(int f, float h, int n) {
    const float u = (float)f * h;
    const float fl = floorf(u);
    float w = u - fl;
    w = rintf(w * 256.0f) * (1.0f / 256.0f);
    int64_t i0 = (int64_t)fl;
    if (w >= 1.0f) {
        i0 += 1;
        w = 0.0f;
    }
    Tap t;
    t.i0 = i0 < 0 ? 0 : (i0 > n - 1 ? n - 1 : i0);
    t.i1 = i0 + 1 > n - 1 ? n - 1 : (i0 + 1 < 0 ? 0 : i0 + 1);
    t.w = w;
    return t;
}

__device__ __forceinline__ float lerp1(float a, float b, float w) {
    return __builtin_fmaf(w, b, (1.0f - w) * a);
}

__device__ __forceinline__ float sample3(const float *c, const Tap &tx, const Tap &ty, const Tap &tz,
                                         int64_t row, int64_t plane) {
    const int64_t z0 = tz.i0 * plane, z1 = tz.i1 * plane, y0 = ty.i0 * row, y1 = ty.i1 * row;
    const float c00 = lerp1(c[z0 + y0 + tx.i0], c[z0 + y0 + tx.i1], tx.w);
    const float c10 = lerp1(c[z0 + y1 + tx.i0], c[z0 + y1 + tx.i1], tx.w);
    const float c01 = lerp1(c[z1 + y0 + tx.i0], c[z1 + y0 + tx.i1], tx.w);
    const float c11 = lerp1(c[z1 + y1 + tx.i0], c[z1 + y1 + tx.i1], tx.w);
    return lerp1(lerp1(c00, c10, ty.w), lerp1(c01, c11, ty.w), tz.w);
}

struct Up {
    int nx, ny, nz;     // coarse
    int fx, fy, fz;     // fine
    float hx, hy, hz;   // fine spacing in coarse texels
};

// mode 0: refine (the sample); 1: grating (cos, sin of the sample)
__global__ void upsample_kernel(const float *c, Up u, float *out, float2 *grating, int mode) {
    const uint64_t n = (uint64_t)u.fx * u.fy * u.fz;
    const int64_t row = u.nx, plane = (int64_t)u.nx * u.ny;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        int x, y, z;
        point_xyz(t, u.fx, u.fy, &x, &y, &z);
        const float v = sample3(c, tap(x, u.hx, u.nx), tap(y, u.hy, u.ny), tap(z, u.hz, u.nz), row,
                                plane);
        if (mode == 0) out[t] = v;
        else grating[t] = make_float2(cos_rad(v), sin_rad(v));
    }
}

// svl += Re(grating * c[index])  (svl_kernel, Gratings.cu:724-752)
__global__ void svl_acc_kernel(float *svl, const float2 *g, uint64_t n, const float2 *coeffs,
                               int index) {
    const float2 c = coeffs[index];
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const float2 a = g[t];
        svl[t] = svl[t] + __builtin_fmaf(a.x, c.x, -(a.y * c.y));
    }
}

// ---------------------------------------------------------------------------
// native fused synthesis: svl[f] = sum over the first H harmonics (i, j, k)
// of Re(c_h e^{i (i Phi_x + j Phi_y + k Phi_z)(f)}), Phi upsampled as above;
// e^{i phi_h} = Ex^i Ey^j Ez^k from the three exponentials and their powers.
// ---------------------------------------------------------------------------
constexpr int kMaxRange = 4;  // (2r+1)^3 <= 729 coefficients

struct Cf {
    float re, im;
};
__device__ __forceinline__ Cf cmulf(Cf a, Cf b) {
    return Cf{__builtin_fmaf(a.re, b.re, -(a.im * b.im)), __builtin_fmaf(a.re, b.im, a.im * b.re)};
}

__global__ void svl_field_kernel(const float *phix, const float *phiy, const float *phiz, Up u,
                                 const float2 *coeffs, int range, int harmonics, float *svl) {
    extern __shared__ float2 s_c[];  // harmonics coefficients, then their (i, j, k)
    int *s_ijk = (int *)(s_c + harmonics);
    const int side = 2 * range + 1;
    for (int h = threadIdx.x; h < harmonics; h += blockDim.x) {
        s_c[h] = coeffs[h];
        const int i = h % side - range, j = (h / side) % side - range, k = h / (side * side) - range;
        s_ijk[h] = (i + 8) | ((j + 8) << 4) | ((k + 8) << 8);
    }
    __syncthreads();
    const uint64_t n = (uint64_t)u.fx * u.fy * u.fz;
    const int64_t row = u.nx, plane = (int64_t)u.nx * u.ny;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        int x, y, z;
        point_xyz(t, u.fx, u.fy, &x, &y, &z);
        const Tap tx = tap(x, u.hx, u.nx), ty = tap(y, u.hy, u.ny), tz = tap(z, u.hz, u.nz);
        const float px = sample3(phix, tx, ty, tz, row, plane);
        const float py = sample3(phiy, tx, ty, tz, row, plane);
        const float pz = sample3(phiz, tx, ty, tz, row, plane);
        // powers E^0..E^range of each axis; negative powers are conjugates
        Cf ex[kMaxRange + 1], ey[kMaxRange + 1], ez[kMaxRange + 1];
        ex[0] = ey[0] = ez[0] = Cf{1.0f, 0.0f};
        ex[1] = Cf{cos_rad(px), sin_rad(px)};
        ey[1] = Cf{cos_rad(py), sin_rad(py)};
        ez[1] = Cf{cos_rad(pz), sin_rad(pz)};
#pragma unroll
        for (int p = 2; p <= kMaxRange; ++p) {
            ex[p] = cmulf(ex[p - 1], ex[1]);
            ey[p] = cmulf(ey[p - 1], ey[1]);
            ez[p] = cmulf(ez[p - 1], ez[1]);
        }
        auto pw = [&](const Cf *e, int k) {
            Cf r = e[0];
#pragma unroll
            for (int p = 1; p <= kMaxRange; ++p) {
                r = (k == p || k == -p) ? e[p] : r;
            }
            return k < 0 ? Cf{r.re, -r.im} : r;
        };
        float acc = 0.0f;
        for (int h = 0; h < harmonics; ++h) {
            const int c3 = s_ijk[h];
            const int i = (c3 & 15) - 8, j = ((c3 >> 4) & 15) - 8, k = ((c3 >> 8) & 15) - 8;
            const Cf e = cmulf(cmulf(pw(ex, i), pw(ey, j)), pw(ez, k));
            const float2 c = s_c[h];
            acc = acc + __builtin_fmaf(e.re, c.x, -(e.im * c.y));
        }
        svl[t] = acc;
    }
}

// ---------------------------------------------------------------------------
// unit-cell Fourier coefficients by direct DFT: coefficient (i, j, k) =
// (1/N) sum f(x,y,z) e^{-2 pi i (i x/nx + j y/ny + k z/nz)}, the forward R2C
// transform scaled as fft_scalar does (Fft_lattice.cu) and picked as
// unit_lattice does (src/main.cu:3578-3711: loop k, j, i over [-r, r])
// ---------------------------------------------------------------------------
__global__ void dft_kernel(const float *cell, int nx, int ny, int nz, int range,
                           const double2 *tw /* (nx+ny+nz) * (2r+1) twiddles */, float2 *out) {
    const int side = 2 * range + 1;
    const int h = blockIdx.x;
    const int i = h % side, j = (h / side) % side, k = h / (side * side);
    const double2 *tx = tw + (uint64_t)i * nx;
    const double2 *ty = tw + (uint64_t)side * nx + (uint64_t)j * ny;
    const double2 *tz = tw + (uint64_t)side * (nx + ny) + (uint64_t)k * nz;
    const uint64_t n = (uint64_t)nx * ny * nz;
    double re = 0.0, im = 0.0;
    for (uint64_t t = threadIdx.x; t < n; t += blockDim.x) {
        int x, y, z;
        point_xyz(t, nx, ny, &x, &y, &z);
        const double2 a = tx[x], b = ty[y], c = tz[z];
        const double abr = a.x * b.x - a.y * b.y, abi = a.x * b.y + a.y * b.x;
        const double er = abr * c.x - abi * c.y, ei = abr * c.y + abi * c.x;
        const double f = cell[t];
        re += f * er;
        im += f * ei;
    }
    __shared__ double sr[256], si[256];
    sr[threadIdx.x] = re;
    si[threadIdx.x] = im;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sr[threadIdx.x] += sr[threadIdx.x + w];
            si[threadIdx.x] += si[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[h] = make_float2((float)(sr[0] / (double)n), (float)(si[0] / (double)n));
}

// fft_fill_kernel (Fft_lattice.cu): the half spectrum (nx/2+1 columns) into
// the full one with its conjugate mirror; the mirror indices use nx on every
// axis, as the reference's do
__global__ void fft_fill_kernel(const float2 *half, float2 *full, int nx, int ny, int nz) {
    const int nx2 = nx / 2 + 1;
    const uint64_t n = (uint64_t)nx2 * ny * nz;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        int x, y, z;
        point_xyz(t, nx2, ny, &x, &y, &z);
        if (t == 0) {
            full[0] = half[0];
            continue;
        }
        const int e = x == 0 ? 0 : nx - x, f = y == 0 ? 0 : nx - y, g = z == 0 ? 0 : nx - z;
        const float2 s = half[t];
        full[(uint64_t)x + (uint64_t)y * nx + (uint64_t)z * nx * ny] = s;
        full[(uint64_t)e + (uint64_t)f * nx + (uint64_t)g * nx * ny] = make_float2(s.x, -s.y);
    }
}

__global__ void fft_scale_kernel(float2 *d, float s, uint64_t n) {
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x) {
        float2 a = d[t];
        a.x /= s;
        a.y /= s;
        d[t] = a;
    }
}

// VecSMultAddKernel_lattice (Gratings.cu:73-94): v = a1 v + a2 w (fused)
__global__ void axpby_kernel(float *v, float a1, const float *w, float a2, uint64_t n) {
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x)
        v[t] = __builtin_fmaf(a2, w[t], a1 * v[t]);
}

// fillfrac_kernel (Gratings.cu:755-772): svl *= fillfrac
__global__ void fillfrac_kernel(float *svl, const float *frac, uint64_t n) {
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * blockDim.x)
        svl[t] = frac[t] * svl[t];
}

// GPUScalar_lattice (Gratings.cu:600-650, 22-70): a.b into result[0]
// (partial sums in double per block, one fixed-order final sum)
__global__ void dot_partial_kernel(const float *a, const float *b, uint64_t n, double *part) {
    double acc = 0.0;
    for (uint64_t t = (uint64_t)blockIdx.x * kCgThreads + threadIdx.x; t < n;
         t += (uint64_t)gridDim.x * kCgThreads)
        acc += (double)a[t] * (double)b[t];
    const double tot = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void dot_final_kernel(const double *part, int G, float *result) {
    const double tot = sum_partials(part, G);
    if (threadIdx.x == 0) result[0] = (float)tot;
}

dim3 stride_grid(uint64_t n) {
    return dim3((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8192)));
}

int check_dims(int nx, int ny, int nz, int minimum) {
    GC_ARG(nx >= minimum && ny >= minimum && nz >= minimum, "grid sizes must be >= %d (got %d %d %d)",
           minimum, nx, ny, nz);
    GC_ARG((uint64_t)nx * ny * nz < (1ull << 32), "grid too large (2^32 points or more)");
    return GPUCAD_OK;
}

PhiArgs phi_args(const gc_svl_params *p, int i, int j, int k) {
    PhiArgs a;
    a.nx = p->nx;
    a.ny = p->ny;
    a.nz = p->nz;
    a.i = i;
    a.j = j;
    a.k = k;
    a.dx = p->dx;
    a.dy = p->dy;
    a.dz = p->dz;
    a.type = p->lattice_type;
    a.uniform_type = p->uniform_type;
    a.const_period = p->const_period;
    a.x_period = p->x_period;
    a.y_period = p->y_period;
    a.z_period = p->z_period;
    a.lcon = p->lcon;
    a.lcon_1 = p->lcon_1;
    a.sinewave_zaxis = p->sinewave_zaxis;
    return a;
}

int check_params(const gc_svl_params *p, const float *period) {
    GC_ARG(p, "params is NULL");
    int st = check_dims(p->nx, p->ny, p->nz, 4);
    if (st) return st;
    GC_ARG(p->lattice_type == 'n' || p->lattice_type == 'b' || p->lattice_type == 'r' ||
               p->lattice_type == 's',
           "lattice_type must be 'n', 'b', 'r' or 's'");
    GC_ARG(p->uniform_type >= 0 && p->uniform_type <= 2, "uniform_type must be 0, 1 or 2");
    GC_ARG(p->uniform_type != 2 || period, "uniform_type 2 needs a period field");
    return GPUCAD_OK;
}

}  // namespace
}  // namespace gc

using namespace gc;

extern "C" {

int gpucad_svl_period_data(float *d_period, int32_t nx, int32_t ny, int32_t nz, float dx, float dy,
                           float dz, float mean_x, float mean_y, float mean_z, char axis) {
    clear_error();
    GC_ARG(d_period, "NULL buffer");
    int st = check_dims(nx, ny, nz, 1);
    if (st) return st;
    const uint64_t n = (uint64_t)nx * ny * nz;
    hipLaunchKernelGGL(period_angle_kernel, stride_grid(n), dim3(256), 0, 0, d_period, nx, ny, nz,
                       dx, dy, dz, mean_x, mean_y, mean_z, (int)axis, 0);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_svl_angle_data(float *d_theta, int32_t nx, int32_t ny, int32_t nz, float dx, float dy,
                          float dz, float mean_x, float mean_y, float mean_z, char axis) {
    clear_error();
    GC_ARG(d_theta, "NULL buffer");
    int st = check_dims(nx, ny, nz, 1);
    if (st) return st;
    const uint64_t n = (uint64_t)nx * ny * nz;
    hipLaunchKernelGGL(period_angle_kernel, stride_grid(n), dim3(256), 0, 0, d_theta, nx, ny, nz,
                       dx, dy, dz, mean_x, mean_y, mean_z, (int)axis, 1);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

static int rescale(const float *in, float *out, uint64_t size, float a1, float b1, int affine) {
    clear_error();
    auto calls = lock_device_calls();  // shared scratch arena
    GC_ARG(in && out, "NULL buffer");
    GC_ARG(size > 0, "size must be positive");
    void *ws = nullptr;
    int st = scratch_arena(1024 * sizeof(float2), &ws);
    if (st) return st;
    float2 *part = (float2 *)ws;
    st = enqueue_minmax0(in, size, part, 0);
    if (st) return st;
    hipLaunchKernelGGL(rescale_kernel, stride_grid(size), dim3(256), 0, 0, in, out, size, part, a1,
                       b1, affine);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_normalise_three(const float *in, float *out, uint64_t size, float a1, float b1) {
    return rescale(in, out, size, a1, b1, 1);
}

int gpucad_normalise_buffer(const float *in, float *out, uint64_t size) {
    return rescale(in, out, size, 0.0f, 0.0f, 0);
}

int gpucad_svl_finding_phi(float *d_phi, const float *d_period, int32_t nx, int32_t ny, int32_t nz,
                           int32_t i, int32_t j, int32_t k, float dx, float dy, float dz,
                           char lattice_type, int32_t uniform_type, float const_period,
                           float x_period, float y_period, float z_period, float lcon,
                           float lcon_1, int32_t sinewave_zaxis) {
    clear_error();
    GC_ARG(d_phi, "NULL buffer");
    gc_svl_params p = {};
    p.nx = nx;
    p.ny = ny;
    p.nz = nz;
    p.dx = dx;
    p.dy = dy;
    p.dz = dz;
    p.lattice_type = lattice_type;
    p.uniform_type = uniform_type;
    p.const_period = const_period;
    p.x_period = x_period;
    p.y_period = y_period;
    p.z_period = z_period;
    p.lcon = lcon;
    p.lcon_1 = lcon_1;
    p.sinewave_zaxis = sinewave_zaxis;
    int st = check_params(&p, d_period);
    if (st) return st;
    const uint64_t n = (uint64_t)nx * ny * nz;
    hipLaunchKernelGGL(finding_phi_kernel, stride_grid(n), dim3(256), 0, 0, d_phi, d_period,
                       phi_args(&p, i, j, k));
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_svl_matvec(const float *d_d, float *d_q, int32_t nx, int32_t ny, int32_t nz) {
    clear_error();
    GC_ARG(d_d && d_q && d_d != d_q, "NULL or aliased buffer");
    int st = check_dims(nx, ny, nz, 4);
    if (st) return st;
    const uint64_t n = (uint64_t)nx * ny * nz;
    hipLaunchKernelGGL(matvec_kernel, stride_grid(n), dim3(256), 0, 0, d_d, d_q, nx, ny, nz);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_svl_cg(float *d_phi, int32_t nx, int32_t ny, int32_t nz, int32_t iter, float end_res,
                  int32_t *final_iter, float *final_res) {
    clear_error();
    auto calls = lock_device_calls();  // shared scratch arena
    GC_ARG(d_phi, "NULL buffer");
    int st = check_dims(nx, ny, nz, 4);
    if (st) return st;
    GC_ARG(iter >= 1, "iter must be >= 1");
    int it = 0;
    float res = 0.f;
    st = cg_solve(d_phi, d_phi, 1, nx, ny, nz, iter, end_res, &it, &res, 0);
    if (st) return st;
    if (final_iter) *final_iter = it;
    if (final_res) *final_res = res;
    return GPUCAD_OK;
}

int gpucad_svl_axpby(float *d_v, float a1, const float *d_w, float a2, uint64_t n) {
    clear_error();
    GC_ARG(d_v && d_w, "NULL buffer");
    hipLaunchKernelGGL(axpby_kernel, stride_grid(n), dim3(256), 0, 0, d_v, a1, d_w, a2, n);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_svl_dot(float *d_result, const float *d_a, const float *d_b, uint64_t n) {
    clear_error();
    auto calls = lock_device_calls();  // shared scratch arena
    GC_ARG(d_result && d_a && d_b, "NULL buffer");
    const int G = (int)std::max<uint64_t>(1, std::min<uint64_t>((n + kCgThreads - 1) / kCgThreads,
                                                                kCgBlocks));
    void *ws = nullptr;
    int st = scratch_arena(G * sizeof(double), &ws);
    if (st) return st;
    hipLaunchKernelGGL(dot_partial_kernel, dim3(G), dim3(kCgThreads), 0, 0, d_a, d_b, n,
                       (double *)ws);
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kCgThreads), 0, 0, (const double *)ws, G,
                       d_result);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_svl_fillfrac(float *d_svl, const float *d_fillfrac, int32_t nx, int32_t ny,
                        int32_t nz) {
    clear_error();
    GC_ARG(d_svl && d_fillfrac, "NULL buffer");
    int st = check_dims(nx, ny, nz, 1);
    if (st) return st;
    const uint64_t n = (uint64_t)nx * ny * nz;
    hipLaunchKernelGGL(fillfrac_kernel, stride_grid(n), dim3(256), 0, 0, d_svl, d_fillfrac, n);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_svl_refine(const float *d_coarse, int32_t nx, int32_t ny, int32_t nz, float *d_fine,
                      int32_t fx, int32_t fy, int32_t fz, float hx, float hy, float hz) {
    clear_error();
    GC_ARG(d_coarse && d_fine, "NULL buffer");
    int st = check_dims(nx, ny, nz, 1);
    if (!st) st = check_dims(fx, fy, fz, 1);
    if (st) return st;
    GC_ARG(hx >= 0.f && hy >= 0.f && hz >= 0.f, "spacings must be non-negative");
    const Up u = {nx, ny, nz, fx, fy, fz, hx, hy, hz};
    const uint64_t n = (uint64_t)fx * fy * fz;
    hipLaunchKernelGGL(upsample_kernel, stride_grid(n), dim3(256), 0, 0, d_coarse, u, d_fine,
                       (float2 *)nullptr, 0);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_svl_grating(const float *d_phi, int32_t nx, int32_t ny, int32_t nz, gc_float2 *d_grating,
                       int32_t fx, int32_t fy, int32_t fz, float hx, float hy, float hz) {
    clear_error();
    GC_ARG(d_phi && d_grating, "NULL buffer");
    int st = check_dims(nx, ny, nz, 1);
    if (!st) st = check_dims(fx, fy, fz, 1);
    if (st) return st;
    GC_ARG(hx >= 0.f && hy >= 0.f && hz >= 0.f, "spacings must be non-negative");
    const Up u = {nx, ny, nz, fx, fy, fz, hx, hy, hz};
    const uint64_t n = (uint64_t)fx * fy * fz;
    hipLaunchKernelGGL(upsample_kernel, stride_grid(n), dim3(256), 0, 0, d_phi, u, (float *)nullptr,
                       (float2 *)d_grating, 1);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_svl_accumulate(float *d_svl, const gc_float2 *d_grating, uint64_t n, int32_t index,
                          const gc_float2 *d_coeffs) {
    clear_error();
    GC_ARG(d_svl && d_grating && d_coeffs, "NULL buffer");
    GC_ARG(n > 0 && index >= 0, "n must be positive and index non-negative");
    hipLaunchKernelGGL(svl_acc_kernel, stride_grid(n), dim3(256), 0, 0, d_svl,
                       (const float2 *)d_grating, n, (const float2 *)d_coeffs, index);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_svl_field(const gc_svl_params *p, const float *d_period, const gc_float2 *d_coeffs,
                     float *d_svl, gc_svl_result *result, void *stream) {
    clear_error();
    auto calls = lock_device_calls();  // shared scratch arena
    int st = check_params(p, d_period);
    if (st) return st;
    GC_ARG(d_coeffs && d_svl, "NULL buffer");
    GC_ARG(p->range >= 0 && p->range <= kMaxRange, "range must be in [0, %d]", kMaxRange);
    const int side = 2 * p->range + 1;
    GC_ARG(p->harmonics >= 0 && p->harmonics <= side * side * side,
           "harmonics must be in [0, (2 range + 1)^3]");
    GC_ARG(p->cg_max_iter >= 1, "cg_max_iter must be >= 1");
    GC_ARG(p->cg_end_res > 0.f, "cg_end_res must be positive");
    st = check_dims(p->fx, p->fy, p->fz, 1);
    if (st) return st;
    hipStream_t s = (hipStream_t)stream;
    const uint64_t n = (uint64_t)p->nx * p->ny * p->nz;
    // the three unit right-hand sides, solved together; a harmonic's residual
    // is at most (|i| + |j| + |k|) times theirs, so they are solved to
    // cg_end_res / (3 range) to meet cg_end_res for every harmonic
    void *ws = nullptr;
    const uint64_t vec = 3 * n * sizeof(float);
    hipError_t he = hipMalloc(&ws, vec);
    if (he != hipSuccess) return fail(GPUCAD_E_OOM, "svl: allocation of %llu bytes failed",
                                      (unsigned long long)vec);
    float *phi = (float *)ws;
    const int unit[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int a = 0; a < 3; ++a) {
        hipLaunchKernelGGL(finding_phi_kernel, stride_grid(n), dim3(256), 0, s, phi + a * n,
                           d_period, phi_args(p, unit[a][0], unit[a][1], unit[a][2]));
    }
    int iters[3] = {0, 0, 0};
    float res[3] = {0.f, 0.f, 0.f};
    he = hipGetLastError();
    if (he == hipSuccess) {
        const float tol = p->cg_end_res / (float)(3 * std::max(1, p->range));
        st = cg_solve(phi, phi, 3, p->nx, p->ny, p->nz, p->cg_max_iter, tol, iters, res, s);
    } else {
        st = fail(GPUCAD_E_HIP, "svl: launch failed: %s", hipGetErrorString(he));
    }
    if (!st) {
        const Up u = {p->nx, p->ny, p->nz, p->fx, p->fy, p->fz, p->fine_dx, p->fine_dy, p->fine_dz};
        const uint64_t fn = (uint64_t)p->fx * p->fy * p->fz;
        hipLaunchKernelGGL(svl_field_kernel, stride_grid(fn), dim3(256),
                           (size_t)std::max(1, p->harmonics) * (sizeof(float2) + sizeof(int)), s,
                           phi, phi + n,
                           phi + 2 * n, u, (const float2 *)d_coeffs, p->range, p->harmonics, d_svl);
        he = hipGetLastError();
        if (he != hipSuccess) st = fail(GPUCAD_E_HIP, "svl: launch failed: %s", hipGetErrorString(he));
        else {
            he = hipStreamSynchronize(s);
            if (he != hipSuccess) st = fail(GPUCAD_E_HIP, "svl: %s", hipGetErrorString(he));
        }
    }
    (void)hipFree(ws);
    if (st) return st;
    if (result) {
        for (int a = 0; a < 3; ++a) {
            result->cg_iters[a] = iters[a];
            result->cg_res[a] = res[a];
        }
    }
    return GPUCAD_OK;
}

int gpucad_unit_spectrum(const float *d_cell, int32_t nx, int32_t ny, int32_t nz, int32_t range,
                         gc_float2 *d_coeffs) {
    clear_error();
    auto calls = lock_device_calls();  // shared scratch arena
    GC_ARG(d_cell && d_coeffs, "NULL buffer");
    int st = check_dims(nx, ny, nz, 1);
    if (st) return st;
    GC_ARG(range >= 0 && range <= kMaxRange, "range must be in [0, %d]", kMaxRange);
    const int side = 2 * range + 1;
    // twiddles e^{-2 pi i m x / n} for m = -r..r on each axis, in double
    std::vector<double2> tw((uint64_t)side * (nx + ny + nz));
    const int dims[3] = {nx, ny, nz};
    uint64_t o = 0;
    for (int a = 0; a < 3; ++a) {
        for (int m = -range; m <= range; ++m) {
            for (int x = 0; x < dims[a]; ++x) {
                // reduce m x mod n first so the angle stays in one period
                const long long q = (((long long)m * x) % dims[a] + dims[a]) % dims[a];
                const double ang = -kTwoPi * (double)q / (double)dims[a];
                tw[o++] = make_double2(std::cos(ang), std::sin(ang));
            }
        }
    }
    void *ws = nullptr;
    st = scratch_arena(tw.size() * sizeof(double2), &ws);
    if (st) return st;
    GC_HIP(hipMemcpy(ws, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(dft_kernel, dim3((unsigned)(side * side * side)), dim3(256), 0, 0, d_cell,
                       nx, ny, nz, range, (const double2 *)ws, (float2 *)d_coeffs);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_fft_fill(const gc_float2 *d_half, gc_float2 *d_full, int32_t nx, int32_t ny,
                    int32_t nz) {
    clear_error();
    GC_ARG(d_half && d_full, "NULL buffer");
    int st = check_dims(nx, ny, nz, 1);
    if (st) return st;
    GC_ARG(nx == ny && ny == nz, "the mirror indices assume a cubic grid");
    const uint64_t n = (uint64_t)(nx / 2 + 1) * ny * nz;
    hipLaunchKernelGGL(fft_fill_kernel, stride_grid(n), dim3(256), 0, 0, (const float2 *)d_half,
                       (float2 *)d_full, nx, ny, nz);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_fft_scale(gc_float2 *d_data, float scalar, uint64_t n) {
    clear_error();
    GC_ARG(d_data, "NULL buffer");
    hipLaunchKernelGGL(fft_scale_kernel, stride_grid(n), dim3(256), 0, 0, (float2 *)d_data, scalar,
                       n);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

}  // extern "C"
