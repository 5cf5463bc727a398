// Unit-cell Fourier transforms of Fft_lattice (src/lattice_files/Fft_lattice.cu:
// 107-126, 255-289): in-place real-to-complex, complex-to-real and inverse
// complex-to-complex 3-D transforms of the shape gpucad_fft_plan3d records
// (the reference's cufftPlan3d calls, src/main.cu:3522-3526), plus the two
// element kernels of the unit-lattice view (add_fft_constants,
// indiviual_grating_sum).
//
// Semantics are cuFFT's (CUDA Toolkit 12.3, the reference's dependency):
// unnormalised; forward e^{-2 pi i jk/n}, inverse e^{+2 pi i jk/n}; n0 the
// slowest and n2 the fastest axis; the in-place real layout pads each row of
// n2 reals to 2 (n2/2 + 1) floats, the half spectrum keeps n2/2 + 1 complex
// values per row.  The complex-to-real pass uses X[n2 - k] = conj(X[k]) and
// ignores the imaginary parts of X[0] (and of X[n2/2] for even n2).
//
// Each axis is a direct DFT: one workgroup per line, the line and the twiddles
// e^{-2 pi i m/n} (double sincospi, m < n) in LDS, every output an FP64 sum
// over the line with the twiddle index jk mod n carried incrementally.  The
// unit cell of the reference is 61^3 (a prime edge): n0 n1 n2 (n0 + n1 + n2)
// complex products, 41.5 M for 61^3.  Axes up to 2048 points.
#include <mutex>

#include "gc_common.hpp"

namespace gc {
namespace {

constexpr int kMaxFftAxis = 2048;
constexpr int kMaxFftDevices = 64;

struct FftShape {
    int n0 = 0, n1 = 0, n2 = 0;  // slowest .. fastest; 0: no plan
};
FftShape g_plan[kMaxFftDevices];
std::mutex g_plan_mu;

int current_plan(FftShape *out) {
    int dev = 0;
    GC_HIP(hipGetDevice(&dev));
    GC_ARG(dev >= 0 && dev < kMaxFftDevices, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_plan_mu);
    GC_ARG(g_plan[dev].n0 > 0, "no transform shape: call gpucad_fft_plan3d first");
    *out = g_plan[dev];
    return GPUCAD_OK;
}

// twiddles w[m] = (cos, sin)(2 pi m / n) for m < n, into LDS
__device__ __forceinline__ void load_twiddles(double2 *w, int n) {
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
        double s, c;
        sincospi(2.0 * (double)m / (double)n, &s, &c);
        w[m] = make_double2(c, s);
    }
}

// One row of n2 reals (row pitch 2 (n2/2 + 1) floats) -> its n2/2 + 1
// forward coefficients, in place.
__global__ void dft_r2c_rows_kernel(float *data, int n2) {
    extern __shared__ double2 lds[];
    double2 *w = lds;
    double *x = (double *)(lds + n2);
    const int h = n2 / 2 + 1;
    float *row = data + (uint64_t)blockIdx.x * (2 * h);
    load_twiddles(w, n2);
    for (int j = threadIdx.x; j < n2; j += blockDim.x) x[j] = row[j];
    __syncthreads();
    float2 *out = (float2 *)row;
    for (int k = threadIdx.x; k < h; k += blockDim.x) {
        double re = 0.0, im = 0.0;
        int m = 0;
        for (int j = 0; j < n2; ++j) {
            re = fma(x[j], w[m].x, re);
            im = fma(-x[j], w[m].y, im);
            m += k;
            if (m >= n2) m -= n2;
        }
        out[k] = make_float2((float)re, (float)im);
    }
}

// n2/2 + 1 coefficients of a row -> the row's n2 reals (unnormalised
// inverse), in place.
__global__ void dft_c2r_rows_kernel(float *data, int n2) {
    extern __shared__ double2 lds[];
    double2 *w = lds;
    double2 *X = lds + n2;
    const int h = n2 / 2 + 1;
    float *row = data + (uint64_t)blockIdx.x * (2 * h);
    load_twiddles(w, n2);
    const float2 *in = (const float2 *)row;
    for (int k = threadIdx.x; k < h; k += blockDim.x)
        X[k] = make_double2((double)in[k].x, (double)in[k].y);
    __syncthreads();
    const int kh = (n2 - 1) / 2;  // coefficients with a distinct conjugate partner
    for (int j = threadIdx.x; j < n2; j += blockDim.x) {
        double acc = 0.0;
        int m = j;  // (j k) mod n2 for k = 1
        for (int k = 1; k <= kh; ++k) {
            // Re(X e^{+i a}) = X.re cos a - X.im sin a
            acc = fma(X[k].x, w[m].x, acc);
            acc = fma(-X[k].y, w[m].y, acc);
            m += j;
            if (m >= n2) m -= n2;
        }
        double v = X[0].x + 2.0 * acc;
        if ((n2 & 1) == 0) v += (j & 1) ? -X[n2 / 2].x : X[n2 / 2].x;
        row[j] = (float)v;  // the row's coefficients were all read into LDS above
    }
}

// Complex lines of length n along an axis of stride `stride` (elements):
// line l starts at (l / stride) n stride + l % stride.  SIGN = -1 forward,
// +1 inverse, in place.
template <int SIGN>
__global__ void dft_axis_kernel(float2 *data, int n, uint32_t stride) {
    extern __shared__ double2 lds[];
    double2 *w = lds;
    double2 *x = lds + n;
    const uint32_t l = blockIdx.x;
    float2 *base = data + (uint64_t)(l / stride) * n * stride + (l % stride);
    load_twiddles(w, n);
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const float2 v = base[(uint64_t)j * stride];
        x[j] = make_double2(v.x, v.y);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        double re = 0.0, im = 0.0;
        int m = 0;
        for (int j = 0; j < n; ++j) {
            // x e^{SIGN i a}: (xr c - SIGN xi s, xi c + SIGN xr s)
            const double c = w[m].x, s = SIGN * w[m].y;
            re = fma(x[j].x, c, re);
            re = fma(-x[j].y, s, re);
            im = fma(x[j].y, c, im);
            im = fma(x[j].x, s, im);
            m += k;
            if (m >= n) m -= n;
        }
        base[(uint64_t)k * stride] = make_float2((float)re, (float)im);
    }
}

int block_for(int n) { return n <= 64 ? 64 : 256; }

template <int SIGN>
int axis_pass(float2 *d, int n, uint64_t lines, uint32_t stride) {
    hipLaunchKernelGGL(dft_axis_kernel<SIGN>, dim3((uint32_t)lines), dim3(block_for(n)),
                       2 * n * sizeof(double2), 0, d, n, stride);
    GC_CHECK_LAUNCH();
    return GPUCAD_OK;
}

__global__ void set_element_kernel(float2 *a, float2 v, uint32_t i) {
    if (threadIdx.x == 0) a[i] = v;
}

__global__ void grating_sum_kernel(const float2 *g, float *out, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        out[i] += g[i].x;
}

}  // namespace
}  // namespace gc

using namespace gc;

extern "C" {

int gpucad_fft_plan3d(int32_t n0, int32_t n1, int32_t n2) {
    clear_error();
    GC_ARG(n0 >= 1 && n1 >= 1 && n2 >= 1, "transform sizes must be positive");
    GC_ARG(n0 <= kMaxFftAxis && n1 <= kMaxFftAxis && n2 <= kMaxFftAxis,
           "transform axes are limited to %d points", kMaxFftAxis);
    GC_ARG((uint64_t)n0 * n1 * (n2 / 2 + 1) < (1ull << 32), "transform too large");
    int dev = 0;
    GC_HIP(hipGetDevice(&dev));
    GC_ARG(dev >= 0 && dev < kMaxFftDevices, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g_plan[dev].n0 = n0;
    g_plan[dev].n1 = n1;
    g_plan[dev].n2 = n2;
    return GPUCAD_OK;
}

int gpucad_fft_destroy(void) {
    clear_error();
    int dev = 0;
    GC_HIP(hipGetDevice(&dev));
    GC_ARG(dev >= 0 && dev < kMaxFftDevices, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g_plan[dev] = FftShape();
    return GPUCAD_OK;
}

int gpucad_fft_r2c(gc_float2 *data) {
    clear_error();
    GC_ARG(data, "NULL buffer");
    FftShape p;
    int r = current_plan(&p);
    if (r) return r;
    const int h = p.n2 / 2 + 1;
    hipLaunchKernelGGL(dft_r2c_rows_kernel, dim3((uint32_t)p.n0 * p.n1), dim3(block_for(h)),
                       p.n2 * (sizeof(double2) + sizeof(double)), 0, (float *)data, p.n2);
    GC_CHECK_LAUNCH();
    float2 *d = (float2 *)data;
    if ((r = axis_pass<-1>(d, p.n1, (uint64_t)p.n0 * h, (uint32_t)h))) return r;
    if ((r = axis_pass<-1>(d, p.n0, (uint64_t)p.n1 * h, (uint32_t)(p.n1 * h)))) return r;
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_fft_c2r(gc_float2 *data) {
    clear_error();
    GC_ARG(data, "NULL buffer");
    FftShape p;
    int r = current_plan(&p);
    if (r) return r;
    const int h = p.n2 / 2 + 1;
    float2 *d = (float2 *)data;
    if ((r = axis_pass<1>(d, p.n0, (uint64_t)p.n1 * h, (uint32_t)(p.n1 * h)))) return r;
    if ((r = axis_pass<1>(d, p.n1, (uint64_t)p.n0 * h, (uint32_t)h))) return r;
    hipLaunchKernelGGL(dft_c2r_rows_kernel, dim3((uint32_t)p.n0 * p.n1), dim3(block_for(p.n2)),
                       (p.n2 + h) * sizeof(double2), 0, (float *)data, p.n2);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_fft_c2c_inverse(gc_float2 *data) {
    clear_error();
    GC_ARG(data, "NULL buffer");
    FftShape p;
    int r = current_plan(&p);
    if (r) return r;
    float2 *d = (float2 *)data;
    const uint64_t n = (uint64_t)p.n0 * p.n1 * p.n2;
    if ((r = axis_pass<1>(d, p.n2, n / p.n2, 1))) return r;
    if ((r = axis_pass<1>(d, p.n1, n / p.n1, (uint32_t)p.n2))) return r;
    if ((r = axis_pass<1>(d, p.n0, n / p.n0, (uint32_t)(p.n1 * p.n2)))) return r;
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_add_fft_constant(gc_float2 *grating, gc_float2 value, int32_t index) {
    clear_error();
    GC_ARG(grating, "NULL buffer");
    GC_ARG(index >= 0, "index must be >= 0");
    hipLaunchKernelGGL(set_element_kernel, dim3(1), dim3(64), 0, 0, (float2 *)grating,
                       make_float2(value.x, value.y), (uint32_t)index);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

int gpucad_grating_sum(const gc_float2 *sum_gratings, float *fft_function, int32_t nx, int32_t ny,
                       int32_t nz) {
    clear_error();
    GC_ARG(sum_gratings && fft_function, "NULL buffer");
    GC_ARG(nx >= 1 && ny >= 1 && nz >= 1, "grid sizes must be positive");
    const uint64_t n = (uint64_t)nx * ny * nz;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(grating_sum_kernel, dim3(blocks), dim3(256), 0, 0,
                       (const float2 *)sum_gratings, fft_function, n);
    GC_CHECK_LAUNCH();
    GC_HIP(hipStreamSynchronize(0));
    return GPUCAD_OK;
}

}  // extern "C"
