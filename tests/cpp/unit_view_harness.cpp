// Replays the reference caller's unit-lattice steps through the drop-in
// classes of gpucad_compat.hpp, in the reference's order and with its buffer
// layouts:
//   fft_init / lattice_init (src/main.cu:3500-3540; the three cufftPlan3d
//   calls become gpucad_fft_plan3d),
//   unit_lattice (3578-3711): create_lattice -> pitched copy -> fft_func ->
//   fft_scalar of a copy -> fft_fill -> ifft_func -> the (2r+1)^3
//   coefficient pick on the host -> pitched copy back -> normalise_buffer ->
//   normalise_four,
//   the approximate view (3720-3895) without the frame drawing: for the first
//   half of the coefficients, a zeroed complex grating, add_fft_constants,
//   ifft_func_complex, indiviual_grating_sum, normalise_buffer; then the
//   final normalise_buffer and normalise_four.
// Writes the outputs as raw little-endian arrays into OUT_DIR:
//   lattice_data.bin   (2r+1)^3 float2 coefficients
//   unit_norm.bin      fft_gratings after unit_lattice (normalised field)
//   unit_ind.bin       d_volumeone_one after unit_lattice (indicator)
//   approx_sum.bin     individual_grating_one after the loop (raw sum)
//   approx_norm.bin    fft_gratings after the approximate view
//   approx_ind.bin     d_volumeone_one after the approximate view
// and prints "coefficients <l_count> gratings <r_count>".
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "gpucad_compat.hpp"

static void *dev_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc of %zu bytes failed\n", bytes);
        std::exit(EXIT_FAILURE);
    }
    return p;
}

static void hip_ok(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s failed: %s\n", what, hipGetErrorString(e));
        std::exit(EXIT_FAILURE);
    }
}

static void dump(const std::string &path, const void *d, size_t bytes) {
    std::vector<char> h(bytes);
    hip_ok(hipMemcpy(h.data(), d, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(h.data(), 1, bytes, f) != bytes) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(EXIT_FAILURE);
    }
    std::fclose(f);
}

// u = mid + i mapped to the unshifted spectrum index (src/main.cu:3638-3684)
static int unshift(int i, int mid) {
    int u = mid + i;
    if (u == mid) return 0;
    if (u > mid) return u - mid;
    return u + (mid + 1);
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s OUT_DIR LATTICE_TYPE\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    const uint lattice_index_type = (uint)std::atoi(argv[2]);
    const int Nxu = 61, Nyu = 61, Nzu = 61;
    const uint sizeone = Nxu * Nyu * Nzu;
    const int range_st = 2, indi_range = 2 * range_st + 1;
    const float bound_isoValone = 0.20f, bound_isoValtwo = 0.30f;
    Fft_lattice fftlattice;
    Gratings lattice;

    float *d_volumeone = (float *)dev_alloc(sizeof(float) * sizeone);
    // fft_init
    float2 *fft_data = (float2 *)dev_alloc(sizeof(float2) * ((Nxu / 2) + 1) * Nyu * Nzu);
    float2 *fft_data_compute = (float2 *)dev_alloc(sizeof(float2) * ((Nxu / 2) + 1) * Nyu * Nzu);
    float2 *fft_data_compute_fill = (float2 *)dev_alloc(sizeof(float2) * Nxu * Nyu * Nzu);
    hip_ok(hipMemset(fft_data, 0, sizeof(float2) * ((Nxu / 2) + 1) * Nyu * Nzu), "memset");
    hip_ok(hipMemset(fft_data_compute, 0, sizeof(float2) * ((Nxu / 2) + 1) * Nyu * Nzu), "memset");
    hip_ok(hipMemset(fft_data_compute_fill, 0, sizeof(float2) * Nxu * Nyu * Nzu), "memset");
    // lattice_init
    gpucad_compat::ensure_init();
    gpucad_compat::check(gpucad_fft_plan3d(Nxu, Nyu, Nzu), "gpucad_fft_plan3d");
    float *fft_gratings = (float *)dev_alloc(sizeof(float) * sizeone);
    float2 *lattice_data = (float2 *)dev_alloc(sizeof(float2) * indi_range * indi_range * indi_range);
    float *d_volumeone_one = (float *)dev_alloc(sizeof(float) * sizeone);
    hip_ok(hipMemset(fft_gratings, 0, sizeof(float) * sizeone), "memset");
    hip_ok(hipMemset(lattice_data, 0, sizeof(float2) * indi_range * indi_range * indi_range),
           "memset");
    hip_ok(hipMemset(d_volumeone_one, 0, sizeof(float) * sizeone), "memset");

    // ---- unit_lattice
    fftlattice.create_lattice(d_volumeone, Nxu, Nyu, Nzu, sizeone, lattice_index_type);
    hip_ok(hipMemcpy2D(fft_data, (Nxu + 1) * sizeof(float), d_volumeone, Nxu * sizeof(float),
                       Nxu * sizeof(float), Nyu * Nzu, hipMemcpyDeviceToDevice),
           "hipMemcpy2D");
    fftlattice.fft_func(fft_data);
    hip_ok(hipMemcpy(fft_data_compute, fft_data, ((Nxu / 2) + 1) * Nyu * Nzu * sizeof(float2),
                     hipMemcpyDeviceToDevice),
           "hipMemcpy");
    fftlattice.fft_scalar(fft_data_compute, sizeone, ((Nxu / 2) + 1) * Nyu * Nzu);
    const uint mid_index = (uint)std::floor((Nxu * Nyu * Nzu) / 2.0);
    fftlattice.fft_fill(fft_data_compute, fft_data_compute_fill, Nxu, Nyu, Nzu, mid_index);
    fftlattice.ifft_func(fft_data);
    std::vector<float2> h_dat((size_t)Nxu * Nyu * Nzu);
    hip_ok(hipMemcpy(h_dat.data(), fft_data_compute_fill, h_dat.size() * sizeof(float2),
                     hipMemcpyDeviceToHost),
           "hipMemcpy");
    std::vector<float2> h_lattice_dat((size_t)indi_range * indi_range * indi_range);
    const int mid = Nxu / 2;
    const int midd2 = (int)std::floor((indi_range * indi_range * indi_range) / 2.0);
    int l_count = 0;
    for (int k = -range_st; k < range_st + 1; ++k)
        for (int j = -range_st; j < range_st + 1; ++j)
            for (int i = -range_st; i < range_st + 1; ++i) {
                if (i == 0 && j == 0 && k == 0) {
                    h_lattice_dat[midd2] = h_dat[0];
                    ++l_count;
                    continue;
                }
                const int indx2 = unshift(i, mid) + unshift(j, mid) * Nxu +
                                  unshift(k, mid) * (Nxu * Nyu);
                h_lattice_dat[l_count++] = h_dat[indx2];
            }
    hip_ok(hipMemcpy2D(fft_gratings, Nxu * sizeof(float), fft_data,
                       (((Nxu / 2) + 1) * 2) * sizeof(float), Nxu * sizeof(float), Nyu * Nzu,
                       hipMemcpyDeviceToDevice),
           "hipMemcpy2D");
    lattice.GPU_buffer_normalise_buffer(fft_gratings, fft_gratings, Nxu * Nyu * Nzu);
    lattice.GPU_buffer_normalise_four(fft_gratings, d_volumeone_one, fft_gratings,
                                      Nxu * Nyu * Nzu, Nxu, Nyu, Nzu, bound_isoValone,
                                      bound_isoValtwo);
    hip_ok(hipMemcpy(lattice_data, h_lattice_dat.data(), sizeof(float2) * h_lattice_dat.size(),
                     hipMemcpyHostToDevice),
           "hipMemcpy");
    dump(out + "/lattice_data.bin", lattice_data, sizeof(float2) * h_lattice_dat.size());
    dump(out + "/unit_norm.bin", fft_gratings, sizeof(float) * sizeone);
    dump(out + "/unit_ind.bin", d_volumeone_one, sizeof(float) * sizeone);

    // ---- the approximate view (its input is d_volumeone, still the unit cell)
    hip_ok(hipMemset(fft_gratings, 0, sizeof(float) * sizeone), "memset");
    float2 *individual_grating = (float2 *)dev_alloc(sizeof(float2) * sizeone);
    float *individual_grating_one = (float *)dev_alloc(sizeof(float) * sizeone);
    hip_ok(hipMemset(individual_grating_one, 0, sizeof(float) * sizeone), "memset");
    hip_ok(hipMemcpy2D(fft_data, (Nxu + 1) * sizeof(float), d_volumeone, Nxu * sizeof(float),
                       Nxu * sizeof(float), Nyu * Nzu, hipMemcpyDeviceToDevice),
           "hipMemcpy2D");
    fftlattice.fft_func(fft_data);
    hip_ok(hipMemcpy(fft_data_compute, fft_data, ((Nxu / 2) + 1) * Nyu * Nzu * sizeof(float2),
                     hipMemcpyDeviceToDevice),
           "hipMemcpy");
    fftlattice.fft_scalar(fft_data_compute, sizeone, ((Nxu / 2) + 1) * Nyu * Nzu);
    fftlattice.fft_fill(fft_data_compute, fft_data_compute_fill, Nxu, Nyu, Nzu, mid_index);
    fftlattice.ifft_func(fft_data);
    hip_ok(hipMemcpy(h_dat.data(), fft_data_compute_fill, h_dat.size() * sizeof(float2),
                     hipMemcpyDeviceToHost),
           "hipMemcpy");
    l_count = 0;
    int r_count = 0;
    for (int k = -range_st; k < range_st + 1; ++k)
        for (int j = -range_st; j < range_st + 1; ++j)
            for (int i = -range_st; i < range_st + 1; ++i) {
                if (l_count < midd2) {
                    hip_ok(hipMemset(individual_grating, 0, sizeof(float2) * sizeone), "memset");
                    const int indx2 = unshift(i, mid) + unshift(j, mid) * Nxu +
                                      unshift(k, mid) * (Nxu * Nyu);
                    h_lattice_dat[l_count] = h_dat[indx2];
                    fftlattice.add_fft_constants(individual_grating, h_dat[indx2], indx2);
                    fftlattice.ifft_func_complex(individual_grating);
                    fftlattice.indiviual_grating_sum(individual_grating, individual_grating_one,
                                                     Nxu, Nyu, Nzu);
                    lattice.GPU_buffer_normalise_buffer(individual_grating_one, fft_gratings,
                                                        Nxu * Nyu * Nzu);
                    ++r_count;
                }
                ++l_count;
            }
    dump(out + "/approx_sum.bin", individual_grating_one, sizeof(float) * sizeone);
    hip_ok(hipMemcpy(fft_gratings, individual_grating_one, sizeof(float) * sizeone,
                     hipMemcpyDeviceToDevice),
           "hipMemcpy");
    lattice.GPU_buffer_normalise_buffer(fft_gratings, fft_gratings, Nxu * Nyu * Nzu);
    lattice.GPU_buffer_normalise_four(fft_gratings, d_volumeone_one, fft_gratings,
                                      Nxu * Nyu * Nzu, Nxu, Nyu, Nzu, bound_isoValone,
                                      bound_isoValtwo);
    dump(out + "/approx_norm.bin", fft_gratings, sizeof(float) * sizeone);
    dump(out + "/approx_ind.bin", d_volumeone_one, sizeof(float) * sizeone);
    std::printf("coefficients %d gratings %d\n", l_count, r_count);
    gpucad_compat::check(gpucad_fft_destroy(), "gpucad_fft_destroy");
    for (void *p : {(void *)d_volumeone, (void *)fft_data, (void *)fft_data_compute,
                    (void *)fft_data_compute_fill, (void *)fft_gratings, (void *)lattice_data,
                    (void *)d_volumeone_one, (void *)individual_grating,
                    (void *)individual_grating_one})
        hip_ok(hipFree(p), "hipFree");
    return 0;
}
