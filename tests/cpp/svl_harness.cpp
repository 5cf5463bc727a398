// Replays the reference caller's spatial-lattice loop through the drop-in
// classes of gpucad_compat.hpp (src/main.cu:3544-3576 init_svl,
// 3578-3711 unit_lattice, 3904-4037 spatial_lattice_run): angle and period
// data, normalise_three, then for each of the first half of the 5^3
// coefficients finding_phi -> GPUCG_lattice(500, 1, 0.01) -> copytotexture ->
// updateTexture -> grating -> svl -> GPU_buffer_normalise_four, and after the
// loop computeIsosurface_lattice and file_write_obj.  The unit cell's
// coefficients come from gpucad_unit_spectrum (the reference takes them
// from a cuFFT of the same cell).
//
// usage: svl_harness OUT_DIR N PERIOD LATTICE_TYPE
// writes OUT_DIR/coeffs.bin (125 complex), OUT_DIR/svl.bin ((2N)^3 float32),
// OUT_DIR/spatial_lattice.obj; prints one line per harmonic
// "h iterations residual" and a final line "active totalVerts".
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
    if (hipMemset(p, 0, bytes) != hipSuccess) std::exit(EXIT_FAILURE);
    return p;
}

static void save(const std::string &path, const void *d, size_t bytes) {
    std::vector<char> h(bytes);
    if (hipMemcpy(h.data(), d, bytes, hipMemcpyDeviceToHost) != hipSuccess) std::exit(1);
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(h.data(), 1, bytes, f) != bytes) std::exit(1);
    std::fclose(f);
}

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s OUT_DIR N PERIOD LATTICE_TYPE\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    const int N = std::atoi(argv[2]);
    const float period_of_grating = (float)std::atof(argv[3]);
    const char latticetype_one = argv[4][0];
    const int NumX = N, NumY = N, NumZ = N, NumX2 = 2 * N, NumY2 = 2 * N, NumZ2 = 2 * N;
    const size_t size = (size_t)NumX * NumY * NumZ, size2 = (size_t)NumX2 * NumY2 * NumZ2;
    const float dx = 1.0f, dy = 1.0f, dz = 1.0f, dx2 = 0.5f * dx, dy2 = 0.5f * dy, dz2 = 0.5f * dz;
    const uint Nxu = 61, Nyu = 61, Nzu = 61;
    const int indi_range = 5, range_st = 2;

    Isosurface isosurf;
    Fft_lattice fftlattice;
    Gratings lattice;
    File_output output_file;
    uint *d_triTable = nullptr, *d_numVertsTable = nullptr;
    isosurf.allocateTextures_s(&d_triTable, &d_numVertsTable);
    lattice.NX = NumX;
    lattice.NY = NumY;
    lattice.NZ = NumZ;

    // init_svl
    float *d_phi = (float *)dev_alloc(sizeof(float) * size);
    float *d_theta = (float *)dev_alloc(sizeof(float) * size);
    float *d_period = (float *)dev_alloc(sizeof(float) * size);
    float2 *d_ga = (float2 *)dev_alloc(sizeof(float2) * size2);
    float *d_svl = (float *)dev_alloc(sizeof(float) * size2);
    float *d_volumethree = (float *)dev_alloc(sizeof(float) * size2);
    float *d_volumethree_one = (float *)dev_alloc(sizeof(float) * size2);
    float *d_volumethree_two = (float *)dev_alloc(sizeof(float) * size2);
    hipPitchedPtr devPitchedPtr;
    if (hipMalloc3D(&devPitchedPtr, make_hipExtent(NumX * sizeof(float), NumY, NumZ)) !=
        hipSuccess)
        return 1;
    lattice.setupTexture(NumX, NumY, NumZ);

    // unit_lattice: the unit cell and its 5^3 coefficients (lattice_data)
    float *d_cell = (float *)dev_alloc(sizeof(float) * Nxu * Nyu * Nzu);
    fftlattice.create_lattice(d_cell, Nxu, Nyu, Nzu, Nxu * Nyu * Nzu, 0);
    float2 *lattice_data = (float2 *)dev_alloc(sizeof(float2) * indi_range * indi_range * indi_range);
    gpucad_compat::check(gpucad_unit_spectrum(d_cell, Nxu, Nyu, Nzu, range_st,
                                              (gc_float2 *)lattice_data),
                         "unit_spectrum");
    save(out + "/coeffs.bin", lattice_data, sizeof(float2) * 125);

    // spatial_lattice_run
    float mn_x = 0, mn_y = 0, mn_z = 0;
    if (latticetype_one == 'r') {
        mn_x = NumX / 2.0f;
        mn_y = NumY / 2.0f;
        mn_z = NumZ / 2.0f;
    }
    lattice.angle_data(d_theta, NumX, NumY, NumZ, dx, dy, dz, mn_x, mn_y, mn_z, 'z');
    lattice.period_data(d_period, NumX, NumY, NumZ, dx, dy, dz, mn_x, mn_y, mn_z, 'z');
    lattice.GPU_buffer_normalise_three(d_period, d_period, size, NumX / 10, NumX / 4);
    const int half_range = (indi_range * indi_range * indi_range) / 2;
    const float bound_isoValone = 0.20f, bound_isoValtwo = 0.30f;
    float bound_isoVal = 0.25f;  // taken by reference, as ImguiApp::bound_isoVal
    int mycount = 0;
    for (int k = -range_st; k < range_st + 1; k++)
        for (int j = -range_st; j < range_st + 1; j++)
            for (int i = -range_st; i < range_st + 1; i++) {
                if (mycount < half_range) {
                    int FinalIter_l = 0;
                    float FinalRes_l = 0.f;
                    lattice.finding_phi(d_phi, d_period, NumX, NumY, NumZ, i, j, k, dx, dy, dz,
                                        latticetype_one, 0, period_of_grating, 30.0f, 30.0f,
                                        30.0f, 0.16f, 0.6f, false);
                    lattice.GPUCG_lattice(d_phi, 500, 1, 0.01f, FinalIter_l, FinalRes_l);
                    lattice.copytotexture(d_phi, devPitchedPtr, NumX, NumY, NumZ);
                    lattice.updateTexture(devPitchedPtr);
                    lattice.grating(d_ga, NumX2, NumY2, NumZ2, dx2, dy2, dz2);
                    lattice.svl(d_svl, d_ga, NumX2, NumY2, NumZ2, mycount, lattice_data);
                    lattice.GPU_buffer_normalise_four(d_svl, d_volumethree_one, d_volumethree,
                                                      size2, NumX2, NumY2, NumZ2,
                                                      bound_isoValone, bound_isoValtwo);
                    std::printf("%d %d %g\n", mycount, FinalIter_l, FinalRes_l);
                }
                mycount++;
            }
    save(out + "/svl.bin", d_svl, sizeof(float) * size2);

    // extraction of the band lattice, as after each harmonic in the reference
    uint3 gridSize = make_uint3(NumX2, NumY2, NumZ2);
    uint3 gridSizeMask = make_uint3(NumX2 - 1, NumY2 - 1, NumZ2 - 1);
    uint3 gridSizeShift = make_uint3(1, NumX2 - 1, (NumX2 - 1) * (NumY2 - 1));
    uint numVoxels = gridSizeMask.x * gridSizeMask.y * gridSizeMask.z;
    float3 voxelSize = make_float3(dx2, dy2, dz2);
    float3 gridcenter = make_float3(0.0f, 0.0f, 0.0f);
    uint maxVerts = (uint)(size2 * 4 > 300000 ? size2 * 4 : 300000);
    uint *d_vv = (uint *)dev_alloc(numVoxels * sizeof(uint));
    uint *d_vvs = (uint *)dev_alloc(numVoxels * sizeof(uint));
    uint *d_vo = (uint *)dev_alloc(numVoxels * sizeof(uint));
    uint *d_vos = (uint *)dev_alloc(numVoxels * sizeof(uint));
    uint *d_comp = (uint *)dev_alloc(numVoxels * sizeof(uint));
    float4 *d_pos = (float4 *)dev_alloc((size_t)maxVerts * sizeof(float4));
    float4 *d_norm = (float4 *)dev_alloc((size_t)maxVerts * sizeof(float4));
    uint activeVoxels = 0, totalVerts = 0;
    isosurf.computeIsosurface_lattice(d_volumethree_one, d_pos, d_norm, bound_isoVal, numVoxels,
                                      d_vv, d_vvs, d_vo, d_vos, gridSize, gridSizeShift,
                                      gridSizeMask, voxelSize, gridcenter, &activeVoxels,
                                      &totalVerts, d_comp, maxVerts, d_volumethree,
                                      d_volumethree_two, bound_isoValone, bound_isoValtwo, 0.0f,
                                      0.0f);
    output_file.file_write_obj(d_pos, totalVerts, (out + "/spatial_lattice.obj").c_str());
    std::printf("%u %u\n", activeVoxels, totalVerts);
    lattice.deleteTexture();
    isosurf.destroyAllTextureObjects();
    return 0;
}
