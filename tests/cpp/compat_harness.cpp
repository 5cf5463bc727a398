// Replays the reference caller's unit-lattice sequence through the drop-in
// classes of gpucad_compat.hpp (src/main.cu:2040-2078 table upload,
// 3578-3711 unit lattice, 4106-4137 display, 4695-4778 export):
//   allocateTextures_s -> Fft_lattice::create_lattice (61^3) ->
//   Gratings::GPU_buffer_normalise_four (normalised in place) ->
//   Isosurface::computeIsosurface_latticeone -> File_output::file_write_obj
// and prints "type active total" per lattice type.  Used by the GPU tests,
// which repeat the sequence through the C-ABI and compare outputs.
#include <cstdio>
#include <string>

#include "gpucad_compat.hpp"

static void *dev_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc of %zu bytes failed\n", bytes);
        std::exit(EXIT_FAILURE);
    }
    return p;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s OUT_DIR\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    const uint Nxu = 61, Nyu = 61, Nzu = 61, size = Nxu * Nyu * Nzu;
    Isosurface isosurf;
    Fft_lattice fftlattice;
    Gratings lattice;
    File_output output_file;
    uint *d_triTable = nullptr, *d_numVertsTable = nullptr;
    isosurf.allocateTextures_s(&d_triTable, &d_numVertsTable);

    // initMC for the unit cell grid
    uint3 gridSize = make_uint3(Nxu, Nyu, Nzu);
    uint3 gridSizeMask = make_uint3(Nxu - 1, Nyu - 1, Nzu - 1);
    uint3 gridSizeShift = make_uint3(1, Nxu - 1, (Nxu - 1) * (Nyu - 1));
    uint numVoxels = gridSizeMask.x * gridSizeMask.y * gridSizeMask.z;
    float3 voxelSize = make_float3(1.0f, 1.0f, 1.0f);
    float3 gridcenter = make_float3(0.0f, 0.0f, 0.0f);
    uint maxVerts = size * 4 > 300000 ? size * 4 : 300000;
    uint *d_voxelVerts, *d_voxelVertsScan, *d_voxelOccupied, *d_voxelOccupiedScan, *d_comp;
    float *d_volumeone, *d_volumeone_one;
    float4 *d_pos, *d_norm;
    d_voxelVerts = (decltype(d_voxelVerts))dev_alloc(numVoxels * sizeof(uint));
    d_voxelVertsScan = (decltype(d_voxelVertsScan))dev_alloc(numVoxels * sizeof(uint));
    d_voxelOccupied = (decltype(d_voxelOccupied))dev_alloc(numVoxels * sizeof(uint));
    d_voxelOccupiedScan = (decltype(d_voxelOccupiedScan))dev_alloc(numVoxels * sizeof(uint));
    d_comp = (decltype(d_comp))dev_alloc(numVoxels * sizeof(uint));
    d_volumeone = (decltype(d_volumeone))dev_alloc(size * sizeof(float));
    d_volumeone_one = (decltype(d_volumeone_one))dev_alloc(size * sizeof(float));
    d_pos = (decltype(d_pos))dev_alloc((size_t)maxVerts * sizeof(float4));
    d_norm = (decltype(d_norm))dev_alloc((size_t)maxVerts * sizeof(float4));
    float bound_isoVal = 0.25f, bound_isoValone = 0.20f, bound_isoValtwo = 0.30f;
    for (uint type = 0; type < 6; ++type) {
        uint activeVoxels = 0, totalVerts = 0;
        fftlattice.create_lattice(d_volumeone, Nxu, Nyu, Nzu, size, type);
        lattice.GPU_buffer_normalise_four(d_volumeone, d_volumeone_one, d_volumeone, size, Nxu,
                                          Nyu, Nzu, bound_isoValone, bound_isoValtwo);
        isosurf.computeIsosurface_latticeone(
            d_volumeone_one, d_pos, d_norm, bound_isoVal, numVoxels, d_voxelVerts,
            d_voxelVertsScan, d_voxelOccupied, d_voxelOccupiedScan, gridSize, gridSizeShift,
            gridSizeMask, voxelSize, gridcenter, &activeVoxels, &totalVerts, d_comp, maxVerts,
            d_volumeone, bound_isoValone, bound_isoValtwo);
        std::string path = out + "/unit_lattice_" + std::to_string(type) + ".obj";
        output_file.file_write_obj(d_pos, totalVerts, path.c_str());
        std::printf("%u %u %u\n", type, activeVoxels, totalVerts);
    }
    isosurf.destroyAllTextureObjects();
    return 0;
}
