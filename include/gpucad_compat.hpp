// gpucad_compat.hpp — the reference's C++ operator API on top of libgpucad.
//
// Header-only C++ classes with the method signatures of the reference's
// hot-path classes, so that the reference's caller (class Multitopo,
// src/main.cu) makes its hot-path method calls unchanged and runs the MI355X
// kernels.  What the header does not replace is the caller's own use of the
// CUDA runtime and libraries around those calls (allocation, copies, events,
// the Vulkan interop, and the three cufftPlan3d / cufftDestroy calls, which
// become gpucad_fft_plan3d / gpucad_fft_destroy):
//
//   MarchingCubeCuda   src/MarchingCubes_kernel.h:34-124 (path methods)
//   Isosurface         src/Isosurface.h:10-80 (lattice and boolean extraction)
//   Modelling          src/Modelling.h:10-48 (primitive fields)
//   Fft_lattice        src/lattice_files/Fft_lattice.h:10-24 (create_lattice,
//                      the unit cell's transforms, fft_scalar, fft_fill,
//                      indiviual_grating_sum, add_fft_constants)
//   Interpolations     src/Interpolations.h:10-40 (the phase texture of the
//                      lattice synthesis)
//   Gratings           src/lattice_files/Gratings.h:10-80 (lattice synthesis,
//                      normalise, masks)
//   File_output        src/File_output.h:35-44 (file_write_obj)
//
// Semantics follow the reference: every call is synchronous; launch-shape
// arguments (dim3 grid, threads) are accepted and ignored (the library picks
// its own shapes); a failing call prints the library's message and exits the
// process, as the reference's checkCudaErrors does
// (src/commons/helper_cuda.h:583-611).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>

#include "gpucad.h"

typedef unsigned int uint;

// src/MarchingCubes_kernel.h:12-18
struct grid_points {
    int val = 0;
    float t_x = 0.0;
    float t_y = 0.0;
    float t_z = 0.0;
};
static_assert(sizeof(grid_points) == 16, "grid_points is a 16-byte record");

namespace gpucad_compat {

inline void check(int status, const char *what) {
    if (status != GPUCAD_OK) {
        std::fprintf(stderr, "gpucad error in %s: %s (%s)\n", what,
                     gpucad_status_string(status), gpucad_last_error());
        std::exit(EXIT_FAILURE);
    }
}

inline gc_uint3 u3(uint3 v) { return gc_uint3{v.x, v.y, v.z}; }
inline gc_float3 f3(float3 v) { return gc_float3{v.x, v.y, v.z}; }

inline uint32_t bool_flags(bool obj_union, bool obj_diff, bool obj_intersect, bool fixed,
                           bool dynamic, bool make_region) {
    return (obj_union ? GC_BOOL_UNION : 0u) | (obj_diff ? GC_BOOL_DIFF : 0u) |
           (obj_intersect ? GC_BOOL_INTERSECT : 0u) | (fixed ? GC_BOOL_FIXED : 0u) |
           (dynamic ? GC_BOOL_DYNAMIC : 0u) | (make_region ? GC_BOOL_MAKE_REGION : 0u);
}

inline void ensure_init() {
    static bool done = false;
    if (!done) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        check(gpucad_init(dev), "gpucad_init");
        done = true;
    }
}

}  // namespace gpucad_compat

class MarchingCubeCuda {
public:
    // Uploads the case tables (uint, 255-terminated rows, as src/tables.h)
    // and returns their device copies; also initialises the library.
    void allocateTextures_s(uint **d_triTable, uint **d_numVertsTable) {
        using namespace gpucad_compat;
        ensure_init();
        uint8_t tri[256 * 16], nv[256];
        check(gpucad_mc_tables(tri, nv, nullptr), "gpucad_mc_tables");
        uint tri32[256 * 16], nv32[256];
        for (int i = 0; i < 256 * 16; ++i) tri32[i] = tri[i];
        for (int i = 0; i < 256; ++i) nv32[i] = nv[i];
        if (hipMalloc((void **)d_triTable, sizeof(tri32)) != hipSuccess ||
            hipMalloc((void **)d_numVertsTable, sizeof(nv32)) != hipSuccess ||
            hipMemcpy(*d_triTable, tri32, sizeof(tri32), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(*d_numVertsTable, nv32, sizeof(nv32), hipMemcpyHostToDevice) != hipSuccess)
            check(GPUCAD_E_HIP, "allocateTextures_s");
    }

    void destroyAllTextureObjects() { gpucad_compat::check(gpucad_shutdown(), "gpucad_shutdown"); }

    void ThrustScanWrapper_lattice(unsigned int *output, unsigned int *input,
                                   unsigned int numElements) {
        gpucad_compat::check(gpucad_exclusive_scan_u32(output, input, numElements),
                             "ThrustScanWrapper_lattice");
    }

    void classifyVoxel_lattice_new(dim3, dim3, uint *voxelVerts, uint *voxelOccupied,
                                   float *volume, uint3 gridSize, uint3 gridSizeShift,
                                   uint3 gridSizeMask, uint numVoxels, float3 voxelSize,
                                   float isoValue) {
        using namespace gpucad_compat;
        ensure_init();
        check(gpucad_classify_voxel_lattice_new(voxelVerts, voxelOccupied, volume, u3(gridSize),
                                                u3(gridSizeShift), u3(gridSizeMask), numVoxels,
                                                f3(voxelSize), isoValue),
              "classifyVoxel_lattice_new");
    }

    void compactVoxels_lattice(dim3, dim3, uint *compactedVoxelArray, uint *voxelOccupied,
                               uint *voxelOccupiedScan, uint numVoxels) {
        gpucad_compat::check(gpucad_compact_voxels(compactedVoxelArray, voxelOccupied,
                                                   voxelOccupiedScan, numVoxels),
                             "compactVoxels_lattice");
    }

    void generateTriangles_lattice_new(dim3, dim3, float4 *pos, float4 *norm,
                                       uint *compactedVoxelArray, uint *numVertsScanned,
                                       float *volume, uint3 gridSize, uint3 gridSizeShift,
                                       uint3 gridSizeMask, float3 voxelSize, float3 gridcenter,
                                       float isoValue, uint activeVoxels, uint maxVerts,
                                       uint totalverts, float *volume_one, float *volume_two,
                                       float isovalue1, float isovalue2, float iso1,
                                       float iso2) {
        using namespace gpucad_compat;
        ensure_init();
        check(gpucad_generate_triangles_lattice_new(
                  (gc_float4 *)pos, (gc_float4 *)norm, compactedVoxelArray, numVertsScanned,
                  volume, u3(gridSize), u3(gridSizeShift), u3(gridSizeMask), f3(voxelSize),
                  f3(gridcenter), isoValue, activeVoxels, maxVerts, totalverts, volume_one,
                  volume_two, isovalue1, isovalue2, iso1, iso2),
              "generateTriangles_lattice_new");
    }

    void generateTriangles_lattice_newone(dim3, dim3, float4 *pos, float4 *norm,
                                          uint *compactedVoxelArray, uint *numVertsScanned,
                                          float *volume, uint3 gridSize, uint3 gridSizeShift,
                                          uint3 gridSizeMask, float3 voxelSize,
                                          float3 gridcenter, float isoValue, uint activeVoxels,
                                          uint maxVerts, uint totalverts, float *volume_one,
                                          float isovalue1, float isovalue2) {
        using namespace gpucad_compat;
        ensure_init();
        check(gpucad_generate_triangles_lattice_newone(
                  (gc_float4 *)pos, (gc_float4 *)norm, compactedVoxelArray, numVertsScanned,
                  volume, u3(gridSize), u3(gridSizeShift), u3(gridSizeMask), f3(voxelSize),
                  f3(gridcenter), isoValue, activeVoxels, maxVerts, totalverts, volume_one,
                  isovalue1, isovalue2),
              "generateTriangles_lattice_newone");
    }

    // src/MarchingCubes_kernel.h:39-45 (boolean meshing stages); primitive,
    // topo and compute_lattice select nothing in these kernels
    void classify_copy_Voxel_lattice(dim3, dim3, uint *, grid_points *vol_one, float *volume_two,
                                     float *vol_lattice, bool fixed, bool dynamic, float iso1,
                                     float iso2, uint3 gridSize, uint3, uint3, uint, float3,
                                     float isoValue, bool obj_union, bool obj_diff,
                                     bool obj_intersect) {
        using namespace gpucad_compat;
        ensure_init();
        check(gpucad_classify_copy_voxel(
                  vol_one, volume_two, vol_lattice, u3(gridSize), iso1, iso2, isoValue,
                  bool_flags(obj_union, obj_diff, obj_intersect, fixed, dynamic, false)),
              "classify_copy_Voxel_lattice");
    }

    void classifyVoxel_lattice(dim3, dim3, float *, uint3, uint *voxelVerts, uint *voxelOccupied,
                               grid_points *primitive_fixed, float *primitive_dynamic, float *,
                               float *lattice_field, uint3 gridSize, uint3 gridSizeShift,
                               uint3 gridSizeMask, uint numVoxels, float iso1, float iso2, float3,
                               float isoValue, bool obj_union, bool obj_diff, bool obj_intersect,
                               bool, bool, bool, bool fixed, bool dynamic, bool make_region) {
        using namespace gpucad_compat;
        ensure_init();
        check(gpucad_classify_voxel_boolean(
                  voxelVerts, voxelOccupied, primitive_fixed, primitive_dynamic, lattice_field,
                  u3(gridSize), u3(gridSizeShift), u3(gridSizeMask), numVoxels, iso1, iso2,
                  isoValue,
                  bool_flags(obj_union, obj_diff, obj_intersect, fixed, dynamic, make_region)),
              "classifyVoxel_lattice");
    }

    void generateTriangles_lattice(dim3, dim3, float4 *pos, float4 *norm, uint *compactVoxelArray,
                                   uint *numVertsScanned, uint3 gridSize, uint3 gridSizeShift,
                                   uint3 gridSizeMask, float3 voxelSize, float3 gridcenter,
                                   float isoValue, uint activeVoxels, uint maxVerts, uint,
                                   grid_points *primitive_fixed, float *primitive_dynamic,
                                   float *, float *lattice_field, float iso1, float iso2, uint *,
                                   bool obj_union, bool obj_diff, bool obj_intersect, bool, bool,
                                   bool, bool fixed, bool dynamic, bool make_region) {
        using namespace gpucad_compat;
        ensure_init();
        check(gpucad_generate_triangles_boolean(
                  (gc_float4 *)pos, (gc_float4 *)norm, compactVoxelArray, numVertsScanned,
                  u3(gridSize), u3(gridSizeShift), u3(gridSizeMask), f3(voxelSize),
                  f3(gridcenter), isoValue, activeVoxels, maxVerts, primitive_fixed,
                  primitive_dynamic, lattice_field, iso1, iso2,
                  bool_flags(obj_union, obj_diff, obj_intersect, fixed, dynamic, make_region)),
              "generateTriangles_lattice");
    }
};

class Isosurface : public MarchingCubeCuda {
public:
    // src/Isosurface.h:14-16, src/Isosurface.cu:15-27
    void copy_parameter(uint *voxel_verts, float isoValue, uint3 gridSize, uint3 gridSizeShift,
                        uint3 gridSizeMask, float3 voxelSize, uint numVoxels,
                        grid_points *vol_one, float *vol_two, float *vol_lattice, bool fixed,
                        bool dynamic, float iso1, float iso2, bool obj_union, bool obj_diff,
                        bool obj_intersect) {
        classify_copy_Voxel_lattice(dim3(), dim3(), voxel_verts, vol_one, vol_two, vol_lattice,
                                    fixed, dynamic, iso1, iso2, gridSize, gridSizeShift,
                                    gridSizeMask, numVoxels, voxelSize, isoValue, obj_union,
                                    obj_diff, obj_intersect);
    }

    // src/Isosurface.h:28-33, src/Isosurface.cu:44-134
    void computeIsosurface(float *, uint3, float4 *pos, float4 *norm, float isoValue,
                           uint numVoxels, uint *d_voxelVerts, uint *d_voxelVertsScan,
                           uint *d_voxelOccupied, uint *d_voxelOccupiedScan, uint3 gridSize,
                           uint3 gridSizeShift, uint3 gridSizeMask, float3 voxelSize,
                           float3 gridcenter, uint *activeVoxels, uint *totalVerts,
                           uint *d_compVoxelArray, uint maxVerts, grid_points *primitive_fixed,
                           float *primitive_dynamic, float *, float *lattice_field, float iso1,
                           float iso2, bool obj_union, bool obj_diff, bool obj_intersect, bool,
                           bool, bool, bool fixed, bool dynamic, bool make_region,
                           size_t *nfacets) {
        using namespace gpucad_compat;
        ensure_init();
        uint64_t nf = nfacets ? (uint64_t)*nfacets : 0;
        check(gpucad_compute_isosurface(
                  (gc_float4 *)pos, (gc_float4 *)norm, isoValue, numVoxels, d_voxelVerts,
                  d_voxelVertsScan, d_voxelOccupied, d_voxelOccupiedScan, u3(gridSize),
                  u3(gridSizeShift), u3(gridSizeMask), f3(voxelSize), f3(gridcenter),
                  activeVoxels, totalVerts, d_compVoxelArray, maxVerts, primitive_fixed,
                  primitive_dynamic, lattice_field, iso1, iso2,
                  bool_flags(obj_union, obj_diff, obj_intersect, fixed, dynamic, make_region),
                  &nf),
              "computeIsosurface");
        if (nfacets) *nfacets = (size_t)nf;
    }
    void computeIsosurface_lattice(float *vol, float4 *pos, float4 *norm, float &isoValue,
                                   uint numVoxels, uint *d_voxelVerts, uint *d_voxelVertsScan,
                                   uint *d_voxelOccupied, uint *d_voxelOccupiedScan,
                                   uint3 gridSize, uint3 gridSizeShift, uint3 gridSizeMask,
                                   float3 voxelSize, float3 gridcenter, uint *activeVoxels,
                                   uint *totalVerts, uint *d_compVoxelArray, uint maxVerts,
                                   float *vol_one, float *vol_two, float isovalue1,
                                   float isovalue2, float iso1, float iso2) {
        using namespace gpucad_compat;
        ensure_init();
        check(gpucad_compute_isosurface_lattice(
                  vol, (gc_float4 *)pos, (gc_float4 *)norm, isoValue, numVoxels, d_voxelVerts,
                  d_voxelVertsScan, d_voxelOccupied, d_voxelOccupiedScan, u3(gridSize),
                  u3(gridSizeShift), u3(gridSizeMask), f3(voxelSize), f3(gridcenter),
                  activeVoxels, totalVerts, d_compVoxelArray, maxVerts, vol_one, vol_two,
                  isovalue1, isovalue2, iso1, iso2),
              "computeIsosurface_lattice");
    }

    void computeIsosurface_latticeone(float *vol, float4 *pos, float4 *norm, float &isoValue,
                                      uint numVoxels, uint *d_voxelVerts, uint *d_voxelVertsScan,
                                      uint *d_voxelOccupied, uint *d_voxelOccupiedScan,
                                      uint3 gridSize, uint3 gridSizeShift, uint3 gridSizeMask,
                                      float3 voxelSize, float3 gridcenter, uint *activeVoxels,
                                      uint *totalVerts, uint *d_compVoxelArray, uint maxVerts,
                                      float *vol_one, float isovalue1, float isovalue2) {
        using namespace gpucad_compat;
        ensure_init();
        check(gpucad_compute_isosurface_latticeone(
                  vol, (gc_float4 *)pos, (gc_float4 *)norm, isoValue, numVoxels, d_voxelVerts,
                  d_voxelVertsScan, d_voxelOccupied, d_voxelOccupiedScan, u3(gridSize),
                  u3(gridSizeShift), u3(gridSizeMask), f3(voxelSize), f3(gridcenter),
                  activeVoxels, totalVerts, d_compVoxelArray, maxVerts, vol_one, isovalue1,
                  isovalue2),
              "computeIsosurface_latticeone");
    }
};

class Modelling {
public:
    static float a;
    static float b;

    Modelling() = default;
    Modelling(int, int, int) {}

    void distance_from_line(float *data_1, float3 center, float3 axis, float radius_1,
                            float thickness_radial, float thickness_axial, int Nx, int Ny, int Nz,
                            float dx, float dy, float dz, bool onetime) {
        gpucad_compat::check(
            gpucad_distance_from_line(data_1, gpucad_compat::f3(center), gpucad_compat::f3(axis),
                                      radius_1, thickness_radial, thickness_axial, Nx, Ny, Nz, dx,
                                      dy, dz, onetime ? 1 : 0),
            "distance_from_line");
    }

    void sphere_with_center(float *data_1, float3 center, float radius_1, float thickness_wall,
                            int Nx, int Ny, int Nz, float dx, float dy, float dz, bool onetime) {
        gpucad_compat::check(gpucad_sphere_with_center(data_1, gpucad_compat::f3(center), radius_1,
                                                       thickness_wall, Nx, Ny, Nz, dx, dy, dz,
                                                       onetime ? 1 : 0),
                             "sphere_with_center");
    }

    void cuboid(float *data_1, float3 center, float3 angles, float x_width, float y_width,
                float z_width, int Nx, int Ny, int Nz, float dx, float dy, float dz) {
        gpucad_compat::check(gpucad_cuboid(data_1, gpucad_compat::f3(center),
                                           gpucad_compat::f3(angles), x_width, y_width, z_width,
                                           Nx, Ny, Nz, dx, dy, dz),
                             "cuboid");
    }

    void cuboid_shell(float *data_1, float3 center, float3 angles, float x_width, float y_width,
                      float z_width, float thickness, int Nx, int Ny, int Nz, float dx, float dy,
                      float dz) {
        gpucad_compat::check(gpucad_cuboid_shell(data_1, gpucad_compat::f3(center),
                                                 gpucad_compat::f3(angles), x_width, y_width,
                                                 z_width, thickness, Nx, Ny, Nz, dx, dy, dz),
                             "cuboid_shell");
    }

    void torus_with_center(float *data_1, float3 center, float3 angles, float torus_radius,
                           float torus_circle_radius, int Nx, int Ny, int Nz, float dx, float dy,
                           float dz) {
        gpucad_compat::check(gpucad_torus_with_center(data_1, gpucad_compat::f3(center),
                                                      gpucad_compat::f3(angles), torus_radius,
                                                      torus_circle_radius, Nx, Ny, Nz, dx, dy,
                                                      dz),
                             "torus_with_center");
    }

    void cone_with_base_radius_height(float *data_1, float3 center, float3 angles,
                                      float base_radius, float cone_height, int Nx, int Ny,
                                      int Nz, float dx, float dy, float dz) {
        gpucad_compat::check(
            gpucad_cone_with_base_radius_height(data_1, gpucad_compat::f3(center),
                                                gpucad_compat::f3(angles), base_radius,
                                                cone_height, Nx, Ny, Nz, dx, dy, dz),
            "cone_with_base_radius_height");
    }

    void cone_frustum(float *data_1, float3 center, float3 angles, float top_radius,
                      float bottom_radius, float cone_frustum_height, int Nx, int Ny, int Nz,
                      float dx, float dy, float dz) {
        gpucad_compat::check(gpucad_cone_frustum(data_1, gpucad_compat::f3(center),
                                                 gpucad_compat::f3(angles), top_radius,
                                                 bottom_radius, cone_frustum_height, Nx, Ny, Nz,
                                                 dx, dy, dz),
                             "cone_frustum");
    }

    void pyramid_frustum(float *data_1, float3 center, float3 angles, float x_width_base,
                         float x_width_top, float y_height, float z_width_base,
                         float z_width_top, int Nx, int Ny, int Nz, float dx, float dy,
                         float dz) {
        gpucad_compat::check(gpucad_pyramid_frustum(data_1, gpucad_compat::f3(center),
                                                    gpucad_compat::f3(angles), x_width_base,
                                                    x_width_top, y_height, z_width_base,
                                                    z_width_top, Nx, Ny, Nz, dx, dy, dz),
                             "pyramid_frustum");
    }
};

inline float Modelling::a = 0.0f;
inline float Modelling::b = 0.0f;

class Fft_lattice {
public:
    void create_lattice(float *d_latticevol, uint NX, uint NY, uint NZ, uint size,
                        uint lattice_type_index) {
        gpucad_compat::ensure_init();
        gpucad_compat::check(
            gpucad_create_lattice(d_latticevol, NX, NY, NZ, size, lattice_type_index),
            "create_lattice");
    }
    // The transforms run on the shape gpucad_fft_plan3d recorded (the
    // caller's cufftPlan3d calls, src/main.cu:3522-3526, become one call of
    // it; cufftDestroy becomes gpucad_fft_destroy).
    void fft_func(float2 *fft_data) {
        gpucad_compat::check(gpucad_fft_r2c((gc_float2 *)fft_data), "fft_func");
    }
    void ifft_func(float2 *fft_data) {
        gpucad_compat::check(gpucad_fft_c2r((gc_float2 *)fft_data), "ifft_func");
    }
    void ifft_func_complex(float2 *fft_data) {
        gpucad_compat::check(gpucad_fft_c2c_inverse((gc_float2 *)fft_data), "ifft_func_complex");
    }
    void indiviual_grating_sum(float2 *sum_gratings, float *fft_function, int Nx, int Ny,
                               int Nz) {
        gpucad_compat::check(
            gpucad_grating_sum((const gc_float2 *)sum_gratings, fft_function, Nx, Ny, Nz),
            "indiviual_grating_sum");
    }
    void add_fft_constants(float2 *indvidual_grating, float2 fft_constant, int index) {
        gpucad_compat::check(gpucad_add_fft_constant((gc_float2 *)indvidual_grating,
                                                     gc_float2{fft_constant.x, fft_constant.y},
                                                     index),
                             "add_fft_constants");
    }
    void fft_scalar(float2 *fft_data_compute, float scalar_val, int size) {
        gpucad_compat::check(gpucad_fft_scale((gc_float2 *)fft_data_compute, scalar_val, size),
                             "fft_scalar");
    }
    void fft_fill(float2 *fft_compute, float2 *fft_compute_fill, int Nx, int Ny, int Nz,
                  uint mid_index) {
        (void)mid_index;
        gpucad_compat::check(gpucad_fft_fill((const gc_float2 *)fft_compute,
                                             (gc_float2 *)fft_compute_fill, Nx, Ny, Nz),
                             "fft_fill");
    }
};

// The reference samples the phase through a 3-D texture (linear filter,
// unnormalised coordinates, src/Interpolations.cu:57-107).  Here the
// "texture" is a plain device copy of the phase taken at updateTexture, and
// Gratings::grating/refine sample it with the same filter arithmetic.
class Interpolations {
protected:
    float *tex_ = nullptr;  // the texture's contents
    int tex_x_ = 0, tex_y_ = 0, tex_z_ = 0;

public:
    void setupTexture(int x, int y, int z) {
        deleteTexture();
        tex_x_ = x;
        tex_y_ = y;
        tex_z_ = z;
        gpucad_compat::check(
            hipMalloc((void **)&tex_, sizeof(float) * (size_t)x * y * z) == hipSuccess
                ? GPUCAD_OK
                : GPUCAD_E_OOM,
            "setupTexture");
    }
    // writes d_phi into the pitched staging buffer, as the reference does
    void copytotexture(float *d_phi, hipPitchedPtr data_ptr, int NX, int NY, int NZ) {
        gpucad_compat::check(
            hipMemcpy2D(data_ptr.ptr, data_ptr.pitch, d_phi, sizeof(float) * NX,
                        sizeof(float) * NX, (size_t)NY * NZ, hipMemcpyDeviceToDevice) == hipSuccess
                ? GPUCAD_OK
                : GPUCAD_E_HIP,
            "copytotexture");
    }
    // the staging buffer becomes the texture's contents
    void updateTexture(hipPitchedPtr data_ptr) {
        gpucad_compat::check(
            hipMemcpy2D(tex_, sizeof(float) * tex_x_, data_ptr.ptr, data_ptr.pitch,
                        sizeof(float) * tex_x_, (size_t)tex_y_ * tex_z_,
                        hipMemcpyDeviceToDevice) == hipSuccess
                ? GPUCAD_OK
                : GPUCAD_E_HIP,
            "updateTexture");
    }
    void deleteTexture() {
        if (tex_) (void)hipFree(tex_);
        tex_ = nullptr;
    }
};

class Gratings : public Interpolations {
public:
    int NX = 0, NY = 0, NZ = 0;

    void angle_data(float *d_theta, int NX_, int NY_, int NZ_, float dx, float dy, float dz,
                    float mean_x, float mean_y, float mean_z, char axis) {
        gpucad_compat::check(gpucad_svl_angle_data(d_theta, NX_, NY_, NZ_, dx, dy, dz, mean_x,
                                                   mean_y, mean_z, axis),
                             "angle_data");
    }
    void period_data(float *d_period, int NX_, int NY_, int NZ_, float dx, float dy, float dz,
                     float mean_x, float mean_y, float mean_z, char axis) {
        gpucad_compat::check(gpucad_svl_period_data(d_period, NX_, NY_, NZ_, dx, dy, dz, mean_x,
                                                    mean_y, mean_z, axis),
                             "period_data");
    }
    void VecSMultAdd_lattice(float *d_v, float a1, float *d_w, const float a2, const int NX_,
                             const int NY_, const int NZ_) {
        gpucad_compat::check(gpucad_svl_axpby(d_v, a1, d_w, a2, (uint64_t)NX_ * NY_ * NZ_),
                             "VecSMultAdd_lattice");
    }
    void finding_phi(float *d_phi, float *d_period, int x_dim, int y_dim, int z_dim, int i,
                     int j, int k, float dx, float dy, float dz, char latticetype_one,
                     int unform_type, float const_peirod, float x_period, float y_period,
                     float z_period, float lcon, float lcon_1, bool sinewave_zaxis) {
        gpucad_compat::check(
            gpucad_svl_finding_phi(d_phi, d_period, x_dim, y_dim, z_dim, i, j, k, dx, dy, dz,
                                   latticetype_one, unform_type, const_peirod, x_period,
                                   y_period, z_period, lcon, lcon_1, sinewave_zaxis ? 1 : 0),
            "finding_phi");
    }
    // on the object's NX, NY, NZ, as the reference's
    void GPUCG_lattice(float *d_phi, const int iter, const int OptIter, const float EndRes,
                       int &FinalIter, float &FinalRes) {
        (void)OptIter;
        int32_t it = 0;
        float res = 0.f;
        gpucad_compat::check(gpucad_svl_cg(d_phi, NX, NY, NZ, iter, EndRes, &it, &res),
                             "GPUCG_lattice");
        FinalIter = it;
        FinalRes = res;
    }
    void GPUScalar_lattice(float *d_result, float *d_vec1, float *d_vec2, int n, int block_num) {
        (void)block_num;
        gpucad_compat::check(gpucad_svl_dot(d_result, d_vec1, d_vec2, (uint64_t)n),
                             "GPUScalar_lattice");
    }
    void GPU_buffer_normalise_buffer(float *d_vec1, float *d_vec2, int n) {
        gpucad_compat::check(gpucad_normalise_buffer(d_vec1, d_vec2, (uint64_t)n),
                             "GPU_buffer_normalise_buffer");
    }
    void GPU_buffer_normalise_three(float *dataone, float *datatwo, size_t size, float a1,
                                    float b1) {
        gpucad_compat::check(gpucad_normalise_three(dataone, datatwo, size, a1, b1),
                             "GPU_buffer_normalise_three");
    }
    void GPUMatvec_lattice(float *d_d, float *d_q, int NX_, int NY_, int NZ_) {
        gpucad_compat::check(gpucad_svl_matvec(d_d, d_q, NX_, NY_, NZ_), "GPUMatvec_lattice");
    }
    void grating(float2 *dvol, int NX2, int NY2, int NZ2, float dx2, float dy2, float dz2) {
        gpucad_compat::check(gpucad_svl_grating(tex_, tex_x_, tex_y_, tex_z_, (gc_float2 *)dvol,
                                                NX2, NY2, NZ2, dx2, dy2, dz2),
                             "grating");
    }
    void refine(float *dvol, int NX2, int NY2, int NZ2, float dx, float dy, float dz) {
        gpucad_compat::check(gpucad_svl_refine(tex_, tex_x_, tex_y_, tex_z_, dvol, NX2, NY2, NZ2,
                                               dx, dy, dz),
                             "refine");
    }
    void svl(float *d_svl, float2 *d_grating, int NX_, int NY_, int NZ_, int indxx,
             float2 *data_fft) {
        gpucad_compat::check(gpucad_svl_accumulate(d_svl, (const gc_float2 *)d_grating,
                                                   (uint64_t)NX_ * NY_ * NZ_, indxx,
                                                   (const gc_float2 *)data_fft),
                             "svl");
    }
    void fillfrac(float *d_svl, float *d_fillfrac, int NX_, int NY_, int NZ_) {
        gpucad_compat::check(gpucad_svl_fillfrac(d_svl, d_fillfrac, NX_, NY_, NZ_), "fillfrac");
    }
    void GPUCleanUp() {}

    void GPU_buffer_normalise_four(float *dataone, float *datatwo, float *datathree, size_t size,
                                   int Nx, int Ny, int Nz, float isoval_1, float isoval_2) {
        gpucad_compat::check(gpucad_normalise_four(dataone, datatwo, datathree, size, Nx, Ny, Nz,
                                                   isoval_1, isoval_2),
                             "GPU_buffer_normalise_four");
    }

    void primitive_field(grid_points *primitive_field, float *primitive_active, float *isosurf,
                         float isoval, bool fixed, bool active, int NX_, int NY_, int NZ_) {
        (void)isoval;
        if (fixed)
            gpucad_compat::check(
                gpucad_primitive_field_fixed(primitive_field, isosurf, NX_, NY_, NZ_),
                "primitive_field");
        else if (active)
            gpucad_compat::check(
                gpucad_primitive_field_dynamic(primitive_active, isosurf, NX_, NY_, NZ_),
                "primitive_field");
    }
};

class File_output {
public:
    void file_write_obj(float4 *d_pos, uint totalVerts, const char *filename) {
        gpucad_compat::check(gpucad_file_write_obj((const gc_float4 *)d_pos, totalVerts, filename),
                             "file_write_obj");
    }
};

// The caller's Vulkan interop helpers (src/main.cu:1164-1211, 2017-2035,
// 739-931) with the CUDA runtime calls replaced: the caller still obtains the
// fd from Vulkan (vkGetMemoryFdKHR / vkGetSemaphoreFdKHR) and passes it in.
// Errors abort like checkCudaErrors.
namespace gpucad_interop {

// importCudaExternalMemory: the whole allocation of `size` bytes, offset 0.
inline void importExternalMemory(void **devPtr, gc_external_buffer *&mem, int fd,
                                 unsigned long long size) {
    gpucad_compat::check(gpucad_import_external_buffer(fd, size, 0, &mem, devPtr),
                         "importExternalMemory");
}
inline void destroyExternalMemory(gc_external_buffer *mem) {
    gpucad_compat::check(gpucad_release_external_buffer(mem), "destroyExternalMemory");
}
// importCudaExternalSemaphore (binary semaphore, opaque fd).
inline void importExternalSemaphore(gc_external_semaphore *&sem, int fd) {
    gpucad_compat::check(gpucad_import_external_semaphore(fd, &sem), "importExternalSemaphore");
}
// cudaSignalExternalSemaphoresAsync / cudaWaitExternalSemaphoresAsync with
// fence value 0 on the default stream, as drawFrame does.
inline void signalExternalSemaphore(gc_external_semaphore *sem) {
    gpucad_compat::check(gpucad_signal_external_semaphore(sem, nullptr), "signalExternalSemaphore");
}
inline void waitExternalSemaphore(gc_external_semaphore *sem) {
    gpucad_compat::check(gpucad_wait_external_semaphore(sem, nullptr), "waitExternalSemaphore");
}
inline void destroyExternalSemaphore(gc_external_semaphore *sem) {
    gpucad_compat::check(gpucad_release_external_semaphore(sem), "destroyExternalSemaphore");
}

}  // namespace gpucad_interop
