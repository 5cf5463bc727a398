/*
 * gpucad.h — C-ABI of the MI355X implicit-geometry engine (libgpucad.so).
 *
 * The library evaluates implicit lattice / primitive fields on the GPU,
 * extracts their isosurface with marching cubes and assembles vertex and
 * index buffers.  Every entry point takes plain pointers and sizes; device
 * pointers are HIP device allocations owned by the caller unless stated.
 *
 * Two families of entry points:
 *   - "drop-in" entry points mirror one method of the reference's C++ classes
 *     (cited per declaration) with the same argument meaning and the same
 *     synchronous completion (outputs are valid when the call returns).
 *     C++ wrappers with the reference's exact class/method signatures live in
 *     gpucad_compat.hpp.
 *   - "native" entry points (gpucad_lattice_field, gpucad_extract, batch,
 *     server) are asynchronous on the caller's HIP stream unless a host
 *     output pointer asks for completion.
 *
 * Every function returns a status code (GPUCAD_OK on success); the message of
 * the last failure on the calling thread is available from gpucad_last_error().
 * Stream arguments are hipStream_t passed as void*; NULL means the null stream.
 */
#ifndef GPUCAD_H_
#define GPUCAD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                              */
/* ------------------------------------------------------------------------ */
enum {
    GPUCAD_OK = 0,
    GPUCAD_E_ARG = 1,       /* invalid argument (sizes, NULL pointers, ranges) */
    GPUCAD_E_OVERFLOW = 2,  /* output capacity too small; counts are still reported */
    GPUCAD_E_OOM = 3,       /* device or host allocation failed / budget exceeded */
    GPUCAD_E_HIP = 4,       /* a HIP runtime call failed */
    GPUCAD_E_BUSY = 5,      /* resource in use (server slots full, not ready) */
    GPUCAD_E_NOTFOUND = 6,  /* unknown handle / ticket */
    GPUCAD_E_IO = 7         /* file could not be written */
};

/* POD mirrors of the HIP/CUDA vector types used by the reference's
 * signatures; layout-identical to uint3/float3/float4 (checked by
 * static_assert in the implementation). */
typedef struct gc_uint3 { uint32_t x, y, z; } gc_uint3;
typedef struct gc_float3 { float x, y, z; } gc_float3;
typedef struct gc_float4 { float x, y, z, w; } gc_float4;
typedef struct gc_float2 { float x, y; } gc_float2;

/* ------------------------------------------------------------------------ */
/* Library                                                                   */
/* ------------------------------------------------------------------------ */
int gpucad_version(void);                  /* 10000*major + 100*minor + patch */
/* sizeof of the ABI structs, for binding checks: sizes[0..n) receives
 * gc_lattice_params, gc_extract_args, gc_mesh_counts, gc_uint3, gc_float3,
 * gc_float4, gc_float2, gc_svl_params, gc_svl_result (in that order);
 * returns the number written. */
int gpucad_abi_sizes(uint32_t *sizes, int32_t n);
const char *gpucad_status_string(int status);
const char *gpucad_last_error(void);       /* thread-local; "" when none */

/* Selects the HIP device for the calling thread and uploads the marching-cubes
 * tables (once per device).  Replaces MarchingCubeCuda::allocateTextures_s
 * (src/MarchingCubes_kernel.cu:24-68), which uploads triTable/numVertsTable
 * into texture objects before any extraction. */
int gpucad_init(int device);
/* Releases per-device state created by gpucad_init.  Replaces
 * MarchingCubeCuda::destroyAllTextureObjects (src/MarchingCubes_kernel.cu:71-75). */
int gpucad_shutdown(void);

/* Host copy of the 256-case tables (no GPU needed): tri_table[256*16]
 * (edge ids, 255 = end), num_verts[256], edge_mask[256] (12-bit crossing
 * mask).  Same data as src/tables.h:11-45, 49-307, 311-569. */
int gpucad_mc_tables(uint8_t *tri_table, uint8_t *num_verts, uint16_t *edge_mask);

/* ------------------------------------------------------------------------ */
/* Native lattice field                                                      */
/* ------------------------------------------------------------------------ */
enum { GC_TPMS_GYROID = 0, GC_TPMS_SCHWARZ_P = 1, GC_TPMS_DIAMOND = 2 };
enum { GC_PRIM_NONE = 0, GC_PRIM_SPHERE = 1, GC_PRIM_BOX = 2 };
enum { GC_OP_NONE = 0, GC_OP_INTERSECT = 1, GC_OP_UNION = 2, GC_OP_SUBTRACT = 3 };
#define GC_MAX_MODES 8

/* A spatially varying TPMS lattice on an (nx, ny, nz)-point grid over the
 * unit cube: point (x,y,z) sits at p = (x*hx, y*hy, z*hz), h = 1/(n-1).
 *
 *  phase coordinates (voxel units, revolutions after division by the period):
 *    no bend:  (u, v) = c + R(theta(z)) * ((x, y) - c),  c = ((nx-1)/2, (ny-1)/2),
 *              theta(z) = rot0 + (rot1 - rot0) * z / (nz-1)   (radians),  w = z
 *    bend:     a = p_xy - C, C = (1/2, 1/2 - R_b), r = |a|, phi = atan2(a.x, a.y),
 *              u = (C.x + R_b*phi)/hx, v = (C.y + r)/hy, w = z
 *    P(x) = period0 + (period1 - period0) * x/(nx-1) voxels;  X = u/P, Y = v/P, Z = w/P
 *  tpms f:   gyroid  sin X cos Y + sin Y cos Z + sin Z cos X
 *            P       cos X + cos Y + cos Z
 *            diamond sXsYsZ + sXcYcZ + cXsYcZ + cXcYsZ        (angles 2*pi*X, ...)
 *  thickness t(p) = t0 + sum_m mode_a[m] cos(2*pi*(k_m . p + mode_phi[m]))
 *  g = sheet ? |f| - t : f
 *  primitive s: sphere |p - c| - radius;  box: exact box SDF with half extents
 *  G = op(g, s): intersect max(g,s), union min(g,s), subtract max(g,-s), none g
 * Inside is G < iso.  Rotation and bend are exclusive. */
typedef struct gc_lattice_params {
    int32_t nx, ny, nz;
    int32_t tpms;
    float period0, period1;
    float rot0, rot1;
    float bend_radius;
    int32_t sheet;
    float t0;
    int32_t n_modes;
    int32_t mode_k[GC_MAX_MODES][3];
    float mode_a[GC_MAX_MODES];
    float mode_phi[GC_MAX_MODES];
    int32_t prim;
    int32_t op;
    float prim_center[3];
    float prim_radius;
    float prim_half[3];
} gc_lattice_params;

/* 64-bit words per sign row: bit b of word w of row r = y + ny*z is
 * [G(64w+b, y, z) < iso]; padding bits are zero. */
uint64_t gpucad_sign_words(int32_t nx, int32_t ny, int32_t nz);

/* Evaluates G into d_field (nx*ny*nz floats, x fastest).  If d_signs is not
 * NULL it also writes the sign words for isovalue iso (gpucad_sign_words()
 * words), which lets gpucad_extract skip its own classification read. */
int gpucad_lattice_field(const gc_lattice_params *p, float *d_field, uint64_t *d_signs,
                         float iso, void *stream);
/* Same, with the number of z planes each workgroup walks (1..128; 0 = default
 * 32).  The batch autotuner sweeps it. */
int gpucad_lattice_field_lz(const gc_lattice_params *p, float *d_field, uint64_t *d_signs,
                            float iso, int32_t lz, void *stream);

/* ------------------------------------------------------------------------ */
/* Native extraction: indexed mesh with shared-edge vertex dedup            */
/* ------------------------------------------------------------------------ */
/* Vertex k of the output is the k-th crossing grid edge in canonical order
 * (edge id 3*pointIndex + axis, pointIndex = x + nx*(y + ny*z)), placed at
 * p0 + t*(p1 - p0), t = (iso - f0)/(f1 - f0), low end first.  Triangle j is
 * the j-th triangle in cell order (cell index x + (nx-1)*(y + (ny-1)*z)) and
 * triTable order within the cell; its three entries are vertex indices. */
typedef struct gc_extract_args {
    const float *field;        /* device, nx*ny*nz */
    const uint64_t *signs;     /* device, optional (NULL: computed into the workspace) */
    int32_t nx, ny, nz;
    float iso;
    gc_float3 origin;          /* position of point (0,0,0) */
    gc_float3 spacing;         /* point spacing per axis */
    float *vertices;           /* device float3[vert_capacity] */
    uint64_t vert_capacity;
    uint32_t *triangles;       /* device uint32[3*tri_capacity] */
    uint64_t tri_capacity;
    float *normals;            /* reserved for per-vertex normals: must be NULL
                                  (GPUCAD_E_ARG otherwise) */
    void *workspace;           /* device, >= gpucad_extract_workspace_bytes() */
    uint64_t workspace_bytes;
    uint64_t *d_counts;        /* device uint64[2] = {vertices, triangles}, or NULL */
    void *emit_event;          /* optional hipEvent_t, recorded on the stream just before
                                  the emit launch (the last and longest of the call's
                                  launches), so that launch can be timed; or NULL */
} gc_extract_args;

typedef struct gc_mesh_counts { uint64_t n_vertices, n_triangles; } gc_mesh_counts;

uint64_t gpucad_extract_workspace_bytes(int32_t nx, int32_t ny, int32_t nz);
/* Enqueues the extraction on stream.  If counts is not NULL the call waits for
 * completion, fills counts and returns GPUCAD_E_OVERFLOW when a capacity was
 * too small (nothing is written past a capacity in any case). */
int gpucad_extract(const gc_extract_args *a, gc_mesh_counts *counts, void *stream);

/* ------------------------------------------------------------------------ */
/* Batched form: many lattice grids of uneven sizes and types per launch    */
/* (no counterpart in the reference)                                        */
/* ------------------------------------------------------------------------ */
/* A batch owns the fields, sign words, extraction workspace and mesh buffers
 * of its grids.  gpucad_batch_run evaluates every field (one workgroup per
 * brick of any grid) and extracts every mesh (one workgroup per block of
 * sign words of any grid); with use_graph the whole run is one replay of a
 * captured HIP graph.  Each grid's mesh is bit-identical to what
 * gpucad_lattice_field + gpucad_extract produce for it alone. */
typedef struct gc_batch gc_batch;

typedef struct gc_batch_options {
    float iso;                 /* isovalue of every grid */
    int32_t brick_x;           /* points along x per wave row: 64, 32, 16, 8; 0 = autotuned */
    int32_t brick_lz;          /* z planes per brick, 1..128; 0 = autotuned */
    int32_t waves_per_cu;      /* must be 0 (GPUCAD_E_ARG otherwise): the field
                                  runs one workgroup per brick, so there is no
                                  resident-wave count to choose; reported as 0 */
    int32_t use_graph;         /* nonzero: a run is one replay of a captured graph */
    int32_t unit_domain;       /* vertex positions: 0 = point indices, 1 = [0,1]^3
                                  (spacing 1/(n-1) per axis) */
    uint64_t vert_capacity;    /* total over the batch; 0 = sized at creation */
    uint64_t tri_capacity;     /* (by a counting run of the batch) */
} gc_batch_options;

typedef struct gc_batch_grid_result {
    uint64_t field_offset;     /* first point of the grid's field (floats) */
    uint64_t vert_offset;      /* the grid's vertices: [vert_offset, +n_vertices) */
    uint64_t n_vertices;
    uint64_t tri_offset;       /* its triangles; corners are vertex numbers */
    uint64_t n_triangles;      /* relative to vert_offset */
    int32_t status;            /* GPUCAD_OK or GPUCAD_E_OVERFLOW (mesh not complete) */
    int32_t pad;
} gc_batch_grid_result;

typedef struct gc_batch_buffers {
    float *fields;             /* device, field_floats */
    uint64_t field_floats;
    float *vertices;           /* device float3[vert_capacity] */
    uint64_t vert_capacity;
    uint32_t *triangles;       /* device uint32[3*tri_capacity] */
    uint64_t tri_capacity;
} gc_batch_buffers;

/* Validates the grids, allocates the batch, autotunes what the options leave
 * at 0 (timing candidates on this batch; choices are cached per device and
 * batch signature, and in the file named by GPUCAD_TUNE_CACHE if set), sizes
 * the mesh buffers and captures the graph.  Grids: 1..65535, each with
 * nx <= 65536 and fewer than 2^32 sign words. */
int gpucad_batch_create(const gc_lattice_params *grids, int32_t n_grids,
                        const gc_batch_options *opt, gc_batch **out);
/* Enqueues one run of the whole batch on stream (asynchronous). */
int gpucad_batch_run(gc_batch *b, void *stream);
/* Waits for the last run and fills res[0..n) (n = the batch's grid count). */
int gpucad_batch_results(gc_batch *b, gc_batch_grid_result *res, int32_t n);
int gpucad_batch_buffers(const gc_batch *b, gc_batch_buffers *out);
/* The options in use after autotuning. */
int gpucad_batch_tuning(const gc_batch *b, gc_batch_options *chosen);
int gpucad_batch_destroy(gc_batch *b);

/* ------------------------------------------------------------------------ */
/* Served form: lattice-edit requests as they arrive (no reference          */
/* counterpart)                                                              */
/* ------------------------------------------------------------------------ */
/* A server owns one device arena of budget_bytes.  Submitted requests wait
 * in a FIFO; each call of gpucad_server_poll advances the loop: requests are
 * admitted into a launch group while their field, sign words and extraction
 * work fit in the arena (up to max_batch per group), a group runs on one of
 * `slots` streams (field + count, then, once its mesh sizes are known and
 * mesh memory is allocated, vertices + triangles), and groups on different
 * streams overlap on the device.  Requests that arrive while groups run are
 * admitted as soon as a stream and memory are free (continuous batching).
 * A finished request's field and mesh stay resident in the arena until
 * gpucad_server_release.  Each mesh is bit-identical to what
 * gpucad_lattice_field + gpucad_extract produce for the request alone
 * (vertex positions in [0,1]^3: spacing 1/(n-1) per axis). */
typedef struct gc_server gc_server;

typedef struct gc_server_options {
    uint64_t budget_bytes;     /* device arena size (fields, meshes, work) */
    int32_t max_batch;         /* requests per launch group, 1..65535; 0 = 256 */
    int32_t slots;             /* launch groups in flight (streams), 1..8; 0 = 3 */
    float iso;
    int32_t brick_x;           /* field brick (as gc_batch_options); 0 = 64 */
    int32_t brick_lz;          /* 0 = 16 */
    int32_t waves_per_cu;      /* must be 0 (see gc_batch_options) */
} gc_server_options;

typedef struct gc_request_result {
    uint64_t ticket;
    int32_t status;            /* GPUCAD_OK, or GPUCAD_E_OOM: larger than the budget */
    int32_t pad;
    const float *field;        /* device, nx*ny*nz floats */
    const float *vertices;     /* device float3[n_vertices] */
    const uint32_t *triangles; /* device uint32[3*n_triangles]; vertex numbers of this mesh */
    uint64_t n_vertices, n_triangles;
    double submit_s, done_s;   /* seconds on the server's clock */
} gc_request_result;

typedef struct gc_server_stats {
    uint64_t submitted, completed, rejected, groups;
    uint64_t bytes_in_use, peak_bytes_in_use, budget_bytes;
    double mean_group_size;
} gc_server_stats;

int gpucad_server_create(const gc_server_options *opt, gc_server **out);
/* Queues a request (the parameters are copied); returns its ticket. */
int gpucad_server_submit(gc_server *s, const gc_lattice_params *p, uint64_t *ticket);
/* Advances the loop without waiting and returns up to max finished requests
 * (each is returned once); *n_done receives how many. */
int gpucad_server_poll(gc_server *s, gc_request_result *done, int32_t max, int32_t *n_done);
/* Frees a finished request's field and mesh. */
int gpucad_server_release(gc_server *s, uint64_t ticket);
int gpucad_server_stats(const gc_server *s, gc_server_stats *out);
/* Waits for all work in flight and frees the arena. */
int gpucad_server_destroy(gc_server *s);
/* Drives the server with n requests arriving at arrival_s[i] seconds after
 * the start (non-decreasing): submits each on time, polls continuously,
 * releases each result when it completes, and writes latency_s[i] = done -
 * arrival.  *wall_s receives the time from start to the last completion.
 * Requests with status other than OK get latency -1. */
int gpucad_server_run_trace(gc_server *s, const gc_lattice_params *reqs,
                            const double *arrival_s, int32_t n, double *latency_s,
                            double *wall_s);
/* The same requests served by plain per-request calls, one at a time in
 * arrival order (gpucad_lattice_field with sign words, then gpucad_extract
 * with its synchronous counts, into buffers allocated once for the largest
 * request): the baseline the served form is compared with. */
int gpucad_plain_run_trace(const gc_lattice_params *reqs, const double *arrival_s, int32_t n,
                           float iso, double *latency_s, double *wall_s);

/* ------------------------------------------------------------------------ */
/* Drop-in field entry points                                               */
/* ------------------------------------------------------------------------ */
/* Fft_lattice::create_lattice (src/lattice_files/Fft_lattice.h:97,
 * kernel src/lattice_files/Fft_lattice.cu:12-66): unit-cell lattice types 0..5. */
int gpucad_create_lattice(float *d_latticevol, uint32_t nx, uint32_t ny, uint32_t nz,
                          uint32_t size, uint32_t type);
/* Modelling::sphere_with_center (src/Modelling.h:28, src/Modelling.cu:314-372) */
int gpucad_sphere_with_center(float *data, gc_float3 center, float radius, float thickness,
                              int32_t nx, int32_t ny, int32_t nz, float dx, float dy, float dz,
                              int32_t shell);
/* Modelling::cuboid (src/Modelling.h:30, src/Modelling.cu:375-431) */
int gpucad_cuboid(float *data, gc_float3 center, gc_float3 angles, float x_width,
                  float y_width, float z_width, int32_t nx, int32_t ny, int32_t nz,
                  float dx, float dy, float dz);
/* Modelling::cuboid_shell (src/Modelling.h:32, src/Modelling.cu:435-498) */
int gpucad_cuboid_shell(float *data, gc_float3 center, gc_float3 angles, float x_width,
                        float y_width, float z_width, float thickness, int32_t nx,
                        int32_t ny, int32_t nz, float dx, float dy, float dz);
/* Modelling::pyramid_frustum (src/Modelling.h:40, src/Modelling.cu:501-567) */
int gpucad_pyramid_frustum(float *data, gc_float3 center, gc_float3 angles,
                           float x_width_base, float x_width_top, float y_height,
                           float z_width_base, float z_width_top, int32_t nx, int32_t ny,
                           int32_t nz, float dx, float dy, float dz);
/* Modelling::torus_with_center (src/Modelling.h:34, src/Modelling.cu:569-627) */
int gpucad_torus_with_center(float *data, gc_float3 center, gc_float3 angles,
                             float torus_radius, float circle_radius, int32_t nx,
                             int32_t ny, int32_t nz, float dx, float dy, float dz);
/* Modelling::cone_with_base_radius_height (src/Modelling.h:36, src/Modelling.cu:631-694) */
int gpucad_cone_with_base_radius_height(float *data, gc_float3 center, gc_float3 angles,
                                        float base_radius, float cone_height, int32_t nx,
                                        int32_t ny, int32_t nz, float dx, float dy, float dz);
/* Modelling::cone_frustum (src/Modelling.h:38, src/Modelling.cu:698-761) */
int gpucad_cone_frustum(float *data, gc_float3 center, gc_float3 angles, float top_radius,
                        float bottom_radius, float height, int32_t nx, int32_t ny,
                        int32_t nz, float dx, float dy, float dz);
/* Modelling::distance_from_line (src/Modelling.h:26, src/Modelling.cu:244-311) */
int gpucad_distance_from_line(float *data, gc_float3 center, gc_float3 axis, float radius,
                              float thickness_radial, float thickness_axial, int32_t nx,
                              int32_t ny, int32_t nz, float dx, float dy, float dz,
                              int32_t cylinder_disc);
/* Gratings::GPU_buffer_normalise_four (src/lattice_files/Gratings.h:50,
 * src/lattice_files/Gratings.cu:1579-1617): a = min(0, min f), b = max(0, max f),
 * k = (f-a)/(b-a); boundary points -> indicator 0, k 0; else indicator
 * (iso1 <= k <= iso2).  normalised may alias in. */
int gpucad_normalise_four(const float *in, float *indicator, float *normalised, uint64_t size,
                          int32_t nx, int32_t ny, int32_t nz, float iso1, float iso2);
/* Gratings::primitive_field (src/lattice_files/Gratings.h:77,
 * src/lattice_files/Gratings.cu:1695-1734), dynamic-primitive form: sets
 * isosurf[i] = FLT_MAX where primitive_active[i] >= 0. */
int gpucad_primitive_field_dynamic(const float *primitive_active, float *isosurf, int32_t nx,
                                   int32_t ny, int32_t nz);
/* Gratings::primitive_field, fixed-primitive form: sets isosurf[i] = FLT_MAX
 * where primitive_fixed[i].val > -1.  primitive_fixed points at the
 * reference's grid_points records {int val; float t_x, t_y, t_z;} (16 bytes,
 * src/MarchingCubes_kernel.h:12-18). */
int gpucad_primitive_field_fixed(const void *primitive_fixed, float *isosurf, int32_t nx,
                                 int32_t ny, int32_t nz);

/* ------------------------------------------------------------------------ */
/* Spatially variant lattice (SVL) synthesis: the step before meshing       */
/* (src/main.cu:3904-4037 spatial_lattice_run, 3578-3711 unit_lattice).     */
/* Synchronous on the null stream unless a stream is taken, like the        */
/* reference's methods.  Coarse grids need >= 4 points per axis.            */
/* ------------------------------------------------------------------------ */
/* Gratings::period_data (src/lattice_files/Gratings.h:36, Gratings.cu:1376,
 * kernel 815-860): x = (i - mean_x) dx, ...; axis 'z': sqrt((x+1)^2 + (y+1)^2),
 * 'y': sqrt((x+1)^2 + (z+1)^2), otherwise sqrt((z+1)^2 + (y+1)^2). */
int gpucad_svl_period_data(float *d_period, int32_t nx, int32_t ny, int32_t nz, float dx, float dy,
                           float dz, float mean_x, float mean_y, float mean_z, char axis);
/* Gratings::angle_data (Gratings.h:33, Gratings.cu:1358, kernel 775-812):
 * atan2(y+1, x+1), or atan2(z+1, x+1) for axis 'y'. */
int gpucad_svl_angle_data(float *d_theta, int32_t nx, int32_t ny, int32_t nz, float dx, float dy,
                          float dz, float mean_x, float mean_y, float mean_z, char axis);
/* Gratings::GPU_buffer_normalise_three (Gratings.h:48, Gratings.cu:1539-1575):
 * a = min(0, min f), b = max(0, max f), out = a1 + b1 (f-a)/(b-a) (the
 * reference's [a1, a1+b1] range).  out may alias in. */
int gpucad_normalise_three(const float *in, float *out, uint64_t size, float a1, float b1);
/* Gratings::GPU_buffer_normalise_buffer (Gratings.h:46, Gratings.cu:1500-1536):
 * out = (f-a)/(b-a), a and b as above.  out may alias in. */
int gpucad_normalise_buffer(const float *in, float *out, uint64_t size);
/* Gratings::finding_phi (Gratings.h:39-40, kernel Gratings.cu:100-417): the
 * right-hand side D^T K of harmonic (i, j, k); lattice_type 'n', 'b', 'r'
 * or 's'; uniform_type 0 (const_period), 1 (x/y/z_period) or 2 (d_period
 * field, nx*ny*nz). */
int gpucad_svl_finding_phi(float *d_phi, const float *d_period, int32_t nx, int32_t ny, int32_t nz,
                           int32_t i, int32_t j, int32_t k, float dx, float dy, float dz,
                           char lattice_type, int32_t uniform_type, float const_period,
                           float x_period, float y_period, float z_period, float lcon,
                           float lcon_1, int32_t sinewave_zaxis);
/* Gratings::GPUMatvec_lattice (Gratings.h:62, kernel Gratings.cu:420-597):
 * q = A d, the phase solve's operator. */
int gpucad_svl_matvec(const float *d_d, float *d_q, int32_t nx, int32_t ny, int32_t nz);
/* Gratings::GPUCG_lattice (Gratings.h:42, Gratings.cu:875-974): conjugate
 * gradients from x = 0 on A x = d_phi, the solution written over d_phi;
 * stops when r.r <= end_res^2 or at `iter` (the reference's iCounter),
 * final_iter/final_res as the reference reports them.  Scalars stay on the
 * device (no host round trip per iteration). */
int gpucad_svl_cg(float *d_phi, int32_t nx, int32_t ny, int32_t nz, int32_t iter, float end_res,
                  int32_t *final_iter, float *final_res);
/* Gratings::VecSMultAdd_lattice (Gratings.h:37, Gratings.cu:73-94, 977-985):
 * v = a1 v + a2 w over n points. */
int gpucad_svl_axpby(float *d_v, float a1, const float *d_w, float a2, uint64_t n);
/* Gratings::GPUScalar_lattice (Gratings.h:44, Gratings.cu:600-650, 1032-1048):
 * result[0] = a.b (binary32 result of a double sum). */
int gpucad_svl_dot(float *d_result, const float *d_a, const float *d_b, uint64_t n);
/* Gratings::fillfrac (Gratings.h:71, Gratings.cu:755-772, 1000-1006): svl *= fillfrac. */
int gpucad_svl_fillfrac(float *d_svl, const float *d_fillfrac, int32_t nx, int32_t ny,
                        int32_t nz);
/* Gratings::refine (Gratings.h:67, Gratings.cu:689-722 with the texture of
 * Interpolations.cu:57-107): trilinear sample of the coarse field at
 * texel coordinate (f*hx, f*hy, f*hz) per fine point f, clamped (the
 * reference passes hx = dx2 = 0.5). */
int gpucad_svl_refine(const float *d_coarse, int32_t nx, int32_t ny, int32_t nz, float *d_fine,
                      int32_t fx, int32_t fy, int32_t fz, float hx, float hy, float hz);
/* Gratings::grating (Gratings.h:65, Gratings.cu:653-687): (cos, sin) of the
 * same sample of phi. */
int gpucad_svl_grating(const float *d_phi, int32_t nx, int32_t ny, int32_t nz, gc_float2 *d_grating,
                       int32_t fx, int32_t fy, int32_t fz, float hx, float hy, float hz);
/* Gratings::svl (Gratings.h:69, Gratings.cu:724-752): svl += Re(grating *
 * coeffs[index]). */
int gpucad_svl_accumulate(float *d_svl, const gc_float2 *d_grating, uint64_t n, int32_t index,
                          const gc_float2 *d_coeffs);
/* Fft_lattice::fft_fill (Fft_lattice.h:21, Fft_lattice.cu:150-225): the
 * half spectrum ((nx/2+1) x ny x nz) into the full one with its conjugate
 * mirror (cubic grids, as the reference's mirror indices assume). */
int gpucad_fft_fill(const gc_float2 *d_half, gc_float2 *d_full, int32_t nx, int32_t ny,
                    int32_t nz);
/* Fft_lattice::fft_scalar (Fft_lattice.h:20, Fft_lattice.cu:124-147): data /= scalar. */
int gpucad_fft_scale(gc_float2 *d_data, float scalar, uint64_t n);

/* The unit cell's transforms.  gpucad_fft_plan3d records the shape on the
 * current device, replacing the caller's three cufftPlan3d(&plan, Nxu, Nyu,
 * Nzu, CUFFT_R2C / C2R / C2C) (src/main.cu:3522-3526): n0 the slowest and n2
 * the fastest axis, as cufftPlan3d's (nx, ny, nz); axes up to 2048 points.
 * gpucad_fft_destroy replaces the cufftDestroy calls (src/main.cu:5070-5074).
 * The transforms are cuFFT's (unnormalised, forward e^{-2 pi i jk/n}, inverse
 * e^{+2 pi i jk/n}), in place, synchronous; the real layout pads rows of n2
 * reals to 2 (n2/2 + 1) floats and the half spectrum keeps n2/2 + 1 complex
 * values per row (direct DFT per axis, FP64 accumulation). */
int gpucad_fft_plan3d(int32_t n0, int32_t n1, int32_t n2);
int gpucad_fft_destroy(void);
/* Fft_lattice::fft_func (Fft_lattice.h:16, Fft_lattice.cu:107-112):
 * cufftExecR2C in place. */
int gpucad_fft_r2c(gc_float2 *data);
/* Fft_lattice::ifft_func (Fft_lattice.h:17, Fft_lattice.cu:114-119):
 * cufftExecC2R in place (the imaginary parts of each row's first coefficient,
 * and of its n2/2-th for even n2, are not used). */
int gpucad_fft_c2r(gc_float2 *data);
/* Fft_lattice::ifft_func_complex (Fft_lattice.h:18, Fft_lattice.cu:121-126):
 * cufftExecC2C(..., CUFFT_INVERSE) in place. */
int gpucad_fft_c2c_inverse(gc_float2 *data);
/* Fft_lattice::add_fft_constants (Fft_lattice.h:22, Fft_lattice.cu:277-289):
 * grating[index] = value. */
int gpucad_add_fft_constant(gc_float2 *grating, gc_float2 value, int32_t index);
/* Fft_lattice::indiviual_grating_sum (Fft_lattice.h:21, Fft_lattice.cu:255-274):
 * fft_function[i] += sum_gratings[i].x for i < nx ny nz. */
int gpucad_grating_sum(const gc_float2 *sum_gratings, float *fft_function, int32_t nx, int32_t ny,
                       int32_t nz);

/* The (2 range + 1)^3 Fourier coefficients of a unit cell that
 * unit_lattice keeps (src/main.cu:3578-3711: forward real transform scaled
 * by 1/(nx ny nz), coefficient (i, j, k) at index (i+r) + (2r+1)((j+r) +
 * (2r+1)(k+r)), loop order k, j, i), by direct DFT with FP64 accumulation.
 * No reference counterpart as one call (the reference runs a full cuFFT and
 * picks these on the host); range <= 4. */
int gpucad_unit_spectrum(const float *d_cell, int32_t nx, int32_t ny, int32_t nz, int32_t range,
                         gc_float2 *d_coeffs);

/* The whole synthesis loop of spatial_lattice_run in one call (native form).
 * The right-hand side of finding_phi is linear in (i, j, k), so the phase of
 * harmonic (i, j, k) is i Phi_x + j Phi_y + k Phi_z with Phi_* the solutions
 * for the unit wave numbers: three solves (together, to cg_end_res/(3 range)
 * so every harmonic meets cg_end_res) instead of one per harmonic, then one
 * pass that upsamples the three phases and sums the first `harmonics`
 * coefficients (reference order) per fine point. */
typedef struct gc_svl_params {
    int32_t nx, ny, nz;          /* coarse grid (NumX, NumY, NumZ), >= 4 each */
    float dx, dy, dz;            /* coarse spacing (the reference: 1) */
    char lattice_type;           /* 'n' 'b' 'r' 's' (latticetype_one) */
    char pad[3];
    int32_t uniform_type;        /* 0 const, 1 per axis, 2 period field */
    float const_period, x_period, y_period, z_period;
    float lcon, lcon_1;          /* the 's' type's sine amplitude and frequency */
    int32_t sinewave_zaxis;
    int32_t range;               /* coefficients (2 range + 1)^3, range <= 4 */
    int32_t harmonics;           /* coefficients summed (the reference: half) */
    int32_t cg_max_iter;         /* the reference: 500 */
    float cg_end_res;            /* the reference: 0.01 (on sqrt(r.r)) */
    int32_t fx, fy, fz;          /* fine grid (2 NumX, ...) */
    float fine_dx, fine_dy, fine_dz;  /* fine spacing in coarse texels (0.5) */
} gc_svl_params;

typedef struct gc_svl_result {
    int32_t cg_iters[3];         /* iterations of the x, y, z unit solves */
    float cg_res[3];             /* their final sqrt(r.r) */
} gc_svl_result;

int gpucad_svl_field(const gc_svl_params *p, const float *d_period, const gc_float2 *d_coeffs,
                     float *d_svl, gc_svl_result *result, void *stream);

/* ------------------------------------------------------------------------ */
/* Drop-in extraction entry points (non-indexed float4 output)              */
/* ------------------------------------------------------------------------ */
/* MarchingCubeCuda::classifyVoxel_lattice_new (src/MarchingCubes_kernel.h:75,
 * src/MarchingCubes_kernel.cu:3207-3265): per cell cube = sum [f_k < iso] 2^k,
 * voxelVerts = numVerts[cube], voxelOccupied = voxelVerts > 0. */
int gpucad_classify_voxel_lattice_new(uint32_t *voxel_verts, uint32_t *voxel_occupied,
                                      const float *volume, gc_uint3 grid_size,
                                      gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                      uint32_t num_voxels, gc_float3 voxel_size, float iso);
/* MarchingCubeCuda::ThrustScanWrapper_lattice (src/MarchingCubes_kernel.h:123,
 * src/MarchingCubes_kernel.cu:3198-3203): exclusive uint32 scan (wrapping). */
int gpucad_exclusive_scan_u32(uint32_t *output, const uint32_t *input, uint32_t n);
/* MarchingCubeCuda::compactVoxels_lattice (src/MarchingCubes_kernel.h:79,
 * src/MarchingCubes_kernel.cu:1594-1615). */
int gpucad_compact_voxels(uint32_t *compacted, const uint32_t *voxel_occupied,
                          const uint32_t *voxel_occupied_scan, uint32_t num_voxels);
/* MarchingCubeCuda::generateTriangles_lattice_new (src/MarchingCubes_kernel.h:108,
 * src/MarchingCubes_kernel.cu:3424-3588): band interpolation over two fields. */
int gpucad_generate_triangles_lattice_new(gc_float4 *pos, gc_float4 *norm,
                                          const uint32_t *compacted,
                                          const uint32_t *num_verts_scanned,
                                          const float *volume, gc_uint3 grid_size,
                                          gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                          gc_float3 voxel_size, gc_float3 grid_center,
                                          float iso, uint32_t active_voxels,
                                          uint32_t max_verts, uint32_t total_verts,
                                          const float *volume_one, const float *volume_two,
                                          float isovalue1, float isovalue2, float iso1,
                                          float iso2);
/* MarchingCubeCuda::generateTriangles_lattice_newone (src/MarchingCubes_kernel.h:113,
 * src/MarchingCubes_kernel.cu:3678-3836): band interpolation over one field. */
int gpucad_generate_triangles_lattice_newone(gc_float4 *pos, gc_float4 *norm,
                                             const uint32_t *compacted,
                                             const uint32_t *num_verts_scanned,
                                             const float *volume, gc_uint3 grid_size,
                                             gc_uint3 grid_size_shift,
                                             gc_uint3 grid_size_mask, gc_float3 voxel_size,
                                             gc_float3 grid_center, float iso,
                                             uint32_t active_voxels, uint32_t max_verts,
                                             uint32_t total_verts, const float *volume_one,
                                             float isovalue1, float isovalue2);
/* Textbook generate (vertexInterp4, src/MarchingCubes_kernel.cu:1619-1629,
 * emission as in 3041-3170) over the classification of
 * gpucad_classify_voxel_lattice_new: interpolates `volume` itself at iso. */
int gpucad_generate_triangles_standard(gc_float4 *pos, gc_float4 *norm,
                                       const uint32_t *compacted,
                                       const uint32_t *num_verts_scanned,
                                       const float *volume, gc_uint3 grid_size,
                                       gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                       gc_float3 voxel_size, gc_float3 grid_center, float iso,
                                       uint32_t active_voxels, uint32_t max_verts);
/* Isosurface::computeIsosurface_lattice (src/Isosurface.h:60-64,
 * src/Isosurface.cu:401-486): classify -> scan -> compact -> scan -> generate
 * (band, two fields).  active_voxels and total_verts are host outputs; writes
 * past max_verts-3 are dropped while *total_verts stays unclamped. */
int gpucad_compute_isosurface_lattice(const float *vol, gc_float4 *pos, gc_float4 *norm,
                                      float iso, uint32_t num_voxels, uint32_t *d_voxel_verts,
                                      uint32_t *d_voxel_verts_scan,
                                      uint32_t *d_voxel_occupied,
                                      uint32_t *d_voxel_occupied_scan, gc_uint3 grid_size,
                                      gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                      gc_float3 voxel_size, gc_float3 grid_center,
                                      uint32_t *active_voxels, uint32_t *total_verts,
                                      uint32_t *d_comp_voxel_array, uint32_t max_verts,
                                      const float *vol_one, const float *vol_two,
                                      float isovalue1, float isovalue2, float iso1, float iso2);
/* gpucad_compute_isosurface_lattice_noalloc: no counterpart in the reference.
 * For callers that cannot allocate during the call (code around a HIP graph
 * capture, tight render or serving loops, harnesses that forbid hipMalloc):
 * the arguments, outputs, status codes and results of
 * gpucad_compute_isosurface_lattice, with the scans' tile sums in a fixed
 * per-device array (2 MiB, held until the process ends) and no allocation of
 * device or host memory.  gpucad_init must have been called on the device
 * (GPUCAD_E_ARG otherwise); num_voxels <= 2^30 (GPUCAD_E_ARG beyond).  Calls
 * on one device share the array and are serialised by a per-device lock. */
int gpucad_compute_isosurface_lattice_noalloc(
    const float *vol, gc_float4 *pos, gc_float4 *norm, float iso, uint32_t num_voxels,
    uint32_t *d_voxel_verts, uint32_t *d_voxel_verts_scan, uint32_t *d_voxel_occupied,
    uint32_t *d_voxel_occupied_scan, gc_uint3 grid_size, gc_uint3 grid_size_shift,
    gc_uint3 grid_size_mask, gc_float3 voxel_size, gc_float3 grid_center, uint32_t *active_voxels,
    uint32_t *total_verts, uint32_t *d_comp_voxel_array, uint32_t max_verts, const float *vol_one,
    const float *vol_two, float isovalue1, float isovalue2, float iso1, float iso2);
/* Isosurface::computeIsosurface_latticeone (src/Isosurface.h:66-70,
 * src/Isosurface.cu:488-572) */
int gpucad_compute_isosurface_latticeone(const float *vol, gc_float4 *pos, gc_float4 *norm,
                                         float iso, uint32_t num_voxels,
                                         uint32_t *d_voxel_verts, uint32_t *d_voxel_verts_scan,
                                         uint32_t *d_voxel_occupied,
                                         uint32_t *d_voxel_occupied_scan, gc_uint3 grid_size,
                                         gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                         gc_float3 voxel_size, gc_float3 grid_center,
                                         uint32_t *active_voxels, uint32_t *total_verts,
                                         uint32_t *d_comp_voxel_array, uint32_t max_verts,
                                         const float *vol_one, float isovalue1,
                                         float isovalue2);
/* Same orchestration with the textbook generate (gpucad_generate_triangles_standard). */
int gpucad_compute_isosurface_standard(const float *vol, gc_float4 *pos, gc_float4 *norm,
                                       float iso, uint32_t num_voxels, uint32_t *d_voxel_verts,
                                       uint32_t *d_voxel_verts_scan,
                                       uint32_t *d_voxel_occupied,
                                       uint32_t *d_voxel_occupied_scan, gc_uint3 grid_size,
                                       gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                       gc_float3 voxel_size, gc_float3 grid_center,
                                       uint32_t *active_voxels, uint32_t *total_verts,
                                       uint32_t *d_comp_voxel_array, uint32_t max_verts);

/* ------------------------------------------------------------------------ */
/* Drop-in boolean meshing (primitives with each other or with a lattice band) */
/* ------------------------------------------------------------------------ */
/* The reference's boolean switches as one bit set (bools of
 * Isosurface::computeIsosurface, src/Isosurface.h:28-33).  primitive_fixed
 * points to grid_points records {int val; float t_x, t_y, t_z;} (16 bytes). */
enum {
    GC_BOOL_UNION = 1,
    GC_BOOL_DIFF = 2,
    GC_BOOL_INTERSECT = 4,
    GC_BOOL_FIXED = 8,
    GC_BOOL_DYNAMIC = 16,
    GC_BOOL_MAKE_REGION = 32
};
/* MarchingCubeCuda::classifyVoxel_lattice (src/MarchingCubes_kernel.h,
 * kernel classifyVoxel src/MarchingCubes_kernel.cu:869-1047): cube index from
 * the boolean occupancy of each corner. */
int gpucad_classify_voxel_boolean(uint32_t *voxel_verts, uint32_t *voxel_occupied,
                                  const void *primitive_fixed, const float *primitive_dynamic,
                                  const float *lattice_field, gc_uint3 grid_size,
                                  gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                  uint32_t num_voxels, float iso1, float iso2, float iso,
                                  uint32_t flags);
/* MarchingCubeCuda::classify_copy_Voxel_lattice (kernel classify_copy_Voxel,
 * src/MarchingCubes_kernel.cu:158-446): rewrites vol_one's val as -1 (inside
 * the boolean) / 1 and blends the crossing parameters of each point's +x/+y/+z
 * edges into t_x/t_y/t_z; grid_size = points per axis; every point but the
 * last. */
int gpucad_classify_copy_voxel(void *vol_one, const float *volume_two, const float *vol_lattice,
                               gc_uint3 grid_size, float iso1, float iso2, float iso,
                               uint32_t flags);
/* MarchingCubeCuda::generateTriangles_lattice (kernel
 * generateTriangles_lattice_kernel, src/MarchingCubes_kernel.cu:1865-2199):
 * non-indexed pos/norm, normal cross(v2 - v0, v1 - v0) with w = 0.5. */
int gpucad_generate_triangles_boolean(gc_float4 *pos, gc_float4 *norm, const uint32_t *compacted,
                                      const uint32_t *num_verts_scanned, gc_uint3 grid_size,
                                      gc_uint3 grid_size_shift, gc_uint3 grid_size_mask,
                                      gc_float3 voxel_size, gc_float3 grid_center, float iso,
                                      uint32_t active_voxels, uint32_t max_verts,
                                      const void *primitive_fixed,
                                      const float *primitive_dynamic, const float *lattice_field,
                                      float iso1, float iso2, uint32_t flags);
/* Isosurface::computeIsosurface (src/Isosurface.h:28-33, src/Isosurface.cu:44-134):
 * classify, scans, compaction and generate; fills *active_voxels,
 * *total_verts and *nfacets = total/3 (left unchanged when nothing is
 * active, as the reference does). */
int gpucad_compute_isosurface(gc_float4 *pos, gc_float4 *norm, float iso, uint32_t num_voxels,
                              uint32_t *d_voxel_verts, uint32_t *d_voxel_verts_scan,
                              uint32_t *d_voxel_occupied, uint32_t *d_voxel_occupied_scan,
                              gc_uint3 grid_size, gc_uint3 grid_size_shift,
                              gc_uint3 grid_size_mask, gc_float3 voxel_size,
                              gc_float3 grid_center, uint32_t *active_voxels,
                              uint32_t *total_verts, uint32_t *d_comp_voxel_array,
                              uint32_t max_verts, const void *primitive_fixed,
                              const float *primitive_dynamic, const float *lattice_field,
                              float iso1, float iso2, uint32_t flags, uint64_t *nfacets);

/* ------------------------------------------------------------------------ */
/* Graphics-API interop (pos/norm buffers and semaphores shared with Vulkan)  */
/* ------------------------------------------------------------------------ */
typedef struct gc_external_buffer gc_external_buffer;       /* opaque */
typedef struct gc_external_semaphore gc_external_semaphore; /* opaque */
/* importCudaExternalMemory (src/main.cu:1164-1191): imports the allocation of
 * `size` bytes exported as an opaque POSIX fd (vkGetMemoryFdKHR) and maps
 * bytes [offset, size) as a device buffer (*d_ptr; the reference maps offset
 * 0).  As with the CUDA call, a successful import owns the fd.  The extraction
 * entry points write into *d_ptr in place. */
int gpucad_import_external_buffer(int32_t fd, uint64_t size, uint64_t offset,
                                  gc_external_buffer **buf, void **d_ptr);
/* cudaDestroyExternalMemory (src/main.cu:754-931): unmaps and releases. */
int gpucad_release_external_buffer(gc_external_buffer *buf);
/* importCudaExternalSemaphore (src/main.cu:1193-1211): binary semaphore
 * exported as an opaque fd. */
int gpucad_import_external_semaphore(int32_t fd, gc_external_semaphore **sem);
/* cudaSignalExternalSemaphoresAsync / cudaWaitExternalSemaphoresAsync with
 * fence value 0 (src/main.cu:2017-2035, 3122-3134), enqueued on `stream`. */
int gpucad_signal_external_semaphore(gc_external_semaphore *sem, void *stream);
int gpucad_wait_external_semaphore(gc_external_semaphore *sem, void *stream);
/* cudaDestroyExternalSemaphore (src/main.cu:739-744). */
int gpucad_release_external_semaphore(gc_external_semaphore *sem);

/* ------------------------------------------------------------------------ */
/* .obj output                                                               */
/* ------------------------------------------------------------------------ */
/* File_output::file_write_obj (src/File_output.h:189, src/File_output.cu:5-81):
 * byte-identical output — weld key int(v*1000)*0.001 per component, vertices
 * numbered in first-seen order, faces with 3 distinct ids kept once per
 * ordered triple, written "f a c b". */
int gpucad_file_write_obj(const gc_float4 *d_pos, uint32_t total_verts, const char *filename);
/* Native writer from indexed device buffers (vertices float3, triangles
 * uint32 x3): "v x y z" lines then "f a b c" (1-based; flip_winding writes
 * "f a c b" like the reference). */
int gpucad_write_obj_indexed(const float *d_vertices, uint64_t n_vertices,
                             const uint32_t *d_triangles, uint64_t n_triangles,
                             const char *filename, int32_t flip_winding);

#ifdef __cplusplus
}
#endif

#endif /* GPUCAD_H_ */
