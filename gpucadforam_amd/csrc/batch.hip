// Batched form: many lattice grids of uneven sizes and types per run.
//
// Field: the grids' bricks (BX x 4*(64/BX) columns x lz planes, lattice_body.hpp)
// are listed per field variant as (grid, bx, by, bz) codes; one launch per
// variant present runs one workgroup per brick.  Extraction: the grids' word
// blocks are numbered globally (each grid a contiguous range), and the count,
// vert and tri stages run one workgroup per global block (extract.hip); the
// block scan produces one vertex and one triangle range per grid in the
// batch's mesh buffers.  A run has no memset node, so it can be captured once
// as a HIP graph and replayed.
//
// The field's brick shape (BX, lz) is picked by timing candidates on the
// batch itself; the choice is cached per device and batch signature (in
// memory, and in the file GPUCAD_TUNE_CACHE names, if set: lines
// "<format tag> <signature> <bx> <lz>", other tags and invalid shapes are
// skipped on load).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "plan.hpp"

namespace gc {

namespace {

int variant_index(const gc_lattice_params &p) { return field_variant(&p).index(); }

template <bool ROT, int TPMS, bool THICK, int BX>
__global__ void __launch_bounds__(64 * kFieldWaves)
batch_field_kernel(const FieldConsts *__restrict__ consts, const uint64_t *__restrict__ items,
                   uint32_t n, float *fields, uint64_t *signs) {
    __shared__ float ztab[kMaxLZ * kZTab];
    const uint32_t it = blockIdx.x;  // one workgroup per brick
    if (it >= n) return;
    const uint64_t code = items[it];
    const uint32_t seg = (uint32_t)(code & 0xffff);
    const int bx = (int)((code >> 16) & 0xffff), by = (int)((code >> 32) & 0xffff),
              bz = (int)(code >> 48);
    const FieldConsts &c = consts[seg];
    lattice_brick<ROT, TPMS, THICK, BX>(c, fields + c.field_off, signs + c.sign_off, bx * BX,
                                        by * kFieldWaves * (64 / BX), bz * c.lz, ztab);
}

typedef void (*BatchFieldKernel)(const FieldConsts *, const uint64_t *, uint32_t, float *,
                                 uint64_t *);

template <bool ROT, int TPMS, bool THICK>
BatchFieldKernel pick_bx(int bx) {
    switch (bx) {
        case 64: return batch_field_kernel<ROT, TPMS, THICK, 64>;
        case 32: return batch_field_kernel<ROT, TPMS, THICK, 32>;
        case 16: return batch_field_kernel<ROT, TPMS, THICK, 16>;
        default: return batch_field_kernel<ROT, TPMS, THICK, 8>;
    }
}

template <bool ROT, int TPMS>
BatchFieldKernel pick_thick(bool thick, int bx) {
    return thick ? pick_bx<ROT, TPMS, true>(bx) : pick_bx<ROT, TPMS, false>(bx);
}

template <bool ROT>
BatchFieldKernel pick_tpms(int tpms, bool thick, int bx) {
    switch (tpms) {
        case GC_TPMS_GYROID: return pick_thick<ROT, GC_TPMS_GYROID>(thick, bx);
        case GC_TPMS_SCHWARZ_P: return pick_thick<ROT, GC_TPMS_SCHWARZ_P>(thick, bx);
        default: return pick_thick<ROT, GC_TPMS_DIAMOND>(thick, bx);
    }
}

BatchFieldKernel field_kernel_for(int variant, int bx) {
    const bool rot = variant >= 6, thick = (variant & 1) != 0;
    const int tpms = (variant % 6) / 2;
    return rot ? pick_tpms<true>(tpms, thick, bx) : pick_tpms<false>(tpms, thick, bx);
}

}  // namespace

uint64_t brick_count(const gc_lattice_params &p, int bx, int lz) {
    const int by = kFieldWaves * (64 / bx);
    return (uint64_t)((p.nx + bx - 1) / bx) * (uint64_t)((p.ny + by - 1) / by) *
           (uint64_t)((p.nz + lz - 1) / lz);
}

int validate_batch_grid(const gc_lattice_params &p, int i) {
    int st = validate_lattice(&p);
    if (st) return fail(st, "grid %d: %s", i, gpucad_last_error());
    GC_ARG(p.nx <= 65536 && p.ny < 65536 && p.nz < 65536, "grid %d too large", i);
    GC_ARG(gpucad_sign_words(p.nx, p.ny, p.nz) < 0xffffffffull, "grid %d too large", i);
    return GPUCAD_OK;
}

// Grids, constants and segments; field_off/sign_off per grid are given.
void plan_segments(const gc_lattice_params *p, int n, const uint64_t *field_off,
                   const uint64_t *sign_off, float iso, int unit_domain, HostPlan &h) {
    h.consts.resize(n);
    h.segs.resize(n);
    h.blk2seg.clear();
    uint32_t block = 0;
    for (int i = 0; i < n; ++i) {
        FieldConsts c = make_field_consts(&p[i], iso, 16);
        c.field_off = field_off[i];
        c.sign_off = sign_off[i];
        h.consts[i] = c;
        Seg s;
        std::memset(&s, 0, sizeof(s));
        s.g = make_grid(p[i].nx, p[i].ny, p[i].nz);
        s.field_off = field_off[i];
        s.sign_off = sign_off[i];
        s.block_off = block;
        s.iso = iso;
        s.sx = unit_domain ? (float)(1.0 / (p[i].nx - 1)) : 1.0f;
        s.sy = unit_domain ? (float)(1.0 / (p[i].ny - 1)) : 1.0f;
        s.sz = unit_domain ? (float)(1.0 / (p[i].nz - 1)) : 1.0f;
        h.segs[i] = s;
        const uint32_t nbg = (uint32_t)((s.g.words + kBlock - 1) / kBlock);
        h.blk2seg.insert(h.blk2seg.end(), nbg, (uint32_t)i);
        block += nbg;
    }
    h.nb = block;
}

// Brick list of a shape, grouped by field variant.
void plan_items(const gc_lattice_params *p, int n, const Shape &sh, HostPlan &h) {
    h.items.clear();
    for (int i = 0; i < n; ++i) h.consts[i].lz = sh.lz;
    const int by = kFieldWaves * (64 / sh.bx);
    for (int v = 0; v < kVariants; ++v) {
        h.vbegin[v] = (uint32_t)h.items.size();
        for (int i = 0; i < n; ++i) {
            if (variant_index(p[i]) != v) continue;
            const int nbx = (p[i].nx + sh.bx - 1) / sh.bx, nby = (p[i].ny + by - 1) / by,
                      nbz = (p[i].nz + sh.lz - 1) / sh.lz;
            for (int z = 0; z < nbz; ++z)
                for (int y = 0; y < nby; ++y)
                    for (int x = 0; x < nbx; ++x)
                        h.items.push_back((uint64_t)i | ((uint64_t)x << 16) |
                                          ((uint64_t)y << 32) | ((uint64_t)z << 48));
        }
    }
    h.vbegin[kVariants] = (uint32_t)h.items.size();
}

// Enqueues the field launches of a plan.
int enqueue_field_plan(const HostPlan &h, const DevPlan &d, const Shape &sh, float *fields,
                       uint64_t *signs, hipStream_t stream) {
    for (int v = 0; v < kVariants; ++v) {
        const uint32_t n = h.vbegin[v + 1] - h.vbegin[v];
        if (!n) continue;
        hipLaunchKernelGGL(field_kernel_for(v, sh.bx), dim3(n), dim3(64 * kFieldWaves), 0,
                           stream, d.consts, d.items + h.vbegin[v], n, fields, signs);
        GC_CHECK_LAUNCH();
    }
    return GPUCAD_OK;
}

BatchExtract make_batch_extract(const HostPlan &h, const DevPlan &d, const float *fields,
                                const uint64_t *signs, float *vertices, uint64_t vcap,
                                uint32_t *triangles, uint64_t tcap) {
    BatchExtract b;
    b.fields = fields;
    b.signs = signs;
    b.segs = d.segs;
    b.blk2seg = d.blk2seg;
    b.nseg = (uint32_t)h.segs.size();
    b.nb = h.nb;
    b.base = d.base;
    b.twave = d.twave;
    b.bsum = d.bsum;
    b.ctl = d.ctl;
    b.res = d.res;
    b.vertices = vertices;
    b.vcap = vertices ? vcap : 0;
    b.triangles = triangles;
    b.tcap = triangles ? tcap : 0;
    return b;
}

int device_cus(int *cus) {
    int dev = 0;
    GC_HIP(hipGetDevice(&dev));
    GC_HIP(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev));
    return GPUCAD_OK;
}

int default_shape(Shape *s) {
    *s = Shape();
    return GPUCAD_OK;
}

// ---------------------------------------------------------------------------
// Tuning cache
// ---------------------------------------------------------------------------
namespace {

std::mutex g_tune_mu;
std::map<std::string, Shape> g_tune;
bool g_tune_loaded = false;

// Tag of the cache file's lines: a file written by another build of the
// tuner (other shapes or kernels) is ignored rather than trusted.
constexpr const char *kTuneFormat = "gpucad-tune-2";

bool valid_shape(const Shape &s) {
    return (s.bx == 64 || s.bx == 32 || s.bx == 16 || s.bx == 8) && s.lz >= 1 && s.lz <= kMaxLZ;
}

void tune_load_locked() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char *path = std::getenv("GPUCAD_TUNE_CACHE");
    if (!path || !*path) return;
    FILE *f = std::fopen(path, "r");
    if (!f) return;
    char line[512], tag[64], key[256];
    while (std::fgets(line, sizeof(line), f)) {
        Shape s;
        if (std::sscanf(line, "%63s %255s %d %d", tag, key, &s.bx, &s.lz) != 4) continue;
        if (std::strcmp(tag, kTuneFormat) != 0 || !valid_shape(s)) continue;
        g_tune[key] = s;
    }
    std::fclose(f);
}

bool tune_lookup(const std::string &key, Shape *s) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_locked();
    auto it = g_tune.find(key);
    if (it == g_tune.end()) return false;
    *s = it->second;
    return true;
}

void tune_store(const std::string &key, const Shape &s) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_locked();
    g_tune[key] = s;
    const char *path = std::getenv("GPUCAD_TUNE_CACHE");
    if (!path || !*path) return;
    FILE *f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s %s %d %d\n", kTuneFormat, key.c_str(), s.bx, s.lz);
    std::fclose(f);
}

// Device name, CU count, grid count, 2-digit total points and variant mask.
std::string batch_signature(const gc_lattice_params *p, int n, int cus) {
    hipDeviceProp_t prop;
    int dev = 0;
    char arch[64] = "unknown";
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        std::snprintf(arch, sizeof(arch), "%s", prop.gcnArchName);
    for (char *c = arch; *c; ++c)
        if (*c == ' ' || *c == ':') *c = '_';
    double pts = 0;
    unsigned mask = 0;
    for (int i = 0; i < n; ++i) {
        pts += (double)p[i].nx * p[i].ny * p[i].nz;
        mask |= 1u << variant_index(p[i]);
    }
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s/cu%d/n%d/pts%.1e/v%x", arch, cus, n, pts, mask);
    return buf;
}

}  // namespace

// ---------------------------------------------------------------------------
// The batch object
// ---------------------------------------------------------------------------
struct Batch {
    int n = 0, cus = 0;
    std::vector<gc_lattice_params> grids;
    gc_batch_options opt = {};
    Shape shape;
    HostPlan plan;
    DevPlan d;
    char *arena = nullptr;
    uint64_t *items_buf = nullptr;
    size_t items_cap = 0;
    float *fields = nullptr;
    uint64_t field_floats = 0;
    uint64_t *signs = nullptr;
    float *vertices = nullptr;
    uint32_t *triangles = nullptr;
    uint64_t vcap = 0, tcap = 0;
    std::vector<uint64_t> field_off;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr, t0 = nullptr, t1 = nullptr;

    int enqueue(hipStream_t s, int phases_extract, bool field = true) {
        int st = GPUCAD_OK;
        if (field) {
            st = enqueue_field_plan(plan, d, shape, fields, signs, s);
            if (st) return st;
        }
        if (phases_extract) {
            const BatchExtract be = make_batch_extract(plan, d, fields, signs,
                                                       vertices, vcap, triangles, tcap);
            st = enqueue_extract_batch(be, phases_extract, s);
        }
        return st;
    }

    int upload_items() {
        if (plan.items.size() > items_cap) return fail(GPUCAD_E_ARG, "brick list too long");
        GC_HIP(hipMemcpy(d.consts, plan.consts.data(), plan.consts.size() * sizeof(FieldConsts),
                         hipMemcpyHostToDevice));
        GC_HIP(hipMemcpy(items_buf, plan.items.data(), plan.items.size() * sizeof(uint64_t),
                         hipMemcpyHostToDevice));
        return GPUCAD_OK;
    }

    // Smallest time (ms) of `reps` runs of enqueue(phases, field).
    int time_runs(int phases, bool field, int reps, float *ms) {
        float best = 1e30f;
        for (int r = 0; r < reps + 1; ++r) {
            GC_HIP(hipEventRecord(t0, cap));
            int st = enqueue(cap, phases, field);
            if (st) return st;
            GC_HIP(hipEventRecord(t1, cap));
            GC_HIP(hipEventSynchronize(t1));
            float e = 0.f;
            GC_HIP(hipEventElapsedTime(&e, t0, t1));
            if (r > 0) best = std::min(best, e);  // the first run warms up
        }
        *ms = best;
        return GPUCAD_OK;
    }

    int tune(const std::string &key) {
        const bool fix_bx = opt.brick_x != 0, fix_lz = opt.brick_lz != 0;
        const std::vector<int> bxs = fix_bx ? std::vector<int>{opt.brick_x}
                                            : std::vector<int>{64, 32, 16, 8};
        const std::vector<int> lzs = fix_lz ? std::vector<int>{opt.brick_lz}
                                            : std::vector<int>{8, 16, 32};
        Shape best = shape;
        float best_ms = 1e30f;
        for (int bx : bxs)
            for (int lz : lzs) {
                Shape s = shape;
                s.bx = bx;
                s.lz = lz;
                shape = s;
                plan_items(grids.data(), n, shape, plan);
                int st = upload_items();
                if (st) return st;
                float ms = 0.f;
                st = time_runs(0, true, 2, &ms);
                if (st) return st;
                if (ms < best_ms) {
                    best_ms = ms;
                    best = s;
                }
            }
        shape = best;
        plan_items(grids.data(), n, shape, plan);
        int st = upload_items();
        if (st) return st;
        tune_store(key, shape);
        return GPUCAD_OK;
    }

    void release() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        if (done) (void)hipEventDestroy(done);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        if (arena) (void)hipFree(arena);
        if (items_buf) (void)hipFree(items_buf);
        if (vertices) (void)hipFree(vertices);
        if (triangles) (void)hipFree(triangles);
        exec = nullptr;
        graph = nullptr;
        cap = nullptr;
        done = t0 = t1 = nullptr;
        arena = nullptr;
        items_buf = nullptr;
        vertices = nullptr;
        triangles = nullptr;
    }
};

int batch_create(const gc_lattice_params *grids, int n, const gc_batch_options *o, Batch **out) {
    GC_ARG(grids && out && o, "NULL argument");
    GC_ARG(n >= 1 && n <= 65535, "grid count must be 1..65535");
    GC_ARG(o->brick_x == 0 || o->brick_x == 64 || o->brick_x == 32 || o->brick_x == 16 ||
               o->brick_x == 8,
           "brick_x must be 0, 64, 32, 16 or 8");
    GC_ARG(o->brick_lz >= 0 && o->brick_lz <= kMaxLZ, "brick_lz must be 0..%d", kMaxLZ);
    GC_ARG(o->waves_per_cu == 0,
           "waves_per_cu must be 0: the field runs one workgroup per brick");
    for (int i = 0; i < n; ++i) {
        int st = validate_batch_grid(grids[i], i);
        if (st) return st;
    }
    const DeviceTables *tabs = nullptr;
    int st = device_tables(&tabs);
    if (st) return st;
    Batch *b = new Batch();
    b->n = n;
    b->grids.assign(grids, grids + n);
    b->opt = *o;
    st = device_cus(&b->cus);
    auto bail = [&](int s) {
        b->release();
        delete b;
        return s;
    };
    if (st) return bail(st);
    // ---- placement
    b->field_off.resize(n);
    std::vector<uint64_t> sign_off(n);
    uint64_t fo = 0, so = 0;
    for (int i = 0; i < n; ++i) {
        b->field_off[i] = fo;
        sign_off[i] = so;
        fo += round_up((uint64_t)grids[i].nx * grids[i].ny * grids[i].nz, 64);
        so += round_up(gpucad_sign_words(grids[i].nx, grids[i].ny, grids[i].nz), 32);
    }
    b->field_floats = fo;
    plan_segments(grids, n, b->field_off.data(), sign_off.data(), o->iso, o->unit_domain,
                  b->plan);
    const uint64_t nb = b->plan.nb;
    // ---- one arena for everything but the brick list and the meshes
    struct Part {
        uint64_t off, bytes;
    };
    uint64_t off = 0;
    auto part = [&](uint64_t bytes) {
        Part p{off, bytes};
        off = round_up(off + bytes, 256);
        return p;
    };
    const Part p_consts = part(n * sizeof(FieldConsts)), p_segs = part(n * sizeof(Seg)),
               p_b2s = part(nb * 4), p_res = part(n * sizeof(SegResult)),
               p_base = part(nb * kBaseBytesPerBlock),
               p_twave = part(nb * kTwaveBytesPerBlock), p_bsum = part(nb * 16),
               p_ctl = part(64), p_fields = part(fo * 4),
               p_signs = part(so * 8);
    hipError_t he = hipMalloc(&b->arena, off);
    if (he != hipSuccess) {
        b->arena = nullptr;
        return bail(fail(GPUCAD_E_OOM, "batch allocation of %llu bytes failed",
                         (unsigned long long)off));
    }
    char *A = b->arena;
    b->d.consts = (FieldConsts *)(A + p_consts.off);
    b->d.segs = (Seg *)(A + p_segs.off);
    b->d.blk2seg = (uint32_t *)(A + p_b2s.off);
    b->d.res = (SegResult *)(A + p_res.off);
    b->d.base = (uint32_t *)(A + p_base.off);
    b->d.twave = (uint32_t *)(A + p_twave.off);
    b->d.bsum = (uint64_t *)(A + p_bsum.off);
    b->d.ctl = (uint64_t *)(A + p_ctl.off);
    b->fields = (float *)(A + p_fields.off);
    b->signs = (uint64_t *)(A + p_signs.off);
    // bricks narrower than 64 leave the tail bytes of a row's last word
    // unwritten: they must read as zero
    if (hipMemset(b->signs, 0, p_signs.bytes) != hipSuccess ||
        hipMemcpy(b->d.segs, b->plan.segs.data(), n * sizeof(Seg), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(b->d.blk2seg, b->plan.blk2seg.data(), nb * 4, hipMemcpyHostToDevice) !=
            hipSuccess)
        return bail(fail(GPUCAD_E_HIP, "batch upload failed"));
    // ---- brick list big enough for every candidate shape
    uint64_t most = 0;
    for (int bx : {64, 32, 16, 8})
        for (int lz : {1, 8, 16, 32, o->brick_lz > 0 ? o->brick_lz : 16}) {
            uint64_t c = 0;
            for (int i = 0; i < n; ++i) c += brick_count(grids[i], bx, lz);
            most = std::max(most, c);
        }
    if (most >= 0xffffffffull) return bail(fail(GPUCAD_E_ARG, "too many bricks"));
    b->items_cap = most;
    if (hipMalloc(&b->items_buf, most * 8) != hipSuccess) {
        b->items_buf = nullptr;
        return bail(fail(GPUCAD_E_OOM, "brick list allocation failed"));
    }
    b->d.items = b->items_buf;
    if (hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->done) != hipSuccess || hipEventCreate(&b->t0) != hipSuccess ||
        hipEventCreate(&b->t1) != hipSuccess)
        return bail(fail(GPUCAD_E_HIP, "stream/event creation failed"));
    // ---- shape: given, cached or tuned
    Shape sh;
    sh.bx = o->brick_x ? o->brick_x : 64;
    sh.lz = o->brick_lz ? o->brick_lz : 16;
    b->shape = sh;
    plan_items(grids, n, b->shape, b->plan);
    st = b->upload_items();
    if (st) return bail(st);
    // ---- mesh buffers: sized by a counting run unless given (before tuning,
    // so that the emit stages are timed with their stores)
    b->vcap = o->vert_capacity;
    b->tcap = o->tri_capacity;
    if (!b->vcap || !b->tcap) {
        st = b->enqueue(b->cap, EXTRACT_COUNT, true);
        if (st) return bail(st);
        uint64_t ctl[2];
        if (hipStreamSynchronize(b->cap) != hipSuccess ||
            hipMemcpy(ctl, b->d.ctl, sizeof(ctl), hipMemcpyDeviceToHost) != hipSuccess)
            return bail(fail(GPUCAD_E_HIP, "counting run failed"));
        if (!b->vcap) b->vcap = std::max<uint64_t>(ctl[0], 1);
        if (!b->tcap) b->tcap = std::max<uint64_t>(ctl[1], 1);
    }
    if (hipMalloc(&b->vertices, b->vcap * 12) != hipSuccess ||
        hipMalloc(&b->triangles, b->tcap * 12) != hipSuccess)
        return bail(fail(GPUCAD_E_OOM, "mesh allocation failed (%llu vertices, %llu triangles)",
                         (unsigned long long)b->vcap, (unsigned long long)b->tcap));
    // ---- shape: given, cached or tuned
    const bool want_tune = !o->brick_x || !o->brick_lz;
    const std::string key = batch_signature(grids, n, b->cus);
    Shape cached;
    if (want_tune && tune_lookup(key, &cached)) {
        b->shape = cached;
        if (o->brick_x) b->shape.bx = o->brick_x;
        if (o->brick_lz) b->shape.lz = o->brick_lz;
    } else if (want_tune) {
        st = b->tune(key);
        if (st) return bail(st);
    }
    plan_items(grids, n, b->shape, b->plan);
    st = b->upload_items();
    if (st) return bail(st);
    // ---- graph
    if (o->use_graph) {
        if (hipStreamBeginCapture(b->cap, hipStreamCaptureModeRelaxed) != hipSuccess)
            return bail(fail(GPUCAD_E_HIP, "graph capture failed to start"));
        st = b->enqueue(b->cap, EXTRACT_ALL, true);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(b->cap, &g);
        if (st) {
            if (g) (void)hipGraphDestroy(g);
            return bail(st);
        }
        if (ce != hipSuccess || !g) return bail(fail(GPUCAD_E_HIP, "graph capture failed"));
        b->graph = g;
        if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            b->exec = nullptr;
            return bail(fail(GPUCAD_E_HIP, "graph instantiation failed"));
        }
    }
    if (hipStreamSynchronize(b->cap) != hipSuccess)
        return bail(fail(GPUCAD_E_HIP, "batch set-up failed"));
    *out = b;
    return GPUCAD_OK;
}

}  // namespace gc

struct gc_batch {
    gc::Batch *impl;
};

extern "C" {

int gpucad_batch_create(const gc_lattice_params *grids, int32_t n_grids,
                        const gc_batch_options *opt, gc_batch **out) {
    gc::clear_error();
    GC_ARG(out, "out is NULL");
    *out = nullptr;
    gc::Batch *b = nullptr;
    int st = gc::batch_create(grids, n_grids, opt, &b);
    if (st) return st;
    *out = new gc_batch{b};
    return GPUCAD_OK;
}

int gpucad_batch_run(gc_batch *h, void *stream) {
    gc::clear_error();
    GC_ARG(h && h->impl, "batch is NULL");
    gc::Batch *b = h->impl;
    hipStream_t s = gc::as_stream(stream);
    if (b->exec) {
        GC_HIP(hipGraphLaunch(b->exec, s));
    } else {
        int st = b->enqueue(s, gc::EXTRACT_ALL, true);
        if (st) return st;
    }
    GC_HIP(hipEventRecord(b->done, s));
    return GPUCAD_OK;
}

int gpucad_batch_results(gc_batch *h, gc_batch_grid_result *res, int32_t n) {
    gc::clear_error();
    GC_ARG(h && h->impl && res, "NULL argument");
    gc::Batch *b = h->impl;
    GC_ARG(n == b->n, "result array holds %d entries, the batch has %d grids", n, b->n);
    GC_HIP(hipEventSynchronize(b->done));
    std::vector<gc::SegResult> r(b->n);
    GC_HIP(hipMemcpy(r.data(), b->d.res, b->n * sizeof(gc::SegResult), hipMemcpyDeviceToHost));
    int worst = GPUCAD_OK;
    for (int i = 0; i < b->n; ++i) {
        gc_batch_grid_result &o = res[i];
        o.field_offset = b->field_off[i];
        o.vert_offset = r[i].v_off;
        o.n_vertices = r[i].nv;
        o.tri_offset = r[i].t_off;
        o.n_triangles = r[i].nt;
        const bool over = r[i].ovf || r[i].v_off + r[i].nv > b->vcap ||
                          r[i].t_off + r[i].nt > b->tcap;
        o.status = over ? GPUCAD_E_OVERFLOW : GPUCAD_OK;
        o.pad = 0;
        if (over) worst = GPUCAD_E_OVERFLOW;
    }
    if (worst) return gc::fail(worst, "some meshes exceed the batch's mesh buffers");
    return GPUCAD_OK;
}

int gpucad_batch_buffers(const gc_batch *h, gc_batch_buffers *out) {
    gc::clear_error();
    GC_ARG(h && h->impl && out, "NULL argument");
    const gc::Batch *b = h->impl;
    out->fields = b->fields;
    out->field_floats = b->field_floats;
    out->vertices = b->vertices;
    out->vert_capacity = b->vcap;
    out->triangles = b->triangles;
    out->tri_capacity = b->tcap;
    return GPUCAD_OK;
}

int gpucad_batch_tuning(const gc_batch *h, gc_batch_options *chosen) {
    gc::clear_error();
    GC_ARG(h && h->impl && chosen, "NULL argument");
    const gc::Batch *b = h->impl;
    *chosen = b->opt;
    chosen->brick_x = b->shape.bx;
    chosen->brick_lz = b->shape.lz;
    chosen->waves_per_cu = 0;  // not a tuning dimension: one workgroup per brick
    chosen->vert_capacity = b->vcap;
    chosen->tri_capacity = b->tcap;
    return GPUCAD_OK;
}

int gpucad_batch_destroy(gc_batch *h) {
    gc::clear_error();
    if (!h) return GPUCAD_OK;
    if (h->impl) {
        if (h->impl->done) (void)hipEventSynchronize(h->impl->done);
        if (h->impl->cap) (void)hipStreamSynchronize(h->impl->cap);
        h->impl->release();
        delete h->impl;
    }
    delete h;
    return GPUCAD_OK;
}

}  // extern "C"
