// Library-level C-ABI: version, status strings, thread-local error messages,
// device initialisation (marching-cubes tables, pinned host words, the fixed
// scratch of the allocation-free drop-in) and the drop-in scratch arena.
#include <cstdarg>
#include <cstring>
#include <mutex>

#include "gc_common.hpp"

namespace gc {
namespace {

thread_local char t_error[512] = "";

constexpr int kMaxDevices = 64;

struct DeviceState {
    DeviceTables tables;
    bool tables_ready = false;
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    uint32_t *host_words = nullptr;  // pinned host memory that kernels write
    // held by every synchronous entry point that uses the arena or the host
    // words, so calls from several host threads on one device take turns
    std::recursive_mutex call_mu;
};

DeviceState g_dev[kMaxDevices];
std::mutex g_mu;

int current_device(int *dev) {
    GC_HIP(hipGetDevice(dev));
    if (*dev < 0 || *dev >= kMaxDevices) return fail(GPUCAD_E_ARG, "device %d out of range", *dev);
    return GPUCAD_OK;
}

}  // namespace

int fail(int status, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_error, sizeof(t_error), fmt, ap);
    va_end(ap);
    return status;
}

void clear_error() { t_error[0] = '\0'; }

int device_tables(const DeviceTables **out) {
    int dev = 0;
    int st = current_device(&dev);
    if (st) return st;
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState &ds = g_dev[dev];
    if (!ds.tables_ready) {
        DeviceTables t;
        GC_HIP(hipMalloc(&t.tri_packed, 256 * sizeof(uint64_t)));
        GC_HIP(hipMalloc(&t.num_verts, 256));
        GC_HIP(hipMalloc(&t.tri_table, 256 * 16));
        GC_HIP(hipMemcpy(t.tri_packed, host_tri_packed(), 256 * sizeof(uint64_t),
                         hipMemcpyHostToDevice));
        GC_HIP(hipMemcpy(t.num_verts, host_num_verts(), 256, hipMemcpyHostToDevice));
        GC_HIP(hipMemcpy(t.tri_table, host_tri_table(), 256 * 16, hipMemcpyHostToDevice));
        ds.tables = t;
        ds.tables_ready = true;
    }
    *out = &ds.tables;
    return GPUCAD_OK;
}

int host_words(uint32_t **out) {
    int dev = 0;
    int st = current_device(&dev);
    if (st) return st;
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState &ds = g_dev[dev];
    if (!ds.host_words) {
        void *p = nullptr;
        hipError_t e = hipHostMalloc(&p, 64, hipHostMallocCoherent);
        if (e != hipSuccess)
            return fail(GPUCAD_E_OOM, "pinned host allocation failed: %s", hipGetErrorString(e));
        ds.host_words = (uint32_t *)p;
    }
    *out = ds.host_words;
    return GPUCAD_OK;
}

int tables_if_ready(const DeviceTables **out) {
    int dev = 0;
    int st = current_device(&dev);
    if (st) return st;
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState &ds = g_dev[dev];
    if (!ds.tables_ready) return fail(GPUCAD_E_ARG, "gpucad_init was not called on device %d", dev);
    *out = &ds.tables;
    return GPUCAD_OK;
}

int host_words_if_ready(uint32_t **out) {
    int dev = 0;
    int st = current_device(&dev);
    if (st) return st;
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState &ds = g_dev[dev];
    if (!ds.host_words) return fail(GPUCAD_E_ARG, "gpucad_init was not called on device %d", dev);
    *out = ds.host_words;
    return GPUCAD_OK;
}

std::unique_lock<std::recursive_mutex> lock_device_calls() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) dev = 0;
    return std::unique_lock<std::recursive_mutex>(g_dev[dev].call_mu);
}

int scratch_arena(size_t bytes, void **out) {
    int dev = 0;
    int st = current_device(&dev);
    if (st) return st;
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceState &ds = g_dev[dev];
    if (ds.scratch_bytes < bytes) {
        if (ds.scratch) {
            // The previous arena may still be read by work queued on any stream.
            GC_HIP(hipDeviceSynchronize());
            GC_HIP(hipFree(ds.scratch));
            ds.scratch = nullptr;
            ds.scratch_bytes = 0;
        }
        size_t want = bytes + bytes / 4 + 4096;
        hipError_t e = hipMalloc(&ds.scratch, want);
        if (e != hipSuccess) {
            ds.scratch = nullptr;
            return fail(GPUCAD_E_OOM, "scratch allocation of %zu bytes failed: %s", want,
                        hipGetErrorString(e));
        }
        ds.scratch_bytes = want;
    }
    *out = ds.scratch;
    return GPUCAD_OK;
}

}  // namespace gc

extern "C" {

int gpucad_version(void) { return 10000; }

int gpucad_abi_sizes(uint32_t *sizes, int32_t n) {
    const uint32_t s[] = {(uint32_t)sizeof(gc_lattice_params), (uint32_t)sizeof(gc_extract_args),
                          (uint32_t)sizeof(gc_mesh_counts),    (uint32_t)sizeof(gc_uint3),
                          (uint32_t)sizeof(gc_float3),         (uint32_t)sizeof(gc_float4),
                          (uint32_t)sizeof(gc_float2),         (uint32_t)sizeof(gc_svl_params),
                          (uint32_t)sizeof(gc_svl_result)};
    int m = 0;
    for (; m < n && m < (int)(sizeof(s) / sizeof(s[0])); ++m) sizes[m] = s[m];
    return m;
}

const char *gpucad_status_string(int status) {
    switch (status) {
        case GPUCAD_OK: return "ok";
        case GPUCAD_E_ARG: return "invalid argument";
        case GPUCAD_E_OVERFLOW: return "output capacity exceeded";
        case GPUCAD_E_OOM: return "out of memory";
        case GPUCAD_E_HIP: return "HIP runtime error";
        case GPUCAD_E_BUSY: return "busy";
        case GPUCAD_E_NOTFOUND: return "not found";
        case GPUCAD_E_IO: return "i/o error";
        default: return "unknown status";
    }
}

const char *gpucad_last_error(void) { return gc::t_error; }

int gpucad_init(int device) {
    gc::clear_error();
    int n = 0;
    GC_HIP(hipGetDeviceCount(&n));
    GC_ARG(device >= 0 && device < n, "device %d not present (%d devices)", device, n);
    GC_HIP(hipSetDevice(device));
    const gc::DeviceTables *t = nullptr;
    int st = gc::device_tables(&t);
    if (st) return st;
    uint32_t *hw = nullptr;
    if ((st = gc::host_words(&hw))) return st;
    return gc::noalloc_setup();
}

int gpucad_shutdown(void) {
    gc::clear_error();
    int dev = 0;
    int st = gc::current_device(&dev);
    if (st) return st;
    auto calls = gc::lock_device_calls();
    std::lock_guard<std::mutex> lk(gc::g_mu);
    gc::DeviceState &ds = gc::g_dev[dev];
    GC_HIP(hipDeviceSynchronize());
    if (ds.tables_ready) {
        GC_HIP(hipFree(ds.tables.tri_packed));
        GC_HIP(hipFree(ds.tables.num_verts));
        GC_HIP(hipFree(ds.tables.tri_table));
        ds.tables = gc::DeviceTables();
        ds.tables_ready = false;
    }
    if (ds.scratch) {
        GC_HIP(hipFree(ds.scratch));
        ds.scratch = nullptr;
        ds.scratch_bytes = 0;
    }
    if (ds.host_words) {
        GC_HIP(hipHostFree(ds.host_words));
        ds.host_words = nullptr;
    }
    return GPUCAD_OK;
}

}  // extern "C"
