// Drives the interop entry points the way the reference's caller does with
// Vulkan memory: an allocation exported as an opaque POSIX fd is imported,
// mapped, and the extraction writes its vertices into it in place.  There is
// no Vulkan device here, so the exporter is HIP's own virtual-memory API
// (hipMemCreate with a POSIX-fd handle type, hipMemExportToShareableHandle);
// the allocation is also mapped directly, which shows that the imported
// pointer aliases it.
//
// Prints "PASS ...", "SKIP <reason>" (exit 0) or "FAIL <reason>" (exit 1).
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "gpucad.h"

#define HIPOK(x)                                                              \
    do {                                                                      \
        hipError_t e_ = (x);                                                  \
        if (e_ != hipSuccess) {                                               \
            std::printf("FAIL %s: %s\n", #x, hipGetErrorString(e_));          \
            return 1;                                                         \
        }                                                                     \
    } while (0)

int main() {
    int vmm = 0;
    HIPOK(hipDeviceGetAttribute(&vmm, hipDeviceAttributeVirtualMemoryManagementSupported, 0));
    if (!vmm) {
        std::printf("SKIP the device has no virtual-memory management (no fd exporter)\n");
        return 0;
    }
    // a sphere field on a 64^3 grid: f = |p - c| - r
    const int n = 64;
    std::vector<float> host((size_t)n * n * n);
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                const float dx = x - 31.5f, dy = y - 30.25f, dz = z - 32.75f;
                host[((size_t)z * n + y) * n + x] = std::sqrt(dx * dx + dy * dy + dz * dz) - 21.0f;
            }
    float *field = nullptr;
    HIPOK(hipMalloc(&field, host.size() * 4));
    HIPOK(hipMemcpy(field, host.data(), host.size() * 4, hipMemcpyHostToDevice));
    const uint64_t vcap = 40000, tcap = 80000;
    const uint64_t wsb = gpucad_extract_workspace_bytes(n, n, n);
    void *ws = nullptr;
    float *vref = nullptr;
    uint32_t *tri = nullptr;
    HIPOK(hipMalloc(&ws, wsb));
    HIPOK(hipMalloc(&vref, vcap * 12));
    HIPOK(hipMalloc(&tri, tcap * 12));

    // ---- the exporter: a VMM allocation with a POSIX-fd handle type
    hipMemAllocationProp prop = {};
    prop.type = hipMemAllocationTypePinned;
    prop.location.type = hipMemLocationTypeDevice;
    prop.location.id = 0;
    prop.requestedHandleType = hipMemHandleTypePosixFileDescriptor;
    size_t gran = 0;
    HIPOK(hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityMinimum));
    const size_t bytes = ((vcap * 12 + gran - 1) / gran) * gran;
    hipMemGenericAllocationHandle_t h;
    HIPOK(hipMemCreate(&h, bytes, &prop, 0));
    int fd = -1;
    HIPOK(hipMemExportToShareableHandle(&fd, h, hipMemHandleTypePosixFileDescriptor, 0));
    void *va = nullptr;
    HIPOK(hipMemAddressReserve(&va, bytes, 0, nullptr, 0));
    HIPOK(hipMemMap(va, bytes, 0, h, 0));
    hipMemAccessDesc acc = {};
    acc.location.type = hipMemLocationTypeDevice;
    acc.location.id = 0;
    acc.flags = hipMemAccessFlagsProtReadWrite;
    HIPOK(hipMemSetAccess(va, bytes, &acc, 1));

    // ---- the import (the library owns fd on success)
    gc_external_buffer *buf = nullptr;
    void *ptr = nullptr;
    int st = gpucad_import_external_buffer(fd, bytes, 0, &buf, &ptr);
    if (st != GPUCAD_OK) {
        std::printf("SKIP import of a HIP-exported fd not supported here: %s\n",
                    gpucad_last_error());
        close(fd);
        return 0;
    }
    // aliasing: bytes written through the import are read through the VMM map
    std::vector<uint32_t> pat(1024), back(1024, 0);
    for (int i = 0; i < 1024; ++i) pat[i] = 0x9e3779b9u * (uint32_t)(i + 1);
    HIPOK(hipMemcpy(ptr, pat.data(), 4096, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(back.data(), va, 4096, hipMemcpyDeviceToHost));
    if (back != pat) {
        std::printf("FAIL the imported buffer does not alias the exported allocation\n");
        return 1;
    }

    // ---- extraction into the imported buffer vs into device memory
    gc_extract_args a;
    std::memset(&a, 0, sizeof(a));
    a.field = field;
    a.nx = a.ny = a.nz = n;
    a.iso = 0.0f;
    a.origin = gc_float3{-1.0f, -1.0f, -1.0f};
    a.spacing = gc_float3{2.0f / (n - 1), 2.0f / (n - 1), 2.0f / (n - 1)};
    a.triangles = tri;
    a.tri_capacity = tcap;
    a.workspace = ws;
    a.workspace_bytes = wsb;
    gc_mesh_counts c1 = {0, 0}, c2 = {0, 0};
    a.vertices = vref;
    a.vert_capacity = vcap;
    if ((st = gpucad_extract(&a, &c1, nullptr)) != GPUCAD_OK) {
        std::printf("FAIL extract (device memory): %s\n", gpucad_last_error());
        return 1;
    }
    a.vertices = (float *)ptr;
    a.vert_capacity = bytes / 12;
    if ((st = gpucad_extract(&a, &c2, nullptr)) != GPUCAD_OK) {
        std::printf("FAIL extract (imported buffer): %s\n", gpucad_last_error());
        return 1;
    }
    if (c1.n_vertices != c2.n_vertices || c1.n_triangles != c2.n_triangles ||
        c1.n_vertices == 0) {
        std::printf("FAIL counts differ: %llu/%llu vs %llu/%llu\n",
                    (unsigned long long)c1.n_vertices, (unsigned long long)c1.n_triangles,
                    (unsigned long long)c2.n_vertices, (unsigned long long)c2.n_triangles);
        return 1;
    }
    std::vector<float> r1(c1.n_vertices * 3), r2(c1.n_vertices * 3);
    HIPOK(hipMemcpy(r1.data(), vref, r1.size() * 4, hipMemcpyDeviceToHost));
    HIPOK(hipMemcpy(r2.data(), va, r2.size() * 4, hipMemcpyDeviceToHost));  // through the map
    if (std::memcmp(r1.data(), r2.data(), r1.size() * 4) != 0) {
        std::printf("FAIL vertices written through the import differ\n");
        return 1;
    }
    if ((st = gpucad_release_external_buffer(buf)) != GPUCAD_OK) {
        std::printf("FAIL release: %s\n", gpucad_last_error());
        return 1;
    }
    HIPOK(hipMemUnmap(va, bytes));
    HIPOK(hipMemAddressFree(va, bytes));
    HIPOK(hipMemRelease(h));
    HIPOK(hipFree(field));
    HIPOK(hipFree(ws));
    HIPOK(hipFree(vref));
    HIPOK(hipFree(tri));
    std::printf("PASS %llu vertices, %llu triangles written through an imported fd\n",
                (unsigned long long)c1.n_vertices, (unsigned long long)c1.n_triangles);
    return 0;
}
