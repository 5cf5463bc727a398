// Graphics-API interop: device buffers and semaphores exported by another API
// (the reference's caller exports Vulkan VkDeviceMemory and VkSemaphore as
// opaque POSIX file descriptors) imported into HIP, so the extraction writes
// `pos/norm` in place and the two APIs order their work on the GPU.
//
// Reference: src/main.cu:1164-1191 importCudaExternalMemory (opaque fd,
// whole allocation mapped at offset 0), 1193-1211 importCudaExternalSemaphore
// (opaque fd), 2017-2035 and 3122-3134 signal/wait with fence value 0 (binary
// semaphores), 739-931 destroy.
#include "gc_common.hpp"

#include <new>

struct gc_external_buffer {
    hipExternalMemory_t mem = nullptr;
    void *ptr = nullptr;
};

struct gc_external_semaphore {
    hipExternalSemaphore_t sem = nullptr;
};

extern "C" {

int gpucad_import_external_buffer(int32_t fd, uint64_t size, uint64_t offset,
                                  gc_external_buffer **out, void **d_ptr) {
    gc::clear_error();
    GC_ARG(out && d_ptr, "out and d_ptr must not be NULL");
    *out = nullptr;
    *d_ptr = nullptr;
    GC_ARG(fd >= 0, "fd must be a file descriptor");
    GC_ARG(size > 0, "size must be > 0");
    GC_ARG(offset < size, "offset %llu outside the %llu-byte allocation",
           (unsigned long long)offset, (unsigned long long)size);
    hipExternalMemoryHandleDesc hd = {};
    hd.type = hipExternalMemoryHandleTypeOpaqueFd;
    hd.handle.fd = fd;
    hd.size = size;
    hd.flags = 0;
    gc_external_buffer *b = new (std::nothrow) gc_external_buffer;
    if (!b) return gc::fail(GPUCAD_E_OOM, "out of host memory");
    hipError_t e = hipImportExternalMemory(&b->mem, &hd);
    if (e != hipSuccess) {
        delete b;
        return gc::fail(GPUCAD_E_HIP, "hipImportExternalMemory(fd %d, %llu bytes) failed: %s",
                        (int)fd, (unsigned long long)size, hipGetErrorString(e));
    }
    hipExternalMemoryBufferDesc bd = {};
    bd.offset = offset;
    bd.size = size - offset;
    bd.flags = 0;
    e = hipExternalMemoryGetMappedBuffer(&b->ptr, b->mem, &bd);
    if (e != hipSuccess) {
        (void)hipDestroyExternalMemory(b->mem);
        delete b;
        return gc::fail(GPUCAD_E_HIP, "hipExternalMemoryGetMappedBuffer failed: %s",
                        hipGetErrorString(e));
    }
    *out = b;
    *d_ptr = b->ptr;
    return GPUCAD_OK;
}

int gpucad_release_external_buffer(gc_external_buffer *b) {
    gc::clear_error();
    if (!b) return GPUCAD_OK;
    // the mapped range is freed like a device allocation, then the import
    hipError_t e1 = b->ptr ? hipFree(b->ptr) : hipSuccess;
    hipError_t e2 = hipDestroyExternalMemory(b->mem);
    delete b;
    if (e1 != hipSuccess || e2 != hipSuccess)
        return gc::fail(GPUCAD_E_HIP, "releasing an external buffer failed: %s / %s",
                        hipGetErrorString(e1), hipGetErrorString(e2));
    return GPUCAD_OK;
}

int gpucad_import_external_semaphore(int32_t fd, gc_external_semaphore **out) {
    gc::clear_error();
    GC_ARG(out, "out must not be NULL");
    *out = nullptr;
    GC_ARG(fd >= 0, "fd must be a file descriptor");
    hipExternalSemaphoreHandleDesc hd = {};
    hd.type = hipExternalSemaphoreHandleTypeOpaqueFd;
    hd.handle.fd = fd;
    hd.flags = 0;
    gc_external_semaphore *s = new (std::nothrow) gc_external_semaphore;
    if (!s) return gc::fail(GPUCAD_E_OOM, "out of host memory");
    hipError_t e = hipImportExternalSemaphore(&s->sem, &hd);
    if (e != hipSuccess) {
        delete s;
        return gc::fail(GPUCAD_E_HIP, "hipImportExternalSemaphore(fd %d) failed: %s", (int)fd,
                        hipGetErrorString(e));
    }
    *out = s;
    return GPUCAD_OK;
}

int gpucad_signal_external_semaphore(gc_external_semaphore *s, void *stream) {
    gc::clear_error();
    GC_ARG(s && s->sem, "semaphore is NULL");
    hipExternalSemaphoreSignalParams p = {};
    p.params.fence.value = 0;  // binary semaphore
    p.flags = 0;
    GC_HIP(hipSignalExternalSemaphoresAsync(&s->sem, &p, 1, gc::as_stream(stream)));
    return GPUCAD_OK;
}

int gpucad_wait_external_semaphore(gc_external_semaphore *s, void *stream) {
    gc::clear_error();
    GC_ARG(s && s->sem, "semaphore is NULL");
    hipExternalSemaphoreWaitParams p = {};
    p.params.fence.value = 0;  // binary semaphore
    p.flags = 0;
    GC_HIP(hipWaitExternalSemaphoresAsync(&s->sem, &p, 1, gc::as_stream(stream)));
    return GPUCAD_OK;
}

int gpucad_release_external_semaphore(gc_external_semaphore *s) {
    gc::clear_error();
    if (!s) return GPUCAD_OK;
    hipError_t e = hipDestroyExternalSemaphore(s->sem);
    delete s;
    if (e != hipSuccess)
        return gc::fail(GPUCAD_E_HIP, "hipDestroyExternalSemaphore failed: %s",
                        hipGetErrorString(e));
    return GPUCAD_OK;
}

}  // extern "C"
