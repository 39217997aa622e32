// Runtime part of the C ABI: status/error reporting, device memory, streams,
// pinned host memory.  Device memory is plain hipMalloc'd HBM owned by the
// caller through the returned pointer; the library never keeps references to
// caller buffers beyond a call, except inside an RK engine object which
// documents what it holds.
#include <stdarg.h>
#include <string.h>

#include "qa_api_internal.h"

namespace qa {
static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what,
              file, line);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return QA_ERR_NOMEM;
    return QA_ERR_HIP;
}
}  // namespace qa

// Fill kernel used instead of hipMemsetAsync so that clears are safe to
// record into HIP graphs.
__global__ void qa_fill_zero_kernel(uint4 *p, size_t n16, unsigned char *tail, size_t ntail) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
    if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

extern "C" {

const char *qa_last_error(void) { return qa::g_err; }

int qa_version(void) { return QA_VERSION; }

int qa_device_count(int *n) {
    QA_REQUIRE(n != nullptr, "qa_device_count: null output");
    QA_HIP(hipGetDeviceCount(n));
    return QA_OK;
}

int qa_set_device(int dev) {
    QA_HIP(hipSetDevice(dev));
    return QA_OK;
}

int qa_device_info(int dev, char *name, int name_len, int *cu_count, size_t *hbm_bytes) {
    hipDeviceProp_t p;
    QA_HIP(hipGetDeviceProperties(&p, dev));
    if (name && name_len > 0) {
        strncpy(name, p.gcnArchName, (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
    return QA_OK;
}

int qa_mem_info(size_t *free_bytes, size_t *total_bytes) {
    QA_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return QA_OK;
}

int qa_malloc(void **ptr, size_t bytes) {
    QA_REQUIRE(ptr != nullptr, "qa_malloc: null output");
    *ptr = nullptr;
    if (bytes == 0) bytes = 256;
    QA_HIP(hipMalloc(ptr, bytes));
    return QA_OK;
}

int qa_free(void *ptr) {
    if (ptr) QA_HIP(hipFree(ptr));
    return QA_OK;
}

int qa_host_malloc(void **ptr, size_t bytes) {
    QA_REQUIRE(ptr != nullptr, "qa_host_malloc: null output");
    if (bytes == 0) bytes = 64;
    QA_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return QA_OK;
}

int qa_host_free(void *ptr) {
    if (ptr) QA_HIP(hipHostFree(ptr));
    return QA_OK;
}

int qa_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return QA_OK;
    QA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    QA_HIP(hipStreamSynchronize((hipStream_t)stream));
    return QA_OK;
}

int qa_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return QA_OK;
    QA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    QA_HIP(hipStreamSynchronize((hipStream_t)stream));
    return QA_OK;
}

int qa_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return QA_OK;
    QA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return QA_OK;
}

int qa_fill_zero(void *dst, size_t bytes, void *stream) {
    if (bytes == 0) return QA_OK;
    QA_REQUIRE(((uintptr_t)dst & 15) == 0, "qa_fill_zero: pointer must be 16-byte aligned");
    size_t n16 = bytes / 16, ntail = bytes % 16;
    unsigned blocks = (unsigned)((n16 + 255) / 256);
    if (blocks == 0) blocks = 1;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(qa_fill_zero_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       (uint4 *)dst, n16, (unsigned char *)dst + n16 * 16, ntail);
    QA_CHECK_LAUNCH();
    return QA_OK;
}

int qa_stream_create(void **stream) {
    QA_REQUIRE(stream != nullptr, "qa_stream_create: null output");
    hipStream_t s;
    QA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return QA_OK;
}

int qa_stream_destroy(void *stream) {
    if (stream) QA_HIP(hipStreamDestroy((hipStream_t)stream));
    return QA_OK;
}

int qa_stream_sync(void *stream) {
    QA_HIP(hipStreamSynchronize((hipStream_t)stream));
    return QA_OK;
}

int qa_event_create(void **event) {
    QA_REQUIRE(event != nullptr, "qa_event_create: null output");
    hipEvent_t e;
    QA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    *event = (void *)e;
    return QA_OK;
}

int qa_event_destroy(void *event) {
    if (event) QA_HIP(hipEventDestroy((hipEvent_t)event));
    return QA_OK;
}

int qa_event_record(void *event, void *stream) {
    QA_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return QA_OK;
}

// *done = 1 when every operation before the record has finished (non-blocking).
int qa_event_query(void *event, int *done) {
    QA_REQUIRE(done != nullptr, "qa_event_query: null output");
    const hipError_t e = hipEventQuery((hipEvent_t)event);
    if (e == hipErrorNotReady) {
        *done = 0;
        return QA_OK;
    }
    QA_HIP(e);
    *done = 1;
    return QA_OK;
}

int qa_device_sync(void) {
    QA_HIP(hipDeviceSynchronize());
    return QA_OK;
}

}  // extern "C"
