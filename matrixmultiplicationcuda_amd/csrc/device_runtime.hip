// device_runtime.hip -- thin wrappers over the HIP runtime this library is
// linked against (allocation, copies, streams, events) and a plain streaming
// copy kernel, so callers without their own HIP bindings share one runtime.
#include <algorithm>

#include "bcsr_internal.h"

namespace {

// Plain streaming copy (16 bytes per lane per step, grid-stride).  Used as a
// known-size reference for hardware byte counters and as a device copy.
__global__ __launch_bounds__(256) void bcsr_stream_copy(u32x4 *__restrict__ dst, const u32x4 *__restrict__ src,
                                                        size_t n16) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}

}  // namespace

// ---------------------------------------------------------------------------
// Device plumbing
// ---------------------------------------------------------------------------
extern "C" {

int bcsr_device_count(int32_t *count) {
    if (!count) return BCSR_ERR_INVALID_ARG;
    int n = 0;
    int rc = hip_status(hipGetDeviceCount(&n));
    *count = rc == BCSR_OK ? n : 0;
    return rc;
}

int bcsr_set_device(int32_t device) { return hip_status(hipSetDevice(device)); }

int bcsr_dev_malloc(void **ptr, size_t bytes) {
    if (!ptr) return BCSR_ERR_INVALID_ARG;
    *ptr = nullptr;
    return hip_status(hipMalloc(ptr, bytes ? bytes : 1));
}

int bcsr_dev_free(void *ptr) { return hip_status(hipFree(ptr)); }

int bcsr_memcpy_htod(void *dst, const void *src, size_t bytes) {
    if (!bytes) return BCSR_OK;
    if (!dst || !src) return BCSR_ERR_INVALID_ARG;
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

int bcsr_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    if (!bytes) return BCSR_OK;
    if (!dst || !src) return BCSR_ERR_INVALID_ARG;
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

int bcsr_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return BCSR_OK;
    if (!dst || !src) return BCSR_ERR_INVALID_ARG;
    return hip_status(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(stream)));
}

int bcsr_memset_async(void *dst, int32_t byte_value, size_t bytes, void *stream) {
    if (!bytes) return BCSR_OK;
    if (!dst) return BCSR_ERR_INVALID_ARG;
    return hip_status(hipMemsetAsync(dst, byte_value, bytes, static_cast<hipStream_t>(stream)));
}

int bcsr_copy_kernel(void *dst, const void *src, size_t bytes, void *stream) {
    if (!bytes) return BCSR_OK;
    if (!dst || !src || (bytes & 15u) || ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15u))
        return BCSR_ERR_INVALID_ARG;
    const size_t n16 = bytes / 16;
    const size_t blocks = std::min<size_t>((n16 + 255) / 256, 8192);
    hipLaunchKernelGGL(bcsr_stream_copy, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<u32x4 *>(dst), static_cast<const u32x4 *>(src),
                       n16);
    return hip_status(hipGetLastError());
}

int bcsr_stream_create(void **stream) {
    if (!stream) return BCSR_ERR_INVALID_ARG;
    hipStream_t s = nullptr;
    int rc = hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return rc;
}

int bcsr_stream_destroy(void *stream) { return hip_status(hipStreamDestroy(static_cast<hipStream_t>(stream))); }

int bcsr_stream_synchronize(void *stream) {
    return hip_status(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
}

int bcsr_device_synchronize(void) { return hip_status(hipDeviceSynchronize()); }

int bcsr_event_create(void **event) {
    if (!event) return BCSR_ERR_INVALID_ARG;
    hipEvent_t e = nullptr;
    int rc = hip_status(hipEventCreate(&e));
    *event = e;
    return rc;
}

int bcsr_event_destroy(void *event) { return hip_status(hipEventDestroy(static_cast<hipEvent_t>(event))); }

int bcsr_event_record(void *event, void *stream) {
    return hip_status(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
}

int bcsr_event_elapsed_ms(void *start, void *end, float *ms) {
    if (!ms) return BCSR_ERR_INVALID_ARG;
    return hip_status(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(end)));
}

// Stream capture into a HIP graph: the launches enqueued on `stream` between
// begin and end are recorded, not run; the instantiated graph replays them.
// Thread-local capture mode, so other threads' HIP calls are unaffected.
int bcsr_graph_begin(void *stream) {
    if (!stream) return BCSR_ERR_INVALID_ARG;  // the null stream cannot be captured
    return hip_status(hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal));
}

int bcsr_graph_end(void *stream, void **graph_exec) {
    if (!stream || !graph_exec) return BCSR_ERR_INVALID_ARG;
    *graph_exec = nullptr;
    hipGraph_t g = nullptr;
    int rc = hip_status(hipStreamEndCapture(static_cast<hipStream_t>(stream), &g));
    if (rc != BCSR_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    hipGraphExec_t e = nullptr;
    rc = hip_status(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
    const int rc_destroy = hip_status(hipGraphDestroy(g));  // the executable graph keeps its own copy
    if (rc != BCSR_OK) return rc;
    if (rc_destroy != BCSR_OK) {
        (void)hipGraphExecDestroy(e);
        return rc_destroy;
    }
    *graph_exec = e;
    return BCSR_OK;
}

int bcsr_graph_launch(void *graph_exec, void *stream) {
    if (!graph_exec) return BCSR_ERR_INVALID_ARG;
    return hip_status(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), static_cast<hipStream_t>(stream)));
}

int bcsr_graph_destroy(void *graph_exec) {
    if (!graph_exec) return BCSR_OK;
    return hip_status(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)));
}

}  // extern "C"

