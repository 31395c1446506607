// bw_probe.hip -- read bandwidth of the CUs' vector-memory path from a buffer
// of a chosen size: from L1 or L2 when it fits there, from the Infinity Cache
// (256 MB) or HBM beyond.  Every thread of a full grid (2048 workgroups of 256
// threads) streams 16-byte loads over the buffer; the rate is bytes loaded
// over the time of one launch (HIP events, after a warm-up launch).  It gives
// the ceiling of the per-row BCSR kernels, which deliver every block's slab
// of B to the CU that multiplies it.
//
// usage: bcsr_bw_probe [KiB ...]   (default 16 2048 65536 1048576)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void probe(const u32x4 *buf, unsigned n16, int iters, unsigned *sink) {
    const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned step = (gridDim.x * blockDim.x) % n16;
    unsigned idx = tid % n16;
    u32x4 acc = {0u, 0u, 0u, 0u};
#pragma unroll 8
    for (int i = 0; i < iters; ++i) {
        acc ^= buf[idx];
        idx += step;
        if (idx >= n16) idx -= n16;
    }
    // never true for the buffer's contents; keeps the loads
    if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9e3779b9u) sink[0] = tid;
}

bool ok(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return false;
}

}  // namespace

int main(int argc, char **argv) {
    std::vector<size_t> kib;
    for (int i = 1; i < argc; ++i) kib.push_back(std::strtoull(argv[i], nullptr, 10));
    if (kib.empty()) kib = {16, 2048, 65536, 1048576};
    const int blocks = 2048, threads = 256, iters = 2000;
    unsigned *sink = nullptr;
    hipEvent_t e0, e1;
    if (!ok(hipMalloc(&sink, 4), "hipMalloc") || !ok(hipEventCreate(&e0), "event") || !ok(hipEventCreate(&e1), "event"))
        return 1;
    for (size_t k : kib) {
        if (k < 16) k = 16;
        const size_t bytes = k * 1024;
        u32x4 *buf = nullptr;
        if (!ok(hipMalloc(&buf, bytes), "hipMalloc") || !ok(hipMemset(buf, 1, bytes), "hipMemset")) return 1;
        const unsigned n16 = static_cast<unsigned>(bytes / 16);
        hipLaunchKernelGGL(probe, dim3(blocks), dim3(threads), 0, 0, buf, n16, 10, sink);
        if (!ok(hipDeviceSynchronize(), "warm-up")) return 1;
        (void)hipEventRecord(e0);
        hipLaunchKernelGGL(probe, dim3(blocks), dim3(threads), 0, 0, buf, n16, iters, sink);
        (void)hipEventRecord(e1);
        if (!ok(hipEventSynchronize(e1), "launch")) return 1;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double moved = double(blocks) * threads * iters * 16;
        std::printf("buffer %zu KiB: %.2f GB loaded in %.3f ms = %.1f TB/s\n", k, moved / 1e9, ms,
                    moved / (ms * 1e-3) / 1e12);
        (void)hipFree(buf);
    }
    (void)hipFree(sink);
    return 0;
}
