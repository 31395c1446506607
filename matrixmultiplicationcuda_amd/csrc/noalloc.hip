// noalloc.hip -- bcsr_spmm_noalloc: bcsr_spmm_cfg for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  No reference counterpart.
//
// The only product that needs memory beyond its inputs and outputs is the
// FP8 mode's quantised B (B8 and its per-slab scales), which bcsr_spmm keeps
// in a workspace the plan allocates.  Here it
// lives in `scratch`, a file-scope device array of kScratchBytes that every
// GPU holds (once this module is loaded) until the process ends.  Its device
// address is taken once per device with hipGetSymbolAddress, with that
// device current; host code uses only that address, never the array's name
// (which on the host is a stand-in, not a device address).
//
// Every call on one GPU shares that GPU's scratch: calls that overlap on one
// GPU (different streams, no ordering) race.  Nothing in it needs clearing:
// the quantiser writes every byte of B8 and every scale the product reads.
#include <array>
#include <mutex>

#include "bcsr_internal.h"

namespace {

constexpr size_t kScratchBytes = BCSR_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::array<void *, kMaxDevices> g_scratch_ptr{};

// The current device's copy of `scratch` (resolved on the first call on it).
int scratch_of(int dev, uint8_t **out) {
    if (dev < 0 || dev >= kMaxDevices) return BCSR_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const int rc = hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (rc != BCSR_OK) return rc;
        if (!p) return BCSR_ERR_HIP;
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<uint8_t *>(g_scratch_ptr[dev]);
    return BCSR_OK;
}

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

}  // namespace

extern "C" int bcsr_spmm_noalloc(const bcsr_plan *plan, const void *B, float *C, int32_t N,
                                 const bcsr_launch_cfg *cfg, void *stream) {
    int rc = check_operands(plan, B, C, N);
    if (rc != BCSR_OK) return rc;
    if (N == 0 || plan->info.rows == 0) return BCSR_OK;
    LaunchShape ls;
    rc = resolve_launch(plan, N, cfg, ls);
    if (rc != BCSR_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (plan->info.precision != BCSR_PREC_FP8_E4M3) {
        // the 16-bit kernels read only the plan, B and C; the panel kernel's
        // schedule is device memory its first launch allocates
        if (ls.variant == 3) {
            if (!plan->panel_mu) return BCSR_ERR_UNSUPPORTED;
            {
                std::lock_guard<std::mutex> lk(*plan->panel_mu);
                if (!plan->d_panel) return BCSR_ERR_UNSUPPORTED;
            }
            return launch_panel(plan, ls, B, C, N, s);
        }
        return launch_pairs(plan, ls, B, C, N, s);
    }
    const int K = plan->info.cols, slabs = K / 16;
    const size_t off_s = align256(static_cast<size_t>(K) * N);
    const size_t need = off_s + b8_scale_bytes(K, N);
    if (need > kScratchBytes) return BCSR_ERR_SCRATCH_LIMIT;
    uint8_t *base = nullptr;
    rc = scratch_of(plan->device, &base);
    if (rc != BCSR_OK) return rc;
    uint8_t *B8 = base, *bscale = base + off_s;
    if (slabs > 0) {
        rc = enqueue_quant(B, K, N, B8, bscale, s);
        if (rc != BCSR_OK) return rc;
    }
    return launch_fp8_product(plan, ls, B8, bscale, C, N, s);
}
