// kernels_batch.hip -- many independent BCSR x dense products in one
// persistent launch (no reference counterpart: the reference launches one
// kernel per product).
//
// A batch binds count products C_i = A_i . B_i (plans of one precision, each
// with its own block shape, N and output tile width).  Their work is cut into
// items (problem, column tile, run of device block-rows taken from the
// problem's longest-first row order), sized by a work target and queued
// problem by problem (largest first), tile by tile.  A persistent grid of
// waves pulls items from a
// device counter (one atomic per item) until the queue is empty, and runs each
// with the per-matrix kernels' own work functions, so the batched results are
// bit-identical to the per-matrix launches.  A run is two kernels: a one-thread
// reset of the counter and the persistent kernel (plus, for FP8, a batched
// B-quantisation kernel first); the sequence can be captured once into a
// hipGraph and replayed.  Memory is never cleared with hipMemsetAsync, so the
// captured graph holds kernel nodes only.
#include <algorithm>
#include <cstdio>
#include <new>
#include <numeric>

#include "fp8_device.h"
#include "pair_device.h"

namespace {

enum { BPREC_F16 = 0, BPREC_BF16 = 1, BPREC_FP8 = 2, BPREC_HILO = 3 };

struct BatchMat {
    SpmmArgs a;
    int key;  // rs * 16 + epl
};

struct QuantJob {
    const _Float16 *B;
    uint8_t *B8;
    uint8_t *scale;
    int N;
    int chunks;
    int slabs;
    int wg0;  // first workgroup index of this job (prefix over slab pairs * chunks)
};

// Block shapes the persistent kernel is compiled for.  Its register
// allocation is the maximum over its branches, so shapes whose work function
// would push a form past 256 VGPRs (one wave per SIMD) are left to the
// per-matrix kernels: 256-column tiles, and for FP8 and hi+lo the 64-row,
// 128-column tile.  fp16 and bf16 keep that tile through the plain work
// function (208 VGPRs; the 96-problem workload takes 0.541 against 0.555 ms
// for fp16 and 0.542 against 0.558 ms for bf16 with 64-column tiles).
__host__ __device__ constexpr bool batch_shape(int prec, int rs, int epl) {
    const int h = prec == BPREC_HILO ? 2 : 1;
    if (rs == 4 && epl == 8) return prec == BPREC_F16 || prec == BPREC_BF16;
    return (rs == 1 || rs == 2 || rs == 4) && (epl == 2 || epl == 4 || epl == 8) && rs * epl * h <= 16;
}

// The prefetching work function where its two load sets fit, else the plain one.
__host__ __device__ constexpr bool batch_prefetch(int prec, int rs, int epl) {
    return rs * epl * (prec == BPREC_HILO ? 2 : 1) <= 16;
}

template <int PREC, int RS, int EPL>
__device__ __forceinline__ void run_rows(const SpmmArgs &a, int tile, int s0, int s1) {
    if constexpr (batch_shape(PREC, RS, EPL)) {
        for (int s = s0; s < s1; ++s) {
            const int br = const_load(a.row_order, s);  // a scalar load (a's pointers come from memory)
            if constexpr (PREC == BPREC_FP8)
                fp8dev::fp8_work<RS, EPL, false>(a, br, tile);
            else
                pairdev::pair_work<RS, EPL, batch_prefetch(PREC, RS, EPL), PREC>(a, br, tile);
        }
    }
}

template <int PREC>
__device__ __forceinline__ void run_item(const SpmmArgs &a, int key, int tile, int s0, int s1) {
    switch (key) {
        case 0x12: run_rows<PREC, 1, 2>(a, tile, s0, s1); break;
        case 0x14: run_rows<PREC, 1, 4>(a, tile, s0, s1); break;
        case 0x18: run_rows<PREC, 1, 8>(a, tile, s0, s1); break;
        case 0x22: run_rows<PREC, 2, 2>(a, tile, s0, s1); break;
        case 0x24: run_rows<PREC, 2, 4>(a, tile, s0, s1); break;
        case 0x28: run_rows<PREC, 2, 8>(a, tile, s0, s1); break;
        case 0x42: run_rows<PREC, 4, 2>(a, tile, s0, s1); break;
        case 0x44: run_rows<PREC, 4, 4>(a, tile, s0, s1); break;
        case 0x48: run_rows<PREC, 4, 8>(a, tile, s0, s1); break;
        default: break;
    }
}

// The queue counters (kMaxQueues) back to zero before a run.
constexpr int kMaxQueues = 8;
__global__ void batch_reset(int *counter) {
    if (threadIdx.x < kMaxQueues) counter[threadIdx.x] = 0;
}

// One wave per workgroup; every wave leaves once the counter passes n_items.
// The FP8 form is held to 256 VGPRs, so two of its waves share each SIMD:
// unbounded it takes 268 (one wave per SIMD), bounded 254 with no spills, and
// it runs the 96-problem workload in 0.477 against 0.497 ms on MI355X.  The
// 16-bit forms already fit two waves (174-208 VGPRs); holding fp16 to three
// (168, before the 64 x 128 shape was added) spilled and was slower (0.564
// against 0.554 ms).
template <int PREC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PREC == BPREC_FP8 ? 2 : 1))) void
bcsr_batch_persistent(const BatchMat *__restrict__ mats, const int4 *__restrict__ items,
                      const int *__restrict__ qoff, int nq, int *counter) {
    // queue q holds items qoff[q] .. qoff[q + 1] - 1.  With eight queues a
    // wave starts at its XCD's (workgroups are dispatched round-robin over
    // the 8 XCDs) and, once that is empty, takes the others' in turn; it
    // leaves after finding every queue empty once
    int q = nq == 1 ? 0 : static_cast<int>(blockIdx.x & 7);
    for (int left = nq; left > 0;) {
        int it = 0;
        if (threadIdx.x == 0) it = atomicAdd(&counter[q], 1);
        it = __builtin_amdgcn_readfirstlane(__shfl(it, 0, 64));
        const int q0 = qoff[q], q1 = qoff[q + 1];
        if (it >= q1 - q0) {
            --left;
            q = q + 1 == nq ? 0 : q + 1;
            continue;
        }
        const int4 w = items[q0 + it];
        const BatchMat &m = mats[w.x];
        const SpmmArgs a = m.a;
        run_item<PREC>(a, m.key, w.y, w.z, w.w);
    }
}

// FP8: every problem's B quantised by one launch (each slab with its own
// scale), workgroups striding over the (job, slab pair, chunk) list;
// job_of[w] names workgroup w's job (one load instead of a search over the
// jobs' prefix sums).
__global__ __launch_bounds__(128) void batch_quant(const QuantJob *__restrict__ jobs, const int *__restrict__ job_of,
                                                   int total_wg) {
    for (int w = blockIdx.x; w < total_wg; w += gridDim.x) {
        const QuantJob &J = jobs[job_of[w]];
        const int local = w - J.wg0;
        fp8dev::quant_slab(J.B, J.B8, J.scale, J.N, J.slabs, local / J.chunks, local % J.chunks);
    }
}

typedef void (*PersistentKernel)(const BatchMat *, const int4 *, const int *, int, int *);

PersistentKernel pick_persistent(int prec) {
    switch (prec) {
        case BPREC_F16: return bcsr_batch_persistent<BPREC_F16>;
        case BPREC_BF16: return bcsr_batch_persistent<BPREC_BF16>;
        case BPREC_FP8: return bcsr_batch_persistent<BPREC_FP8>;
        case BPREC_HILO: return bcsr_batch_persistent<BPREC_HILO>;
        default: return nullptr;
    }
}

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

}  // namespace

struct bcsr_batch {
    int device = 0;
    int precision = 0;
    int count = 0;
    int n_items = 0;
    int n_queues = 1;
    int grid = 0;
    int n_jobs = 0;
    int quant_wg = 0;
    int quant_grid = 0;
    double flops = 0.0;
    size_t bytes = 0;
    void *d_base = nullptr;
    BatchMat *d_mats = nullptr;
    int4 *d_items = nullptr;
    int *d_qoff = nullptr;
    int *d_counter = nullptr;
    QuantJob *d_jobs = nullptr;
    int *d_job_of = nullptr;
    std::vector<int32_t> tiles;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int batch_enqueue(const bcsr_batch *b, hipStream_t stream) {
    if (b->n_jobs) {
        hipLaunchKernelGGL(batch_quant, dim3(b->quant_grid), dim3(128), 0, stream, b->d_jobs, b->d_job_of,
                           b->quant_wg);
        int rc = hip_status(hipGetLastError());
        if (rc != BCSR_OK) return rc;
    }
    hipLaunchKernelGGL(batch_reset, dim3(1), dim3(64), 0, stream, b->d_counter);
    int rc = hip_status(hipGetLastError());
    if (rc != BCSR_OK) return rc;
    if (!b->n_items) return BCSR_OK;
    PersistentKernel k = pick_persistent(b->precision);
    hipLaunchKernelGGL(k, dim3(b->grid), dim3(64), 0, stream, b->d_mats, b->d_items, b->d_qoff, b->n_queues,
                       b->d_counter);
    return hip_status(hipGetLastError());
}

int auto_batch_tile(const bcsr_plan *p, int N) {
    int tile = N >= 128 ? 128 : (N >= 64 ? 64 : 32);
    if (p->rs == 4 && N < 512) tile = std::min(tile, 64);  // as the per-matrix rule
    while (tile > 32 && !batch_shape(p->info.precision, p->rs, tile / 16)) tile /= 2;
    return tile;
}

}  // namespace

extern "C" int bcsr_batch_tile_supported(int32_t precision, int32_t dev_bh, int32_t tile_n) {
    if (dev_bh % 16 || tile_n % 16 || tile_n <= 0) return 0;
    return batch_shape(precision, dev_bh / 16, tile_n / 16) ? 1 : 0;
}

static int batch_create(const bcsr_plan *const *plans, const void *const *B, float *const *C, const int32_t *N,
                        const int32_t *tile_n, int32_t count, const bcsr_batch_cfg *cfg, bcsr_batch **out) {
    if (!out || count < 0 || (count > 0 && (!plans || !B || !C || !N || !plans[0]))) return BCSR_ERR_INVALID_ARG;
    *out = nullptr;
    int dev = 0;
    int rc = hip_status(hipGetDevice(&dev));
    if (rc != BCSR_OK) return rc;
    int n_cu = 0;
    rc = hip_status(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    if (rc != BCSR_OK) return rc;
    const int precision = count ? plans[0]->info.precision : BCSR_PREC_F16;
    const bool fp8 = precision == BCSR_PREC_FP8_E4M3;
    // defaults, fastest over the 96-problem batched workload on MI355X (graph
    // replays): FP8 8 waves per CU and items of up to 64 units (0.434 ms;
    // 6 and 64: 0.433); the 16-bit modes 6 and up to 128 (fp16 0.489 against
    // 0.526 ms at 8 and 64, bf16 0.492 against 0.521)
    const int wpc = cfg && cfg->waves_per_cu > 0 ? cfg->waves_per_cu : (fp8 ? 8 : 6);
    if (wpc > 32) return BCSR_ERR_INVALID_ARG;
    const int qmode = cfg ? cfg->queues : 0;
    if (qmode < 0 || qmode > 2) return BCSR_ERR_INVALID_ARG;

    std::vector<BatchMat> mats(count);
    std::vector<int32_t> tiles(count);
    std::vector<QuantJob> jobs;
    std::vector<size_t> ws_off(count, 0);
    size_t ws_total = 0;
    int quant_wg = 0;
    double flops = 0.0;
    for (int i = 0; i < count; ++i) {
        const bcsr_plan *p = plans[i];
        if (!p) return BCSR_ERR_INVALID_ARG;
        if (p->info.precision != precision) return BCSR_ERR_UNSUPPORTED;
        if (p->device != dev) return BCSR_ERR_INVALID_ARG;
        const int n = N[i];
        if (n < 0) return BCSR_ERR_INVALID_ARG;
        if (n % 16) return BCSR_ERR_SHAPE;
        if (n && p->info.rows && !C[i]) return BCSR_ERR_INVALID_ARG;
        if (n && p->info.cols && !B[i]) return BCSR_ERR_INVALID_ARG;
        if ((reinterpret_cast<uintptr_t>(B[i]) & 15u) || (reinterpret_cast<uintptr_t>(C[i]) & 15u))
            return BCSR_ERR_SHAPE;
        const int64_t kn = static_cast<int64_t>(p->info.cols) * n;
        if (kn * 2 > INT32_MAX) return BCSR_ERR_UNSUPPORTED;  // 32-bit buffer offsets
        if (fp8 && p->info.dev_pairs * p->rs * 2048 > INT32_MAX) return BCSR_ERR_UNSUPPORTED;
        int t = tile_n && tile_n[i] > 0 ? tile_n[i] : auto_batch_tile(p, n);
        if (t % 16 || !batch_shape(precision, p->rs, t / 16)) return BCSR_ERR_UNSUPPORTED;
        tiles[i] = t;
        flops += 2.0 * static_cast<double>(p->info.nnzb) * p->info.bh * p->info.bw * n;
        SpmmArgs &a = mats[i].a;
        a = SpmmArgs{};
        a.grp_ptr = p->d_grp_ptr;
        a.grp_col = p->d_grp_col;
        a.vals = static_cast<const u32x4 *>(p->d_vals);
        a.ascale = p->d_ascale;
        a.row_order = p->d_row_order;
        a.B = static_cast<const _Float16 *>(B[i]);
        a.C = C[i];
        a.N = n;
        a.n_tiles = n ? (n + t - 1) / t : 0;
        a.dev_rows = p->dev_rows;
        a.b_bytes = static_cast<int>(kn * 2);
        a.wpb = 1;
        mats[i].key = p->rs * 16 + t / 16;
        if (fp8 && n && p->info.cols) {
            const int slabs = p->info.cols / 16, chunks = (n + 255) / 256;
            ws_off[i] = ws_total;
            ws_total += align256(static_cast<size_t>(kn)) + align256(b8_scale_bytes(p->info.cols, n));
            a.b_bytes = static_cast<int>(kn);
            a.b_chunks = chunks;
            a.b_slabs = slabs;
            a.a_bytes = static_cast<int>(p->info.dev_pairs * p->rs * 2048);
            QuantJob J{};
            J.B = static_cast<const _Float16 *>(B[i]);
            J.N = n;
            J.chunks = chunks;
            J.slabs = slabs;
            J.wg0 = quant_wg;
            quant_wg += (slabs + 1) / 2 * chunks;
            jobs.push_back(J);
        }
    }

    // work items: per problem and column tile, runs of rows (longest first)
    // holding about `bin` units of work; a unit is one pair (or quad) step
    // of a 16-column MFMA tile, and every row costs one more for its store
    struct Item {
        int4 w;
        double work;
    };
    std::vector<Item> items;
    double total_work = 0.0;
    for (int i = 0; i < count; ++i) {
        const bcsr_plan *p = plans[i];
        const int nt = mats[i].a.n_tiles;
        if (!nt) continue;
        for (int r = 0; r < p->dev_rows; ++r)
            total_work += (p->h_grp_ptr[r + 1] - p->h_grp_ptr[r] + 1.0) * p->rs * (tiles[i] / 16) * nt;
    }
    const int grid = n_cu * wpc;
    // default item size: the cap above, fewer units only when the batch is
    // small (at least four items per wave)
    const double bin_cap = fp8 ? 64.0 : 128.0;
    double bin = cfg && cfg->bin_work > 0 ? cfg->bin_work : std::clamp(total_work / (4.0 * grid), 8.0, bin_cap);
    for (int i = 0; i < count; ++i) {
        const bcsr_plan *p = plans[i];
        const int nt = mats[i].a.n_tiles;
        if (!nt || !p->dev_rows) continue;
        const double unit = p->rs * (tiles[i] / 16);
        for (int t = 0; t < nt; ++t) {
            int s0 = 0;
            double acc = 0.0;
            for (int s = 0; s < p->dev_rows; ++s) {
                const int r = p->h_row_order[s];
                acc += (p->h_grp_ptr[r + 1] - p->h_grp_ptr[r] + 1.0) * unit;
                if (acc >= bin || s + 1 == p->dev_rows) {
                    items.push_back({int4{i, t, s0, s + 1}, acc});
                    s0 = s + 1;
                    acc = 0.0;
                }
            }
        }
    }
    // Queue order: problem by problem, largest total work first, and within
    // a problem column tile by column tile (each tile's runs of rows in
    // longest-first order).  At any time the waves then share one problem's
    // B slices in L2.  Measured on MI355X over the 96-problem workload:
    // 0.552 ms (fp16) and 0.496 ms (FP8) against 0.687 and 0.511 ms for the
    // items sorted by their own work alone; rows before tiles is 6 % slower.
    std::vector<double> pwork(count, 0.0);
    for (const Item &it : items) pwork[it.w.x] += it.work;
    std::stable_sort(items.begin(), items.end(), [&](const Item &x, const Item &y) {
        if (x.w.x != y.w.x) return pwork[x.w.x] > pwork[y.w.x] || (pwork[x.w.x] == pwork[y.w.x] && x.w.x < y.w.x);
        if (x.w.y != y.w.y) return x.w.y < y.w.y;
        return x.w.z < y.w.z;
    });
    if (items.size() > static_cast<size_t>(INT32_MAX)) return BCSR_ERR_UNSUPPORTED;

    // Queues.  One global queue, or one per XCD: a problem's column tiles
    // pinned to XCDs as in the per-matrix kernels (tile t of nt on the XCDs
    // x with x % nt == t, its runs of rows dealt over them in turn), so each
    // tile's slice of B is fetched into 8 / nt L2s instead of all eight;
    // tile counts that do not divide 8 deal their items over the XCDs in
    // turn.  Every queue keeps the global order.
    const int nq = qmode == 2 ? kMaxQueues : 1;
    std::vector<std::vector<int4>> qitems(nq);
    {
        int prev_i = -1, prev_t = -1, run = 0, rr = 0;
        for (const Item &it : items) {
            const int i = it.w.x, t = it.w.y, nt = mats[i].a.n_tiles;
            if (i != prev_i || t != prev_t) run = 0;
            prev_i = i;
            prev_t = t;
            int x = 0;
            if (nq > 1) {
                if (nt <= kMaxQueues && kMaxQueues % nt == 0)
                    x = t + nt * (run % (kMaxQueues / nt));
                else if (nt % kMaxQueues == 0)
                    x = t % kMaxQueues;
                else
                    x = rr++ % kMaxQueues;
            }
            ++run;
            qitems[x].push_back(it.w);
        }
    }
    std::vector<int> qoff(nq + 1, 0);
    for (int q = 0; q < nq; ++q) qoff[q + 1] = qoff[q] + static_cast<int>(qitems[q].size());

    bcsr_batch *b = new (std::nothrow) bcsr_batch();
    if (!b) return BCSR_ERR_OUT_OF_MEMORY;
    const size_t off_m = 0;
    const size_t off_i = align256(off_m + sizeof(BatchMat) * std::max(count, 1));
    const size_t off_q = align256(off_i + sizeof(int4) * std::max<size_t>(items.size(), 1));
    const size_t off_c = align256(off_q + sizeof(int) * (kMaxQueues + 1));
    const size_t off_j = align256(off_c + sizeof(int) * kMaxQueues);
    std::vector<int> job_of(quant_wg);
    for (size_t jn = 0; jn < jobs.size(); ++jn) {
        const int end = jn + 1 < jobs.size() ? jobs[jn + 1].wg0 : quant_wg;
        for (int w = jobs[jn].wg0; w < end; ++w) job_of[w] = static_cast<int>(jn);
    }
    const size_t off_jo = align256(off_j + sizeof(QuantJob) * std::max<size_t>(jobs.size(), 1));
    const size_t off_w = align256(off_jo + sizeof(int) * std::max<size_t>(job_of.size(), 1));
    const size_t total = off_w + ws_total;
    rc = hip_status(hipMalloc(&b->d_base, total));
    if (rc != BCSR_OK) {
        delete b;
        return rc;
    }
    char *base = static_cast<char *>(b->d_base);
    // FP8: point each problem at its slice of the quantised-B workspace
    for (int i = 0, jn = 0; i < count && fp8; ++i) {
        if (!mats[i].a.b_chunks) continue;
        uint8_t *B8 = reinterpret_cast<uint8_t *>(base + off_w + ws_off[i]);
        uint8_t *sc = B8 + align256(static_cast<size_t>(plans[i]->info.cols) * N[i]);
        mats[i].a.B8 = B8;
        mats[i].a.b8scale = sc;
        jobs[jn].B8 = B8;
        jobs[jn].scale = sc;
        ++jn;
    }
    std::vector<int4> witems;
    witems.reserve(items.size());
    for (const auto &qv : qitems) witems.insert(witems.end(), qv.begin(), qv.end());
    struct Copy {
        size_t off;
        const void *src;
        size_t bytes;
    } copies[] = {{off_m, mats.data(), sizeof(BatchMat) * count},
                  {off_i, witems.data(), sizeof(int4) * witems.size()},
                  {off_q, qoff.data(), sizeof(int) * qoff.size()},
                  {off_j, jobs.data(), sizeof(QuantJob) * jobs.size()},
                  {off_jo, job_of.data(), sizeof(int) * job_of.size()}};
    for (const Copy &c : copies) {
        if (!c.bytes) continue;
        rc = hip_status(hipMemcpy(base + c.off, c.src, c.bytes, hipMemcpyHostToDevice));
        if (rc != BCSR_OK) {
            (void)hipFree(b->d_base);
            delete b;
            return rc;
        }
    }
    b->device = dev;
    b->precision = precision;
    b->count = count;
    b->n_items = static_cast<int>(items.size());
    b->n_queues = nq;
    b->grid = grid;
    b->n_jobs = static_cast<int>(jobs.size());
    b->quant_wg = quant_wg;
    b->quant_grid = std::max(1, std::min(quant_wg, n_cu * 16));
    b->flops = flops;
    b->bytes = total;
    b->d_mats = reinterpret_cast<BatchMat *>(base + off_m);
    b->d_items = reinterpret_cast<int4 *>(base + off_i);
    b->d_qoff = reinterpret_cast<int *>(base + off_q);
    b->d_counter = reinterpret_cast<int *>(base + off_c);
    b->d_jobs = reinterpret_cast<QuantJob *>(base + off_j);
    b->d_job_of = reinterpret_cast<int *>(base + off_jo);
    b->tiles = std::move(tiles);
    *out = b;
    return BCSR_OK;
}

// the host-side work lists are std::vectors: an allocation failure is a
// status, never an exception through the C ABI
extern "C" int bcsr_batch_create(const bcsr_plan *const *plans, const void *const *B, float *const *C,
                                 const int32_t *N, const int32_t *tile_n, int32_t count,
                                 const bcsr_batch_cfg *cfg, bcsr_batch **out) {
    try {
        return batch_create(plans, B, C, N, tile_n, count, cfg, out);
    } catch (const std::bad_alloc &) {
        return BCSR_ERR_OUT_OF_MEMORY;
    }
}

extern "C" int bcsr_batch_run(bcsr_batch *batch, void *stream) {
    if (!batch) return BCSR_ERR_INVALID_ARG;
    return batch_enqueue(batch, static_cast<hipStream_t>(stream));
}

extern "C" int bcsr_batch_run_graph(bcsr_batch *batch, void *stream) {
    if (!batch || !stream) return BCSR_ERR_INVALID_ARG;  // the null stream cannot be captured
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!batch->exec) {
        int rc = hip_status(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        if (rc != BCSR_OK) return rc;
        rc = batch_enqueue(batch, s);
        hipGraph_t g = nullptr;
        const int rc_end = hip_status(hipStreamEndCapture(s, &g));
        if (rc != BCSR_OK || rc_end != BCSR_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc != BCSR_OK ? rc : rc_end;
        }
        rc = hip_status(hipGraphInstantiate(&batch->exec, g, nullptr, nullptr, 0));
        if (rc != BCSR_OK) {
            (void)hipGraphDestroy(g);
            batch->exec = nullptr;
            return rc;
        }
        batch->graph = g;
    }
    return hip_status(hipGraphLaunch(batch->exec, s));
}

extern "C" int bcsr_batch_time(bcsr_batch *batch, int32_t graph, int32_t warmup, int32_t reps, void *stream,
                               float *ms) {
    if (!batch || !ms || reps <= 0 || warmup < 0) return BCSR_ERR_INVALID_ARG;
    auto run = [&]() { return graph ? bcsr_batch_run_graph(batch, stream) : bcsr_batch_run(batch, stream); };
    for (int i = 0; i < warmup; ++i) {
        int rc = run();
        if (rc != BCSR_OK) return rc;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = hip_status(hipEventCreate(&e0));
    if (rc == BCSR_OK) rc = hip_status(hipEventCreate(&e1));
    if (rc == BCSR_OK) rc = hip_status(hipEventRecord(e0, s));
    for (int i = 0; i < reps && rc == BCSR_OK; ++i) rc = run();
    if (rc == BCSR_OK) rc = hip_status(hipEventRecord(e1, s));
    if (rc == BCSR_OK) rc = hip_status(hipEventSynchronize(e1));
    float t = 0.0f;
    if (rc == BCSR_OK) rc = hip_status(hipEventElapsedTime(&t, e0, e1));
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc == BCSR_OK) *ms = t / static_cast<float>(reps);
    return rc;
}

extern "C" int bcsr_batch_get_info(const bcsr_batch *batch, bcsr_batch_info *info) {
    if (!batch || !info) return BCSR_ERR_INVALID_ARG;
    *info = bcsr_batch_info{};
    info->count = batch->count;
    info->items = batch->n_items;
    info->grid = batch->grid;
    info->precision = batch->precision;
    info->flops = batch->flops;
    info->device_bytes = static_cast<int64_t>(batch->bytes);
    return BCSR_OK;
}

extern "C" int bcsr_batch_tiles(const bcsr_batch *batch, int32_t *tiles) {
    if (!batch || (!tiles && batch->count)) return BCSR_ERR_INVALID_ARG;
    std::copy(batch->tiles.begin(), batch->tiles.end(), tiles);
    return BCSR_OK;
}

extern "C" int bcsr_batch_destroy(bcsr_batch *batch) {
    if (!batch) return BCSR_OK;
    int rc = BCSR_OK;
    if (batch->exec) rc = hip_status(hipGraphExecDestroy(batch->exec));
    if (batch->graph) {
        const int r2 = hip_status(hipGraphDestroy(batch->graph));
        if (rc == BCSR_OK) rc = r2;
    }
    if (batch->d_base) {
        const int r3 = hip_status(hipFree(batch->d_base));
        if (rc == BCSR_OK) rc = r3;
    }
    delete batch;
    return rc;
}
