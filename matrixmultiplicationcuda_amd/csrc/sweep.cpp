// sweep.cpp -- bcsr_sweep, the axis-sweep benchmark runner: the counterpart of
// the reference's NVBench harness (nvbench_main.cu).  Every benchmark of that
// harness is a name here with the same axes M, K, N, SPARS (percent) and PAT
// (index into the reference's pattern list, nvbench_main.cu:56-64), the same
// inputs (A = generate_matrix<float>(M, K, SPARS/100, pattern, bh, 123),
// B = generate_matrix<float>(K, N, 0, "random", bh, 456), both rounded to
// fp16, nvbench_main.cu:139-175) and the same measurement: the output is
// cleared outside the timer and each sample times one launch with events
// (nvbench_main.cu:396-404).  Like NVBench's cold measurement each sample
// starts with cold caches: a flush buffer larger than the MI355X's L2 and
// Infinity Cache is rewritten before it.  Samples are taken until both
// --min-samples and --min-time are reached (or --max-samples).
//
// Each state is checked against the dense product (rocBLAS fp16 GEMM with
// fp32 accumulation): the reference's per-entry bound |C - C_ref| <= 1e-2
// (nvbench_main.cu:357-364) for the fp16 products, and for the
// reduced-precision storage modes max |C - C_ref| <= tol * max (|A| |B|)
// (tol 2^-7 for bf16, 2^-3 for FP8): normwise, because a value far below its
// FP8 scale group's maximum is stored as a subnormal with a large relative
// error, which a single-product entry (a diagonal A) shows unaveraged.  The
// largest per-entry ratio |C - C_ref| / (|A| |B|) is reported as well.  A
// failed check is reported and makes the exit status 1 instead of aborting
// the run.
//
// Output: a markdown table per benchmark on stdout (as NVBench prints), and
// optionally CSV (--csv) and JSON (--json) with one record per state.
//
// Usage:
//   bcsr_sweep --list
//   bcsr_sweep [-b NAME|INDEX]... [-a AXIS=VALUES]... [--csv FILE] [--json FILE]
//              [--min-samples N] [--max-samples N] [--min-time SEC]
//              [--flush-mib MiB] [--no-verify] [--device D]
// With no -b it runs the benchmarks the reference registers, on the axes it
// registers them with (nvbench_main.cu:1417-1451).  VALUES is "v", "v1,v2",
// "[v1,v2]" or "[start:stop:step]" (inclusive); an -a replaces that axis in
// every axis set of the selected benchmarks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "BCSRMatrix.h"
#include "CSRMatrix.h"
#include "Matrix.h"
#include "bcsr_mi355x.h"
#include "matrix_generator.h"

namespace {

const std::vector<std::string> kPatterns = {"random",        "checkerboard",  "diagonal",     "blockdiagonal",
                                            "blockrandom",   "pattern64by16", "pattern32by16"};

enum class Kind {
    Entry,          // a sparseMatrixMulTensor* entry point (registered device arrays)
    Entry1,         // sparseMatrixMulTensor1 (square n)
    Adaptive,       // sparseMatrixMulTensorAdaptive* (host colBlocksPerRow rule)
    Mode,           // this library's storage-precision modes through a plan
    Csr1Co,         // CSR scalar product
    Csr1,           // CSR scalar product summed onto C
    DenseBlas,      // rocBLAS fp16 GEMM (the reference's cuBLAS rows)
    RocsparseBsr,   // rocSPARSE generic SpMM, BSR
    RocsparseCsr    // rocSPARSE generic SpMM, CSR
};

using EntryFn = int (*)(const int *, const int *, const void *, const void *, float *, unsigned, unsigned, void *);
using AdaptiveFn = int (*)(const int *, const int *, const void *, const void *, float *, unsigned, unsigned, unsigned,
                           void *);

struct Axes {
    std::vector<long> N, M, K, SPARS, PAT;
};

// the three axis sets the reference registers each benchmark with
std::vector<Axes> reference_axes(int pat_1024) {
    return {Axes{{32, 64, 128}, {1024}, {256}, {60, 65, 70, 75, 80, 85, 90}, {pat_1024}},
            Axes{{32, 64, 128}, {512}, {512}, {60, 70, 80, 90}, {0, 1, 2, 3, 4, 5, 6}},
            Axes{{32, 64, 128}, {2048}, {256}, {60, 70, 80, 90}, {0, 1, 2, 3, 4, 5, 6}}};
}

struct Bench {
    std::string name;
    Kind kind;
    int bh, bw;
    bool registered;     // registered by the reference's harness
    const char *source;  // the reference benchmark or kernel it stands for
    EntryFn entry = nullptr;
    AdaptiveFn adaptive = nullptr;
    int precision = BCSR_PREC_F16;
    std::vector<Axes> axes;
};

std::vector<Bench> registry() {
    std::vector<Bench> b;
    auto add = [&](std::string name, Kind kind, int bh, bool reg, const char *src, int pat) -> Bench & {
        Bench x{std::move(name), kind, bh, 16, reg, src};
        x.axes = reference_axes(pat);
        b.push_back(std::move(x));
        return b.back();
    };
    // registered by the reference (nvbench_main.cu:1417-1451)
    add("cuBLAS_GEMM_TENSOR", Kind::DenseBlas, 16, true, "nvbench_main.cu:1129", 5);
    add("sparseMatrixMulTensor_v2", Kind::Entry, 16, true, "nvbench_main.cu:380", 5).entry = sparseMatrixMulTensor_v2;
    add("sparseMatrixMulTensor32x16", Kind::Entry, 32, true, "nvbench_main.cu:636", 6).entry =
        sparseMatrixMulTensor32x16;
    add("sparseMatrixMulTensor32x16_v2", Kind::Entry, 32, true, "nvbench_main.cu:828", 6).entry =
        sparseMatrixMulTensor32x16_v2;
    add("sparseMatrixMulTensorAdaptive_16x16", Kind::Adaptive, 16, true, "nvbench_main.cu:1233", 5).adaptive =
        sparseMatrixMulTensorAdaptive;
    add("sparseMatrixMulTensorAdaptive_32x16", Kind::Adaptive, 32, true, "nvbench_main.cu:1294", 6).adaptive =
        sparseMatrixMulTensorAdaptive32x16;
    add("sparseMatrixMulTensor32x16_shared_vectorized", Kind::Entry, 32, true, "nvbench_main.cu:927", 6).entry =
        sparseMatrixMulTensor32x16_shared_vectorized;
    add("sparseMatrixMulTensor32x16_v2_shared_vectorized", Kind::Entry, 32, true, "nvbench_main.cu:978", 6).entry =
        sparseMatrixMulTensor32x16_v2_shared_vectorized;
    // defined but not registered by the reference
    add("cuBLAS_GEMM", Kind::DenseBlas, 16, false, "nvbench_main.cu:1078", 5);
    add("sparseMatrixMult1", Kind::Csr1, 16, false, "nvbench_main.cu:279", 5);
    add("sparseMatrixMulTensor", Kind::Entry, 16, false, "nvbench_main.cu:326", 5).entry = sparseMatrixMulTensor;
    add("sparseMatrixMulTensor_v3", Kind::Entry, 16, false, "nvbench_main.cu:433", 5).entry = sparseMatrixMulTensor_v3;
    add("sparseMatrixMulTensor_v1_improved", Kind::Entry, 16, false, "nvbench_main.cu:584", 5).entry =
        sparseMatrixMulTensor_v1_improved;
    add("sparseMatrixMulTensor_option2_ldmatrix_sm80", Kind::Entry, 16, false, "nvbench_main.cu:1185", 5).entry =
        sparseMatrixMulTensor_option2_ldmatrix_sm80;
    add("sparseMatrixMulTensor64x16", Kind::Entry, 64, false, "nvbench_main.cu:531", 5).entry =
        sparseMatrixMulTensor64x16;
    add("sparseMatrixMulTensor64x16_v2", Kind::Entry, 64, false, "nvbench_main.cu:1029", 5).entry =
        sparseMatrixMulTensor64x16_v2;
    add("sparseMatrixMulTensor32x16_shared", Kind::Entry, 32, false, "nvbench_main.cu:876", 6).entry =
        sparseMatrixMulTensor32x16_shared;
    add("sparseMatrixMulTensorAdaptive_64x16", Kind::Adaptive, 64, false, "nvbench_main.cu:1355", 5).adaptive =
        sparseMatrixMulTensorAdaptive64x16;
    // legacy-driver rows and this library's extensions
    add("sparseMatrixMulTensor1", Kind::Entry1, 16, false, "main.cu:496-501", 5);
    add("sparseMatrixMult1Co", Kind::Csr1Co, 16, false, "main.cu:436-444", 5);
    add("rocSPARSE_CSR", Kind::RocsparseCsr, 16, false, "main.cu:530-563 (cuSPARSE CSR)", 5);
    add("rocSPARSE_BSR", Kind::RocsparseBsr, 16, false, "library comparison", 5);
    add("bcsr_spmm_bf16", Kind::Mode, 16, false, "storage mode", 5).precision = BCSR_PREC_BF16;
    add("bcsr_spmm_fp8_e4m3", Kind::Mode, 16, false, "storage mode", 5).precision = BCSR_PREC_FP8_E4M3;
    add("bcsr_spmm_hilo", Kind::Mode, 16, false, "storage mode", 5).precision = BCSR_PREC_F16_HILO;
    return b;
}

struct Options {
    std::vector<std::string> benches;
    std::map<std::string, std::vector<long>> axis_over;
    std::string csv, json;
    int min_samples = 10, max_samples = 1000, device = 0;
    double min_time = 0.02;  // seconds of timed GPU work per state
    long flush_mib = 512;
    bool verify = true;
};

[[noreturn]] void usage_error(const std::string &msg) {
    std::fprintf(stderr, "bcsr_sweep: %s\n(run with --list for the benchmarks, --help for the options)\n", msg.c_str());
    std::exit(2);
}

long parse_long(const std::string &s) {
    char *end = nullptr;
    const long v = std::strtol(s.c_str(), &end, 10);
    if (s.empty() || *end) usage_error("not an integer: '" + s + "'");
    return v;
}

std::vector<long> parse_values(std::string v) {
    if (!v.empty() && v.front() == '[') {
        if (v.back() != ']') usage_error("unterminated '[' in " + v);
        v = v.substr(1, v.size() - 2);
    }
    std::vector<long> out;
    if (v.find(':') != std::string::npos) {
        std::vector<std::string> p;
        std::stringstream ss(v);
        for (std::string t; std::getline(ss, t, ':');) p.push_back(t);
        if (p.size() < 2 || p.size() > 3) usage_error("range must be [start:stop] or [start:stop:step]");
        const long a = parse_long(p[0]), z = parse_long(p[1]), step = p.size() == 3 ? parse_long(p[2]) : 1;
        if (step <= 0 || z < a) usage_error("empty range " + v);
        for (long x = a; x <= z; x += step) out.push_back(x);
        return out;
    }
    std::stringstream ss(v);
    for (std::string t; std::getline(ss, t, ',');) out.push_back(parse_long(t));
    if (out.empty()) usage_error("no values in " + v);
    return out;
}

void print_help() {
    std::printf(
        "usage: bcsr_sweep [--list] [-b NAME|INDEX]... [-a AXIS=VALUES]... [--csv FILE] [--json FILE]\n"
        "                  [--min-samples N] [--max-samples N] [--min-time SEC] [--flush-mib MiB]\n"
        "                  [--no-verify] [--device D]\n"
        "axes: M, K, N, SPARS (percent zero), PAT (0 random, 1 checkerboard, 2 diagonal, 3 blockdiagonal,\n"
        "      4 blockrandom, 5 pattern64by16, 6 pattern32by16; taken modulo 7)\n"
        "VALUES: v | v1,v2 | [v1,v2] | [start:stop:step]\n");
}

Options parse(int argc, char **argv, const std::vector<Bench> &reg, bool &list_only) {
    Options o;
    list_only = false;
    auto need = [&](int &i) -> std::string {
        if (i + 1 >= argc) usage_error(std::string("missing value after ") + argv[i]);
        return argv[++i];
    };
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--list" || a == "-l") {
            list_only = true;
        } else if (a == "--help" || a == "-h") {
            print_help();
            std::exit(0);
        } else if (a == "-b" || a == "--benchmark") {
            o.benches.push_back(need(i));
        } else if (a == "-a" || a == "--axis") {
            const std::string v = need(i);
            const size_t eq = v.find('=');
            if (eq == std::string::npos) usage_error("axis needs NAME=VALUES: " + v);
            std::string name = v.substr(0, eq);
            const size_t br = name.find('[');  // NVBench's "N[pow2]" style flags are not supported
            if (br != std::string::npos) usage_error("axis flags are not supported: " + name);
            if (name != "M" && name != "K" && name != "N" && name != "SPARS" && name != "PAT")
                usage_error("unknown axis " + name);
            o.axis_over[name] = parse_values(v.substr(eq + 1));
        } else if (a == "--csv") {
            o.csv = need(i);
        } else if (a == "--json") {
            o.json = need(i);
        } else if (a == "--min-samples") {
            o.min_samples = static_cast<int>(parse_long(need(i)));
        } else if (a == "--max-samples") {
            o.max_samples = static_cast<int>(parse_long(need(i)));
        } else if (a == "--min-time") {
            o.min_time = std::atof(need(i).c_str());
        } else if (a == "--flush-mib") {
            o.flush_mib = parse_long(need(i));
        } else if (a == "--no-verify") {
            o.verify = false;
        } else if (a == "--device") {
            o.device = static_cast<int>(parse_long(need(i)));
        } else {
            usage_error("unknown option " + a);
        }
    }
    if (o.min_samples < 1 || o.max_samples < o.min_samples) usage_error("need 1 <= min-samples <= max-samples");
    if (o.flush_mib < 0 || o.flush_mib > 8192) usage_error("flush-mib must be in [0, 8192]");
    for (const std::string &b : o.benches) {
        bool ok = false;
        for (size_t k = 0; k < reg.size(); ++k)
            ok |= reg[k].name == b || (std::all_of(b.begin(), b.end(), ::isdigit) && std::to_string(k) == b);
        if (!ok) usage_error("unknown benchmark " + b);
    }
    return o;
}

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void bcsr_check(int rc, const char *what) {
    if (rc != BCSR_OK) throw std::runtime_error(std::string(what) + ": " + bcsr_status_string(rc));
}

struct DevBuf {
    void *p = nullptr;
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { hip_check(hipMalloc(&p, bytes ? bytes : 16), "hipMalloc"); }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

struct Result {
    std::string bench, pattern, status, note;
    long M = 0, K = 0, N = 0, spars = 0, pat = 0;
    int bh = 16, bw = 16;
    int64_t stored = 0;  // stored entries (nnzb * bh * bw, nnz or M * K)
    int samples = 0;
    double mean_ms = 0, min_ms = 0, max_ms = 0, noise = 0;
    double flops = 0, bytes = 0, max_abs = 0, max_rel = 0;
};

// host inputs of one (M, K, N, SPARS, PAT, bh) state
struct Inputs {
    std::unique_ptr<Matrix> A, B;
    std::vector<float> absAB;  // |A| |B| (for the relative bounds), filled on demand
};

void generate_inputs(long M, long K, long N, double sp, const std::string &pattern, int bh, Inputs &in) {
    mg::GenSpec sa;
    sa.pattern = pattern;
    sa.sparsity = sp;
    sa.bh = sa.bw = bh;
    in.A = std::make_unique<Matrix>(static_cast<int>(M), static_cast<int>(K));
    mg::generate_half_into(in.A->data, static_cast<int>(M), static_cast<int>(K), sa, 123u);
    mg::GenSpec sb;
    sb.pattern = "random";
    sb.sparsity = 0.0;
    sb.bh = sb.bw = bh;
    in.B = std::make_unique<Matrix>(static_cast<int>(K), static_cast<int>(N));
    mg::generate_half_into(in.B->data, static_cast<int>(K), static_cast<int>(N), sb, 456u);
    in.absAB.clear();
}

class Runner {
public:
    Runner(const Options &o) : opt_(o) {
        hip_check(hipSetDevice(o.device), "hipSetDevice");
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, o.device), "hipGetDeviceProperties");
        device_ = std::string("[") + std::to_string(o.device) + "] " + (prop.name[0] ? prop.name : "AMD GPU") + " (" +
                  prop.gcnArchName + ", " + std::to_string(prop.multiProcessorCount) + " CUs)";
        hip_check(hipStreamCreate(&stream_), "hipStreamCreate");
        hip_check(hipEventCreate(&t0_), "hipEventCreate");
        hip_check(hipEventCreate(&t1_), "hipEventCreate");
        if (o.flush_mib) flush_ = std::make_unique<DevBuf>(static_cast<size_t>(o.flush_mib) << 20);
    }
    ~Runner() {
        (void)hipEventDestroy(t0_);
        (void)hipEventDestroy(t1_);
        (void)hipStreamDestroy(stream_);
    }
    const std::string &device() const { return device_; }

    Result run(const Bench &b, long M, long K, long N, long spars, long pat);

private:
    // times call() until the sample targets are met; C is cleared before each
    void measure(const std::function<void()> &call, float *C, size_t c_bytes, Result &r);
    void dense_reference(const Inputs &in, long M, long K, long N, std::vector<float> &C);

    const Options &opt_;
    std::string device_;
    hipStream_t stream_ = nullptr;
    hipEvent_t t0_ = nullptr, t1_ = nullptr;
    std::unique_ptr<DevBuf> flush_;
    unsigned flush_round_ = 0;
};

void Runner::measure(const std::function<void()> &call, float *C, size_t c_bytes, Result &r) {
    call();  // first-use costs (plan registry, library handles, code objects)
    hip_check(hipStreamSynchronize(stream_), "warm-up");
    std::vector<double> t;
    double total = 0;
    while (static_cast<int>(t.size()) < opt_.max_samples &&
           (static_cast<int>(t.size()) < opt_.min_samples || total < opt_.min_time)) {
        hip_check(hipMemsetAsync(C, 0, c_bytes, stream_), "clear C");
        if (flush_)
            hip_check(hipMemsetAsync(flush_->p, static_cast<int>(++flush_round_ & 0xff),
                                     static_cast<size_t>(opt_.flush_mib) << 20, stream_),
                      "cache flush");
        hip_check(hipEventRecord(t0_, stream_), "hipEventRecord");
        call();
        hip_check(hipEventRecord(t1_, stream_), "hipEventRecord");
        hip_check(hipEventSynchronize(t1_), "hipEventSynchronize");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, t0_, t1_), "hipEventElapsedTime");
        t.push_back(ms);
        total += ms * 1e-3;
    }
    double s = 0, s2 = 0;
    for (double x : t) s += x;
    r.samples = static_cast<int>(t.size());
    r.mean_ms = s / t.size();
    for (double x : t) s2 += (x - r.mean_ms) * (x - r.mean_ms);
    r.min_ms = *std::min_element(t.begin(), t.end());
    r.max_ms = *std::max_element(t.begin(), t.end());
    r.noise = t.size() > 1 ? std::sqrt(s2 / (t.size() - 1)) / r.mean_ms : 0.0;
}

void Runner::dense_reference(const Inputs &in, long M, long K, long N, std::vector<float> &C) {
    DevBuf dA(sizeof(half) * M * K), dB(sizeof(half) * K * N), dC(sizeof(float) * M * N);
    hip_check(hipMemcpy(dA.p, in.A->data, sizeof(half) * M * K, hipMemcpyHostToDevice), "copy A");
    hip_check(hipMemcpy(dB.p, in.B->data, sizeof(half) * K * N, hipMemcpyHostToDevice), "copy B");
    bcsr_check(bcsr_dense_gemm_f16(dA.p, dB.p, dC.as<float>(), M, K, N, stream_), "dense reference");
    C.resize(static_cast<size_t>(M) * N);
    hip_check(hipStreamSynchronize(stream_), "dense reference");
    hip_check(hipMemcpy(C.data(), dC.p, sizeof(float) * M * N, hipMemcpyDeviceToHost), "copy C_ref");
}

// |A| |B| on the device, as the scale of the relative bounds
void abs_product(const Inputs &in, long M, long K, long N, hipStream_t s, std::vector<float> &out) {
    std::vector<half> a(in.A->data, in.A->data + M * K), b(in.B->data, in.B->data + K * N);
    for (half &x : a) x = float2half_rn(std::fabs(half2float(x)));
    for (half &x : b) x = float2half_rn(std::fabs(half2float(x)));
    DevBuf dA(sizeof(half) * M * K), dB(sizeof(half) * K * N), dC(sizeof(float) * M * N);
    hip_check(hipMemcpy(dA.p, a.data(), sizeof(half) * M * K, hipMemcpyHostToDevice), "copy |A|");
    hip_check(hipMemcpy(dB.p, b.data(), sizeof(half) * K * N, hipMemcpyHostToDevice), "copy |B|");
    bcsr_check(bcsr_dense_gemm_f16(dA.p, dB.p, dC.as<float>(), M, K, N, s), "|A||B|");
    out.resize(static_cast<size_t>(M) * N);
    hip_check(hipStreamSynchronize(s), "|A||B|");
    hip_check(hipMemcpy(out.data(), dC.p, sizeof(float) * M * N, hipMemcpyDeviceToHost), "copy |A||B|");
}

double mode_tolerance(int prec) {
    switch (prec) {
    case BCSR_PREC_BF16:
        return std::ldexp(1.0, -7);
    case BCSR_PREC_FP8_E4M3:
        return 0.125;  // A and B normal values each within 2^-4 relative (3 mantissa bits)
    default:
        return 1e-3;
    }
}

Result Runner::run(const Bench &b, long M, long K, long N, long spars, long pat) {
    Result r;
    r.bench = b.name;
    r.M = M;
    r.K = K;
    r.N = N;
    r.spars = spars;
    r.pat = pat;
    r.bh = b.bh;
    r.bw = b.bw;
    r.pattern = kPatterns[static_cast<size_t>(((pat % 7) + 7) % 7)];
    // shapes the products need (the reference has no partial-tile guards)
    std::string why;
    if (M <= 0 || K <= 0 || N <= 0) why = "M, K, N must be positive";
    else if (N % 16) why = "N % 16 != 0";
    else if (M % b.bh || K % b.bw) why = "M or K not a whole number of blocks";
    else if (b.kind == Kind::Entry1 && (M != N || K != N)) why = "square n only";
    else if (spars < 0 || spars > 100) why = "SPARS outside [0, 100]";
    else if (M * K > (1L << 30) || K * N > (1L << 30) || M * N > (1L << 30)) why = "operand above 2^30 entries";
    if (!why.empty()) {
        r.status = "skip";
        r.note = why;
        return r;
    }
    Inputs in;
    generate_inputs(M, K, N, spars / 100.0, r.pattern, b.bh, in);
    const size_t mn = static_cast<size_t>(M) * N;
    DevBuf dB(sizeof(half) * K * N), dC(sizeof(float) * mn);
    hip_check(hipMemcpy(dB.p, in.B->data, sizeof(half) * K * N, hipMemcpyHostToDevice), "copy B");
    float *C = dC.as<float>();
    const double q_bc = 2.0 * K * N + 4.0 * mn;  // B read, C written
    std::function<void()> call;
    std::function<void()> cleanup = [] {};
    std::unique_ptr<BCSRMatrix> bcsr;
    std::unique_ptr<CSRMatrix> csr;
    std::unique_ptr<DevBuf> dA, ph, pi, pd;
    int *gh = nullptr, *gi = nullptr, *ch = nullptr, *ci = nullptr;
    half *gd = nullptr, *cd = nullptr;
    bcsr_plan *plan = nullptr;
    bcsr_sparse_ref *sref = nullptr;
    void *s = stream_;
    const unsigned uM = static_cast<unsigned>(M), uN = static_cast<unsigned>(N);
    double sA = 2.0;
    switch (b.kind) {
    case Kind::DenseBlas:
        dA = std::make_unique<DevBuf>(sizeof(half) * M * K);
        hip_check(hipMemcpy(dA->p, in.A->data, sizeof(half) * M * K, hipMemcpyHostToDevice), "copy A");
        r.stored = M * K;
        r.flops = 2.0 * M * K * N;
        r.bytes = 2.0 * M * K + q_bc;
        call = [&] { bcsr_check(bcsr_dense_gemm_f16(dA->p, dB.p, C, M, K, N, s), b.name.c_str()); };
        break;
    case Kind::Csr1Co:
    case Kind::Csr1:
    case Kind::RocsparseCsr:
        csr = std::make_unique<CSRMatrix>(*in.A);
        csr->copyToDevice(&ch, &ci, &cd);
        cleanup = [&] {
            (void)hipFree(ch);
            (void)hipFree(ci);
            (void)hipFree(cd);
        };
        r.stored = csr->nnz();
        r.flops = 2.0 * r.stored * N;
        r.bytes = 6.0 * r.stored + 4.0 * (M + 1) + q_bc;
        if (b.kind == Kind::RocsparseCsr) {
            bcsr_check(bcsr_rocsparse_csr_create(ch, ci, cd, M, K, r.stored, dB.p, C, N, &sref), "rocsparse csr");
            call = [&] { bcsr_check(bcsr_rocsparse_run(sref, s), b.name.c_str()); };
        } else {
            const int acc = b.kind == Kind::Csr1 ? 2 : 0;
            call = [&, acc] { bcsr_check(bcsr_csr_spmm(ch, ci, cd, dB.p, C, M, K, N, acc, s), b.name.c_str()); };
        }
        break;
    default: {
        bcsr = std::make_unique<BCSRMatrix>(*in.A, b.bh, b.bw);
        const int64_t nnzb = bcsr->nnzb();
        r.stored = nnzb * b.bh * b.bw;
        r.flops = 2.0 * r.stored * N;
        if (b.kind == Kind::Mode) {
            std::vector<float> f32(static_cast<size_t>(r.stored));
            for (size_t i = 0; i < f32.size(); ++i) f32[i] = half2float(bcsr->data[i]);
            bcsr_check(bcsr_plan_create_f32(bcsr->hdr, bcsr->idx, f32.data(), bcsr->blockRows, bcsr->blockCols,
                                            b.bh, b.bw, b.precision, &plan),
                       "plan");
            sA = b.precision == BCSR_PREC_FP8_E4M3 ? 1.0 : b.precision == BCSR_PREC_F16_HILO ? 4.0 : 2.0;
            call = [&] { bcsr_check(bcsr_spmm(plan, dB.p, C, static_cast<int32_t>(N), s), b.name.c_str()); };
        } else if (b.kind == Kind::RocsparseBsr) {
            ph = std::make_unique<DevBuf>(sizeof(int) * (bcsr->blockRows + 1));
            pi = std::make_unique<DevBuf>(sizeof(int) * std::max<int64_t>(nnzb, 1));
            pd = std::make_unique<DevBuf>(sizeof(half) * std::max<int64_t>(r.stored, 1));
            hip_check(hipMemcpy(ph->p, bcsr->hdr, sizeof(int) * (bcsr->blockRows + 1), hipMemcpyHostToDevice), "hdr");
            hip_check(hipMemcpy(pi->p, bcsr->idx, sizeof(int) * nnzb, hipMemcpyHostToDevice), "idx");
            hip_check(hipMemcpy(pd->p, bcsr->data, sizeof(half) * r.stored, hipMemcpyHostToDevice), "data");
            bcsr_check(bcsr_rocsparse_create(ph->as<int>(), pi->as<int>(), pd->p, bcsr->blockRows, bcsr->blockCols,
                                             nnzb, b.bh, dB.p, C, N, &sref),
                       "rocsparse bsr");
            call = [&] { bcsr_check(bcsr_rocsparse_run(sref, s), b.name.c_str()); };
        } else {
            bcsr->copyToDevice(&gh, &gi, &gd);
            cleanup = [&] { bcsr_release_device(gh); };
            if (b.kind == Kind::Entry) {
                call = [&] { bcsr_check(b.entry(gh, gi, gd, dB.p, C, uM, uN, s), b.name.c_str()); };
            } else if (b.kind == Kind::Entry1) {
                call = [&] { bcsr_check(sparseMatrixMulTensor1(gh, gi, gd, dB.p, C, uN, s), b.name.c_str()); };
            } else {
                // the reference's host rule for colBlocksPerRow (nvbench_main.cu:1245-1252)
                const double ar = static_cast<double>(N) / static_cast<double>(M);
                const unsigned cbpr = ar >= 3.0 ? 3u : ar >= 2.0 ? 2u : 1u;
                call = [&, cbpr] { bcsr_check(b.adaptive(gh, gi, gd, dB.p, C, uM, uN, cbpr, s), b.name.c_str()); };
            }
        }
        r.bytes = r.stored * sA + 4.0 * nnzb + 4.0 * (M / b.bh + 1) + q_bc +
                  (b.precision == BCSR_PREC_FP8_E4M3 ? 4.0 * nnzb : 0.0);
        break;
    }
    }
    try {
        measure(call, C, sizeof(float) * mn, r);
        if (opt_.verify) {
            std::vector<float> got(mn), ref;
            hip_check(hipStreamSynchronize(stream_), "sync");
            hip_check(hipMemcpy(got.data(), C, sizeof(float) * mn, hipMemcpyDeviceToHost), "copy C");
            if (b.kind == Kind::DenseBlas) {
                // the dense row is checked against this library's BCSR product
                BCSRMatrix bc(*in.A, 16, 16);
                int *h2 = nullptr, *i2 = nullptr;
                half *d2 = nullptr;
                bc.copyToDevice(&h2, &i2, &d2);
                DevBuf dR(sizeof(float) * mn);
                const int rc = sparseMatrixMulTensor(h2, i2, d2, dB.p, dR.as<float>(), uM, uN, s);
                hip_check(hipStreamSynchronize(stream_), "sync");
                ref.resize(mn);
                hip_check(hipMemcpy(ref.data(), dR.p, sizeof(float) * mn, hipMemcpyDeviceToHost), "copy ref");
                bcsr_release_device(h2);
                bcsr_check(rc, "check product");
            } else {
                dense_reference(in, M, K, N, ref);
            }
            abs_product(in, M, K, N, stream_, in.absAB);
            bool ok = true;
            const bool f16 = b.kind != Kind::Mode || b.precision == BCSR_PREC_F16_HILO;
            double scale = 0.0;
            for (size_t i = 0; i < mn; ++i) {
                const double d = std::fabs(static_cast<double>(got[i]) - ref[i]);
                const double sc = in.absAB[i];
                scale = std::max(scale, sc);
                r.max_abs = std::max(r.max_abs, d);
                if (sc > 0) r.max_rel = std::max(r.max_rel, d / sc);
                if (!(d == d) || (f16 && !(d <= 1e-2))) ok = false;  // NaN fails too
            }
            if (!f16 && !(r.max_abs <= mode_tolerance(b.precision) * scale)) ok = false;
            r.status = ok ? "pass" : "FAIL";
        } else {
            r.status = "unchecked";
        }
    } catch (...) {
        if (plan) bcsr_plan_destroy(plan);
        if (sref) bcsr_rocsparse_destroy(sref);
        cleanup();
        throw;
    }
    if (plan) bcsr_plan_destroy(plan);
    if (sref) bcsr_rocsparse_destroy(sref);
    cleanup();
    return r;
}

std::string fmt_time(double ms) {
    char buf[64];
    if (ms < 1.0) std::snprintf(buf, sizeof buf, "%.3f us", ms * 1e3);
    else std::snprintf(buf, sizeof buf, "%.3f ms", ms);
    return buf;
}

std::string json_escape(const std::string &s) {
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o;
}

constexpr double kHbmPeak = 8.0e12;     // bytes/s
constexpr double kF16Peak = 2.5e15;     // dense fp16/bf16 MFMA flop/s
constexpr double kFp8Peak = 5.0e15;     // dense fp8 MFMA flop/s
constexpr double kFp32VecPeak = 157.3e12;

double roofline(const Bench &b, const Result &r) {
    double peak = kF16Peak;
    if (b.kind == Kind::Mode && b.precision == BCSR_PREC_FP8_E4M3) peak = kFp8Peak;
    if (b.kind == Kind::Csr1Co || b.kind == Kind::Csr1) peak = kFp32VecPeak;
    const double t_star = std::max(r.flops / peak, r.bytes / kHbmPeak);
    return r.mean_ms > 0 ? t_star / (r.mean_ms * 1e-3) : 0.0;
}

}  // namespace

int main(int argc, char **argv) {
    const std::vector<Bench> reg = registry();
    bool list_only = false;
    const Options opt = parse(argc, argv, reg, list_only);
    if (list_only) {
        std::printf("| # | benchmark | block | registered | stands for |\n|---|---|---|---|---|\n");
        for (size_t k = 0; k < reg.size(); ++k)
            std::printf("| %zu | %s | %dx%d | %s | %s |\n", k, reg[k].name.c_str(), reg[k].bh, reg[k].bw,
                        reg[k].registered ? "yes" : "no", reg[k].source);
        return 0;
    }
    std::vector<const Bench *> sel;
    if (opt.benches.empty()) {
        for (const Bench &b : reg)
            if (b.registered) sel.push_back(&b);
    } else {
        for (const std::string &n : opt.benches)
            for (size_t k = 0; k < reg.size(); ++k)
                if (reg[k].name == n || std::to_string(k) == n) sel.push_back(&reg[k]);
    }
    std::vector<std::pair<const Bench *, Result>> results;
    int failures = 0;
    try {
        Runner run(opt);
        for (const Bench *b : sel) {
            std::vector<Axes> sets = b->axes;
            for (Axes &ax : sets)
                for (const auto &kv : opt.axis_over) {
                    if (kv.first == "M") ax.M = kv.second;
                    if (kv.first == "K") ax.K = kv.second;
                    if (kv.first == "N") ax.N = kv.second;
                    if (kv.first == "SPARS") ax.SPARS = kv.second;
                    if (kv.first == "PAT") ax.PAT = kv.second;
                }
            // identical sets after the overrides run once
            std::vector<Axes> uniq;
            for (const Axes &ax : sets) {
                bool dup = false;
                for (const Axes &u : uniq)
                    dup |= u.M == ax.M && u.K == ax.K && u.N == ax.N && u.SPARS == ax.SPARS && u.PAT == ax.PAT;
                if (!dup) uniq.push_back(ax);
            }
            std::printf("\n# %s\n\n## %s\n\n", b->name.c_str(), run.device().c_str());
            std::printf("| N | M | K | SPARS | PAT | Samples | GPU Time | Noise | Elem/s | TFLOP/s | GB/s | Roofline | "
                        "Check |\n|---|---|---|---|---|---|---|---|---|---|---|---|---|\n");
            for (const Axes &ax : uniq)
                for (long pat : ax.PAT)
                    for (long sp : ax.SPARS)
                        for (long K : ax.K)
                            for (long M : ax.M)
                                for (long N : ax.N) {
                                    Result r = run.run(*b, M, K, N, sp, pat);
                                    if (r.status == "FAIL") ++failures;
                                    if (r.status == "skip") {
                                        std::printf("| %ld | %ld | %ld | %ld | %ld | - | - | - | - | - | - | - | "
                                                    "skip: %s |\n",
                                                    N, M, K, sp, pat, r.note.c_str());
                                    } else {
                                        const double t = r.mean_ms * 1e-3;
                                        std::printf("| %ld | %ld | %ld | %ld | %ld | %dx | %s | %.2f%% | %.3e | %.2f | "
                                                    "%.1f | %.3f | %s |\n",
                                                    N, M, K, sp, pat, r.samples, fmt_time(r.mean_ms).c_str(),
                                                    r.noise * 100.0, static_cast<double>(M) * N / t,
                                                    r.flops / t * 1e-12, r.bytes / t * 1e-9, roofline(*b, r),
                                                    r.status.c_str());
                                    }
                                    std::fflush(stdout);
                                    results.emplace_back(b, r);
                                }
        }
        if (!opt.csv.empty()) {
            std::ofstream f(opt.csv);
            if (!f) throw std::runtime_error("cannot write " + opt.csv);
            f << "benchmark,device,M,K,N,SPARS,PAT,pattern,bh,bw,stored,samples,gpu_time_mean_ms,gpu_time_min_ms,"
                 "gpu_time_max_ms,noise,elem_per_s,tflops,gbps,roofline_frac,max_abs_err,max_rel_err,check\n";
            char line[1024];
            for (const auto &br : results) {
                const Result &r = br.second;
                const double t = r.mean_ms * 1e-3;
                std::snprintf(line, sizeof line,
                              "%s,\"%s\",%ld,%ld,%ld,%ld,%ld,%s,%d,%d,%lld,%d,%.6f,%.6f,%.6f,%.5f,%.6e,%.4f,%.3f,%.4f,"
                              "%.6e,%.6e,%s\n",
                              r.bench.c_str(), run.device().c_str(), r.M, r.K, r.N, r.spars, r.pat,
                              r.pattern.c_str(), r.bh, r.bw, static_cast<long long>(r.stored), r.samples, r.mean_ms,
                              r.min_ms, r.max_ms, r.noise, t > 0 ? static_cast<double>(r.M) * r.N / t : 0.0,
                              t > 0 ? r.flops / t * 1e-12 : 0.0, t > 0 ? r.bytes / t * 1e-9 : 0.0,
                              roofline(*br.first, r), r.max_abs, r.max_rel,
                              r.status == "skip" ? ("skip: " + r.note).c_str() : r.status.c_str());
                f << line;
            }
        }
        if (!opt.json.empty()) {
            std::ofstream f(opt.json);
            if (!f) throw std::runtime_error("cannot write " + opt.json);
            f << "{\"device\": \"" << json_escape(run.device()) << "\", \"settings\": {\"min_samples\": "
              << opt.min_samples << ", \"max_samples\": " << opt.max_samples << ", \"min_time_s\": " << opt.min_time
              << ", \"flush_mib\": " << opt.flush_mib << ", \"verify\": " << (opt.verify ? "true" : "false")
              << "}, \"states\": [";
            char rec[2048];
            for (size_t i = 0; i < results.size(); ++i) {
                const Result &r = results[i].second;
                const double t = r.mean_ms * 1e-3;
                std::snprintf(rec, sizeof rec,
                              "%s\n  {\"benchmark\": \"%s\", \"axes\": {\"M\": %ld, \"K\": %ld, \"N\": %ld, \"SPARS\": "
                              "%ld, \"PAT\": %ld}, \"pattern\": \"%s\", \"block\": [%d, %d], \"stored\": %lld, "
                              "\"samples\": %d, \"gpu_time_mean_ms\": %.6f, \"gpu_time_min_ms\": %.6f, "
                              "\"gpu_time_max_ms\": %.6f, \"noise\": %.5f, \"tflops\": %.4f, \"gbps\": %.3f, "
                              "\"roofline_frac\": %.4f, \"max_abs_err\": %.6e, \"max_rel_err\": %.6e, "
                              "\"check\": \"%s\", \"note\": \"%s\"}",
                              i ? "," : "", r.bench.c_str(), r.M, r.K, r.N, r.spars, r.pat, r.pattern.c_str(), r.bh,
                              r.bw, static_cast<long long>(r.stored), r.samples, r.mean_ms, r.min_ms, r.max_ms,
                              r.noise, t > 0 ? r.flops / t * 1e-12 : 0.0, t > 0 ? r.bytes / t * 1e-9 : 0.0,
                              roofline(*results[i].first, r), r.max_abs, r.max_rel, r.status.c_str(),
                              json_escape(r.note).c_str());
                f << rec;
            }
            f << "\n]}\n";
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    if (failures) std::fprintf(stderr, "%d state(s) failed the check\n", failures);
    return failures ? 1 : 0;
}
