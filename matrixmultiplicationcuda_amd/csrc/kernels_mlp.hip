// kernels_mlp.hip -- the learned dispatcher's cost model: a 3-layer MLP
// (in -> hidden -> hidden -> out, GELU between layers, fp32) with its forward
// pass, backward pass and Adam step as HIP kernels (no reference counterpart:
// the reference's only dispatch rule is the Adaptive kernels' fixed
// colBlocksPerRow choice, nvbench_main.cu:1245-1252).
//
// Layout: activations row-major [n x width]; layer l has W_l [out x in]
// (row o holds the weights of output o) and b_l [out]; all parameters live in
// one flat buffer W1 b1 W2 b2 W3 b3, gradients and the Adam moments in
// buffers of the same layout.  The loss is the masked mean squared error
// sum_{r,t} M (P - Y)^2 / sum M, so unsupported (row, output) pairs carry no
// gradient.  Every gradient element is produced by one thread with a fixed
// summation order: the backward pass is deterministic.
#include <cmath>
#include <cstdint>
#include <new>
#include <random>
#include <vector>

#include "bcsr_internal.h"

namespace {

constexpr float kInvSqrt2 = 0.70710678118654752f;
constexpr float kInvSqrt2Pi = 0.39894228040143268f;

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * kInvSqrt2)); }

// d gelu / dz = Phi(z) + z phi(z)
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * kInvSqrt2)) + z * kInvSqrt2Pi * expf(-0.5f * z * z);
}

// Z = X W^T + b; A = gelu(Z) (act) or Z.  One thread per output element.
__global__ __launch_bounds__(256) void mlp_dense_fwd(const float *__restrict__ X, const float *__restrict__ W,
                                                     const float *__restrict__ b, float *__restrict__ Z,
                                                     float *__restrict__ A, int n, int in, int out, int act) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<int64_t>(n) * out) return;
    const int r = static_cast<int>(t / out), o = static_cast<int>(t % out);
    const float *x = X + static_cast<int64_t>(r) * in;
    const float *w = W + static_cast<int64_t>(o) * in;
    float z = b[o];
    for (int i = 0; i < in; ++i) z = fmaf(x[i], w[i], z);
    Z[t] = z;
    A[t] = act ? gelu(z) : z;
}

// dP = 2 M (P - Y) * inv_count; loss += sum M (P - Y)^2 * inv_count (one
// partial per workgroup, summed by mlp_sum).
__global__ __launch_bounds__(256) void mlp_loss_grad(const float *__restrict__ P, const float *__restrict__ Y,
                                                     const float *__restrict__ M, float *__restrict__ dP,
                                                     float *__restrict__ partial, int64_t count, float inv_count) {
    __shared__ float red[256];
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (t < count) {
        const float m = M[t];
        const float d = m > 0.0f ? P[t] - Y[t] : 0.0f;
        dP[t] = 2.0f * d * inv_count;
        l = d * d * inv_count;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[0] = sum of in[0 .. n) (one workgroup, fixed order).
__global__ __launch_bounds__(256) void mlp_sum(const float *__restrict__ in, float *__restrict__ out, int n) {
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) s += in[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// dZ = dA * gelu'(Z) (act) or dA.
__global__ __launch_bounds__(256) void mlp_act_bwd(const float *__restrict__ dA, const float *__restrict__ Z,
                                                   float *__restrict__ dZ, int64_t count, int act) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= count) return;
    dZ[t] = act ? dA[t] * gelu_grad(Z[t]) : dA[t];
}

// dX = dZ W: dX[r, i] = sum_o dZ[r, o] W[o, i].
__global__ __launch_bounds__(256) void mlp_bwd_x(const float *__restrict__ dZ, const float *__restrict__ W,
                                                 float *__restrict__ dX, int n, int in, int out) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<int64_t>(n) * in) return;
    const int r = static_cast<int>(t / in), i = static_cast<int>(t % in);
    const float *dz = dZ + static_cast<int64_t>(r) * out;
    float s = 0.0f;
    for (int o = 0; o < out; ++o) s = fmaf(dz[o], W[static_cast<int64_t>(o) * in + i], s);
    dX[t] = s;
}

// dW[o, i] = sum_r dZ[r, o] X[r, i] for i < in; thread i == in gives db[o].
__global__ __launch_bounds__(256) void mlp_bwd_w(const float *__restrict__ dZ, const float *__restrict__ X,
                                                 float *__restrict__ dW, float *__restrict__ db, int n, int in,
                                                 int out) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<int64_t>(out) * (in + 1)) return;
    const int o = static_cast<int>(t / (in + 1)), i = static_cast<int>(t % (in + 1));
    float s = 0.0f;
    if (i < in) {
        for (int r = 0; r < n; ++r)
            s = fmaf(dZ[static_cast<int64_t>(r) * out + o], X[static_cast<int64_t>(r) * in + i], s);
        dW[static_cast<int64_t>(o) * in + i] = s;
    } else {
        for (int r = 0; r < n; ++r) s += dZ[static_cast<int64_t>(r) * out + o];
        db[o] = s;
    }
}

// Adam with bias correction factors c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t).
__global__ __launch_bounds__(256) void mlp_adam(float *__restrict__ p, const float *__restrict__ g,
                                                float *__restrict__ m, float *__restrict__ v, int count, float lr,
                                                float b1, float b2, float eps, float c1, float c2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.0f - b1) * gt;
    const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    p[t] -= lr * (mt * c1) / (sqrtf(vt * c2) + eps);
}

unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

// Rows per call: 2^24 rows x 4096 units keeps every grid within 2^28 blocks.
constexpr int32_t kMaxRows = 1 << 24;

}  // namespace

struct bcsr_costmodel {
    int in = 0, hid = 0, out = 0;
    int n_params = 0;
    float *d_params = nullptr, *d_grads = nullptr, *d_m = nullptr, *d_v = nullptr;
    float *d_loss = nullptr;  // [1]
    int step = 0;
    int cap = 0;  // rows the workspace holds
    float *d_ws = nullptr;
    // parameter offsets
    int oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
};

namespace {

struct Acts {
    float *Z1, *A1, *Z2, *A2, *P, *dP, *dA2, *dZ2, *dA1, *dZ1, *dZ3, *partial;
};

int ensure_ws(bcsr_costmodel *cm, int n) {
    if (n <= cm->cap) return BCSR_OK;
    if (cm->d_ws) {
        int rc = hip_status(hipFree(cm->d_ws));
        cm->d_ws = nullptr;
        cm->cap = 0;
        if (rc != BCSR_OK) return rc;
    }
    const size_t per_row = 8 * static_cast<size_t>(cm->hid) + 3 * static_cast<size_t>(cm->out);
    const size_t bytes = sizeof(float) * (per_row * n + blocks_for(static_cast<int64_t>(n) * cm->out) + 64);
    int rc = hip_status(hipMalloc(&cm->d_ws, bytes));
    if (rc != BCSR_OK) return rc;
    cm->cap = n;
    return BCSR_OK;
}

Acts acts(const bcsr_costmodel *cm) {
    const size_t n = cm->cap, h = cm->hid, o = cm->out;
    float *p = cm->d_ws;
    Acts a{};
    a.Z1 = p; p += n * h;
    a.A1 = p; p += n * h;
    a.Z2 = p; p += n * h;
    a.A2 = p; p += n * h;
    a.dA2 = p; p += n * h;
    a.dZ2 = p; p += n * h;
    a.dA1 = p; p += n * h;
    a.dZ1 = p; p += n * h;
    a.P = p; p += n * o;
    a.dP = p; p += n * o;
    a.dZ3 = p; p += n * o;
    a.partial = p;
    return a;
}

int forward(bcsr_costmodel *cm, const float *X, int n, float *P, hipStream_t s) {
    Acts a = acts(cm);
    const float *w = cm->d_params;
    hipLaunchKernelGGL(mlp_dense_fwd, dim3(blocks_for(static_cast<int64_t>(n) * cm->hid)), dim3(256), 0, s, X,
                       w + cm->oW1, w + cm->ob1, a.Z1, a.A1, n, cm->in, cm->hid, 1);
    hipLaunchKernelGGL(mlp_dense_fwd, dim3(blocks_for(static_cast<int64_t>(n) * cm->hid)), dim3(256), 0, s, a.A1,
                       w + cm->oW2, w + cm->ob2, a.Z2, a.A2, n, cm->hid, cm->hid, 1);
    // the output layer has no activation (its Z copy goes to scratch)
    hipLaunchKernelGGL(mlp_dense_fwd, dim3(blocks_for(static_cast<int64_t>(n) * cm->out)), dim3(256), 0, s, a.A2,
                       w + cm->oW3, w + cm->ob3, a.dZ3, P, n, cm->hid, cm->out, 0);
    return hip_status(hipGetLastError());
}

int backward(bcsr_costmodel *cm, const float *X, const float *Y, const float *M, int n, float inv_count,
             hipStream_t s) {
    Acts a = acts(cm);
    int rc = forward(cm, X, n, a.P, s);
    if (rc != BCSR_OK) return rc;
    const float *w = cm->d_params;
    float *g = cm->d_grads;
    const int64_t no = static_cast<int64_t>(n) * cm->out, nh = static_cast<int64_t>(n) * cm->hid;
    const unsigned lb = blocks_for(no);
    hipLaunchKernelGGL(mlp_loss_grad, dim3(lb), dim3(256), 0, s, a.P, Y, M, a.dP, a.partial, no, inv_count);
    hipLaunchKernelGGL(mlp_sum, dim3(1), dim3(256), 0, s, a.partial, cm->d_loss, static_cast<int>(lb));
    // layer 3 (no activation): dZ3 = dP
    hipLaunchKernelGGL(mlp_bwd_w, dim3(blocks_for(static_cast<int64_t>(cm->out) * (cm->hid + 1))), dim3(256), 0, s,
                       a.dP, a.A2, g + cm->oW3, g + cm->ob3, n, cm->hid, cm->out);
    hipLaunchKernelGGL(mlp_bwd_x, dim3(blocks_for(nh)), dim3(256), 0, s, a.dP, w + cm->oW3, a.dA2, n, cm->hid,
                       cm->out);
    // layer 2
    hipLaunchKernelGGL(mlp_act_bwd, dim3(blocks_for(nh)), dim3(256), 0, s, a.dA2, a.Z2, a.dZ2, nh, 1);
    hipLaunchKernelGGL(mlp_bwd_w, dim3(blocks_for(static_cast<int64_t>(cm->hid) * (cm->hid + 1))), dim3(256), 0, s,
                       a.dZ2, a.A1, g + cm->oW2, g + cm->ob2, n, cm->hid, cm->hid);
    hipLaunchKernelGGL(mlp_bwd_x, dim3(blocks_for(nh)), dim3(256), 0, s, a.dZ2, w + cm->oW2, a.dA1, n, cm->hid,
                       cm->hid);
    // layer 1
    hipLaunchKernelGGL(mlp_act_bwd, dim3(blocks_for(nh)), dim3(256), 0, s, a.dA1, a.Z1, a.dZ1, nh, 1);
    hipLaunchKernelGGL(mlp_bwd_w, dim3(blocks_for(static_cast<int64_t>(cm->hid) * (cm->in + 1))), dim3(256), 0, s,
                       a.dZ1, X, g + cm->oW1, g + cm->ob1, n, cm->in, cm->hid);
    return hip_status(hipGetLastError());
}

}  // namespace

extern "C" int bcsr_cm_create(int32_t n_in, int32_t n_hidden, int32_t n_out, uint32_t seed, bcsr_costmodel **out) {
    if (!out || n_in <= 0 || n_hidden <= 0 || n_out <= 0 || n_in > 4096 || n_hidden > 4096 || n_out > 4096)
        return BCSR_ERR_INVALID_ARG;
    *out = nullptr;
    bcsr_costmodel *cm = new (std::nothrow) bcsr_costmodel();
    if (!cm) return BCSR_ERR_OUT_OF_MEMORY;
    cm->in = n_in;
    cm->hid = n_hidden;
    cm->out = n_out;
    cm->oW1 = 0;
    cm->ob1 = cm->oW1 + n_hidden * n_in;
    cm->oW2 = cm->ob1 + n_hidden;
    cm->ob2 = cm->oW2 + n_hidden * n_hidden;
    cm->oW3 = cm->ob2 + n_hidden;
    cm->ob3 = cm->oW3 + n_out * n_hidden;
    cm->n_params = cm->ob3 + n_out;
    // initial weights U(-1/sqrt(fan_in), 1/sqrt(fan_in)), biases zero
    std::vector<float> h;
    try {
        h.assign(cm->n_params, 0.0f);
    } catch (const std::bad_alloc &) {
        delete cm;
        return BCSR_ERR_OUT_OF_MEMORY;
    }
    std::mt19937 rng(seed);
    auto fill = [&](int off, int rows, int fan_in) {
        std::uniform_real_distribution<float> u(-1.0f / std::sqrt(static_cast<float>(fan_in)),
                                                1.0f / std::sqrt(static_cast<float>(fan_in)));
        for (int k = 0; k < rows * fan_in; ++k) h[off + k] = u(rng);
    };
    fill(cm->oW1, n_hidden, n_in);
    fill(cm->oW2, n_hidden, n_hidden);
    fill(cm->oW3, n_out, n_hidden);
    const size_t pb = sizeof(float) * cm->n_params;
    float *base = nullptr;
    int rc = hip_status(hipMalloc(&base, 4 * pb + 256));
    if (rc != BCSR_OK) {
        delete cm;
        return rc;
    }
    cm->d_params = base;
    cm->d_grads = base + cm->n_params;
    cm->d_m = base + 2 * cm->n_params;
    cm->d_v = base + 3 * cm->n_params;
    cm->d_loss = base + 4 * cm->n_params;
    rc = hip_status(hipMemcpy(cm->d_params, h.data(), pb, hipMemcpyHostToDevice));
    if (rc == BCSR_OK) rc = hip_status(hipMemset(cm->d_grads, 0, 3 * pb + sizeof(float)));
    if (rc != BCSR_OK) {
        (void)hipFree(base);
        delete cm;
        return rc;
    }
    *out = cm;
    return BCSR_OK;
}

extern "C" int bcsr_cm_destroy(bcsr_costmodel *cm) {
    if (!cm) return BCSR_OK;
    int rc = BCSR_OK;
    if (cm->d_ws) rc = hip_status(hipFree(cm->d_ws));
    if (cm->d_params) {
        const int r2 = hip_status(hipFree(cm->d_params));
        if (rc == BCSR_OK) rc = r2;
    }
    delete cm;
    return rc;
}

extern "C" int bcsr_cm_num_params(const bcsr_costmodel *cm, int32_t *n) {
    if (!cm || !n) return BCSR_ERR_INVALID_ARG;
    *n = cm->n_params;
    return BCSR_OK;
}

extern "C" int bcsr_cm_set_params(bcsr_costmodel *cm, const float *host, int32_t n) {
    if (!cm || !host || n != cm->n_params) return BCSR_ERR_INVALID_ARG;
    int rc = hip_status(hipMemcpy(cm->d_params, host, sizeof(float) * n, hipMemcpyHostToDevice));
    if (rc == BCSR_OK) rc = hip_status(hipMemset(cm->d_m, 0, 2 * sizeof(float) * cm->n_params));
    cm->step = 0;
    return rc;
}

extern "C" int bcsr_cm_get_params(const bcsr_costmodel *cm, float *host, int32_t n) {
    if (!cm || !host || n != cm->n_params) return BCSR_ERR_INVALID_ARG;
    return hip_status(hipMemcpy(host, cm->d_params, sizeof(float) * n, hipMemcpyDeviceToHost));
}

extern "C" int bcsr_cm_get_grads(const bcsr_costmodel *cm, float *host, int32_t n) {
    if (!cm || !host || n != cm->n_params) return BCSR_ERR_INVALID_ARG;
    return hip_status(hipMemcpy(host, cm->d_grads, sizeof(float) * n, hipMemcpyDeviceToHost));
}

extern "C" int bcsr_cm_forward(bcsr_costmodel *cm, const float *X, int32_t n, float *P, void *stream) {
    if (!cm || n < 0 || n > kMaxRows || (n && (!X || !P))) return BCSR_ERR_INVALID_ARG;
    if (!n) return BCSR_OK;
    int rc = ensure_ws(cm, n);
    if (rc != BCSR_OK) return rc;
    return forward(cm, X, n, P, static_cast<hipStream_t>(stream));
}

// Gradients of the masked MSE at the current parameters (into the model's
// gradient buffer) and the loss (*loss, synchronises the stream).
extern "C" int bcsr_cm_backward(bcsr_costmodel *cm, const float *X, const float *Y, const float *M, int32_t n,
                                float inv_count, float *loss, void *stream) {
    if (!cm || n <= 0 || n > kMaxRows || !X || !Y || !M || !(inv_count > 0.0f)) return BCSR_ERR_INVALID_ARG;
    int rc = ensure_ws(cm, n);
    if (rc != BCSR_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    rc = backward(cm, X, Y, M, n, inv_count, s);
    if (rc != BCSR_OK || !loss) return rc;
    rc = hip_status(hipMemcpyAsync(loss, cm->d_loss, sizeof(float), hipMemcpyDeviceToHost, s));
    if (rc == BCSR_OK) rc = hip_status(hipStreamSynchronize(s));
    return rc;
}

// `steps` full-batch Adam steps on (X, Y, M); *loss = the loss before the
// last step (synchronises the stream when loss is given).
extern "C" int bcsr_cm_train(bcsr_costmodel *cm, const float *X, const float *Y, const float *M, int32_t n,
                             float inv_count, int32_t steps, float lr, float beta1, float beta2, float eps,
                             float *loss, void *stream) {
    if (!cm || n <= 0 || n > kMaxRows || steps < 0 || !X || !Y || !M || !(inv_count > 0.0f) || !(lr > 0.0f))
        return BCSR_ERR_INVALID_ARG;
    int rc = ensure_ws(cm, n);
    if (rc != BCSR_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (int k = 0; k < steps; ++k) {
        rc = backward(cm, X, Y, M, n, inv_count, s);
        if (rc != BCSR_OK) return rc;
        ++cm->step;
        const float c1 = 1.0f / (1.0f - std::pow(beta1, static_cast<float>(cm->step)));
        const float c2 = 1.0f / (1.0f - std::pow(beta2, static_cast<float>(cm->step)));
        hipLaunchKernelGGL(mlp_adam, dim3(blocks_for(cm->n_params)), dim3(256), 0, s, cm->d_params, cm->d_grads,
                           cm->d_m, cm->d_v, cm->n_params, lr, beta1, beta2, eps, c1, c2);
        rc = hip_status(hipGetLastError());
        if (rc != BCSR_OK) return rc;
    }
    if (!loss) return BCSR_OK;
    rc = hip_status(hipMemcpyAsync(loss, cm->d_loss, sizeof(float), hipMemcpyDeviceToHost, s));
    if (rc == BCSR_OK) rc = hip_status(hipStreamSynchronize(s));
    return rc;
}
