// mlp.hip -- training and inference kernels of the learned BFS direction
// policy (FP32 MLP 16 -> 64 (ReLU) -> 2 logits).
//
// No counterpart in the reference (its BFS is top-down only,
// src/traversals/bfs.jl:108-122); the policy replaces the alpha/beta
// direction rule of the device BFS and never changes any output.
//
// Parameter layout (LGMI_POLICY_NWEIGHTS floats, PolicyWeights): W1[64][16],
// b1[64], W2[2][64], b2[2], feature mean mu[16], feature scale sd[16].
// Inputs are standardised as (x - mu) / sd; mu and sd are fixed, the other
// 1218 parameters are trained.
//
// Loss: weighted softmax cross-entropy, L = sum_i w_i CE(z_i, y_i) / sum_i w_i.
// Mapping: one wavefront per sample, lane j = hidden unit j.  Each lane
// keeps the gradient of its own row of W1 and b1 and of column j of W2 in
// registers over the samples of its wavefront; the block then sums its
// wavefronts in LDS and writes a per-block gradient; a second kernel sums
// the blocks in order and applies Adam.  Every sum has a fixed order, so a
// training run is bit-reproducible.
#include "lgmi_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace lgmi {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kTrain = kHidden * kFeat + kHidden + 2 * kHidden + 2;   // 1218 trained parameters
static_assert(kHidden == 64, "one lane per hidden unit");

struct MlpArgs {
    const float* x;        // [n][16]
    const int32_t* y;      // [n]
    const float* w;        // [n] sample weights (nullptr: 1)
    int32_t n;
    const float* p;        // parameters (PolicyWeights layout)
    float inv_wsum;        // 1 / sum of weights
    float* gpart;          // [blocks][kTrain + 1] (last entry: loss)
};

// Forward for one sample held by one wavefront; returns the hidden unit of
// this lane, its pre-activation sign, the standardised input and the logits.
__device__ inline void mlp_forward(const float* p, const float* xrow, int lane, float* xs, float& h, float& z0,
                                   float& z1) {
    const float* W1 = p;
    const float* b1 = p + kHidden * kFeat;
    const float* W2 = b1 + kHidden;
    const float* b2 = W2 + 2 * kHidden;
    const float* mu = b2 + 2;
    const float* sd = mu + kFeat;
    float acc = b1[lane];
    #pragma unroll
    for (int i = 0; i < kFeat; i++) {
        xs[i] = (xrow[i] - mu[i]) / sd[i];
        acc += W1[lane * kFeat + i] * xs[i];
    }
    h = acc > 0.f ? acc : 0.f;
    float a0 = W2[lane] * h, a1 = W2[kHidden + lane] * h;
    // butterfly sums: identical in every lane, fixed order
    for (int o = 32; o >= 1; o >>= 1) {
        a0 += __shfl_xor(a0, o);
        a1 += __shfl_xor(a1, o);
    }
    z0 = a0 + b2[0];
    z1 = a1 + b2[1];
}

__global__ __launch_bounds__(kBlock) void mlp_grad_kernel(MlpArgs a) {
    __shared__ float s_red[kWaves][kFeat + 3];     // per wave, per lane staged below
    __shared__ float s_g[kWaves][kTrain + 1];
    const int tid = threadIdx.x, lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const float* W2 = a.p + kHidden * kFeat + kHidden;
    float gW1[kFeat];
    #pragma unroll
    for (int i = 0; i < kFeat; i++) gW1[i] = 0.f;
    float gb1 = 0.f, gW2a = 0.f, gW2b = 0.f, gb2a = 0.f, gb2b = 0.f, loss = 0.f;
    const int gw = blockIdx.x * kWaves + wid, nw = gridDim.x * kWaves;
    for (int s = gw; s < a.n; s += nw) {
        float xs[kFeat];
        float h, z0, z1;
        mlp_forward(a.p, a.x + (size_t)s * kFeat, lane, xs, h, z0, z1);
        const float m = fmaxf(z0, z1);
        const float e0 = expf(z0 - m), e1 = expf(z1 - m);
        const float den = e0 + e1;
        const float p0 = e0 / den, p1 = e1 / den;
        const int yi = a.y[s];
        const float wi = (a.w ? a.w[s] : 1.f) * a.inv_wsum;
        const float d0 = wi * (p0 - (yi == 0 ? 1.f : 0.f));
        const float d1 = wi * (p1 - (yi == 1 ? 1.f : 0.f));
        if (lane == 0) loss += wi * (logf(den) + m - (yi == 0 ? z0 : z1));
        gW2a += d0 * h;
        gW2b += d1 * h;
        gb2a += d0;
        gb2b += d1;
        const float dh = h > 0.f ? (W2[lane] * d0 + W2[kHidden + lane] * d1) : 0.f;
        gb1 += dh;
        #pragma unroll
        for (int i = 0; i < kFeat; i++) gW1[i] += dh * xs[i];
    }
    (void)s_red;
    // stage this wave's gradient in LDS (layout of the trained parameters)
    float* g = s_g[wid];
    #pragma unroll
    for (int i = 0; i < kFeat; i++) g[lane * kFeat + i] = gW1[i];
    g[kHidden * kFeat + lane] = gb1;
    g[kHidden * kFeat + kHidden + lane] = gW2a;
    g[kHidden * kFeat + kHidden + kHidden + lane] = gW2b;
    if (lane == 0) {
        g[kTrain - 2] = gb2a;
        g[kTrain - 1] = gb2b;
        g[kTrain] = loss;
    }
    __syncthreads();
    // b2 and the loss were kept only by lane 0 for the loss; gb2 is identical
    // in every lane (d0/d1 are wave-uniform), so lane 0's copy is the sum.
    for (int k = tid; k < kTrain + 1; k += kBlock) {
        float t = 0.f;
        for (int q = 0; q < kWaves; q++) t += s_g[q][k];
        a.gpart[(size_t)blockIdx.x * (kTrain + 1) + k] = t;
    }
}

// Sums per-block gradients in block order; optionally applies one Adam step.
__global__ void mlp_reduce_adam(const float* gpart, int32_t nblocks, float* p, float* m, float* v, float* grad_out,
                                float* loss_out, int32_t do_step, float lr, float beta1, float beta2, float eps,
                                float bc1, float bc2) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < kTrain + 1; k += gridDim.x * blockDim.x) {
        float t = 0.f;
        for (int b = 0; b < nblocks; b++) t += gpart[(size_t)b * (kTrain + 1) + k];
        if (k == kTrain) {
            if (loss_out) *loss_out = t;
            continue;
        }
        if (grad_out) grad_out[k] = t;
        if (do_step) {
            const float mk = beta1 * m[k] + (1.f - beta1) * t;
            const float vk = beta2 * v[k] + (1.f - beta2) * t * t;
            m[k] = mk;
            v[k] = vk;
            p[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
        }
    }
}

__global__ __launch_bounds__(kBlock) void mlp_logits_kernel(const float* x, int32_t n, const float* p, float* out) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int s = blockIdx.x * kWaves + wid; s < n; s += gridDim.x * kWaves) {
        float xs[kFeat];
        float h, z0, z1;
        mlp_forward(p, x + (size_t)s * kFeat, lane, xs, h, z0, z1);
        if (lane == 0) { out[2 * (size_t)s] = z0; out[2 * (size_t)s + 1] = z1; }
    }
}

struct MlpData {
    float* x = nullptr;
    int32_t* y = nullptr;
    float* w = nullptr;
    float* p = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    float* gpart = nullptr;
    float* grad = nullptr;
    float* loss = nullptr;
    int blocks = 0;
    hipStream_t s = nullptr;
    ~MlpData() {
        dev_free(x); dev_free(y); dev_free(w); dev_free(p); dev_free(m); dev_free(v); dev_free(gpart);
        dev_free(grad); dev_free(loss);
        if (s) (void)hipStreamDestroy(s);
    }
};

void mlp_upload(MlpData& d, const float* x, const int32_t* y, const float* w, int32_t n, const float* params) {
    ensure_device();
    LGMI_HIP(hipStreamCreateWithFlags(&d.s, hipStreamNonBlocking));
    d.x = dev_alloc_n<float>((size_t)std::max(n, 1) * kFeat);
    d.y = dev_alloc_n<int32_t>((size_t)std::max(n, 1));
    if (w) d.w = dev_alloc_n<float>((size_t)std::max(n, 1));
    d.p = dev_alloc_n<float>(LGMI_POLICY_NWEIGHTS);
    d.m = dev_alloc_n<float>(kTrain);
    d.v = dev_alloc_n<float>(kTrain);
    d.blocks = std::max(1, std::min(256, (n + kWaves * 16 - 1) / (kWaves * 16)));
    d.gpart = dev_alloc_n<float>((size_t)d.blocks * (kTrain + 1));
    d.grad = dev_alloc_n<float>(kTrain);
    d.loss = dev_alloc_n<float>(1);
    if (n > 0) {
        LGMI_HIP(hipMemcpyAsync(d.x, x, sizeof(float) * (size_t)n * kFeat, hipMemcpyHostToDevice, d.s));
        LGMI_HIP(hipMemcpyAsync(d.y, y, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, d.s));
        if (w) LGMI_HIP(hipMemcpyAsync(d.w, w, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, d.s));
    }
    LGMI_HIP(hipMemcpyAsync(d.p, params, sizeof(float) * LGMI_POLICY_NWEIGHTS, hipMemcpyHostToDevice, d.s));
    LGMI_HIP(hipMemsetAsync(d.m, 0, sizeof(float) * kTrain, d.s));
    LGMI_HIP(hipMemsetAsync(d.v, 0, sizeof(float) * kTrain, d.s));
    LGMI_HIP(hipStreamSynchronize(d.s));
}

MlpArgs mlp_args(const MlpData& d, int32_t n, double wsum) {
    MlpArgs a{};
    a.x = d.x; a.y = d.y; a.w = d.w; a.n = n; a.p = d.p;
    a.inv_wsum = wsum > 0 ? (float)(1.0 / wsum) : 0.f;
    a.gpart = d.gpart;
    return a;
}

void check_params(const float* params) {
    const float* sd = params + kTrain + kFeat;
    for (int i = 0; i < kFeat; i++)
        if (!(sd[i] > 0.f)) throw Error(LGMI_E_ARG, "feature scales must be positive");
}

double weight_sum(const float* w, const int32_t* y, int32_t n) {
    double s = 0.0;
    for (int32_t i = 0; i < n; i++) {
        if (y[i] != 0 && y[i] != 1) throw Error(LGMI_E_ARG, "labels must be 0 or 1");
        const double wi = w ? (double)w[i] : 1.0;
        if (!(wi >= 0.0)) throw Error(LGMI_E_ARG, "sample weights must be >= 0");
        s += wi;
    }
    return s;
}

}  // namespace
}  // namespace lgmi

using namespace lgmi;

#define API_TRY try {
#define API_CATCH                                                  \
    }                                                              \
    catch (const Error& e) {                                       \
        set_last_error(e.what());                                  \
        return e.code;                                             \
    }                                                              \
    catch (const std::exception& e) {                              \
        set_last_error(e.what());                                  \
        return LGMI_E_HIP;                                         \
    }                                                              \
    return LGMI_OK;

extern "C" int lgmi_mlp_loss_grad(const float* x, const int32_t* y, const float* w, int32_t n, const float* params,
                                  float* loss, float* grad) {
    API_TRY
    if (n < 0 || (n > 0 && (!x || !y)) || !params) throw Error(LGMI_E_ARG, "bad arguments");
    check_params(params);
    const double wsum = weight_sum(w, y, n);
    MlpData d;
    mlp_upload(d, x, y, w, n, params);
    MlpArgs a = mlp_args(d, n, wsum);
    mlp_grad_kernel<<<d.blocks, kBlock, 0, d.s>>>(a);
    mlp_reduce_adam<<<8, 256, 0, d.s>>>(d.gpart, d.blocks, d.p, d.m, d.v, d.grad, d.loss, 0, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f);
    LGMI_HIP(hipGetLastError());
    if (loss) LGMI_HIP(hipMemcpyAsync(loss, d.loss, sizeof(float), hipMemcpyDeviceToHost, d.s));
    if (grad) {
        LGMI_HIP(hipMemcpyAsync(grad, d.grad, sizeof(float) * kTrain, hipMemcpyDeviceToHost, d.s));
        std::fill(grad + kTrain, grad + LGMI_POLICY_NWEIGHTS, 0.f);
    }
    LGMI_HIP(hipStreamSynchronize(d.s));
    API_CATCH
}

extern "C" int lgmi_mlp_train(const float* x, const int32_t* y, const float* w, int32_t n, float* params,
                              int32_t steps, float lr, float* loss_out) {
    API_TRY
    if (n <= 0 || !x || !y || !params || steps < 0 || !(lr > 0.f)) throw Error(LGMI_E_ARG, "bad arguments");
    check_params(params);
    const double wsum = weight_sum(w, y, n);
    MlpData d;
    mlp_upload(d, x, y, w, n, params);
    MlpArgs a = mlp_args(d, n, wsum);
    const float b1 = 0.9f, b2 = 0.999f;
    double p1 = 1.0, p2 = 1.0;
    for (int32_t t = 0; t < steps; t++) {
        p1 *= b1;
        p2 *= b2;
        mlp_grad_kernel<<<d.blocks, kBlock, 0, d.s>>>(a);
        mlp_reduce_adam<<<8, 256, 0, d.s>>>(d.gpart, d.blocks, d.p, d.m, d.v, nullptr, d.loss, 1, lr, b1, b2, 1e-8f,
                                            (float)(1.0 - p1), (float)(1.0 - p2));
    }
    LGMI_HIP(hipGetLastError());
    // final loss at the trained parameters
    mlp_grad_kernel<<<d.blocks, kBlock, 0, d.s>>>(a);
    mlp_reduce_adam<<<8, 256, 0, d.s>>>(d.gpart, d.blocks, d.p, d.m, d.v, nullptr, d.loss, 0, 0.f, 0.f, 0.f, 0.f, 1.f,
                                        1.f);
    LGMI_HIP(hipMemcpyAsync(params, d.p, sizeof(float) * LGMI_POLICY_NWEIGHTS, hipMemcpyDeviceToHost, d.s));
    if (loss_out) LGMI_HIP(hipMemcpyAsync(loss_out, d.loss, sizeof(float), hipMemcpyDeviceToHost, d.s));
    LGMI_HIP(hipStreamSynchronize(d.s));
    API_CATCH
}

extern "C" int lgmi_mlp_forward(const float* x, int32_t n, const float* params, float* logits) {
    API_TRY
    if (n < 0 || (n > 0 && (!x || !logits)) || !params) throw Error(LGMI_E_ARG, "bad arguments");
    check_params(params);
    if (n == 0) return LGMI_OK;
    std::vector<int32_t> y((size_t)n, 0);
    MlpData d;
    mlp_upload(d, x, y.data(), nullptr, n, params);
    float* out = dev_alloc_n<float>((size_t)n * 2);
    try {
        const int blocks = std::max(1, std::min(1024, (n + kWaves - 1) / kWaves));
        mlp_logits_kernel<<<blocks, kBlock, 0, d.s>>>(d.x, n, d.p, out);
        LGMI_HIP(hipGetLastError());
        LGMI_HIP(hipMemcpyAsync(logits, out, sizeof(float) * (size_t)n * 2, hipMemcpyDeviceToHost, d.s));
        LGMI_HIP(hipStreamSynchronize(d.s));
    } catch (...) {
        dev_free(out);
        throw;
    }
    dev_free(out);
    API_CATCH
}
