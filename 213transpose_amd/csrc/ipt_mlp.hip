// ipt_mlp.hip — training and inference kernels of the selector's cost model
// (MLP IPT_MLP_IN -> 64 -> 64 -> IPT_NUM_CANDIDATES, ReLU, fp32).
//
// Loss: masked mean squared error of predicted vs measured log2 times,
//   L = (1 / sum(M)) * sum_{r,k} M[r,k] * (Yhat[r,k] - Y[r,k])^2
// Backward is written out by hand (no autodiff); the gradients are checked
// against a float64 CPU computation in the tests.  Every sum runs in a fixed
// order, so results are deterministic run to run.
#include "ipt_internal.h"
#include "ipt_select.h"

#include <cmath>
#include <vector>

namespace ipt {

// Y[r][o] = act(b[o] + sum_i W[o][i] * X[r][i]), act = ReLU or identity.
__global__ void __launch_bounds__(256)
k_mlp_linear_fwd(const float* __restrict__ X, int in, const float* __restrict__ W,
                 const float* __restrict__ b, float* __restrict__ Y, int out, int64_t n, int relu) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * out) return;
    const int64_t r = idx / out;
    const int o = (int)(idx - r * out);
    float a = b[o];
    for (int i = 0; i < in; i++) a = fmaf(W[o * in + i], X[r * in + i], a);
    Y[idx] = relu ? (a > 0.0f ? a : 0.0f) : a;
}

// dY = 2 * scale * M * (Yhat - Y); sq = (M * (Yhat - Y))^2.
__global__ void __launch_bounds__(256)
k_mlp_mse_grad(const float* __restrict__ Yh, const float* __restrict__ Y,
               const float* __restrict__ M, int64_t total, float scale, float* __restrict__ dY,
               float* __restrict__ sq) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const float d = (Yh[idx] - Y[idx]) * M[idx];
    dY[idx] = 2.0f * scale * d;
    sq[idx] = d * d;
}

// out[0] = scale * sum(v) (one workgroup, fixed order).
__global__ void __launch_bounds__(256)
k_mlp_sum(const float* __restrict__ v, int64_t n, float scale, float* __restrict__ out) {
    __shared__ double part[256];
    double a = 0;
    for (int64_t i = threadIdx.x; i < n; i += 256) a += v[i];
    part[threadIdx.x] = a;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = (float)(part[0] * scale);
}

// dW[o][i] = sum_r dY[r][o] * X[r][i]; db[o] = sum_r dY[r][o].  One
// workgroup per weight (blockIdx.x = o * (in + 1) + i, i == in for the bias):
// each thread sums the rows r = tid, tid + 256, ... and the workgroup adds
// the 256 partial sums as a fixed tree, so the result is the same run to run.
constexpr int kBwdBlock = 256;
__global__ void __launch_bounds__(kBwdBlock)
k_mlp_linear_bwd_w(const float* __restrict__ dY, const float* __restrict__ X, float* __restrict__ dW,
                   float* __restrict__ db, int64_t n, int in, int out) {
    __shared__ float part[kBwdBlock];
    const int o = blockIdx.x / (in + 1), i = blockIdx.x % (in + 1);
    float a = 0.0f;
    if (i == in) {
        for (int64_t r = threadIdx.x; r < n; r += kBwdBlock) a += dY[r * out + o];
    } else {
        for (int64_t r = threadIdx.x; r < n; r += kBwdBlock) a = fmaf(dY[r * out + o], X[r * in + i], a);
    }
    part[threadIdx.x] = a;
    __syncthreads();
    for (int s = kBwdBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (i == in) db[o] = part[0];
        else dW[o * in + i] = part[0];
    }
}

// dX[r][i] = relu'(H[r][i]) * sum_o dY[r][o] * W[o][i].
__global__ void __launch_bounds__(256)
k_mlp_linear_bwd_x(const float* __restrict__ dY, const float* __restrict__ W,
                   const float* __restrict__ H, float* __restrict__ dX, int64_t n, int in, int out) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * in) return;
    const int64_t r = idx / in;
    const int i = (int)(idx - r * in);
    float a = 0.0f;
    for (int o = 0; o < out; o++) a = fmaf(dY[r * out + o], W[o * in + i], a);
    dX[idx] = H[idx] > 0.0f ? a : 0.0f;
}

// Adam step t (1-based) with bias correction.
__global__ void __launch_bounds__(256)
k_mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
           float* __restrict__ v, int np, float lr, float b1, float b2, float eps, float c1,
           float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

namespace {

inline int blocks(int64_t n) { return (int)((n + 255) / 256); }

struct MlpBuffers {
    float *x = nullptr, *y = nullptr, *mask = nullptr, *p = nullptr, *g = nullptr, *m = nullptr,
          *v = nullptr, *h1 = nullptr, *h2 = nullptr, *yh = nullptr, *dy = nullptr,
          *dh2 = nullptr, *dh1 = nullptr, *sq = nullptr, *loss = nullptr;
    std::vector<float*> all;
    hipError_t alloc(float** ptr, size_t count) {
        hipError_t e = hipMalloc(ptr, std::max<size_t>(count, 1) * sizeof(float));
        if (e == hipSuccess) all.push_back(*ptr);
        return e;
    }
    ~MlpBuffers() {
        for (float* q : all) (void)hipFree(q);
    }
};

// Parameter views.
struct Views {
    const float *W1, *b1, *W2, *b2, *W3, *b3;
    float *gW1, *gb1, *gW2, *gb2, *gW3, *gb3;
    Views(float* p, float* g)
        : W1(p), b1(p + kW1), W2(p + kW1 + kB1), b2(p + kW1 + kB1 + kW2),
          W3(p + kW1 + kB1 + kW2 + kB2), b3(p + kW1 + kB1 + kW2 + kB2 + kW3), gW1(g),
          gb1(g + kW1), gW2(g + kW1 + kB1), gb2(g + kW1 + kB1 + kW2),
          gW3(g + kW1 + kB1 + kW2 + kB2), gb3(g + kW1 + kB1 + kW2 + kB2 + kW3) {}
};

void forward(const MlpBuffers& B, const Views& P, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_linear_fwd, dim3(blocks(n * kHid)), dim3(256), 0, s, B.x, kFeat, P.W1,
                       P.b1, B.h1, kHid, n, 1);
    hipLaunchKernelGGL(k_mlp_linear_fwd, dim3(blocks(n * kHid)), dim3(256), 0, s, B.h1, kHid, P.W2,
                       P.b2, B.h2, kHid, n, 1);
    hipLaunchKernelGGL(k_mlp_linear_fwd, dim3(blocks(n * kOut)), dim3(256), 0, s, B.h2, kHid, P.W3,
                       P.b3, B.yh, kOut, n, 0);
}

void backward(const MlpBuffers& B, const Views& P, int64_t n, float scale, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_mse_grad, dim3(blocks(n * kOut)), dim3(256), 0, s, B.yh, B.y, B.mask,
                       n * kOut, scale, B.dy, B.sq);
    hipLaunchKernelGGL(k_mlp_linear_bwd_w, dim3(kOut * (kHid + 1)), dim3(kBwdBlock), 0, s, B.dy,
                       B.h2, P.gW3, P.gb3, n, kHid, kOut);
    hipLaunchKernelGGL(k_mlp_linear_bwd_x, dim3(blocks(n * kHid)), dim3(256), 0, s, B.dy, P.W3,
                       B.h2, B.dh2, n, kHid, kOut);
    hipLaunchKernelGGL(k_mlp_linear_bwd_w, dim3(kHid * (kHid + 1)), dim3(kBwdBlock), 0, s, B.dh2,
                       B.h1, P.gW2, P.gb2, n, kHid, kHid);
    hipLaunchKernelGGL(k_mlp_linear_bwd_x, dim3(blocks(n * kHid)), dim3(256), 0, s, B.dh2, P.W2,
                       B.h1, B.dh1, n, kHid, kHid);
    hipLaunchKernelGGL(k_mlp_linear_bwd_w, dim3(kHid * (kFeat + 1)), dim3(kBwdBlock), 0, s,
                       B.dh1, B.x, P.gW1, P.gb1, n, kFeat, kHid);
}

// Allocate and upload; `training` adds gradient / optimiser / backward state.
ipt_status setup(MlpBuffers& B, const float* x, const float* y, const float* mask, int64_t n,
                 const float* params, bool training, float* mask_sum) {
    hipError_t e = B.alloc(&B.x, n * kFeat);
    if (e == hipSuccess) e = B.alloc(&B.p, kParams);
    if (e == hipSuccess) e = B.alloc(&B.h1, n * kHid);
    if (e == hipSuccess) e = B.alloc(&B.h2, n * kHid);
    if (e == hipSuccess) e = B.alloc(&B.yh, n * kOut);
    if (training) {
        if (e == hipSuccess) e = B.alloc(&B.y, n * kOut);
        if (e == hipSuccess) e = B.alloc(&B.mask, n * kOut);
        if (e == hipSuccess) e = B.alloc(&B.g, kParams);
        if (e == hipSuccess) e = B.alloc(&B.m, kParams);
        if (e == hipSuccess) e = B.alloc(&B.v, kParams);
        if (e == hipSuccess) e = B.alloc(&B.dy, n * kOut);
        if (e == hipSuccess) e = B.alloc(&B.dh2, n * kHid);
        if (e == hipSuccess) e = B.alloc(&B.dh1, n * kHid);
        if (e == hipSuccess) e = B.alloc(&B.sq, n * kOut);
        if (e == hipSuccess) e = B.alloc(&B.loss, 1);
    }
    if (e == hipSuccess) e = hipMemcpy(B.x, x, n * kFeat * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B.p, params, kParams * sizeof(float), hipMemcpyHostToDevice);
    if (training) {
        if (e == hipSuccess) e = hipMemcpy(B.y, y, n * kOut * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(B.mask, mask, n * kOut * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(B.m, 0, kParams * sizeof(float));
        if (e == hipSuccess) e = hipMemset(B.v, 0, kParams * sizeof(float));
        double ms = 0;
        for (int64_t i = 0; i < n * kOut; i++) ms += mask[i];
        *mask_sum = (float)ms;
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("selector buffers", e);
        return e == hipErrorOutOfMemory ? IPT_ERR_OUT_OF_MEMORY : IPT_ERR_HIP;
    }
    return IPT_OK;
}

bool finite_all(const float* v, int64_t n) {
    for (int64_t i = 0; i < n; i++)
        if (!std::isfinite(v[i])) return false;
    return true;
}

}  // namespace
}  // namespace ipt

using namespace ipt;

extern "C" {

ipt_status ipt_selector_train(const float* x, const float* y, const float* mask, int64_t n,
                              int epochs, float lr, float* params, float* loss_out) {
    if (!x || !y || !mask || !params || n <= 0 || epochs < 0 || !(lr > 0.0f))
        return IPT_ERR_INVALID_ARG;
    if (!finite_all(x, n * kFeat) || !finite_all(params, kParams)) return IPT_ERR_INVALID_ARG;
    DeviceCtx* ctx;
    ipt_status st = get_ctx(&ctx);
    if (st != IPT_OK) return st;
    MlpBuffers B;
    float msum = 0;
    if ((st = setup(B, x, y, mask, n, params, true, &msum)) != IPT_OK) return st;
    if (!(msum > 0)) return IPT_ERR_INVALID_ARG;
    const float scale = 1.0f / msum;
    Views P(B.p, B.g);
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.0f, c2 = 1.0f;
    for (int t = 1; t <= epochs; t++) {
        forward(B, P, n, nullptr);
        backward(B, P, n, scale, nullptr);
        c1 *= b1;
        c2 *= b2;
        hipLaunchKernelGGL(k_mlp_adam, dim3(blocks(kParams)), dim3(256), 0, nullptr, B.p, B.g, B.m,
                           B.v, kParams, lr, b1, b2, 1e-8f, 1.0f - c1, 1.0f - c2);
    }
    forward(B, P, n, nullptr);
    backward(B, P, n, scale, nullptr);
    hipLaunchKernelGGL(k_mlp_sum, dim3(1), dim3(256), 0, nullptr, B.sq, n * kOut, scale, B.loss);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(params, B.p, kParams * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && loss_out) e = hipMemcpy(loss_out, B.loss, sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        set_error("ipt_selector_train", e);
        return IPT_ERR_HIP;
    }
    return IPT_OK;
}

ipt_status ipt_selector_grad(const float* x, const float* y, const float* mask, int64_t n,
                             const float* params, float* loss_out, float* grad_out) {
    if (!x || !y || !mask || !params || !grad_out || n <= 0) return IPT_ERR_INVALID_ARG;
    DeviceCtx* ctx;
    ipt_status st = get_ctx(&ctx);
    if (st != IPT_OK) return st;
    MlpBuffers B;
    float msum = 0;
    if ((st = setup(B, x, y, mask, n, params, true, &msum)) != IPT_OK) return st;
    if (!(msum > 0)) return IPT_ERR_INVALID_ARG;
    Views P(B.p, B.g);
    forward(B, P, n, nullptr);
    backward(B, P, n, 1.0f / msum, nullptr);
    hipLaunchKernelGGL(k_mlp_sum, dim3(1), dim3(256), 0, nullptr, B.sq, n * kOut, 1.0f / msum, B.loss);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(grad_out, B.g, kParams * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && loss_out) e = hipMemcpy(loss_out, B.loss, sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        set_error("ipt_selector_grad", e);
        return IPT_ERR_HIP;
    }
    return IPT_OK;
}

ipt_status ipt_selector_infer(const float* params, const float* x, int64_t n, float* y) {
    if (!params || !x || !y || n <= 0) return IPT_ERR_INVALID_ARG;
    DeviceCtx* ctx;
    ipt_status st = get_ctx(&ctx);
    if (st != IPT_OK) return st;
    MlpBuffers B;
    if ((st = setup(B, x, nullptr, nullptr, n, params, false, nullptr)) != IPT_OK) return st;
    Views P(B.p, B.p);
    forward(B, P, n, nullptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(y, B.yh, n * kOut * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        set_error("ipt_selector_infer", e);
        return IPT_ERR_HIP;
    }
    return IPT_OK;
}

}  // extern "C"
