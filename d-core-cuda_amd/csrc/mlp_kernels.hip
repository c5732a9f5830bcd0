// Cost-model MLP for the variant scheduler: 8 inputs, two hidden layers of
// 32 ReLU units, one linear output, fp32.  Hand-written forward, backward
// and Adam kernels (no library): the network is tiny, so one thread owns
// one sample end to end and gradients are summed per workgroup in LDS
// before one global atomic per parameter and workgroup.
//
// Parameter layout (kMlpParams floats): W1[32][8], b1[32], W2[32][32],
// b2[32], W3[32], b3.
#include <hip/hip_runtime.h>

#include "dcx_internal.hpp"

namespace dcx {
namespace {

constexpr int kIn = kMlpIn, kH = kMlpHidden;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kH;
static_assert(oB3 + 1 == kMlpParams, "parameter layout");

struct Act {
    float h1[kH], h2[kH];
};

__device__ float forward(const float* __restrict__ p, const float* __restrict__ x, Act& a) {
#pragma unroll 4
    for (int j = 0; j < kH; ++j) {
        float s = p[oB1 + j];
#pragma unroll
        for (int i = 0; i < kIn; ++i) s += p[oW1 + j * kIn + i] * x[i];
        a.h1[j] = s > 0.f ? s : 0.f;
    }
#pragma unroll 4
    for (int j = 0; j < kH; ++j) {
        float s = p[oB2 + j];
        for (int i = 0; i < kH; ++i) s += p[oW2 + j * kH + i] * a.h1[i];
        a.h2[j] = s > 0.f ? s : 0.f;
    }
    float o = p[oB3];
    for (int i = 0; i < kH; ++i) o += p[oW3 + i] * a.h2[i];
    return o;
}

__global__ void k_mlp_forward(const float* p, const float* X, uint32_t n, float* y) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        float x[kIn];
        for (int i = 0; i < kIn; ++i) x[i] = X[(size_t)s * kIn + i];
        Act a;
        y[s] = forward(p, x, a);
    }
}

// grad += d/dp of 0.5 * (f(x) - t)^2 summed over the samples, loss += the
// same sum.  The caller divides by n when it applies the step.
__global__ __launch_bounds__(256) void k_mlp_backward(const float* p, const float* X,
                                                      const float* T, uint32_t n, float* grad,
                                                      float* loss) {
    __shared__ float g[kMlpParams];
    __shared__ float l_sum;
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) g[i] = 0.f;
    if (threadIdx.x == 0) l_sum = 0.f;
    __syncthreads();
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        float x[kIn];
        for (int i = 0; i < kIn; ++i) x[i] = X[(size_t)s * kIn + i];
        Act a;
        const float e = forward(p, x, a) - T[s];
        atomicAdd(&l_sum, 0.5f * e * e);
        // output layer
        atomicAdd(&g[oB3], e);
        float d2[kH];
        for (int j = 0; j < kH; ++j) {
            atomicAdd(&g[oW3 + j], e * a.h2[j]);
            d2[j] = a.h2[j] > 0.f ? e * p[oW3 + j] : 0.f;
        }
        // second hidden layer
        float d1[kH];
        for (int i = 0; i < kH; ++i) d1[i] = 0.f;
        for (int j = 0; j < kH; ++j) {
            if (d2[j] == 0.f) continue;
            atomicAdd(&g[oB2 + j], d2[j]);
            for (int i = 0; i < kH; ++i) {
                atomicAdd(&g[oW2 + j * kH + i], d2[j] * a.h1[i]);
                d1[i] += d2[j] * p[oW2 + j * kH + i];
            }
        }
        // first hidden layer
        for (int i = 0; i < kH; ++i) {
            const float d = a.h1[i] > 0.f ? d1[i] : 0.f;
            if (d == 0.f) continue;
            atomicAdd(&g[oB1 + i], d);
            for (int q = 0; q < kIn; ++q) atomicAdd(&g[oW1 + i * kIn + q], d * x[q]);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x)
        if (g[i] != 0.f) atomicAdd(&grad[i], g[i]);
    if (threadIdx.x == 0) atomicAdd(loss, l_sum);
}

// One Adam step on the mean gradient (grad / n); step counter t >= 1.
__global__ void k_mlp_adam(float* p, const float* grad, float* m, float* v, uint32_t n, float lr,
                           float b1, float b2, float eps, int t) {
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kMlpParams;
         i += gridDim.x * blockDim.x) {
        const float gi = grad[i] / (float)n;
        m[i] = b1 * m[i] + (1.f - b1) * gi;
        v[i] = b2 * v[i] + (1.f - b2) * gi * gi;
        p[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
    }
}

unsigned blocks_for(uint32_t n) {
    unsigned b = (n + 255) / 256;
    return b ? (b < 1024 ? b : 1024) : 1;
}

}  // namespace

hipError_t launch_mlp_forward(const float* p, const float* X, uint32_t n, float* y,
                              hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks_for(n)), dim3(256), 0, st, p, X, n, y);
    return hipGetLastError();
}

hipError_t launch_mlp_backward(const float* p, const float* X, const float* T, uint32_t n,
                               float* grad, float* loss, hipStream_t st) {
    hipError_t e = hipMemsetAsync(grad, 0, kMlpParams * sizeof(float), st);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(loss, 0, sizeof(float), st);
    if (e != hipSuccess || !n) return e;
    hipLaunchKernelGGL(k_mlp_backward, dim3(blocks_for(n)), dim3(256), 0, st, p, X, T, n, grad,
                       loss);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float* p, const float* grad, float* m, float* v, uint32_t n, float lr,
                           int t, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_adam, dim3(6), dim3(256), 0, st, p, grad, m, v, n ? n : 1u, lr, 0.9f,
                       0.999f, 1e-8f, t);
    return hipGetLastError();
}

}  // namespace dcx
