// Kernels of the learned tile selector: a 3-layer MLP (4 -> 64 ReLU -> 64 ReLU
// -> nout) over FP32 weights laid out W1[64][4], b1[64], W2[64][64], b2[64],
// W3[nout][64], b3[nout] in one buffer.
//
// forward : one 64-thread block per sample (thread j owns hidden unit j).
// backward: per-sample kernel for the activation gradients, then one thread per
//           parameter summing over the batch in sample order (deterministic).
// adam    : elementwise.
// infer   : one wave, one sample, arg-min over the valid candidates.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mmsk {

namespace {
constexpr int H = kMlpHidden;
constexpr int IN = kMlpIn;
constexpr int kMaxOut = 64;

struct Offsets {
    int w1, b1, w2, b2, w3, b3, total;
};
__host__ __device__ constexpr Offsets offsets(int nout) {
    return Offsets{0, H * IN, H * IN + H, H * IN + H + H * H, H * IN + 2 * H + H * H,
                   H * IN + 2 * H + H * H + nout * H, H * IN + 2 * H + H * H + nout * H + nout};
}
}  // namespace

int mlp_num_params(int nout) { return offsets(nout).total; }

// h1/h2 of one sample into shared memory; returns via y (nout values) when y != nullptr.
__device__ __forceinline__ void forward_one(const float* __restrict__ P, const float* __restrict__ x, int nout,
                                            float* s_h1, float* s_h2, float* h1_out, float* h2_out,
                                            float* y_out) {
    const Offsets o = offsets(nout);
    const int j = threadIdx.x;
    float a = P[o.b1 + j];
#pragma unroll
    for (int i = 0; i < IN; ++i) a += P[o.w1 + j * IN + i] * x[i];
    a = a > 0.f ? a : 0.f;
    s_h1[j] = a;
    if (h1_out) h1_out[j] = a;
    __syncthreads();
    float b = P[o.b2 + j];
    for (int k = 0; k < H; ++k) b += P[o.w2 + j * H + k] * s_h1[k];
    b = b > 0.f ? b : 0.f;
    s_h2[j] = b;
    if (h2_out) h2_out[j] = b;
    __syncthreads();
    if (j < nout && y_out) {
        float y = P[o.b3 + j];
        for (int k = 0; k < H; ++k) y += P[o.w3 + j * H + k] * s_h2[k];
        y_out[j] = y;
    }
}

__global__ __launch_bounds__(64) void mlp_forward_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                         int nout, float* __restrict__ H1, float* __restrict__ H2,
                                                         float* __restrict__ Y) {
    __shared__ float s_h1[H], s_h2[H];
    const int b = blockIdx.x;
    forward_one(P, X + (size_t)b * IN, nout, s_h1, s_h2, H1 ? H1 + (size_t)b * H : nullptr,
                H2 ? H2 + (size_t)b * H : nullptr, Y + (size_t)b * nout);
}

// Masked mean-squared error: loss = sum(m (y - t)^2) / sum(m); also the mask count.
__global__ __launch_bounds__(256) void mlp_loss_kernel(const float* __restrict__ Y, const float* __restrict__ T,
                                                       const float* __restrict__ Mk, int n, float* __restrict__ out) {
    __shared__ float s_l[256], s_c[256];
    float l = 0.f, c = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) {
        const float d = Y[i] - T[i];
        l += Mk[i] * d * d;
        c += Mk[i];
    }
    s_l[threadIdx.x] = l;
    s_c[threadIdx.x] = c;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_l[threadIdx.x] += s_l[threadIdx.x + s];
            s_c[threadIdx.x] += s_c[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s_c[0] > 0.f ? s_l[0] / s_c[0] : 0.f;
        out[1] = s_c[0];
    }
}

// dH2 and dH1 of one sample (pre-activation gradients), from dY = 2 m (y - t) / count.
__global__ __launch_bounds__(64) void mlp_backprop_kernel(const float* __restrict__ P, int nout,
                                                          const float* __restrict__ H1, const float* __restrict__ H2,
                                                          const float* __restrict__ Y, const float* __restrict__ T,
                                                          const float* __restrict__ Mk, const float* __restrict__ stats,
                                                          float* __restrict__ dZ2, float* __restrict__ dZ1) {
    __shared__ float s_dy[kMaxOut], s_dz2[H];
    const Offsets o = offsets(nout);
    const int b = blockIdx.x, j = threadIdx.x;
    const float inv = stats[1] > 0.f ? 2.f / stats[1] : 0.f;
    if (j < nout) {
        const size_t i = (size_t)b * nout + j;
        s_dy[j] = inv * Mk[i] * (Y[i] - T[i]);
    }
    __syncthreads();
    float g = 0.f;
    for (int q = 0; q < nout; ++q) g += s_dy[q] * P[o.w3 + q * H + j];
    g = H2[(size_t)b * H + j] > 0.f ? g : 0.f;
    s_dz2[j] = g;
    dZ2[(size_t)b * H + j] = g;
    __syncthreads();
    float g1 = 0.f;
    for (int q = 0; q < H; ++q) g1 += s_dz2[q] * P[o.w2 + q * H + j];
    g1 = H1[(size_t)b * H + j] > 0.f ? g1 : 0.f;
    dZ1[(size_t)b * H + j] = g1;
}

// One thread per parameter: sum over the batch in sample order.
__global__ __launch_bounds__(256) void mlp_param_grad_kernel(const float* __restrict__ X, int batch, int nout,
                                                             const float* __restrict__ H1, const float* __restrict__ H2,
                                                             const float* __restrict__ Y, const float* __restrict__ T,
                                                             const float* __restrict__ Mk,
                                                             const float* __restrict__ stats,
                                                             const float* __restrict__ dZ2,
                                                             const float* __restrict__ dZ1, float* __restrict__ G) {
    const Offsets o = offsets(nout);
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= o.total) return;
    const float inv = stats[1] > 0.f ? 2.f / stats[1] : 0.f;
    float s = 0.f;
    if (p < o.b1) {  // W1[j][i]
        const int j = p / IN, i = p % IN;
        for (int b = 0; b < batch; ++b) s += dZ1[(size_t)b * H + j] * X[(size_t)b * IN + i];
    } else if (p < o.w2) {  // b1[j]
        const int j = p - o.b1;
        for (int b = 0; b < batch; ++b) s += dZ1[(size_t)b * H + j];
    } else if (p < o.b2) {  // W2[j][k]
        const int j = (p - o.w2) / H, k = (p - o.w2) % H;
        for (int b = 0; b < batch; ++b) s += dZ2[(size_t)b * H + j] * H1[(size_t)b * H + k];
    } else if (p < o.w3) {  // b2[j]
        const int j = p - o.b2;
        for (int b = 0; b < batch; ++b) s += dZ2[(size_t)b * H + j];
    } else if (p < o.b3) {  // W3[q][k]
        const int q = (p - o.w3) / H, k = (p - o.w3) % H;
        for (int b = 0; b < batch; ++b) {
            const size_t i = (size_t)b * nout + q;
            s += inv * Mk[i] * (Y[i] - T[i]) * H2[(size_t)b * H + k];
        }
    } else {  // b3[q]
        const int q = p - o.b3;
        for (int b = 0; b < batch; ++b) {
            const size_t i = (size_t)b * nout + q;
            s += inv * Mk[i] * (Y[i] - T[i]);
        }
    }
    G[p] = s;
}

__global__ __launch_bounds__(256) void mlp_adam_kernel(float* __restrict__ P, const float* __restrict__ G,
                                                       float* __restrict__ Mm, float* __restrict__ V, int n,
                                                       float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    const float m = b1 * Mm[i] + (1.f - b1) * g;
    const float v = b2 * V[i] + (1.f - b2) * g * g;
    Mm[i] = m;
    V[i] = v;
    const float mh = m / c1, vh = v / c2;
    P[i] -= lr * mh / (sqrtf(vh) + eps);
}

// One sample: predictions into y[nout].
__global__ __launch_bounds__(64) void mlp_infer_kernel(const float* __restrict__ P, const float* __restrict__ x,
                                                       int nout, float* __restrict__ y) {
    __shared__ float s_h1[H], s_h2[H];
    forward_one(P, x, nout, s_h1, s_h2, nullptr, nullptr, y);
}

hipError_t launch_mlp_forward(const float* params, const float* X, int batch, int nout, float* H1, float* H2,
                              float* Y, hipStream_t stream) {
    if (batch <= 0 || nout <= 0 || nout > kMaxOut) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(batch), dim3(64), 0, stream, params, X, nout, H1, H2, Y);
    return hipGetLastError();
}

hipError_t launch_mlp_backward(const float* params, const float* X, int batch, int nout, const float* H1,
                               const float* H2, const float* Y, const float* T, const float* Mask,
                               float* dparams, float* loss, hipStream_t stream, float* scratch) {
    // scratch: 2 (stats) + 2 * batch * H floats
    if (batch <= 0 || nout <= 0 || nout > kMaxOut) return hipErrorInvalidValue;
    float* stats = scratch;
    float* dZ2 = scratch + 64;
    float* dZ1 = dZ2 + (size_t)batch * H;
    hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(256), 0, stream, Y, T, Mask, batch * nout, stats);
    hipLaunchKernelGGL(mlp_backprop_kernel, dim3(batch), dim3(64), 0, stream, params, nout, H1, H2, Y, T, Mask,
                       stats, dZ2, dZ1);
    const int np = offsets(nout).total;
    hipLaunchKernelGGL(mlp_param_grad_kernel, dim3((np + 255) / 256), dim3(256), 0, stream, X, batch, nout, H1, H2,
                       Y, T, Mask, stats, dZ2, dZ1, dparams);
    if (loss) {
        hipError_t e = hipMemcpyAsync(loss, stats, sizeof(float), hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float* params, const float* g, float* m, float* v, int n, int step, float lr, float b1,
                           float b2, float eps, hipStream_t stream) {
    if (n <= 0 || step < 1) return hipErrorInvalidValue;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, params, g, m, v, n, lr, b1, b2,
                       eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const float* params, const float* x, int nout, float* y, hipStream_t stream) {
    if (nout <= 0 || nout > kMaxOut) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(1), dim3(64), 0, stream, params, x, nout, y);
    return hipGetLastError();
}

}  // namespace mmsk
