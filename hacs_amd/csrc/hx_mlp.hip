// hx_mlp.hip -- forward / backward / Adam for the kernel selector's 2-layer MLP (FP32).
//
//   hidden = relu(X W1^T + b1)        X: B x F, W1: H x F, b1: H
//   y      = hidden W2^T + b2         W2: O x H, b2: O, y: B x O
//   loss   = sum mask (y - t)^2 / sum mask            (mask == NULL: all ones)
// The network only ranks kernel variants / tiles; it never touches a value of the exact path, so a
// misprediction can cost time but cannot change a result.  Every reduction runs in a fixed order
// (one thread per output, serial over the batch), so gradients are deterministic run to run.
#include "hx_common.h"

namespace hx {
namespace {

__global__ void mlp_hidden_kernel(const float* X, uint32_t B, uint32_t F, const float* W1, const float* b1, uint32_t H,
                                  float* hidden) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * H) return;
    const uint32_t b = idx / H, h = idx % H;
    float s = b1[h];
    for (uint32_t f = 0; f < F; ++f) s += X[b * F + f] * W1[h * F + f];
    hidden[idx] = s > 0.f ? s : 0.f;
}

__global__ void mlp_out_kernel(const float* hidden, uint32_t B, uint32_t H, const float* W2, const float* b2,
                               uint32_t O, float* y) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * O) return;
    const uint32_t b = idx / O, o = idx % O;
    float s = b2[o];
    for (uint32_t h = 0; h < H; ++h) s += hidden[b * H + h] * W2[o * H + h];
    y[idx] = s;
}

// dy = 2 mask (y - t) / count, and the loss; one block of 256 threads, fixed-order tree reduction.
__global__ void mlp_dy_kernel(const float* y, const float* t, const float* mask, uint32_t N, float* dy, float* loss) {
    __shared__ float s_l[256], s_c[256];
    float l = 0.f, c = 0.f;
    for (uint32_t i = threadIdx.x; i < N; i += 256) {
        const float mk = mask ? mask[i] : 1.f;
        const float d = y[i] - t[i];
        l += mk * d * d;
        c += mk;
    }
    s_l[threadIdx.x] = l;
    s_c[threadIdx.x] = c;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            s_l[threadIdx.x] += s_l[threadIdx.x + st];
            s_c[threadIdx.x] += s_c[threadIdx.x + st];
        }
        __syncthreads();
    }
    const float cnt = s_c[0] > 0.f ? s_c[0] : 1.f;
    if (threadIdx.x == 0) *loss = s_l[0] / cnt;
    for (uint32_t i = threadIdx.x; i < N; i += 256) {
        const float mk = mask ? mask[i] : 1.f;
        dy[i] = 2.f * mk * (y[i] - t[i]) / cnt;
    }
}

// dW2[o][h] = sum_b dy[b][o] hidden[b][h]; db2[o] = sum_b dy[b][o]  (one thread per output)
__global__ void mlp_dw2_kernel(const float* dy, const float* hidden, uint32_t B, uint32_t H, uint32_t O, float* dW2,
                               float* db2) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < O * H) {
        const uint32_t o = idx / H, h = idx % H;
        float s = 0.f;
        for (uint32_t b = 0; b < B; ++b) s += dy[b * O + o] * hidden[b * H + h];
        dW2[idx] = s;
    } else if (idx < O * H + O) {
        const uint32_t o = idx - O * H;
        float s = 0.f;
        for (uint32_t b = 0; b < B; ++b) s += dy[b * O + o];
        db2[o] = s;
    }
}

// dh[b][h] = (sum_o dy[b][o] W2[o][h]) * [hidden[b][h] > 0]
__global__ void mlp_dh_kernel(const float* dy, const float* W2, const float* hidden, uint32_t B, uint32_t H,
                              uint32_t O, float* dh) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= B * H) return;
    const uint32_t b = idx / H, h = idx % H;
    float s = 0.f;
    for (uint32_t o = 0; o < O; ++o) s += dy[b * O + o] * W2[o * H + h];
    dh[idx] = hidden[idx] > 0.f ? s : 0.f;
}

// dW1[h][f] = sum_b dh[b][h] X[b][f]; db1[h] = sum_b dh[b][h]
__global__ void mlp_dw1_kernel(const float* dh, const float* X, uint32_t B, uint32_t F, uint32_t H, float* dW1,
                               float* db1) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < H * F) {
        const uint32_t h = idx / F, f = idx % F;
        float s = 0.f;
        for (uint32_t b = 0; b < B; ++b) s += dh[b * H + h] * X[b * F + f];
        dW1[idx] = s;
    } else if (idx < H * F + H) {
        const uint32_t h = idx - H * F;
        float s = 0.f;
        for (uint32_t b = 0; b < B; ++b) s += dh[b * H + h];
        db1[h] = s;
    }
}

__global__ void adam_kernel(float* w, const float* g, float* m1, float* m2, uint32_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

inline uint32_t blocks(uint64_t n) { return (uint32_t)((n + 255) / 256); }

}  // namespace
}  // namespace hx

using namespace hx;

extern "C" {

HX_API int hx_mlp_forward(const float* X, uint32_t B, uint32_t F, const float* W1, const float* b1, uint32_t H,
                          const float* W2, const float* b2, uint32_t O, float* hidden, float* y, hx_stream s) {
    return guard("hx_mlp_forward", [&] {
        HX_CHECK(X && W1 && b1 && W2 && b2 && hidden && y, "NULL argument");
        if (B == 0) return;
        mlp_hidden_kernel<<<blocks((uint64_t)B * H), 256, 0, (hipStream_t)s>>>(X, B, F, W1, b1, H, hidden);
        after_launch("mlp_hidden_kernel");
        mlp_out_kernel<<<blocks((uint64_t)B * O), 256, 0, (hipStream_t)s>>>(hidden, B, H, W2, b2, O, y);
        after_launch("mlp_out_kernel");
    });
}

HX_API int hx_mlp_backward(const float* X, uint32_t B, uint32_t F, const float* W1, uint32_t H, const float* W2,
                           uint32_t O, const float* hidden, const float* y, const float* target, const float* mask,
                           float* dW1, float* db1, float* dW2, float* db2, float* loss, hx_stream s) {
    return guard("hx_mlp_backward", [&] {
        HX_CHECK(X && W1 && W2 && hidden && y && target && dW1 && db1 && dW2 && db2 && loss, "NULL argument");
        (void)W1;
        const hipStream_t st = (hipStream_t)s;
        float* dy = (float*)stream_scratch((hipStream_t)s, SCRATCH_MLP, ((size_t)B * O + (size_t)B * H + 1) * sizeof(float));
        HX_CHECK(dy != nullptr, "hx_mlp_backward: its scratch cannot grow inside a stream capture (run it once before capturing)");
        float* dh = dy + (size_t)B * O;
        mlp_dy_kernel<<<1, 256, 0, st>>>(y, target, mask, B * O, dy, loss);
        after_launch("mlp_dy_kernel");
        mlp_dw2_kernel<<<blocks((uint64_t)O * H + O), 256, 0, st>>>(dy, hidden, B, H, O, dW2, db2);
        after_launch("mlp_dw2_kernel");
        mlp_dh_kernel<<<blocks((uint64_t)B * H), 256, 0, st>>>(dy, W2, hidden, B, H, O, dh);
        after_launch("mlp_dh_kernel");
        mlp_dw1_kernel<<<blocks((uint64_t)H * F + H), 256, 0, st>>>(dh, X, B, F, H, dW1, db1);
        after_launch("mlp_dw1_kernel");
    });
}

HX_API int hx_adam_step(float* param, const float* grad, float* m1, float* m2, uint32_t n, float lr, float beta1,
                        float beta2, float eps, uint32_t t, hx_stream s) {
    return guard("hx_adam_step", [&] {
        HX_CHECK(param && grad && m1 && m2 && t >= 1, "bad argument");
        if (n == 0) return;
        const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
        adam_kernel<<<blocks(n), 256, 0, (hipStream_t)s>>>(param, grad, m1, m2, n, lr, beta1, beta2, eps, c1, c2);
        after_launch("adam_kernel");
    });
}

}  // extern "C"
