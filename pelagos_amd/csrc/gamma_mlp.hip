// gamma_mlp.hip -- training kernels of the learned memory-strength proposal.
//
// The network maps 8 per-variable features (taken at the end of a leg) to a
// shift of that variable's memory strength in the next relay legs:
//     h1 = relu(W1 x + b1)      W1 [32][8]
//     h2 = relu(W2 h1 + b2)     W2 [32][32]
//     z  = w3 . h2 + b3         w3 [32]
//     shift = (hi - lo) (sigmoid(z) - 1/2), gamma = clamp(u + shift, lo, hi)
// with u the leg's own uniform draw on [lo, hi].  Parameters are one fp32
// vector in that order (kMlpParams floats).  Training fits sigmoid(z) to soft
// targets in [0, 1] with a weighted logistic loss (the targets and weights
// come from reward-weighted regression over candidate shifts, in learned.py).
//
// Kernels: a fused forward/backward per sample whose parameter gradients are
// summed in a fixed order (per-block partial sums, then a fixed-order sum over
// blocks: bit-reproducible), an Adam step, and a plain fp32 forward.  The
// inference used while decoding is the bf16-MFMA version inside
// relaybp.hip (learned_gamma).  No counterpart in the reference, which has no
// learned component.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <string>

#include "pelagos_relaybp.h"

namespace {

constexpr int kIn = PEL_MLP_FEATURES, kH = PEL_MLP_HIDDEN;
constexpr int kW1 = 0, kB1 = kW1 + kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH, kB3 = kW3 + kH;
constexpr int kMlpParams = kB3 + 1;
static_assert(kMlpParams == PEL_MLP_PARAMS, "parameter layout");

constexpr int kGradBlock = 256;
constexpr int kTile = 64;             // samples per forward/backward tile
constexpr int kOwn = (PEL_MLP_PARAMS + kGradBlock - 1) / kGradBlock;  // parameters per thread
constexpr int kGradBlocks = 120;      // fixed, so the summation order never depends on the device
constexpr int kVec = kIn + 4 * kH + 1;  // per-sample vectors kept for the parameter sums

thread_local std::string g_err;

int fail(int code, const char *msg) {
    g_err = msg;
    return code;
}

// Forward and backward of one sample; fills x, h1, h2, dh1, dh2 (pre-activation
// gradients) and dz in vec, returns the sample's weighted loss.
__device__ float sample_fwd_bwd(const float *__restrict__ p, const float *__restrict__ x_in, float y, float w,
                                float *vec) {
    float x[kIn], h1[kH], h2[kH];
#pragma unroll
    for (int i = 0; i < kIn; ++i) x[i] = x_in[i];
#pragma unroll 4
    for (int j = 0; j < kH; ++j) {
        float s = p[kB1 + j];
#pragma unroll
        for (int i = 0; i < kIn; ++i) s += p[kW1 + j * kIn + i] * x[i];
        h1[j] = fmaxf(s, 0.f);
    }
    float z = p[kB3];
#pragma unroll 4
    for (int j = 0; j < kH; ++j) {
        float s = p[kB2 + j];
        for (int i = 0; i < kH; ++i) s += p[kW2 + j * kH + i] * h1[i];
        h2[j] = fmaxf(s, 0.f);
        z += p[kW3 + j] * h2[j];
    }
    // weighted logistic loss: w * (softplus(z) - y z)
    const float sp = fmaxf(z, 0.f) + log1pf(expf(-fabsf(z)));
    const float loss = w * (sp - y * z);
    const float sig = 1.f / (1.f + expf(-z));
    const float dz = w * (sig - y);
    float dh2[kH], dh1[kH];
#pragma unroll 4
    for (int j = 0; j < kH; ++j) dh2[j] = h2[j] > 0.f ? dz * p[kW3 + j] : 0.f;
#pragma unroll 4
    for (int i = 0; i < kH; ++i) {
        float s = 0.f;
        for (int j = 0; j < kH; ++j) s += p[kW2 + j * kH + i] * dh2[j];
        dh1[i] = h1[i] > 0.f ? s : 0.f;
    }
    float *o = vec;
#pragma unroll
    for (int i = 0; i < kIn; ++i) *o++ = x[i];
    for (int j = 0; j < kH; ++j) *o++ = h1[j];
    for (int j = 0; j < kH; ++j) *o++ = h2[j];
    for (int j = 0; j < kH; ++j) *o++ = dh1[j];
    for (int j = 0; j < kH; ++j) *o++ = dh2[j];
    *o = dz;
    return loss;
}

// d(sum of losses)/d(param q) over the kTile samples staged in LDS.
__device__ float param_grad(int q, const float *vecs, int count) {
    // vec layout: x[0,8) h1[8,40) h2[40,72) dh1[72,104) dh2[104,136) dz[136]
    float g = 0.f;
    if (q < kB1) {  // W1[j][i] = dh1[j] * x[i]
        const int j = q / kIn, i = q % kIn;
        for (int s = 0; s < count; ++s) g += vecs[s * kVec + kIn + 2 * kH + j] * vecs[s * kVec + i];
    } else if (q < kW2) {
        const int j = q - kB1;
        for (int s = 0; s < count; ++s) g += vecs[s * kVec + kIn + 2 * kH + j];
    } else if (q < kB2) {  // W2[j][i] = dh2[j] * h1[i]
        const int j = (q - kW2) / kH, i = (q - kW2) % kH;
        for (int s = 0; s < count; ++s) g += vecs[s * kVec + kIn + 3 * kH + j] * vecs[s * kVec + kIn + i];
    } else if (q < kW3) {
        const int j = q - kB2;
        for (int s = 0; s < count; ++s) g += vecs[s * kVec + kIn + 3 * kH + j];
    } else if (q < kB3) {  // w3[j] = dz * h2[j]
        const int j = q - kW3;
        for (int s = 0; s < count; ++s) g += vecs[s * kVec + 4 * kH + kIn] * vecs[s * kVec + kIn + kH + j];
    } else {
        for (int s = 0; s < count; ++s) g += vecs[s * kVec + 4 * kH + kIn];
    }
    return g;
}

// Block b handles samples [b * per, min(n, (b + 1) * per)) in tiles of kTile;
// thread t owns parameters t, t + kGradBlock, ...  partial[b][q] = sum of the
// block's per-sample gradients in sample order; partial[b][P] = loss sum.
__global__ void __launch_bounds__(kGradBlock) mlp_grad_kernel(const float *__restrict__ params,
                                                              const float *__restrict__ x,
                                                              const float *__restrict__ y,
                                                              const float *__restrict__ w, int64_t n, int64_t per,
                                                              float *__restrict__ partial) {
    __shared__ float p[kMlpParams];
    __shared__ float vecs[kTile * kVec];
    __shared__ float losses[kTile];
    const int tid = threadIdx.x;
    for (int q = tid; q < kMlpParams; q += kGradBlock) p[q] = params[q];
    __syncthreads();
    const int64_t lo = (int64_t)blockIdx.x * per;
    const int64_t hi = lo + per < n ? lo + per : n;
    float g[kOwn], lsum = 0.f;
#pragma unroll
    for (int k = 0; k < kOwn; ++k) g[k] = 0.f;
    for (int64_t t0 = lo; t0 < hi; t0 += kTile) {
        const int count = (int)(hi - t0 < kTile ? hi - t0 : kTile);
        if (tid < count) {
            const int64_t s = t0 + tid;
            losses[tid] = sample_fwd_bwd(p, x + s * kIn, y[s], w[s], vecs + tid * kVec);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kOwn; ++k)
            if (tid + k * kGradBlock < kMlpParams) g[k] += param_grad(tid + k * kGradBlock, vecs, count);
        if (tid == 0)
            for (int s = 0; s < count; ++s) lsum += losses[s];
        __syncthreads();
    }
    float *out = partial + (int64_t)blockIdx.x * (kMlpParams + 1);
#pragma unroll
    for (int k = 0; k < kOwn; ++k)
        if (tid + k * kGradBlock < kMlpParams) out[tid + k * kGradBlock] = g[k];
    if (tid == 0) out[kMlpParams] = lsum;
}

// grad[q] = (sum over blocks in order) / n; loss = mean loss.
__global__ void mlp_reduce_kernel(const float *__restrict__ partial, int nblocks, int64_t n, float *__restrict__ grad,
                                  float *__restrict__ loss) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > kMlpParams) return;
    float s = 0.f;
    for (int b = 0; b < nblocks; ++b) s += partial[(int64_t)b * (kMlpParams + 1) + q];
    const float v = s / (float)n;
    if (q < kMlpParams) grad[q] = v;
    else if (loss) *loss = v;
}

__global__ void mlp_adam_kernel(float *__restrict__ params, float *__restrict__ m1, float *__restrict__ m2,
                                const float *__restrict__ grad, float lr, float b1, float b2, float eps,
                                float c1, float c2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= kMlpParams) return;
    const float g = grad[q];
    const float a = b1 * m1[q] + (1.f - b1) * g;
    const float v = b2 * m2[q] + (1.f - b2) * g * g;
    m1[q] = a;
    m2[q] = v;
    params[q] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

__global__ void mlp_forward_kernel(const float *__restrict__ params, const float *__restrict__ x, int64_t n,
                                   float *__restrict__ z_out) {
    __shared__ float p[kMlpParams];
    for (int q = threadIdx.x; q < kMlpParams; q += blockDim.x) p[q] = params[q];
    __syncthreads();
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x) {
        float h1[kH];
        for (int j = 0; j < kH; ++j) {
            float a = p[kB1 + j];
            for (int i = 0; i < kIn; ++i) a += p[kW1 + j * kIn + i] * x[s * kIn + i];
            h1[j] = fmaxf(a, 0.f);
        }
        float z = p[kB3];
        for (int j = 0; j < kH; ++j) {
            float a = p[kB2 + j];
            for (int i = 0; i < kH; ++i) a += p[kW2 + j * kH + i] * h1[i];
            z += p[kW3 + j] * fmaxf(a, 0.f);
        }
        z_out[s] = z;
    }
}

}  // namespace

extern "C" {

int pel_mlp_param_count(void) { return kMlpParams; }

const char *pel_mlp_last_error(void) { return g_err.c_str(); }

int64_t pel_mlp_grad_workspace_bytes(void) { return (int64_t)kGradBlocks * (kMlpParams + 1) * 4; }

int pel_mlp_grad(const float *params, const float *x, const float *y, const float *w, int64_t n, float *grad,
                 float *loss, void *workspace, int64_t workspace_bytes, void *stream) {
    if (!params || !x || !y || !w || !grad || !workspace) return fail(PEL_EINVAL, "pel_mlp_grad: NULL argument");
    if (n <= 0) return fail(PEL_EINVAL, "pel_mlp_grad: empty batch");
    if (workspace_bytes < pel_mlp_grad_workspace_bytes()) return fail(PEL_EINVAL, "pel_mlp_grad: workspace too small");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t per = ((n + kGradBlocks - 1) / kGradBlocks + kTile - 1) / kTile * kTile;
    float *partial = static_cast<float *>(workspace);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(kGradBlocks), dim3(kGradBlock), 0, s, params, x, y, w, n, per, partial);
    if (hipGetLastError() != hipSuccess) return fail(PEL_EDEVICE, "pel_mlp_grad: launch failed");
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, s, partial, kGradBlocks, n,
                       grad, loss);
    if (hipGetLastError() != hipSuccess) return fail(PEL_EDEVICE, "pel_mlp_grad: launch failed");
    return PEL_OK;
}

int pel_mlp_adam(float *params, float *m1, float *m2, const float *grad, float lr, float beta1, float beta2,
                 float eps, int32_t step, void *stream) {
    if (!params || !m1 || !m2 || !grad) return fail(PEL_EINVAL, "pel_mlp_adam: NULL argument");
    if (step < 1) return fail(PEL_EINVAL, "pel_mlp_adam: step counts from 1");
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                       params, m1, m2, grad, lr, beta1, beta2, eps, c1, c2);
    if (hipGetLastError() != hipSuccess) return fail(PEL_EDEVICE, "pel_mlp_adam: launch failed");
    return PEL_OK;
}

int pel_mlp_forward(const float *params, const float *x, int64_t n, float *z, void *stream) {
    if (!params || !x || !z) return fail(PEL_EINVAL, "pel_mlp_forward: NULL argument");
    if (n <= 0) return PEL_OK;
    const int64_t blocks = (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                       params, x, n, z);
    if (hipGetLastError() != hipSuccess) return fail(PEL_EDEVICE, "pel_mlp_forward: launch failed");
    return PEL_OK;
}

}  // extern "C"
