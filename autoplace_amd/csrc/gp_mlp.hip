// Inference and training kernels of the learned schedule controller's MLP (gp_mlp.h): batched forward, the
// gradient of a weighted squared loss, and Adam. The controller itself runs inside the engine's control kernel.
//
// Gradient: one 64-lane workgroup per chunk of 64 samples. Lane s runs sample s forward and backward through LDS
// (rows padded to 65 floats: lanes reading the same column hit different banks), then the workgroup forms the
// chunk's parameter gradient as sums over its 64 samples in sample order, written as one partial per chunk; a second
// kernel adds the partials in chunk order. The result does not depend on the launch (deterministic).
#include "gp_common.h"
#include "gp_mlp.h"

namespace ap {

constexpr int MB = 64;        // samples per chunk (one wavefront)
constexpr int ROW = MLP_H + 1;

__global__ __launch_bounds__(256) void mlp_forward_kernel(const float* P, const float* x, int n, float* out) {
    __shared__ float W[MLP_PARAMS];
    for (int k = threadIdx.x; k < MLP_PARAMS; k += 256) W[k] = P[k];
    __syncthreads();
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    float xi[MLP_IN], h1[MLP_H];
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) xi[i] = x[(long long)s * MLP_IN + i];
#pragma unroll
    for (int j = 0; j < MLP_H; j++) {
        float a = W[MLP_B1 + j];
#pragma unroll
        for (int i = 0; i < MLP_IN; i++) a += W[MLP_W1 + j * MLP_IN + i] * xi[i];
        h1[j] = fmaxf(a, 0.f);
    }
    float o0 = W[MLP_B3], o1 = W[MLP_B3 + 1];
    for (int j = 0; j < MLP_H; j++) {
        float b = W[MLP_B2 + j];
#pragma unroll
        for (int k = 0; k < MLP_H; k++) b += W[MLP_W2 + j * MLP_H + k] * h1[k];
        const float h2 = fmaxf(b, 0.f);
        o0 += W[MLP_W3 + j] * h2;
        o1 += W[MLP_W3 + MLP_H + j] * h2;
    }
    out[2LL * s] = o0;
    out[2LL * s + 1] = o1;
}

// loss = sum_s w_s |o_s - t_s|^2; partial[chunk] = (d loss / d params [MLP_PARAMS], loss)
__global__ __launch_bounds__(MB) void mlp_grad_kernel(const float* P, const float* x, const float* t, const float* w,
                                                      int n, float* partial) {
    extern __shared__ float lds[];
    float* W = lds;                              // MLP_PARAMS
    float* X = W + MLP_PARAMS;                   // [MB][MLP_IN]
    float* H1 = X + MB * MLP_IN;                 // [MB][ROW]
    float* H2 = H1 + MB * ROW;                   // [MB][ROW]
    float* D1 = H2 + MB * ROW;                   // [MB][ROW]
    float* D2 = D1 + MB * ROW;                   // [MB][ROW]
    float* DO = D2 + MB * ROW;                   // [MB][2]
    __shared__ double lred[1];
    const int s = threadIdx.x;
    const long long g = (long long)blockIdx.x * MB + s;
    const bool live = g < n;
    for (int k = s; k < MLP_PARAMS; k += MB) W[k] = P[k];
    for (int i = 0; i < MLP_IN; i++) X[s * MLP_IN + i] = live ? x[g * MLP_IN + i] : 0.f;
    __syncthreads();
    // forward (lane s: sample s)
    for (int j = 0; j < MLP_H; j++) {
        float a = W[MLP_B1 + j];
        for (int i = 0; i < MLP_IN; i++) a += W[MLP_W1 + j * MLP_IN + i] * X[s * MLP_IN + i];
        H1[s * ROW + j] = fmaxf(a, 0.f);
    }
    float o0 = W[MLP_B3], o1 = W[MLP_B3 + 1];
    for (int j = 0; j < MLP_H; j++) {
        float b = W[MLP_B2 + j];
        for (int k = 0; k < MLP_H; k++) b += W[MLP_W2 + j * MLP_H + k] * H1[s * ROW + k];
        const float h2 = fmaxf(b, 0.f);
        H2[s * ROW + j] = h2;
        o0 += W[MLP_W3 + j] * h2;
        o1 += W[MLP_W3 + MLP_H + j] * h2;
    }
    // backward
    const float ws = live ? w[g] : 0.f;
    const float e0 = live ? o0 - t[2 * g] : 0.f, e1 = live ? o1 - t[2 * g + 1] : 0.f;
    double loss = (double)ws * ((double)e0 * e0 + (double)e1 * e1);
    const float do0 = 2.f * ws * e0, do1 = 2.f * ws * e1;
    DO[2 * s] = do0;
    DO[2 * s + 1] = do1;
    for (int j = 0; j < MLP_H; j++) {
        const float d = W[MLP_W3 + j] * do0 + W[MLP_W3 + MLP_H + j] * do1;
        D2[s * ROW + j] = H2[s * ROW + j] > 0.f ? d : 0.f;
    }
    for (int k = 0; k < MLP_H; k++) {
        float d = 0.f;
        for (int j = 0; j < MLP_H; j++) d += W[MLP_W2 + j * MLP_H + k] * D2[s * ROW + j];
        D1[s * ROW + k] = H1[s * ROW + k] > 0.f ? d : 0.f;
    }
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
    if (s == 0) lred[0] = loss;
    __syncthreads();
    // the chunk's parameter gradient: sums over its samples, in sample order
    float* out = partial + (long long)blockIdx.x * (MLP_PARAMS + 1);
    for (int p = s; p < MLP_PARAMS; p += MB) {
        float acc = 0.f;
        if (p < MLP_B1) {                                  // W1[j][i]
            const int j = p / MLP_IN, i = p % MLP_IN;
            for (int q = 0; q < MB; q++) acc += D1[q * ROW + j] * X[q * MLP_IN + i];
        } else if (p < MLP_W2) {                           // b1[j]
            const int j = p - MLP_B1;
            for (int q = 0; q < MB; q++) acc += D1[q * ROW + j];
        } else if (p < MLP_B2) {                           // W2[j][k]
            const int j = (p - MLP_W2) / MLP_H, k = (p - MLP_W2) % MLP_H;
            for (int q = 0; q < MB; q++) acc += D2[q * ROW + j] * H1[q * ROW + k];
        } else if (p < MLP_W3) {                           // b2[j]
            const int j = p - MLP_B2;
            for (int q = 0; q < MB; q++) acc += D2[q * ROW + j];
        } else if (p < MLP_B3) {                           // W3[o][j]
            const int o = (p - MLP_W3) / MLP_H, j = (p - MLP_W3) % MLP_H;
            for (int q = 0; q < MB; q++) acc += DO[2 * q + o] * H2[q * ROW + j];
        } else {                                           // b3[o]
            const int o = p - MLP_B3;
            for (int q = 0; q < MB; q++) acc += DO[2 * q + o];
        }
        out[p] = acc;
    }
    if (s == 0) out[MLP_PARAMS] = (float)lred[0];
}

// grad[p] = sum over chunks of partial[chunk][p], in chunk order (FP64 accumulation); loss likewise
__global__ __launch_bounds__(256) void mlp_sum_kernel(const float* partial, int chunks, float* grad, double* loss) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p > MLP_PARAMS) return;
    double acc = 0.0;
    for (int b = 0; b < chunks; b++) acc += partial[(long long)b * (MLP_PARAMS + 1) + p];
    if (p < MLP_PARAMS) grad[p] = (float)acc;
    else if (loss) *loss = acc;
}

__global__ __launch_bounds__(256) void mlp_adam_kernel(float* P, float* m, float* v, const float* grad, float lr,
                                                       float b1, float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= MLP_PARAMS) return;
    const float gp = grad[p];
    const float mp = b1 * m[p] + (1.f - b1) * gp;
    const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    P[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

}  // namespace ap

using namespace ap;

extern "C" {

int ap_mlp_param_count(void) { return MLP_PARAMS; }

long long ap_mlp_work_floats(int n) { return (long long)grid_of(n > 0 ? n : 1, MB) * (MLP_PARAMS + 1); }

int ap_mlp_forward(const float* params, const float* x, int n, float* out, void* stream) {
    AP_REQUIRE(params && x && out && n >= 0, "mlp_forward: bad arguments");
    if (n == 0) return AP_OK;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(grid_of(n, 256)), dim3(256), 0, (hipStream_t)stream, params, x, n,
                       out);
    AP_LAUNCHED();
    return AP_OK;
}

int ap_mlp_grad(const float* params, const float* x, const float* target, const float* weight, int n, float* grad,
                double* loss, float* work, void* stream) {
    AP_REQUIRE(params && x && target && weight && grad && work && n > 0, "mlp_grad: bad arguments");
    const int chunks = (int)grid_of(n, MB);
    const size_t lds = (MLP_PARAMS + MB * MLP_IN + 4 * MB * ROW + 2 * MB) * sizeof(float);
    static bool attr = false;
    if (!attr) {
        AP_HIP(hipFuncSetAttribute((const void*)mlp_grad_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
        attr = true;
    }
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(chunks), dim3(MB), lds, (hipStream_t)stream, params, x, target, weight,
                       n, work);
    AP_LAUNCHED();
    hipLaunchKernelGGL(mlp_sum_kernel, dim3(grid_of(MLP_PARAMS + 1, 256)), dim3(256), 0, (hipStream_t)stream, work,
                       chunks, grad, loss);
    AP_LAUNCHED();
    return AP_OK;
}

int ap_mlp_adam(float* params, float* m, float* v, const float* grad, int step, float lr, float beta1, float beta2,
                float eps, void* stream) {
    AP_REQUIRE(params && m && v && grad && step >= 1, "mlp_adam: bad arguments");
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3(grid_of(MLP_PARAMS, 256)), dim3(256), 0, (hipStream_t)stream, params, m,
                       v, grad, lr, beta1, beta2, eps, c1, c2);
    AP_LAUNCHED();
    return AP_OK;
}

}  // extern "C"
