// The learned schedule controller: a 12 -> 64 -> 64 -> 2 ReLU MLP in FP32 whose two outputs, clipped to +-0.2, are
// the logs of factors on the density-weight multiplier and on the smoothing of each pass (ScheduleAdjust). Its
// features are per-pass statistics of the run (mlp_features). It has no counterpart in the reference, whose rule it
// adjusts (gp_cuda_kernel.cu:198-209); with zero outputs, or after `window` passes without an overflow improvement
// once the density share is 0.95 or more (the fallback), the rule runs unchanged.
#pragma once
#include "gp_common.h"
#include "gp_control.h"

namespace ap {

constexpr int MLP_IN = 12, MLP_H = 64, MLP_OUT = 2;
// parameter layout (floats): W1 [64][12], b1 [64], W2 [64][64], b2 [64], W3 [2][64], b3 [2]
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_H * MLP_IN, MLP_W2 = MLP_B1 + MLP_H, MLP_B2 = MLP_W2 + MLP_H * MLP_H,
              MLP_W3 = MLP_B2 + MLP_H, MLP_B3 = MLP_W3 + MLP_OUT * MLP_H, MLP_PARAMS = MLP_B3 + MLP_OUT;
constexpr float MLP_CLIP = 0.2f;
constexpr int MLP_LOG = 16;    // floats per logged pass: 12 features, 2 applied outputs, overflow, HPWL

// controller state per level (double X[NUM_X], device)
enum { X_STARTED, X_PREV_OVF, X_HPWL0, X_ALPHA0, X_BEST_OVF, X_LAST_IMPROVE, X_FALLBACK, X_COUNT, NUM_X };

struct Ctl {
    const float* params;        // null: the plain rule
    double* X;                  // controller state
    float sigma;                // exploration noise added to the outputs (0: none)
    unsigned long long seed;
    float* log;                 // null or [log_cap][MLP_LOG]
    int log_cap;
    int window;                 // fallback after this many passes without a 0.1 % overflow improvement once
                                // the density share is 0.95 or more (the rule's own stall test starts there)
    int hold;                   // the exploration draw changes every `hold` learned passes
    float phase;                // 1: a coarse (cluster) level, 0: the cell level
};

// Features of a pass, from the state before the schedule update and the pass's statistics; each clamped to
// [-10, 10]. wl2 / dn2: squared norms of the wirelength and (weighted) density parts of the gradient.
__device__ __forceinline__ void mlp_features(const double* S, const double* c, const double* X, double moved,
                                             double turned, double hpwl, double ovf, double wa, double wl2, double dn2,
                                             float phase, float* x) {
    const double alpha = turned > 0 ? moved / turned : S[ALPHA];
    const double dens_part = S[LAM] * c[C_TOTAL_AREA];
    const double hp_prev = S[HPWL_PREV];
    double f[MLP_IN];
    f[0] = ovf;
    f[1] = 100.0 * (ovf - X[X_PREV_OVF]);
    f[2] = log(fmax(hpwl, 1e-30) / fmax(X[X_HPWL0], 1e-30));
    f[3] = hp_prev < 1e29 ? 10.0 * (hpwl - hp_prev) / fmax(hp_prev, 1e-30) : 0.0;
    f[4] = 1.0 - wa / fmax(hpwl, 1e-30);
    f[5] = log(fmax(S[GAM], 1e-30) / fmax(c[C_BASE_GAMMA], 1e-30));
    f[6] = dens_part / (c[C_TOTAL_PINS] + dens_part);
    f[7] = log(fmax(alpha, 1e-30) / fmax(X[X_ALPHA0], 1e-30));
    f[8] = 0.2 * log(sqrt(fmax(wl2, 1e-60) / fmax(dn2, 1e-60)));
    f[9] = S[ITER] / 1000.0;
    f[10] = log1p(moved / c[C_MOTION_UNIT]);
    f[11] = phase;
#pragma unroll
    for (int k = 0; k < MLP_IN; k++) x[k] = (float)fmin(10.0, fmax(-10.0, isfinite(f[k]) ? f[k] : 0.0));
}

// counter-based normal draw (splitmix64 + Box-Muller): the same (seed, pass, k) gives the same number
__device__ __forceinline__ float normal_draw(unsigned long long seed, unsigned long long pass, int k) {
    auto mix = [](unsigned long long z) {
        z += 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const unsigned long long a = mix(seed ^ mix(pass * 4 + (unsigned long long)k));
    const unsigned long long b = mix(a);
    const double u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0);   // (0, 1]
    const double u2 = (b >> 11) * (1.0 / 9007199254740992.0);
    return (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

// The MLP on one feature vector by one 64-lane workgroup: lane j computes hidden unit j of both layers; the two
// outputs are wavefront sums. Every lane returns the outputs (unclipped). h: LDS [MLP_H].
__device__ __forceinline__ void mlp_forward_block(const float* P, const float* x, float* h, float& o0, float& o1) {
    const int j = threadIdx.x;
    float a = P[MLP_B1 + j];
#pragma unroll
    for (int k = 0; k < MLP_IN; k++) a += P[MLP_W1 + j * MLP_IN + k] * x[k];
    h[j] = fmaxf(a, 0.f);
    __syncthreads();
    float b = P[MLP_B2 + j];
    for (int k = 0; k < MLP_H; k++) b += P[MLP_W2 + j * MLP_H + k] * h[k];
    const float h2 = fmaxf(b, 0.f);
    float s0 = P[MLP_W3 + j] * h2, s1 = P[MLP_W3 + MLP_H + j] * h2;
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, 64);
        s1 += __shfl_xor(s1, o, 64);
    }
    o0 = s0 + P[MLP_B3];
    o1 = s1 + P[MLP_B3 + 1];
    __syncthreads();
}

}  // namespace ap
