// Shared definitions of the HIP kernels of the global-placement iteration (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/autoplace_gp.h"

namespace ap {

// fixed point of the density maps: 2^32 per bin area (integer sums do not depend on their order)
constexpr double kDensityScale = 4294967296.0;

// floor(w 2^32) of a float w, as (unsigned long long)((double)w * 2^32) gives it, in single-precision operations
// without branches: f = w 2^32 is exact (a power-of-two scale), h = trunc(f 2^-32) its high word, and f - h 2^32 is
// exact (both are multiples of f's unit in the last place, the difference below 2^32), so its truncation is the low
// word. For every w in [0, 2^31); w <= 0, w >= 2^31 and NaN give 0.
__host__ __device__ __forceinline__ unsigned long long fixed_point32(float w) {
    const bool ok = w > 0.f && w < 2147483648.f;
    const float f = (ok ? w : 0.f) * 4294967296.f;
    const float h = truncf(f * 2.3283064365386963e-10f);
    const unsigned hi = (unsigned)h, lo = (unsigned)(f - h * 4294967296.f);
    return ((unsigned long long)hi << 32) | lo;
}

// control state layout (S, float64) and constants (c, float64)
enum { LAM, GAM, ALPHA, A, ITER, HPWL_PREV, MOTION, WW, MU, SETTLED, BEST_HPWL, BEST_OVF, IS_BEST, IMPROVED, STOP,
       ACCEPT, LOOP_IT, EVALS, RES_HPWL, RES_OVF, HAS_BEST, NUM_S };
enum { C_BASE_GAMMA, C_MOTION_UNIT, C_TOTAL_PINS, C_TOTAL_AREA, C_PITCH, C_MIN_ITER, C_STOP_OVF, C_SETTLE, C_STALL,
       C_INNER, C_REFINE, C_ACCURACY, C_HANDOVER, NUM_C };

void set_error(const char* fmt, ...);

// Device buffers of the library (gp_devmem.hip): dev_malloc as hipMalloc, from a cache of released blocks of the same
// size class on the current device; dev_free synchronises the device (as hipFree does) and keeps the block for the
// next request of its class; dev_free_synced is for callers that have just synchronised the device. Null is ignored.
hipError_t dev_malloc(void** p, size_t bytes);
void dev_free(void* p);
void dev_free_synced(void* p);
template <typename T>
hipError_t dev_malloc(T** p, size_t bytes) {
    return dev_malloc((void**)p, bytes);
}

#define AP_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            ::ap::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_));         \
            return AP_ERR_HIP;                                                                         \
        }                                                                                              \
    } while (0)

#define AP_LAUNCHED()                                                                                  \
    do {                                                                                               \
        hipError_t e_ = hipGetLastError();                                                             \
        if (e_ != hipSuccess) {                                                                        \
            ::ap::set_error("%s:%d launch: %s", __FILE__, __LINE__, hipGetErrorString(e_));            \
            return AP_ERR_HIP;                                                                         \
        }                                                                                              \
    } while (0)

#define AP_REQUIRE(cond, ...)                                                                          \
    do {                                                                                               \
        if (!(cond)) {                                                                                 \
            ::ap::set_error(__VA_ARGS__);                                                              \
            return AP_ERR_ARG;                                                                         \
        }                                                                                              \
    } while (0)

inline unsigned grid_of(long long n, int block) { return (unsigned)((n + block - 1) / block); }

// extent of a node along one axis in bin units, and its first / last bin (clamped to the grid)
struct Span {
    float lo, hi;
    int b0, b1;
};
__device__ __forceinline__ Span make_span(float centre, float size, float inv_bin, int nb) {
    Span s;
    s.lo = (centre - 0.5f * size) * inv_bin;
    s.hi = (centre + 0.5f * size) * inv_bin;
    s.b0 = max((int)floorf(s.lo), 0);
    s.b1 = min((int)floorf(s.hi), nb - 1);
    return s;
}
__device__ __forceinline__ float span_overlap(const Span& s, int b) {
    return fmaxf(fminf(s.hi, (float)b + 1.f) - fmaxf(s.lo, (float)b), 0.f);
}

// wavefront-wide sum of a double (64 lanes), result in every lane
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// block-wide sum of up to 4 doubles per thread in a fixed order; blockDim.x must be a multiple of 64 (<= 1024).
// The result is valid in thread 0.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double* lds /* [16 * K] */) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < K; k++) lds[wave * K + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            double s = 0;
            for (int w = 0; w < nw; w++) s += lds[w * K + k];
            v[k] = s;
        }
    }
    __syncthreads();
}

}  // namespace ap
