// WA wirelength of the engine: one lane per pin, nets packed whole into 64-lane wavefronts.
//
// The drop-in kernel (gp_kernels.hip) gives each net one lane that loops over the net's pins: the lanes of a
// wavefront then wait for its longest net (degrees 2..50 here, up to --ignore_net_degree), and the per-net phase is
// VALU-bound. Here the nets of a level are packed, once, into wavefronts of at most 64 pins (no net straddles two
// wavefronts; nets of more than 64 pins go to a workgroup-wide path). Every lane holds one pin; each net is a
// segment of lanes, and max / min / the four exponential sums per axis are segmented tree reductions over the
// lanes (shuffles), with the results broadcast from the segment's first lane. The sums are taken in a fixed tree
// order, so the results are deterministic (and within FP32 rounding of the sequential sums).
#include <vector>

#include "gp_common.h"
#include "gp_kernels.h"
#include "gp_wa.h"
#include "gp_wa_common.h"

namespace ap {

struct WaveArgs {
    const float2* pos;
    const float2* pos_other;
    const int* flag;
    const int* pin_cell;
    const float2* pin_off;
    const int* net_start;
    const int4* waves;       // pin_begin, pin_count, net_begin, net_count
    int num_waves;
    int num_mov;
    const float* gamma;
    const int* pin_slot;
    float2* pin_grad;
    double* partial;         // per workgroup (WA, HPWL)
};

__device__ __forceinline__ float seg_max(float v, int lane, int end) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const float o = __shfl_down(v, off, 64);
        if (lane + off < end) v = fmaxf(v, o);
    }
    return v;
}
__device__ __forceinline__ float seg_min(float v, int lane, int end) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const float o = __shfl_down(v, off, 64);
        if (lane + off < end) v = fminf(v, o);
    }
    return v;
}
__device__ __forceinline__ float seg_sum(float v, int lane, int end) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const float o = __shfl_down(v, off, 64);
        if (lane + off < end) v += o;
    }
    return v;
}

__global__ __launch_bounds__(256) void wa_wave_kernel(WaveArgs a) {
    __shared__ double red[16 * 2];
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    const float2* pos = a.flag ? (*a.flag == 0 ? a.pos_other : a.pos) : a.pos;
    double acc[2] = {0.0, 0.0};
    if (w < a.num_waves) {
        const int4 wv = a.waves[w];
        const int pin_begin = wv.x, npin = wv.y, net_begin = wv.z, nnet = wv.w;
        // segment heads: lane j (< nnet) marks the first pin of net j
        unsigned long long mask = 0ull;
        {
            int s = lane < nnet ? a.net_start[net_begin + lane] - pin_begin : 64;
            const bool head = s < npin;             // an empty last net has no pin to mark
            // OR of (1 << s) over the head lanes
            unsigned long long bit = head ? (1ull << s) : 0ull;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const unsigned lo = __shfl_xor((unsigned)bit, off, 64), hi = __shfl_xor((unsigned)(bit >> 32), off, 64);
                bit |= ((unsigned long long)hi << 32) | lo;
            }
            mask = bit;
        }
        const bool live = lane < npin;
        const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
        const int seg_start = 63 - __clzll(mask & upto);
        const unsigned long long above = mask & ~upto;
        const int seg_end = above ? __ffsll(above) - 1 : npin;
        const int p = pin_begin + lane;
        float2 x = make_float2(0.f, 0.f);
        if (live) {
            const int c = a.pin_cell[p];
            x = a.pin_off[p];
            if (c < a.num_mov) {
                const float2 q = pos[c];
                x.x += q.x;
                x.y += q.y;
            }
        }
        const float ig = 1.0f / *a.gamma;
        const float hx = __shfl(seg_max(x.x, lane, seg_end), seg_start, 64);
        const float hy = __shfl(seg_max(x.y, lane, seg_end), seg_start, 64);
        const float lx = __shfl(seg_min(x.x, lane, seg_end), seg_start, 64);
        const float ly = __shfl(seg_min(x.y, lane, seg_end), seg_start, 64);
        const float epx = live ? expf((x.x - hx) * ig) : 0.f, epy = live ? expf((x.y - hy) * ig) : 0.f;
        const float enx = live ? expf((lx - x.x) * ig) : 0.f, eny = live ? expf((ly - x.y) * ig) : 0.f;
        const float spx = __shfl(seg_sum(epx, lane, seg_end), seg_start, 64);
        const float spy = __shfl(seg_sum(epy, lane, seg_end), seg_start, 64);
        const float snx = __shfl(seg_sum(enx, lane, seg_end), seg_start, 64);
        const float sny = __shfl(seg_sum(eny, lane, seg_end), seg_start, 64);
        const float xpx = __shfl(seg_sum(x.x * epx, lane, seg_end), seg_start, 64) / spx;
        const float xpy = __shfl(seg_sum(x.y * epy, lane, seg_end), seg_start, 64) / spy;
        const float xnx = __shfl(seg_sum(x.x * enx, lane, seg_end), seg_start, 64) / snx;
        const float xny = __shfl(seg_sum(x.y * eny, lane, seg_end), seg_start, 64) / sny;
        if (live) {
            float2 gr;
            gr.x = epx * (1 + (x.x - xpx) * ig) / spx - enx * (1 - (x.x - xnx) * ig) / snx;
            gr.y = epy * (1 + (x.y - xpy) * ig) / spy - eny * (1 - (x.y - xny) * ig) / sny;
            const int slot = a.pin_slot ? a.pin_slot[p] : p;
            if (slot >= 0) a.pin_grad[slot] = gr;
            if (lane == seg_start) {
                acc[0] = (double)((xpx - xnx) + (xpy - xny));
                acc[1] = (double)((hx - lx) + (hy - ly));
            }
        }
    }
    block_sum<2>(acc, red);
    if (threadIdx.x == 0) {
        a.partial[2 * blockIdx.x] = acc[0];
        a.partial[2 * blockIdx.x + 1] = acc[1];
    }
}

// nets of more than 64 pins: one workgroup per net, the drop-in kernel's workgroup-wide path
__global__ __launch_bounds__(WA_BLOCK) void wa_big_nets_kernel(WaArgs a, const int* big, int nbig, double* partial) {
    __shared__ float red[4 * WA_BLOCK];
    __shared__ double dred[16 * 2];
    double acc[2] = {0.0, 0.0};
    if (a.flag && *a.flag == 0) a.pos = a.pos_other;
    wa_big_net_path(a, big[blockIdx.x], acc[0], acc[1], red);
    block_sum<2>(acc, dred);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = acc[0];
        partial[2 * blockIdx.x + 1] = acc[1];
    }
}

int WaPlan::create(const int* d_net_start, int num_nets) {
    std::vector<int> ns(num_nets + 1);
    AP_HIP(hipMemcpy(ns.data(), d_net_start, (num_nets + 1) * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int4> waves;
    std::vector<int> big;
    int cur_pins = 0, cur_begin = -1;
    auto flush = [&](int next_net) {
        if (cur_begin >= 0) waves.push_back(make_int4(ns[cur_begin], cur_pins, cur_begin, next_net - cur_begin));
        cur_begin = -1;
        cur_pins = 0;
    };
    for (int i = 0; i < num_nets; i++) {
        const int deg = ns[i + 1] - ns[i];
        if (deg > 64) {
            flush(i);
            big.push_back(i);
            continue;
        }
        if (cur_begin < 0 || cur_pins + deg > 64 || i - cur_begin >= 64) {
            flush(i);
            cur_begin = i;
        }
        cur_pins += deg;
    }
    flush(num_nets);
    num_waves = (int)waves.size();
    num_big = (int)big.size();
    if (num_waves) {
        AP_HIP(hipMalloc(&d_waves, waves.size() * sizeof(int4)));
        AP_HIP(hipMemcpy(d_waves, waves.data(), waves.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    if (num_big) {
        AP_HIP(hipMalloc(&d_big, big.size() * sizeof(int)));
        AP_HIP(hipMemcpy(d_big, big.data(), big.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    return AP_OK;
}

void WaPlan::destroy() {
    if (d_waves) (void)hipFree(d_waves);
    if (d_big) (void)hipFree(d_big);
    d_waves = nullptr;
    d_big = nullptr;
}

int WaPlan::launch(const WaArgs& a, double* partial, hipStream_t s) const {
    if (num_waves) {
        WaveArgs w{a.pos, a.pos_other, a.flag, a.pin_cell, a.pin_off, a.net_start, d_waves, num_waves, a.num_mov,
                   a.gamma, a.pin_slot, a.pin_grad, partial};
        hipLaunchKernelGGL(wa_wave_kernel, dim3(wave_blocks()), dim3(256), 0, s, w);
        AP_LAUNCHED();
    }
    if (num_big) {
        hipLaunchKernelGGL(wa_big_nets_kernel, dim3(num_big), dim3(WA_BLOCK), 0, s, a, d_big, num_big,
                           partial + 2LL * wave_blocks());
        AP_LAUNCHED();
    }
    return AP_OK;
}

}  // namespace ap
