// Device helpers shared by the WA kernels (drop-in and engine).
#pragma once
#include "gp_common.h"
#include "gp_kernels.h"

namespace ap {

// smallest i in [lo, hi) with a[i] >= key (hi if none)
__device__ __forceinline__ int lower_bound(const int* a, int lo, int hi, int key) {
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ float2 pin_xy(const float2* pos, const int* pin_cell, const float2* pin_off, int p,
                                         int num_mov) {
    int c = pin_cell[p];
    float2 o = pin_off[p];
    if (c < num_mov) {
        float2 q = pos[c];
        o.x += q.x;
        o.y += q.y;
    }
    return o;
}

// One net whose pins do not fit the LDS stage: the whole workgroup loops over its pins (three passes over global
// memory), with fixed-order block reductions. Rare (degree > WA_CAP).
__device__ __forceinline__ void wa_big_net_path(const WaArgs& a, int net, double& wa_acc, double& hpwl_acc, float* red /* [4*256] */) {
    const int p0 = a.net_start[net], p1 = a.net_start[net + 1], t = threadIdx.x, nt = blockDim.x;
    const float ig = 1.0f / *a.gamma;
    float hx = -1e30f, hy = -1e30f, lx = 1e30f, ly = 1e30f;
    for (int p = p0 + t; p < p1; p += nt) {
        float2 x = pin_xy(a.pos, a.pin_cell, a.pin_off, p, a.num_mov);
        hx = fmaxf(hx, x.x); hy = fmaxf(hy, x.y); lx = fminf(lx, x.x); ly = fminf(ly, x.y);
    }
    red[t] = hx; red[nt + t] = hy; red[2 * nt + t] = lx; red[3 * nt + t] = ly;
    __syncthreads();
    for (int s = nt / 2; s > 0; s >>= 1) {
        if (t < s) {
            red[t] = fmaxf(red[t], red[t + s]);
            red[nt + t] = fmaxf(red[nt + t], red[nt + t + s]);
            red[2 * nt + t] = fminf(red[2 * nt + t], red[2 * nt + t + s]);
            red[3 * nt + t] = fminf(red[3 * nt + t], red[3 * nt + t + s]);
        }
        __syncthreads();
    }
    hx = red[0]; hy = red[nt]; lx = red[2 * nt]; ly = red[3 * nt];
    __syncthreads();
    float sums[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // sp.x sp.y sn.x sn.y (x ep).x (x ep).y (x en).x (x en).y
    for (int p = p0 + t; p < p1; p += nt) {
        float2 x = pin_xy(a.pos, a.pin_cell, a.pin_off, p, a.num_mov);
        float epx = expf((x.x - hx) * ig), epy = expf((x.y - hy) * ig);
        float enx = expf((lx - x.x) * ig), eny = expf((ly - x.y) * ig);
        sums[0] += epx; sums[1] += epy; sums[2] += enx; sums[3] += eny;
        sums[4] += x.x * epx; sums[5] += x.y * epy; sums[6] += x.x * enx; sums[7] += x.y * eny;
    }
    float tot[8];
    for (int k = 0; k < 8; k += 4) {
        for (int j = 0; j < 4; j++) red[j * nt + t] = sums[k + j];
        __syncthreads();
        for (int s = nt / 2; s > 0; s >>= 1) {
            if (t < s)
                for (int j = 0; j < 4; j++) red[j * nt + t] += red[j * nt + t + s];
            __syncthreads();
        }
        for (int j = 0; j < 4; j++) tot[k + j] = red[j * nt];
        __syncthreads();
    }
    const float xpx = tot[4] / tot[0], xpy = tot[5] / tot[1], xnx = tot[6] / tot[2], xny = tot[7] / tot[3];
    for (int p = p0 + t; p < p1; p += nt) {
        float2 x = pin_xy(a.pos, a.pin_cell, a.pin_off, p, a.num_mov);
        float epx = expf((x.x - hx) * ig), epy = expf((x.y - hy) * ig);
        float enx = expf((lx - x.x) * ig), eny = expf((ly - x.y) * ig);
        float2 gr;
        gr.x = epx * (1 + (x.x - xpx) * ig) / tot[0] - enx * (1 - (x.x - xnx) * ig) / tot[2];
        gr.y = epy * (1 + (x.y - xpy) * ig) / tot[1] - eny * (1 - (x.y - xny) * ig) / tot[3];
        const int slot = a.pin_slot ? a.pin_slot[p] : p;
        if (slot >= 0) a.pin_grad[slot] = gr;
    }
    if (t == 0) {
        float wa = (xpx - xnx) + (xpy - xny), hp = (hx - lx) + (hy - ly);
        if (a.out) a.out[net] = make_float2(wa, hp);
        wa_acc += wa;
        hpwl_acc += hp;
    }
}

}  // namespace ap
