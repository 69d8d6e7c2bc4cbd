// HIP kernels of the GP iteration for gfx950 and the drop-in C ABI (include/autoplace_gp.h, part 1).
//
//   WA        wa_net_kernel: a workgroup owns the nets whose first pin falls in its 256-pin window. Pin loads
//             (cell, offset, position gather) are one per lane, coalesced; the pins are staged in LDS; one lane per
//             net reduces max / min and the exponential sums over its pins in pin order (the same order as a
//             sequential pass, so the per-net values are deterministic); one lane per pin writes its gradient.
//             Block partial sums of (WA, HPWL) in FP64. wa_cell_sum_kernel: per-cell sum over the cell CSR.
//   step      Nesterov step and select: one node (two coordinates) per lane, 8-byte loads and stores.
//   control   one lane, double precision, the decision order of the reference's control kernel.
// The drop-in density map and gather are the GP engine's tile binning and gather (gp_engine.hip).
#include <stdarg.h>
#include <stdio.h>

#include <initializer_list>
#include <mutex>
#include <vector>

#include "gp_common.h"
#include "gp_control.h"
#include "gp_kernels.h"
#include "gp_wa_common.h"

namespace ap {

static thread_local char g_err[512];

void set_error(const char* fmt, ...) {
    va_list ap_;
    va_start(ap_, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap_);
    va_end(ap_);
}

// ------------------------------------------------------------------------------------------- WA wirelength

__global__ __launch_bounds__(WA_BLOCK) void wa_net_kernel(WaArgs a) {
    if (a.flag && *a.flag == 0) a.pos = a.pos_other;
    __shared__ float2 sx[WA_CAP];       // pin coordinates
    __shared__ float2 sep[WA_CAP];      // e^{(x - hi) / gamma}
    __shared__ float2 sen[WA_CAP];      // e^{(lo - x) / gamma}
    __shared__ int snet[WA_CAP];        // local net of the pin
    __shared__ float4 sacc[WA_CAP];     // per net: sp.x, sp.y, sn.x, sn.y
    __shared__ float4 smean[WA_CAP];    // per net: xp.x, xp.y, xn.x, xn.y
    __shared__ int sbound[2];
    __shared__ double sred[16 * 2];

    const int t = threadIdx.x;
    const long long win0 = (long long)blockIdx.x * WA_SPAN;
    if (t < 2) {
        long long key = win0 + (t ? WA_SPAN : 0);
        int k = key > 0x7fffffff ? 0x7fffffff : (int)key;
        sbound[t] = a.block_net ? a.block_net[blockIdx.x + t] : lower_bound(a.net_start, 0, a.num_nets, k);
    }
    __syncthreads();
    const int net_lo = sbound[0], net_hi = sbound[1];
    const float ig = 1.0f / *a.gamma;
    double wa_acc = 0, hpwl_acc = 0;

    int net_cur = net_lo;
    while (net_cur < net_hi) {
        const int base = a.net_start[net_cur];
        // nets [net_cur, net_stop) whose pins fit the stage
        __syncthreads();
        if (t == 0) {
            // largest stop in [net_cur, min(net_hi, net_cur + WA_CAP)] with pins <= WA_CAP (the cap on nets keeps
            // the per-net LDS arrays in bounds when nets are empty)
            int lo = net_cur + 1, hi = min(net_hi, net_cur + WA_CAP) + 1;
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (a.net_start[mid] - base <= WA_CAP) lo = mid + 1;
                else hi = mid;
            }
            sbound[0] = lo - 1;
        }
        __syncthreads();
        const int net_stop = sbound[0];
        if (net_stop == net_cur) {            // a single net larger than the stage
            wa_big_net_path(a, net_cur, wa_acc, hpwl_acc, (float*)sx);
            __syncthreads();
            net_cur++;
            continue;
        }
        const int npin = a.net_start[net_stop] - base, nnet = net_stop - net_cur;
        for (int q = t; q < npin; q += WA_BLOCK) sx[q] = pin_xy(a.pos, a.pin_cell, a.pin_off, base + q, a.num_mov);
        __syncthreads();
        // nets go round-robin over the four wavefronts (net 4k + w to wavefront w), so a long net holds up only the
        // quarter of the nets that share its wavefront
        for (int j0 = 0; j0 < nnet; j0 += WA_BLOCK) {
            const int j = j0 + (t & 63) * (WA_BLOCK / 64) + (t >> 6);
            if (j >= nnet) continue;
            const int s = a.net_start[net_cur + j] - base, e = a.net_start[net_cur + j + 1] - base;
            float hx = -1e30f, hy = -1e30f, lx = 1e30f, ly = 1e30f;
            for (int q = s; q < e; q++) {
                float2 x = sx[q];
                hx = fmaxf(hx, x.x); hy = fmaxf(hy, x.y); lx = fminf(lx, x.x); ly = fminf(ly, x.y);
            }
            float spx = 0, spy = 0, snx = 0, sny = 0, xpx = 0, xpy = 0, xnx = 0, xny = 0;
            for (int q = s; q < e; q++) {
                float2 x = sx[q];
                float epx = expf((x.x - hx) * ig), epy = expf((x.y - hy) * ig);
                float enx = expf((lx - x.x) * ig), eny = expf((ly - x.y) * ig);
                sep[q] = make_float2(epx, epy);
                sen[q] = make_float2(enx, eny);
                snet[q] = j;
                spx += epx; spy += epy; snx += enx; sny += eny;
                xpx += x.x * epx; xpy += x.y * epy; xnx += x.x * enx; xny += x.y * eny;
            }
            if (e > s) {
                xpx /= spx; xpy /= spy; xnx /= snx; xny /= sny;
                sacc[j] = make_float4(spx, spy, snx, sny);
                smean[j] = make_float4(xpx, xpy, xnx, xny);
                float wa = (xpx - xnx) + (xpy - xny), hp = (hx - lx) + (hy - ly);
                if (a.out) a.out[net_cur + j] = make_float2(wa, hp);
                wa_acc += wa;
                hpwl_acc += hp;
            } else if (a.out) {
                a.out[net_cur + j] = make_float2(0.f, 0.f);
            }
        }
        __syncthreads();
        for (int q = t; q < npin; q += WA_BLOCK) {
            const int j = snet[q];
            const float2 x = sx[q], ep = sep[q], en = sen[q];
            const float4 acc = sacc[j], m = smean[j];
            float2 gr;
            gr.x = ep.x * (1 + (x.x - m.x) * ig) / acc.x - en.x * (1 - (x.x - m.z) * ig) / acc.z;
            gr.y = ep.y * (1 + (x.y - m.y) * ig) / acc.y - en.y * (1 - (x.y - m.w) * ig) / acc.w;
            const int slot = a.pin_slot ? a.pin_slot[base + q] : base + q;
            if (slot >= 0) a.pin_grad[slot] = gr;
        }
        net_cur = net_stop;
    }
    if (a.partial) {
        double v[2] = {wa_acc, hpwl_acc};
        block_sum<2>(v, sred);
        if (t == 0) {
            a.partial[2 * blockIdx.x] = v[0];
            a.partial[2 * blockIdx.x + 1] = v[1];
        }
    }
}

// A cell's WA gradient: the gradients of its pins (pin order in pin_grad, the cell's pins listed by the cell CSR),
// summed in CSR order, one cell per lane; four pins' index and gradient loads are issued before the first of them is
// added (the same sums as one pin at a time, with a quarter of the dependent memory waits).
__global__ __launch_bounds__(256) void wa_cell_sum_kernel(const float2* pin_grad, const int* cell_start,
                                                          const int* cell_pin, int num_mov, float2* grad) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= num_mov) return;
    const int k1 = cell_start[c + 1];
    int k = cell_start[c];
    float2 sum = make_float2(0.f, 0.f);
    for (; k < k1; k += 4) {
        int p[4];
#pragma unroll
        for (int j = 0; j < 4; j++) p[j] = k + j < k1 ? cell_pin[k + j] : -1;
        float2 q[4];
#pragma unroll
        for (int j = 0; j < 4; j++) q[j] = p[j] >= 0 ? pin_grad[p[j]] : make_float2(0.f, 0.f);
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (p[j] >= 0) {
                sum.x += q[j].x;
                sum.y += q[j].y;
            }
    }
    grad[c] = sum;
}

__global__ __launch_bounds__(256) void wa_block_table_kernel(const int* net_start, int num_nets, int blocks,
                                                             int* block_net) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > blocks) return;
    const long long key = (long long)b * WA_SPAN;
    block_net[b] = lower_bound(net_start, 0, num_nets, key > 0x7fffffff ? 0x7fffffff : (int)key);
}

int launch_wa_block_table(const int* net_start, int num_nets, int blocks, int* block_net, hipStream_t s) {
    hipLaunchKernelGGL(wa_block_table_kernel, dim3(grid_of((long long)blocks + 1, 256)), dim3(256), 0, s, net_start,
                       num_nets, blocks, block_net);
    AP_LAUNCHED();
    return AP_OK;
}

int launch_wa_cell_sum(const float2* pin_grad, const int* cell_start, const int* cell_pin, int num_mov, float2* grad,
                       hipStream_t s) {
    if (num_mov <= 0) return AP_OK;
    hipLaunchKernelGGL(wa_cell_sum_kernel, dim3(grid_of(num_mov, 256)), dim3(256), 0, s, pin_grad, cell_start,
                       cell_pin, num_mov, grad);
    AP_LAUNCHED();
    return AP_OK;
}

int launch_wa(const WaArgs& a, int num_pins, hipStream_t s) {
    if (a.num_nets <= 0) return AP_OK;
    unsigned grid = grid_of((long long)num_pins + 1, WA_SPAN);   // +1: a trailing empty net still has an owner
    hipLaunchKernelGGL(wa_net_kernel, dim3(grid), dim3(WA_BLOCK), 0, s, a);
    AP_LAUNCHED();
    return AP_OK;
}

// ------------------------------------------------------------------------------------------- Nesterov, select
// (the density map and the gather of the drop-in entry points run on the engine's tile binning and gather:
// gp_engine.hip, "drop-in density map and gather")

// u' = v - alpha g, v' = clamp(u' + coef (u' - u), lo, hi). PAIRS: one node (its x and y) per lane with 8-byte
// loads, an odd trailing coordinate by the first lane; else (a pointer not 8-byte aligned) one coordinate per lane.
template <bool PAIRS>
__global__ __launch_bounds__(256) void nesterov_pair_kernel(const float* u, const float* v, const float* g,
                                                            const float* lo, const float* hi, const float* alpha_p,
                                                            const float* coef_p, int n, float* u_next,
                                                            float* v_next) {
    const float alpha = *alpha_p, coef = *coef_p;
    auto step = [&](float uu, float vv, float gg, float l, float h, float& un, float& vn) {
        un = vv - alpha * gg;
        vn = fminf(fmaxf(un + coef * (un - uu), l), h);
    };
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (!PAIRS) {
        if (k < n) step(u[k], v[k], g[k], lo[k], hi[k], u_next[k], v_next[k]);
        return;
    }
    if (k < n / 2) {
        const float2 uu = ((const float2*)u)[k], vv = ((const float2*)v)[k], gg = ((const float2*)g)[k];
        const float2 l = ((const float2*)lo)[k], h = ((const float2*)hi)[k];
        float2 un, vn;
        step(uu.x, vv.x, gg.x, l.x, h.x, un.x, vn.x);
        step(uu.y, vv.y, gg.y, l.y, h.y, un.y, vn.y);
        ((float2*)u_next)[k] = un;
        ((float2*)v_next)[k] = vn;
    }
    if ((n & 1) && k == 0) step(u[n - 1], v[n - 1], g[n - 1], lo[n - 1], hi[n - 1], u_next[n - 1], v_next[n - 1]);
}

// an accepted step moves u, v, g to the pass's point; a best one is also kept (PAIRS as above)
template <bool PAIRS>
__global__ __launch_bounds__(256) void select_pair_kernel(const double* S, float* u, float* v, float* g,
                                                          const float* u_next, const float* v_next,
                                                          const float* g_next, float* best, int n) {
    if (S[ACCEPT] == 0) return;                      // uniform: the whole launch returns
    const bool keep = S[IS_BEST] != 0;
    const int k = blockIdx.x * 256 + threadIdx.x;
    auto one = [&](int i) {
        u[i] = u_next[i];
        v[i] = v_next[i];
        g[i] = g_next[i];
        if (keep) best[i] = v_next[i];
    };
    if (!PAIRS) {
        if (k < n) one(k);
        return;
    }
    if (k < n / 2) {
        const float2 vn = ((const float2*)v_next)[k];
        ((float2*)u)[k] = ((const float2*)u_next)[k];
        ((float2*)v)[k] = vn;
        ((float2*)g)[k] = ((const float2*)g_next)[k];
        if (keep) ((float2*)best)[k] = vn;
    }
    if ((n & 1) && k == 0) one(n - 1);
}

static bool aligned8(std::initializer_list<const void*> ps) {
    for (const void* p : ps)
        if ((uintptr_t)p & 7) return false;
    return true;
}

// ------------------------------------------------------------------------------------------- control

__global__ void control_kernel(double* S, float* kp, const double* c, const float* moved, const float* turned,
                               const float* hpwl, const float* ovf, const float* wa) {
    if (threadIdx.x == 0 && blockIdx.x == 0) control_step(S, kp, c, *moved, *turned, *hpwl, *ovf, *wa);
}

// ------------------------------------------------------------------------------------------- exact HPWL

// HPWL over every net (evaluator.py:4-8): one lane per net, FP64 per-workgroup partials, then one workgroup sums the
// partials in a fixed order.
__global__ __launch_bounds__(256) void net_hpwl_kernel(const float2* pos, const int* pin_cell, const float2* pin_off,
                                                       const int* net_start, int num_nets, int num_mov, double* part) {
    __shared__ double red[16];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double h[1] = {0.0};
    if (i < num_nets) {
        const int p0 = net_start[i], p1 = net_start[i + 1];
        if (p1 > p0) {
            float hx = -INFINITY, hy = -INFINITY, lx = INFINITY, ly = INFINITY;
            for (int p = p0; p < p1; p++) {
                float2 x = pin_xy(pos, pin_cell, pin_off, p, num_mov);
                hx = fmaxf(hx, x.x); hy = fmaxf(hy, x.y); lx = fminf(lx, x.x); ly = fminf(ly, x.y);
            }
            h[0] = ((double)hx - lx) + ((double)hy - ly);
        }
    }
    block_sum<1>(h, red);
    if (threadIdx.x == 0) part[blockIdx.x] = h[0];
}

__global__ __launch_bounds__(256) void sum_partials_kernel(const double* part, int n, double* out) {
    __shared__ double red[16];
    double v[1] = {0.0};
    for (int k = threadIdx.x; k < n; k += blockDim.x) v[0] += part[k];
    block_sum<1>(v, red);
    if (threadIdx.x == 0) *out = v[0];
}

}  // namespace ap

using namespace ap;

// ------------------------------------------------------------------------------------------- C ABI, part 1

extern "C" {

const char* ap_last_error(void) { return ap::g_err; }
int ap_version(void) { return 100; }

int ap_wa_wirelength(const float* pos, const int* pin_cell, const float* pin_off, const int* net_start,
                     const int* cell_start, const int* cell_pin, int num_nets, int num_mov, const float* gamma,
                     float* pin_grad, float* out, float* grad, int num_pins, void* stream) {
    AP_REQUIRE(num_nets >= 0 && num_mov >= 0 && num_pins >= 0, "wa: negative size");
    AP_REQUIRE(num_nets == 0 || (net_start && gamma && pin_grad), "wa: null pointer");
    hipStream_t s = (hipStream_t)stream;
    WaArgs a{(const float2*)pos, pin_cell, (const float2*)pin_off, net_start, num_nets, num_mov, gamma,
             (float2*)pin_grad, (float2*)out, nullptr, nullptr, nullptr, nullptr, nullptr};
    int st = launch_wa(a, num_pins, s);
    if (st) return st;
    return launch_wa_cell_sum((const float2*)pin_grad, cell_start, cell_pin, num_mov, (float2*)grad, s);
}

int ap_nesterov_step(const float* u, const float* v, const float* g, const float* lo, const float* hi,
                     const float* alpha, const float* coef, int n, float* u_next, float* v_next, void* stream) {
    AP_REQUIRE(n >= 0, "nesterov_step: negative size");
    if (n == 0) return AP_OK;
    AP_REQUIRE(u && v && g && lo && hi && alpha && coef && u_next && v_next, "nesterov_step: null pointer");
    if (aligned8({u, v, g, lo, hi, u_next, v_next}))
        hipLaunchKernelGGL(nesterov_pair_kernel<true>, dim3(grid_of(n / 2 + 1, 256)), dim3(256), 0,
                           (hipStream_t)stream, u, v, g, lo, hi, alpha, coef, n, u_next, v_next);
    else
        hipLaunchKernelGGL(nesterov_pair_kernel<false>, dim3(grid_of(n, 256)), dim3(256), 0, (hipStream_t)stream, u,
                           v, g, lo, hi, alpha, coef, n, u_next, v_next);
    AP_LAUNCHED();
    return AP_OK;
}

int ap_control(double* S, float* kp, const double* c, const float* moved, const float* turned, const float* hpwl,
               const float* ovf, const float* wa, void* stream) {
    AP_REQUIRE(S && kp && c && moved && turned && hpwl && ovf && wa, "control: null pointer");
    hipLaunchKernelGGL(control_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, S, kp, c, moved, turned, hpwl, ovf, wa);
    AP_LAUNCHED();
    return AP_OK;
}

int ap_net_hpwl(const float* pos, const int* pin_cell, const float* pin_off, const int* net_start, int num_nets,
                int num_mov, double* work, double* out, void* stream) {
    AP_REQUIRE(num_nets >= 0 && out && (num_nets == 0 || (work && net_start)), "net_hpwl: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    const int nblk = num_nets > 0 ? (int)grid_of(num_nets, 256) : 0;
    if (nblk) {
        hipLaunchKernelGGL(net_hpwl_kernel, dim3(nblk), dim3(256), 0, s, (const float2*)pos, pin_cell,
                           (const float2*)pin_off, net_start, num_nets, num_mov, work);
        AP_LAUNCHED();
    }
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, work, nblk, out);
    AP_LAUNCHED();
    return AP_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------ exact HPWL without an allocation
// ap_net_hpwl_noalloc keeps its FP64 workgroup partials in this file-scope device array instead of a caller's work
// buffer: one partial per 256 nets, so at most AP_NOALLOC_HPWL_PARTIALS * 256 nets. Each device that runs the call
// holds its own copy of the array until the process ends. In host code `scratch` only names the symbol; the device
// address of each device's copy comes from hipGetSymbolAddress, taken once per device on its first call.
static __device__ __align__(256) double scratch[AP_NOALLOC_HPWL_PARTIALS];

namespace {
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_of_device;   // device index -> address of that device's `scratch`

// The current device's address of `scratch` (taken on the device's first call, with the device current).
int scratch_address(double** out) {
    int dev = -1;
    AP_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_of_device.size() <= dev) g_scratch_of_device.resize(dev + 1, nullptr);
    if (!g_scratch_of_device[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("net_hpwl_noalloc: no device address of the scratch array on device %d: %s", dev,
                      hipGetErrorString(e));
            return AP_ERR_HIP;
        }
        g_scratch_of_device[dev] = (double*)p;
    }
    *out = g_scratch_of_device[dev];
    return AP_OK;
}
}  // namespace

extern "C" {

int ap_net_hpwl_noalloc(const float* pos, const int* pin_cell, const float* pin_off, const int* net_start,
                        int num_nets, int num_mov, double* out, void* stream) {
    AP_REQUIRE(num_nets >= 0 && out && (num_nets == 0 || (net_start && pos && pin_cell && pin_off)),
               "net_hpwl_noalloc: bad arguments");
    const long long nblk = num_nets > 0 ? (long long)grid_of(num_nets, 256) : 0;
    AP_REQUIRE(nblk <= AP_NOALLOC_HPWL_PARTIALS,
               "net_hpwl_noalloc: %d nets need %lld workgroup partials, above its fixed scratch of %d (at most %lld "
               "nets): use ap_net_hpwl with a work buffer",
               num_nets, nblk, AP_NOALLOC_HPWL_PARTIALS, (long long)AP_NOALLOC_HPWL_PARTIALS * 256);
    double* work = nullptr;
    const int st = scratch_address(&work);
    if (st) return st;
    hipStream_t s = (hipStream_t)stream;
    // every partial the sum reads is written by its workgroup first: nothing to clear between calls
    if (nblk) {
        hipLaunchKernelGGL(net_hpwl_kernel, dim3((unsigned)nblk), dim3(256), 0, s, (const float2*)pos, pin_cell,
                           (const float2*)pin_off, net_start, num_nets, num_mov, work);
        AP_LAUNCHED();
    }
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, work, (int)nblk, out);
    AP_LAUNCHED();
    return AP_OK;
}

int ap_select_step(const double* S, float* u, float* v, float* g, const float* u_next, const float* v_next,
                   const float* g_next, float* best, int n, void* stream) {
    AP_REQUIRE(n >= 0, "select: negative size");
    if (n == 0) return AP_OK;
    AP_REQUIRE(S && u && v && g && u_next && v_next && g_next && best, "select: null pointer");
    if (aligned8({u, v, g, u_next, v_next, g_next, best}))
        hipLaunchKernelGGL(select_pair_kernel<true>, dim3(grid_of(n / 2 + 1, 256)), dim3(256), 0,
                           (hipStream_t)stream, S, u, v, g, u_next, v_next, g_next, best, n);
    else
        hipLaunchKernelGGL(select_pair_kernel<false>, dim3(grid_of(n, 256)), dim3(256), 0, (hipStream_t)stream, S, u,
                           v, g, u_next, v_next, g_next, best, n);
    AP_LAUNCHED();
    return AP_OK;
}

}  // extern "C"
