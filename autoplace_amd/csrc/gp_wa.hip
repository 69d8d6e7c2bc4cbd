// WA wirelength of the engine: one lane per pin, every net in an aligned group of 2^k lanes.
//
// The drop-in kernel (gp_kernels.hip) gives each net one lane that loops over the net's pins: the lanes of a
// wavefront then wait for its longest net (degrees 2..50 here, up to --ignore_net_degree), and the per-net phase is
// VALU-bound. Here the nets of a level are sorted once, on the host, into degree classes: a net of d <= 64 pins gets
// a group of G = 2^k >= d lanes (G >= 2), aligned to G inside a wavefront, and each wavefront holds 64 / G nets of one
// class, in net order. Every lane holds one pin (or nothing, past the net's degree); max / min and the four
// exponential sums per axis are butterfly reductions over the group whose steps inside a row of 16 lanes are DPP
// lane permutations (quad_perm, row_half_mirror, row_mirror: plain VALU operands, no LDS crossbar), and the steps
// across rows (G = 32, 64) are shuffles. After a butterfly every lane of the group holds the same bits (each step
// adds the same two partial results, in either order), so no broadcast is needed and the results are
// deterministic. Nets of more than 64 pins take a workgroup-wide path.
#include <algorithm>
#include <cstring>
#include <vector>

#include <rocprim/device/device_scan.hpp>

#include "gp_common.h"
#include "gp_kernels.h"
#include "gp_wa.h"
#include "gp_wa_common.h"

namespace ap {

// DPP lane permutations (gfx9 encodings of the dpp_ctrl operand)
constexpr int DPP_QUAD_XOR1 = 0xB1;        // quad_perm [1, 0, 3, 2]
constexpr int DPP_QUAD_XOR2 = 0x4E;        // quad_perm [2, 3, 0, 1]
constexpr int DPP_ROW_HALF_MIRROR = 0x141; // lane i of each 8 <- lane 7 - i
constexpr int DPP_ROW_MIRROR = 0x140;      // lane i of each 16 <- lane 15 - i

// Every pattern used reads a lane of the same row, so no lane takes the bound value; with bound_ctrl set the compiler
// folds the move into the operation that consumes it (v_max_f32_dpp, v_add_f32_dpp) instead of a separate mov.
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

struct MaxOp {
    __device__ __forceinline__ float operator()(float a, float b) const { return fmaxf(a, b); }
};
struct MinOp {
    __device__ __forceinline__ float operator()(float a, float b) const { return fminf(a, b); }
};
struct AddOp {
    __device__ __forceinline__ float operator()(float a, float b) const { return a + b; }
};

// reduction over the aligned group of 2^LG lanes, the result in every lane of the group. After the two quad steps
// every lane of a quad holds the quad's value, so the mirror partner (in the other quad of the half row, the other
// half of the row) holds the other half's value.
template <int LG, typename Op>
__device__ __forceinline__ float group_reduce(float v, Op op) {
    v = op(v, dpp<DPP_QUAD_XOR1>(v));
    if (LG >= 2) v = op(v, dpp<DPP_QUAD_XOR2>(v));
    if (LG >= 3) v = op(v, dpp<DPP_ROW_HALF_MIRROR>(v));
    if (LG >= 4) v = op(v, dpp<DPP_ROW_MIRROR>(v));
    if (LG >= 5) v = op(v, __shfl_xor(v, 16, 64));
    if (LG >= 6) v = op(v, __shfl_xor(v, 32, 64));
    return v;
}

// e^t for t <= 0, a pin's weight against its net's extreme pin: 2^(t log2 e) on the hardware exp2, two instructions
// against the thirteen of the library's expf (which splits t log2 e for a result within 1 ulp for any t). The
// product's rounding costs a relative error of at most |t| 2^-24, which the weight e^t itself shrinks (|t| e^t <=
// 1/e), so every sum of weights and weighted positions keeps a relative error of about 2^-23.
__device__ __forceinline__ float exp_weight(float t) { return __builtin_amdgcn_exp2f(t * 1.44269504f); }

// q: the position of the pin's cell (zero for a fixed pin or an empty lane), loaded by the caller; ig = 1 / gamma,
// divided once per thread by the caller. OUT: the group's first lane also writes the net's (WA, HPWL) to a.out[net]
// ((0, 0) for a net without pins)
template <int LG, bool OUT = false>
__device__ __forceinline__ void wa_group(const WaveArgs& a, float2 q, const int4& ln, int lane, float ig,
                                         double (&acc)[2], int net = -1) {
    constexpr int G = 1 << LG;
    const int sub = lane & (G - 1);
    const int c = ln.x, slot = ln.y;
    const bool live = c >= 0;
    float2 x = make_float2(__int_as_float(ln.z), __int_as_float(ln.w));
    x.x += q.x;
    x.y += q.y;
    const float NEG = -3.402823466e38f, POS = 3.402823466e38f;
    const float hx = group_reduce<LG>(live ? x.x : NEG, MaxOp()), hy = group_reduce<LG>(live ? x.y : NEG, MaxOp());
    const float lx = group_reduce<LG>(live ? x.x : POS, MinOp()), ly = group_reduce<LG>(live ? x.y : POS, MinOp());
    const float epx = live ? exp_weight((x.x - hx) * ig) : 0.f, epy = live ? exp_weight((x.y - hy) * ig) : 0.f;
    const float enx = live ? exp_weight((lx - x.x) * ig) : 0.f, eny = live ? exp_weight((ly - x.y) * ig) : 0.f;
    // the sums are >= 1 (the extreme pin contributes e^0): hardware reciprocals (1 ulp) replace the 8 divisions
    const float rpx = __builtin_amdgcn_rcpf(group_reduce<LG>(epx, AddOp()));
    const float rpy = __builtin_amdgcn_rcpf(group_reduce<LG>(epy, AddOp()));
    const float rnx = __builtin_amdgcn_rcpf(group_reduce<LG>(enx, AddOp()));
    const float rny = __builtin_amdgcn_rcpf(group_reduce<LG>(eny, AddOp()));
    const float xpx = group_reduce<LG>(x.x * epx, AddOp()) * rpx, xpy = group_reduce<LG>(x.y * epy, AddOp()) * rpy;
    const float xnx = group_reduce<LG>(x.x * enx, AddOp()) * rnx, xny = group_reduce<LG>(x.y * eny, AddOp()) * rny;
    if (live) {
        float2 gr;
        gr.x = epx * (1 + (x.x - xpx) * ig) * rpx - enx * (1 - (x.x - xnx) * ig) * rnx;
        gr.y = epy * (1 + (x.y - xpy) * ig) * rpy - eny * (1 - (x.y - xny) * ig) * rny;
        if (slot >= 0) a.pin_grad[slot] = gr;
        if (sub == 0) {
            acc[0] += (double)((xpx - xnx) + (xpy - xny));
            acc[1] += (double)((hx - lx) + (hy - ly));
        }
    }
    if (OUT && sub == 0 && net >= 0)
        a.out[net] = live ? make_float2((xpx - xnx) + (xpy - xny), (hx - lx) + (hy - ly)) : make_float2(0.f, 0.f);
}

// Each wavefront of the launch runs PER consecutive wavefronts of the plan, their lane data and cell positions loaded
// before the first is reduced (their memory waits overlap). Each plan wavefront's (WA, HPWL) sums go to its own
// partial (a wavefront sum in a fixed order), so the partials do not depend on PER or the workgroup size.
template <bool OUT = false, int PER = WA_PLAN_PER_WAVE>
__device__ __forceinline__ void wa_wave_body(const WaveArgs& a, int wg, int nwg) {
    const int lane = threadIdx.x & 63;
    const int w0 = (xcd_remap(wg, nwg) * (int)(blockDim.x >> 6) + (threadIdx.x >> 6)) * PER;
    const float2* pos = a.flag ? (*a.flag == 0 ? a.pos_other : a.pos) : a.pos;
    double acc[PER][2];
    int4 ln[PER];
    int lg[PER], nt[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        acc[k][0] = acc[k][1] = 0.0;
        const bool in = w0 + k < a.num_waves;       // uniform over the wavefront
        ln[k] = in ? a.lanes[(long long)(w0 + k) * 64 + lane] : make_int4(-1, -1, 0, 0);
        lg[k] = in ? a.lg[w0 + k] : 0;
        nt[k] = OUT && in ? a.lane_net[(long long)(w0 + k) * 64 + lane] : -1;
    }
    float2 q[PER];
#pragma unroll
    for (int k = 0; k < PER; k++)
        q[k] = ln[k].x >= 0 && ln[k].x < a.num_mov ? pos[ln[k].x] : make_float2(0.f, 0.f);
    const float ig = 1.0f / *a.gamma;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        switch (lg[k]) {   // uniform over the wavefront
            case 0: break;
            case 1: wa_group<1, OUT>(a, q[k], ln[k], lane, ig, acc[k], nt[k]); break;
            case 2: wa_group<2, OUT>(a, q[k], ln[k], lane, ig, acc[k], nt[k]); break;
            case 3: wa_group<3, OUT>(a, q[k], ln[k], lane, ig, acc[k], nt[k]); break;
            case 4: wa_group<4, OUT>(a, q[k], ln[k], lane, ig, acc[k], nt[k]); break;
            case 5: wa_group<5, OUT>(a, q[k], ln[k], lane, ig, acc[k], nt[k]); break;
            default: wa_group<6, OUT>(a, q[k], ln[k], lane, ig, acc[k], nt[k]); break;
        }
    }
    if (OUT) return;                                 // per-net outputs only: no partials
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const double wa = wave_sum(acc[k][0]), hp = wave_sum(acc[k][1]);
        if (lane == 0 && w0 + k < a.num_waves) {
            a.partial[2 * (w0 + k)] = wa;
            a.partial[2 * (w0 + k) + 1] = hp;
        }
    }
}

template <int PER>
__global__ __launch_bounds__(512) void wa_wave_kernel(WaveArgs a) {
    wa_wave_body<false, PER>(a, blockIdx.x, gridDim.x);
}
// the drop-in's planned WA: also every net's (WA, HPWL)
template <int PER>
__global__ __launch_bounds__(512) void wa_wave_nets_kernel(WaveArgs a) {
    wa_wave_body<true, PER>(a, blockIdx.x, gridDim.x);
}

// many levels per launch: workgroup b does map[b] = (level, workgroup, workgroups of the level, 0); a level whose
// control state says stopped (S[STOP] != 0) is skipped
__global__ __launch_bounds__(256) void wa_wave_grouped(const WaveArgs* tab, const int4* map,
                                                       const double* const* S) {
    const int4 m = map[blockIdx.x];
    if (S[m.x][STOP] != 0) return;
    const WaveArgs a = tab[m.x];
    wa_wave_body<false, WA_PLAN_PER_WAVE>(a, m.y, m.z);   // 256-lane workgroups
}

int launch_wa_grouped(const WaveArgs* d_tab, const int4* d_map, int count, const double* const* d_S, hipStream_t s) {
    if (!count) return AP_OK;
    hipLaunchKernelGGL(wa_wave_grouped, dim3(count), dim3(WA_LANES), 0, s, d_tab, d_map, d_S);
    AP_LAUNCHED();
    return AP_OK;
}

WaveArgs WaPlan::wave_args(const WaArgs& a, double* partial) const {
    return WaveArgs{a.pos,     a.pos_other, a.flag, d_lg,  d_lanes, num_waves, a.num_mov, a.gamma, a.pin_grad,
                    partial,   a.out,       d_net};
}

// nets of more than 64 pins: one workgroup per net, the drop-in kernel's workgroup-wide path
__global__ __launch_bounds__(WA_BLOCK) void wa_big_nets_kernel(WaArgs a, const int* big, int nbig, double* partial) {
    __shared__ float red[4 * WA_BLOCK];
    __shared__ double dred[16 * 2];
    double acc[2] = {0.0, 0.0};
    if (a.flag && *a.flag == 0) a.pos = a.pos_other;
    wa_big_net_path(a, big[blockIdx.x], acc[0], acc[1], red);
    if (!partial) return;                            // per-net outputs only
    block_sum<2>(acc, dred);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = acc[0];
        partial[2 * blockIdx.x + 1] = acc[1];
    }
}

// set-up: the lane-ordered copy of the pins' cell, gradient slot and offset
__global__ __launch_bounds__(256) void wa_lanes_fill(const int* lane_pin, long long n, const int* pin_cell,
                                                     const float2* pin_off, const int* pin_slot, int4* lanes) {
    const long long l = (long long)blockIdx.x * 256 + threadIdx.x;
    if (l >= n) return;
    const int p = lane_pin[l];
    int4 v = make_int4(-1, -1, 0, 0);
    if (p >= 0) {
        const float2 o = pin_off[p];
        v = make_int4(pin_cell[p], pin_slot ? pin_slot[p] : p, __float_as_int(o.x), __float_as_int(o.y));
    }
    lanes[l] = v;
}

// The plan, built on the device. Nets are taken in net order; each degree class (2^k lanes per net, k = 1 .. 6)
// keeps one open wavefront of 64 / 2^k nets, emitted when its last net joins it; at the end the classes' partial
// wavefronts follow in class order. So the wavefronts stay close to net order (and to the cells' order) within every
// class. Nets of more than 64 pins go to the block-wide path. Construction, without a host loop over the nets:
//   wa_plan_class   every net's class (0: more than 64 pins) as a one-hot count; an exclusive scan gives each net
//                   its rank within its class and the class sizes;
//   wa_plan_rank    the nets that complete a wavefront (rank + 1 a multiple of 64 / 2^k); their exclusive scan is the
//                   full wavefronts' order; the nets by (class, rank) and the big nets in order;
//   wa_plan_lanes   every net's pins into the lanes of its wavefront: the wavefront of its completing net, or its
//                   class's partial one.
// The pins' data is then copied into lane order once, so that a wavefront reads one contiguous 1 KB block.
struct ClassCount {
    int c[7];                  // [0]: nets of more than 64 pins, [k]: nets of 2^k lanes
};
struct ClassCountPlus {
    __host__ __device__ ClassCount operator()(const ClassCount& a, const ClassCount& b) const {
        ClassCount r;
        for (int k = 0; k < 7; k++) r.c[k] = a.c[k] + b.c[k];
        return r;
    }
};

__device__ __forceinline__ int net_class(int deg) {
    if (deg > 64) return 0;
    int lg = 1;
    while ((1 << lg) < deg) lg++;
    return lg;
}

__global__ __launch_bounds__(256) void wa_plan_class(const int* net_start, int num_nets, unsigned char* cls,
                                                     ClassCount* one) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= num_nets) return;
    const int k = net_class(net_start[i + 1] - net_start[i]);
    cls[i] = (unsigned char)k;
    ClassCount v;
    for (int j = 0; j < 7; j++) v.c[j] = j == k;
    one[i] = v;
}

struct ClassOffsets {
    int off[7];                // start of each class in the nets-by-rank array
};

__global__ __launch_bounds__(256) void wa_plan_rank(const unsigned char* cls, const ClassCount* prefix, int num_nets,
                                                    ClassOffsets o, int* rank, int* completes, int* by_rank,
                                                    int* big) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= num_nets) return;
    const int k = cls[i];
    const int r = ((const int*)&prefix[i])[k];
    rank[i] = r;
    if (k == 0) {
        big[r] = i;
        completes[i] = 0;
        return;
    }
    const int per_wave = 64 >> k;
    completes[i] = (r + 1) % per_wave == 0;
    by_rank[o.off[k] + r] = i;
}

__global__ __launch_bounds__(256) void wa_plan_lanes(const int* net_start, const unsigned char* cls, const int* rank,
                                                     int num_nets, ClassOffsets o, ClassOffsets count,
                                                     ClassOffsets partial_wave, const int* by_rank,
                                                     const int* wave_of_completing, int* lane_pin,
                                                     unsigned char* lg, int* lane_net) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= num_nets) return;
    const int k = cls[i];
    if (k == 0) return;
    const int per_wave = 64 >> k, r = rank[i];
    const int last = (r / per_wave + 1) * per_wave - 1;      // the rank that completes this net's wavefront
    const int wave = last < count.off[k] ? wave_of_completing[by_rank[o.off[k] + last]] : partial_wave.off[k];
    const int base = wave * 64 + (r % per_wave) * (1 << k);
    for (int q = net_start[i], q1 = net_start[i + 1]; q < q1; q++) lane_pin[base + (q - net_start[i])] = q;
    if (r % per_wave == 0) lg[wave] = (unsigned char)k;
    if (lane_net) lane_net[base] = i;                        // the group's first lane, also for a net of no pins
}

__global__ __launch_bounds__(256) void wa_plan_clear(int* a, long long n) {
    const long long l = (long long)blockIdx.x * 256 + threadIdx.x;
    if (l < n) a[l] = -1;
}

int WaPlan::create(const int* d_net_start, int num_nets, const int* d_pin_cell, const float* d_pin_off,
                   const int* d_pin_slot, int num_pins, bool with_nets) {
    num_waves = num_big = 0;
    if (num_nets <= 0) return AP_OK;
    int first = 0, end = 0;
    AP_HIP(hipMemcpy(&first, d_net_start, sizeof(int), hipMemcpyDeviceToHost));
    AP_HIP(hipMemcpy(&end, d_net_start + num_nets, sizeof(int), hipMemcpyDeviceToHost));
    AP_REQUIRE(first >= 0 && end <= num_pins, "wa plan: net_start out of the pin range");
    // scratch of the construction, freed at the end
    unsigned char* cls = nullptr;
    ClassCount *one = nullptr, *prefix = nullptr;
    int *rank = nullptr, *completes = nullptr, *order = nullptr, *by_rank = nullptr, *lane_pin = nullptr;
    void* temp = nullptr;
    std::vector<void*> scratch;
    auto take = [&](void** p, size_t bytes) {
        const hipError_t e = dev_malloc(p, bytes > 0 ? bytes : 1);
        if (e == hipSuccess) scratch.push_back(*p);
        return e;
    };
    auto release = [&]() {
        (void)hipDeviceSynchronize();
        for (void* p : scratch) (void)dev_free(p);
        scratch.clear();
    };
    const dim3 g(grid_of(num_nets, 256)), b(256);
    hipError_t e = hipSuccess;
    if ((e = take((void**)&cls, num_nets)) || (e = take((void**)&one, num_nets * sizeof(ClassCount))) ||
        (e = take((void**)&prefix, num_nets * sizeof(ClassCount))) || (e = take((void**)&rank, num_nets * 4ull)) ||
        (e = take((void**)&completes, num_nets * 4ull)) || (e = take((void**)&order, num_nets * 4ull)) ||
        (e = take((void**)&by_rank, num_nets * 4ull))) {
        release();
        AP_HIP(e);
    }
    hipLaunchKernelGGL(wa_plan_class, g, b, 0, nullptr, d_net_start, num_nets, cls, one);
    ClassCount zero{};
    size_t need = 0, need2 = 0;
    e = rocprim::exclusive_scan(nullptr, need, one, prefix, zero, (size_t)num_nets, ClassCountPlus(), nullptr);
    if (e == hipSuccess)
        e = rocprim::exclusive_scan(nullptr, need2, completes, order, 0, (size_t)num_nets, rocprim::plus<int>(),
                                    nullptr);
    if (e == hipSuccess) e = take(&temp, std::max(need, need2));
    if (e == hipSuccess)
        e = rocprim::exclusive_scan(temp, need, one, prefix, zero, (size_t)num_nets, ClassCountPlus(), nullptr);
    ClassCount last_one{}, last_prefix{};
    if (e == hipSuccess) e = hipMemcpy(&last_one, one + num_nets - 1, sizeof(ClassCount), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(&last_prefix, prefix + num_nets - 1, sizeof(ClassCount), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        release();
        AP_HIP(e);
    }
    ClassOffsets off{}, count{}, partial{};
    int acc = 0;
    for (int k = 0; k < 7; k++) {
        count.off[k] = last_prefix.c[k] + last_one.c[k];
        off.off[k] = acc;
        acc += k ? count.off[k] : 0;
    }
    num_big = count.off[0];
    if (num_big && (e = dev_malloc(&d_big, num_big * sizeof(int))) != hipSuccess) {
        release();
        AP_HIP(e);
    }
    hipLaunchKernelGGL(wa_plan_rank, g, b, 0, nullptr, cls, prefix, num_nets, off, rank, completes, by_rank, d_big);
    size_t tb = std::max(need, need2);
    e = rocprim::exclusive_scan(temp, tb, completes, order, 0, (size_t)num_nets, rocprim::plus<int>(), nullptr);
    int last_order = 0, last_completes = 0;
    if (e == hipSuccess) e = hipMemcpy(&last_order, order + num_nets - 1, sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(&last_completes, completes + num_nets - 1, sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        release();
        AP_HIP(e);
    }
    num_waves = last_order + last_completes;     // the full wavefronts, in the order their last nets come
    for (int k = 1; k <= 6; k++) {               // then each class's partial one
        partial.off[k] = -1;
        if (count.off[k] % (64 >> k)) partial.off[k] = num_waves++;
    }
    if (num_waves) {
        const long long n = (long long)num_waves * 64;
        if ((e = dev_malloc(&d_lg, num_waves)) != hipSuccess || (e = dev_malloc(&d_lanes, n * sizeof(int4))) !=
            hipSuccess || (e = take((void**)&lane_pin, n * sizeof(int))) != hipSuccess) {
            release();
            AP_HIP(e);
        }
        if (with_nets && (e = dev_malloc(&d_net, n * sizeof(int))) != hipSuccess) {
            release();
            AP_HIP(e);
        }
        hipLaunchKernelGGL(wa_plan_clear, dim3(grid_of(n, 256)), b, 0, nullptr, lane_pin, n);
        if (d_net) hipLaunchKernelGGL(wa_plan_clear, dim3(grid_of(n, 256)), b, 0, nullptr, d_net, n);
        hipLaunchKernelGGL(wa_plan_lanes, g, b, 0, nullptr, d_net_start, cls, rank, num_nets, off, count, partial,
                           by_rank, order, lane_pin, d_lg, d_net);
        hipLaunchKernelGGL(wa_lanes_fill, dim3(grid_of(n, 256)), b, 0, nullptr, lane_pin, n, d_pin_cell,
                           (const float2*)d_pin_off, d_pin_slot, d_lanes);
    }
    e = hipGetLastError();
    release();
    AP_HIP(e);
    return AP_OK;
}

void WaPlan::destroy() {
    if (d_lg) (void)dev_free(d_lg);
    if (d_lanes) (void)dev_free(d_lanes);
    if (d_big) (void)dev_free(d_big);
    if (d_net) (void)dev_free(d_net);
    d_lg = nullptr;
    d_lanes = nullptr;
    d_big = nullptr;
    d_net = nullptr;
}

int WaPlan::launch(const WaArgs& a, double* partial, hipStream_t s) const {
    if (num_waves) {
        const WaveArgs w = wave_args(a, partial);
        AP_REQUIRE(shape_ok(per, lanes), "wa plan: launch shape %d x %d", per, lanes);
        AP_REQUIRE(!a.out || d_net, "wa plan: per-net outputs need a plan made with its nets");
        const dim3 g(wave_blocks()), b(lanes);
        if (a.out) {
            if (per == 1) hipLaunchKernelGGL(wa_wave_nets_kernel<1>, g, b, 0, s, w);
            else if (per == 2) hipLaunchKernelGGL(wa_wave_nets_kernel<2>, g, b, 0, s, w);
            else hipLaunchKernelGGL(wa_wave_nets_kernel<4>, g, b, 0, s, w);
        } else {
            if (per == 1) hipLaunchKernelGGL(wa_wave_kernel<1>, g, b, 0, s, w);
            else if (per == 2) hipLaunchKernelGGL(wa_wave_kernel<2>, g, b, 0, s, w);
            else hipLaunchKernelGGL(wa_wave_kernel<4>, g, b, 0, s, w);
        }
        AP_LAUNCHED();
    }
    if (num_big) {
        hipLaunchKernelGGL(wa_big_nets_kernel, dim3(num_big), dim3(WA_BLOCK), 0, s, a, d_big, num_big,
                           partial ? partial + 2LL * num_waves : nullptr);
        AP_LAUNCHED();
    }
    return AP_OK;
}

}  // namespace ap

// ------------------------------------------------------------------------------ C ABI: the drop-in WA with a plan
using namespace ap;

struct ap_wa_plan {
    WaPlan plan;
    int num_nets = 0, num_pins = 0;
};

extern "C" {

int ap_wa_plan_create(const int* pin_cell, const float* pin_off, const int* net_start, int num_nets, int num_pins,
                      ap_wa_plan** plan) {
    AP_REQUIRE(plan && num_nets >= 0 && num_pins >= 0, "wa_plan_create: bad arguments");
    AP_REQUIRE(num_nets == 0 || (pin_cell && pin_off && net_start), "wa_plan_create: null pointer");
    ap_wa_plan* h = new ap_wa_plan();
    h->num_nets = num_nets;
    h->num_pins = num_pins;
    const int st = h->plan.create(net_start, num_nets, pin_cell, pin_off, nullptr, num_pins, true);
    if (st) {
        h->plan.destroy();
        delete h;
        return st;
    }
    *plan = h;
    return AP_OK;
}

int ap_wa_plan_destroy(ap_wa_plan* plan) {
    if (plan) {
        (void)hipDeviceSynchronize();
        plan->plan.destroy();
        delete plan;
    }
    return AP_OK;
}

int ap_wa_wirelength_planned(const ap_wa_plan* plan, const float* pos, const int* pin_cell, const float* pin_off,
                             const int* net_start, const int* cell_start, const int* cell_pin, int num_nets,
                             int num_mov, const float* gamma, float* pin_grad, float* out, float* grad, int num_pins,
                             void* stream) {
    AP_REQUIRE(plan, "wa_planned: null plan");
    AP_REQUIRE(num_nets == plan->num_nets && num_pins == plan->num_pins,
               "wa_planned: the plan is of %d nets and %d pins, not %d and %d", plan->num_nets, plan->num_pins,
               num_nets, num_pins);
    AP_REQUIRE(num_mov >= 0, "wa_planned: negative size");
    AP_REQUIRE(num_nets == 0 || (net_start && gamma && pin_grad && out), "wa_planned: null pointer");
    hipStream_t s = (hipStream_t)stream;
    const WaArgs a{(const float2*)pos, pin_cell, (const float2*)pin_off, net_start, num_nets, num_mov, gamma,
                   (float2*)pin_grad, (float2*)out, nullptr, nullptr, nullptr, nullptr, nullptr};
    int st = plan->plan.launch(a, nullptr, s);
    if (st) return st;
    return launch_wa_cell_sum((const float2*)pin_grad, cell_start, cell_pin, num_mov, (float2*)grad, s);
}

}  // extern "C"
