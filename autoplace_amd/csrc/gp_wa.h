// The engine's WA plan: nets in aligned lane groups of 2^k lanes, one degree class per wavefront (gp_wa.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "gp_kernels.h"

namespace ap {

constexpr int WA_PLAN_PER_WAVE = 2;   // wavefronts of the plan run by one launched wavefront (default)
constexpr int WA_LANES = 256;         // lanes per workgroup of the WA launch (default)

struct WaveArgs {
    const float2* pos;
    const float2* pos_other;
    const int* flag;
    const unsigned char* lg;   // per wavefront: log2 of its lanes per net
    const int4* lanes;         // per lane: cell (-1: none), gradient slot, offset x, y
    int num_waves;
    int num_mov;
    const float* gamma;
    float2* pin_grad;
    double* partial;           // per workgroup (WA, HPWL)
    float2* out;               // per net (WA, HPWL) (the drop-in's planned WA; null in the engine)
    const int* lane_net;       // with out: per lane, the net of the lane's group (-1: none)
};

// grouped launch of many levels' WA (tables on the device; see wa_wave_grouped)
int launch_wa_grouped(const WaveArgs* d_tab, const int4* d_map, int count, const double* const* d_S, hipStream_t s);

struct WaPlan {
    int num_waves = 0, num_big = 0;
    // the launch shape of launch(): plan wavefronts per launched wavefront (1, 2, 4) and lanes per workgroup (64 ..
    // 512). Each plan wavefront writes its own partial, so the results do not depend on the shape.
    int per = WA_PLAN_PER_WAVE, lanes = WA_LANES;
    unsigned char* d_lg = nullptr;   // per wavefront: log2 of its lanes per net
    int4* d_lanes = nullptr;         // per lane, in wavefront order: pin cell (-1: no pin), gradient slot (-1: none),
                                     // pin offset x, y (float bits)
    int* d_big = nullptr;            // nets of more than 64 pins
    int* d_net = nullptr;            // with_nets: per lane, the net of the lane's group (-1: none)
    // pin_slot: where pin p's gradient goes (null: p); with_nets: the per-lane nets, for per-net outputs
    int create(const int* d_net_start, int num_nets, const int* d_pin_cell, const float* d_pin_off,
               const int* d_pin_slot, int num_pins, bool with_nets = false);
    void destroy();
    int wave_blocks(int p, int l) const { return (num_waves + (l / 64) * p - 1) / ((l / 64) * p); }
    int wave_blocks() const { return wave_blocks(per, lanes); }
    // (WA, HPWL) pairs written per launch: one per plan wavefront, then one per big net
    int partial_count() const { return num_waves + num_big; }
    static bool shape_ok(int p, int l) { return (p == 1 || p == 2 || p == 4) && l >= 64 && l <= 512 && l % 64 == 0; }
    // a.out set: also every net's (WA, HPWL) (a plan made with_nets)
    int launch(const WaArgs& a, double* partial, hipStream_t s) const;
    WaveArgs wave_args(const WaArgs& a, double* partial) const;
};

// Work item of workgroup b among nb when consecutive workgroups land on the 8 XCDs in turn: XCD x gets a contiguous
// range of work items, so the workgroups sharing an XCD's L2 work on neighbouring data.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

}  // namespace ap
