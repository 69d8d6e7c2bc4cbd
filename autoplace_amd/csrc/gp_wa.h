// The engine's WA plan: nets in aligned lane groups of 2^k lanes, one degree class per wavefront (gp_wa.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "gp_kernels.h"

namespace ap {

constexpr int WA_PLAN_PER_WAVE = 2;   // wavefronts of the plan run by one launched wavefront

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
    unsigned char* d_lg = nullptr;   // per wavefront: log2 of its lanes per net
    int4* d_lanes = nullptr;         // per lane, in wavefront order: pin cell (-1: no pin), gradient slot (-1: none),
                                     // pin offset x, y (float bits)
    int* d_big = nullptr;            // nets of more than 64 pins
    int* d_net = nullptr;            // with_nets: per lane, the net of the lane's group (-1: none)
    // pin_slot: where pin p's gradient goes (null: p); with_nets: the per-lane nets, for per-net outputs
    int create(const int* d_net_start, int num_nets, const int* d_pin_cell, const float* d_pin_off,
               const int* d_pin_slot, int num_pins, bool with_nets = false);
    void destroy();
    int wave_blocks() const { return (num_waves + 4 * WA_PLAN_PER_WAVE - 1) / (4 * WA_PLAN_PER_WAVE); }
    int partial_count() const { return wave_blocks() + num_big; }   // (WA, HPWL) pairs written per launch
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
