// Kernel argument blocks shared by the drop-in entry points and the GP engine.
#pragma once
#include <hip/hip_runtime.h>

namespace ap {

constexpr int WA_BLOCK = 256;   // lanes per WA workgroup
constexpr int WA_SPAN = 256;    // a workgroup owns the nets whose first pin is in its 256-pin window
constexpr int WA_CAP = 512;     // pins staged in LDS at once

struct WaArgs {
    const float2* pos;
    const int* pin_cell;
    const float2* pin_off;
    const int* net_start;
    int num_nets, num_mov;
    const float* gamma;
    float2* pin_grad;
    float2* out;        // per-net (WA, HPWL), may be null
    double* partial;    // per-workgroup (WA, HPWL) in FP64, may be null
    const float2* pos_other;   // with flag: the position is pos_other when *flag == 0, else pos
    const int* flag;           // null: pos as given
    const int* block_net;      // optional [workgroups + 1]: first net owned by each workgroup (else searched)
    const int* pin_slot;       // optional [pins]: write pin p's gradient to pin_grad[pin_slot[p]] (< 0: skip)
};

// block_net[b] = first net whose first pin is at or after b * WA_SPAN, b in [0, blocks]
int launch_wa_block_table(const int* net_start, int num_nets, int blocks, int* block_net, hipStream_t s);

int launch_wa(const WaArgs& a, int num_pins, hipStream_t s);
// grad[c] = the sum of the gradients of cell c's pins (pin_grad in pin order, the cell's pins in cell_pin)
int launch_wa_cell_sum(const float2* pin_grad, const int* cell_start, const int* cell_pin, int num_mov, float2* grad,
                       hipStream_t s);

}  // namespace ap
