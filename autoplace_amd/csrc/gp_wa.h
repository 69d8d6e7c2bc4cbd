// The engine's WA plan: nets packed whole into 64-lane wavefronts (gp_wa.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "gp_kernels.h"

namespace ap {

struct WaPlan {
    int num_waves = 0, num_big = 0;
    int4* d_waves = nullptr;   // per wavefront: first pin, pins, first net, nets
    int* d_big = nullptr;      // nets of more than 64 pins
    int create(const int* d_net_start, int num_nets);
    void destroy();
    int wave_blocks() const { return (num_waves + 3) / 4; }
    int partial_count() const { return wave_blocks() + num_big; }   // (WA, HPWL) pairs written per launch
    int launch(const WaArgs& a, double* partial, hipStream_t s) const;
};

}  // namespace ap
