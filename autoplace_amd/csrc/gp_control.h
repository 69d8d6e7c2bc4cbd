// One pass of the run's control in double precision: the decisions of the reference's control_kernel
// (cpp_to_py/gp_cuda_kernel.cu:183-240) -- NesterovOptimizer.step's step length, ParamScheduler.step and
// ParamScheduler.need_to_stop (src/nesterov_optimizer.py:34-53, src/param_scheduler.py:63-116) -- in the same order.
#pragma once
#include "gp_common.h"

namespace ap {

__device__ __forceinline__ void control_step(double* S, float* kp, const double* c, double moved, double turned, double hpwl,
                             double ovf, double wa) {
    S[ACCEPT] = 0;
    S[IS_BEST] = 0;
    if (S[STOP] != 0) return;                                    // stopped: later passes change nothing
    if (!(isfinite(moved) && isfinite(turned) && isfinite(hpwl) && isfinite(ovf))) {
        S[STOP] = 3;                                             // diverged: the step is not taken
        return;
    }
    const double alpha_next = turned > 0 ? moved / turned : S[ALPHA];
    S[ALPHA] = alpha_next;
    S[EVALS] += 1;
    S[ACCEPT] = 1;
    S[A] = (1 + sqrt(4 * S[A] * S[A] + 1)) / 2;
    double iter = S[ITER];
    S[MOTION] = moved / c[C_MOTION_UNIT];
    const double part = S[LAM] * c[C_TOTAL_AREA];
    S[WW] = part / (c[C_TOTAL_PINS] + part);
    const double mu = hpwl < S[HPWL_PREV] ? 1.05 * 0.98
                                          : 1.05 * fmax(pow(1.05, -S[MOTION] * c[C_PITCH] / S[GAM]), 0.95);
    S[HPWL_PREV] = hpwl;
    S[MU] = mu;
    S[LAM] *= mu;
    if (c[C_ACCURACY] != 0) {
        const double err = fmax(1e-6, 1 - wa / fmax(hpwl, 1e-12)), m = fmax(mu, 1 / mu);
        const double gm = S[GAM] * fmin(m, fmax(1 / m, sqrt(fmax(ovf, 1e-3) / c[C_REFINE] / err)));
        S[GAM] = fmin(100 * c[C_BASE_GAMMA], fmax(0.01 * c[C_BASE_GAMMA], gm));
    } else {
        S[GAM] = c[C_BASE_GAMMA] * pow(10.0, (ovf - 0.1) * 20 / 9 - 1);
    }
    S[ITER] = iter = iter + 1;
    double stop = 0, is_best = 0;
    if (!(hpwl < 1e30 && ovf < 1e30)) {
        stop = 3;
    } else {
        if (iter >= c[C_MIN_ITER] && ovf < c[C_STOP_OVF]) {
            S[SETTLED] += 1;
            is_best = hpwl < S[BEST_HPWL];
            S[BEST_HPWL] = fmin(hpwl, S[BEST_HPWL]);
            if (S[SETTLED] > c[C_SETTLE]) stop = 1;
        } else if (S[SETTLED] == 0) {
            is_best = ovf < S[BEST_OVF];
            if (ovf < S[BEST_OVF] * 0.999) S[IMPROVED] = iter;
            S[BEST_OVF] = fmin(ovf, S[BEST_OVF]);
        }
        if (stop == 0) {
            if ((is_best != 0 && S[SETTLED] > 0) || S[WW] < 0.95) S[IMPROVED] = iter;
            if (iter - S[IMPROVED] > c[C_STALL]) stop = 2;
        }
        if (stop == 0 && c[C_HANDOVER] > 0 && S[LAM] > 0 && S[GAM] <= c[C_HANDOVER]) stop = 5;
    }
    S[IS_BEST] = is_best;
    if (is_best != 0) {
        S[RES_HPWL] = hpwl;
        S[RES_OVF] = ovf;
        S[HAS_BEST] = 1;
    }
    S[LOOP_IT] += 1;
    if (stop == 0 && S[LOOP_IT] >= c[C_INNER]) stop = 4;
    S[STOP] = stop;
    if (S[HAS_BEST] == 0) {
        S[RES_HPWL] = hpwl;
        S[RES_OVF] = ovf;
    }
    const double a_next = (1 + sqrt(4 * S[A] * S[A] + 1)) / 2;
    kp[0] = (float)S[GAM];
    kp[1] = (float)S[LAM];
    kp[2] = (float)S[LAM];
    kp[3] = (float)alpha_next;
    kp[4] = (float)((S[A] - 1) / a_next);
}

}  // namespace ap
