// One pass of the run's control in double precision. The decisions are the reference's control_kernel
// (cpp_to_py/gp_cuda_kernel.cu:183-240): NesterovOptimizer.step's step length, ParamScheduler.step and
// ParamScheduler.need_to_stop (src/nesterov_optimizer.py:34-53, src/param_scheduler.py:63-116), in that order, so
// that the same statistics give the same state bit for bit. The schedule update accepts a learned adjustment:
// factors on the density-weight multiplier and on the smoothing (both exactly 1 for the plain rule).
#pragma once
#include "gp_common.h"

namespace ap {

struct ScheduleAdjust {
    double lam_factor = 1.0;   // multiplies the rule's density-weight growth mu
    double gam_factor = 1.0;   // multiplies the rule's new smoothing gamma (before its clamp)
};

// Barzilai-Borwein step length of the next pass; every finite step is taken
__device__ __forceinline__ double next_step_length(double* S, double moved, double turned) {
    const double alpha = turned > 0 ? moved / turned : S[ALPHA];
    S[ALPHA] = alpha;
    S[EVALS] += 1;
    S[ACCEPT] = 1;
    S[A] = (1 + sqrt(4 * S[A] * S[A] + 1)) / 2;
    return alpha;
}

// density weight (growth mu from the HPWL trend and the motion), smoothing (accuracy or overflow rule), iteration
__device__ __forceinline__ void update_schedule(double* S, const double* c, double moved, double hpwl, double ovf,
                                                double wa, const ScheduleAdjust& adj) {
    S[MOTION] = moved / c[C_MOTION_UNIT];
    const double dens_part = S[LAM] * c[C_TOTAL_AREA];
    S[WW] = dens_part / (c[C_TOTAL_PINS] + dens_part);
    double mu;
    if (hpwl < S[HPWL_PREV]) {
        mu = 1.05 * 0.98;
    } else {
        const double damp = pow(1.05, -S[MOTION] * c[C_PITCH] / S[GAM]);
        mu = 1.05 * fmax(damp, 0.95);
    }
    mu *= adj.lam_factor;
    S[HPWL_PREV] = hpwl;
    S[MU] = mu;
    S[LAM] *= mu;
    double gam;
    if (c[C_ACCURACY] != 0) {
        const double model_err = fmax(1e-6, 1 - wa / fmax(hpwl, 1e-12));
        const double m = fmax(mu, 1 / mu);
        const double want = sqrt(fmax(ovf, 1e-3) / c[C_REFINE] / model_err);
        gam = S[GAM] * fmin(m, fmax(1 / m, want)) * adj.gam_factor;
        gam = fmin(100 * c[C_BASE_GAMMA], fmax(0.01 * c[C_BASE_GAMMA], gam));
    } else {
        gam = c[C_BASE_GAMMA] * pow(10.0, (ovf - 0.1) * 20 / 9 - 1) * adj.gam_factor;
    }
    S[GAM] = gam;
    S[ITER] += 1;
}

// need_to_stop: 0 = go on, 1 converged, 2 stalled, 3 diverged, 5 handover (4, max_iter, is the caller's)
__device__ __forceinline__ double stop_test(double* S, const double* c, double hpwl, double ovf, double& is_best) {
    const double iter = S[ITER];
    is_best = 0;
    if (!(hpwl < 1e30 && ovf < 1e30)) return 3;
    double stop = 0;
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
    return stop;
}

// the kernels' parameters of the next pass
__device__ __forceinline__ void publish(const double* S, float* kp, double alpha) {
    const double a_next = (1 + sqrt(4 * S[A] * S[A] + 1)) / 2;
    kp[0] = (float)S[GAM];
    kp[1] = (float)S[LAM];
    kp[2] = (float)S[LAM];
    kp[3] = (float)alpha;
    kp[4] = (float)((S[A] - 1) / a_next);
}

__device__ __forceinline__ void control_step(double* S, float* kp, const double* c, double moved, double turned,
                                             double hpwl, double ovf, double wa,
                                             const ScheduleAdjust& adj = ScheduleAdjust()) {
    S[ACCEPT] = 0;
    S[IS_BEST] = 0;
    if (S[STOP] != 0) return;                                    // stopped: later passes change nothing
    if (!(isfinite(moved) && isfinite(turned) && isfinite(hpwl) && isfinite(ovf))) {
        S[STOP] = 3;                                             // diverged: the step is not taken
        return;
    }
    const double alpha = next_step_length(S, moved, turned);
    update_schedule(S, c, moved, hpwl, ovf, wa, adj);
    double is_best;
    double stop = stop_test(S, c, hpwl, ovf, is_best);
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
    publish(S, kp, alpha);
}

}  // namespace ap
