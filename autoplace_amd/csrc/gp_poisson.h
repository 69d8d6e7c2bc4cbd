// Poisson plan shared by the C ABI and the GP engine.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/autoplace_gp.h"

namespace ap {

void poisson_weights(int N, int M, double bx, double by, std::vector<float>& wx, std::vector<float>& wy);

// One grid's FFT solve as an entry of the argument table of grouped launches (many grids per launch): the buffers
// of a Poisson plan of the FFT method and the map / field it is applied to.
struct FftTab {
    const double* S;             // control state of the grid's block: skipped once S[STOP] != 0 (null: never)
    const float* rho;
    int rho_ld;
    float *fx, *fy;
    int fld;
    float2* f2;                  // non-null: the field interleaved, (x, y) pairs with row stride fld (fx, fy unused)
    float *w0, *w1, *w2, *w3;
    const float2 *twN, *phN, *twM, *phM;
    const float *omega, *nu;
    int N, M, lgN, lgM;
    float scale, bx;
};

// work maps of the five grouped FFT launches: per workgroup (grid, x, y, z)
struct FftGroup {
    int4* map[5] = {};
    int count[5] = {};
    size_t lds_ac = 0, lds_b = 0;
    int create(const std::vector<FftTab>& tabs);
    void destroy();
    int launch(const FftTab* d_tab, hipStream_t s) const;
};

struct Poisson {
    int N = 0, M = 0, method = 0;
    double bx = 1, by = 1;
    float *d_wx = nullptr, *d_wy = nullptr, *d_work = nullptr, *d_mean = nullptr;
    double* d_part = nullptr;
    // GEMM method: Cx, Cx^T, Sx^T (N x N); Cy, Sy, Cy^T (M x M)
    float *d_cx = nullptr, *d_cxt = nullptr, *d_sxt = nullptr, *d_cy = nullptr, *d_sy = nullptr, *d_cyt = nullptr;
    // FFT method: twiddles
    float2* d_tw = nullptr;
    float* d_omega = nullptr;   // FFT method: pi k / (N bx) then pi l / (M by)
    static constexpr int fft_min = 64;       // AUTO picks the FFT from this side up
    static constexpr int fft_max = 4096;     // largest side the LDS FFT holds
    static bool fft_supported(int n);
    static bool fft_max_ok();

    int create(int nbx, int nby, double bx, double by, int method);
    void destroy();
    int solve(const float* rho, int rho_ld, float* fx, float* fy, int fld, hipStream_t s);
    // FFT method only: the field as (x, y) pairs, f[n * fld + j] (one 8-byte store per bin, one load per bin for a
    // reader that needs both)
    int solve_pairs(const float* rho, int rho_ld, float2* f, int fld, hipStream_t s);
    int energy(const float* rho, int rho_ld, double* out, hipStream_t s);

    int fft_create();
    void fft_destroy();
    int fft_solve(const float* rho, int rho_ld, float* fx, float* fy, float2* f2, int fld, hipStream_t s);
    FftTab fft_tab(const float* rho, int rho_ld, float* fx, float* fy, float2* f2, int fld, const double* S) const;
};

// the plan inside a C ABI handle (gp_poisson.hip)
const Poisson& poisson_of(const ap_poisson* plan);

}  // namespace ap
