// Poisson solve of the electrostatic density (include/autoplace_gp.h, part 2).
//
// The reference takes the field from torch.fft on the map mirrored in both axes (4x the data,
// src/core/electronic_density_layer.py:62-77,133). The same field is a cosine / sine series of the unmirrored map:
//     C  = Cx rho Cy^T                                   (2-D DCT-II, unnormalised)
//     Ax = wx o C,  Ay = wy o C                          (spectral weights, see poisson_weights)
//     field_x = Sx^T Ax Cy,   field_y = Cx^T Ay Sy
// with Cx[k][n] = cos(pi k (2n + 1) / 2N), Sx[k][n] = sin(...). Method GEMM runs it as six FP32 GEMMs on the MFMA
// units (v_mfma_f32_32x32x2f32): 12 N^3 flops for an N x N grid. Method FFT (gp_fft.hip) runs the same series as
// LDS-resident FFTs for power-of-two sizes.
#include <math.h>

#include <vector>

#include "gp_common.h"
#include "gp_poisson.h"

namespace ap {

typedef float floatx16 __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------------------------- MFMA GEMM
// C[M][N] = A[M][K] B[K][N], row-major FP32 with leading dimensions. Workgroup tile 128 x 128, K tile 16,
// 4 wavefronts in 2 x 2, each 64 x 64 = 2 x 2 MFMA tiles of 32 x 32. Loads are bounds-checked (zero padding), so
// any M, N, K works. Epilogue: plain store, or two stores scaled by W1 / W2 (the spectral weights).
constexpr int GB_M = 128, GB_N = 128, GB_K = 16;

struct GemmProblem {
    const float* A;
    const float* B;
    float* C;
    float* C2;             // epilogue 1: C = acc * W1, C2 = acc * W2
    const float* W1;
    const float* W2;
    int lda, ldb, ldc;
    const float* a_shift;  // optional device scalar subtracted from every element of A as it is loaded
};
struct GemmBatch {
    GemmProblem p[2];
    int M, N, K;
    int epilogue;
};

__global__ __launch_bounds__(256) void gemm_mfma_kernel(GemmBatch batch) {
    const GemmProblem& pr = batch.p[blockIdx.z];
    const int M = batch.M, N = batch.N, K = batch.K;
    __shared__ float As[GB_K][GB_M + 4];   // A tile stored k-major: As[k][m]
    __shared__ float Bs[GB_K][GB_N + 4];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int m0 = blockIdx.y * GB_M, n0 = blockIdx.x * GB_N;
    const float shift = pr.a_shift ? *pr.a_shift : 0.f;
    floatx16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;

    for (int k0 = 0; k0 < K; k0 += GB_K) {
        // A tile: 128 rows x 16 k; thread loads 8 values (rows t / 2, k chunk (t & 1) * 8)
        {
            const int r = t >> 1, kc = (t & 1) * 8;
            const int gm = m0 + r;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int gk = k0 + kc + q;
                As[kc + q][r] = (gm < M && gk < K) ? pr.A[(long long)gm * pr.lda + gk] - shift : 0.f;
            }
        }
        // B tile: 16 k x 128 cols; thread loads 8 values (row t / 16, cols (t & 15) * 8)
        {
            const int kr = t >> 4, cc = (t & 15) * 8;
            const int gk = k0 + kr;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int gn = n0 + cc + q;
                Bs[kr][cc + q] = (gk < K && gn < N) ? pr.B[(long long)gk * pr.ldb + gn] : 0.f;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GB_K; kk += 2) {
            const int kl = kk + (lane >> 5);
            float a0 = As[kl][wm + (lane & 31)], a1 = As[kl][wm + 32 + (lane & 31)];
            float b0 = Bs[kl][wn + (lane & 31)], b1 = Bs[kl][wn + 32 + (lane & 31)];
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    // accumulator layout of a 32 x 32 tile: lane l, register r -> row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int row = m0 + wm + 32 * i + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int col = n0 + wn + 32 * j + (lane & 31);
                if (row < M && col < N) {
                    const float v = acc[i][j][r];
                    if (batch.epilogue == 1) {
                        const long long w = (long long)row * N + col;
                        pr.C[(long long)row * pr.ldc + col] = v * pr.W1[w];
                        pr.C2[(long long)row * pr.ldc + col] = v * pr.W2[w];
                    } else {
                        pr.C[(long long)row * pr.ldc + col] = v;
                    }
                }
            }
}

// mean of the map (two fixed-order stages, FP64): the cosine series of the zero-mean map has the same field (the
// spectral weight of the constant mode is 0) and far less cancellation in FP32 at low frequencies
constexpr int MEAN_BLOCKS = 256;
__global__ __launch_bounds__(256) void map_sum_kernel(const float* rho, int ld, int N, int M, double* part) {
    __shared__ double red[16];
    double v[1] = {0.0};
    const long long total = (long long)N * M;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
        const int r = (int)(e / M), c = (int)(e - (long long)r * M);
        v[0] += rho[(long long)r * ld + c];
    }
    block_sum<1>(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}
__global__ __launch_bounds__(256) void map_mean_kernel(const double* part, int n, double inv_count, float* mean) {
    __shared__ double red[16];
    double v[1] = {0.0};
    for (int k = threadIdx.x; k < n; k += 256) v[0] += part[k];
    block_sum<1>(v, red);
    if (threadIdx.x == 0) *mean = (float)(v[0] * inv_count);
}

static int launch_gemm(const GemmBatch& b, int nbatch, hipStream_t s) {
    dim3 grid(grid_of(b.N, GB_N), grid_of(b.M, GB_M), nbatch);
    hipLaunchKernelGGL(gemm_mfma_kernel, grid, dim3(256), 0, s, b);
    AP_LAUNCHED();
    return AP_OK;
}

// --------------------------------------------------------------------------------------------- the plan

// spectral weights: Ax = wx C, Ay = wy C (row-major [N][M]); closed form of the mirrored-FFT multiplier
void poisson_weights(int N, int M, double bx, double by, std::vector<float>& wx, std::vector<float>& wy) {
    wx.assign((size_t)N * M, 0.f);
    wy.assign((size_t)N * M, 0.f);
    const double s = 4.0 / ((double)N * M);
    for (int k = 0; k < N; k++)
        for (int l = 0; l < M; l++) {
            const double w = M_PI * k / (N * bx), v = M_PI * l / (M * by), den = (w * w + v * v) * bx;
            const double inv = den > 0 ? 1.0 / den : 0.0;
            wx[(size_t)k * M + l] = (float)(s * w * inv * (l == 0 ? 0.5 : 1.0));
            wy[(size_t)k * M + l] = (float)(s * v * inv * (k == 0 ? 0.5 : 1.0));
        }
}

static void trig(int n, std::vector<float>& c, std::vector<float>& st, std::vector<float>& ct_t,
                 std::vector<float>& st_t) {
    c.resize((size_t)n * n);
    st.resize((size_t)n * n);
    ct_t.resize((size_t)n * n);
    st_t.resize((size_t)n * n);
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++) {
            // angle reduced exactly in integers: pi k (2j + 1) / 2n, k (2j + 1) mod 4n
            long long q = ((long long)k * (2 * j + 1)) % (4LL * n);
            double ang = M_PI * (double)q / (2.0 * n);
            float cv = (float)cos(ang), sv = (float)sin(ang);
            c[(size_t)k * n + j] = cv;
            st[(size_t)k * n + j] = sv;
            ct_t[(size_t)j * n + k] = cv;
            st_t[(size_t)j * n + k] = sv;
        }
}

static int upload(const std::vector<float>& h, float** d) {
    AP_HIP(dev_malloc(d, h.size() * sizeof(float)));
    AP_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    return AP_OK;
}

int Poisson::create(int nbx, int nby, double bx_, double by_, int method_) {
    N = nbx;
    M = nby;
    bx = bx_;
    by = by_;
    const bool pow2 = (N & (N - 1)) == 0 && (M & (M - 1)) == 0 && N >= 2 && M >= 2;
    if (method_ == AP_POISSON_AUTO) method_ = (pow2 && N >= fft_min && M >= fft_min && fft_max_ok()) ? AP_POISSON_FFT
                                                                                                    : AP_POISSON_GEMM;
    if (method_ == AP_POISSON_FFT && !(pow2 && fft_supported(N) && fft_supported(M))) {
        set_error("poisson: the FFT method needs power-of-two sides in [%d, %d], got %d x %d", 2, fft_max, N, M);
        return AP_ERR_UNSUPPORTED;
    }
    method = method_;
    std::vector<float> wx, wy;
    poisson_weights(N, M, bx, by, wx, wy);
    int st;
    if ((st = upload(wx, &d_wx)) || (st = upload(wy, &d_wy))) return st;
    const size_t nm = (size_t)N * M;
    AP_HIP(dev_malloc(&d_work, 4 * nm * sizeof(float)));
    AP_HIP(dev_malloc(&d_part, MEAN_BLOCKS * sizeof(double)));
    AP_HIP(dev_malloc(&d_mean, sizeof(float)));
    if (method == AP_POISSON_GEMM) {
        std::vector<float> cx, sx, cxt, sxt, cy, sy, cyt, syt;
        trig(N, cx, sx, cxt, sxt);
        trig(M, cy, sy, cyt, syt);
        if ((st = upload(cx, &d_cx)) || (st = upload(cxt, &d_cxt)) || (st = upload(sxt, &d_sxt)) ||
            (st = upload(cy, &d_cy)) || (st = upload(sy, &d_sy)) || (st = upload(cyt, &d_cyt)))
            return st;
    } else {
        if ((st = fft_create())) return st;
    }
    return AP_OK;
}

void Poisson::destroy() {
    void* ptrs[] = {d_wx, d_wy, d_work, d_cx, d_cxt, d_sxt, d_cy, d_sy, d_cyt, d_part, d_mean};
    for (void* p : ptrs)
        if (p) (void)dev_free(p);
    fft_destroy();
}

// E = 1/2 sum_kl w_phi[k][l] C[k][l]^2 with w_phi = 4/(NM) h_k h_l / ((w_k^2 + v_l^2) bx), h_0 = 1/2: the
// electrostatic energy 1/2 sum rho phi of the map, in the units in which its position gradient is the field gather.
__global__ __launch_bounds__(256) void energy_partial_kernel(const float* C, int N, int M, double bx, double by,
                                                             double* part) {
    __shared__ double red[16];
    double v[1] = {0.0};
    const long long total = (long long)N * M;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
        const int k = (int)(e / M), l = (int)(e - (long long)k * M);
        if (k == 0 && l == 0) continue;
        const double w = M_PI * k / (N * bx), nu = M_PI * l / (M * by);
        const double wphi = 4.0 / ((double)N * M) * (k ? 1.0 : 0.5) * (l ? 1.0 : 0.5) / ((w * w + nu * nu) * bx);
        const double c = C[e];
        v[0] += 0.5 * wphi * c * c;
    }
    block_sum<1>(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}
__global__ __launch_bounds__(256) void energy_final_kernel(const double* part, int n, double* out) {
    __shared__ double red[16];
    double v[1] = {0.0};
    for (int k = threadIdx.x; k < n; k += 256) v[0] += part[k];
    block_sum<1>(v, red);
    if (threadIdx.x == 0) *out = v[0];
}

int Poisson::energy(const float* rho, int rho_ld, double* out, hipStream_t s) {
    if (method != AP_POISSON_GEMM) {
        set_error("poisson_energy: needs a plan of the GEMM method");
        return AP_ERR_UNSUPPORTED;
    }
    const size_t nm = (size_t)N * M;
    float *T1 = d_work, *C = d_work + nm;
    hipLaunchKernelGGL(map_sum_kernel, dim3(MEAN_BLOCKS), dim3(256), 0, s, rho, rho_ld, N, M, d_part);
    AP_LAUNCHED();
    hipLaunchKernelGGL(map_mean_kernel, dim3(1), dim3(256), 0, s, d_part, MEAN_BLOCKS, 1.0 / ((double)N * M), d_mean);
    AP_LAUNCHED();
    GemmBatch b{};
    b.M = N; b.N = M; b.K = M; b.epilogue = 0;
    b.p[0] = GemmProblem{rho, d_cyt, T1, nullptr, nullptr, nullptr, rho_ld, M, M, d_mean};
    int st;
    if ((st = launch_gemm(b, 1, s))) return st;
    b.M = N; b.N = M; b.K = N;
    b.p[0] = GemmProblem{d_cx, T1, C, nullptr, nullptr, nullptr, N, M, M};
    if ((st = launch_gemm(b, 1, s))) return st;
    hipLaunchKernelGGL(energy_partial_kernel, dim3(MEAN_BLOCKS), dim3(256), 0, s, C, N, M, bx, by, d_part);
    AP_LAUNCHED();
    hipLaunchKernelGGL(energy_final_kernel, dim3(1), dim3(256), 0, s, d_part, MEAN_BLOCKS, out);
    AP_LAUNCHED();
    return AP_OK;
}

int Poisson::solve_pairs(const float* rho, int rho_ld, float2* f, int fld, hipStream_t s) {
    AP_REQUIRE(method == AP_POISSON_FFT, "poisson: the paired field needs the FFT method");
    return fft_solve(rho, rho_ld, nullptr, nullptr, f, fld, s);
}

int Poisson::solve(const float* rho, int rho_ld, float* fx, float* fy, int fld, hipStream_t s) {
    if (method == AP_POISSON_FFT) return fft_solve(rho, rho_ld, fx, fy, nullptr, fld, s);
    const size_t nm = (size_t)N * M;
    float *T1 = d_work, *Ax = d_work + nm, *Ay = d_work + 2 * nm, *U = d_work + 3 * nm;
    float* V = T1;   // T1 is dead once Ax / Ay exist
    int st;
    hipLaunchKernelGGL(map_sum_kernel, dim3(MEAN_BLOCKS), dim3(256), 0, s, rho, rho_ld, N, M, d_part);
    AP_LAUNCHED();
    hipLaunchKernelGGL(map_mean_kernel, dim3(1), dim3(256), 0, s, d_part, MEAN_BLOCKS, 1.0 / ((double)N * M), d_mean);
    AP_LAUNCHED();
    GemmBatch b{};
    // T1 = (rho - mean) Cy^T  [N x M] = [N x M][M x M]
    b.M = N; b.N = M; b.K = M; b.epilogue = 0;
    b.p[0] = GemmProblem{rho, d_cyt, T1, nullptr, nullptr, nullptr, rho_ld, M, M, d_mean};
    if ((st = launch_gemm(b, 1, s))) return st;
    // C = Cx T1 [N x M] = [N x N][N x M]; Ax = wx C, Ay = wy C
    b.M = N; b.N = M; b.K = N; b.epilogue = 1;
    b.p[0] = GemmProblem{d_cx, T1, Ax, Ay, d_wx, d_wy, N, M, M};
    if ((st = launch_gemm(b, 1, s))) return st;
    // U = Sx^T Ax, V = Cx^T Ay   [N x M] = [N x N][N x M]
    b.epilogue = 0;
    b.p[0] = GemmProblem{d_sxt, Ax, U, nullptr, nullptr, nullptr, N, M, M};
    b.p[1] = GemmProblem{d_cxt, Ay, V, nullptr, nullptr, nullptr, N, M, M};
    if ((st = launch_gemm(b, 2, s))) return st;
    // field_x = U Cy, field_y = V Sy   [N x M] = [N x M][M x M]
    b.M = N; b.N = M; b.K = M;
    b.p[0] = GemmProblem{U, d_cy, fx, nullptr, nullptr, nullptr, M, M, fld};
    b.p[1] = GemmProblem{V, d_sy, fy, nullptr, nullptr, nullptr, M, M, fld};
    return launch_gemm(b, 2, s);
}

}  // namespace ap

using namespace ap;

struct ap_poisson {
    Poisson p;
};

namespace ap {
const Poisson& poisson_of(const ap_poisson* plan) { return plan->p; }
}  // namespace ap

extern "C" {

int ap_poisson_create(int nbx, int nby, double bx, double by, int method, ap_poisson** plan) {
    AP_REQUIRE(plan && nbx > 0 && nby > 0 && bx > 0 && by > 0, "poisson_create: bad arguments");
    AP_REQUIRE(method >= AP_POISSON_AUTO && method <= AP_POISSON_FFT, "poisson_create: unknown method %d", method);
    ap_poisson* h = new ap_poisson();
    int st = h->p.create(nbx, nby, bx, by, method);
    if (st) {
        h->p.destroy();
        delete h;
        return st;
    }
    *plan = h;
    return AP_OK;
}

int ap_poisson_destroy(ap_poisson* plan) {
    if (plan) {
        plan->p.destroy();
        delete plan;
    }
    return AP_OK;
}

int ap_poisson_method(const ap_poisson* plan) { return plan ? plan->p.method : -1; }

int ap_poisson_energy(ap_poisson* plan, const float* rho, int rho_ld, double* energy, void* stream) {
    AP_REQUIRE(plan && rho && energy, "poisson_energy: null argument");
    AP_REQUIRE(rho_ld >= plan->p.M, "poisson_energy: leading dimension below %d", plan->p.M);
    return plan->p.energy(rho, rho_ld, energy, (hipStream_t)stream);
}

int ap_poisson_solve(ap_poisson* plan, const float* rho, int rho_ld, float* field_x, float* field_y, int field_ld,
                     void* stream) {
    AP_REQUIRE(plan && rho && field_x && field_y, "poisson_solve: null argument");
    AP_REQUIRE(rho_ld >= plan->p.M && field_ld >= plan->p.M, "poisson_solve: leading dimension below %d", plan->p.M);
    return plan->p.solve(rho, rho_ld, field_x, field_y, field_ld, (hipStream_t)stream);
}

}  // extern "C"
