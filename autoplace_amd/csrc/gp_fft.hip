// FFT form of the Poisson solve for power-of-two grids (2 .. 4096 per side).
//
// The cosine / sine series of gp_poisson.hip evaluated with LDS-resident complex FFTs, one workgroup per row:
//   pass A  rows of rho (length M):   T = rho Cy^T                 DCT-II of two real rows in one complex FFT
//   pass B  columns of T (length N), read in place by pairs:
//                                     C = DCT-II along x; Ax = wx C, Ay = wy C (weights computed in place);
//                                     U^T = Sx^T Ax, V^T = Cx^T Ay  (sine and cosine series of one column from
//                                     one complex inverse FFT: both outputs are real, so they share it)
//   pass C  columns of U^T, V^T (rows of U, V, length M), read in place by pairs:
//                                     field_x = U Cy, field_y = V Sy  (one inverse FFT per row)
// Passes B and C run consecutive workgroups on one XCD, so the lines the strided column reads share meet in one L2
// (this replaced two transposes through LDS tiles). The batch's grouped launches keep the transposes.
// DCT-II via Makhoul's reordering (v[n] = x[2n], v[L-1-n] = x[2n+1]; X[k] = Re(e^{-i pi k/2L} DFT(v)[k])); the
// cosine series y[n] = sum_k a[k] cos(pi k (2n+1)/2L) via its inverse (V[k] = e^{i pi k/2L}(X'[k] - i X'[L-k]),
// X'[0] = 2a[0]; y[2n] = v[n]/2, y[2n+1] = v[L-1-n]/2); the sine series as (-1)^n times the cosine series of the
// reversed coefficients. Each pass reads and writes every element once; the map stays unmirrored.
#include <math.h>

#include <algorithm>
#include <vector>

#include "gp_common.h"
#include "gp_poisson.h"
#include "gp_wa.h"   // xcd_remap

namespace ap {

constexpr int FB = 512;   // lanes per FFT workgroup

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ int bitrev(int n, int lg) { return (int)(__brev((unsigned)n) >> (32 - lg)); }
// LDS place of FFT element i: bits 0-3 flipped by bits 3-6 (a bijection inside every 128-element block) and one pad
// slot per block. The butterflies' accesses -- runs of 1, 4, 16, 64, ... consecutive elements 4 runs apart -- then
// spread over the banks: by a model of 8-byte accesses on 32 or 64 four-byte banks, about half the cycles of one
// pad slot per 8 elements, whose slot breaks every 64-byte run. An FFT of L elements takes lds_len(L) slots.
__device__ __forceinline__ int lpad(int i) { return (i ^ ((i >> 3) & 15)) + (i >> 7); }
__host__ __device__ __forceinline__ int lds_len(int L) { return L + (L >> 7) + 1; }
// The twiddles in LDS: tw[j] (j < L / 2, for a last radix-2 stage) at lds_len(L), then, 16-byte aligned at
// tws_off(L), the per-stage table (float4 entries), stage by stage: a radix-8 stage of half-size h (h = 1, 8, 64,
// ...) holds (W_2h^j, W_4h^j) for j < h, then (W_8h^j, W_8h^(j+h)) for j < h; a last radix-4 stage (W_2h^j, W_4h^j)
// for j < h. A butterfly reads its twiddles in one or two 16-byte accesses, and consecutive lanes read consecutive
// entries, where the plain table's stride-L / 2h reads hit one bank.
__host__ __device__ __forceinline__ int stage_len(int lg) {
    int n = 0, st = 1;
    for (; st + 2 <= lg; st += 3) n += 2 << (st - 1);
    if (st + 1 <= lg) n += 1 << (st - 1);
    return n;
}
__host__ __device__ __forceinline__ int tws_off(int L) { return (lds_len(L) + (L >> 1) + 1) & ~1; }
// float2 slots of a workgroup's LDS for an FFT of L = 2^lg elements
__host__ __device__ __forceinline__ int fft_slots(int L, int lg) { return tws_off(L) + 2 * stage_len(lg); }
// position of input element m in Makhoul's order: even m -> m / 2, odd m -> L - 1 - (m - 1) / 2
__device__ __forceinline__ int makhoul(int m, int L) { return (m & 1) ? L - 1 - (m >> 1) : (m >> 1); }

// the radix-2 butterfly (x, y) <- (x + w y, x - w y)
__device__ __forceinline__ void bfly(float2& x, float2& y, float2 w) {
    const float2 t = cmul(y, w);
    y = make_float2(x.x - t.x, x.y - t.y);
    x = make_float2(x.x + t.x, x.y + t.y);
}
// w W_4^1: W_4^1 = -i (forward) or +i (inverse, sg = -1), exact
__device__ __forceinline__ float2 quarter(float2 w, float sg) { return make_float2(sg * w.y, -sg * w.x); }

// In-place decimation in time on bit-reversed input, the twiddles from the stage table tws (and tw for a last radix-2
// stage, load_twiddles); inverse: conjugate twiddles (no 1 / L scale). Three radix-2 stages (half-sizes h, 2h, 4h) are
// done together as one radix-8 butterfly in registers on x[i + k h], k < 8: one barrier and one LDS round trip per
// three stages (an FFT of 2048 points: 3 radix-8 stages and one radix-4, 4 round trips instead of 11); the twiddles at
// j + h of the second stage and at j + 2h, j + 3h of the third are the ones at j and j + h times W_4^1 (exact).
// One or two stages left end as one radix-4 or radix-2 stage.
__device__ void fft_lds(float2* s, const float2* tw, const float4* tws, int L, int lg, bool inverse) {
    const float sg = inverse ? -1.f : 1.f;
    int st = 1, base = 0;
    for (; st + 2 <= lg; base += 2 << (st - 1), st += 3) {
        const int h = 1 << (st - 1);
        for (int b = threadIdx.x; b < (L >> 3); b += blockDim.x) {
            const int j = b & (h - 1);
            const int i0 = ((b >> (st - 1)) << (st + 2)) + j;
            const float4 ta = tws[base + j], tb = tws[base + h + j];
            const float2 w1 = make_float2(ta.x, sg * ta.y), w2 = make_float2(ta.z, sg * ta.w);
            const float2 w4 = make_float2(tb.x, sg * tb.y), w5 = make_float2(tb.z, sg * tb.w);
            const float2 w3 = quarter(w2, sg), w6 = quarter(w4, sg), w7 = quarter(w5, sg);
            float2 x[8];
#pragma unroll
            for (int k = 0; k < 8; k++) x[k] = s[lpad(i0 + k * h)];
#pragma unroll
            for (int k = 0; k < 8; k += 2) bfly(x[k], x[k + 1], w1);
            bfly(x[0], x[2], w2);
            bfly(x[1], x[3], w3);
            bfly(x[4], x[6], w2);
            bfly(x[5], x[7], w3);
            bfly(x[0], x[4], w4);
            bfly(x[1], x[5], w5);
            bfly(x[2], x[6], w6);
            bfly(x[3], x[7], w7);
#pragma unroll
            for (int k = 0; k < 8; k++) s[lpad(i0 + k * h)] = x[k];
        }
        __syncthreads();
    }
    if (st + 1 <= lg) {
        const int h = 1 << (st - 1);
        for (int b = threadIdx.x; b < (L >> 2); b += blockDim.x) {
            const int j = b & (h - 1);
            const int i0 = ((b >> (st - 1)) << (st + 1)) + j;
            const float4 t = tws[base + j];
            const float2 w1 = make_float2(t.x, sg * t.y), w2 = make_float2(t.z, sg * t.w);
            const float2 w3 = quarter(w2, sg);
            float2 x[4];
#pragma unroll
            for (int k = 0; k < 4; k++) x[k] = s[lpad(i0 + k * h)];
            bfly(x[0], x[1], w1);
            bfly(x[2], x[3], w1);
            bfly(x[0], x[2], w2);
            bfly(x[1], x[3], w3);
#pragma unroll
            for (int k = 0; k < 4; k++) s[lpad(i0 + k * h)] = x[k];
        }
        __syncthreads();
        st += 2;
    }
    if (st == lg) {
        const int half = 1 << (st - 1), stride = L >> st;
        for (int b = threadIdx.x; b < (L >> 1); b += blockDim.x) {
            const int j = b & (half - 1);
            const int i0 = ((b >> (st - 1)) << st) + j, i1 = i0 + half;
            float2 w = tw[j * stride];
            w.y *= sg;
            float2 x0 = s[lpad(i0)], x1 = s[lpad(i1)];
            bfly(x0, x1, w);
            s[lpad(i0)] = x0;
            s[lpad(i1)] = x1;
        }
        __syncthreads();
    }
}

// the stage table tws (and tw_s[j] = tw[j], j < L / 2, for a last radix-2 stage) from the twiddles in global memory
__device__ __forceinline__ void load_twiddles(float2* tw_s, float4* tws, const float2* tw, int L, int lg) {
    int st = 1, base = 0;
    for (; st + 2 <= lg; base += 2 << (st - 1), st += 3) {
        const int h = 1 << (st - 1), s1 = L >> st, s2 = L >> (st + 1), s3 = L >> (st + 2);
        for (int j = threadIdx.x; j < h; j += blockDim.x) {
            const float2 a = tw[j * s1], b = tw[j * s2], c = tw[j * s3], d = tw[(j + h) * s3];
            tws[base + j] = make_float4(a.x, a.y, b.x, b.y);
            tws[base + h + j] = make_float4(c.x, c.y, d.x, d.y);
        }
    }
    if (st + 1 <= lg) {
        const int h = 1 << (st - 1), s1 = L >> st, s2 = L >> (st + 1);
        for (int j = threadIdx.x; j < h; j += blockDim.x) {
            const float2 a = tw[j * s1], b = tw[j * s2];
            tws[base + j] = make_float4(a.x, a.y, b.x, b.y);
        }
        st += 2;
    }
    if (st == lg)
        for (int j = threadIdx.x; j < (L >> 1); j += blockDim.x) tw_s[j] = tw[j];
}

// DCT-II values X0[k], X1[k] of the row pair packed in one forward FFT Z (Zr = Z[(L - k) mod L], p = ph[k] =
// (cos, sin)(pi k / 2L)): V0 = (Z + conj Zr) / 2, V1 = (Z - conj Zr) / 2i, X = Re(conj(p) V)
__device__ __forceinline__ void dct_pair_at(float2 Z, float2 Zr, float2 p, float& X0, float& X1) {
    const float v0r = 0.5f * (Z.x + Zr.x), v0i = 0.5f * (Z.y - Zr.y);
    const float v1r = 0.5f * (Z.y + Zr.y), v1i = -0.5f * (Z.x - Zr.x);
    X0 = p.x * v0r + p.y * v0i;
    X1 = p.x * v1r + p.y * v1i;
}

// DCT-II of the rows r0 = 2 blockIdx.x, r0 + 1 (real, length L) -> out rows; spectrum of the pair from one FFT
__device__ void dct2_pair_to_lds(const float* in0, const float* in1, float2* s, const float2* tw_s, const float4* tws,
                                 const float2* ph, int L, int lg, float* X0, float* X1) {
    for (int m = threadIdx.x; m < L; m += blockDim.x)
        s[lpad(bitrev(makhoul(m, L), lg))] = make_float2(in0[m], in1[m]);
    __syncthreads();
    fft_lds(s, tw_s, tws, L, lg, false);
    for (int k = threadIdx.x; k < L; k += blockDim.x)
        dct_pair_at(s[lpad(k)], s[lpad((L - k) & (L - 1))], ph[k], X0[k], X1[k]);
}

__device__ __forceinline__ void pass_a_body(const float* rho, int ld_in, float* T, int ld_out, const float2* tw,
                                            const float2* ph, int L, int lg, int wg) {
    extern __shared__ float2 lds[];
    float2* s = lds;
    float2* tw_s = lds + lds_len(L);
    float4* tws = (float4*)(lds + tws_off(L));
    load_twiddles(tw_s, tws, tw, L, lg);
    const int r0 = 2 * wg;
    __syncthreads();
    dct2_pair_to_lds(rho + (long long)r0 * ld_in, rho + (long long)(r0 + 1) * ld_in, s, tw_s, tws, ph, L, lg,
                     T + (long long)r0 * ld_out, T + (long long)(r0 + 1) * ld_out);
}

__global__ __launch_bounds__(FB) void fft_pass_a(const float* rho, int ld_in, float* T, int ld_out, const float2* tw,
                                                 const float2* ph, int L, int lg) {
    pass_a_body(rho, ld_in, T, ld_out, tw, ph, L, lg, blockIdx.x);
}

// The cosine series of a_cos and the sine series of a_sin (both real, length L) from one inverse FFT: its input at k
// is h[k] = Vc[k] + i Vs[k], Vc / Vs Makhoul's inverse spectra of a_cos and of the reversed a_sin, which need the
// coefficients at k and at L - k. A lane therefore owns pair units u = (k, L - k), k = lane + u FB <= L / 2, and
// keeps the coefficients of its units in registers: no LDS beyond the FFT's own.
constexpr int FFT_UNITS = (4096 / 2 + FB) / FB;   // pair units per lane for L <= 4096

// h[k] from X'[k] - i X'[L-k] of the cosine input (xc, xcr) and of the reversed sine input (xs, xsr)
__device__ __forceinline__ void cs_put(float2* s, const float2* ph, int lg, int k, float xc, float xcr, float xs,
                                       float xsr) {
    const float2 p = ph[k];   // e^{i pi k / 2L}
    const float2 vc = cmul(p, make_float2(xc, -xcr)), vs = cmul(p, make_float2(xs, -xsr));
    s[lpad(bitrev(k, lg))] = make_float2(vc.x - vs.y, vc.y + vs.x);   // vc + i vs
}
// the unit (k, L - k): c / cr = a_cos[k] / a_cos[L-k], sv / sr = a_sin[k] / a_sin[L-k] (X'[0] = 2 a_cos[0] and 0 for
// the reversed sine input; at k = L / 2 both indices are the same)
__device__ __forceinline__ void cs_put_unit(float2* s, const float2* ph, int L, int lg, int k, float c, float cr,
                                            float sv, float sr) {
    if (k == 0) {
        cs_put(s, ph, lg, 0, 2.f * c, 0.f, 0.f, 0.f);
    } else if (2 * k == L) {
        cs_put(s, ph, lg, k, c, c, sv, sv);
    } else {
        cs_put(s, ph, lg, k, c, cr, sr, sv);
        cs_put(s, ph, lg, L - k, cr, c, sv, sr);
    }
}
// after the inverse FFT of h: y_cos[j], y_sin[j] (Makhoul's output order), or the pairs (y_cos[j], y_sin[j]) into
// y_pair when it is given
__device__ __forceinline__ void cs_store(const float2* s, int L, float* y_cos, float* y_sin, float2* y_pair = nullptr) {
    for (int j = threadIdx.x; j < L; j += blockDim.x) {
        const float2 y = s[lpad(makhoul(j, L))];
        const float c = 0.5f * y.x, sn = (j & 1) ? -0.5f * y.y : 0.5f * y.y;
        if (y_pair) {
            y_pair[j] = make_float2(c, sn);
        } else {
            y_cos[j] = c;
            y_sin[j] = sn;
        }
    }
}

struct PassB {
    const float* Tt;       // [M][N]
    float* Ut;             // [M][N]  sine series along x of Ax
    float* Vt;             // [M][N]  cosine series along x of Ay
    const float* omega;    // [N]  pi k / (N bx)
    const float* nu;       // [M]  pi l / (M by)
    float scale;           // 4 / (N M)
    float bx;
    int N, lgN;
    const float2* tw;
    const float2* ph;
    int col;               // > 0: Tt is T itself, [N][col] (pass A's rows), read by column pairs (no transpose)
};

// the spectral weights of row l at x-frequency k: Ay = nu / den C, Ax = omega / den C
__device__ __forceinline__ void weights_at(const PassB& a, int k, int l, float nu, float C, float& ay, float& ax) {
    const float w = a.omega[k], den = (w * w + nu * nu) * a.bx;
    const float inv = den > 0.f ? a.scale / den : 0.f;
    ay = nu * inv * (k == 0 ? 0.5f : 1.f) * C;
    ax = w * inv * (l == 0 ? 0.5f : 1.f) * C;
}

template <int U>
__device__ __forceinline__ void pass_b_body(const PassB& a, int wg) {
    extern __shared__ float2 lds[];
    const int L = a.N, lg = a.lgN, half = L >> 1;
    float2* s = lds;
    float2* tw_s = lds + lds_len(L);
    float4* tws = (float4*)(lds + tws_off(L));
    load_twiddles(tw_s, tws, a.tw, L, lg);
    const int l0 = 2 * wg;
    if (a.col) {
        // columns l0, l0 + 1 of T: one 8-byte load per row
        for (int m = threadIdx.x; m < L; m += blockDim.x)
            s[lpad(bitrev(makhoul(m, L), lg))] = *(const float2*)(a.Tt + (long long)m * a.col + l0);
    } else {
        const float* in0 = a.Tt + (long long)l0 * L;
        const float* in1 = in0 + L;
        for (int m = threadIdx.x; m < L; m += blockDim.x)
            s[lpad(bitrev(makhoul(m, L), lg))] = make_float2(in0[m], in1[m]);
    }
    __syncthreads();
    fft_lds(s, tw_s, tws, L, lg, false);
    // DCT-II of both rows at this lane's units: c[u][row][0 at k, 1 at L - k]
    float c[U][2][2];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int k = threadIdx.x + u * FB;
        c[u][0][0] = c[u][0][1] = c[u][1][0] = c[u][1][1] = 0.f;
        if (k > half) continue;
        const float2 Zk = s[lpad(k)], Zr = s[lpad((L - k) & (L - 1))];
        dct_pair_at(Zk, Zr, a.ph[k], c[u][0][0], c[u][1][0]);
        if (k && 2 * k != L) dct_pair_at(Zr, Zk, a.ph[L - k], c[u][0][1], c[u][1][1]);
    }
    __syncthreads();   // s is the inverse FFT's from here on
#pragma unroll
    for (int q = 0; q < 2; q++) {   // unrolled: c stays in registers
        const int l = l0 + q;
        const float nu = a.nu[l];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = threadIdx.x + u * FB;
            if (k > half) continue;
            float ay, ax, ayr = 0.f, axr = 0.f;
            weights_at(a, k, l, nu, c[u][q][0], ay, ax);
            if (k && 2 * k != L) weights_at(a, L - k, l, nu, c[u][q][1], ayr, axr);
            cs_put_unit(s, a.ph, L, lg, k, ay, ayr, ax, axr);   // cosine series of Ay, sine series of Ax
        }
        __syncthreads();
        fft_lds(s, tw_s, tws, L, lg, true);
        cs_store(s, L, a.Vt + (long long)l * L, a.Ut + (long long)l * L);
        __syncthreads();
    }
}

// the column-reading passes run consecutive workgroups on the same XCD, so the rows' lines they share meet in one L2
// (U: pair units per lane, (L / 2) / FB + 1 for the grid's side)
template <int U>
__global__ __launch_bounds__(FB) void fft_pass_b(PassB a) {
    pass_b_body<U>(a, a.col ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x);
}

// row n of U, V (the batch's grouped launches, after their transpose): one inverse FFT
__device__ __forceinline__ void pass_c_body(const float* U, const float* V, float* fx, float* fy, float2* f2,
                                            int ld_out, const float2* tw, const float2* ph, int L, int lg, int n) {
    extern __shared__ float2 lds[];
    const int half = L >> 1;
    float2* s = lds;
    float2* tw_s = lds + lds_len(L);
    float4* tws = (float4*)(lds + tws_off(L));
    load_twiddles(tw_s, tws, tw, L, lg);
    const float* u_row = U + (long long)n * L;
    const float* v_row = V + (long long)n * L;
#pragma unroll
    for (int u = 0; u < FFT_UNITS; u++) {
        const int k = threadIdx.x + u * FB;
        if (k > half) continue;
        const bool two = k && 2 * k != L;
        cs_put_unit(s, ph, L, lg, k, u_row[k], two ? u_row[L - k] : 0.f, v_row[k], two ? v_row[L - k] : 0.f);
    }
    __syncthreads();
    fft_lds(s, tw_s, tws, L, lg, true);
    if (f2)
        cs_store(s, L, nullptr, nullptr, f2 + (long long)n * ld_out);
    else
        cs_store(s, L, fx + (long long)n * ld_out, fy + (long long)n * ld_out);
}

// Rows n0 = 2 wg, n0 + 1 of U, V read in place as columns n0, n0 + 1 of pass B's outputs Ut, Vt [L][col] (no
// transpose): one 8-byte load per element pair, kept in registers; then one inverse FFT per row, as pass B does its two
template <int NU>
__device__ __forceinline__ void pass_c_pair_body(const float* U, const float* V, float* fx, float* fy, float2* f2,
                                                 int ld_out, const float2* tw, const float2* ph, int L, int lg, int wg,
                                                 int col) {
    extern __shared__ float2 lds[];
    const int half = L >> 1;
    float2* s = lds;
    float2* tw_s = lds + lds_len(L);
    float4* tws = (float4*)(lds + tws_off(L));
    load_twiddles(tw_s, tws, tw, L, lg);
    const int n0 = 2 * wg;
    float2 cu[NU][2], cv[NU][2];   // rows (n0, n0 + 1) of U and V at k [0] and at L - k [1]
#pragma unroll
    for (int u = 0; u < NU; u++) {
        const int k = threadIdx.x + u * FB;
        cu[u][0] = cu[u][1] = cv[u][0] = cv[u][1] = make_float2(0.f, 0.f);
        if (k > half) continue;
        const long long a0 = (long long)k * col + n0;
        cu[u][0] = *(const float2*)(U + a0);
        cv[u][0] = *(const float2*)(V + a0);
        if (k && 2 * k != L) {
            const long long a1 = (long long)(L - k) * col + n0;
            cu[u][1] = *(const float2*)(U + a1);
            cv[u][1] = *(const float2*)(V + a1);
        }
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {   // unrolled: the pairs stay in registers
#pragma unroll
        for (int u = 0; u < NU; u++) {
            const int k = threadIdx.x + u * FB;
            if (k > half) continue;
            cs_put_unit(s, ph, L, lg, k, q ? cu[u][0].y : cu[u][0].x, q ? cu[u][1].y : cu[u][1].x,
                        q ? cv[u][0].y : cv[u][0].x, q ? cv[u][1].y : cv[u][1].x);
        }
        __syncthreads();
        fft_lds(s, tw_s, tws, L, lg, true);
        const int n = n0 + q;
        if (f2)
            cs_store(s, L, nullptr, nullptr, f2 + (long long)n * ld_out);
        else
            cs_store(s, L, fx + (long long)n * ld_out, fy + (long long)n * ld_out);
        __syncthreads();
    }
}

// U, V: pass B's outputs Ut, Vt [L][col], each workgroup two columns
template <int NU>
__global__ __launch_bounds__(FB) void fft_pass_c(const float* U, const float* V, float* fx, float* fy, float2* f2,
                                                 int ld_out, const float2* tw, const float2* ph, int L, int lg,
                                                 int col) {
    pass_c_pair_body<NU>(U, V, fx, fy, f2, ld_out, tw, ph, L, lg, xcd_remap(blockIdx.x, gridDim.x), col);
}

// out[c][r] = in[r][c] for 2 planes (blockIdx.z), 32 x 32 tiles through LDS
constexpr int TT = 32;
__device__ __forceinline__ void transpose_body(const float* in0, const float* in1, float* out0, float* out1, int R,
                                               int C, int bx, int by, int bz) {
    __shared__ float tile[TT][TT + 1];
    const float* in = bz ? in1 : in0;
    float* out = bz ? out1 : out0;
    const int c0 = bx * TT, r0 = by * TT;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int k = ty; k < TT; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < R && c < C) tile[k][tx] = in[(long long)r * C + c];
    }
    __syncthreads();
    for (int k = ty; k < TT; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < R && c < C) out[(long long)c * R + r] = tile[tx][k];
    }
}

__global__ __launch_bounds__(256) void transpose_kernel(const float* in0, const float* in1, float* out0, float* out1,
                                                        int R, int C) {
    transpose_body(in0, in1, out0, out1, R, C, blockIdx.x, blockIdx.y, blockIdx.z);
}

// ------------------------------------------------------------------------------------------------ grouped launches
// Many grids (of any power-of-two sizes) per launch: workgroup b does map[b] = (grid, x, y, z) of its grid's pass.
__device__ __forceinline__ bool fft_skip(const FftTab& t) { return t.S && t.S[STOP] != 0; }

__global__ __launch_bounds__(FB) void fft_a_grouped(const FftTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    const FftTab t = tab[m.x];
    if (fft_skip(t)) return;
    pass_a_body(t.rho, t.rho_ld, t.w0, t.M, t.twM, t.phM, t.M, t.lgM, m.y);
}
__global__ __launch_bounds__(256) void fft_t1_grouped(const FftTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    const FftTab t = tab[m.x];
    if (fft_skip(t)) return;
    transpose_body(t.w0, t.w0, t.w1, t.w1, t.N, t.M, m.y, m.z, 0);
}
__global__ __launch_bounds__(FB) void fft_b_grouped(const FftTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    const FftTab t = tab[m.x];
    if (fft_skip(t)) return;
    const PassB b{t.w1, t.w2, t.w3, t.omega, t.nu, t.scale, t.bx, t.N, t.lgN, t.twN, t.phN, 0};
    pass_b_body<FFT_UNITS>(b, m.y);
}
__global__ __launch_bounds__(256) void fft_t2_grouped(const FftTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    const FftTab t = tab[m.x];
    if (fft_skip(t)) return;
    transpose_body(t.w2, t.w3, t.w0, t.w1, t.M, t.N, m.y, m.z, m.w);
}
__global__ __launch_bounds__(FB) void fft_c_grouped(const FftTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    const FftTab t = tab[m.x];
    if (fft_skip(t)) return;
    pass_c_body(t.w0, t.w1, t.fx, t.fy, t.f2, t.fld, t.twM, t.phM, t.M, t.lgM, m.y);
}

// The FFT kernels' dynamic LDS limit: the largest side's need (above the 64 KB default from side 4096 on), set once
static int fft_lds_limit() {
    static bool done = false;
    if (done) return AP_OK;
    int lg = 0;
    while ((1 << lg) < Poisson::fft_max) lg++;
    const int bytes = fft_slots(Poisson::fft_max, lg) * (int)sizeof(float2);
    const void* fn[12] = {(const void*)fft_pass_a,      (const void*)fft_pass_b<1>,  (const void*)fft_pass_b<2>,
                          (const void*)fft_pass_b<3>,   (const void*)fft_pass_b<5>,  (const void*)fft_pass_c<1>,
                          (const void*)fft_pass_c<2>,   (const void*)fft_pass_c<3>,  (const void*)fft_pass_c<5>,
                          (const void*)fft_a_grouped,   (const void*)fft_b_grouped,  (const void*)fft_c_grouped};
    for (const void* f : fn) AP_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    done = true;
    return AP_OK;
}

int FftGroup::create(const std::vector<FftTab>& tabs) {
    if (int st = fft_lds_limit()) return st;
    std::vector<int4> m[5];
    for (int g = 0; g < (int)tabs.size(); g++) {
        const FftTab& t = tabs[g];
        AP_REQUIRE(t.N >= 2 && t.M >= 2 && (t.N & (t.N - 1)) == 0 && (t.M & (t.M - 1)) == 0,
                   "fft group: grid %d is not a power of two", g);
        for (int y = 0; y < t.N / 2; y++) m[0].push_back(make_int4(g, y, 0, 0));
        for (int y = 0; y < (int)grid_of(t.N, TT); y++)
            for (int x = 0; x < (int)grid_of(t.M, TT); x++) m[1].push_back(make_int4(g, x, y, 0));
        for (int y = 0; y < t.M / 2; y++) m[2].push_back(make_int4(g, y, 0, 0));
        for (int z = 0; z < 2; z++)
            for (int y = 0; y < (int)grid_of(t.M, TT); y++)
                for (int x = 0; x < (int)grid_of(t.N, TT); x++) m[3].push_back(make_int4(g, x, y, z));
        for (int y = 0; y < t.N; y++) m[4].push_back(make_int4(g, y, 0, 0));
        lds_ac = std::max(lds_ac, (size_t)fft_slots(t.M, t.lgM) * sizeof(float2));
        lds_b = std::max(lds_b, (size_t)fft_slots(t.N, t.lgN) * sizeof(float2));
    }
    for (int k = 0; k < 5; k++) {
        count[k] = (int)m[k].size();
        if (!count[k]) continue;
        AP_HIP(dev_malloc(&map[k], m[k].size() * sizeof(int4)));
        AP_HIP(hipMemcpy(map[k], m[k].data(), m[k].size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    return AP_OK;
}

void FftGroup::destroy() {
    for (int k = 0; k < 5; k++) {
        if (map[k]) (void)dev_free(map[k]);
        map[k] = nullptr;
        count[k] = 0;
    }
}

int FftGroup::launch(const FftTab* d_tab, hipStream_t s) const {
    if (!count[0]) return AP_OK;
    hipLaunchKernelGGL(fft_a_grouped, dim3(count[0]), dim3(FB), lds_ac, s, d_tab, map[0]);
    AP_LAUNCHED();
    hipLaunchKernelGGL(fft_t1_grouped, dim3(count[1]), dim3(256), 0, s, d_tab, map[1]);
    AP_LAUNCHED();
    hipLaunchKernelGGL(fft_b_grouped, dim3(count[2]), dim3(FB), lds_b, s, d_tab, map[2]);
    AP_LAUNCHED();
    hipLaunchKernelGGL(fft_t2_grouped, dim3(count[3]), dim3(256), 0, s, d_tab, map[3]);
    AP_LAUNCHED();
    hipLaunchKernelGGL(fft_c_grouped, dim3(count[4]), dim3(FB), lds_ac, s, d_tab, map[4]);
    AP_LAUNCHED();
    return AP_OK;
}

FftTab Poisson::fft_tab(const float* rho, int rho_ld, float* fx, float* fy, float2* f2, int fld,
                        const double* S) const {
    const size_t nm = (size_t)N * M;
    FftTab t{};
    t.S = S;
    t.rho = rho;
    t.rho_ld = rho_ld;
    t.fx = fx;
    t.fy = fy;
    t.f2 = f2;
    t.fld = fld;
    t.w0 = d_work;
    t.w1 = d_work + nm;
    t.w2 = d_work + 2 * nm;
    t.w3 = d_work + 3 * nm;
    t.twN = d_tw;
    t.phN = d_tw + N / 2;
    t.twM = d_tw + N / 2 + N;
    t.phM = t.twM + M / 2;
    t.omega = d_omega;
    t.nu = d_omega + N;
    t.N = N;
    t.M = M;
    int l = 0;
    while ((1 << l) < N) l++;
    t.lgN = l;
    l = 0;
    while ((1 << l) < M) l++;
    t.lgM = l;
    t.scale = (float)(4.0 / ((double)N * M));
    t.bx = (float)bx;
    return t;
}

// ------------------------------------------------------------------------------------------------ the plan

bool Poisson::fft_supported(int n) { return n >= 2 && n <= fft_max && (n & (n - 1)) == 0; }
bool Poisson::fft_max_ok() { return true; }

static int ilog2(int n) {
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}

int Poisson::fft_create() {
    if (int st = fft_lds_limit()) return st;
    // per side: twiddles e^{-2 pi i j / L} (L / 2) then phases e^{i pi k / 2L} (L); omega / nu tables
    std::vector<float2> h;
    auto add = [&](int L) {
        for (int j = 0; j < L / 2; j++)
            h.push_back(make_float2((float)cos(-2 * M_PI * j / L), (float)sin(-2 * M_PI * j / L)));
        for (int k = 0; k < L; k++)
            h.push_back(make_float2((float)cos(M_PI * k / (2.0 * L)), (float)sin(M_PI * k / (2.0 * L))));
    };
    add(N);
    add(M);
    std::vector<float> om(N), nu(M);
    for (int k = 0; k < N; k++) om[k] = (float)(M_PI * k / (N * bx));
    for (int l = 0; l < M; l++) nu[l] = (float)(M_PI * l / (M * by));
    AP_HIP(dev_malloc(&d_tw, h.size() * sizeof(float2)));
    AP_HIP(hipMemcpy(d_tw, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice));
    AP_HIP(dev_malloc(&d_omega, (N + M) * sizeof(float)));
    AP_HIP(hipMemcpy(d_omega, om.data(), N * sizeof(float), hipMemcpyHostToDevice));
    AP_HIP(hipMemcpy(d_omega + N, nu.data(), M * sizeof(float), hipMemcpyHostToDevice));
    return AP_OK;
}

void Poisson::fft_destroy() {
    if (d_tw) (void)dev_free(d_tw);
    if (d_omega) (void)dev_free(d_omega);
    d_tw = nullptr;
    d_omega = nullptr;
}

int Poisson::fft_solve(const float* rho, int rho_ld, float* fx, float* fy, float2* f2, int fld, hipStream_t s) {
    const size_t nm = (size_t)N * M;
    const int lgN = ilog2(N), lgM = ilog2(M);
    const float2* twN = d_tw;
    const float2* phN = d_tw + N / 2;
    const float2* twM = d_tw + N / 2 + N;
    const float2* phM = twM + M / 2;
    float *w0 = d_work, *w1 = d_work + nm, *w2 = d_work + 2 * nm, *w3 = d_work + 3 * nm;
    // pass A: T = rho Cy^T (rows of length M), into w0
    const size_t lds_m = (size_t)fft_slots(M, lgM) * sizeof(float2);
    const size_t lds_n = (size_t)fft_slots(N, lgN) * sizeof(float2);
    hipLaunchKernelGGL(fft_pass_a, dim3(N / 2), dim3(FB), lds_m, s, rho, rho_ld, w0, M, twM, phM, M, lgM);
    AP_LAUNCHED();
    // pass B: columns of T (length N), read in place by pairs -> Ut (w2), Vt (w3); no transpose
    PassB b{w0, w2, w3, d_omega, d_omega + N, (float)(4.0 / ((double)N * M)), (float)bx, N, lgN, twN, phN, M};
    // the pair units per lane a side needs ((L / 2) / FB + 1), as a template argument: registers for those only
    auto units = [](int L) { return (L / 2) / FB + 1; };
    const int ub = units(N), uc = units(M);
    if (ub <= 1)
        hipLaunchKernelGGL(fft_pass_b<1>, dim3(M / 2), dim3(FB), lds_n, s, b);
    else if (ub == 2)
        hipLaunchKernelGGL(fft_pass_b<2>, dim3(M / 2), dim3(FB), lds_n, s, b);
    else if (ub == 3)
        hipLaunchKernelGGL(fft_pass_b<3>, dim3(M / 2), dim3(FB), lds_n, s, b);
    else
        hipLaunchKernelGGL(fft_pass_b<5>, dim3(M / 2), dim3(FB), lds_n, s, b);
    AP_LAUNCHED();
    // pass C: columns of Ut, Vt (rows of U, V) -> field; no transpose
    if (uc <= 1)
        hipLaunchKernelGGL(fft_pass_c<1>, dim3(N / 2), dim3(FB), lds_m, s, w2, w3, fx, fy, f2, fld, twM, phM, M, lgM, N);
    else if (uc == 2)
        hipLaunchKernelGGL(fft_pass_c<2>, dim3(N / 2), dim3(FB), lds_m, s, w2, w3, fx, fy, f2, fld, twM, phM, M, lgM, N);
    else if (uc == 3)
        hipLaunchKernelGGL(fft_pass_c<3>, dim3(N / 2), dim3(FB), lds_m, s, w2, w3, fx, fy, f2, fld, twM, phM, M, lgM, N);
    else
        hipLaunchKernelGGL(fft_pass_c<5>, dim3(N / 2), dim3(FB), lds_m, s, w2, w3, fx, fy, f2, fld, twM, phM, M, lgM, N);
    AP_LAUNCHED();
    return AP_OK;
}

}  // namespace ap
