// Problem preparation on the device (C ABI part 5): the free-area fraction of a density grid
// (src/database.py:103-134 of the reference).
//
// The usable area is the union of the rows minus the union of the blockers, given as a mask over the grid of all
// box edges (xs x ys, computed once per design on the host). A bin's usable area is the sum over the edge-grid
// cells it overlaps of (cell usable) x (x overlap) x (y overlap); each bin finds the few edge intervals under it by
// binary search and sums them in a fixed order (x interval, then y interval), in FP64. The bin edges and the
// interval overlaps are the host form's, operation for operation; the sums run in another order than the host's
// matrix products, so the two agree to rounding (about 1e-16 of a bin).
#include "gp_common.h"

namespace ap {

constexpr int PREP_BLOCK = 256;

// the k-th of nb equal bin edges of [lo, hi], as the host computes it: lo + (hi - lo) * k / nb
__device__ __forceinline__ double bin_edge(double lo, double hi, int k, int nb) {
    return lo + (hi - lo) * (double)k / (double)nb;
}

// first index i with edges[i] > v (edges ascending), in [0, n]
__device__ __forceinline__ int upper_index(const long long* edges, int n, double v) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((double)edges[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(PREP_BLOCK) void free_fraction_kernel(const long long* xs, int nx, const long long* ys,
                                                                   int ny, const unsigned char* usable, double x0,
                                                                   double y0, double x1, double y1, int nbx, int nby,
                                                                   double* out) {
    const int bin = blockIdx.x * PREP_BLOCK + threadIdx.x;
    if (bin >= nbx * nby) return;
    const int k = bin / nby, l = bin % nby;
    const double bx0 = bin_edge(x0, x1, k, nbx), bx1 = bin_edge(x0, x1, k + 1, nbx);
    const double by0 = bin_edge(y0, y1, l, nby), by1 = bin_edge(y0, y1, l + 1, nby);
    // edge intervals [xs[i], xs[i+1]) that can overlap [bx0, bx1]: i from the last edge <= bx0 to the first >= bx1
    const int i0 = max(upper_index(xs, nx, bx0) - 1, 0), i1 = min(upper_index(xs, nx, bx1), nx - 1);
    const int j0 = max(upper_index(ys, ny, by0) - 1, 0), j1 = min(upper_index(ys, ny, by1), ny - 1);
    double area = 0.0;
    for (int i = i0; i < i1; i++) {
        const double ox = fmax(fmin((double)xs[i + 1], bx1) - fmax((double)xs[i], bx0), 0.0);
        if (ox == 0.0) continue;
        double col = 0.0;
        for (int j = j0; j < j1; j++) {
            if (!usable[(size_t)i * (ny - 1) + j]) continue;
            col += fmax(fmin((double)ys[j + 1], by1) - fmax((double)ys[j], by0), 0.0);
        }
        area += ox * col;
    }
    const double bin_area = (x1 - x0) / (double)nbx * (y1 - y0) / (double)nby;   // the host's order
    out[bin] = fmin(fmax(1.0 - area / bin_area, 0.0), 1.0);
}

}  // namespace ap

using namespace ap;

extern "C" {

int ap_free_fraction(const long long* xs, int nx, const long long* ys, int ny, const unsigned char* usable, double x0,
                     double y0, double x1, double y1, int nbx, int nby, double* out, void* stream) {
    AP_REQUIRE(xs && ys && usable && out && nx >= 2 && ny >= 2, "free_fraction: bad edge grid");
    AP_REQUIRE(nbx > 0 && nby > 0 && (long long)nbx * nby < (1ll << 31), "free_fraction: bad grid %d x %d", nbx, nby);
    AP_REQUIRE(x1 > x0 && y1 > y0, "free_fraction: empty region");
    hipLaunchKernelGGL(free_fraction_kernel, dim3(grid_of((long long)nbx * nby, PREP_BLOCK)), dim3(PREP_BLOCK), 0,
                       (hipStream_t)stream, xs, nx, ys, ny, usable, x0, y0, x1, y1, nbx, nby, out);
    AP_LAUNCHED();
    return AP_OK;
}

}  // extern "C"
