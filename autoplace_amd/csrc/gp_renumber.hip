// Spatial renumbering of a level for the engine: the nodes along the Hilbert curve of their bins, the nets by their
// first cell in the new order, and the pins of every cell. No reference counterpart: the reference's kernels read
// nodes in the problem's order (the engine's scatter, WA and gather run faster when index neighbours are spatial
// neighbours).
//
//   keys     one thread per node: the bin (x / bx as x * (1 / bx) in single precision, floored and clamped to the
//            grid), its index along the Hilbert curve of the 2^order square grid, and the filler flag above it
//   perm     a stable radix sort of (key, node): cells in curve order, then fillers in curve order
//   nets     new cell of every pin; per net the smallest one (num_mov + 1 for a net without pins); a stable sort
//            of (key, net) orders the nets, a scan of their degrees gives the new net starts, and one thread per
//            new pin finds its net by binary search and copies the pin's new cell and offset
//   cells    a stable sort of (new cell, new pin) lists the pins of every cell in pin order; cell_start[i] is
//            the first entry of cell i (a binary search), the fixed pins (cell num_mov) last
// Every sort is a stable radix sort of integer keys, so the result is one permutation, the same on every run.
#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "gp_common.h"

namespace ap {

constexpr int RN_BLOCK = 256;

// index of bin (x, y) along the Hilbert curve of a 2^order x 2^order grid (the xy -> d walk: per level, the
// quadrant's place, then the quadrant rotated into the standard orientation)
__device__ __forceinline__ unsigned hilbert_index(long long x, long long y, int order) {
    unsigned long long d = 0;
    for (long long s = 1LL << (order - 1); s > 0; s >>= 1) {
        const long long rx = (x & s) > 0, ry = (y & s) > 0;
        d += (unsigned long long)(s * s * ((3 * rx) ^ ry));
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - x;
                y = s - 1 - y;
            }
            const long long t = x;
            x = y;
            y = t;
        }
    }
    return (unsigned)d;
}

__global__ __launch_bounds__(RN_BLOCK) void rn_node_keys(const float2* pos, int num_nodes, int num_mov, int nbx,
                                                         int nby, float ibx, float iby, int order, unsigned* key,
                                                         int* idx) {
    const int i = blockIdx.x * RN_BLOCK + threadIdx.x;
    if (i >= num_nodes) return;
    const float2 p = pos[i];
    const float fx = fminf(fmaxf(floorf(p.x * ibx), 0.f), (float)(nbx - 1));
    const float fy = fminf(fmaxf(floorf(p.y * iby), 0.f), (float)(nby - 1));
    key[i] = hilbert_index((long long)fx, (long long)fy, order) | (i >= num_mov ? 1u << (2 * order) : 0u);
    idx[i] = i;
}

// inv[old cell] = new cell; the fixed pins' index num_mov stays
__global__ __launch_bounds__(RN_BLOCK) void rn_inverse(const int* perm, int num_mov, int* inv) {
    const int i = blockIdx.x * RN_BLOCK + threadIdx.x;
    if (i < num_mov) inv[perm[i]] = i;
    if (i == 0) inv[num_mov] = num_mov;
}

__global__ __launch_bounds__(RN_BLOCK) void rn_new_cell(const int* pin_cell, int num_pins, const int* inv,
                                                        int* new_cell) {
    const int p = blockIdx.x * RN_BLOCK + threadIdx.x;
    if (p < num_pins) new_cell[p] = inv[pin_cell[p]];
}

__global__ __launch_bounds__(RN_BLOCK) void rn_net_keys(const int* net_start, int num_nets, const int* new_cell,
                                                        int num_mov, unsigned* key, int* idx) {
    const int e = blockIdx.x * RN_BLOCK + threadIdx.x;
    if (e >= num_nets) return;
    int k = num_mov + 1;
    for (int p = net_start[e], p1 = net_start[e + 1]; p < p1; p++) k = min(k, new_cell[p]);
    key[e] = (unsigned)k;
    idx[e] = e;
}

// degree of every new net; entry num_nets 0 (the scan's last value is then the pin count)
__global__ __launch_bounds__(RN_BLOCK) void rn_degrees(const int* net_start, const int* order, int num_nets,
                                                       int* deg) {
    const int k = blockIdx.x * RN_BLOCK + threadIdx.x;
    if (k < num_nets) {
        const int e = order[k];
        deg[k] = net_start[e + 1] - net_start[e];
    } else if (k == num_nets) {
        deg[k] = 0;
    }
}

// new pin q: its net k is the last with start[k] <= q (empty nets before it share its start), its source pin the
// same place in the old net
__global__ __launch_bounds__(RN_BLOCK) void rn_copy_pins(const int* start, int num_nets, const int* order,
                                                         const int* net_start, const int* new_cell,
                                                         const float2* off, int num_pins, int* cell_out,
                                                         float2* off_out, unsigned* ckey, int* pidx) {
    const int q = blockIdx.x * RN_BLOCK + threadIdx.x;
    if (q >= num_pins) return;
    int lo = 0, hi = num_nets;               // start[lo] <= q < start[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start[mid] <= q) lo = mid; else hi = mid;
    }
    const int src = net_start[order[lo]] + (q - start[lo]);
    const int c = new_cell[src];
    cell_out[q] = c;
    off_out[q] = off[src];
    ckey[q] = (unsigned)c;
    pidx[q] = q;
}

// cell_start[i] = the first sorted entry with cell >= i, for i = 0 .. num_mov
__global__ __launch_bounds__(RN_BLOCK) void rn_cell_starts(const unsigned* sorted_cell, int num_pins, int num_mov,
                                                           int* cell_start) {
    const int i = blockIdx.x * RN_BLOCK + threadIdx.x;
    if (i > num_mov) return;
    int lo = 0, hi = num_pins;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (sorted_cell[mid] < (unsigned)i) lo = mid + 1; else hi = mid;
    }
    cell_start[i] = lo;
}

inline unsigned key_bits(unsigned long long v) {   // bits that hold every value 0 .. v
    unsigned b = 1;
    while (b < 32 && (v >> b)) b++;
    return b;
}

// workspace carved from one cached block (with no block: the bytes it takes)
struct RnWork {
    char* base = nullptr;
    size_t used = 0;
    template <typename T>
    T* take(long long count) {
        T* p = base ? (T*)(base + used) : nullptr;
        used += ((size_t)(count > 0 ? count : 1) * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

}  // namespace ap

using namespace ap;

extern "C" int ap_level_renumber(const float* pos, int num_nodes, int num_mov, int nbx, int nby, double bx,
                                 double by, const int* pin_cell, const float* pin_off, const int* net_start,
                                 int num_nets, int num_pins, int* perm, int* pin_cell_out, float* pin_off_out,
                                 int* net_start_out, int* cell_start, int* cell_pin, void* stream) {
    AP_REQUIRE(pos && perm && net_start && net_start_out && cell_start && num_nodes >= 1 && num_mov >= 0 &&
                   num_mov <= num_nodes && num_nets >= 0 && num_pins >= 0,
               "level_renumber: bad arguments");
    AP_REQUIRE(num_pins == 0 || (pin_cell && pin_off && pin_cell_out && pin_off_out && cell_pin && num_nets > 0),
               "level_renumber: null pin arrays, or pins without nets");
    AP_REQUIRE(nbx >= 1 && nby >= 1 && nbx <= 32768 && nby <= 32768 && bx > 0 && by > 0,
               "level_renumber: grid %d x %d of bins %g x %g outside 1 .. 32768 bins of positive size", nbx, nby, bx,
               by);
    hipStream_t s = (hipStream_t)stream;
    int order = 1;
    while (order < 31 && ((long long)std::max(nbx, nby) - 1) >> order) order++;
    const unsigned node_bits = 2 * order + 1;
    const unsigned cell_bits = key_bits((unsigned long long)num_mov + 1);

    // rocPRIM's temporary storage: the largest of the three sorts and the scan
    size_t t1 = 0, t2 = 0, t3 = 0, t4 = 0;
    AP_HIP(rocprim::radix_sort_pairs(nullptr, t1, (unsigned*)nullptr, (unsigned*)nullptr, (int*)nullptr,
                                     (int*)nullptr, (size_t)num_nodes, 0, node_bits, s));
    AP_HIP(rocprim::radix_sort_pairs(nullptr, t2, (unsigned*)nullptr, (unsigned*)nullptr, (int*)nullptr,
                                     (int*)nullptr, (size_t)num_nets, 0, cell_bits, s));
    AP_HIP(rocprim::radix_sort_pairs(nullptr, t3, (unsigned*)nullptr, (unsigned*)nullptr, (int*)nullptr,
                                     (int*)nullptr, (size_t)num_pins, 0, cell_bits, s));
    AP_HIP(rocprim::exclusive_scan(nullptr, t4, (int*)nullptr, (int*)nullptr, 0, (size_t)num_nets + 1,
                                   rocprim::plus<int>(), s));
    const size_t temp_bytes = std::max(std::max(t1, t2), std::max(t3, t4));
    const long long N = num_nodes, P = num_pins, E = num_nets;
    unsigned *nkey, *nkey_alt, *ekey, *ekey_alt, *ckey, *ckey_alt;
    int *nidx, *inv, *new_cell, *eidx, *eorder, *deg, *pidx;
    void* temp;
    RnWork w;
    auto carve = [&]() {
        w.used = 0;
        nkey = w.take<unsigned>(N);
        nkey_alt = w.take<unsigned>(N);
        nidx = w.take<int>(N);
        inv = w.take<int>((long long)num_mov + 1);
        new_cell = w.take<int>(P);
        ekey = w.take<unsigned>(E);
        ekey_alt = w.take<unsigned>(E);
        eidx = w.take<int>(E);
        eorder = w.take<int>(E);
        deg = w.take<int>(E + 1);
        ckey = w.take<unsigned>(P);
        ckey_alt = w.take<unsigned>(P);
        pidx = w.take<int>(P);
        temp = (void*)w.take<char>((long long)temp_bytes);
    };
    carve();
    AP_HIP(dev_malloc((void**)&w.base, w.used));
    carve();
    const dim3 b(RN_BLOCK);
    int st = AP_OK;
    auto run = [&]() -> int {
        size_t tb = temp_bytes;
        const float ibx = 1.0f / (float)bx, iby = 1.0f / (float)by;
        hipLaunchKernelGGL(rn_node_keys, dim3(grid_of(N, RN_BLOCK)), b, 0, s, (const float2*)pos, num_nodes, num_mov,
                           nbx, nby, ibx, iby, order, nkey, nidx);
        AP_LAUNCHED();
        AP_HIP(rocprim::radix_sort_pairs(temp, tb, nkey, nkey_alt, nidx, perm, (size_t)N, 0, node_bits, s));
        hipLaunchKernelGGL(rn_inverse, dim3(grid_of((long long)num_mov + 1, RN_BLOCK)), b, 0, s, perm, num_mov, inv);
        AP_LAUNCHED();
        if (P > 0) {
            hipLaunchKernelGGL(rn_new_cell, dim3(grid_of(P, RN_BLOCK)), b, 0, s, pin_cell, num_pins, inv, new_cell);
            AP_LAUNCHED();
        }
        if (E > 0) {
            hipLaunchKernelGGL(rn_net_keys, dim3(grid_of(E, RN_BLOCK)), b, 0, s, net_start, num_nets, new_cell,
                               num_mov, ekey, eidx);
            AP_LAUNCHED();
            tb = temp_bytes;
            AP_HIP(rocprim::radix_sort_pairs(temp, tb, ekey, ekey_alt, eidx, eorder, (size_t)E, 0, cell_bits, s));
        }
        hipLaunchKernelGGL(rn_degrees, dim3(grid_of(E + 1, RN_BLOCK)), b, 0, s, net_start, eorder, num_nets, deg);
        AP_LAUNCHED();
        tb = temp_bytes;
        AP_HIP(rocprim::exclusive_scan(temp, tb, deg, net_start_out, 0, (size_t)E + 1, rocprim::plus<int>(), s));
        if (P > 0) {
            hipLaunchKernelGGL(rn_copy_pins, dim3(grid_of(P, RN_BLOCK)), b, 0, s, net_start_out, num_nets, eorder,
                               net_start, new_cell, (const float2*)pin_off, num_pins, pin_cell_out,
                               (float2*)pin_off_out, ckey, pidx);
            AP_LAUNCHED();
            tb = temp_bytes;
            AP_HIP(rocprim::radix_sort_pairs(temp, tb, ckey, ckey_alt, pidx, cell_pin, (size_t)P, 0, cell_bits, s));
        }
        hipLaunchKernelGGL(rn_cell_starts, dim3(grid_of((long long)num_mov + 1, RN_BLOCK)), b, 0, s, ckey_alt,
                           num_pins, num_mov, cell_start);
        AP_LAUNCHED();
        return AP_OK;
    };
    st = run();
    // the workspace back to the cache once the stream is through with it
    const hipError_t e = hipStreamSynchronize(s);
    dev_free_synced(w.base);
    if (st) return st;
    AP_HIP(e);
    return AP_OK;
}
