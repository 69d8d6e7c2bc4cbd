// The GP iteration engine (include/autoplace_gp.h, part 3): whole passes of src/gpu_loop.py:58-64 on the device.
//
// One pass = Nesterov step -> WA -> density (force grid + overflow grid) -> Poisson -> gather -> control.
// Differences from the reference's pass, all in the data movement:
//   * u / v / g are double-buffered; the control kernel flips a device flag instead of select_kernel's 3 copies,
//     and the best placement is copied by the next pass's Nesterov kernel (which reads v anyway);
//   * the density map is built per tile of bins: a counting sort puts every node into the buckets of the tiles its
//     bins fall in, and the tile kernel sums each tile's buckets with LDS integer atomics and writes the tile's
//     density and overflow once. No global atomics, no mirrored copies, no clearing pass (the reference adds every
//     contribution to global memory with an atomic and writes the map 4x for the FFT), and the cost does not depend
//     on the node order;
//   * |v' - v| and |g' - g| are accumulated in FP64 by the kernels that write v' and g', the WA / HPWL totals by the
//     WA kernel and the overflow by the tile kernel, as per-workgroup partials the control kernel sums in a fixed
//     order (deterministic);
//   * the pass runs on two streams: the overflow grid, WA and the cells' WA sums beside the force grid's binning,
//     tiles and Poisson solve (they read the same positions and write buffers of their own); the gather joins them.
// Every kernel reads its parameters (gamma, weight, step) from device memory, so a run of passes has no host sync
// and replays from a HIP graph.
#include <stdio.h>

#include <stdlib.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <functional>
#include <vector>

#include "gp_common.h"
#include "gp_control.h"
#include "gp_kernels.h"
#include "gp_mlp.h"
#include "gp_poisson.h"
#include "gp_wa.h"

namespace ap {

constexpr int EB = 256;   // lanes per workgroup of the element kernels

// pointer pair selected by a device flag: flag == null -> a; *flag == 0 -> a; else b
struct Pick {
    const float2* a;
    const float2* b;
    const int* flag;
    __device__ __forceinline__ const float2* get(bool other) const {
        if (!flag) return a;
        const int f = *flag;
        // select by arithmetic: a select between the two fields becomes an indexed load from a stack copy of the
        // struct (scratch memory)
        const uintptr_t ua = (uintptr_t)a, ub = (uintptr_t)b;
        const uintptr_t m = (other ? f == 0 : f != 0) ? ~(uintptr_t)0 : 0;
        return (const float2*)(ua ^ ((ua ^ ub) & m));
    }
};

struct Bufs {          // the two point sets: set 0 = the caller's u / v / g, set 1 = internal
    float2* u[2];
    float2* v[2];
    float2* g[2];
    float2* best;
    int* flags;        // [0] set holding the current point, [1] best copy pending
};

// ------------------------------------------------------------------------------ Nesterov (+ deferred best copy)
struct NestArgs {
    Bufs B;
    const float2* lo;
    const float2* hi;
    const float* kp;
    int n;
    double* part;
};

// one of two pointers by a device flag, selected by arithmetic: an index into the pointer pair of a struct copy
// becomes an indexed load from a stack copy (scratch memory)
template <typename T>
__device__ __forceinline__ T* pick_ptr(T* a, T* b, bool second) {
    const uintptr_t ua = (uintptr_t)a, ub = (uintptr_t)b, m = second ? ~(uintptr_t)0 : 0;
    return (T*)(ua ^ ((ua ^ ub) & m));
}

__device__ __forceinline__ void nesterov_body(const NestArgs& a, int wg) {
    __shared__ double red[16];
    const Bufs& B = a.B;
    const bool cur = B.flags[0] != 0;
    const bool copy_best = B.flags[1] != 0;
    const float alpha = a.kp[3], coef = a.kp[4];
    const int i = wg * EB + threadIdx.x;
    double m2[1] = {0.0};
    if (i < a.n) {
        const float2 u = pick_ptr(B.u[0], B.u[1], cur)[i], v = pick_ptr(B.v[0], B.v[1], cur)[i];
        const float2 g = pick_ptr(B.g[0], B.g[1], cur)[i], l = a.lo[i], h = a.hi[i];
        float2 un, vn;
        un.x = v.x - alpha * g.x;
        un.y = v.y - alpha * g.y;
        vn.x = fminf(fmaxf(un.x + coef * (un.x - u.x), l.x), h.x);
        vn.y = fminf(fmaxf(un.y + coef * (un.y - u.y), l.y), h.y);
        pick_ptr(B.u[0], B.u[1], !cur)[i] = un;
        pick_ptr(B.v[0], B.v[1], !cur)[i] = vn;
        if (copy_best) B.best[i] = v;
        const double dx = (double)vn.x - v.x, dy = (double)vn.y - v.y;
        m2[0] = dx * dx + dy * dy;
    }
    block_sum<1>(m2, red);
    if (threadIdx.x == 0) a.part[wg] = m2[0];
}

// ------------------------------------------------------------------------------ density: tile binning
// A grid of nbx x nby bins is cut into tiles of T x T bins. A node whose bins fall in at most 2 x 2 tiles has an
// entry (centre, size, ratio) in the bucket of each of those tiles; any other node goes to the grid's list of big
// nodes. A bucket holds the nodes of one tile and one plane (cells, fillers). Per tile (G tiles per workgroup), the
// tile kernel sums the contributions of its buckets' entries and of the big nodes, clipped to the tile, with LDS
// integer atomics, then writes the tile's density and overflow partial sum. The contributions are the reference's
// 2^32 fixed-point integers (contribution() below), so the sums, and the map, do not depend on the order of the
// entries. The entries are placed in one of two forms:
//   fill form (the single-level passes, one launch per pass and grid):
//     eng_bin_fill   every node's entries straight into their buckets: an LDS histogram ranks the workgroup's
//                    entries, one global atomic per workgroup and bucket takes its share of the bucket, and the
//                    entries go into the bucket's room [start[b], start[b+1]). The rooms were set before the pass
//                    from the buckets' counts in the pass before (rooms_body, run by two extra workgroups of the
//                    Nesterov kernel). An entry past its bucket's room adds its contributions, clipped to its tile,
//                    to the grid's spill map with global integer atomics, and the tile kernel folds them in.
//   counting-sort form (the batch's grouped launches): count (the same histogram and atomics), an exclusive scan of
//     the counts, then a scatter of each node's entries to start[b] + its place.
// (eng_over_* : the overflow grid's launches, beside the force grid's)
constexpr int BIN_PER_LANE = 8;
constexpr int BIN_NODES = EB * BIN_PER_LANE;
constexpr int BIN_MAX_BUCKETS = 16384;   // planes x tiles of a grid: the count / fill kernels' LDS histogram, 64 KB
constexpr int SPLIT_SLICE = 4096;        // entries per workgroup above which a tile's bucket is split
constexpr int SPLIT_MAX = 16;            // workgroups that may sum one tile
constexpr int SPLIT_HELPERS = 256;       // helper workgroups of a grid's tile launch
constexpr int SPLIT_FACTOR = 4;          // a split tile holds at least this many times the mean entries per tile
constexpr int SPLIT_LIST = 256;          // marked tiles a pass can split
constexpr int SCAN_T = 1024;

struct BinJob {
    const float2* size;
    const float* ratio;
    int num_nodes, num_mov;    // nodes of the grid; nodes from num_mov on are plane 1
    float inv_bx, inv_by;
    int nbx, nby;
    int T, tiles_y, tiles;     // tile side in bins, tiles per column, tiles
    int T0;                    // the side chosen from the node density at set-up (the autotuner searches around it)
    int G, wgs;                // tiles per workgroup of the tile kernel (1, 2, 4, 8: TB / G lanes each), its
                               // workgroups
    int planes;                // 2: cells + fillers, 1: cells only
    int blocks;                // workgroups of the count / fill and scatter kernels for this grid
    int* count;                // [planes tiles] entries asked of each bucket in the pass. Counting-sort form: zero
                               // between passes. Fill form: kept until the rooms of the next pass are set from them
    int* start;                // [planes tiles + 1] bucket starts (fill form: room of bucket b = start[b + 1] -
                               // start[b])
    int4* off;                 // [num_nodes] place of the node's entry in the bucket of tile (+0,+0) (+0,+1) (+1,+0)
                               // (+1,+1) from its first tile; -1: no entry (counting-sort form)
    float4* geo;               // [room] entries: centre x, y, size x, y
    float* gr;                 // [room] entries: ratio
    int* big_count;            // [1] big nodes listed in the pass (cleared by the scan / by rooms_body)
    int* big_n;                // [1] big nodes of the pass (written by the scan; counting-sort form)
    int* big;                  // [num_nodes] big node ids
    const float* fixed;        // [nbx][nby]
    float* rho;                // density out (null: overflow only)
    int count_over;            // 1: this grid's overflow feeds the control
    int fill;                  // 1: fill form, 0: counting-sort form
    int room;                  // entries geo / gr hold
    int room_pct, room_add;    // fill form: the room of a bucket for the next pass, count x room_pct / 100 + room_add
    unsigned long long* spill; // [planes][nbx][nby] fill form: spilled contributions (zero between passes)
    int* spill_tile;           // [tiles] fill form: 1 when the tile has spilled contributions (zero between passes)
    unsigned long long* spill_n;  // [1] fill form: entries spilled since the level was created
    float* rho_mirror;         // drop-in map (null: none): the total density mirrored in both axes, [2 nbx][2 nby]
    float* over_map;           // drop-in map (null: none): max(cells + blocked - target, 0) per bin, [nbx][nby]
    // fill form, split tiles: a tile whose bucket reaches split_need entries in a pass (marked by the fill kernel as
    // it counts) is summed by its own workgroup and helper workgroups (leading the tile launch), a slice of split_slice
    // entries or more each, at most SPLIT_MAX workgroups a tile
    int split_slice;           // 0: no splitting
    int split_factor;          // split_need: the larger of 2 x split_slice and split_factor x the mean per tile
    int split_need;            // entries of one bucket that mark its tile (0: none)
    int helpers;               // helper workgroups leading the tile launch (0: none)
    int* heavy_n;              // [1] tiles marked in the pass (cleared by the rooms step)
    int* heavy_list;           // [SPLIT_LIST] the marked tiles, in the order they were marked
    int* heavy_pos;            // [tiles] a marked tile's place in heavy_list (SPLIT_LIST: not listed); -1 unmarked
    unsigned* arrive;          // [tiles] participants of a split tile done (zero between passes)
    unsigned long long* split_n;  // [1] split tiles summed since the level was created
};

void split_threshold(BinJob& J);

__device__ __forceinline__ unsigned long long contribution(float wx, const Span& sy, int by) {
    return fixed_point32(wx * span_overlap(sy, by));
}

// first tile of a node and whether its bins stay within 2 x 2 tiles
__device__ __forceinline__ bool small_node(const Span& sx, const Span& sy, int T, int tiles_y, int& t0, int& wx,
                                           int& wy) {
    const int tx0 = sx.b0 / T, ty0 = sy.b0 / T;
    wx = sx.b1 / T - tx0;
    wy = sy.b1 / T - ty0;
    t0 = tx0 * tiles_y + ty0;
    return wx <= 1 && wy <= 1;
}

// the count kernel of the counting-sort form (the batch's grouped launches) and of the fill form's set-up: every
// small node's place in the bucket of each tile its bins fall in (an LDS histogram per workgroup, then one global
// atomic per workgroup and bucket), big nodes to the grid's list
__device__ __forceinline__ void bin_count_body(const Pick& pos_p, const BinJob& f, const BinJob& j, int blk) {
    extern __shared__ int hist[];   // planes x tiles of the grid
    const bool judge = blk >= f.blocks;
    const BinJob J = judge ? j : f;    // by value: a reference to a kernel argument would go to scratch
    const int base = (blk - (judge ? f.blocks : 0)) * BIN_NODES;
    const int nbuck = J.planes * J.tiles;
    for (int t = threadIdx.x; t < nbuck; t += EB) hist[t] = 0;
    __syncthreads();
    const float2* pos = pos_p.get(true);
    int rank[BIN_PER_LANE][4], b0[BIN_PER_LANE];
    float4 nd[BIN_PER_LANE];    // every node's centre and size loaded before the first is used (one memory wait)
#pragma unroll
    for (int q = 0; q < BIN_PER_LANE; q++) {
        const int i = base + q * EB + threadIdx.x;
        const int k = i < J.num_nodes ? i : 0;
        const float2 p = pos[k], sz = J.size[k];
        nd[q] = make_float4(p.x, p.y, sz.x, sz.y);
    }
#pragma unroll
    for (int q = 0; q < BIN_PER_LANE; q++) {
        rank[q][0] = rank[q][1] = rank[q][2] = rank[q][3] = -1;
        b0[q] = 0;
        const int i = base + q * EB + threadIdx.x;
        if (i >= J.num_nodes) continue;
        const float2 p = make_float2(nd[q].x, nd[q].y), sz = make_float2(nd[q].z, nd[q].w);
        const Span sx = make_span(p.x, sz.x, J.inv_bx, J.nbx), sy = make_span(p.y, sz.y, J.inv_by, J.nby);
        if (sx.b0 > sx.b1 || sy.b0 > sy.b1) continue;
        int t0, wx, wy;
        if (!small_node(sx, sy, J.T, J.tiles_y, t0, wx, wy)) {
            J.big[atomicAdd(J.big_count, 1)] = i;
            continue;
        }
        b0[q] = t0 + (i >= J.num_mov ? J.tiles : 0);
        rank[q][0] = atomicAdd(&hist[b0[q]], 1);
        if (wy) rank[q][1] = atomicAdd(&hist[b0[q] + 1], 1);
        if (wx) rank[q][2] = atomicAdd(&hist[b0[q] + J.tiles_y], 1);
        if (wx && wy) rank[q][3] = atomicAdd(&hist[b0[q] + J.tiles_y + 1], 1);
    }
    __syncthreads();
    // this workgroup's share of every bucket it touched: hist[b] becomes its first place in bucket b
    for (int t = threadIdx.x; t < nbuck; t += EB) {
        const int c = hist[t];
        if (c) hist[t] = atomicAdd(&J.count[t], c);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < BIN_PER_LANE; q++) {
        const int i = base + q * EB + threadIdx.x;
        if (i >= J.num_nodes) continue;
        int4 o;
        o.x = rank[q][0] >= 0 ? hist[b0[q]] + rank[q][0] : -1;
        o.y = rank[q][1] >= 0 ? hist[b0[q] + 1] + rank[q][1] : -1;
        o.z = rank[q][2] >= 0 ? hist[b0[q] + J.tiles_y] + rank[q][2] : -1;
        o.w = rank[q][3] >= 0 ? hist[b0[q] + J.tiles_y + 1] + rank[q][3] : -1;
        J.off[i] = o;
    }
}
// the fill form's set-up count (one grid per launch): the counts the first rooms are set from
__global__ __launch_bounds__(EB) void eng_rooms_count(Pick pos_p, BinJob f, BinJob j, int which) {
    bin_count_body(pos_p, f, j, (which ? f.blocks : 0) + xcd_remap(blockIdx.x, gridDim.x));
}

// counting-sort form: the bucket starts from the counts of the pass (exclusive scan by the NT lanes of one
// workgroup, 4 NT buckets a round, coalesced), the counts back to zero for the next pass
template <int NT>
__device__ __forceinline__ void scan_buckets(const BinJob& J) {
    __shared__ int wsum[NT / 64];
    const int n = J.planes * J.tiles;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int carry = 0;                                  // the same in every lane
    for (int base = 0; base < n; base += 4 * NT) {
        const int k0 = base + 4 * threadIdx.x;
        int c[4], sum = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            c[i] = k0 + i < n ? J.count[k0 + i] : 0;
            sum += c[i];
        }
        int x = sum;                                // inclusive scan of the lanes' sums
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        int run = carry + x - sum, total = 0;
        for (int w = 0; w < NT / 64; w++) {
            const int v = wsum[w];
            if (w < wave) run += v;
            total += v;
        }
        __syncthreads();                            // wsum is rewritten by the next round
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (k0 + i >= n) break;
            J.start[k0 + i] = run;
            run += c[i];
            J.count[k0 + i] = 0;
        }
        carry += total;
    }
    if (threadIdx.x == 0) {
        J.start[n] = carry;
        *J.big_n = *J.big_count;
        *J.big_count = 0;
    }
}

// the batch's scan launch (its count launches do not scan): one workgroup per grid
__device__ __forceinline__ void bin_scan_body(const BinJob& f, const BinJob& j, int which) {
    const BinJob J = which ? j : f;
    scan_buckets<SCAN_T>(J);
}

// Fill form: the rooms of the pass from the counts of the last one, by the EB lanes of one workgroup, 4 EB buckets a
// round. Bucket b gets count x room_pct / 100 + room_add places, the starts are the rooms' prefix sums, clamped to
// the entry arrays (sized so that the default rooms always fit; a room cut short only spills more), and the counts
// and the big-node count go back to zero for the pass's fill.
__device__ __forceinline__ void rooms_body(const BinJob& J) {
    __shared__ long long wsum[EB / 64];
    const int n = J.planes * J.tiles;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long cap = J.room;
    long long carry = 0;                            // the same in every lane
    for (int base = 0; base < n; base += 4 * EB) {
        const int k0 = base + 4 * threadIdx.x;      // a multiple of 4: the int4 accesses are aligned
        int4 c = make_int4(0, 0, 0, 0);
        if (k0 + 3 < n) {
            c = *(const int4*)(J.count + k0);
        } else if (k0 < n) {
            c.x = J.count[k0];
            c.y = k0 + 1 < n ? J.count[k0 + 1] : 0;
            c.z = k0 + 2 < n ? J.count[k0 + 2] : 0;
        }
        const int cs[4] = {c.x, c.y, c.z, c.w};
        long long v[4], sum = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            v[i] = k0 + i < n ? (long long)cs[i] * J.room_pct / 100 + J.room_add : 0;
            sum += v[i];
        }
        long long x = sum;                          // inclusive scan of the lanes' sums
        for (int o = 1; o < 64; o <<= 1) {
            const long long y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        long long run = carry + x - sum, total = 0;
        for (int w = 0; w < EB / 64; w++) {
            const long long t = wsum[w];
            if (w < wave) run += t;
            total += t;
        }
        __syncthreads();                            // wsum is rewritten by the next round
        int st[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            st[i] = (int)(run < cap ? run : cap);
            run += v[i];
        }
        if (k0 + 3 < n) {
            *(int4*)(J.start + k0) = make_int4(st[0], st[1], st[2], st[3]);
            *(int4*)(J.count + k0) = make_int4(0, 0, 0, 0);
        } else {
            for (int i = 0; i < 4 && k0 + i < n; i++) {
                J.start[k0 + i] = st[i];
                J.count[k0 + i] = 0;
            }
        }
        carry += total;
    }
    if (threadIdx.x == 0) {
        J.start[n] = (int)(carry < cap ? carry : cap);
        *J.big_count = 0;
        if (J.helpers) *J.heavy_n = 0;
    }
}
// the rooms of both grids before an evaluation that does not follow a Nesterov step (workgroup 0: force grid,
// 1: overflow grid)
__global__ __launch_bounds__(EB) void eng_rooms(BinJob f, BinJob j) { rooms_body(blockIdx.x ? j : f); }
// the Nesterov step of a pass; its first `rooms` workgroups set the rooms of the pass's fill (force grid, then
// overflow grid), beside the step
__global__ __launch_bounds__(EB) void eng_nesterov(NestArgs a, BinJob f, BinJob j, int rooms) {
    if ((int)blockIdx.x < rooms) {
        rooms_body(blockIdx.x ? j : f);
        return;
    }
    nesterov_body(a, blockIdx.x - rooms);
}

__device__ __forceinline__ void bin_scatter_body(const Pick& pos_p, const BinJob& f, const BinJob& j, int blk) {
    const bool judge = blk >= f.blocks;
    const BinJob J = judge ? j : f;    // by value: a reference to a kernel argument would go to scratch
    const int base = (blk - (judge ? f.blocks : 0)) * BIN_NODES;
    const float2* pos = pos_p.get(true);
#pragma unroll 2
    for (int q = 0; q < BIN_PER_LANE; q++) {
        const int i = base + q * EB + threadIdx.x;
        if (i >= J.num_nodes) continue;
        const int4 o = J.off[i];
        if (o.x < 0) continue;                                  // big node, or no bins
        const float2 p = pos[i], sz = J.size[i];
        const Span sx = make_span(p.x, sz.x, J.inv_bx, J.nbx), sy = make_span(p.y, sz.y, J.inv_by, J.nby);
        int t0, wx, wy;
        small_node(sx, sy, J.T, J.tiles_y, t0, wx, wy);
        const int b = t0 + (i >= J.num_mov ? J.tiles : 0);
        const float4 g = make_float4(p.x, p.y, sz.x, sz.y);
        const float r = J.ratio[i];
        int e = J.start[b] + o.x;
        J.geo[e] = g;
        J.gr[e] = r;
        if (o.y >= 0) {
            e = J.start[b + 1] + o.y;
            J.geo[e] = g;
            J.gr[e] = r;
        }
        if (o.z >= 0) {
            e = J.start[b + J.tiles_y] + o.z;
            J.geo[e] = g;
            J.gr[e] = r;
        }
        if (o.w >= 0) {
            e = J.start[b + J.tiles_y + 1] + o.w;
            J.geo[e] = g;
            J.gr[e] = r;
        }
    }
}

// a node's contributions to the bins of the tile [X0, X1] x [Y0, Y1], into the tile's LDS accumulator a (row stride T)
__device__ __forceinline__ void add_clipped(unsigned long long* a, int T, int X0, int Y0, int X1, int Y1,
                                            const Span& sx, const Span& sy, float r) {
    const int x0 = max(sx.b0, X0), x1 = min(sx.b1, X1), y0 = max(sy.b0, Y0), y1 = min(sy.b1, Y1);
    if (x1 - x0 < 3 && y1 - y0 < 3) {
        // at most 3 x 3 bins (every node whose smoothed size is under 2 bins): straight-line code over a fixed
        // 3 x 3 window, the weights of each axis computed once; the same float products as the loop below
        float wx[3], wy[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            wx[k] = x0 + k <= x1 ? r * span_overlap(sx, x0 + k) : 0.f;
            wy[k] = y0 + k <= y1 ? span_overlap(sy, y0 + k) : 0.f;
        }
        unsigned long long* p = a + (x0 - X0) * T + (y0 - Y0);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const unsigned long long c = fixed_point32(wx[i] * wy[j]);
                if (c) atomicAdd(p + i * T + j, c);
            }
        return;
    }
    for (int bx = x0; bx <= x1; bx++) {
        const float wx = r * span_overlap(sx, bx);
        unsigned long long* row = a + (bx - X0) * T - Y0;
        for (int by = y0; by <= y1; by++) {
            const unsigned long long c = contribution(wx, sy, by);
            if (c) atomicAdd(row + by, c);
        }
    }
}

// Fill form: a spilled entry's contributions, clipped to the tile of bucket bk, into the grid's spill map
__device__ __forceinline__ void spill_entry(const BinJob& J, int bk, const Span& sx, const Span& sy, float r) {
    const int plane = bk >= J.tiles, t = bk - plane * J.tiles;
    const int X0 = (t / J.tiles_y) * J.T, Y0 = (t % J.tiles_y) * J.T;
    const int X1 = min(X0 + J.T, J.nbx) - 1, Y1 = min(Y0 + J.T, J.nby) - 1;
    add_clipped(J.spill + ((long long)plane * J.nbx + X0) * J.nby + Y0, J.nby, X0, Y0, X1, Y1, sx, sy, r);
    J.spill_tile[t] = 1;
}

// Fill form, one workgroup of BIN_NODES nodes: the ranks of its entries in an LDS histogram (rank + 1 of the entries
// in tiles (+0,+0) (+0,+1) and (+1,+0) (+1,+1) in two words, 16 bits each: a rank is below BIN_NODES; 0: no entry);
// then hist[b] becomes the workgroup's first place in b's room, or ~place when its entries of b do not all fit (they
// then test their own place against the room); then the entries are written, the ones past their room spilled after
// the others (one bit per entry in a lane mask, so the spill path is one copy of its code).
__device__ __forceinline__ void bin_fill_body(const Pick& pos_p, const BinJob& f, const BinJob& j, int blk) {
    extern __shared__ int hist[];   // planes x tiles of the grid
    const bool judge = blk >= f.blocks;
    const BinJob J = judge ? j : f;    // by value: a reference to a kernel argument would go to scratch
    const int base = (blk - (judge ? f.blocks : 0)) * BIN_NODES;
    const int nbuck = J.planes * J.tiles;
    for (int t = threadIdx.x; t < nbuck; t += EB) hist[t] = 0;
    __syncthreads();
    const float2* pos = pos_p.get(true);
    unsigned rk[BIN_PER_LANE][2];
    float4 nd[BIN_PER_LANE];   // the nodes' centre and size, kept for the second phase (LDS, not registers, bounds the
                               // force grid's occupancy)
    float nr[BIN_PER_LANE];    // and their ratio
    // every node's loads issued before the first is used (one memory wait for the eight, not eight in a row; a lane
    // past the grid's nodes reads node 0 and drops it)
#pragma unroll
    for (int q = 0; q < BIN_PER_LANE; q++) {
        const int i = base + q * EB + threadIdx.x;
        const int k = i < J.num_nodes ? i : 0;
        const float2 p = pos[k], sz = J.size[k];
        nd[q] = make_float4(p.x, p.y, sz.x, sz.y);
        nr[q] = J.ratio[k];
    }
#pragma unroll
    for (int q = 0; q < BIN_PER_LANE; q++) {
        rk[q][0] = rk[q][1] = 0;
        const int i = base + q * EB + threadIdx.x;
        if (i >= J.num_nodes) continue;
        const float4 g = nd[q];
        const Span sx = make_span(g.x, g.z, J.inv_bx, J.nbx), sy = make_span(g.y, g.w, J.inv_by, J.nby);
        if (sx.b0 > sx.b1 || sy.b0 > sy.b1) continue;
        int t0, wx, wy;
        if (!small_node(sx, sy, J.T, J.tiles_y, t0, wx, wy)) {
            J.big[atomicAdd(J.big_count, 1)] = i;
            continue;
        }
        const int b = t0 + (i >= J.num_mov ? J.tiles : 0);
        rk[q][0] = (unsigned)atomicAdd(&hist[b], 1) + 1u;
        if (wy) rk[q][0] |= ((unsigned)atomicAdd(&hist[b + 1], 1) + 1u) << 16;
        if (wx) rk[q][1] = (unsigned)atomicAdd(&hist[b + J.tiles_y], 1) + 1u;
        if (wx && wy) rk[q][1] |= ((unsigned)atomicAdd(&hist[b + J.tiles_y + 1], 1) + 1u) << 16;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nbuck; t += EB) {
        const int c = hist[t];
        if (!c) continue;
        const int at = atomicAdd(&J.count[t], c), lo = J.start[t], hi = J.start[t + 1];
        hist[t] = hi - lo - at >= c ? lo + at : ~(lo + at);
        if (J.helpers && J.split_need && at < J.split_need && at + c >= J.split_need) {   // this bucket's tile is heavy
            const int tt = t >= J.tiles ? t - J.tiles : t;
            if (atomicCAS(J.heavy_pos + tt, -1, -2) == -1) {
                const int idx = atomicAdd(J.heavy_n, 1);
                if (idx < SPLIT_LIST) J.heavy_list[idx] = tt;
                J.heavy_pos[tt] = min(idx, SPLIT_LIST);
            }
        }
    }
    __syncthreads();
    unsigned spilled = 0;   // bit 4 q + k: entry k of node q goes to the spill map
#pragma unroll
    for (int q = 0; q < BIN_PER_LANE; q++) {
        if (!rk[q][0]) continue;                                // no entries: out of range, no bins or a big node
        const int i = base + q * EB + threadIdx.x;
        const float4 g = nd[q];
        const float r = nr[q];
        const Span sx = make_span(g.x, g.z, J.inv_bx, J.nbx), sy = make_span(g.y, g.w, J.inv_by, J.nby);
        int t0, wx, wy;
        small_node(sx, sy, J.T, J.tiles_y, t0, wx, wy);
        const int b = t0 + (i >= J.num_mov ? J.tiles : 0);
        const unsigned rr[4] = {rk[q][0] & 0xFFFFu, rk[q][0] >> 16, rk[q][1] & 0xFFFFu, rk[q][1] >> 16};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!rr[k]) continue;
            const int bk = b + (k & 1) + (k >> 1) * J.tiles_y;
            const int h = hist[bk];
            const int e = (h >= 0 ? h : ~h) + (int)rr[k] - 1;
            if (h >= 0 || e < J.start[bk + 1]) {
                J.geo[e] = g;
                J.gr[e] = r;
            } else {
                spilled |= 1u << (4 * q + k);
            }
        }
    }
    while (spilled) {
        const int bit = __builtin_ctz(spilled);
        spilled &= spilled - 1;
        const int q = bit >> 2, k = bit & 3;
        const int i = base + q * EB + threadIdx.x;
        const float2 p = pos[i], sz = J.size[i];
        const Span sx = make_span(p.x, sz.x, J.inv_bx, J.nbx), sy = make_span(p.y, sz.y, J.inv_by, J.nby);
        int t0, wx, wy;
        small_node(sx, sy, J.T, J.tiles_y, t0, wx, wy);
        const int b = t0 + (i >= J.num_mov ? J.tiles : 0);
        spill_entry(J, b + (k & 1) + (k >> 1) * J.tiles_y, sx, sy, J.ratio[i]);
        atomicAdd(J.spill_n, 1ull);
    }
}
__global__ __launch_bounds__(EB) void eng_bin_fill(Pick pos_p, BinJob f, BinJob j) {
    bin_fill_body(pos_p, f, j, xcd_remap(blockIdx.x, gridDim.x));
}
__global__ __launch_bounds__(EB) void eng_over_fill(Pick pos_p, BinJob f, BinJob j) {
    bin_fill_body(pos_p, f, j, f.blocks + xcd_remap(blockIdx.x, gridDim.x));
}

// Lanes per workgroup of the tile kernel: the single-level passes' (TB_LEVEL) and the batch's grouped launches' (TB).
// A crowded tile's bucket spreads over 8 waves.
constexpr int TB_LEVEL = 512;
constexpr int TB = 512;
constexpr int TILE_PF = 4;  // fixed-map values per lane loaded ahead (the whole tile up to T x T = 4 x lanes)

// One workgroup of NT lanes sums J.G tiles (G = 1, 2, 4, 8), each with its own NT / G lanes (whole waves) and its
// own LDS accumulator, so that small tiles do not each hold a workgroup (and the launch's LDS reservation) to
// themselves. The sub-groups pass the same barriers; a sub-group past the last tile only takes part in them.
// Split tiles (fill form, J.helpers): the tile's workgroup sums the first slice of its bucket's entries, and helper
// workgroups (the launch's first J.helpers, all lanes on one tile, dispatched before the tiles' own so that they run
// beside them) sum the others, as the split plan of the pass's
// rooms step assigned them. Each adds its sums to the grid's spill map with integer atomics and counts itself in
// arrive[t]; the one that arrives last takes the tile's sums back (clearing them) and writes the tile as an unsplit
// tile's workgroup does. Integer sums, so the map does not depend on the split; no workgroup waits for another.
// (wg: the workgroup within the grid's launch; part: the grid's overflow partials)
template <int NT>
__device__ __forceinline__ void bin_tiles_body(const Pick& pos_p, const BinJob& job, float target, double* part,
                                               int wg) {
    extern __shared__ unsigned long long acc_all[];   // G x planes x T x T
    __shared__ double red[NT / 64];
    __shared__ int last_of[8];         // per sub-group: this workgroup writes its split tile
    const BinJob J = job;              // by value: a reference to a kernel argument would go to scratch
    const bool split_on = J.helpers && J.fill;
    const bool helper = split_on && wg < J.helpers;
    // the marked tiles share the helpers evenly, R each (up to SPLIT_MAX - 1); helper h takes slice h % R + 1 of
    // marked tile h / R
    int R = 0;
    if (split_on) {
        const int nh = min(*J.heavy_n, SPLIT_LIST);
        R = nh ? min(SPLIT_MAX - 1, J.helpers / nh) : 0;
        if (helper && (R == 0 || wg / R >= nh)) return;   // an idle helper leaves before any barrier
    }
    const int S = helper ? NT : NT / J.G;   // lanes per tile
    const int sub = threadIdx.x / S, l = threadIdx.x - sub * S;
    const int t = helper ? J.heavy_list[wg / R] : (wg - (split_on ? J.helpers : 0)) * J.G + sub;
    const bool live = t < J.tiles;
    const int hp = live && split_on ? J.heavy_pos[t] : -1;   // its place among the marked tiles
    const int slice = helper ? wg % R + 1 : 0;
    const int T = J.T, TT = T * T;
    const int X0 = (t / J.tiles_y) * T, Y0 = (t % J.tiles_y) * T;
    const int X1 = min(X0 + T, J.nbx) - 1, Y1 = min(Y0 + T, J.nby) - 1;
    unsigned long long* acc = acc_all + sub * J.planes * TT;
    // issued before anything waits on memory: the grid's big-node count, the bucket bounds of both planes (fill form:
    // with their counts and the tile's spill flag) and this lane's share of the fixed map (the first TILE_PF bins of
    // its stride), so that a tile pays two dependent memory latencies, not four
    int lo0 = 0, n0 = 0, lo1 = 0, n1 = 0, c0 = 0, c1 = 0, spilled = 0, nbig = 0;
    if (live) {
        if (slice == 0) nbig = J.fill ? *J.big_count : *J.big_n;   // big nodes: the tile's own workgroup
        lo0 = J.start[t];
        n0 = J.start[t + 1];
        if (J.planes == 2) {
            lo1 = J.start[J.tiles + t];
            n1 = J.start[J.tiles + t + 1];
        }
        if (J.fill) {
            c0 = J.count[t];
            if (J.planes == 2) c1 = J.count[J.tiles + t];
        }
    }
    // workgroups that sum tile t: a listed tile is split into slices of split_slice entries or more, as many as its
    // share of helpers allows
    const int k = hp >= 0 && hp < SPLIT_LIST && J.split_slice > 0
                      ? 1 + min(R, max(0, (c0 + c1) / J.split_slice - 1)) : 1;
    if (helper && slice >= k) return;              // more helpers than slices: uniform over the workgroup
    if (live && J.fill && k == 1) spilled = J.spill_tile[t];   // a split tile's spilled sums stay until its end
    float fix[TILE_PF];
#pragma unroll
    for (int i = 0; i < TILE_PF; i++) {
        const int e = l + i * S, bx = X0 + e / T, by = Y0 + e % T;
        fix[i] = live && e < TT && bx <= X1 && by <= Y1 ? J.fixed[(long long)bx * J.nby + by] : 0.f;
    }
    if (live)
        for (int e = l; e < J.planes * TT; e += S) acc[e] = 0ull;
    n0 -= lo0;
    n1 -= lo1;
    if (J.fill) {                      // the entries in the rooms (the rest were spilled)
        n0 = min(n0, c0);
        n1 = min(n1, c1);
    }
    __syncthreads();
    // the entries of both planes as one range (plane 0 first), two per lane per round, both loaded before either is
    // used
    const int n_all = live ? n0 + n1 : 0;
    // this workgroup's slice of the range (all of it for an unsplit tile)
    const int q0 = (int)((long long)n_all * slice / k), n = (int)((long long)n_all * (slice + 1) / k);
    // A bucket lists its nodes about in curve order, so consecutive entries hit the same bins. The lanes of one wave
    // take entries S / 64 apart (lane i of wave w: entry i S / 64 + w), which spreads a wave's atomics over the tile.
    const int lp = (l & 63) * (S >> 6) + (l >> 6);
    for (int q = q0 + lp; q < n; q += 2 * S) {
        float4 g[2];
        float r[2];
        int pl[2];
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int x = q + k * S;
            pl[k] = x >= n0;
            const int e = x < n0 ? lo0 + x : lo1 + x - n0;
            if (x < n) {
                g[k] = J.geo[e];
                r[k] = J.gr[e];
            }
        }
#pragma unroll
        for (int k = 0; k < 2; k++)
            if (q + k * S < n)
                add_clipped(acc + pl[k] * TT, T, X0, Y0, X1, Y1, make_span(g[k].x, g[k].z, J.inv_bx, J.nbx),
                            make_span(g[k].y, g[k].w, J.inv_by, J.nby), r[k]);
    }
    // big nodes: one at a time, their bins in this tile spread over the tile's lanes
    if (nbig) {
        const float2* pos = pos_p.get(true);
        for (int k = 0; k < nbig; k++) {
            const int i = J.big[k];
            const float2 p = pos[i], sz = J.size[i];
            const Span sx = make_span(p.x, sz.x, J.inv_bx, J.nbx), sy = make_span(p.y, sz.y, J.inv_by, J.nby);
            const int x0 = max(sx.b0, X0), x1 = min(sx.b1, X1), y0 = max(sy.b0, Y0), y1 = min(sy.b1, Y1);
            if (x0 > x1 || y0 > y1) continue;                   // uniform over the tile's lanes
            const float r = J.ratio[i];
            unsigned long long* a = acc + (i >= J.num_mov ? TT : 0);
            const int h = y1 - y0 + 1, n = (x1 - x0 + 1) * h;
            for (int e = l; e < n; e += S) {
                const int bx = x0 + e / h, by = y0 + e % h;
                const unsigned long long c = contribution(r * span_overlap(sx, bx), sy, by);
                if (c) atomicAdd(a + (bx - X0) * T + (by - Y0), c);
            }
        }
    }
    // fill form: the spilled entries' sums of this tile's bins, added and cleared for the next pass
    if (spilled) {
        for (int e = l; e < J.planes * TT; e += S) {
            const int pl = e >= TT, bx = X0 + (e - pl * TT) / T, by = Y0 + (e - pl * TT) % T;
            if (bx > X1 || by > Y1) continue;
            unsigned long long* sp = J.spill + ((long long)pl * J.nbx + bx) * J.nby + by;
            const unsigned long long v = *sp;
            if (v) {
                atomicAdd(acc + e, v);
                *sp = 0ull;
            }
        }
        if (l == 0) J.spill_tile[t] = 0;
    }
    __syncthreads();
    // a split tile: this workgroup's sums into the spill map, then the count of the tile's workgroups done; the last
    // one takes every sum back
    bool fin = live && k == 1;
    if (live && k > 1) {
        for (int e = l; e < J.planes * TT; e += S) {
            const int pl = e >= TT, bx = X0 + (e - pl * TT) / T, by = Y0 + (e - pl * TT) % T;
            const unsigned long long v = acc[e];
            if (v && bx <= X1 && by <= Y1) atomicAdd(J.spill + ((long long)pl * J.nbx + bx) * J.nby + by, v);
        }
        // the sums performed before the count: every access of the split protocol is a device-scope atomic (done
        // at the point of coherence), so waiting for this lane's vector memory operations to complete orders them;
        // an agent-scope fence would also write the L2 back (s_waitcnt: vmcnt 0, expcnt and lgkmcnt unchecked)
        __builtin_amdgcn_s_waitcnt(0x0F70);
    }
    __syncthreads();
    if (live && k > 1 && l == 0) last_of[sub] = atomicAdd(J.arrive + t, 1u) == (unsigned)(k - 1);
    __syncthreads();
    if (live && k > 1 && last_of[sub]) {   // the others' sums are in the map: read and cleared by atomics
        for (int e = l; e < J.planes * TT; e += S) {
            const int pl = e >= TT, bx = X0 + (e - pl * TT) / T, by = Y0 + (e - pl * TT) % T;
            acc[e] = bx <= X1 && by <= Y1 ? atomicExch(J.spill + ((long long)pl * J.nbx + bx) * J.nby + by, 0ull)
                                          : 0ull;
        }
        if (l == 0) {
            atomicExch(J.arrive + t, 0u);
            J.spill_tile[t] = 0;
            J.heavy_pos[t] = -1;
            atomicAdd(J.split_n, 1ull);
        }
        fin = true;
    }
    if (live && k == 1 && hp != -1 && l == 0) J.heavy_pos[t] = -1;   // marked, not split: unmarked for the next pass
    __syncthreads();
    double o = 0.0;
    auto post = [&](int e, float fixed) {
        const int bx = X0 + e / T, by = Y0 + e % T;
        if (bx > X1 || by > Y1) return;
        const long long bin = (long long)bx * J.nby + by;
        const float mov = (float)((double)acc[e] / kDensityScale) + fixed;
        const float r = J.planes == 2 ? mov + (float)((double)acc[TT + e] / kDensityScale) : mov;
        if (J.rho) J.rho[bin] = r;
        if (J.count_over) o += (double)fmaxf(mov - target, 0.f);
        if (J.rho_mirror) {                       // the reference's FFT input: the bin and its three mirror images
            const long long w = 2LL * J.nby, mx = 2LL * J.nbx - 1 - bx, my = w - 1 - by;
            J.rho_mirror[bx * w + by] = r;
            J.rho_mirror[bx * w + my] = r;
            J.rho_mirror[mx * w + by] = r;
            J.rho_mirror[mx * w + my] = r;
        }
        if (J.over_map) J.over_map[bin] = fmaxf(mov - target, 0.f);
    };
#pragma unroll
    for (int i = 0; i < TILE_PF; i++) {
        const int e = l + i * S;
        if (fin && e < TT) post(e, fix[i]);
    }
    for (int e = l + TILE_PF * S; fin && e < TT; e += S) {
        const int bx = X0 + e / T, by = Y0 + e % T;
        post(e, bx <= X1 && by <= Y1 ? J.fixed[(long long)bx * J.nby + by] : 0.f);
    }
    // the tile's overflow: wave sums, then the tile's waves in order (G = 1: the order of block_sum)
    o = wave_sum(o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = o;
    __syncthreads();
    if (fin && l == 0) {
        double s = 0;
        for (int w = 0; w < S / 64; w++) s += red[sub * (S / 64) + w];
        part[t] = s;
    }
}
// the overflow partials: the force grid's tiles, then the overflow grid's
__global__ __launch_bounds__(TB_LEVEL, 8) void eng_bin_tiles(Pick pos_p, BinJob f, float target, double* part) {
    bin_tiles_body<TB_LEVEL>(pos_p, f, target, part, blockIdx.x);
}
__global__ __launch_bounds__(TB_LEVEL, 8) void eng_over_tiles(Pick pos_p, BinJob j, float target, double* part,
                                                              int off) {
    bin_tiles_body<TB_LEVEL>(pos_p, j, target, part + off, blockIdx.x);
}

// ------------------------------------------------------------------------------ gather (+ cell WA sum, |g' - g|)
struct GatherArgs {
    Pick pos;                 // evaluation point (the other set)
    const float2* g_prev;     // null: no |g' - g|
    const float2* g_prev_b;
    float2* g_out_a;
    float2* g_out_b;
    const int* flags;         // null: g_prev / g_out_a used as given
    const float2* size;
    const float* ratio;
    const float* fx;
    const float* fy;
    const float2* f2;         // non-null: the field as (x, y) pairs (fx / fy unused)
    int ld;
    const float2* pin_grad;
    const float2* cell_wl;    // the cells' WA gradients summed by eng_cell_wl (null: the gather sums the pins)
    const int* cell_start;
    const int* cell_pin;
    const float* pins;
    const float* area;
    const float* kp;          // weight kp[1], precond kp[2]
    int num_nodes, num_mov;
    float inv_bx, inv_by;
    int nbx, nby;
    int use_density;          // 0: grad = wl / max(pins, 1) (the weight-0 objective)
    const float* weight_p;    // drop-in gather: the weight and the preconditioner's weight as separate scalars
    const float* precond_p;
};

// a cell's WA gradient: the sum of its pins' gradients, which the WA kernel wrote in cell order (contiguous)
// (four pins' loads issued before the first of them is added, then added in pin order: the sum of the one-at-a-time
// loop, bit for bit, with a quarter of its dependent memory waits)
__device__ __forceinline__ float2 cell_pin_sum(const GatherArgs& a, int i) {
    float wx = 0, wy = 0;
    int k = a.cell_start[i];
    const int e = a.cell_start[i + 1];
    for (; k + 4 <= e; k += 4) {
        float2 q[4];
#pragma unroll
        for (int j = 0; j < 4; j++) q[j] = a.pin_grad[k + j];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            wx += q[j].x;
            wy += q[j].y;
        }
    }
    if (k < e) {                                    // the last 1 .. 3 pins, their loads also issued together
        float2 q[3];
#pragma unroll
        for (int j = 0; j < 3; j++) q[j] = k + j < e ? a.pin_grad[k + j] : make_float2(0.f, 0.f);
#pragma unroll
        for (int j = 0; j < 3; j++)
            if (k + j < e) {
                wx += q[j].x;
                wy += q[j].y;
            }
    }
    return make_float2(wx, wy);
}

// The cells' WA gradients and part_wl[w] (their squared norms summed over the wavefront of nodes w = i / 64, as the
// gather would sum them): the part of the gather that needs only WA, run beside the density stage, over the nodes.
__global__ __launch_bounds__(EB) void eng_cell_wl(GatherArgs a, float2* cell_wl, double* part_wl) {
    const int i = xcd_remap(blockIdx.x, gridDim.x) * EB + threadIdx.x;
    double t2 = 0.0;
    if (i < a.num_mov) {
        const float2 w = cell_pin_sum(a, i);
        cell_wl[i] = w;
        t2 = (double)w.x * w.x + (double)w.y * w.y;
    }
    t2 = wave_sum(t2);
    if ((threadIdx.x & 63) == 0 && i < a.num_nodes) part_wl[i >> 6] = t2;
}

// part[w]: |g' - g|^2 of the wavefront of nodes w = i / 64; part_wl[w], part_dn[w]: squared norms of the wirelength
// part and of the weighted density part of the (unpreconditioned) gradient, the learned controller's gradient-ratio
// feature (part_wl is eng_cell_wl's when a.cell_wl is set). Partials per wavefront of nodes: the same sums for any
// workgroup size (the autotuner's choice).
// DROPIN: the drop-in density_grad -- cell_wl is the caller's wirelength gradient of the first num_mov nodes (no
// extra row), the weights come from weight_p / precond_p, and no partials are written.
template <bool DROPIN = false>
__device__ __forceinline__ void gather_body(const GatherArgs& a, double* part, double* part_wl, double* part_dn, int wg,
                                            int nwg) {
    const int i = xcd_remap(wg, nwg) * (int)blockDim.x + threadIdx.x;
    double t2[3] = {0.0, 0.0, 0.0};
    if (i < a.num_nodes) {
        int cur = a.flags ? a.flags[0] : 0;
        const float2* pos = a.pos.get(true);
        // issued with the node's position, before the field's loads wait: the rest of what the node's gradient needs
        // (behind the uniform conditions below the compiler would not issue them early)
        const float pins = a.pins[i];
        const float area = a.use_density ? a.area[i] : 0.f;
        float2 w_cell = make_float2(0.f, 0.f);
        if (DROPIN) {
            if (a.cell_wl && i < a.num_mov) w_cell = a.cell_wl[i];
        } else if (a.cell_wl) {
            w_cell = a.cell_wl[min(i, a.num_mov)];   // num_mov + 1 entries
        }
        float2 gp = make_float2(0.f, 0.f);
        if (a.g_prev) gp = ((a.flags && cur == 1) ? a.g_prev_b : a.g_prev)[i];
        float gx = 0, gy = 0;
        if (a.use_density) {
            const float weight = DROPIN ? *a.weight_p : a.kp[1];
            const float2 p = pos[i], sz = a.size[i];
            const float r = a.ratio[i];
            Span sx = make_span(p.x, sz.x, a.inv_bx, a.nbx), sy = make_span(p.y, sz.y, a.inv_by, a.nby);
            if (sx.b1 - sx.b0 < 3 && sy.b1 - sy.b0 < 3) {
                // at most 3 x 3 bins: a fixed window, each axis's weights computed once, the field loads issued
                // together; the terms in the loop's order (outside bins have weight 0 and add nothing)
                float wx[3], wy[3], ex[3][3], ey[3][3];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    wx[k] = sx.b0 + k <= sx.b1 ? r * span_overlap(sx, sx.b0 + k) : 0.f;
                    wy[k] = sy.b0 + k <= sy.b1 ? span_overlap(sy, sy.b0 + k) : 0.f;
                }
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        const bool in = sx.b0 + i <= sx.b1 && sy.b0 + j <= sy.b1;
                        const long long o = (long long)(sx.b0 + i) * a.ld + sy.b0 + j;
                        const float2 f = !in ? make_float2(0.f, 0.f)
                                         : a.f2 ? a.f2[o] : make_float2(a.fx[o], a.fy[o]);
                        ex[i][j] = f.x;
                        ey[i][j] = f.y;
                    }
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        const float w = wx[i] * wy[j];
                        gx -= w * ex[i][j];
                        gy -= w * ey[i][j];
                    }
            } else {
                for (int bx = sx.b0; bx <= sx.b1; bx++) {
                    const float wx = r * span_overlap(sx, bx);
                    const long long row = (long long)bx * a.ld;
                    for (int by = sy.b0; by <= sy.b1; by++) {
                        const float w = wx * span_overlap(sy, by);
                        const float2 f = a.f2 ? a.f2[row + by] : make_float2(a.fx[row + by], a.fy[row + by]);
                        gx -= w * f.x;
                        gy -= w * f.y;
                    }
                }
            }
            gx *= weight;
            gy *= weight;
            t2[2] = (double)gx * gx + (double)gy * gy;
        }
        if (i < a.num_mov) {
            const float2 w = a.cell_wl ? w_cell : cell_pin_sum(a, i);
            gx += w.x;
            gy += w.y;
            t2[1] = (double)w.x * w.x + (double)w.y * w.y;
        }
        float scale;
        if (a.use_density) {
            const float precond = DROPIN ? *a.precond_p : a.kp[2];
            scale = precond < 0 ? 1.f : 1.f / fmaxf(pins + precond * area, 1.f);
        } else {
            scale = 1.f / fmaxf(pins, 1.f);
        }
        const float2 gn = make_float2(gx * scale, gy * scale);
        float2* out = (a.flags && cur == 0) ? a.g_out_b : a.g_out_a;
        out[i] = gn;
        if (a.g_prev) {
            const double dx = (double)gn.x - gp.x, dy = (double)gn.y - gp.y;
            t2[0] = dx * dx + dy * dy;
        }
    }
    if (DROPIN) return;
#pragma unroll
    for (int k = 0; k < 3; k++) t2[k] = wave_sum(t2[k]);
    if ((threadIdx.x & 63) == 0 && i < a.num_nodes) {
        part[i >> 6] = t2[0];
        if (!a.cell_wl) part_wl[i >> 6] = t2[1];
        part_dn[i >> 6] = t2[2];
    }
}
__global__ __launch_bounds__(512) void eng_gather(GatherArgs a, double* part, double* part_wl, double* part_dn) {
    gather_body(a, part, part_wl, part_dn, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------------------------ control (+ reductions)
struct Segs {
    const double* moved;
    int n_moved;
    const double* wa;        // pairs (WA, HPWL)
    int n_wa;
    const double* over;
    int n_over;
    const double* turned;
    int n_turned;
    const double* wl2;       // squared norms of the gradient's wirelength / density parts
    const double* dn2;
    int n_norm;
};

// First stage of the reduction of the partials: workgroup b sums the b-th fixed slice of every segment
// (moved^2, turned^2, WA, HPWL, overflow, wl^2, density^2) -> out[b][RED_VALS]. The slices and the order inside
// them do not depend on the launch, so the sums are deterministic.
constexpr int RED_BLOCKS = 64, RED_VALS = 7;
constexpr int RED_LANES = EB;   // lanes per workgroup of the first stage

__device__ __forceinline__ void slice(int n, int wg, int& lo, int& hi) {
    const int per = (n + RED_BLOCKS - 1) / RED_BLOCKS;
    lo = min(n, wg * per);
    hi = min(n, lo + per);
}

// (each lane's first RED_AHEAD elements of every segment are loaded before any is added -- one memory wait, not one
// per element and segment -- and any further ones follow in the loops; each lane adds its elements in slice order)
constexpr int RED_AHEAD = 4;
__device__ __forceinline__ void reduce_body(const Segs& s, double* out, int wg) {
    __shared__ double red[16 * RED_VALS];
    const int t = threadIdx.x;
    int lo[5], hi[5];
    slice(s.n_moved, wg, lo[0], hi[0]);
    slice(s.n_turned, wg, lo[1], hi[1]);
    slice(s.n_wa, wg, lo[2], hi[2]);
    slice(s.n_over, wg, lo[3], hi[3]);
    slice(s.n_norm, wg, lo[4], hi[4]);
    double x[RED_AHEAD][RED_VALS];
#pragma unroll
    for (int j = 0; j < RED_AHEAD; j++) {
        bool in[5];
#pragma unroll
        for (int q = 0; q < 5; q++) in[q] = lo[q] + t + j * RED_LANES < hi[q];
        const int o = t + j * RED_LANES;
        x[j][0] = in[0] ? s.moved[lo[0] + o] : 0.0;
        x[j][1] = in[1] ? s.turned[lo[1] + o] : 0.0;
        x[j][2] = in[2] ? s.wa[2 * (lo[2] + o)] : 0.0;
        x[j][3] = in[2] ? s.wa[2 * (lo[2] + o) + 1] : 0.0;
        x[j][4] = in[3] ? s.over[lo[3] + o] : 0.0;
        x[j][5] = in[4] ? s.wl2[lo[4] + o] : 0.0;
        x[j][6] = in[4] ? s.dn2[lo[4] + o] : 0.0;
    }
    double v[RED_VALS];
#pragma unroll
    for (int k = 0; k < RED_VALS; k++) {
        v[k] = x[0][k];
#pragma unroll
        for (int j = 1; j < RED_AHEAD; j++) v[k] += x[j][k];   // elements past the slice add 0.0: no change
    }
    const int more = t + RED_AHEAD * RED_LANES;
    for (int k = lo[0] + more; k < hi[0]; k += RED_LANES) v[0] += s.moved[k];
    for (int k = lo[1] + more; k < hi[1]; k += RED_LANES) v[1] += s.turned[k];
    for (int k = lo[2] + more; k < hi[2]; k += RED_LANES) {
        v[2] += s.wa[2 * k];
        v[3] += s.wa[2 * k + 1];
    }
    for (int k = lo[3] + more; k < hi[3]; k += RED_LANES) v[4] += s.over[k];
    for (int k = lo[4] + more; k < hi[4]; k += RED_LANES) {
        v[5] += s.wl2[k];
        v[6] += s.dn2[k];
    }
    block_sum<RED_VALS>(v, red);
    if (t == 0)
        for (int k = 0; k < RED_VALS; k++) out[wg * RED_VALS + k] = v[k];
}
__global__ __launch_bounds__(RED_LANES) void eng_reduce(Segs s, double* out) { reduce_body(s, out, blockIdx.x); }

// Second stage and the control. With a learned controller (ctl.params) the 64 lanes evaluate the MLP on the pass's
// features first; its clipped outputs (plus the exploration noise, when asked) scale the rule's density-weight
// growth and smoothing, until the overflow stalls for ctl.window passes (then the rule alone, for the level).
__device__ __forceinline__ void control_body(const double* red, double over_scale, double* S_g, float* kp_g,
                                             const double* c_g, int* flags, double* stats, const Ctl& ctl) {
    __shared__ double r[RED_VALS];
    __shared__ double rb[RED_BLOCKS * RED_VALS];
    __shared__ float x[MLP_IN];
    __shared__ float h[MLP_H];
    __shared__ int learn;
    // the first-stage sums, and the control state, into LDS by all lanes at once (one memory latency, not a chain of
    // them in lane 0); the sums are added by one lane in block order (a fixed order), the state written back after
    __shared__ double S_l[NUM_S], c_l[NUM_C];
    __shared__ float kp_l[5];
    for (int i = threadIdx.x; i < RED_BLOCKS * RED_VALS; i += blockDim.x) rb[i] = red[i];
    if (S_g) {
        for (int i = threadIdx.x; i < NUM_S; i += blockDim.x) S_l[i] = S_g[i];
        for (int i = threadIdx.x; i < NUM_C; i += blockDim.x) c_l[i] = c_g[i];
        if (threadIdx.x < 5) kp_l[threadIdx.x] = kp_g[threadIdx.x];
    }
    double* S = S_g ? S_l : nullptr;
    float* kp = kp_l;
    const double* c = c_l;
    __syncthreads();
    if (threadIdx.x < RED_VALS) {          // the second stage: sum k by lane k, in block order
        double a = 0.0;
        for (int b = 0; b < RED_BLOCKS; b++) a += rb[b * RED_VALS + threadIdx.x];
        r[threadIdx.x] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        learn = 0;
        if (S && ctl.params && S[STOP] == 0) {
            const double moved = sqrt(r[0]), turned = sqrt(r[1]), wa = r[2], hpwl = r[3], ovf = r[4] * over_scale;
            double* X = ctl.X;
            if (X[X_STARTED] == 0) {
                X[X_STARTED] = 1;
                X[X_PREV_OVF] = ovf;
                X[X_HPWL0] = hpwl;
                X[X_ALPHA0] = turned > 0 ? moved / turned : S[ALPHA];
                X[X_BEST_OVF] = ovf;
                X[X_LAST_IMPROVE] = S[ITER];
            }
            if (X[X_FALLBACK] == 0 && isfinite(moved) && isfinite(turned) && isfinite(hpwl) && isfinite(ovf)) {
                mlp_features(S, c, X, moved, turned, hpwl, ovf, wa, r[5], r[6], ctl.phase, x);
                learn = 1;
            }
        }
    }
    __syncthreads();
    float o0 = 0.f, o1 = 0.f;
    if (learn) mlp_forward_block(ctl.params, x, h, o0, o1);      // uniform: learn is one LDS word
    if (threadIdx.x == 0) {
        const double moved = sqrt(r[0]), turned = sqrt(r[1]), wa = r[2], hpwl = r[3], ovf = r[4] * over_scale;
        if (stats) {
            stats[0] = wa; stats[1] = hpwl; stats[2] = ovf; stats[3] = moved; stats[4] = turned;
            stats[5] = sqrt(r[5]); stats[6] = sqrt(r[6]);
        }
        if (S) {
            ScheduleAdjust adj;
            if (learn) {
                double* X = ctl.X;
                const unsigned long long n = (unsigned long long)X[X_COUNT];
                if (ctl.sigma > 0) {
                    const unsigned long long d = n / (unsigned long long)ctl.hold;
                    o0 += ctl.sigma * normal_draw(ctl.seed, d, 0);
                    o1 += ctl.sigma * normal_draw(ctl.seed, d, 1);
                }
                o0 = fminf(MLP_CLIP, fmaxf(-MLP_CLIP, o0));
                o1 = fminf(MLP_CLIP, fmaxf(-MLP_CLIP, o1));
                adj.lam_factor = exp((double)o0);
                adj.gam_factor = exp((double)o1);
                if (ctl.log && n < (unsigned long long)ctl.log_cap) {
                    float* L = ctl.log + n * MLP_LOG;
                    for (int k = 0; k < MLP_IN; k++) L[k] = x[k];
                    L[12] = o0;
                    L[13] = o1;
                    L[14] = (float)ovf;
                    L[15] = (float)hpwl;
                }
                X[X_COUNT] = (double)(n + 1);
            }
            control_step(S, kp, c, moved, turned, hpwl, ovf, wa, adj);
            if (ctl.params && S[ACCEPT] != 0) {
                double* X = ctl.X;
                X[X_PREV_OVF] = ovf;
                // a stall counts, as in the rule's own stall test, once the density share is 0.95 or more
                const bool better = ovf < X[X_BEST_OVF] * 0.999;
                if (better) X[X_BEST_OVF] = ovf;
                if (better || S[WW] < 0.95) X[X_LAST_IMPROVE] = S[ITER];
                else if (S[ITER] - X[X_LAST_IMPROVE] >= ctl.window) X[X_FALLBACK] = 1;
            }
            if (S[ACCEPT] != 0) flags[0] = 1 - flags[0];
            flags[1] = S[IS_BEST] != 0 ? 1 : 0;
        }
    }
    if (S_g) {
        __syncthreads();
        for (int i = threadIdx.x; i < NUM_S; i += blockDim.x) S_g[i] = S_l[i];
        if (threadIdx.x < 5) kp_g[threadIdx.x] = kp_l[threadIdx.x];
    }
}
__global__ __launch_bounds__(64) void eng_control(const double* red, double over_scale, double* S, float* kp,
                                                  const double* c, int* flags, double* stats, Ctl ctl) {
    control_body(red, over_scale, S, kp, c, flags, stats, ctl);
}

// The wirelength-only start's control (src/nesterov_optimizer.py:34-53 and param_scheduler.py:38-48, as
// run_placement.py:41-76 drives them before the density weight is on): every step is taken unless a norm or the HPWL
// is not finite (diverged: the point stays); the BB step alpha = moved / turned (kept when turned is 0); the
// momentum a' = (1 + sqrt(4 a^2 + 1)) / 2 and the next pass's alpha and coef (a' - 1) / a'' in kp; the pass's (WA,
// HPWL) appended to hist; the phase ends when WA fell by at most tol x window x WA over the last `window` passes
// (done 1), or at the pass limit (done 3). A done state freezes the point for the rest of the run.
enum { SS_DONE, SS_COUNT, SS_TOL, SS_WINDOW, SS_MAX, START_STATE = 8 };
__global__ __launch_bounds__(64) void eng_start_control(const double* red, double* S, float* kp, int* flags,
                                                        double* st, double* hist) {
    __shared__ double r[RED_VALS];
    __shared__ double rb[RED_BLOCKS * RED_VALS];
    for (int i = threadIdx.x; i < RED_BLOCKS * RED_VALS; i += blockDim.x) rb[i] = red[i];
    __syncthreads();
    if (threadIdx.x < RED_VALS) {          // the second stage: sum k by lane k, in block order
        double a = 0.0;
        for (int b = 0; b < RED_BLOCKS; b++) a += rb[b * RED_VALS + threadIdx.x];
        r[threadIdx.x] = a;
    }
    __syncthreads();
    if (threadIdx.x != 0 || st[SS_DONE] != 0) return;
    const double moved = sqrt(r[0]), turned = sqrt(r[1]), wa = r[2], hpwl = r[3];
    if (!isfinite(moved) || !isfinite(turned) || !isfinite(hpwl)) {
        st[SS_DONE] = 2;
        return;
    }
    flags[0] = 1 - flags[0];                                  // the step is taken: the pass's point is current
    const double alpha = turned > 0 ? moved / turned : S[ALPHA];
    const double a = S[A], a1 = (1.0 + sqrt(4.0 * a * a + 1.0)) / 2.0, a2 = (1.0 + sqrt(4.0 * a1 * a1 + 1.0)) / 2.0;
    S[ALPHA] = alpha;
    S[A] = a1;
    kp[3] = (float)alpha;
    kp[4] = (float)((a1 - 1.0) / a2);
    const int n = (int)st[SS_COUNT], window = (int)st[SS_WINDOW];
    hist[2 * n] = wa;
    hist[2 * n + 1] = hpwl;
    st[SS_COUNT] = n + 1;
    S[EVALS] += 1;
    if (n + 1 > window && !(hist[2 * (n - window)] - wa > st[SS_TOL] * window * wa)) st[SS_DONE] = 1;
    else if (n + 1 >= (int)st[SS_MAX]) st[SS_DONE] = 3;
}

// level set-up: pin -> its position in the cell CSR, so the WA kernel writes pin gradients in cell order
__global__ __launch_bounds__(EB) void eng_pin_slot_clear(int* slot, int num_pins) {
    const int p = blockIdx.x * EB + threadIdx.x;
    if (p < num_pins) slot[p] = -1;
}
__global__ __launch_bounds__(EB) void eng_pin_slot_fill(int* slot, const int* cell_start, const int* cell_pin,
                                                        int num_mov, int num_pins) {
    const int k = blockIdx.x * EB + threadIdx.x;
    if (k < cell_start[num_mov]) {
        const int p = cell_pin[k];
        if (p >= 0 && p < num_pins) slot[p] = k;
    }
}

// end of a run: the pending best copy, and the current point back into set 0
__global__ __launch_bounds__(EB) void eng_finalize_copy(Bufs B, int n) {
    const int i = blockIdx.x * EB + threadIdx.x;
    if (i >= n) return;
    const int cur = B.flags[0];
    if (B.flags[1]) B.best[i] = B.v[cur][i];
    if (cur == 1) {
        B.u[0][i] = B.u[1][i];
        B.v[0][i] = B.v[1][i];
        B.g[0][i] = B.g[1][i];
    }
}

__global__ void eng_finalize_flags(int* flags) {
    if (threadIdx.x == 0) {
        flags[0] = 0;
        flags[1] = 0;
    }
}

// ------------------------------------------------------------------------------ the level
// Tile side of a grid: the power of two from 8 to 64 at or below the side that gives BIN_TILE_NODES nodes per tile
// on average (2048 x 2048 bins and 2.9M nodes: 32; 512 x 512 bins and 2M cells: 8), raised while more than
// max(64, nodes / 1024) nodes could span more than T + 1 bins along an axis (those take the big-node path, which
// every tile's workgroup scans) or while the buckets do not fit the count / fill kernels' histogram. 0: no side fits.
constexpr int BIN_TILE_NODES = 1024;
// Nodes whose reach along an axis (bins a node can touch: floor(size / bin) + 2 at most) exceeds T + 1 for the tile
// sides T = 8, 16, 32, 64: integer counts, so the same in any order.
__global__ __launch_bounds__(256) void eng_reach_counts(const float2* size, int n, double bx, double by,
                                                        unsigned long long* over) {
    __shared__ unsigned cnt[4];
    if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const float2 s = size[i];
        const double reach = fmin(1e9, fmax(floor((double)s.x / bx), floor((double)s.y / by)) + 2.0);
        for (int k = 0; k < 4; k++)
            if (reach > (double)((8 << k) + 1)) atomicAdd(&cnt[k], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 4 && cnt[threadIdx.x]) atomicAdd(&over[threadIdx.x], (unsigned long long)cnt[threadIdx.x]);
}

// The tile side of a grid: the power of two 8 .. 64 near sqrt(BIN_TILE_NODES bins / node) (about BIN_TILE_NODES
// nodes per tile), enlarged while more than max(64, n / 1024) nodes reach past a tile's neighbours; sides whose
// bucket count exceeds the binning's histogram are skipped. over[k]: the nodes reaching past side 8 << k.
static int tile_side(long long n, const unsigned long long* over, int nbx, int nby, int planes) {
    const double want = std::sqrt((double)BIN_TILE_NODES * nbx * nby / (double)std::max(n, 1LL));
    int T = 8;
    while (T < 64 && 2 * T <= want) T *= 2;
    const long long allow = std::max(64LL, n / 1024);
    int fits = 0;
    for (; T <= 64; T *= 2) {
        if ((long long)planes * grid_of(nbx, T) * grid_of(nby, T) > BIN_MAX_BUCKETS) continue;
        fits = T;
        int k = 0;
        while ((8 << k) < T) k++;
        if ((long long)over[k] <= allow) return T;
    }
    return fits;   // the largest side that fits; its big nodes take the slower path
}

struct GraphEntry {            // a captured run of `passes` passes
    int passes = 0;
    int timing = -1;           // kernel class bracketed by events (-1: none)
    bool start = false;        // passes of the wirelength-only start
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> ev;   // timing: 2 per timed pass (before the class, after it)
    void clear() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        ev.clear();
        exec = nullptr;
        graph = nullptr;
        passes = 0;
        timing = -1;
        start = false;
    }
};
constexpr int GRAPH_CACHE = 4;

struct Level {
    ap_level_desc d{};
    Poisson poisson;
    bool poisson_borrowed = false;   // the plan is the caller's (ap_level_desc.poisson_plan): not destroyed here
    WaPlan wa_plan;
    float2 *u1 = nullptr, *v1 = nullptr, *g1 = nullptr, *pin_grad = nullptr;
    float2* cell_wl = nullptr;   // the cells' WA gradients (eng_cell_wl, two-stream passes)
    int* block_net = nullptr;  // WA workgroup -> first owned net
    int* pin_slot = nullptr;   // pin -> position in the cell CSR (-1: fixed pin)
    BinJob fj{}, jj{};         // density binning of the force grid and of the overflow grid
    bool rooms_stale = true;   // the bucket counts are not the last evaluation's at the current layout (set-up, a
                               // new layout): init_rooms counts before the next evaluation
    size_t hist_f = 0, hist_j = 0, acc_f = 0, acc_j = 0;   // dynamic LDS of the count / tile kernels per grid
    size_t hist_lds = 0, acc_lds = 0;                       // the larger of the two (the batch's grouped launches)
    std::vector<void*> bin_mem;
    float *rho = nullptr, *fx = nullptr, *fy = nullptr;
    float2* f2 = nullptr;      // the field as (x, y) pairs (FFT method; fx / fy otherwise)
    double* part = nullptr;
    double* stats = nullptr;
    double* red = nullptr;     // first-stage sums, RED_BLOCKS x RED_VALS
    double* ctl_x = nullptr;   // learned controller state, NUM_X
    Ctl ctl{};                 // the learned controller (params null: the plain rule)
    int* flags = nullptr;
    int nblk_n = 0, nblk_wa = 0;
    int wav_n = 0;             // wavefronts of 64 nodes: the gather's and the cells' WA sums' partials
    int gather_lanes = EB;     // lanes per workgroup of the gather (the autotuner's choice; results do not depend on it)
    double *p_moved = nullptr, *p_wa = nullptr, *p_over = nullptr, *p_turned = nullptr, *p_wl2 = nullptr,
           *p_dn2 = nullptr;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;                             // WA and the overflow grid, beside the force grid
    hipEvent_t ev_in = nullptr, ev_out = nullptr, t0 = nullptr, t1 = nullptr, ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_cells = nullptr;                          // the side stream's cell WA sums done (the gather's input)
    // the fork / join events of graph captures, kept apart from the ones direct launches record: an event is used
    // either inside captures or outside them, never both
    hipEvent_t cap_fork = nullptr, cap_join = nullptr, cap_cells = nullptr;
    bool capturing = false;
    GraphEntry graphs[GRAPH_CACHE];
    int graph_next = 0;
    // the wirelength-only start (ap_level_start_*): passes without the density stages whose control applies the
    // start rule; its state st f64[START_STATE] and the (WA, HPWL) of every start pass, hist f64[2 start_cap]
    bool start_mode = false;
    double* start_st = nullptr;
    double* start_hist = nullptr;
    int start_cap = 0;
    // event timing of one kernel class: events before and after it in every pass of a run (graph or direct)
    int timing_cls = -1;
    int time_every = 1;             // events in passes 0, every, 2 every, ... of a run only
    int timed_passes(int passes) const { return (passes + time_every - 1) / time_every; }
    hipEvent_t* rec_ev = nullptr;   // events of the run being captured or launched (null: none)
    bool rec_capture = false;       // true: added to the graph being captured as event-record nodes
    int rec_pass = 0;
    hipError_t rec_err = hipSuccess;
    std::vector<hipEvent_t> direct_ev;                 // events of direct runs
    const std::vector<hipEvent_t>* last_ev = nullptr;  // events of the last run (null: it had none)
    int last_ev_passes = 0;
    // profiling: one event before each kernel class and one at the end of the pass
    bool prof = false;
    hipEvent_t pev[AP_PROF_CLASSES + 1] = {};
    double prof_ms[AP_PROF_CLASSES] = {};
    void mark(int k, hipStream_t s) {   // profiled passes: the start of class k (and the end of the one before)
        if (prof) (void)hipEventRecord(pev[k], s);
    }
    // the timed class's events, on the stream the class runs on
    void tbegin(int k, hipStream_t s) { tevent(k, 0, s); }
    void tend(int k, hipStream_t s) { tevent(k, 1, s); }
    void tevent(int k, int end, hipStream_t s) {
        if (!rec_ev || k != timing_cls || rec_pass % time_every) return;
        hipEvent_t ev = rec_ev[2 * (rec_pass / time_every) + end];
        const hipError_t e = rec_capture ? record_in_capture(ev, s) : hipEventRecord(ev, s);
        if (e != hipSuccess && rec_err == hipSuccess) rec_err = e;
    }
    // an event-record node after the nodes captured so far, which the next captured work depends on: the event of a
    // stream-ordered hipEventRecord, inside the graph
    static hipError_t record_in_capture(hipEvent_t ev, hipStream_t s) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        unsigned long long id = 0;
        hipGraph_t g = nullptr;
        const hipGraphNode_t* deps = nullptr;
        size_t nd = 0;
        hipError_t e = hipStreamGetCaptureInfo_v2(s, &cs, &id, &g, &deps, &nd);
        if (e != hipSuccess) return e;
        if (cs != hipStreamCaptureStatusActive || !g) return hipErrorIllegalState;
        hipGraphNode_t node = nullptr;
        if ((e = hipGraphAddEventRecordNode(&node, g, deps, nd, ev)) != hipSuccess) return e;
        return hipStreamUpdateCaptureDependencies(s, &node, 1, hipStreamSetCaptureDependencies);
    }

    Bufs bufs() const {
        Bufs B;
        B.u[0] = (float2*)d.u; B.u[1] = u1;
        B.v[0] = (float2*)d.v; B.v[1] = v1;
        B.g[0] = (float2*)d.g; B.g[1] = g1;
        B.best = (float2*)d.best;
        B.flags = flags;
        return B;
    }

    template <typename T>
    int alloc(T** p, long long n) {
        AP_HIP(dev_malloc((void**)p, (size_t)(n > 0 ? n : 1) * sizeof(T)));
        bin_mem.push_back((void*)*p);
        return AP_OK;
    }
    int bin_job(BinJob& J, const float* size, const float* ratio, int num_nodes, int num_mov, double bx, double by,
                int nbx, int nby, int planes, const float* fixed, float* out, int count_over);
    bool side_ok(const BinJob& J, int T) const;
    void set_side(BinJob& J, int T, int G = 0);
    void update_lds();
    int set_lds() const;
    int init_rooms(const Pick& pos, hipStream_t s);
    int solve_field(hipStream_t s);
    int bin_grid(int which, const Pick& pos, hipStream_t s);
    int tiles_grid(int which, const Pick& pos, hipStream_t s);
    size_t state_bytes() const;
    int state_copy(char* keep, bool restore, hipStream_t s);
    int time_passes(int reps, hipStream_t s, float& ms);
    int autotune(int reps, hipStream_t s, int* sides, bool online = false, bool shapes = true);
    int create(const ap_level_desc& desc);
    void destroy();
    int objective(const Pick& pos, const float* kp_ptr, bool with_state, const float2* g_prev, const float2* g_prev_b,
                  float2* g_out_a, float2* g_out_b, hipStream_t s, double* S, float* kp_out, int use_density,
                  bool rooms_set = false);
    // the kernels' arguments (shared by the single-level launches and the batch's tables)
    WaArgs wa_args(const Pick& pos, const float* kp_ptr) const;
    GatherArgs gather_args(const Pick& pos, const float* kp_ptr, const float2* g_prev, const float2* g_prev_b,
                           float2* g_out_a, float2* g_out_b, int use_density) const;
    Segs segs(bool with_state, bool has_prev, int use_density) const;
    double over_scale() const;
    NestArgs nest_args() const;
    Pick pass_pos() const;
    Pick cur_pos() const;
    int pass(hipStream_t s);
    int finalize(hipStream_t s);
    int graph_for(int passes, int& slot);
};

int Level::bin_job(BinJob& J, const float* size, const float* ratio, int num_nodes, int num_mov, double bx, double by,
                   int nbx, int nby, int planes, const float* fixed, float* out, int count_over) {
    J = BinJob{};
    J.size = (const float2*)size;
    J.ratio = ratio;
    J.num_nodes = num_nodes;
    J.num_mov = num_mov;
    J.inv_bx = 1.f / (float)bx;   // the same float32 reciprocal as the reference's host code: 1.f / (float)bx
    J.inv_by = 1.f / (float)by;
    J.nbx = nbx;
    J.nby = nby;
    J.planes = planes;
    unsigned long long over[4] = {0, 0, 0, 0};
    {
        unsigned long long* d_over = nullptr;
        AP_HIP(dev_malloc(&d_over, sizeof(over)));
        hipError_t e = hipMemset(d_over, 0, sizeof(over));
        if (e == hipSuccess && num_nodes > 0) {
            hipLaunchKernelGGL(eng_reach_counts, dim3(grid_of(num_nodes, 256)), dim3(256), 0, nullptr, J.size,
                               num_nodes, bx, by, d_over);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpy(over, d_over, sizeof(over), hipMemcpyDeviceToHost);
        (void)dev_free(d_over);
        AP_HIP(e);
    }
    J.T = J.T0 = tile_side(num_nodes, over, nbx, nby, planes);
    if (!J.T) {
        set_error("level: density grid %d x %d is beyond the tile binning's %d buckets", nbx, nby, BIN_MAX_BUCKETS);
        return AP_ERR_UNSUPPORTED;
    }
    AP_REQUIRE(num_nodes <= (1 << 27), "level: more than 2^27 nodes on a density grid");
    J.blocks = (int)grid_of(num_nodes, BIN_NODES);
    J.fixed = fixed;
    J.rho = out;
    J.count_over = count_over;
    set_side(J, J.T);
    // Entry arrays for the fill form's rooms: a node has at most 4 entries, so the rooms of the default percentage
    // (150 % of a count + 64) fit in 6 entries per node + 64 per bucket of the side with the most buckets (8).
    // Bucket arrays for the largest bucket count any side may use, so that a new side needs no allocation.
    J.fill = 1;
    J.room_pct = 150;
    J.room_add = 64;
    const long long nbuck_max =
        std::min<long long>(BIN_MAX_BUCKETS, (long long)planes * grid_of(nbx, 8) * grid_of(nby, 8));
    J.room = (int)(6LL * num_nodes + 64LL * nbuck_max);
    int st;
    if ((st = alloc(&J.count, BIN_MAX_BUCKETS)) || (st = alloc(&J.start, BIN_MAX_BUCKETS + 1)) ||
        (st = alloc(&J.off, num_nodes)) || (st = alloc(&J.geo, J.room)) || (st = alloc(&J.gr, J.room)) ||
        (st = alloc(&J.big_count, 1)) || (st = alloc(&J.big_n, 1)) || (st = alloc(&J.big, num_nodes)) ||
        (st = alloc(&J.spill, (long long)planes * nbx * nby)) || (st = alloc(&J.spill_tile, BIN_MAX_BUCKETS)) ||
        (st = alloc(&J.spill_n, 1)) || (st = alloc(&J.heavy_n, 1)) || (st = alloc(&J.heavy_list, SPLIT_LIST)) ||
        (st = alloc(&J.heavy_pos, BIN_MAX_BUCKETS)) || (st = alloc(&J.arrive, BIN_MAX_BUCKETS)) ||
        (st = alloc(&J.split_n, 1)))
        return st;
    J.split_slice = SPLIT_SLICE;
    J.split_factor = SPLIT_FACTOR;
    J.helpers = SPLIT_HELPERS;
    split_threshold(J);
    AP_HIP(hipMemset(J.heavy_n, 0, sizeof(int)));
    AP_HIP(hipMemset(J.heavy_pos, 0xff, BIN_MAX_BUCKETS * sizeof(int)));    // -1: unmarked
    AP_HIP(hipMemset(J.arrive, 0, BIN_MAX_BUCKETS * sizeof(unsigned)));
    AP_HIP(hipMemset(J.split_n, 0, sizeof(unsigned long long)));
    AP_HIP(hipMemset(J.spill_n, 0, sizeof(unsigned long long)));
    AP_HIP(hipMemset(J.count, 0, BIN_MAX_BUCKETS * sizeof(int)));
    AP_HIP(hipMemset(J.start, 0, (BIN_MAX_BUCKETS + 1) * sizeof(int)));
    AP_HIP(hipMemset(J.big_count, 0, sizeof(int)));
    AP_HIP(hipMemset(J.big_n, 0, sizeof(int)));
    AP_HIP(hipMemset(J.spill, 0, (size_t)planes * nbx * nby * sizeof(unsigned long long)));
    AP_HIP(hipMemset(J.spill_tile, 0, BIN_MAX_BUCKETS * sizeof(int)));
    update_lds();
    return AP_OK;
}

// a tile side the binning can use for J: 8..64 and the buckets within the count / fill kernels' histogram
bool Level::side_ok(const BinJob& J, int T) const {
    return T >= 8 && T <= 64 && (long long)J.planes * grid_of(J.nbx, T) * grid_of(J.nby, T) <= BIN_MAX_BUCKETS;
}

// tiles per workgroup of the tile kernel, by default: one (a tile's bucket spread over all the workgroup's lanes; on
// the 2M-cell design this measured best or within 2 % of the best for every side of both grids)
static int default_group(int, int) { return 1; }

// a group the tile kernel can use for side T: 1, 2, 4 or 8 tiles per workgroup (each tile whole waves of the
// single-level kernel's lanes), within 64 KB of LDS
static bool group_ok(int planes, int T, int G) {
    return (G == 1 || G == 2 || G == 4 || G == 8) && TB_LEVEL / G >= 64 &&
           (size_t)G * planes * T * T * sizeof(unsigned long long) <= 65536;
}

void Level::set_side(BinJob& J, int T, int G) {
    rooms_stale = true;
    J.T = T;
    J.tiles_y = (int)grid_of(J.nby, T);
    J.tiles = (int)grid_of(J.nbx, T) * J.tiles_y;
    J.G = G ? G : default_group(J.planes, T);
    J.wgs = (J.tiles + J.G - 1) / J.G;
    split_threshold(J);
}

// the bucket count that marks a tile for splitting: a tile needs 2 x split_slice entries and split_factor x the mean
// per tile (nodes / tiles, about one entry per node); either of its planes' buckets reaching its share marks it
void split_threshold(BinJob& J) {
    if (!J.helpers || !J.split_slice || !J.tiles) {
        J.split_need = 0;
        return;
    }
    const long long need = std::max<long long>(2LL * J.split_slice, (long long)J.split_factor * J.num_nodes / J.tiles);
    J.split_need = (int)std::min<long long>(need / J.planes, 1LL << 30);
}

void Level::update_lds() {
    auto hist = [](const BinJob& J) { return J.count ? (size_t)J.planes * J.tiles * sizeof(int) : (size_t)0; };
    auto acc = [](const BinJob& J) {
        return J.count ? (size_t)J.G * J.planes * J.T * J.T * sizeof(unsigned long long) : (size_t)0;
    };
    hist_f = hist(fj);
    hist_j = hist(jj);
    acc_f = acc(fj);
    acc_j = acc(jj);
    hist_lds = std::max(hist_f, hist_j);
    acc_lds = std::max(acc_f, acc_j);
}

// the dynamic LDS limits of the single-level fill, count and tile kernels, for the current layout
int Level::set_lds() const {
    const void* fn[5] = {(const void*)eng_bin_fill, (const void*)eng_over_fill, (const void*)eng_bin_tiles,
                         (const void*)eng_over_tiles, (const void*)eng_rooms_count};
    const size_t lds[5] = {hist_f, hist_j, acc_f, acc_j, hist_lds};
    for (int k = 0; k < 5; k++)
        if (lds[k]) AP_HIP(hipFuncSetAttribute(fn[k], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds[k]));
    return AP_OK;
}

// The fill form's counts at pos for the current layout (set-up, or a new layout), outside any graph: the rooms of
// the next evaluation are set from them. Spill map and flags are zero between evaluations whatever the layout.
int Level::init_rooms(const Pick& pos, hipStream_t s) {
    if (!rooms_stale) return AP_OK;
    int st;
    if ((st = set_lds())) return st;
    for (int w = 0; w < 2; w++) {
        const BinJob& J = w ? jj : fj;
        if (!J.count) continue;
        AP_HIP(hipMemsetAsync(J.count, 0, BIN_MAX_BUCKETS * sizeof(int), s));
        AP_HIP(hipMemsetAsync(J.big_count, 0, sizeof(int), s));
        hipLaunchKernelGGL(eng_rooms_count, dim3(J.blocks), dim3(EB), w ? hist_j : hist_f, s, pos, fj, jj, w);
        AP_LAUNCHED();
    }
    rooms_stale = false;
    return AP_OK;
}

// the Poisson solve of the pass: the field as (x, y) pairs with the FFT method, as two planes with the GEMM one
int Level::solve_field(hipStream_t s) {
    return f2 ? poisson.solve_pairs(rho, d.nby, f2, d.nby, s) : poisson.solve(rho, d.nby, fx, fy, d.nby, s);
}

// the fill of one grid (0: force, 1: overflow); its rooms are set
int Level::bin_grid(int which, const Pick& pos, hipStream_t s) {
    if (!which)
        hipLaunchKernelGGL(eng_bin_fill, dim3(fj.blocks), dim3(EB), hist_f, s, pos, fj, jj);
    else
        hipLaunchKernelGGL(eng_over_fill, dim3(jj.blocks), dim3(EB), hist_j, s, pos, fj, jj);
    AP_LAUNCHED();
    return AP_OK;
}

// the tile sums of one grid: density and overflow partials
int Level::tiles_grid(int which, const Pick& pos, hipStream_t s) {
    if (!which)
        hipLaunchKernelGGL(eng_bin_tiles, dim3(fj.wgs + fj.helpers), dim3(TB_LEVEL), acc_f, s, pos, fj, d.target,
                           p_over);
    else
        hipLaunchKernelGGL(eng_over_tiles, dim3(jj.wgs + jj.helpers), dim3(TB_LEVEL), acc_j, s, pos, jj, d.target,
                           p_over, fj.tiles);
    AP_LAUNCHED();
    return AP_OK;
}

// The state whole passes change: both point sets (u, v, g), the best copy, the control state S / kp, the set flags
// and the learned controller's state. Everything else a pass writes it recomputes before reading (gradients, maps,
// field, partials), or changes only the time (the bucket rooms).
// (state_copy's layout: 7 point-set arrays, then 4 small arrays of at most 256 bytes, each start 256-byte aligned)
size_t Level::state_bytes() const {
    const size_t n2 = ((size_t)d.num_nodes * sizeof(float2) + 255) & ~(size_t)255;
    return 7 * n2 + 4 * 256;
}

int Level::state_copy(char* keep, bool restore, hipStream_t s) {
    const size_t n2 = (size_t)d.num_nodes * sizeof(float2);
    struct Part {
        void* p;
        size_t n;
    } parts[] = {{d.u, n2}, {d.v, n2}, {d.g, n2}, {u1, n2}, {v1, n2}, {g1, n2}, {d.best, n2},
                 {d.S, NUM_S * sizeof(double)}, {d.kp, 5 * sizeof(float)}, {flags, 4 * sizeof(int)},
                 {ctl_x, NUM_X * sizeof(double)}};
    static_assert(NUM_S * sizeof(double) <= 256 && NUM_X * sizeof(double) <= 256, "state_bytes: small arrays");
    size_t off = 0;
    for (const Part& q : parts) {
        void* dst = restore ? q.p : (void*)(keep + off);
        const void* src = restore ? (const void*)(keep + off) : q.p;
        AP_HIP(hipMemcpyAsync(dst, src, q.n, hipMemcpyDeviceToDevice, s));
        off += (q.n + 255) & ~(size_t)255;
    }
    return AP_OK;
}

// `reps` whole passes as direct runs launch them (the two streams), from the current state: ms per pass. The caller
// saves the state before and restores it after (state_copy).
int Level::time_passes(int reps, hipStream_t s, float& ms) {
    int st;
    if ((st = set_lds()) || (st = init_rooms(cur_pos(), s))) return st;
    AP_HIP(hipEventRecord(t0, s));
    for (int r = 0; r < reps; r++)
        if ((st = pass(s))) return st;
    AP_HIP(hipEventRecord(t1, s));
    AP_HIP(hipEventSynchronize(t1));
    AP_HIP(hipEventElapsedTime(&ms, t0, t1));
    ms /= reps;
    return AP_OK;
}

// Autotuner of the tile layout: for each grid (force, then overflow), the admissible sides from half to twice the
// set-up side (8 .. 64) with 1 or 2 tiles per workgroup are timed on whole passes and the fastest kept. The map does
// not depend on either (integer sums), so the choice changes the time only, never the results. Offline, every timing
// starts from the current state (saved once, restored before every timing and at the end). Online, the timings are
// the level's own next passes: they advance the level as any pass does (the control counts them), so the tuning
// costs no extra passes -- the passes end bit for bit where untuned ones end, as the layout changes no result.
int Level::autotune(int reps, hipStream_t s, int* sides, bool online, bool shapes) {
    AP_HIP(hipDeviceSynchronize());
    int st;
    char* keep = nullptr;
    if (!online) {
        AP_HIP(dev_malloc(&keep, state_bytes()));
        if ((st = state_copy(keep, false, s))) {
            (void)dev_free(keep);
            return st;
        }
    }
    auto timed = [&](int n, float& ms) -> int {
        int e = keep ? state_copy(keep, true, s) : AP_OK;
        return e ? e : time_passes(n, s, ms);
    };
    // coordinate descent: the force grid, the overflow grid, then the force grid again with the overflow grid's
    // choice in place (the two share the GPU in a pass, so each one's best depends on the other's)
    BinJob* order[3] = {&fj, &jj, &fj};
    for (int sweep = 0; sweep < (jj.count ? 3 : 1); sweep++) {
        BinJob* J = order[sweep];
        int best = J->T, best_g = J->G;
        float best_ms = 1e30f;
        for (int T = std::max(8, J->T0 / 2); T <= std::min(64, 2 * J->T0); T *= 2) {
            if (!side_ok(*J, T)) continue;
            for (int G = 1; G <= 2; G *= 2) {
                if (!group_ok(J->planes, T, G)) break;
                set_side(*J, T, G);
                update_lds();
                float warm = 0, ms = 0, again = 0;
                if ((st = timed(1, warm))) break;                 // warm-up at this layout
                if (warm > 1.5f * best_ms) continue;              // clearly slower: not timed further
                if ((st = timed(reps, ms))) break;
                if (ms < 1.1f * best_ms) {                        // a contender: the better of two timings
                    if ((st = timed(reps, again))) break;
                    ms = std::min(ms, again);
                }
                if (ms < best_ms) {
                    best_ms = ms;
                    best = T;
                    best_g = G;
                }
            }
            if (st) break;
        }
        if (st) break;
        set_side(*J, best, best_g);
        update_lds();
    }
    // then the launch shapes of the per-net and per-node kernels (shapes: single-level passes only; the batch's
    // grouped launches keep theirs): WA's plan wavefronts per launched wavefront and lanes per workgroup, then the
    // gather's lanes per workgroup. Their partials are per wavefront, so the results do not depend on them either.
    auto sweep = [&](const std::vector<std::pair<int, int>>& cand, std::function<void(int, int)> apply,
                     std::pair<int, int> cur) -> int {
        std::pair<int, int> best = cur;
        float best_ms = 1e30f;
        for (const auto& c : cand) {
            apply(c.first, c.second);
            float warm = 0, ms = 0, again = 0;
            int e;
            if ((e = timed(1, warm))) return e;
            if (warm > 1.5f * best_ms) continue;
            if ((e = timed(reps, ms))) return e;
            if (ms < 1.1f * best_ms) {
                if ((e = timed(reps, again))) return e;
                ms = std::min(ms, again);
            }
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
        apply(best.first, best.second);
        return AP_OK;
    };
    if (!st && shapes && wa_plan.num_waves) {
        std::vector<std::pair<int, int>> cand;
        for (int p : {1, 2, 4})
            for (int l : {128, 256, 512}) cand.push_back({p, l});
        st = sweep(cand, [&](int p, int l) { wa_plan.per = p; wa_plan.lanes = l; }, {wa_plan.per, wa_plan.lanes});
    }
    if (!st && shapes)
        st = sweep({{128, 0}, {256, 0}, {512, 0}}, [&](int l, int) { gather_lanes = l; }, {gather_lanes, 0});
    // offline: the state the timing passes changed, back; then the keep buffer goes
    if (!st && keep) st = state_copy(keep, true, s);
    (void)hipStreamSynchronize(s);
    if (keep) (void)dev_free(keep);
    if (st) return st;
    // the timing runs leave the fill form's counts behind: back to zero, the state the counting-sort form (the
    // batch's levels) starts from; the fill form counts again at the new layout anyway (rooms_stale)
    AP_HIP(hipDeviceSynchronize());
    for (BinJob* J : {&fj, &jj}) {
        if (!J->count) continue;
        AP_HIP(hipMemset(J->count, 0, BIN_MAX_BUCKETS * sizeof(int)));
        AP_HIP(hipMemset(J->big_count, 0, sizeof(int)));
    }
    if ((st = set_lds())) return st;
    for (GraphEntry& G : graphs) {                    // the kernels' arguments change: capture again
        if (last_ev == &G.ev) last_ev = nullptr;
        G.clear();
    }
    if (sides) {
        sides[0] = fj.T;
        sides[1] = fj.G;
        sides[2] = jj.count ? jj.T : 0;
        sides[3] = jj.count ? jj.G : 0;
    }
    AP_HIP(hipDeviceSynchronize());
    return AP_OK;
}

// With AP_TRACE_CREATE set in the environment, a level's set-up prints the time of each phase (device synchronised)
// to stderr: instrumentation, off by default.
struct CreateTrace {
    bool on = getenv("AP_TRACE_CREATE") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void operator()(const char* phase) {
        if (!on) return;
        (void)hipDeviceSynchronize();
        const auto now = std::chrono::steady_clock::now();
        fprintf(stderr, "level create: %-20s %.3f ms\n", phase,
                std::chrono::duration<double, std::milli>(now - t).count());
        t = now;
    }
};

int Level::create(const ap_level_desc& desc) {
    CreateTrace trace;
    d = desc;
    AP_REQUIRE(d.num_nodes > 0 && d.num_mov >= 0 && d.num_mov <= d.num_nodes && d.num_pins >= 0 && d.num_nets >= 0,
               "level: bad sizes");
    AP_REQUIRE(d.nbx > 0 && d.nby > 0 && d.bx > 0 && d.by > 0, "level: bad force grid");
    AP_REQUIRE(!d.has_judge || (d.njx > 0 && d.njy > 0 && d.jbx > 0 && d.jby > 0 && d.judge_size && d.judge_ratio &&
                                d.judge_fixed),
               "level: bad overflow grid");
    AP_REQUIRE(d.u && d.v && d.g && d.best && d.S && d.kp && d.c && d.size && d.ratio && d.pins && d.area && d.lo &&
                   d.hi && d.fixed,
               "level: null pointer");
    const Poisson* given = d.poisson_plan ? &poisson_of(d.poisson_plan) : nullptr;
    int st;
    if (given && given->N == d.nbx && given->M == d.nby && given->bx == d.bx && given->by == d.by &&
        (d.poisson_method == AP_POISSON_AUTO || d.poisson_method == given->method)) {
        poisson = *given;                          // the caller's plan: used, not destroyed
        poisson_borrowed = true;
    } else if ((st = poisson.create(d.nbx, d.nby, d.bx, d.by, d.poisson_method))) {
        return st;
    }
    trace("poisson plan");
    const long long N = d.num_nodes, nb = (long long)d.nbx * d.nby;
    AP_HIP(dev_malloc(&u1, N * sizeof(float2)));
    AP_HIP(dev_malloc(&v1, N * sizeof(float2)));
    AP_HIP(dev_malloc(&g1, N * sizeof(float2)));
    AP_HIP(dev_malloc(&pin_grad, (d.num_pins + 1) * sizeof(float2)));
    AP_HIP(dev_malloc(&cell_wl, ((size_t)d.num_mov + 1) * sizeof(float2)));
    AP_HIP(dev_malloc(&pin_slot, (d.num_pins + 1) * sizeof(int)));
    if (d.num_pins > 0) {
        hipLaunchKernelGGL(eng_pin_slot_clear, dim3(grid_of(d.num_pins, EB)), dim3(EB), 0, nullptr, pin_slot,
                           d.num_pins);
        AP_LAUNCHED();
        hipLaunchKernelGGL(eng_pin_slot_fill, dim3(grid_of(d.num_pins, EB)), dim3(EB), 0, nullptr, pin_slot,
                           d.cell_start, d.cell_pin, d.num_mov, d.num_pins);
        AP_LAUNCHED();
    }
    AP_HIP(dev_malloc(&rho, nb * sizeof(float)));
    if (poisson.method == AP_POISSON_FFT) {        // the field as (x, y) pairs: one load per bin in the gather
        AP_HIP(dev_malloc(&f2, nb * sizeof(float2)));
    } else {
        AP_HIP(dev_malloc(&fx, nb * sizeof(float)));
        AP_HIP(dev_malloc(&fy, nb * sizeof(float)));
    }
    nblk_n = grid_of(N, EB);
    wav_n = grid_of(N, 64);
    trace("buffers, pin slots");
    if ((st = wa_plan.create(d.net_start, d.num_nets, d.pin_cell, d.pin_off, pin_slot, d.num_pins))) return st;
    nblk_wa = wa_plan.partial_count();
    trace("wa plan");
    if ((st = bin_job(fj, d.size, d.ratio, d.num_nodes, d.num_mov, d.bx, d.by, d.nbx, d.nby, 2, d.fixed, rho,
                      d.has_judge ? 0 : 1)))
        return st;
    if (d.has_judge &&
        (st = bin_job(jj, d.judge_size, d.judge_ratio, d.num_mov, d.num_mov, d.jbx, d.jby, d.njx, d.njy, 1,
                      d.judge_fixed, nullptr, 1)))
        return st;
    trace("bin jobs");
    if ((st = set_lds())) return st;
    // one overflow partial per tile: room for the most tiles any side the autotuner may pick gives (a grid has at
    // most BIN_MAX_BUCKETS tiles)
    const long long n_over = 2LL * BIN_MAX_BUCKETS;
    const long long np = nblk_n + 2LL * nblk_wa + n_over + 3LL * wav_n;
    AP_HIP(dev_malloc(&part, np * sizeof(double)));
    AP_HIP(hipMemset(part, 0, np * sizeof(double)));
    p_moved = part;
    p_wa = p_moved + nblk_n;
    p_over = p_wa + 2LL * nblk_wa;
    p_turned = p_over + n_over;
    p_wl2 = p_turned + wav_n;
    p_dn2 = p_wl2 + wav_n;
    AP_HIP(dev_malloc(&stats, 8 * sizeof(double)));
    AP_HIP(dev_malloc(&red, RED_BLOCKS * RED_VALS * sizeof(double)));
    AP_HIP(dev_malloc(&ctl_x, NUM_X * sizeof(double)));
    AP_HIP(hipMemset(ctl_x, 0, NUM_X * sizeof(double)));
    ctl.X = ctl_x;
    AP_HIP(dev_malloc(&flags, 4 * sizeof(int)));
    AP_HIP(hipMemset(flags, 0, 4 * sizeof(int)));
    AP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    AP_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    AP_HIP(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
    AP_HIP(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
    AP_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    AP_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    AP_HIP(hipEventCreateWithFlags(&ev_cells, hipEventDisableTiming));
    AP_HIP(hipEventCreateWithFlags(&cap_fork, hipEventDisableTiming));
    AP_HIP(hipEventCreateWithFlags(&cap_join, hipEventDisableTiming));
    AP_HIP(hipEventCreateWithFlags(&cap_cells, hipEventDisableTiming));
    AP_HIP(hipEventCreate(&t0));
    AP_HIP(hipEventCreate(&t1));
    AP_HIP(hipDeviceSynchronize());
    trace("partials, streams");
    return AP_OK;
}

void Level::destroy() {
    (void)hipDeviceSynchronize();          // then every buffer below is idle
    for (GraphEntry& G : graphs) G.clear();
    for (hipEvent_t e : direct_ev) (void)hipEventDestroy(e);
    direct_ev.clear();
    last_ev = nullptr;
    for (hipEvent_t& ev : pev)
        if (ev) (void)hipEventDestroy(ev);
    void* ptrs[] = {u1, v1, g1, pin_grad, cell_wl, rho, fx, fy, f2, part, stats, flags, red, ctl_x, block_net,
                    pin_slot, start_st, start_hist};
    for (void* p : ptrs) dev_free_synced(p);
    for (void* p : bin_mem) dev_free_synced(p);
    bin_mem.clear();
    if (!poisson_borrowed) poisson.destroy();
    poisson = Poisson{};
    poisson_borrowed = false;
    wa_plan.destroy();
    if (stream) (void)hipStreamDestroy(stream);
    if (side) (void)hipStreamDestroy(side);
    hipEvent_t evs[] = {ev_in, ev_out, t0, t1, ev_fork, ev_join, ev_cells, cap_fork, cap_join, cap_cells};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
}

WaArgs Level::wa_args(const Pick& pos, const float* kp_ptr) const {
    // the evaluation point: set 1 when the current point is set 0, and the reverse (resolved on the device)
    return WaArgs{pos.a, d.pin_cell, (const float2*)d.pin_off, d.net_start, d.num_nets, d.num_mov, kp_ptr,
                  pin_grad, nullptr, p_wa, pos.b, pos.flag, block_net, pin_slot};
}

GatherArgs Level::gather_args(const Pick& pos, const float* kp_ptr, const float2* g_prev, const float2* g_prev_b,
                              float2* g_out_a, float2* g_out_b, int use_density) const {
    GatherArgs g{};
    g.pos = pos;
    g.g_prev = g_prev;
    g.g_prev_b = g_prev_b;
    g.g_out_a = g_out_a;
    g.g_out_b = g_out_b;
    g.flags = pos.flag;
    g.size = (const float2*)d.size;
    g.ratio = d.ratio;
    g.fx = fx;
    g.fy = fy;
    g.f2 = f2;
    g.ld = d.nby;
    g.pin_grad = pin_grad;
    g.cell_start = d.cell_start;
    g.cell_pin = d.cell_pin;
    g.pins = d.pins;
    g.area = d.area;
    g.kp = kp_ptr;
    g.num_nodes = d.num_nodes;
    g.num_mov = d.num_mov;
    g.inv_bx = 1.f / (float)d.bx;
    g.inv_by = 1.f / (float)d.by;
    g.nbx = d.nbx;
    g.nby = d.nby;
    g.use_density = use_density;
    return g;
}

Segs Level::segs(bool with_state, bool has_prev, int use_density) const {
    return Segs{p_moved, with_state ? nblk_n : 0, p_wa, nblk_wa, p_over, use_density ? fj.tiles + jj.tiles : 0,
                p_turned, has_prev ? wav_n : 0, p_wl2, p_dn2, wav_n};
}

double Level::over_scale() const { return d.has_judge ? d.judge_bin_area / d.mov_area : d.bin_area / d.mov_area; }

NestArgs Level::nest_args() const {
    return NestArgs{bufs(), (const float2*)d.lo, (const float2*)d.hi, d.kp, d.num_nodes, p_moved};
}

Pick Level::pass_pos() const { return Pick{(const float2*)d.v, (const float2*)v1, flags}; }
// the current point, as the kernels that read get(true) see it (the sets swapped against pass_pos: before a pass,
// the point the pass's Nesterov step starts from)
Pick Level::cur_pos() const { return Pick{(const float2*)v1, (const float2*)d.v, flags}; }

// objective at pos with the parameters kp; g_prev (optional) gives |g' - g|. With S the control pass follows.
// The objective's kernels. Passes that are not profiled run two streams: the caller's stream s takes the force grid
// (binning, tiles, Poisson solve), and a side stream forked from it takes WA and the overflow grid, which read the
// same positions and write buffers of their own. The gather waits only for the side stream's cell WA sums (done
// long before the force grid's Poisson solve ends, so the wait costs no cross-queue latency), the reduction for the
// whole side stream (its overflow grid), by when the gather has given that stream time to finish. Profiled
// passes run every kernel class in order on s, so that the events between the classes time each of them alone.
int Level::objective(const Pick& pos, const float* kp_ptr, bool with_state, const float2* g_prev,
                     const float2* g_prev_b, float2* g_out_a, float2* g_out_b, hipStream_t s, double* S, float* kp_out,
                     int use_density, bool rooms_set) {
    int st;
    const bool two = !prof;
    const bool over = use_density && jj.count;
    hipStream_t w = s;                                       // the stream of WA and the overflow grid
    if (use_density && !rooms_set) {                         // not after a Nesterov step (which sets them)
        hipLaunchKernelGGL(eng_rooms, dim3(jj.count ? 2 : 1), dim3(EB), 0, s, fj, jj);
        AP_LAUNCHED();
    }
    if (two) {
        w = side;
        AP_HIP(hipEventRecord(capturing ? cap_fork : ev_fork, s));
        AP_HIP(hipStreamWaitEvent(side, capturing ? cap_fork : ev_fork, 0));
    }
    // the side stream: WA and the cells' WA sums first, then the overflow grid (measured 4 % faster per pass than the
    // other order on the 2M-cell design: WA's streaming overlaps the force grid's binning, and the overflow grid's
    // latency-bound tiles the force grid's); profiled passes keep the class order on s
    GatherArgs g = gather_args(pos, kp_ptr, g_prev, g_prev_b, g_out_a, g_out_b, use_density);
    // two streams: the force grid's fill is captured first, so that the graph's first branch after the fork -- the
    // one that keeps the caller's hardware queue, with no cross-queue wait to start -- is the force grid's chain
    // (fill, tiles, Poisson: the longer one), and the side stream's chain takes the other queue
    if (two && use_density) {
        tbegin(AP_PROF_MAP, s);
        if ((st = bin_grid(0, pos, s))) return st;
        tend(AP_PROF_MAP, s);
    }
    mark(AP_PROF_WA, s);
    tbegin(AP_PROF_WA, w);
    if ((st = wa_plan.launch(wa_args(pos, kp_ptr), p_wa, w))) return st;
    tend(AP_PROF_WA, w);
    // the cells' WA gradients, off the gather's critical path
    mark(AP_PROF_WA_CELLS, s);
    tbegin(AP_PROF_WA_CELLS, w);
    hipLaunchKernelGGL(eng_cell_wl, dim3(nblk_n), dim3(EB), 0, w, g, cell_wl, p_wl2);
    AP_LAUNCHED();
    tend(AP_PROF_WA_CELLS, w);
    g.cell_wl = cell_wl;
    if (two) AP_HIP(hipEventRecord(capturing ? cap_cells : ev_cells, side));
    if (use_density) {
        mark(AP_PROF_MAP, s);
        if (over && (st = bin_grid(1, pos, w))) return st;
        if (!two && (st = bin_grid(0, pos, s))) return st;
        mark(AP_PROF_POST, s);
        tbegin(AP_PROF_POST, s);
        if ((st = tiles_grid(0, pos, s))) return st;
        tend(AP_PROF_POST, s);
        mark(AP_PROF_OVER_TILES, s);
        tbegin(AP_PROF_OVER_TILES, w);
        if (over && (st = tiles_grid(1, pos, w))) return st;
        tend(AP_PROF_OVER_TILES, w);
        mark(AP_PROF_POISSON, s);
        tbegin(AP_PROF_POISSON, s);
        if ((st = solve_field(s))) return st;
        tend(AP_PROF_POISSON, s);
    }
    if (two) AP_HIP(hipStreamWaitEvent(s, capturing ? cap_cells : ev_cells, 0));
    mark(AP_PROF_GATHER, s);
    tbegin(AP_PROF_GATHER, s);
    hipLaunchKernelGGL(eng_gather, dim3(grid_of(d.num_nodes, gather_lanes)), dim3(gather_lanes), 0, s, g, p_turned,
                       p_wl2, p_dn2);
    AP_LAUNCHED();
    tend(AP_PROF_GATHER, s);
    if (two) {
        AP_HIP(hipEventRecord(capturing ? cap_join : ev_join, side));
        AP_HIP(hipStreamWaitEvent(s, capturing ? cap_join : ev_join, 0));
    }
    mark(AP_PROF_CONTROL, s);
    tbegin(AP_PROF_CONTROL, s);
    hipLaunchKernelGGL(eng_reduce, dim3(RED_BLOCKS), dim3(RED_LANES), 0, s, segs(with_state, g_prev != nullptr, use_density),
                       red);
    AP_LAUNCHED();
    Ctl ctl_pass = ctl;
    if (!S) ctl_pass.params = nullptr;                       // a plain objective evaluation: no controller
    if (start_mode && S)
        hipLaunchKernelGGL(eng_start_control, dim3(1), dim3(64), 0, s, red, S, kp_out, flags, start_st, start_hist);
    else
        hipLaunchKernelGGL(eng_control, dim3(1), dim3(64), 0, s, red, use_density ? over_scale() : 0.0, S, kp_out,
                           d.c, flags, stats, ctl_pass);
    AP_LAUNCHED();
    tend(AP_PROF_CONTROL, s);
    mark(AP_PROF_CLASSES, s);
    return AP_OK;
}

int Level::pass(hipStream_t s) {
    mark(AP_PROF_NESTEROV, s);
    tbegin(AP_PROF_NESTEROV, s);
    // the rooms of the pass's fill, beside the step (none for a start pass: no binning)
    const int rooms = start_mode ? 0 : jj.count ? 2 : 1;
    hipLaunchKernelGGL(eng_nesterov, dim3(rooms + nblk_n), dim3(EB), 0, s, nest_args(), fj, jj, rooms);
    AP_LAUNCHED();
    tend(AP_PROF_NESTEROV, s);
    return objective(pass_pos(), d.kp, true, (const float2*)d.g, g1, (float2*)d.g, g1, s, d.S, d.kp,
                     start_mode ? 0 : 1, true);
}

int Level::finalize(hipStream_t s) {
    hipLaunchKernelGGL(eng_finalize_copy, dim3(nblk_n), dim3(EB), 0, s, bufs(), d.num_nodes);
    AP_LAUNCHED();
    hipLaunchKernelGGL(eng_finalize_flags, dim3(1), dim3(64), 0, s, flags);
    AP_LAUNCHED();
    return AP_OK;
}

// the captured graph of `passes` passes with the current timing class, captured now when not cached
int Level::graph_for(int passes, int& slot) {
    for (int k = 0; k < GRAPH_CACHE; k++)
        if (graphs[k].exec && graphs[k].passes == passes && graphs[k].timing == timing_cls &&
            graphs[k].start == start_mode) {
            slot = k;
            return AP_OK;
        }
    slot = graph_next;
    graph_next = (graph_next + 1) % GRAPH_CACHE;
    GraphEntry& G = graphs[slot];
    if (G.exec) {
        // the evicted graph may still run (a replay launched without a host sync since): destroying its exec, its
        // graph or its timing events while it runs frees what the runtime still reads, so wait for both streams
        AP_HIP(hipStreamSynchronize(stream));
        AP_HIP(hipStreamSynchronize(side));
    }
    if (last_ev == &G.ev) last_ev = nullptr;
    G.clear();
    if (timing_cls >= 0) {
        G.ev.resize(2 * (size_t)timed_passes(passes), nullptr);
        for (hipEvent_t& e : G.ev) AP_HIP(hipEventCreate(&e));
    }
    rec_ev = timing_cls >= 0 ? G.ev.data() : nullptr;
    rec_capture = true;
    rec_err = hipSuccess;
    AP_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    capturing = true;
    int st = AP_OK;
    for (int k = 0; k < passes && !st; k++) {
        rec_pass = k;
        st = pass(stream);
    }
    if (!st) st = finalize(stream);
    rec_ev = nullptr;
    capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(stream, &g);
    if (st || e != hipSuccess || rec_err != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        G.clear();
        (void)hipGetLastError();
        if (st) return st;
        set_error("level: graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : rec_err));
        return AP_ERR_HIP;
    }
    G.graph = g;
    AP_HIP(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
    G.passes = passes;
    G.timing = timing_cls;
    G.start = start_mode;
    return AP_OK;
}

// ------------------------------------------------------------------------------ batch: many levels per launch
// The passes of many independent levels (blocks) run as one sequence of grouped launches: every stage of the pass
// is one launch over the workgroups of all blocks, each workgroup reading its block's arguments from a table on the
// device through a work map (block, workgroup, workgroups of the block's stage). Each block keeps its own control
// state; a block that has stopped (S[STOP] != 0) is skipped by every stage but the control (which freezes it) and
// the end-of-run copy -- exactly what its own passes would leave. A run of passes is one HIP graph.
struct PassTab {
    NestArgs nest;
    Pick pos;
    BinJob fj, jj;
    float target;
    double* p_over;
    GatherArgs gather;
    double *p_turned, *p_wl2, *p_dn2;
    Segs segs;
    double* red;
    double over_scale;
    double* S;
    float* kp;
    const double* c;
    int* flags;
    double* stats;
    Ctl ctl;
    Bufs bufs;
    int num_nodes;
};

enum { ST_NEST, ST_COUNT, ST_SCAN, ST_SCATTER, ST_TILES, ST_GATHER, ST_REDUCE, ST_CONTROL, ST_FINAL, ST_FLAGS, ST_WA,
       NUM_ST };

__device__ __forceinline__ bool stopped(const PassTab* tab, int b) { return tab[b].S[STOP] != 0; }

__global__ __launch_bounds__(EB) void eng_nesterov_grouped(const PassTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    if (stopped(tab, m.x)) return;
    const NestArgs a = tab[m.x].nest;
    nesterov_body(a, m.y);
}
__global__ __launch_bounds__(EB) void eng_bin_count_grouped(const PassTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    if (stopped(tab, m.x)) return;
    const PassTab& T = tab[m.x];
    bin_count_body(T.pos, T.fj, T.jj, m.y);
}
__global__ __launch_bounds__(SCAN_T) void eng_bin_scan_grouped(const PassTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    if (stopped(tab, m.x)) return;
    const PassTab& T = tab[m.x];
    bin_scan_body(T.fj, T.jj, m.y);
}
__global__ __launch_bounds__(EB) void eng_bin_scatter_grouped(const PassTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    if (stopped(tab, m.x)) return;
    const PassTab& T = tab[m.x];
    bin_scatter_body(T.pos, T.fj, T.jj, m.y);
}
__global__ __launch_bounds__(TB) void eng_bin_tiles_grouped(const PassTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    if (stopped(tab, m.x)) return;
    const PassTab& T = tab[m.x];
    const bool judge = m.y >= T.fj.wgs;
    bin_tiles_body<TB>(T.pos, judge ? T.jj : T.fj, T.target, T.p_over + (judge ? T.fj.tiles : 0),
                   m.y - (judge ? T.fj.wgs : 0));
}
__global__ __launch_bounds__(EB) void eng_gather_grouped(const PassTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    if (stopped(tab, m.x)) return;
    const PassTab& T = tab[m.x];
    const GatherArgs g = T.gather;
    gather_body(g, T.p_turned, T.p_wl2, T.p_dn2, m.y, m.z);
}
__global__ __launch_bounds__(RED_LANES) void eng_reduce_grouped(const PassTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    if (stopped(tab, m.x)) return;
    const PassTab& T = tab[m.x];
    const Segs sg = T.segs;
    reduce_body(sg, T.red, m.y);
}
__global__ __launch_bounds__(64) void eng_control_grouped(const PassTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    const PassTab& T = tab[m.x];
    const Ctl ctl = T.ctl;
    control_body(T.red, T.over_scale, T.S, T.kp, T.c, T.flags, T.stats, ctl);
}
__global__ __launch_bounds__(EB) void eng_finalize_copy_grouped(const PassTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    const PassTab& T = tab[m.x];
    const Bufs B = T.bufs;
    const int i = m.y * EB + threadIdx.x;
    if (i >= T.num_nodes) return;
    const int cur = B.flags[0];
    if (B.flags[1]) B.best[i] = B.v[cur][i];
    if (cur == 1) {
        B.u[0][i] = B.u[1][i];
        B.v[0][i] = B.v[1][i];
        B.g[0][i] = B.g[1][i];
    }
}
__global__ void eng_finalize_flags_grouped(const PassTab* tab, const int4* map) {
    const int4 m = map[blockIdx.x];
    if (threadIdx.x == 0) {
        int* f = tab[m.x].flags;
        f[0] = 0;
        f[1] = 0;
    }
}

struct Batch {
    std::vector<Level*> levels;
    PassTab* d_tab = nullptr;
    WaveArgs* d_wave = nullptr;
    FftTab* d_fft = nullptr;
    double** d_S = nullptr;
    int4* map[NUM_ST] = {};
    int count[NUM_ST] = {};
    FftGroup fft;
    size_t hist_lds = 0, acc_lds = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr, t0 = nullptr, t1 = nullptr;
    GraphEntry graphs[GRAPH_CACHE];
    int graph_next = 0;

    int build();          // tables and maps from the levels' current set-up
    void release_tables();
    void destroy();
    int pass(hipStream_t s);
    int finalize(hipStream_t s);
    int graph_for(int passes, int& slot);
};

void Batch::release_tables() {
    void* ptrs[] = {d_tab, d_wave, d_fft, d_S};
    for (void* p : ptrs)
        if (p) (void)dev_free(p);
    d_tab = nullptr;
    d_wave = nullptr;
    d_fft = nullptr;
    d_S = nullptr;
    for (int k = 0; k < NUM_ST; k++) {
        if (map[k]) (void)dev_free(map[k]);
        map[k] = nullptr;
        count[k] = 0;
    }
    fft.destroy();
    for (GraphEntry& G : graphs) G.clear();
}

int Batch::build() {
    release_tables();
    const int nb = (int)levels.size();
    std::vector<PassTab> tab(nb);
    std::vector<WaveArgs> wave(nb);
    std::vector<FftTab> ffts(nb);
    std::vector<double*> S(nb);
    std::vector<int4> m[NUM_ST];
    hist_lds = acc_lds = 0;
    for (int b = 0; b < nb; b++) {
        Level& L = *levels[b];
        AP_REQUIRE(L.poisson.method == AP_POISSON_FFT, "batch: block %d does not use the FFT Poisson solve", b);
        AP_REQUIRE(L.wa_plan.num_big == 0, "batch: block %d has nets of more than 64 pins", b);
        const Pick pos = L.pass_pos();
        PassTab& T = tab[b];
        T.nest = L.nest_args();
        T.pos = pos;
        T.fj = L.fj;
        T.jj = L.jj;
        T.fj.fill = 0;                                 // the grouped launches: counting-sort form, no split tiles
        T.jj.fill = 0;
        T.fj.helpers = T.fj.split_need = 0;
        T.jj.helpers = T.jj.split_need = 0;
        T.target = L.d.target;
        T.p_over = L.p_over;
        T.gather = L.gather_args(pos, L.d.kp, (const float2*)L.d.g, L.g1, (float2*)L.d.g, L.g1, 1);
        T.p_turned = L.p_turned;
        T.p_wl2 = L.p_wl2;
        T.p_dn2 = L.p_dn2;
        T.segs = L.segs(true, true, 1);
        T.red = L.red;
        T.over_scale = L.over_scale();
        T.S = L.d.S;
        T.kp = L.d.kp;
        T.c = L.d.c;
        T.flags = L.flags;
        T.stats = L.stats;
        T.ctl = L.ctl;
        T.bufs = L.bufs();
        T.num_nodes = L.d.num_nodes;
        wave[b] = L.wa_plan.wave_args(L.wa_args(pos, L.d.kp), L.p_wa);
        ffts[b] = L.poisson.fft_tab(L.rho, L.d.nby, L.fx, L.fy, L.f2, L.d.nby, L.d.S);
        S[b] = L.d.S;
        hist_lds = std::max(hist_lds, L.hist_lds);
        acc_lds = std::max(acc_lds, L.acc_lds);
        auto add = [&](int st, int n) {
            for (int k = 0; k < n; k++) m[st].push_back(make_int4(b, k, n, 0));
        };
        add(ST_NEST, L.nblk_n);
        add(ST_COUNT, L.fj.blocks + L.jj.blocks);
        add(ST_SCAN, L.d.has_judge ? 2 : 1);
        add(ST_SCATTER, L.fj.blocks + L.jj.blocks);
        add(ST_TILES, L.fj.wgs + L.jj.wgs);
        add(ST_GATHER, L.nblk_n);
        add(ST_REDUCE, RED_BLOCKS);
        add(ST_CONTROL, 1);
        add(ST_FINAL, L.nblk_n);
        add(ST_FLAGS, 1);
        add(ST_WA, L.wa_plan.wave_blocks(WA_PLAN_PER_WAVE, WA_LANES));   // the grouped launch's shape
    }
    AP_HIP(dev_malloc(&d_tab, nb * sizeof(PassTab)));
    AP_HIP(hipMemcpy(d_tab, tab.data(), nb * sizeof(PassTab), hipMemcpyHostToDevice));
    AP_HIP(dev_malloc(&d_wave, nb * sizeof(WaveArgs)));
    AP_HIP(hipMemcpy(d_wave, wave.data(), nb * sizeof(WaveArgs), hipMemcpyHostToDevice));
    AP_HIP(dev_malloc(&d_fft, nb * sizeof(FftTab)));
    AP_HIP(hipMemcpy(d_fft, ffts.data(), nb * sizeof(FftTab), hipMemcpyHostToDevice));
    AP_HIP(dev_malloc(&d_S, nb * sizeof(double*)));
    AP_HIP(hipMemcpy(d_S, S.data(), nb * sizeof(double*), hipMemcpyHostToDevice));
    for (int k = 0; k < NUM_ST; k++) {
        count[k] = (int)m[k].size();
        if (!count[k]) continue;
        AP_HIP(dev_malloc(&map[k], m[k].size() * sizeof(int4)));
        AP_HIP(hipMemcpy(map[k], m[k].data(), m[k].size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    int st = fft.create(ffts);
    if (st) return st;
    AP_HIP(hipFuncSetAttribute((const void*)eng_bin_count_grouped, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)hist_lds));
    AP_HIP(hipFuncSetAttribute((const void*)eng_bin_tiles_grouped, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)acc_lds));
    return AP_OK;
}

void Batch::destroy() {
    release_tables();
    for (Level* L : levels) {
        L->destroy();
        delete L;
    }
    levels.clear();
    if (stream) (void)hipStreamDestroy(stream);
    hipEvent_t evs[] = {ev_in, ev_out, t0, t1};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
}

int Batch::pass(hipStream_t s) {
    int st;
#define AP_GROUPED(kernel, stage, threads, lds)                                                       \
    if (count[stage]) {                                                                            \
        hipLaunchKernelGGL(kernel, dim3(count[stage]), dim3(threads), lds, s, d_tab, map[stage]);  \
        AP_LAUNCHED();                                                                             \
    }
    AP_GROUPED(eng_nesterov_grouped, ST_NEST, EB, 0);
    if ((st = launch_wa_grouped(d_wave, map[ST_WA], count[ST_WA], d_S, s))) return st;
    AP_GROUPED(eng_bin_count_grouped, ST_COUNT, EB, hist_lds);
    AP_GROUPED(eng_bin_scan_grouped, ST_SCAN, SCAN_T, 0);
    AP_GROUPED(eng_bin_scatter_grouped, ST_SCATTER, EB, 0);
    AP_GROUPED(eng_bin_tiles_grouped, ST_TILES, TB, acc_lds);
    if ((st = fft.launch(d_fft, s))) return st;
    AP_GROUPED(eng_gather_grouped, ST_GATHER, EB, 0);
    AP_GROUPED(eng_reduce_grouped, ST_REDUCE, RED_LANES, 0);
    AP_GROUPED(eng_control_grouped, ST_CONTROL, 64, 0);
    return AP_OK;
}

int Batch::finalize(hipStream_t s) {
    AP_GROUPED(eng_finalize_copy_grouped, ST_FINAL, EB, 0);
    AP_GROUPED(eng_finalize_flags_grouped, ST_FLAGS, 64, 0);
#undef AP_GROUPED
    return AP_OK;
}

int Batch::graph_for(int passes, int& slot) {
    for (int k = 0; k < GRAPH_CACHE; k++)
        if (graphs[k].exec && graphs[k].passes == passes) {
            slot = k;
            return AP_OK;
        }
    slot = graph_next;
    graph_next = (graph_next + 1) % GRAPH_CACHE;
    GraphEntry& G = graphs[slot];
    if (G.exec) AP_HIP(hipStreamSynchronize(stream));   // the evicted graph may still run: wait before destroying it
    G.clear();
    AP_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int st = AP_OK;
    for (int k = 0; k < passes && !st; k++) st = pass(stream);
    if (!st) st = finalize(stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(stream, &g);
    if (st || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (st) return st;
        AP_HIP(e);
    }
    G.graph = g;
    AP_HIP(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
    G.passes = passes;
    return AP_OK;
}

// ------------------------------------------------------------------------------ drop-in density map and gather
// The reference's gp_cuda.density_map and density_grad (gp_cuda.cpp:41-65) on the engine's own machinery, so that a
// caller of the six drop-in functions (the reference's gpu_loop.py) gets the same design as the engine: the map by
// the tile binning in its counting-sort form (no global float or 64-bit atomics on the map: LDS integer sums per
// tile, the reference's fixed-point contributions, bit-exact), the gradient by the engine's gather. Stateless: one
// call's buckets live in a workspace (its size from dropin_map_plan) and every launch goes to the given stream.
__global__ __launch_bounds__(EB) void dm_clear(BinJob J) {
    const int n = J.planes * J.tiles;
    for (int k = blockIdx.x * EB + threadIdx.x; k < n; k += gridDim.x * EB) J.count[k] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *J.big_count = 0;
}
__global__ __launch_bounds__(EB) void dm_count(Pick pos, BinJob J) { bin_count_body(pos, J, J, blockIdx.x); }
__global__ __launch_bounds__(SCAN_T) void dm_scan(BinJob J) { scan_buckets<SCAN_T>(J); }
__global__ __launch_bounds__(EB) void dm_scatter(Pick pos, BinJob J) { bin_scatter_body(pos, J, J, blockIdx.x); }
__global__ __launch_bounds__(TB_LEVEL, 8) void dm_tiles(Pick pos, BinJob J, float target, double* part) {
    bin_tiles_body<TB_LEVEL>(pos, J, target, part, blockIdx.x);
}
__global__ __launch_bounds__(EB) void dm_gather(GatherArgs a) {
    gather_body<true>(a, nullptr, nullptr, nullptr, blockIdx.x, gridDim.x);
}

// The drop-in map's tile side and workspace layout: the power of two 8 .. 64 at or below the side that gives
// DROPIN_TILE_NODES nodes per tile (no reach count: a node over 2 x 2 tiles takes the big-node path, which every
// tile clips), one tile per workgroup, each array on a 256-byte boundary. A stateless call has no autotuner; on the
// 200k-cell design (force grid 1024 x 1024, overflow grid 512 x 512) sides 16 and 16 measured fastest of every side
// and group (a whole call 100 and 62 us; 113 and 96 us at sides 32 and 32, which 1024 nodes per tile would give),
// and more tiles per workgroup were slower at every side.
constexpr int DROPIN_TILE_NODES = 256;
struct DropMap {
    int T = 8, planes = 1;
    size_t count = 0, start = 0, big_count = 0, big_n = 0, off = 0, big = 0, geo = 0, gr = 0, part = 0, bytes = 0;
    long long room = 0;
};

static DropMap dropin_map_plan(int num_nodes, int num_mov, int nbx, int nby) {
    DropMap m;
    m.planes = num_mov < num_nodes ? 2 : 1;
    const double want = std::sqrt((double)DROPIN_TILE_NODES * nbx * nby / (double)std::max(num_nodes, 1));
    while (m.T < 64 && 2 * m.T <= want) m.T *= 2;
    while (m.T < 64 && (long long)m.planes * grid_of(nbx, m.T) * grid_of(nby, m.T) > BIN_MAX_BUCKETS) m.T *= 2;
    m.room = 4LL * num_nodes + 1;                  // a small node has at most 2 x 2 entries
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t o = at;
        at += (bytes + 255) & ~(size_t)255;
        return o;
    };
    m.count = take(BIN_MAX_BUCKETS * sizeof(int));
    m.start = take((BIN_MAX_BUCKETS + 1) * sizeof(int));
    m.big_count = take(sizeof(int));
    m.big_n = take(sizeof(int));
    m.off = take((size_t)num_nodes * sizeof(int4));
    m.big = take((size_t)num_nodes * sizeof(int));
    m.geo = take((size_t)m.room * sizeof(float4));
    m.gr = take((size_t)m.room * sizeof(float));
    m.part = take(BIN_MAX_BUCKETS * sizeof(double));
    m.bytes = at;
    return m;
}

static int dropin_map(const float* pos, const float* size, const float* ratio, int num_nodes, int num_mov, float bx,
                      float by, int nbx, int nby, const float* fixed, float target, float* rho, float* over,
                      void* work, long long work_bytes, hipStream_t s) {
    const DropMap m = dropin_map_plan(num_nodes, num_mov, nbx, nby);
    AP_REQUIRE((long long)m.planes * grid_of(nbx, m.T) * grid_of(nby, m.T) <= BIN_MAX_BUCKETS,
               "density_map: grid %d x %d has more tiles than the binning's %d buckets", nbx, nby, BIN_MAX_BUCKETS);
    AP_REQUIRE(work && work_bytes >= (long long)m.bytes, "density_map: workspace of %lld bytes, %zu needed",
               work_bytes, m.bytes);
    char* w = (char*)work;
    BinJob J{};
    J.size = (const float2*)size;
    J.ratio = ratio;
    J.num_nodes = num_nodes;
    J.num_mov = num_mov;
    J.inv_bx = 1.f / bx;          // the same float32 reciprocal as the reference's host code
    J.inv_by = 1.f / by;
    J.nbx = nbx;
    J.nby = nby;
    J.T = J.T0 = m.T;
    J.tiles_y = (int)grid_of(nby, m.T);
    J.tiles = (int)grid_of(nbx, m.T) * J.tiles_y;
    J.G = 1;
    J.wgs = J.tiles;
    J.planes = m.planes;
    J.blocks = (int)grid_of(num_nodes, BIN_NODES);
    J.count = (int*)(w + m.count);
    J.start = (int*)(w + m.start);
    J.big_count = (int*)(w + m.big_count);
    J.big_n = (int*)(w + m.big_n);
    J.off = (int4*)(w + m.off);
    J.big = (int*)(w + m.big);
    J.geo = (float4*)(w + m.geo);
    J.gr = (float*)(w + m.gr);
    J.fixed = fixed;
    J.room = (int)std::min<long long>(m.room, 0x7fffffff);
    J.rho_mirror = rho;
    J.over_map = over;
    const Pick p{(const float2*)pos, nullptr, nullptr};
    const size_t hist = (size_t)J.planes * J.tiles * sizeof(int);
    const size_t acc = (size_t)J.planes * J.T * J.T * sizeof(unsigned long long);
    AP_HIP(hipFuncSetAttribute((const void*)dm_count, hipFuncAttributeMaxDynamicSharedMemorySize, (int)hist));
    AP_HIP(hipFuncSetAttribute((const void*)dm_tiles, hipFuncAttributeMaxDynamicSharedMemorySize, (int)acc));
    hipLaunchKernelGGL(dm_clear, dim3(std::max(1, (int)grid_of((long long)J.planes * J.tiles, EB))), dim3(EB), 0, s,
                       J);
    AP_LAUNCHED();
    if (num_nodes > 0) {
        hipLaunchKernelGGL(dm_count, dim3(J.blocks), dim3(EB), hist, s, p, J);
        AP_LAUNCHED();
    }
    hipLaunchKernelGGL(dm_scan, dim3(1), dim3(SCAN_T), 0, s, J);
    AP_LAUNCHED();
    if (num_nodes > 0) {
        hipLaunchKernelGGL(dm_scatter, dim3(J.blocks), dim3(EB), 0, s, p, J);
        AP_LAUNCHED();
    }
    hipLaunchKernelGGL(dm_tiles, dim3(J.wgs), dim3(TB_LEVEL), acc, s, p, J, target, (double*)(w + m.part));
    AP_LAUNCHED();
    return AP_OK;
}

static int dropin_grad(const float* pos, const float* size, const float* ratio, const float* field,
                       const float* wl_grad, const float* pins, const float* area, int num_nodes, int num_wl,
                       const float* weight, const float* precond, float bx, float by, int nbx, int nby, float* grad,
                       hipStream_t s) {
    GatherArgs a{};
    a.pos = Pick{(const float2*)pos, nullptr, nullptr};
    a.g_out_a = (float2*)grad;
    a.size = (const float2*)size;
    a.ratio = ratio;
    a.fx = field;                             // the reference's field [2][2 nbx][2 nby]: the [:nbx, :nby] quadrants
    a.fy = field + 4LL * nbx * nby;
    a.ld = 2 * nby;
    a.cell_wl = num_wl > 0 ? (const float2*)wl_grad : nullptr;
    a.pins = pins;
    a.area = area;
    a.num_nodes = num_nodes;
    a.num_mov = num_wl;
    a.inv_bx = 1.f / bx;
    a.inv_by = 1.f / by;
    a.nbx = nbx;
    a.nby = nby;
    a.use_density = 1;
    a.weight_p = weight;
    a.precond_p = precond;
    hipLaunchKernelGGL(dm_gather, dim3(grid_of(num_nodes, EB)), dim3(EB), 0, s, a);
    AP_LAUNCHED();
    return AP_OK;
}

}  // namespace ap

using namespace ap;

struct ap_level {
    Level L;
};

extern "C" {

int ap_level_create(const ap_level_desc* desc, ap_level** level) {
    AP_REQUIRE(desc && level, "level_create: null argument");
    ap_level* h = new ap_level();
    int st = h->L.create(*desc);
    if (st) {
        h->L.destroy();
        delete h;
        return st;
    }
    *level = h;
    return AP_OK;
}

int ap_level_destroy(ap_level* level) {
    if (level) {
        (void)hipDeviceSynchronize();
        level->L.destroy();
        delete level;
    }
    return AP_OK;
}

int ap_level_prepare(ap_level* level, int passes) {
    AP_REQUIRE(level && passes > 0, "level_prepare: bad arguments");
    int slot;
    return level->L.graph_for(passes, slot);
}

static int run_passes(Level& L, int passes, int use_graph, void* stream);

int ap_level_run(ap_level* level, int passes, int use_graph, void* stream) {
    AP_REQUIRE(level && passes >= 0, "level_run: bad arguments");
    level->L.start_mode = false;
    return run_passes(level->L, passes, use_graph, stream);
}

int ap_level_start_begin(ap_level* level, double tol, int window, int max_passes) {
    AP_REQUIRE(level && tol >= 0 && window >= 1 && max_passes >= 1,
               "level_start_begin: tol %g, window %d, max_passes %d", tol, window, max_passes);
    Level& L = level->L;
    AP_HIP(hipStreamSynchronize(L.stream));
    if (!L.start_st) AP_HIP(dev_malloc((void**)&L.start_st, START_STATE * sizeof(double)));
    if (L.start_cap < max_passes) {
        dev_free(L.start_hist);
        L.start_hist = nullptr;
        L.start_cap = 0;
        AP_HIP(dev_malloc((void**)&L.start_hist, 2 * (size_t)max_passes * sizeof(double)));
        L.start_cap = max_passes;
    }
    const double st[START_STATE] = {0.0, 0.0, tol, (double)window, (double)max_passes, 0.0, 0.0, 0.0};
    AP_HIP(hipMemcpy(L.start_st, st, sizeof(st), hipMemcpyHostToDevice));
    return AP_OK;
}

int ap_level_start_run(ap_level* level, int passes, int use_graph, void* stream) {
    AP_REQUIRE(level && passes >= 0, "level_start_run: bad arguments");
    Level& L = level->L;
    AP_REQUIRE(L.start_st, "level_start_run: no start state (ap_level_start_begin)");
    L.start_mode = true;
    const int st = run_passes(L, passes, use_graph, stream);
    L.start_mode = false;
    L.rooms_stale = true;           // the points moved with no binning: the next density pass counts afresh
    return st;
}

int ap_level_start_state(ap_level* level, double* state, double* history, int n) {
    AP_REQUIRE(level && state && n >= 0 && (n == 0 || history), "level_start_state: bad arguments");
    Level& L = level->L;
    AP_REQUIRE(L.start_st, "level_start_state: no start state (ap_level_start_begin)");
    AP_HIP(hipStreamSynchronize(L.stream));
    AP_HIP(hipMemcpy(state, L.start_st, START_STATE * sizeof(double), hipMemcpyDeviceToHost));
    const int count = (int)state[SS_COUNT];
    AP_REQUIRE(n <= L.start_cap, "level_start_state: %d passes asked, %d kept", n, L.start_cap);
    if (n) AP_HIP(hipMemcpy(history, L.start_hist, 2 * (size_t)(n < count ? n : count) * sizeof(double),
                            hipMemcpyDeviceToHost));
    return AP_OK;
}

static int run_passes(Level& L, int passes, int use_graph, void* stream) {
    hipStream_t caller = (hipStream_t)stream;
    int st = AP_OK, slot = -1;
    if (use_graph && passes > 0 && (st = L.graph_for(passes, slot))) return st;
    AP_HIP(hipEventRecord(L.ev_in, caller));
    AP_HIP(hipStreamWaitEvent(L.stream, L.ev_in, 0));
    if (!L.start_mode && (st = L.init_rooms(L.cur_pos(), L.stream))) return st;
    AP_HIP(hipEventRecord(L.t0, L.stream));
    if (slot >= 0) {
        GraphEntry& G = L.graphs[slot];
        AP_HIP(hipGraphLaunch(G.exec, L.stream));
        L.last_ev = G.timing >= 0 ? &G.ev : nullptr;
        L.last_ev_passes = (int)(G.ev.size() / 2);
    } else {
        const bool timed = L.timing_cls >= 0 && passes > 0;
        if (timed) {
            if (L.last_ev == &L.direct_ev) L.last_ev = nullptr;
            while (L.direct_ev.size() < 2 * (size_t)L.timed_passes(passes)) {
                hipEvent_t e = nullptr;
                AP_HIP(hipEventCreate(&e));
                L.direct_ev.push_back(e);
            }
            L.rec_ev = L.direct_ev.data();
            L.rec_capture = false;
            L.rec_err = hipSuccess;
        }
        for (int k = 0; k < passes && !st; k++) {
            L.rec_pass = k;
            st = L.pass(L.stream);
        }
        L.rec_ev = nullptr;
        if (!st) st = L.finalize(L.stream);
        if (st) return st;
        if (L.rec_err != hipSuccess) {
            set_error("level_run: timing event: %s", hipGetErrorString(L.rec_err));
            return AP_ERR_HIP;
        }
        L.last_ev = timed ? &L.direct_ev : nullptr;
        L.last_ev_passes = timed ? L.timed_passes(passes) : 0;
    }
    AP_HIP(hipEventRecord(L.t1, L.stream));
    AP_HIP(hipEventRecord(L.ev_out, L.stream));
    AP_HIP(hipStreamWaitEvent(caller, L.ev_out, 0));
    return AP_OK;
}

int ap_level_time_class(ap_level* level, int cls) {
    AP_REQUIRE(level && cls >= -1 && cls < AP_PROF_CLASSES, "level_time_class: bad arguments");
    level->L.timing_cls = cls;
    return AP_OK;
}

int ap_level_time_every(ap_level* level, int every) {
    AP_REQUIRE(level && every > 0, "level_time_every: bad arguments");
    Level& L = level->L;
    AP_HIP(hipDeviceSynchronize());
    L.time_every = every;
    for (GraphEntry& G : L.graphs) {                 // the timed passes change: capture again
        if (L.last_ev == &G.ev) L.last_ev = nullptr;
        G.clear();
    }
    return AP_OK;
}

int ap_level_class_ms(ap_level* level, double* ms, int* passes) {
    AP_REQUIRE(level && ms && passes, "level_class_ms: null argument");
    Level& L = level->L;
    AP_REQUIRE(L.last_ev && L.last_ev_passes > 0, "level_class_ms: the last run had no timed class");
    const std::vector<hipEvent_t>& ev = *L.last_ev;
    const int n = L.last_ev_passes;
    AP_HIP(hipEventSynchronize(ev[2 * n - 1]));
    double sum = 0;
    for (int k = 0; k < n; k++) {
        float t = 0;
        AP_HIP(hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]));
        sum += t;
    }
    *ms = sum / n;
    *passes = n;
    return AP_OK;
}

int ap_level_profile(ap_level* level, int passes, double* ms, void* stream) {
    AP_REQUIRE(level && ms && passes > 0, "level_profile: bad arguments");
    Level& L = level->L;
    hipStream_t caller = (hipStream_t)stream;
    for (int k = 0; k <= AP_PROF_CLASSES; k++)
        if (!L.pev[k]) AP_HIP(hipEventCreate(&L.pev[k]));
    for (int k = 0; k < AP_PROF_CLASSES; k++) L.prof_ms[k] = 0;
    AP_HIP(hipEventRecord(L.ev_in, caller));
    AP_HIP(hipStreamWaitEvent(L.stream, L.ev_in, 0));
    int st = L.init_rooms(L.cur_pos(), L.stream);
    if (st) return st;
    L.prof = true;
    for (int p = 0; p < passes && !st; p++) {
        st = L.pass(L.stream);
        if (st) break;
        if (hipEventSynchronize(L.pev[AP_PROF_CLASSES]) != hipSuccess) {
            st = AP_ERR_HIP;
            break;
        }
        for (int k = 0; k < AP_PROF_CLASSES; k++) {
            float t = 0;
            (void)hipEventElapsedTime(&t, L.pev[k], L.pev[k + 1]);
            L.prof_ms[k] += t;
        }
    }
    L.prof = false;
    if (!st) st = L.finalize(L.stream);
    if (st) {
        set_error("level_profile: a pass failed");
        return st;
    }
    AP_HIP(hipEventRecord(L.ev_out, L.stream));
    AP_HIP(hipStreamWaitEvent(caller, L.ev_out, 0));
    for (int k = 0; k < AP_PROF_CLASSES; k++) ms[k] = L.prof_ms[k] / passes;
    return AP_OK;
}

int ap_level_last_ms(ap_level* level, float* ms) {
    AP_REQUIRE(level && ms, "level_last_ms: null argument");
    AP_HIP(hipEventSynchronize(level->L.t1));
    AP_HIP(hipEventElapsedTime(ms, level->L.t0, level->L.t1));
    return AP_OK;
}

int ap_level_objective(ap_level* level, const float* pos, float* grad, double* stats, void* stream) {
    AP_REQUIRE(level && pos && grad && stats, "level_objective: null argument");
    Level& L = level->L;
    hipStream_t s = (hipStream_t)stream;
    Pick p{(const float2*)pos, nullptr, nullptr};
    int st = L.init_rooms(p, s);
    if (st) return st;
    st = L.objective(p, L.d.kp, false, nullptr, nullptr, (float2*)grad, (float2*)grad, s, nullptr, nullptr, 1);
    if (st) return st;
    AP_HIP(hipMemcpyAsync(stats, L.stats, 5 * sizeof(double), hipMemcpyDeviceToDevice, s));
    return AP_OK;
}

int ap_level_density(ap_level* level, float* rho, void* stream) {
    AP_REQUIRE(level && rho, "level_density: null argument");
    Level& L = level->L;
    AP_HIP(hipMemcpyAsync(rho, L.rho, (size_t)L.d.nbx * L.d.nby * sizeof(float), hipMemcpyDeviceToDevice,
                          (hipStream_t)stream));
    return AP_OK;
}

int ap_level_set_controller(ap_level* level, const float* params, float sigma, unsigned long long seed, int hold,
                            float phase, float* log, int log_cap, int window) {
    AP_REQUIRE(level && sigma >= 0 && hold > 0 && log_cap >= 0 && window > 0 && (!log || log_cap > 0),
               "level_set_controller: bad arguments");
    Level& L = level->L;
    AP_HIP(hipDeviceSynchronize());
    for (GraphEntry& G : L.graphs) {                 // the kernels' arguments change: capture again
        if (L.last_ev == &G.ev) L.last_ev = nullptr;
        G.clear();
    }
    L.ctl.params = params;
    L.ctl.sigma = sigma;
    L.ctl.seed = seed;
    L.ctl.hold = hold;
    L.ctl.phase = phase;
    L.ctl.log = log;
    L.ctl.log_cap = log ? log_cap : 0;
    L.ctl.window = window;
    AP_HIP(hipMemset(L.ctl_x, 0, NUM_X * sizeof(double)));
    AP_HIP(hipDeviceSynchronize());
    return AP_OK;
}

struct ap_batch {
    Batch B;
};

int ap_batch_create(const ap_level_desc* blocks, int count, ap_batch** batch) {
    AP_REQUIRE(blocks && batch && count > 0, "batch_create: bad arguments");
    ap_batch* h = new ap_batch();
    Batch& B = h->B;
    int st = AP_OK;
    for (int b = 0; b < count && !st; b++) {
        Level* L = new Level();
        st = L->create(blocks[b]);
        if (st) {
            L->destroy();
            delete L;
            break;
        }
        B.levels.push_back(L);
    }
    if (!st) st = B.build();
    if (!st) {
        const hipError_t e = hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking);
        if (e == hipSuccess && hipEventCreateWithFlags(&B.ev_in, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&B.ev_out, hipEventDisableTiming) == hipSuccess &&
            hipEventCreate(&B.t0) == hipSuccess && hipEventCreate(&B.t1) == hipSuccess) {
            st = AP_OK;
        } else {
            set_error("batch_create: stream / events");
            st = AP_ERR_HIP;
        }
    }
    if (st) {
        B.destroy();
        delete h;
        return st;
    }
    *batch = h;
    return AP_OK;
}

int ap_batch_destroy(ap_batch* batch) {
    if (batch) {
        (void)hipDeviceSynchronize();
        batch->B.destroy();
        delete batch;
    }
    return AP_OK;
}

int ap_batch_prepare(ap_batch* batch, int passes) {
    AP_REQUIRE(batch && passes > 0, "batch_prepare: bad arguments");
    int slot;
    return batch->B.graph_for(passes, slot);
}

int ap_batch_run(ap_batch* batch, int passes, int use_graph, void* stream) {
    AP_REQUIRE(batch && passes >= 0, "batch_run: bad arguments");
    Batch& B = batch->B;
    hipStream_t caller = (hipStream_t)stream;
    int st = AP_OK, slot = -1;
    if (use_graph && passes > 0 && (st = B.graph_for(passes, slot))) return st;
    AP_HIP(hipEventRecord(B.ev_in, caller));
    AP_HIP(hipStreamWaitEvent(B.stream, B.ev_in, 0));
    AP_HIP(hipEventRecord(B.t0, B.stream));
    if (slot >= 0) {
        AP_HIP(hipGraphLaunch(B.graphs[slot].exec, B.stream));
    } else {
        for (int k = 0; k < passes && !st; k++) st = B.pass(B.stream);
        if (!st) st = B.finalize(B.stream);
        if (st) return st;
    }
    AP_HIP(hipEventRecord(B.t1, B.stream));
    AP_HIP(hipEventRecord(B.ev_out, B.stream));
    AP_HIP(hipStreamWaitEvent(caller, B.ev_out, 0));
    return AP_OK;
}

int ap_batch_last_ms(ap_batch* batch, float* ms) {
    AP_REQUIRE(batch && ms, "batch_last_ms: null argument");
    AP_HIP(hipEventSynchronize(batch->B.t1));
    AP_HIP(hipEventElapsedTime(ms, batch->B.t0, batch->B.t1));
    return AP_OK;
}

int ap_level_autotune(ap_level* level, int reps, int flags, int* sides) {
    AP_REQUIRE(level && reps > 0 && (flags & ~AP_AUTOTUNE_ONLINE) == 0, "level_autotune: bad arguments");
    Level& L = level->L;
    return L.autotune(reps, L.stream, sides, (flags & AP_AUTOTUNE_ONLINE) != 0);
}

int ap_level_set_tile_layout(ap_level* level, int force_side, int force_group, int overflow_side,
                             int overflow_group) {
    AP_REQUIRE(level, "level_set_tile_layout: null argument");
    Level& L = level->L;
    AP_REQUIRE(L.side_ok(L.fj, force_side) &&
                   (!force_group || group_ok(L.fj.planes, force_side, force_group)),
               "level_set_tile_layout: side %d, group %d do not fit the force grid", force_side, force_group);
    AP_REQUIRE(!L.jj.count || (L.side_ok(L.jj, overflow_side) &&
                               (!overflow_group || group_ok(L.jj.planes, overflow_side, overflow_group))),
               "level_set_tile_layout: side %d, group %d do not fit the overflow grid", overflow_side,
               overflow_group);
    AP_HIP(hipDeviceSynchronize());
    L.set_side(L.fj, force_side, force_group);
    if (L.jj.count) L.set_side(L.jj, overflow_side, overflow_group);
    L.update_lds();
    const int st = L.set_lds();
    if (st) return st;
    for (GraphEntry& G : L.graphs) {
        if (L.last_ev == &G.ev) L.last_ev = nullptr;
        G.clear();
    }
    return AP_OK;
}

int ap_level_launch_shape(ap_level* level, int* shape) {
    AP_REQUIRE(level && shape, "level_launch_shape: null argument");
    const Level& L = level->L;
    shape[0] = L.wa_plan.per;
    shape[1] = L.wa_plan.lanes;
    shape[2] = L.gather_lanes;
    return AP_OK;
}

int ap_level_set_launch_shape(ap_level* level, int wa_per, int wa_lanes, int gather_lanes) {
    AP_REQUIRE(level, "level_set_launch_shape: null argument");
    AP_REQUIRE(WaPlan::shape_ok(wa_per, wa_lanes) && gather_lanes >= 64 && gather_lanes <= 512 &&
                   gather_lanes % 64 == 0,
               "level_set_launch_shape: WA %d x %d lanes, gather %d lanes outside the kernels' shapes", wa_per,
               wa_lanes, gather_lanes);
    Level& L = level->L;
    AP_HIP(hipDeviceSynchronize());
    L.wa_plan.per = wa_per;
    L.wa_plan.lanes = wa_lanes;
    L.gather_lanes = gather_lanes;
    for (GraphEntry& G : L.graphs) {                 // the launches change: capture again
        if (L.last_ev == &G.ev) L.last_ev = nullptr;
        G.clear();
    }
    return AP_OK;
}

int ap_level_set_bucket_room(ap_level* level, int percent, int extra) {
    AP_REQUIRE(level && percent >= 0 && percent <= 150 && extra >= 0 && extra <= 64,
               "level_set_bucket_room: percent %d, extra %d outside 0..150, 0..64", percent, extra);
    Level& L = level->L;
    AP_HIP(hipDeviceSynchronize());
    for (BinJob* J : {&L.fj, &L.jj}) {
        J->room_pct = percent;
        J->room_add = extra;
    }
    for (GraphEntry& G : L.graphs) {                 // the kernels' arguments change: capture again
        if (L.last_ev == &G.ev) L.last_ev = nullptr;
        G.clear();
    }
    return AP_OK;
}

int ap_level_set_tile_split(ap_level* level, int slice, int factor) {
    AP_REQUIRE(level && (slice == 0 || (slice >= 64 && slice <= (1 << 24))) && factor >= 0 && factor <= 1024,
               "level_set_tile_split: slice %d outside 0 (off) or 64 .. 2^24 entries, or factor %d outside 0 .. 1024",
               slice, factor);
    Level& L = level->L;
    AP_HIP(hipDeviceSynchronize());
    for (BinJob* J : {&L.fj, &L.jj})
        if (J->helpers) {
            J->split_slice = slice;
            J->split_factor = factor;
            split_threshold(*J);
        }
    for (GraphEntry& G : L.graphs) {                 // the kernels' arguments change: capture again
        if (L.last_ev == &G.ev) L.last_ev = nullptr;
        G.clear();
    }
    return AP_OK;
}

int ap_level_tile_splits(ap_level* level, unsigned long long* splits) {
    AP_REQUIRE(level && splits, "level_tile_splits: null argument");
    Level& L = level->L;
    AP_HIP(hipDeviceSynchronize());
    splits[0] = splits[1] = 0;
    if (L.fj.split_n) AP_HIP(hipMemcpy(&splits[0], L.fj.split_n, sizeof(unsigned long long), hipMemcpyDeviceToHost));
    if (L.jj.split_n) AP_HIP(hipMemcpy(&splits[1], L.jj.split_n, sizeof(unsigned long long), hipMemcpyDeviceToHost));
    return AP_OK;
}

int ap_level_bucket_spills(ap_level* level, unsigned long long* spills) {
    AP_REQUIRE(level && spills, "level_bucket_spills: null argument");
    Level& L = level->L;
    AP_HIP(hipDeviceSynchronize());
    spills[0] = spills[1] = 0;
    if (L.fj.spill_n) AP_HIP(hipMemcpy(&spills[0], L.fj.spill_n, sizeof(unsigned long long), hipMemcpyDeviceToHost));
    if (L.jj.spill_n) AP_HIP(hipMemcpy(&spills[1], L.jj.spill_n, sizeof(unsigned long long), hipMemcpyDeviceToHost));
    return AP_OK;
}

int ap_batch_autotune(ap_batch* batch, int reps, int* layout) {
    AP_REQUIRE(batch && reps > 0, "batch_autotune: bad arguments");
    Batch& B = batch->B;
    for (size_t b = 0; b < B.levels.size(); b++) {
        int four[4];
        const int st = B.levels[b]->autotune(reps, B.stream, four, false, false);
        if (st) return st;
        if (layout) {
            layout[2 * b] = four[0];
            layout[2 * b + 1] = four[1];
        }
    }
    return B.build();
}

int ap_batch_tile_layout(ap_batch* batch, int* layout) {
    AP_REQUIRE(batch && layout, "batch_tile_layout: null argument");
    for (size_t b = 0; b < batch->B.levels.size(); b++) {
        layout[2 * b] = batch->B.levels[b]->fj.T;
        layout[2 * b + 1] = batch->B.levels[b]->fj.G;
    }
    return AP_OK;
}

int ap_level_controller_state(ap_level* level, double* x) {
    AP_REQUIRE(level && x, "level_controller_state: null argument");
    Level& L = level->L;
    AP_HIP(hipStreamSynchronize(L.stream));
    AP_HIP(hipMemcpy(x, L.ctl_x, NUM_X * sizeof(double), hipMemcpyDeviceToHost));
    return AP_OK;
}

// ---------------------------------------------------------------------- drop-in density map and gather (part 1)

long long ap_density_map_work_bytes(int num_nodes, int num_mov, int nbx, int nby) {
    if (num_nodes < 0 || num_mov < 0 || nbx <= 0 || nby <= 0) return -1;
    return (long long)dropin_map_plan(num_nodes, num_mov, nbx, nby).bytes;
}

int ap_density_map_ws(const float* pos, const float* size, const float* ratio, int num_nodes, int num_mov, float bx,
                      float by, int nbx, int nby, const float* fixed, float target, float* rho, float* over,
                      void* work, long long work_bytes, void* stream) {
    AP_REQUIRE(nbx > 0 && nby > 0 && num_nodes >= 0 && num_mov >= 0 && num_mov <= num_nodes,
               "density_map: bad grid %d x %d", nbx, nby);
    AP_REQUIRE(fixed && rho && over && (num_nodes == 0 || (pos && size && ratio)), "density_map: null pointer");
    AP_REQUIRE(((uintptr_t)work & 15) == 0, "density_map: workspace not 16-byte aligned");
    return dropin_map(pos, size, ratio, num_nodes, num_mov, bx, by, nbx, nby, fixed, target, rho, over, work,
                      work_bytes, (hipStream_t)stream);
}

int ap_density_map(const float* pos, const float* size, const float* ratio, int num_nodes, int num_mov, float bx,
                   float by, int nbx, int nby, unsigned long long* map, const float* fixed, float target, float* rho,
                   float* over, void* stream) {
    AP_REQUIRE(nbx > 0 && nby > 0 && num_nodes >= 0 && num_mov >= 0 && num_mov <= num_nodes,
               "density_map: bad grid %d x %d", nbx, nby);
    hipStream_t s = (hipStream_t)stream;
    const long long need = (long long)dropin_map_plan(num_nodes, num_mov, nbx, nby).bytes;
    const long long have = map ? 16LL * nbx * nby : 0;       // the caller's int64 [2][nbx][nby] map
    if (have >= need)
        return ap_density_map_ws(pos, size, ratio, num_nodes, num_mov, bx, by, nbx, nby, fixed, target, rho, over,
                                 map, have, stream);
    void* work = nullptr;                                    // stream-ordered, from the device's pool
    AP_HIP(hipMallocAsync(&work, (size_t)need, s));
    const int st = ap_density_map_ws(pos, size, ratio, num_nodes, num_mov, bx, by, nbx, nby, fixed, target, rho,
                                     over, work, need, stream);
    const hipError_t e = hipFreeAsync(work, s);
    if (st) return st;
    AP_HIP(e);
    return AP_OK;
}

int ap_density_grad(const float* pos, const float* size, const float* ratio, const float* field, const float* wl_grad,
                    const float* pins, const float* area, int num_nodes, int num_wl, const float* weight,
                    const float* precond, float bx, float by, int nbx, int nby, float* grad, void* stream) {
    AP_REQUIRE(nbx > 0 && nby > 0 && num_nodes >= 0 && num_wl >= 0 && num_wl <= num_nodes, "density_grad: bad sizes");
    if (num_nodes == 0) return AP_OK;
    AP_REQUIRE(pos && size && ratio && field && pins && area && weight && precond && grad && (num_wl == 0 || wl_grad),
               "density_grad: null pointer");
    return dropin_grad(pos, size, ratio, field, wl_grad, pins, area, num_nodes, num_wl, weight, precond, bx, by, nbx,
                       nby, grad, (hipStream_t)stream);
}

}  // extern "C"
