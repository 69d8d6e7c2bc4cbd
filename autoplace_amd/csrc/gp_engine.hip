// The GP iteration engine (include/autoplace_gp.h, part 3): whole passes of src/gpu_loop.py:58-64 on the device.
//
// One pass = Nesterov step -> WA -> density scatter (force grid + overflow grid) -> post -> Poisson -> gather ->
// control. Differences from the reference's pass, all in the data movement:
//   * u / v / g are double-buffered; the control kernel flips a device flag instead of select_kernel's 3 copies,
//     and the best placement is copied by the next pass's Nesterov kernel (which reads v anyway);
//   * the map is kept unmirrored (the reference writes it 4x for the FFT), and cleared by the kernel that reads it;
//   * |v' - v| and |g' - g| are accumulated in FP64 by the kernels that write v' and g', the WA / HPWL totals by the
//     WA kernel and the overflow by the post kernel, as per-workgroup partials the control kernel sums in a fixed
//     order (deterministic);
//   * the per-cell WA gradient sum is fused into the gather.
// Every kernel reads its parameters (gamma, weight, step) from device memory, so a run of passes has no host sync
// and replays from a HIP graph.
#include <stdio.h>

#include "gp_common.h"
#include "gp_control.h"
#include "gp_kernels.h"
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
        int f = *flag;
        return (other ? !f : f) ? b : a;
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
__global__ __launch_bounds__(EB) void eng_nesterov(Bufs B, const float2* lo, const float2* hi, const float* kp, int n,
                                                   double* part) {
    __shared__ double red[16];
    const int cur = B.flags[0], nx = 1 - cur;
    const bool copy_best = B.flags[1] != 0;
    const float alpha = kp[3], coef = kp[4];
    const int i = blockIdx.x * EB + threadIdx.x;
    double m2[1] = {0.0};
    if (i < n) {
        const float2 u = B.u[cur][i], v = B.v[cur][i], g = B.g[cur][i], l = lo[i], h = hi[i];
        float2 un, vn;
        un.x = v.x - alpha * g.x;
        un.y = v.y - alpha * g.y;
        vn.x = fminf(fmaxf(un.x + coef * (un.x - u.x), l.x), h.x);
        vn.y = fminf(fmaxf(un.y + coef * (un.y - u.y), l.y), h.y);
        B.u[nx][i] = un;
        B.v[nx][i] = vn;
        if (copy_best) B.best[i] = v;
        const double dx = (double)vn.x - v.x, dy = (double)vn.y - v.y;
        m2[0] = dx * dx + dy * dy;
    }
    block_sum<1>(m2, red);
    if (threadIdx.x == 0) part[blockIdx.x] = m2[0];
}

// ------------------------------------------------------------------------------ density scatter, both grids
struct MapJob {
    const float2* size;
    const float* ratio;
    int num_nodes, num_mov;
    float inv_bx, inv_by;
    int nbx, nby;
    unsigned long long* map;   // fallback accumulation (global integer atomics), read and cleared by the merge
    int blocks;                // workgroups of this job
    int first_block;           // index of the job's first workgroup among all map workgroups
    unsigned long long* slots; // per workgroup: its LDS tile (MAP_LDS values), plain stores
    int4* boxes;               // per workgroup: x0, y0, w, h | plane << 20 (w = 0: nothing in a slot)
    int* fallback;             // set to 1 when a workgroup used the global atomics
};

// A workgroup takes MAP_NODES consecutive nodes (spatial neighbours after the engine's Morton renumbering). When
// their bins fit MAP_LDS and they share a plane, the contributions are summed with integer LDS atomics and each
// touched bin is flushed with one global atomic; otherwise every contribution goes to global memory directly.
// Integer sums either way: the map is the same whatever the path and the order.
constexpr int MAP_PER_LANE = 4;
constexpr int MAP_NODES = EB * MAP_PER_LANE;
constexpr int MAP_LDS = 16384;   // 128 KB: a workgroup's nodes where the density is low (fillers) still fit

__device__ __forceinline__ unsigned long long contribution(float wx, const Span& sy, int by) {
    return (unsigned long long)((double)(wx * span_overlap(sy, by)) * kDensityScale);
}

__global__ __launch_bounds__(EB) void eng_density_map(Pick pos_p, MapJob f, MapJob j) {
    extern __shared__ unsigned long long tile[];   // MAP_LDS entries (dynamic: above the static LDS limit)
    __shared__ int red[4][EB / 64];
    const bool judge = (int)blockIdx.x >= f.blocks;
    const MapJob& J = judge ? j : f;
    const int base = ((int)blockIdx.x - (judge ? f.blocks : 0)) * MAP_NODES;
    if (base >= J.num_nodes) {                                     // uniform over the workgroup
        if (threadIdx.x == 0) J.boxes[blockIdx.x] = make_int4(0, 0, 0, 0);
        return;
    }
    const float2* pos = pos_p.get(true);
    Span sx[MAP_PER_LANE], sy[MAP_PER_LANE];
    float r[MAP_PER_LANE];
    int x0 = 0x7fffffff, y0 = 0x7fffffff, x1 = -1, y1 = -1;
    // lane l of wave w takes nodes base + 16 l + 4 w + q: the lanes of one instruction are 16 nodes apart along the
    // curve, so their LDS atomics rarely hit the same bin (same-address lanes serialise)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int first = base + lane * (MAP_NODES / 64) + wv * MAP_PER_LANE;
#pragma unroll
    for (int q = 0; q < MAP_PER_LANE; q++) {
        const int i = first + q;
        r[q] = 0.f;
        sx[q].b0 = 1;
        sx[q].b1 = 0;                                             // empty unless a node is here
        sy[q].b0 = 1;
        sy[q].b1 = 0;
        if (i < J.num_nodes) {
            const float2 p = pos[i], sz = J.size[i];
            r[q] = J.ratio[i];
            sx[q] = make_span(p.x, sz.x, J.inv_bx, J.nbx);
            sy[q] = make_span(p.y, sz.y, J.inv_by, J.nby);
            if (sx[q].b0 <= sx[q].b1 && sy[q].b0 <= sy[q].b1) {
                x0 = min(x0, sx[q].b0); x1 = max(x1, sx[q].b1);
                y0 = min(y0, sy[q].b0); y1 = max(y1, sy[q].b1);
            }
        }
    }
    // bounding box of the workgroup's bins
    for (int o = 32; o > 0; o >>= 1) {
        x0 = min(x0, __shfl_xor(x0, o, 64)); y0 = min(y0, __shfl_xor(y0, o, 64));
        x1 = max(x1, __shfl_xor(x1, o, 64)); y1 = max(y1, __shfl_xor(y1, o, 64));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wave] = x0; red[1][wave] = y0; red[2][wave] = x1; red[3][wave] = y1;
    }
    __syncthreads();
    x0 = red[0][0]; y0 = red[1][0]; x1 = red[2][0]; y1 = red[3][0];
    for (int w = 1; w < EB / 64; w++) {
        x0 = min(x0, red[0][w]); y0 = min(y0, red[1][w]); x1 = max(x1, red[2][w]); y1 = max(y1, red[3][w]);
    }
    const int last = min(base + MAP_NODES, J.num_nodes) - 1;
    const bool one_plane = (base < J.num_mov) == (last < J.num_mov);
    const long long plane = base < J.num_mov ? 0 : (long long)J.nbx * J.nby;
    const int w = x1 - x0 + 1, h = y1 - y0 + 1;
    const int gb = (int)blockIdx.x;                                // slot / box of this workgroup
    if (x1 < x0) {                                                 // no bins at all (uniform)
        if (threadIdx.x == 0) J.boxes[gb] = make_int4(0, 0, 0, 0);
        return;
    }
    if (one_plane && (long long)w * h <= MAP_LDS) {
        for (int e = threadIdx.x; e < w * h; e += EB) tile[e] = 0ull;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < MAP_PER_LANE; q++)
            for (int bx = sx[q].b0; bx <= sx[q].b1; bx++) {
                const float wx = r[q] * span_overlap(sx[q], bx);
                unsigned long long* row = tile + (bx - x0) * h - y0;
                for (int by = sy[q].b0; by <= sy[q].b1; by++) {
                    const unsigned long long c = contribution(wx, sy[q], by);
                    if (c) atomicAdd(row + by, c);
                }
            }
        __syncthreads();
        // the tile goes to this workgroup's own slot with plain stores; the merge kernel sums the slots per bin
        unsigned long long* slot = J.slots + (long long)gb * MAP_LDS;
        for (int e = threadIdx.x; e < w * h; e += EB) slot[e] = tile[e];
        if (threadIdx.x == 0) J.boxes[gb] = make_int4(x0, y0, w, h | (plane ? (1 << 20) : 0));
        return;
    }
    if (threadIdx.x == 0) {
        J.boxes[gb] = make_int4(0, 0, 0, 0);
        *J.fallback = 1;
    }
#pragma unroll
    for (int q = 0; q < MAP_PER_LANE; q++) {
        const int i = first + q;
        unsigned long long* m = J.map + (i < J.num_mov ? 0 : (long long)J.nbx * J.nby);
        for (int bx = sx[q].b0; bx <= sx[q].b1; bx++) {
            const float wx = r[q] * span_overlap(sx[q], bx);
            unsigned long long* row = m + (long long)bx * J.nby;
            for (int by = sy[q].b0; by <= sy[q].b1; by++) {
                const unsigned long long c = contribution(wx, sy[q], by);
                if (c) atomicAdd(row + by, c);
            }
        }
    }
}

// maps -> unmirrored total density (force grid) and the overflow partial sums; the maps are cleared for the next pass
// Merge + post, one workgroup per 32 x 32 tile of bins: the tile's value is the sum of the slots of the map
// workgroups whose box covers it (found by scanning the boxes in workgroup order: a fixed order, and integer sums
// anyway) plus the fallback map when any workgroup used it; then density, overflow, and the map cleared.
constexpr int MT = 32;
struct PostJob {
    unsigned long long* map;
    int planes;                // 2: cells + fillers (force grid), 1: cells only
    const float* fixed;
    float* rho;                // null: overflow only
    int nbx, nby;
    int tiles_y;               // tiles per column of tiles
    int blocks;                // tiles of this grid
    int count_over;            // 1: this grid's overflow feeds the control
    const unsigned long long* slots;
    const int4* boxes;
    int first_box, num_boxes;  // the map workgroups of this grid
    const int* fallback;
};

__global__ __launch_bounds__(EB) void eng_density_merge(PostJob f, PostJob j, float target, double* part) {
    __shared__ double red[16];
    __shared__ int list[EB + 1];
    __shared__ int nlist;
    __shared__ int wcount[EB / 64];
    const bool judge = (int)blockIdx.x >= f.blocks;
    const PostJob& J = judge ? j : f;
    const int tile = (int)blockIdx.x - (judge ? f.blocks : 0);
    const int X0 = (tile / J.tiles_y) * MT, Y0 = (tile % J.tiles_y) * MT;
    const long long nb = (long long)J.nbx * J.nby;
    const bool fb = *J.fallback != 0;
    // each lane owns 4 bins of the tile: local e = lane + 256 k -> (e / 32, e % 32)
    unsigned long long acc0[4], acc1[4];
    long long bin[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int e = threadIdx.x + EB * k, bx = X0 + e / MT, by = Y0 + e % MT;
        bin[k] = (bx < J.nbx && by < J.nby) ? (long long)bx * J.nby + by : -1;
        acc0[k] = acc1[k] = 0ull;
        if (fb && bin[k] >= 0) {
            acc0[k] = J.map[bin[k]];
            J.map[bin[k]] = 0ull;
            if (J.planes == 2) {
                acc1[k] = J.map[nb + bin[k]];
                J.map[nb + bin[k]] = 0ull;
            }
        }
    }
    // boxes overlapping the tile, in workgroup order, EB at a time
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c0 = 0; c0 < J.num_boxes; c0 += EB) {
        const int wgi = c0 + threadIdx.x;
        bool hit = false;
        int4 bx4 = make_int4(0, 0, 0, 0);
        if (wgi < J.num_boxes) {
            bx4 = J.boxes[J.first_box + wgi];
            const int h = bx4.w & 0xFFFFF;
            hit = bx4.z > 0 && bx4.x < X0 + MT && bx4.x + bx4.z > X0 && bx4.y < Y0 + MT && bx4.y + h > Y0;
        }
        const unsigned long long m = __ballot(hit);
        if (lane == 0) wcount[wave] = __popcll(m);
        __syncthreads();
        int before = 0, total = 0;
        for (int w = 0; w < EB / 64; w++) {
            if (w < wave) before += wcount[w];
            total += wcount[w];
        }
        if (hit) list[before + __popcll(m & ((1ull << lane) - 1))] = J.first_box + wgi;
        if (threadIdx.x == 0) nlist = total;
        __syncthreads();
        for (int q = 0; q < nlist; q++) {
            const int g = list[q];
            const int4 b4 = J.boxes[g];
            const int h = b4.w & 0xFFFFF, plane = b4.w >> 20;
            const unsigned long long* slot = J.slots + (long long)g * MAP_LDS;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (bin[k] < 0) continue;
                const int e = threadIdx.x + EB * k, bx = X0 + e / MT, by = Y0 + e % MT;
                if (bx >= b4.x && bx < b4.x + b4.z && by >= b4.y && by < b4.y + h) {
                    const unsigned long long v = slot[(bx - b4.x) * h + (by - b4.y)];
                    if (plane) acc1[k] += v;
                    else acc0[k] += v;
                }
            }
        }
        __syncthreads();
    }
    double o[1] = {0.0};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (bin[k] < 0) continue;
        const float mov = (float)((double)acc0[k] / kDensityScale) + J.fixed[bin[k]];
        const float r = J.planes == 2 ? mov + (float)((double)acc1[k] / kDensityScale) : mov;
        if (J.rho) J.rho[bin[k]] = r;
        if (J.count_over) o[0] += (double)fmaxf(mov - target, 0.f);
    }
    block_sum<1>(o, red);
    if (threadIdx.x == 0) part[blockIdx.x] = o[0];
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
    int ld;
    const float2* pin_grad;
    const int* cell_start;
    const int* cell_pin;
    const float* pins;
    const float* area;
    const float* kp;          // weight kp[1], precond kp[2]
    int num_nodes, num_mov;
    float inv_bx, inv_by;
    int nbx, nby;
    int use_density;          // 0: grad = wl / max(pins, 1) (the weight-0 objective)
};

__global__ __launch_bounds__(EB) void eng_gather(GatherArgs a, double* part) {
    __shared__ double red[16];
    const int i = blockIdx.x * EB + threadIdx.x;
    double t2[1] = {0.0};
    if (i < a.num_nodes) {
        int cur = a.flags ? a.flags[0] : 0;
        const float2* pos = a.pos.get(true);
        float gx = 0, gy = 0;
        if (a.use_density) {
            const float weight = a.kp[1];
            const float2 p = pos[i], sz = a.size[i];
            const float r = a.ratio[i];
            Span sx = make_span(p.x, sz.x, a.inv_bx, a.nbx), sy = make_span(p.y, sz.y, a.inv_by, a.nby);
            for (int bx = sx.b0; bx <= sx.b1; bx++) {
                const float wx = r * span_overlap(sx, bx);
                const float* fxr = a.fx + (long long)bx * a.ld;
                const float* fyr = a.fy + (long long)bx * a.ld;
                for (int by = sy.b0; by <= sy.b1; by++) {
                    const float w = wx * span_overlap(sy, by);
                    gx -= w * fxr[by];
                    gy -= w * fyr[by];
                }
            }
            gx *= weight;
            gy *= weight;
        }
        if (i < a.num_mov) {
            float wx = 0, wy = 0;
            // the WA kernel wrote the pin gradients in cell order: a cell's pins are contiguous
            for (int k = a.cell_start[i], e = a.cell_start[i + 1]; k < e; k++) {
                const float2 q = a.pin_grad[k];
                wx += q.x;
                wy += q.y;
            }
            gx += wx;
            gy += wy;
        }
        float scale;
        if (a.use_density) {
            const float precond = a.kp[2];
            scale = precond < 0 ? 1.f : 1.f / fmaxf(a.pins[i] + precond * a.area[i], 1.f);
        } else {
            scale = 1.f / fmaxf(a.pins[i], 1.f);
        }
        const float2 gn = make_float2(gx * scale, gy * scale);
        float2* out = (a.flags && cur == 0) ? a.g_out_b : a.g_out_a;
        out[i] = gn;
        if (a.g_prev) {
            const float2 gp = (a.flags && cur == 1) ? a.g_prev_b[i] : a.g_prev[i];
            const double dx = (double)gn.x - gp.x, dy = (double)gn.y - gp.y;
            t2[0] = dx * dx + dy * dy;
        }
    }
    block_sum<1>(t2, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t2[0];
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
    double over_scale;       // bin_area / mov_area of the grid the overflow is measured on
};

// First stage of the reduction of the partials: workgroup b sums the b-th fixed slice of every segment
// (moved^2, turned^2, WA, HPWL, overflow) -> out[b][5]. The slices and the order inside them do not depend on the
// launch, so the sums are deterministic.
constexpr int RED_BLOCKS = 64;

__device__ __forceinline__ void slice(int n, int& lo, int& hi) {
    const int per = (n + RED_BLOCKS - 1) / RED_BLOCKS;
    lo = min(n, (int)blockIdx.x * per);
    hi = min(n, lo + per);
}

__global__ __launch_bounds__(EB) void eng_reduce(Segs s, double* out) {
    __shared__ double red[16 * 5];
    double v[5] = {0, 0, 0, 0, 0};
    int lo, hi;
    slice(s.n_moved, lo, hi);
    for (int k = lo + threadIdx.x; k < hi; k += EB) v[0] += s.moved[k];
    slice(s.n_turned, lo, hi);
    for (int k = lo + threadIdx.x; k < hi; k += EB) v[1] += s.turned[k];
    slice(s.n_wa, lo, hi);
    for (int k = lo + threadIdx.x; k < hi; k += EB) {
        v[2] += s.wa[2 * k];
        v[3] += s.wa[2 * k + 1];
    }
    slice(s.n_over, lo, hi);
    for (int k = lo + threadIdx.x; k < hi; k += EB) v[4] += s.over[k];
    block_sum<5>(v, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < 5; k++) out[blockIdx.x * 5 + k] = v[k];
}

__global__ __launch_bounds__(64) void eng_control(const double* red5, double over_scale, double* S, float* kp,
                                                  const double* c, int* flags, double* stats, int* fallback) {
    if (threadIdx.x == 0) {
        *fallback = 0;             // the merge of this pass has read it
        double r[5] = {0, 0, 0, 0, 0};
        for (int b = 0; b < RED_BLOCKS; b++)
            for (int k = 0; k < 5; k++) r[k] += red5[b * 5 + k];
        const Segs s{nullptr, 0, nullptr, 0, nullptr, 0, nullptr, 0, over_scale};
        const double moved = sqrt(r[0]), turned = sqrt(r[1]), wa = r[2], hpwl = r[3], ovf = r[4] * s.over_scale;
        if (stats) {
            stats[0] = wa; stats[1] = hpwl; stats[2] = ovf; stats[3] = moved; stats[4] = turned;
        }
        if (S) {
            control_step(S, kp, c, moved, turned, hpwl, ovf, wa);
            if (S[ACCEPT] != 0) flags[0] = 1 - flags[0];
            flags[1] = S[IS_BEST] != 0 ? 1 : 0;
        }
    }
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

struct Level {
    ap_level_desc d{};
    Poisson poisson;
    WaPlan wa_plan;
    float2 *u1 = nullptr, *v1 = nullptr, *g1 = nullptr, *pin_grad = nullptr;
    int* block_net = nullptr;  // WA workgroup -> first owned net
    int* pin_slot = nullptr;   // pin -> position in the cell CSR (-1: fixed pin)
    unsigned long long *map = nullptr, *jmap = nullptr;
    unsigned long long* slots = nullptr;   // per map workgroup: its LDS tile
    int4* boxes = nullptr;                 // per map workgroup: the tile's box
    int* fallback = nullptr;               // a map workgroup used the global atomics in this pass
    float *rho = nullptr, *fx = nullptr, *fy = nullptr;
    double* part = nullptr;
    double* stats = nullptr;
    double* red5 = nullptr;    // first-stage sums, RED_BLOCKS x 5
    int* flags = nullptr;
    int nblk_n = 0, nblk_wa = 0, nblk_fpost = 0, nblk_jpost = 0, nblk_fmap = 0, nblk_jmap = 0;
    double *p_moved = nullptr, *p_wa = nullptr, *p_over = nullptr, *p_turned = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr, t0 = nullptr, t1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_passes = 0;
    // profiling: one event before each kernel class and one at the end of the pass
    bool prof = false;
    hipEvent_t pev[AP_PROF_CLASSES + 1] = {};
    double prof_ms[AP_PROF_CLASSES] = {};
    void mark(int k, hipStream_t s) {
        if (prof) (void)hipEventRecord(pev[k], s);
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

    int create(const ap_level_desc& desc);
    void destroy();
    int objective(const Pick& pos, const float* kp_ptr, bool with_state, const float2* g_prev, const float2* g_prev_b,
                  float2* g_out_a, float2* g_out_b, hipStream_t s, double* S, float* kp_out, int use_density);
    int pass(hipStream_t s);
    int finalize(hipStream_t s);
};

int Level::create(const ap_level_desc& desc) {
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
    int st = poisson.create(d.nbx, d.nby, d.bx, d.by, d.poisson_method);
    if (st) return st;
    const long long N = d.num_nodes, nb = (long long)d.nbx * d.nby, nj = d.has_judge ? (long long)d.njx * d.njy : 0;
    AP_HIP(hipMalloc(&u1, N * sizeof(float2)));
    AP_HIP(hipMalloc(&v1, N * sizeof(float2)));
    AP_HIP(hipMalloc(&g1, N * sizeof(float2)));
    AP_HIP(hipMalloc(&pin_grad, (d.num_pins + 1) * sizeof(float2)));
    AP_HIP(hipMalloc(&pin_slot, (d.num_pins + 1) * sizeof(int)));
    if (d.num_pins > 0) {
        hipLaunchKernelGGL(eng_pin_slot_clear, dim3(grid_of(d.num_pins, EB)), dim3(EB), 0, nullptr, pin_slot,
                           d.num_pins);
        AP_LAUNCHED();
        hipLaunchKernelGGL(eng_pin_slot_fill, dim3(grid_of(d.num_pins, EB)), dim3(EB), 0, nullptr, pin_slot,
                           d.cell_start, d.cell_pin, d.num_mov, d.num_pins);
        AP_LAUNCHED();
    }
    AP_HIP(hipMalloc(&map, 2 * nb * sizeof(unsigned long long)));
    AP_HIP(hipMemset(map, 0, 2 * nb * sizeof(unsigned long long)));
    if (nj) {
        AP_HIP(hipMalloc(&jmap, nj * sizeof(unsigned long long)));
        AP_HIP(hipMemset(jmap, 0, nj * sizeof(unsigned long long)));
    }
    AP_HIP(hipMalloc(&rho, nb * sizeof(float)));
    AP_HIP(hipMalloc(&fx, nb * sizeof(float)));
    AP_HIP(hipMalloc(&fy, nb * sizeof(float)));
    nblk_n = grid_of(N, EB);
    if ((st = wa_plan.create(d.net_start, d.num_nets))) return st;
    AP_HIP(hipFuncSetAttribute((const void*)eng_density_map, hipFuncAttributeMaxDynamicSharedMemorySize,
                               MAP_LDS * sizeof(unsigned long long)));
    nblk_wa = wa_plan.partial_count();
    nblk_fmap = grid_of(N, MAP_NODES);
    nblk_jmap = nj ? grid_of(d.num_mov, MAP_NODES) : 0;
    nblk_fpost = grid_of(d.nbx, MT) * grid_of(d.nby, MT);
    nblk_jpost = nj ? grid_of(d.njx, MT) * grid_of(d.njy, MT) : 0;
    AP_HIP(hipMalloc(&slots, (long long)(nblk_fmap + nblk_jmap) * MAP_LDS * sizeof(unsigned long long)));
    AP_HIP(hipMalloc(&boxes, (nblk_fmap + nblk_jmap + 1) * sizeof(int4)));
    AP_HIP(hipMalloc(&fallback, sizeof(int)));
    AP_HIP(hipMemset(fallback, 0, sizeof(int)));
    const long long np = nblk_n + 2LL * nblk_wa + nblk_fpost + nblk_jpost + nblk_n;
    AP_HIP(hipMalloc(&part, np * sizeof(double)));
    AP_HIP(hipMemset(part, 0, np * sizeof(double)));
    p_moved = part;
    p_wa = p_moved + nblk_n;
    p_over = p_wa + 2LL * nblk_wa;
    p_turned = p_over + nblk_fpost + nblk_jpost;
    AP_HIP(hipMalloc(&stats, 8 * sizeof(double)));
    AP_HIP(hipMalloc(&red5, RED_BLOCKS * 5 * sizeof(double)));
    AP_HIP(hipMalloc(&flags, 4 * sizeof(int)));
    AP_HIP(hipMemset(flags, 0, 4 * sizeof(int)));
    AP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    AP_HIP(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
    AP_HIP(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
    AP_HIP(hipEventCreate(&t0));
    AP_HIP(hipEventCreate(&t1));
    AP_HIP(hipDeviceSynchronize());
    return AP_OK;
}

void Level::destroy() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t& ev : pev)
        if (ev) (void)hipEventDestroy(ev);
    void* ptrs[] = {u1,    v1,    g1,       pin_grad, map,      jmap,  rho,   fx, fy, part, stats, flags, red5,
                    block_net, pin_slot, slots, boxes, fallback};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    poisson.destroy();
    wa_plan.destroy();
    if (stream) (void)hipStreamDestroy(stream);
    hipEvent_t evs[] = {ev_in, ev_out, t0, t1};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
}

// objective at pos with the parameters kp; g_prev (optional) gives |g' - g|. With S the control pass follows.
int Level::objective(const Pick& pos, const float* kp_ptr, bool with_state, const float2* g_prev,
                     const float2* g_prev_b, float2* g_out_a, float2* g_out_b, hipStream_t s, double* S, float* kp_out,
                     int use_density) {
    int st;
    // WA
    {
        // the evaluation point: set 1 when the current point is set 0, and the reverse (resolved on the device)
        WaArgs w{pos.a, d.pin_cell, (const float2*)d.pin_off, d.net_start, d.num_nets, d.num_mov, kp_ptr,
                 pin_grad, nullptr, p_wa, pos.b, pos.flag, block_net, pin_slot};
        mark(AP_PROF_WA, s);
        if ((st = wa_plan.launch(w, p_wa, s))) return st;
    }
    int n_over = 0;
    double over_scale = 0;
    if (use_density) {
        // the same float32 reciprocal as the reference's host code: 1.f / (float)bx
        MapJob f{(const float2*)d.size, d.ratio, d.num_nodes, d.num_mov, 1.f / (float)d.bx, 1.f / (float)d.by,
                 d.nbx, d.nby, map, nblk_fmap, 0, slots, boxes, fallback};
        MapJob j{(const float2*)d.judge_size, d.judge_ratio, d.num_mov, d.num_mov, 0.f, 0.f, d.njx, d.njy, jmap,
                 nblk_jmap, nblk_fmap, slots, boxes, fallback};
        if (d.has_judge) {
            j.inv_bx = 1.f / (float)d.jbx;
            j.inv_by = 1.f / (float)d.jby;
        }
        mark(AP_PROF_MAP, s);
        hipLaunchKernelGGL(eng_density_map, dim3(nblk_fmap + nblk_jmap), dim3(EB),
                           MAP_LDS * sizeof(unsigned long long), s, pos, f, j);
        AP_LAUNCHED();
        PostJob pf{map, 2, d.fixed, rho, d.nbx, d.nby, (int)grid_of(d.nby, MT), nblk_fpost, d.has_judge ? 0 : 1,
                   slots, boxes, 0, nblk_fmap, fallback};
        PostJob pj{jmap, 1, d.judge_fixed, nullptr, d.njx, d.njy, d.has_judge ? (int)grid_of(d.njy, MT) : 1,
                   nblk_jpost, 1, slots, boxes, nblk_fmap, nblk_jmap, fallback};
        mark(AP_PROF_POST, s);
        hipLaunchKernelGGL(eng_density_merge, dim3(nblk_fpost + nblk_jpost), dim3(EB), 0, s, pf, pj, d.target,
                           p_over);
        AP_LAUNCHED();
        n_over = nblk_fpost + nblk_jpost;
        over_scale = d.has_judge ? d.judge_bin_area / d.mov_area : d.bin_area / d.mov_area;
        mark(AP_PROF_POISSON, s);
        if ((st = poisson.solve(rho, d.nby, fx, fy, d.nby, s))) return st;
    }
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
    mark(AP_PROF_GATHER, s);
    hipLaunchKernelGGL(eng_gather, dim3(nblk_n), dim3(EB), 0, s, g, p_turned);
    AP_LAUNCHED();
    mark(AP_PROF_CONTROL, s);
    Segs sg{p_moved, with_state ? nblk_n : 0, p_wa, nblk_wa, p_over, n_over, p_turned, g_prev ? nblk_n : 0,
            over_scale};
    hipLaunchKernelGGL(eng_reduce, dim3(RED_BLOCKS), dim3(EB), 0, s, sg, red5);
    AP_LAUNCHED();
    hipLaunchKernelGGL(eng_control, dim3(1), dim3(64), 0, s, red5, over_scale, S, kp_out, d.c, flags, stats,
                       fallback);
    AP_LAUNCHED();
    mark(AP_PROF_CLASSES, s);
    return AP_OK;
}

int Level::pass(hipStream_t s) {
    Bufs B = bufs();
    mark(AP_PROF_NESTEROV, s);
    hipLaunchKernelGGL(eng_nesterov, dim3(nblk_n), dim3(EB), 0, s, B, (const float2*)d.lo, (const float2*)d.hi, d.kp,
                       d.num_nodes, p_moved);
    AP_LAUNCHED();
    Pick pos{(const float2*)d.v, (const float2*)v1, flags};
    return objective(pos, d.kp, true, (const float2*)d.g, g1, (float2*)d.g, g1, s, d.S, d.kp, 1);
}

int Level::finalize(hipStream_t s) {
    hipLaunchKernelGGL(eng_finalize_copy, dim3(nblk_n), dim3(EB), 0, s, bufs(), d.num_nodes);
    AP_LAUNCHED();
    hipLaunchKernelGGL(eng_finalize_flags, dim3(1), dim3(64), 0, s, flags);
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

int ap_level_run(ap_level* level, int passes, int use_graph, void* stream) {
    AP_REQUIRE(level && passes >= 0, "level_run: bad arguments");
    Level& L = level->L;
    hipStream_t caller = (hipStream_t)stream;
    AP_HIP(hipEventRecord(L.ev_in, caller));
    AP_HIP(hipStreamWaitEvent(L.stream, L.ev_in, 0));
    AP_HIP(hipEventRecord(L.t0, L.stream));
    int st = AP_OK;
    if (use_graph) {
        if (!L.exec || L.graph_passes != passes) {
            if (L.exec) {
                (void)hipGraphExecDestroy(L.exec);
                L.exec = nullptr;
            }
            if (L.graph) {
                (void)hipGraphDestroy(L.graph);
                L.graph = nullptr;
            }
            AP_HIP(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < passes && !st; k++) st = L.pass(L.stream);
            if (!st) st = L.finalize(L.stream);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(L.stream, &g);
            if (st) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            AP_HIP(e);
            L.graph = g;
            AP_HIP(hipGraphInstantiate(&L.exec, L.graph, nullptr, nullptr, 0));
            L.graph_passes = passes;
        }
        AP_HIP(hipGraphLaunch(L.exec, L.stream));
    } else {
        for (int k = 0; k < passes && !st; k++) st = L.pass(L.stream);
        if (!st) st = L.finalize(L.stream);
        if (st) return st;
    }
    AP_HIP(hipEventRecord(L.t1, L.stream));
    AP_HIP(hipEventRecord(L.ev_out, L.stream));
    AP_HIP(hipStreamWaitEvent(caller, L.ev_out, 0));
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
    L.prof = true;
    int st = AP_OK;
    for (int p = 0; p < passes && !st; p++) {
        st = L.pass(L.stream);
        if (st) break;
        if (hipEventSynchronize(L.pev[AP_PROF_CLASSES]) != hipSuccess) {
            st = AP_ERR_HIP;
            break;
        }
        for (int k = 0; k < AP_PROF_CLASSES; k++) {
            float t = 0;
            int nxt = k + 1;
            if (k == AP_PROF_POISSON) nxt = AP_PROF_GATHER;
            (void)hipEventElapsedTime(&t, L.pev[k], L.pev[nxt]);
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
    int st = L.objective(p, L.d.kp, false, nullptr, nullptr, (float2*)grad, (float2*)grad, s, nullptr, nullptr, 1);
    if (st) return st;
    AP_HIP(hipMemcpyAsync(stats, L.stats, 5 * sizeof(double), hipMemcpyDeviceToDevice, s));
    return AP_OK;
}

}  // extern "C"
