// Clustering for the two-level placement on the device (src/cluster.py:18-85 of the reference), C ABI part 4.
//
// Cells are merged by rounds of star matching on the weighted cell graph of the small nets:
//   edges   every net of 2 .. max_degree pins gives each ordered pair of its pins on two different movable cells
//           the integer weight 420 / (degree - 1) (degree counts every pin of the net). The pairs are sorted by
//           (cell, cell) key and equal keys summed (rocPRIM radix sort + reduce by key): a row-sorted edge CSR.
//   round   with the heads drawn by the caller (the reference draws them from torch's CPU generator, so the caller
//           does), every non-head cluster below the size limit picks the head neighbour of the highest weight per
//           merged area (ties: the lowest head index); every head takes at most `per_head` of the clusters that
//           picked it, highest score first (ties: the lowest tail index). Taken clusters join their head; the kept
//           clusters are renumbered in index order, and the cells, areas and edges follow (the edges re-sorted and
//           merged, self-loops dropped).
// Weights and areas are integers and the scores one IEEE division each, so the clusters are those of the
// reference's torch operators on any device, bit for bit. Each step is a thread per cluster or per edge; the
// sorts and scans are rocPRIM's. The counts of a round (clusters and edges after it) stay on the device, read by the
// kernels that follow; the host reads them once per round, at its end, to decide on the next.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>
#include <rocprim/device/device_scan.hpp>

#include "gp_common.h"

namespace ap {

constexpr int CLU_BLOCK = 256;
constexpr int CLU_MAX_DEGREE = 64;

// ordered pin pairs (p != q) of a net whose cells are movable and different
__device__ __forceinline__ int clu_net_pairs(const int* pin_cell, int p0, int p1, int n) {
    int c = 0;
    for (int p = p0; p < p1; p++) {
        const int a = pin_cell[p];
        if (a >= n) continue;
        for (int q = p0; q < p1; q++) {
            const int b = pin_cell[q];
            c += (q != p && b < n && b != a);
        }
    }
    return c;
}

__global__ __launch_bounds__(CLU_BLOCK) void clu_count_pairs(const int* net_start, int num_nets, const int* pin_cell,
                                                             int n, int max_degree, unsigned* count) {
    const int net = blockIdx.x * CLU_BLOCK + threadIdx.x;
    if (net >= num_nets) return;
    const int p0 = net_start[net], p1 = net_start[net + 1];
    const int d = p1 - p0;
    count[net] = (d >= 2 && d <= max_degree) ? (unsigned)clu_net_pairs(pin_cell, p0, p1, n) : 0u;
}

__global__ __launch_bounds__(CLU_BLOCK) void clu_emit_pairs(const int* net_start, int num_nets, const int* pin_cell,
                                                            int n, int max_degree, const unsigned* offset,
                                                            unsigned long long* key, int* weight) {
    const int net = blockIdx.x * CLU_BLOCK + threadIdx.x;
    if (net >= num_nets) return;
    const int p0 = net_start[net], p1 = net_start[net + 1];
    const int d = p1 - p0;
    if (d < 2 || d > max_degree) return;
    const int w = 420 / (d - 1);
    unsigned o = offset[net];
    for (int p = p0; p < p1; p++) {
        const int a = pin_cell[p];
        if (a >= n) continue;
        for (int q = p0; q < p1; q++) {
            const int b = pin_cell[q];
            if (q != p && b < n && b != a) {
                key[o] = (unsigned long long)a * (unsigned long long)n + (unsigned long long)b;
                weight[o] = w;
                o++;
            }
        }
    }
}

// count[0]: edges, count[1]: clusters, on the device. The merged keys end with the self-loops' key num^2 when a
// round made self-loops: that entry is not an edge.
__global__ void clu_drop_self_loops(const unsigned long long* key, unsigned* count) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const unsigned c = count[0];
        const unsigned long long num = count[1];
        if (c > 0 && key[c - 1] == num * num) count[0] = c - 1;
    }
}

// edges from the merged keys: row (cluster), column and weight of every edge
__global__ __launch_bounds__(CLU_BLOCK) void clu_split_keys(const unsigned long long* key, const int* weight,
                                                            const unsigned* count, int* ei, int* ej, int* ew) {
    const unsigned e = blockIdx.x * CLU_BLOCK + threadIdx.x;
    if (e >= count[0]) return;
    const unsigned long long num = count[1];
    const unsigned long long k = key[e];
    ei[e] = (int)(k / num);
    ej[e] = (int)(k % num);
    ew[e] = weight[e];
}

// row_start[r] = first edge of row r: a binary search of the sorted keys for r * num
__global__ __launch_bounds__(CLU_BLOCK) void clu_row_starts(const unsigned long long* key, const unsigned* count,
                                                            int* row_start) {
    const int r = blockIdx.x * CLU_BLOCK + threadIdx.x;
    const int num = (int)count[1];
    if (r > num) return;
    const unsigned long long target = (unsigned long long)r * (unsigned long long)num;
    unsigned lo = 0, hi = count[0];
    while (lo < hi) {
        const unsigned mid = (lo + hi) >> 1;
        if (key[mid] < target) lo = mid + 1; else hi = mid;
    }
    row_start[r] = (int)lo;
}

// every non-head cluster below `big`: its best head neighbour by weight / merged area (first of equal scores, i.e.
// the lowest head index, as the row's columns ascend); -1 when none qualifies
__global__ __launch_bounds__(CLU_BLOCK) void clu_best_head(const int* row_start, const int* ej, const int* ew, int num,
                                                           const unsigned char* head, const long long* area,
                                                           long long big, float cap, int* best_to, double* best_score) {
    const int t = blockIdx.x * CLU_BLOCK + threadIdx.x;
    if (t >= num) return;
    int to = -1;
    double score = 0.0;
    const long long at = area[t];
    if (!head[t] && at < big) {
        for (int e = row_start[t], e1 = row_start[t + 1]; e < e1; e++) {
            const int j = ej[e];
            const long long aj = area[j];
            if (!head[j] || aj >= big) continue;
            const long long total = at + aj;
            if (!((float)total <= cap)) continue;
            const double s = (double)ew[e] / (double)total;
            if (to < 0 || s > score) {
                to = j;
                score = s;
            }
        }
    }
    best_to[t] = to;
    best_score[t] = score;
}

__global__ __launch_bounds__(CLU_BLOCK) void clu_identity(int* parent, int num) {
    const int c = blockIdx.x * CLU_BLOCK + threadIdx.x;
    if (c < num) parent[c] = c;
}

// the cells' first clusters (one per cell), and count[1] = the clusters
__global__ __launch_bounds__(CLU_BLOCK) void clu_start(int* cid, int n, unsigned* count) {
    const int c = blockIdx.x * CLU_BLOCK + threadIdx.x;
    if (c < n) cid[c] = c;
    if (c == 0) count[1] = (unsigned)n;
}

// every head takes the PER_HEAD best of the clusters that picked it: score descending, then the lowest tail index
// (the row's columns ascend, and an entry only passes the ones strictly below it)
template <int PER_HEAD>
__global__ __launch_bounds__(CLU_BLOCK) void clu_take(const int* row_start, const int* ej, int num,
                                                      const unsigned char* head, const int* best_to,
                                                      const double* best_score, int* parent) {
    const int h = blockIdx.x * CLU_BLOCK + threadIdx.x;
    if (h >= num || !head[h]) return;
    double ss[PER_HEAD];
    int ts[PER_HEAD];
#pragma unroll
    for (int k = 0; k < PER_HEAD; k++) {
        ss[k] = -1.0;                // scores are positive: -1 marks a free slot
        ts[k] = -1;
    }
    for (int e = row_start[h], e1 = row_start[h + 1]; e < e1; e++) {
        const int t = ej[e];
        if (best_to[t] != h) continue;
        const double s = best_score[t];
        if (!(s > ss[PER_HEAD - 1])) continue;
        // insert below every entry >= s: slot k keeps its entry, takes the new one, or takes slot k - 1's
#pragma unroll
        for (int k = PER_HEAD - 1; k >= 0; k--) {
            if (ss[k] >= s) continue;
            const bool here = (k == 0) || (ss[k - 1] >= s);
            ss[k] = here ? s : ss[k > 0 ? k - 1 : 0];
            ts[k] = here ? t : ts[k > 0 ? k - 1 : 0];
        }
    }
#pragma unroll
    for (int k = 0; k < PER_HEAD; k++)
        if (ts[k] >= 0) parent[ts[k]] = h;
}

// the kept clusters (their own parents), and the merged areas' sums cleared (no more clusters after the round)
__global__ __launch_bounds__(CLU_BLOCK) void clu_kept(const int* parent, int num, int* kept, long long* area_out) {
    const int c = blockIdx.x * CLU_BLOCK + threadIdx.x;
    if (c < num) {
        kept[c] = parent[c] == c;
        area_out[c] = 0;
    }
}

// count[1] = the clusters after the round: the rank of the last cluster plus its own flag
__global__ void clu_count_kept(const int* rank, const int* kept, int num, unsigned* count) {
    if (blockIdx.x == 0 && threadIdx.x == 0) count[1] = (unsigned)(rank[num - 1] + kept[num - 1]);
}

// new index of every old cluster: the rank of its parent among the kept ones
__global__ __launch_bounds__(CLU_BLOCK) void clu_new_index(const int* parent, const int* rank, int num, int* new_id) {
    const int c = blockIdx.x * CLU_BLOCK + threadIdx.x;
    if (c < num) new_id[c] = rank[parent[c]];
}

__global__ __launch_bounds__(CLU_BLOCK) void clu_relabel_cells(int* cid, int n, const int* new_id) {
    const int x = blockIdx.x * CLU_BLOCK + threadIdx.x;
    if (x < n) cid[x] = new_id[cid[x]];
}

// integer sums: the same whatever the order of the atomics
__global__ __launch_bounds__(CLU_BLOCK) void clu_merge_area(const long long* area, int num, const int* new_id,
                                                            long long* area_out) {
    const int c = blockIdx.x * CLU_BLOCK + threadIdx.x;
    if (c < num)
        atomicAdd((unsigned long long*)&area_out[new_id[c]], (unsigned long long)area[c]);
}

// the edges between new clusters (count[1] of them); a self-loop gets the key one past every valid one, so it sorts
// last
__global__ __launch_bounds__(CLU_BLOCK) void clu_relabel_edges(const int* ei, const int* ej, const int* ew,
                                                               const unsigned* count, const int* new_id,
                                                               unsigned long long* key, int* weight) {
    const unsigned e = blockIdx.x * CLU_BLOCK + threadIdx.x;
    if (e >= count[0]) return;
    const unsigned long long num_new = count[1];
    const unsigned long long a = (unsigned long long)new_id[ei[e]], b = (unsigned long long)new_id[ej[e]];
    key[e] = a == b ? num_new * num_new : a * num_new + b;
    weight[e] = ew[e];
}

__global__ __launch_bounds__(CLU_BLOCK) void clu_copy_out(const int* cid, int n, const long long* area, int num,
                                                          long long* cid_out, long long* area_out) {
    const int i = blockIdx.x * CLU_BLOCK + threadIdx.x;
    if (i < n) cid_out[i] = cid[i];
    if (i < num) area_out[i] = area[i];
}

inline unsigned bits_for(unsigned long long v) {   // bits that hold every value 0 .. v
    unsigned b = 1;
    while (b < 64 && (v >> b)) b++;
    return b;
}

// The heads' draws: torch.rand(num, generator=...) < 0.5 of the reference (cluster.py:65). torch's CPU generator is
// MT19937 (Matsumoto and Nishimura, 1998) seeded by its standard initialisation, and a float32 draw keeps the low
// 24 bits of one 32-bit output as k / 2^24, so a draw is below 0.5 exactly when bit 23 of the output is clear. One
// stream serves every round in order. The flags are made ahead: a round needs as many as there are clusters before
// it, at most the count of the round before, so while the device runs a round the host draws that many for the next
// (the surplus carries over to the round after).
struct HeadStream {
    uint32_t mt[624];
    int idx = 624;
    std::vector<unsigned char> ahead;   // drawn flags not yet handed to a round: ahead[used, size)
    size_t used = 0;
    unsigned char* pinned = nullptr;    // staging for the copy to the device (num_cells bytes)
    unsigned char* dev = nullptr;       // the round's heads on the device
    bool seeded = false;

    void seed(uint32_t s) {
        mt[0] = s;
        for (uint32_t i = 1; i < 624; i++) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i;
        idx = 624;
        ahead.clear();
        used = 0;
        seeded = true;
    }
    static uint32_t mix(uint32_t a, uint32_t b, uint32_t c) {     // word i from words i, i + 1 and i + 397
        const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
        return c ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
    }
    void twist() {       // the next 624 words, in three loops without wrap-around indices (each vectorises)
        int i = 0;
        for (; i < 624 - 397; i++) mt[i] = mix(mt[i], mt[i + 1], mt[i + 397]);
        for (; i < 623; i++) mt[i] = mix(mt[i], mt[i + 1], mt[i + 397 - 624]);
        mt[623] = mix(mt[623], mt[0], mt[396]);
        idx = 0;
    }
    // `count` more flags (1: head) at the end of `ahead`
    void draw(size_t count) {
        const size_t at = ahead.size();
        ahead.resize(at + count);
        unsigned char* out = ahead.data() + at;
        size_t k = 0;
        while (k < count) {
            if (idx >= 624) twist();
            const size_t m = std::min<size_t>(624 - idx, count - k);
            for (size_t j = 0; j < m; j++) {
                uint32_t y = mt[idx + j];
                y ^= y >> 11;
                y ^= (y << 7) & 0x9d2c5680u;
                y ^= (y << 15) & 0xefc60000u;
                y ^= y >> 18;
                out[k + j] = (unsigned char)(((y >> 23) & 1u) ^ 1u);
            }
            idx += (int)m;
            k += m;
        }
    }
    // hand `count` flags to a round: drawn now if not ahead, staged and copied to `dev` on the stream
    int take(size_t count, hipStream_t s) {
        if (ahead.size() - used < count) draw(count - (ahead.size() - used));
        std::memcpy(pinned, ahead.data() + used, count);
        used += count;
        if (used == ahead.size()) {
            ahead.clear();
            used = 0;
        } else if (used > (ahead.size() >> 1)) {     // keep the carried surplus at the front
            ahead.erase(ahead.begin(), ahead.begin() + (long)used);
            used = 0;
        }
        AP_HIP(hipMemcpyAsync(dev, pinned, count, hipMemcpyHostToDevice, s));
        return AP_OK;
    }
    // draw ahead up to `count` unused flags (the next round's need is at most this)
    void prepare(size_t count) {
        if (ahead.size() - used < count) draw(count - (ahead.size() - used));
    }
};

struct Cluster {
    int n = 0, num = 0;
    unsigned long long cap_edges = 0;   // pair entries at the start: no later edge list is longer
    unsigned num_edges = 0;
    unsigned long long *key = nullptr, *key_alt = nullptr;
    int *weight = nullptr, *weight_alt = nullptr;
    int *ei = nullptr, *ej = nullptr, *ew = nullptr, *row_start = nullptr;
    int *cid = nullptr, *parent = nullptr, *kept = nullptr, *rank = nullptr, *new_id = nullptr, *best_to = nullptr;
    double* best_score = nullptr;
    long long *area = nullptr, *area_alt = nullptr;
    unsigned *dev_count = nullptr;      // [0]: edges, [1]: clusters (both on the device, ahead of the host's copies)
    unsigned* host_count = nullptr;     // pinned: dev_count read back at the end of a round
    void* temp = nullptr;
    size_t temp_bytes = 0;
    std::vector<void*> mem;
    HeadStream heads;

    template <typename T>
    int alloc(T** p, unsigned long long count) {
        AP_HIP(dev_malloc((void**)p, (size_t)(count > 0 ? count : 1) * sizeof(T)));
        mem.push_back((void*)*p);
        return AP_OK;
    }
    void destroy() {
        (void)hipDeviceSynchronize();      // then every buffer below is idle
        for (void* p : mem) dev_free_synced(p);
        mem.clear();
        dev_free_synced(temp);
        temp = nullptr;
        if (heads.pinned) (void)hipHostFree(heads.pinned);
        heads.pinned = nullptr;
        if (host_count) (void)hipHostFree(host_count);
        host_count = nullptr;
    }
    int temp_for(size_t need) {
        if (need > temp_bytes) {
            dev_free(temp);
            temp = nullptr;
            AP_HIP(dev_malloc(&temp, need));
            temp_bytes = need;
        }
        return AP_OK;
    }
    // the counts on the device, read back: the host's num_edges and num
    int read_counts(hipStream_t s) {
        AP_HIP(hipMemcpyAsync(host_count, dev_count, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
        AP_HIP(hipStreamSynchronize(s));
        num_edges = host_count[0];
        num = (int)host_count[1];
        return AP_OK;
    }
    int sort_and_merge(unsigned long long entries, unsigned long long max_keys, hipStream_t s);
    int create(int cells, const int* pin_cell, const int* net_start, int num_nets, const long long* area_in,
               int max_degree, hipStream_t s);
    // while_busy: host work run once the round's first kernels are queued, before the host waits for them
    int round(const unsigned char* head, float cap, long long big, int per_head, hipStream_t s,
              HeadStream* while_busy = nullptr);
};

// sort `entries` (key, weight) pairs by key, sum equal keys into the edge list, drop the self-loop key num^2 when
// present (num = count[1], the clusters, at most max_keys) and rebuild the rows; no wait for the device
int Cluster::sort_and_merge(unsigned long long entries, unsigned long long max_keys, hipStream_t s) {
    const unsigned end_bit = bits_for(max_keys * max_keys);
    if (entries > 0) {
        size_t need = 0, need2 = 0;
        AP_HIP(rocprim::radix_sort_pairs(nullptr, need, key, key_alt, weight, weight_alt, (size_t)entries, 0, end_bit,
                                         s));
        AP_HIP(rocprim::reduce_by_key(nullptr, need2, key_alt, weight_alt, (size_t)entries, key, weight, dev_count,
                                      rocprim::plus<int>(), rocprim::equal_to<unsigned long long>(), s));
        int st = temp_for(need > need2 ? need : need2);
        if (st) return st;
        size_t b = temp_bytes;
        AP_HIP(rocprim::radix_sort_pairs(temp, b, key, key_alt, weight, weight_alt, (size_t)entries, 0, end_bit, s));
        b = temp_bytes;
        AP_HIP(rocprim::reduce_by_key(temp, b, key_alt, weight_alt, (size_t)entries, key, weight, dev_count,
                                      rocprim::plus<int>(), rocprim::equal_to<unsigned long long>(), s));
        hipLaunchKernelGGL(clu_drop_self_loops, dim3(1), dim3(64), 0, s, key, dev_count);
        AP_LAUNCHED();
        hipLaunchKernelGGL(clu_split_keys, dim3(grid_of((long long)entries, CLU_BLOCK)), dim3(CLU_BLOCK), 0, s, key,
                           weight, dev_count, ei, ej, ew);
        AP_LAUNCHED();
    } else {
        AP_HIP(hipMemsetAsync(dev_count, 0, sizeof(unsigned), s));
    }
    hipLaunchKernelGGL(clu_row_starts, dim3(grid_of((long long)max_keys + 1, CLU_BLOCK)), dim3(CLU_BLOCK), 0, s, key,
                       dev_count, row_start);
    AP_LAUNCHED();
    return AP_OK;
}

int Cluster::create(int cells, const int* pin_cell, const int* net_start, int num_nets, const long long* area_in,
                    int max_degree, hipStream_t s) {
    n = num = cells;
    int st;
    unsigned* count = nullptr;
    unsigned* offset = nullptr;
    AP_HIP(hipHostMalloc((void**)&host_count, 2 * sizeof(unsigned), hipHostMallocDefault));
    if ((st = alloc(&count, (unsigned long long)num_nets + 1)) || (st = alloc(&offset, (unsigned long long)num_nets + 1)))
        return st;
    if (num_nets > 0) {
        hipLaunchKernelGGL(clu_count_pairs, dim3(grid_of(num_nets, CLU_BLOCK)), dim3(CLU_BLOCK), 0, s, net_start,
                           num_nets, pin_cell, n, max_degree, count);
        AP_LAUNCHED();
    }
    AP_HIP(hipMemsetAsync(count + num_nets, 0, sizeof(unsigned), s));
    size_t need = 0;
    AP_HIP(rocprim::exclusive_scan(nullptr, need, count, offset, 0u, (size_t)num_nets + 1, rocprim::plus<unsigned>(),
                                   s));
    if ((st = temp_for(need > 0 ? need : 1))) return st;
    size_t b = temp_bytes;
    AP_HIP(rocprim::exclusive_scan(temp, b, count, offset, 0u, (size_t)num_nets + 1, rocprim::plus<unsigned>(), s));
    AP_HIP(hipMemcpyAsync(host_count, offset + num_nets, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    AP_HIP(hipStreamSynchronize(s));
    const unsigned total = host_count[0];
    cap_edges = total;
    if ((st = alloc(&key, cap_edges)) || (st = alloc(&key_alt, cap_edges)) || (st = alloc(&weight, cap_edges)) ||
        (st = alloc(&weight_alt, cap_edges)) || (st = alloc(&ei, cap_edges)) || (st = alloc(&ej, cap_edges)) ||
        (st = alloc(&ew, cap_edges)) || (st = alloc(&row_start, (unsigned long long)n + 1)) ||
        (st = alloc(&cid, n)) || (st = alloc(&parent, n)) || (st = alloc(&kept, n)) || (st = alloc(&rank, n)) ||
        (st = alloc(&new_id, n)) || (st = alloc(&best_to, n)) || (st = alloc(&best_score, n)) ||
        (st = alloc(&area, n)) || (st = alloc(&area_alt, n)) || (st = alloc(&dev_count, 2)))
        return st;
    if (num_nets > 0 && total > 0) {
        hipLaunchKernelGGL(clu_emit_pairs, dim3(grid_of(num_nets, CLU_BLOCK)), dim3(CLU_BLOCK), 0, s, net_start,
                           num_nets, pin_cell, n, max_degree, offset, key, weight);
        AP_LAUNCHED();
    }
    hipLaunchKernelGGL(clu_start, dim3(grid_of(n, CLU_BLOCK)), dim3(CLU_BLOCK), 0, s, cid, n, dev_count);
    AP_LAUNCHED();
    AP_HIP(hipMemcpyAsync(area, area_in, (size_t)n * sizeof(long long), hipMemcpyDeviceToDevice, s));
    if ((st = sort_and_merge(total, (unsigned long long)n, s))) return st;
    return read_counts(s);
}

// The round's steps run over the clusters before it (num, known to the host) and read the clusters after it from
// count[1] on the device; when nothing merges, every cluster keeps its index and the edges their keys, so the state
// is rebuilt unchanged. One wait, at the end, for the counts.
int Cluster::round(const unsigned char* head, float cap, long long big, int per_head, hipStream_t s,
                   HeadStream* while_busy) {
    const dim3 g(grid_of(num, CLU_BLOCK)), b(CLU_BLOCK);
    hipLaunchKernelGGL(clu_best_head, g, b, 0, s, row_start, ej, ew, num, head, area, big, cap, best_to, best_score);
    AP_LAUNCHED();
    hipLaunchKernelGGL(clu_identity, g, b, 0, s, parent, num);
    AP_LAUNCHED();
    switch (per_head) {
        case 1: hipLaunchKernelGGL(clu_take<1>, g, b, 0, s, row_start, ej, num, head, best_to, best_score, parent); break;
        case 2: hipLaunchKernelGGL(clu_take<2>, g, b, 0, s, row_start, ej, num, head, best_to, best_score, parent); break;
        case 3: hipLaunchKernelGGL(clu_take<3>, g, b, 0, s, row_start, ej, num, head, best_to, best_score, parent); break;
        default: hipLaunchKernelGGL(clu_take<4>, g, b, 0, s, row_start, ej, num, head, best_to, best_score, parent);
    }
    AP_LAUNCHED();
    hipLaunchKernelGGL(clu_kept, g, b, 0, s, parent, num, kept, area_alt);
    AP_LAUNCHED();
    size_t need = 0;
    AP_HIP(rocprim::exclusive_scan(nullptr, need, kept, rank, 0, (size_t)num, rocprim::plus<int>(), s));
    int st = temp_for(need);
    if (st) return st;
    size_t tb = temp_bytes;
    AP_HIP(rocprim::exclusive_scan(temp, tb, kept, rank, 0, (size_t)num, rocprim::plus<int>(), s));
    hipLaunchKernelGGL(clu_count_kept, dim3(1), dim3(64), 0, s, rank, kept, num, dev_count);
    AP_LAUNCHED();
    hipLaunchKernelGGL(clu_new_index, g, b, 0, s, parent, rank, num, new_id);
    AP_LAUNCHED();
    hipLaunchKernelGGL(clu_relabel_cells, dim3(grid_of(n, CLU_BLOCK)), b, 0, s, cid, n, new_id);
    AP_LAUNCHED();
    hipLaunchKernelGGL(clu_merge_area, g, b, 0, s, area, num, new_id, area_alt);
    AP_LAUNCHED();
    std::swap(area, area_alt);
    const unsigned entries = num_edges;
    if (entries)
        hipLaunchKernelGGL(clu_relabel_edges, dim3(grid_of(entries, CLU_BLOCK)), b, 0, s, ei, ej, ew, dev_count,
                           new_id, key, weight);
    AP_LAUNCHED();
    if ((st = sort_and_merge(entries, (unsigned long long)num, s))) return st;
    if (while_busy) while_busy->prepare((size_t)num);         // the next round's heads, beside this round
    return read_counts(s);
}

}  // namespace ap

using namespace ap;

struct ap_cluster {
    Cluster C;
};

extern "C" {

int ap_cluster_create(int num_cells, const int* pin_cell, const int* net_start, int num_nets, const long long* area,
                      int max_degree, void* stream, ap_cluster** out) {
    AP_REQUIRE(out && num_cells > 0 && num_nets >= 0 && area && (num_nets == 0 || (pin_cell && net_start)),
               "cluster_create: bad arguments");
    AP_REQUIRE(max_degree >= 2 && max_degree <= CLU_MAX_DEGREE, "cluster_create: max_degree %d outside 2 .. %d",
               max_degree, CLU_MAX_DEGREE);
    AP_REQUIRE((unsigned long long)num_cells * (unsigned long long)num_cells < (1ull << 62),
               "cluster_create: %d cells are too many for the pair keys", num_cells);
    ap_cluster* h = new ap_cluster();
    const int st = h->C.create(num_cells, pin_cell, net_start, num_nets, area, max_degree, (hipStream_t)stream);
    if (st) {
        (void)hipDeviceSynchronize();
        h->C.destroy();
        delete h;
        return st;
    }
    *out = h;
    return AP_OK;
}

int ap_cluster_round(ap_cluster* c, const unsigned char* head, float cap, long long big, int per_head, int* num_out,
                     void* stream) {
    AP_REQUIRE(c && head && num_out, "cluster_round: null argument");
    AP_REQUIRE(per_head >= 1 && per_head <= 4, "cluster_round: per_head %d outside 1 .. 4", per_head);
    const int st = c->C.round(head, cap, big, per_head, (hipStream_t)stream);
    if (st) return st;
    *num_out = c->C.num;
    return AP_OK;
}

int ap_cluster_seed_heads(ap_cluster* c, unsigned int seed) {
    AP_REQUIRE(c, "cluster_seed_heads: null argument");
    Cluster& C = c->C;
    if (!C.heads.pinned) AP_HIP(hipHostMalloc((void**)&C.heads.pinned, (size_t)C.n, hipHostMallocDefault));
    if (!C.heads.dev) {
        const int st = C.alloc(&C.heads.dev, (unsigned long long)C.n);
        if (st) return st;
    }
    C.heads.seed(seed);
    return AP_OK;
}

int ap_cluster_round_seeded(ap_cluster* c, float cap, long long big, int per_head, int* num_out, void* stream) {
    AP_REQUIRE(c && num_out, "cluster_round_seeded: null argument");
    AP_REQUIRE(c->C.heads.seeded, "cluster_round_seeded: no head stream (ap_cluster_seed_heads)");
    AP_REQUIRE(per_head >= 1 && per_head <= 4, "cluster_round_seeded: per_head %d outside 1 .. 4", per_head);
    Cluster& C = c->C;
    hipStream_t s = (hipStream_t)stream;
    int st = C.heads.take((size_t)C.num, s);
    if (st) return st;
    st = C.round(C.heads.dev, cap, big, per_head, s, &C.heads);
    if (st) return st;
    *num_out = C.num;
    return AP_OK;
}

int ap_head_draws(unsigned int seed, int count, unsigned char* out) {
    AP_REQUIRE(count >= 0 && (count == 0 || out), "head_draws: bad arguments");
    HeadStream h;
    h.seed(seed);
    h.draw((size_t)count);
    if (count) std::memcpy(out, h.ahead.data(), (size_t)count);
    return AP_OK;
}

int ap_cluster_num(const ap_cluster* c) { return c ? c->C.num : -1; }

int ap_cluster_edges(const ap_cluster* c) { return c ? (int)c->C.num_edges : -1; }

int ap_cluster_result(ap_cluster* c, long long* cid, long long* area, void* stream) {
    AP_REQUIRE(c && cid && area, "cluster_result: null argument");
    const Cluster& C = c->C;
    const int m = C.n > C.num ? C.n : C.num;
    hipLaunchKernelGGL(clu_copy_out, dim3(grid_of(m, CLU_BLOCK)), dim3(CLU_BLOCK), 0, (hipStream_t)stream, C.cid, C.n,
                       C.area, C.num, cid, area);
    AP_LAUNCHED();
    return AP_OK;
}

int ap_cluster_destroy(ap_cluster* c) {
    if (c) {
        (void)hipDeviceSynchronize();
        c->C.destroy();
        delete c;
    }
    return AP_OK;
}

}  // extern "C"
