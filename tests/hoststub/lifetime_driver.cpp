// Drives the host side of the level and batch engines (graph cache, captures, timing events, layouts, autotuner,
// profile, controller, destroy) over the HIP stub (hip_stub.cpp): no kernel runs, so the device buffers hold zeros
// and the results mean nothing; what is checked is that every stream, event, graph, executable and buffer is used
// only while alive and destroyed once, and that no executable graph is destroyed while a launch of it may still run.
// Exits 0 after "lifetime driver: ok"; any failure aborts.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "autoplace_gp.h"

#define CHECK(x)                                                                                     \
    do {                                                                                             \
        const int st_ = (x);                                                                         \
        if (st_ != AP_OK) {                                                                          \
            fprintf(stderr, "%s:%d %s -> %d: %s\n", __FILE__, __LINE__, #x, st_, ap_last_error());  \
            abort();                                                                                 \
        }                                                                                            \
    } while (0)

namespace {

template <typename T>
T* buf(std::vector<std::vector<char>>& keep, size_t n, T v = T()) {
    keep.emplace_back(sizeof(T) * (n ? n : 1));
    T* p = (T*)keep.back().data();
    for (size_t i = 0; i < n; i++) p[i] = v;
    return p;
}

// a small level: `cells` cells and as many fillers, nets of 2 and 3 pins, an nb x nb force grid and (judge) an
// nb / 2 overflow grid
ap_level_desc level(std::vector<std::vector<char>>& keep, int cells, int nb, bool judge) {
    ap_level_desc d;
    std::memset(&d, 0, sizeof(d));
    const int nodes = 2 * cells, nets = cells;
    std::vector<int> pin_cell, net_start{0};
    for (int k = 0; k < nets; k++) {
        const int deg = 2 + k % 2;
        for (int q = 0; q < deg; q++) pin_cell.push_back((k + 7 * q) % cells);
        net_start.push_back((int)pin_cell.size());
    }
    const int pins = (int)pin_cell.size();
    std::vector<int> cell_start(cells + 1, 0), cell_pin(pins);
    for (int p : pin_cell) cell_start[p + 1]++;
    for (int c = 0; c < cells; c++) cell_start[c + 1] += cell_start[c];
    std::vector<int> at(cell_start.begin(), cell_start.end() - 1);
    for (int p = 0; p < pins; p++) cell_pin[at[pin_cell[p]]++] = p;
    d.num_nodes = nodes;
    d.num_mov = cells;
    d.num_pins = pins;
    d.num_nets = nets;
    int* pc = buf<int>(keep, pins);
    std::memcpy(pc, pin_cell.data(), pins * sizeof(int));
    int* ns = buf<int>(keep, nets + 1);
    std::memcpy(ns, net_start.data(), (nets + 1) * sizeof(int));
    int* cs = buf<int>(keep, cells + 1);
    std::memcpy(cs, cell_start.data(), (cells + 1) * sizeof(int));
    int* cp = buf<int>(keep, pins);
    std::memcpy(cp, cell_pin.data(), pins * sizeof(int));
    d.pin_cell = pc;
    d.pin_off = buf<float>(keep, 2 * (size_t)pins, 0.f);
    d.net_start = ns;
    d.cell_start = cs;
    d.cell_pin = cp;
    d.size = buf<float>(keep, 2 * (size_t)nodes, 1.5f);
    d.ratio = buf<float>(keep, nodes, 1.f);
    d.pins = buf<float>(keep, nodes, 2.f);
    d.area = buf<float>(keep, nodes, 2.25f);
    d.lo = buf<float>(keep, 2 * (size_t)nodes, 0.75f);
    d.hi = buf<float>(keep, 2 * (size_t)nodes, (float)nb - 0.75f);
    d.nbx = d.nby = nb;
    d.bx = d.by = 1.0;
    d.fixed = buf<float>(keep, (size_t)nb * nb, 0.f);
    d.target = 1.f;
    d.bin_area = 1.0;
    d.mov_area = 2.25 * cells;
    if (judge) {
        d.has_judge = 1;
        d.njx = d.njy = nb / 2;
        d.jbx = d.jby = 2.0;
        d.judge_bin_area = 4.0;
        d.judge_size = buf<float>(keep, 2 * (size_t)cells, 1.f);
        d.judge_ratio = buf<float>(keep, cells, 1.f);
        d.judge_fixed = buf<float>(keep, (size_t)(nb / 2) * (nb / 2), 0.f);
    }
    d.poisson_method = AP_POISSON_FFT;
    d.u = buf<float>(keep, 2 * (size_t)nodes, 4.f);
    d.v = buf<float>(keep, 2 * (size_t)nodes, 4.f);
    d.g = buf<float>(keep, 2 * (size_t)nodes, 0.f);
    d.best = buf<float>(keep, 2 * (size_t)nodes, 4.f);
    d.S = buf<double>(keep, 21, 0.0);
    d.kp = buf<float>(keep, 5, 1.f);
    d.c = buf<double>(keep, 13, 1.0);
    return d;
}

}  // namespace

int main() {
    std::vector<std::vector<char>> keep;
    keep.reserve(256);
    // one level: more distinct pass counts than the graph cache holds, run back to back with no host
    // synchronisation between them (each new count evicts a cached graph whose last launch may still run)
    ap_level_desc d = level(keep, 3000, 64, true);
    ap_level* L = nullptr;
    CHECK(ap_level_create(&d, &L));
    const int counts[] = {1, 2, 3, 5, 8, 13, 2, 21};
    for (int round = 0; round < 3; round++)
        for (int n : counts) CHECK(ap_level_run(L, n, 1, nullptr));
    // timing classes: graphs with event-record nodes, evicted and re-captured
    for (int cls = 0; cls < 9; cls++) {
        CHECK(ap_level_time_class(L, cls));
        CHECK(ap_level_time_every(L, 1 + cls % 3));
        CHECK(ap_level_run(L, 4 + cls % 2, 1, nullptr));
        double ms = 0;
        int timed = 0;
        CHECK(ap_level_class_ms(L, &ms, &timed));
    }
    CHECK(ap_level_time_class(L, -1));
    CHECK(ap_level_prepare(L, 7));
    CHECK(ap_level_run(L, 7, 1, nullptr));
    CHECK(ap_level_run(L, 6, 0, nullptr));              // direct launches
    float last = 0;
    CHECK(ap_level_last_ms(L, &last));
    double prof[16];
    CHECK(ap_level_profile(L, 2, prof, nullptr));
    // layouts, rooms and the autotuner (offline, then online) between runs that leave launches pending
    CHECK(ap_level_run(L, 3, 1, nullptr));
    CHECK(ap_level_set_tile_layout(L, 16, 1, 8, 1));
    CHECK(ap_level_run(L, 3, 1, nullptr));
    CHECK(ap_level_set_bucket_room(L, 120, 8));
    CHECK(ap_level_run(L, 3, 1, nullptr));
    int sides[4];
    CHECK(ap_level_autotune(L, 2, 0, sides));
    CHECK(ap_level_run(L, 5, 1, nullptr));
    CHECK(ap_level_autotune(L, 2, AP_AUTOTUNE_ONLINE, sides));
    CHECK(ap_level_run(L, 5, 1, nullptr));
    unsigned long long spills[2];
    CHECK(ap_level_bucket_spills(L, spills));
    // the learned controller attached and detached between runs
    std::vector<float> params((size_t)ap_mlp_param_count(), 0.f);
    float* ctl_log = buf<float>(keep, 64 * 16, 0.f);
    CHECK(ap_level_set_controller(L, params.data(), 0.1f, 7, 2, 0.f, ctl_log, 64, 50));
    CHECK(ap_level_run(L, 4, 1, nullptr));
    double state[8];
    CHECK(ap_level_controller_state(L, state));
    CHECK(ap_level_set_controller(L, nullptr, 0.f, 0, 1, 0.f, nullptr, 0, 50));
    CHECK(ap_level_run(L, 9, 1, nullptr));             // destroyed with that run's launch pending
    CHECK(ap_level_destroy(L));

    // a batch of three levels (one without an overflow grid), evictions and the autotuner as above
    std::vector<ap_level_desc> blocks = {level(keep, 2000, 32, true), level(keep, 5000, 64, false),
                                         level(keep, 1200, 32, true)};
    ap_batch* B = nullptr;
    CHECK(ap_batch_create(blocks.data(), (int)blocks.size(), &B));
    for (int round = 0; round < 2; round++)
        for (int n : counts) CHECK(ap_batch_run(B, n, 1, nullptr));
    int layout[6];
    CHECK(ap_batch_autotune(B, 2, layout));
    CHECK(ap_batch_run(B, 4, 1, nullptr));
    CHECK(ap_batch_tile_layout(B, layout));
    CHECK(ap_batch_run(B, 3, 0, nullptr));
    CHECK(ap_batch_destroy(B));

    // a Poisson plan's life, and the drop-in map's stream-ordered workspace
    ap_poisson* P = nullptr;
    CHECK(ap_poisson_create(64, 64, 1.0, 1.0, AP_POISSON_FFT, &P));
    float* rho = buf<float>(keep, 64 * 64, 0.5f);
    float* fx = buf<float>(keep, 64 * 64);
    float* fy = buf<float>(keep, 64 * 64);
    CHECK(ap_poisson_solve(P, rho, 64, fx, fy, 64, nullptr));
    CHECK(ap_poisson_destroy(P));
    const ap_level_desc& q = blocks[0];
    float* mirror = buf<float>(keep, 4 * 32 * 32);
    float* over = buf<float>(keep, 32 * 32);
    CHECK(ap_density_map(q.v, q.size, q.ratio, q.num_nodes, q.num_mov, 1.f, 1.f, 32, 32, nullptr, q.fixed, 1.f, mirror,
                         over, nullptr));
    printf("lifetime driver: ok\n");
    return 0;
}
