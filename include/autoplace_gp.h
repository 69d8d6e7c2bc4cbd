/* C ABI of the MI355X (gfx950) global-placement library, libautoplace_gp.so.
 *
 * Plain pointers (device memory) and sizes; no framework types. Every call is asynchronous on the given HIP stream
 * (`stream` is a hipStream_t; NULL = the null stream of the current device) and returns AP_OK or an error code;
 * ap_last_error() describes the last error of the calling thread.
 *
 * Part 1 replaces the six launchers the reference's binding declares (cpp_to_py/gp_cuda.cpp:4-13, kernels in
 * cpp_to_py/gp_cuda_kernel.cu) with the same arguments, in the same order, plus the stream (and, for the WA, the pin
 * count the reference reads from its tensor). Part 2 is the Poisson solve the reference leaves to torch.fft
 * (src/core/electronic_density_layer.py:133). Part 3 runs whole passes of the GP iteration on the device, the work of
 * src/gpu_loop.py:58-64 (nesterov_step, Objective.on_gpu, control, select) as one native engine.
 */
#ifndef AUTOPLACE_GP_H
#define AUTOPLACE_GP_H

#if defined(__GNUC__)
#define AP_API __attribute__((visibility("default")))
#else
#define AP_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define AP_OK 0
#define AP_ERR_ARG 1          /* invalid argument (shape, size, null pointer) */
#define AP_ERR_HIP 2          /* HIP runtime error */
#define AP_ERR_UNSUPPORTED 3  /* valid request this build does not support */

/* Text of the last error of the calling thread (empty if none). */
AP_API const char* ap_last_error(void);
/* Library ABI version (major * 100 + minor). */
AP_API int ap_version(void);
/* The library keeps the device memory of the engines, plans and clusterings it destroys, per device and size class,
 * for the next ones (no reference counterpart: the reference's buffers are torch tensors, which torch's caching
 * allocator keeps the same way). ap_cached_device_bytes: the bytes kept and not in use; ap_release_cached_memory hands
 * them back to HIP (the library also does so by itself when an allocation runs out of memory). */
AP_API long long ap_cached_device_bytes(void);
AP_API int ap_release_cached_memory(void);

/* ---------------------------------------------------------------- part 1: drop-in kernels */

/* WA wirelength (gp_cuda.cpp:4-5 `wa_wirelength`; kernels gp_cuda_kernel.cu:12-65).
 * pos f32[>=num_mov][2]; pin_cell i32[num_pins] (num_mov = fixed pin at off); pin_off f32[num_pins][2];
 * net_start i32[num_nets+1] (pins grouped by net); cell_start i32[num_mov+1], cell_pin i32[>=cell_start[num_mov]];
 * gamma f32[1] on the device. Writes pin_grad f32[num_pins][2], out f32[num_nets][2] = per net (WA, HPWL) and
 * grad f32[num_mov][2]. */
AP_API int ap_wa_wirelength(const float* pos, const int* pin_cell, const float* pin_off, const int* net_start,
                     const int* cell_start, const int* cell_pin, int num_nets, int num_mov, const float* gamma,
                     float* pin_grad, float* out, float* grad, int num_pins, void* stream);

/* WA wirelength with a plan of its netlist: the same arguments and outputs as ap_wa_wirelength (gp_cuda.cpp:4-5),
 * plus the plan. The plan (no counterpart in the reference, whose kernel gives each net one thread) lays the netlist
 * out once, on the device, as the GP engine's WA: every net of d <= 64 pins in an aligned group of 2^ceil(log2 d)
 * lanes, one degree class per wavefront, each lane one pin, the net's max / min and exponential sums as butterfly
 * reductions over its group (nets of more pins: one workgroup each). It copies what it needs of pin_cell, pin_off and
 * net_start, so a plan describes the netlist as it was at ap_wa_plan_create: a changed netlist needs a new plan.
 * create is synchronous (the netlist must be written before the call); destroy synchronises the device. The per-net
 * values and gradients agree with ap_wa_wirelength's to float rounding (the sums are taken in another order). */
typedef struct ap_wa_plan ap_wa_plan;
AP_API int ap_wa_plan_create(const int* pin_cell, const float* pin_off, const int* net_start, int num_nets,
                             int num_pins, ap_wa_plan** plan);
AP_API int ap_wa_plan_destroy(ap_wa_plan* plan);
AP_API int ap_wa_wirelength_planned(const ap_wa_plan* plan, const float* pos, const int* pin_cell,
                                    const float* pin_off, const int* net_start, const int* cell_start,
                                    const int* cell_pin, int num_nets, int num_mov, const float* gamma,
                                    float* pin_grad, float* out, float* grad, int num_pins, void* stream);

/* Density map (gp_cuda.cpp:6-7 `density_map`; kernels gp_cuda_kernel.cu:84-95,130-146), computed by the GP
 * engine's tile binning (per-tile LDS sums of the reference's 2^32 fixed-point contributions: bit-exact).
 * map: the caller's u64[2][nbx][nby] (the reference's binding allocates it zeroed), here only a workspace: when it
 * holds ap_density_map_work_bytes() bytes it is used as such (its contents on return are unspecified), else the call
 * takes a stream-ordered workspace from the device's memory pool (hipMallocAsync) and frees it on the same stream;
 * fixed f32[nbx][nby]. Writes rho f32[2 nbx][2 nby] (total density mirrored in both axes) and over f32[nbx][nby]. */
AP_API int ap_density_map(const float* pos, const float* size, const float* ratio, int num_nodes, int num_mov, float bx,
                   float by, int nbx, int nby, unsigned long long* map, const float* fixed, float target, float* rho,
                   float* over, void* stream);

/* Workspace bytes of ap_density_map_ws for these sizes (-1: bad sizes). */
AP_API long long ap_density_map_work_bytes(int num_nodes, int num_mov, int nbx, int nby);

/* ap_density_map with the caller's workspace (work_bytes >= ap_density_map_work_bytes(); any contents, 16-byte
 * aligned): allocates nothing. Calls that share a workspace must not overlap. */
AP_API int ap_density_map_ws(const float* pos, const float* size, const float* ratio, int num_nodes, int num_mov,
                      float bx, float by, int nbx, int nby, const float* fixed, float target, float* rho, float* over,
                      void* work, long long work_bytes, void* stream);

/* Nesterov step (gp_cuda.cpp:8-9 `nesterov_step`; kernel gp_cuda_kernel.cu:157-170); n = coordinates. */
AP_API int ap_nesterov_step(const float* u, const float* v, const float* g, const float* lo, const float* hi,
                     const float* alpha, const float* coef, int n, float* u_next, float* v_next, void* stream);

/* Preconditioned gradient from the field (gp_cuda.cpp:10-11 `density_grad`; kernel gp_cuda_kernel.cu:102-125,148-153),
 * computed by the GP engine's gather.
 * field f32[2][2 nbx][2 nby] (only [:, :nbx, :nby] is read); wl_grad f32[num_wl][2]; weight, precond f32[1]. */
AP_API int ap_density_grad(const float* pos, const float* size, const float* ratio, const float* field, const float* wl_grad,
                    const float* pins, const float* area, int num_nodes, int num_wl, const float* weight,
                    const float* precond, float bx, float by, int nbx, int nby, float* grad, void* stream);

/* One pass of the run's control (gp_cuda.cpp:12 `control`; kernel gp_cuda_kernel.cu:172-240).
 * S f64[21], kp f32[5], c f64[13]; the five scalars are f32[1] on the device. */
AP_API int ap_control(double* S, float* kp, const double* c, const float* moved, const float* turned, const float* hpwl,
               const float* ovf, const float* wa, void* stream);

/* Take an accepted step, keep the best placement (gp_cuda.cpp:13 `select_step`; kernel gp_cuda_kernel.cu:242-259). */
AP_API int ap_select_step(const double* S, float* u, float* v, float* g, const float* u_next, const float* v_next,
                   const float* g_next, float* best, int n, void* stream);

/* Exact HPWL over every net (src/evaluator.py:4-8; src/core/wa_wirelength_hpwl.py:58-70), reduced in FP64 on the
 * device: work f64[ceil(num_nets / 256)], out f64[1]. */
AP_API int ap_net_hpwl(const float* pos, const int* pin_cell, const float* pin_off, const int* net_start, int num_nets,
                       int num_mov, double* work, double* out, void* stream);

/* ap_net_hpwl without a work buffer and without any allocation. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Its partials live in a fixed device array of AP_NOALLOC_HPWL_PARTIALS doubles (512 KiB per device, held until the
 * process ends), so num_nets is at most AP_NOALLOC_HPWL_PARTIALS * 256 (16,777,216); above that it returns
 * AP_ERR_ARG. Calls on one device share that array: calls that overlap on one device (concurrent streams) race. */
#define AP_NOALLOC_HPWL_PARTIALS 65536
AP_API int ap_net_hpwl_noalloc(const float* pos, const int* pin_cell, const float* pin_off, const int* net_start,
                               int num_nets, int num_mov, double* out, void* stream);

/* ---------------------------------------------------------------- part 2: Poisson solve */

#define AP_POISSON_AUTO 0
#define AP_POISSON_GEMM 1   /* cosine / sine series as FP32 MFMA GEMMs */
#define AP_POISSON_FFT 2    /* LDS FFTs (powers of two only) */

typedef struct ap_poisson ap_poisson;

/* Plan for an nbx x nby grid of bins bx x by; method as above. */
AP_API int ap_poisson_create(int nbx, int nby, double bx, double by, int method, ap_poisson** plan);
AP_API int ap_poisson_destroy(ap_poisson* plan);
/* Method the plan uses (AP_POISSON_GEMM or AP_POISSON_FFT). */
AP_API int ap_poisson_method(const ap_poisson* plan);
/* field of the map rho (row stride rho_ld floats; rho_ld = nby for an unmirrored map, 2 nby for the top-left
 * quadrant of a mirrored one) into field_x / field_y (row stride field_ld). Equal to the reference's
 * irfft2(rfft2(mirror(rho)) * to_field)[:, :nbx, :nby]. */
AP_API int ap_poisson_solve(ap_poisson* plan, const float* rho, int rho_ld, float* field_x, float* field_y, int field_ld,
                     void* stream);

/* Electrostatic energy of the map, E = 1/2 sum_b rho_b phi_b (phi the potential whose field ap_poisson_solve
 * returns), reduced in FP64 into energy f64[1] on the device. The reference never evaluates it (src/calculator.py:4-5);
 * it is the value whose position gradient the field gather approximates. Needs a plan of the GEMM method. */
AP_API int ap_poisson_energy(ap_poisson* plan, const float* rho, int rho_ld, double* energy, void* stream);

/* ---------------------------------------------------------------- part 3: the GP iteration engine */

typedef struct ap_level ap_level;

/* One level's problem and state, all device pointers (the caller owns them and keeps them alive). */
typedef struct ap_level_desc {
    int num_nodes, num_mov, num_pins, num_nets;
    /* netlist: optimized nets, pins grouped by net; cell CSR of the movable pins */
    const int* pin_cell;
    const float* pin_off;
    const int* net_start;
    const int* cell_start;
    const int* cell_pin;
    /* nodes (cells then fillers): smoothed size f32[N][2], density ratio f32[N], pins f32[N], area f32[N],
     * position bounds f32[N][2] */
    const float* size;
    const float* ratio;
    const float* pins;
    const float* area;
    const float* lo;
    const float* hi;
    /* force grid */
    int nbx, nby;
    double bx, by;
    const float* fixed;      /* f32[nbx][nby] blocked density */
    float target;
    double bin_area, mov_area;
    /* overflow grid (has_judge = 0: the overflow is measured on the force grid) */
    int has_judge, njx, njy;
    double jbx, jby, judge_bin_area;
    const float* judge_size;   /* f32[num_mov][2] */
    const float* judge_ratio;  /* f32[num_mov] */
    const float* judge_fixed;  /* f32[njx][njy] */
    int poisson_method;        /* AP_POISSON_* */
    /* optimizer buffers f32[N][2]: u, v, g of the current point; best placement */
    float* u;
    float* v;
    float* g;
    float* best;
    /* control: S f64[21], kp f32[5], c f64[13] (layouts of src/gpu_loop.py:20-21, gp_cuda_kernel.cu:178-181) */
    double* S;
    float* kp;
    const double* c;
    /* optional: a Poisson plan of the force grid (same grid, bin sizes and a compatible method) that the level uses
     * instead of creating its own; the caller keeps it alive while the level lives, and the level does not destroy
     * it. Null: the level creates its own plan. */
    const ap_poisson* poisson_plan;
} ap_level_desc;

AP_API int ap_level_create(const ap_level_desc* desc, ap_level** level);
AP_API int ap_level_destroy(ap_level* level);
/* Run `passes` passes (each: Nesterov step, objective, control, select). With use_graph the passes are replayed from
 * a HIP graph captured on the first call with that pass count (the last 4 captures are kept). On return
 * (stream-ordered) u / v / g / best / S / kp hold the state after the passes, as the reference's loop leaves them. */
AP_API int ap_level_run(ap_level* level, int passes, int use_graph, void* stream);
/* Capture (and instantiate) the graph of `passes` passes now, so that the first ap_level_run with that count only
 * launches it. Optional. */
AP_API int ap_level_prepare(ap_level* level, int passes);
/* The wirelength-only start of a level (src/run_placement.py:41-76 while the density weight is 0, with the steps of
 * nesterov_optimizer.py:34-53 and the rule of param_scheduler.py:38-48) as engine passes: each pass is the Nesterov
 * step with kp's alpha and coef, WA, the cells' WA gradient divided by max(1, pins) (fillers 0), the FP64 norms, and
 * a control that takes every step (unless a norm or the HPWL is not finite: diverged), sets the BB step and the
 * momentum (S[ALPHA], S[A], kp[3], kp[4]) and records the pass's (WA, HPWL); the phase ends once WA fell by at most
 * tol x window x WA over the last `window` passes, or after max_passes passes. A run's passes after the end leave
 * the point unchanged. begin: resets the start state (synchronises the level's stream). start_run: `passes` such
 * passes (use_graph as ap_level_run). start_state: state f64[8] = (done: 0 running, 1 the rule, 2 diverged, 3 the
 * pass limit; passes taken; tol; window; max_passes) and history f64[n][2] = (WA, HPWL) of the first n passes taken
 * (waits for the level's stream). */
AP_API int ap_level_start_begin(ap_level* level, double tol, int window, int max_passes);
AP_API int ap_level_start_run(ap_level* level, int passes, int use_graph, void* stream);
AP_API int ap_level_start_state(ap_level* level, double* state, double* history, int n);
/* Kernel classes of a pass, in launch order, for ap_level_profile and ap_level_time_class. */
#define AP_PROF_NESTEROV 0
#define AP_PROF_WA 1
#define AP_PROF_WA_CELLS 2    /* the cells' WA gradients (sums of their pins') */
#define AP_PROF_MAP 3         /* density: tile binning of both grids (the fill of each grid's buckets) */
#define AP_PROF_POST 4        /* density: per-tile sums of the force grid, density and overflow */
#define AP_PROF_OVER_TILES 5  /* density: per-tile sums of the overflow grid (when it has one) */
#define AP_PROF_POISSON 6
#define AP_PROF_GATHER 7
#define AP_PROF_CONTROL 8
#define AP_PROF_CLASSES 9
/* Run `passes` passes with direct launches and HIP events around every kernel class; ms[AP_PROF_CLASSES] receives
 * the average milliseconds per pass of each class. Blocks the calling thread until the passes are done. */
AP_API int ap_level_profile(ap_level* level, int passes, double* ms, void* stream);
/* Graphs captured from now on record a HIP event before and after the kernels of class cls in every timed pass (-1:
 * none). After a run, ap_level_class_ms gives the class's average ms over the run's timed passes and their number
 * (blocks until the run ends). */
AP_API int ap_level_time_class(ap_level* level, int cls);
/* No counterpart in the reference. The timed passes of a run: 0, every, 2 every, ... (default 1: all). An event
 * record in a graph holds its queue for a few microseconds, so sparse samples disturb a timed run less. Drops the
 * level's captured graphs. */
AP_API int ap_level_time_every(ap_level* level, int every);
AP_API int ap_level_class_ms(ap_level* level, double* ms, int* passes);
/* Timing of the last ap_level_run (ms between its first and last kernel), for the caller's instrumentation. */
AP_API int ap_level_last_ms(ap_level* level, float* ms);

/* One gradient evaluation of the objective at pos (Objective.on_gpu, src/calculator.py:37-43) with the parameters
 * in kp (gamma, weight, precond): grad f32[N][2]; stats f64[5] = (WA, HPWL, overflow, 0, 0). */
AP_API int ap_level_objective(ap_level* level, const float* pos, float* grad, double* stats, void* stream);
/* The force grid's total density f32[nbx][nby] (unmirrored) of the last objective evaluated (pass or
 * ap_level_objective): the top-left quadrant of the reference's density_map result. */
AP_API int ap_level_density(ap_level* level, float* rho, void* stream);

/* Autotuner of the density tile layout and of the per-net and per-node kernels' launch shapes: the admissible tile
 * sides from half to twice the set-up side (8 .. 64 bins) with 1 or 2 tiles per workgroup of the tile kernel, for
 * the force grid, then the overflow grid, then the force grid again; then the WA kernel's plan wavefronts per
 * launched wavefront (1, 2, 4) and lanes per workgroup (128, 256, 512), then the gather's lanes per workgroup (128,
 * 256, 512); each timed on whole passes (reps each), the fastest kept. flags 0: timed from the level's current
 * state, which is saved first and restored after every timing (no pass of the level runs). flags
 * AP_AUTOTUNE_ONLINE: timed on the level's own next passes, which advance it as ap_level_run's passes would (the
 * control counts them): no extra passes. layout i32[4] (optional): force side, force group, overflow side, overflow
 * group (0, 0: no overflow grid); ap_level_launch_shape gives the shapes. The density map is an integer sum and the
 * WA and gather partials are per wavefront, so the results are the same whatever the choice. Drops the level's
 * captured graphs. */
#define AP_AUTOTUNE_ONLINE 1
AP_API int ap_level_autotune(ap_level* level, int reps, int flags, int* layout);
/* Set the density tile layout directly (an admissible side and group of each grid, group 0: the default for the
 * side; the overflow pair is ignored without an overflow grid): the autotuner's choice, replayed. Drops the level's
 * captured graphs. */
AP_API int ap_level_set_tile_layout(ap_level* level, int force_side, int force_group, int overflow_side,
                                    int overflow_group);
/* No counterpart in the reference. The launch shapes of the level's WA and gather kernels, shape i32[3]: WA plan
 * wavefronts per launched wavefront, WA lanes per workgroup, gather lanes per workgroup (the autotuner's choice, or
 * the defaults 2, 256, 256). set: those shapes directly (WA: 1, 2 or 4 and a multiple of 64 up to 512; gather: a
 * multiple of 64 up to 512), the autotuner's choice replayed; the results do not depend on them. Drops the level's
 * captured graphs. */
AP_API int ap_level_launch_shape(ap_level* level, int* shape);
AP_API int ap_level_set_launch_shape(ap_level* level, int wa_per, int wa_lanes, int gather_lanes);
/* No counterpart in the reference (its kernels read nodes in the problem's order). The spatial renumbering a level's
 * engine runs on: perm i32 [num_nodes] (new -> old) lists the cells, then the fillers, each in a stable order of the
 * Hilbert index of their bin (pos f32 [num_nodes][2] over an nbx x nby grid of bx x by bins, nbx, nby <= 32768; the
 * bin of x is floor(x * (1 / bx)) in single precision, clamped to the grid). The netlist (pin_cell i32 [num_pins]
 * grouped by net, cell num_mov = fixed pin; pin_off f32 [num_pins][2]; net_start i32 [num_nets + 1]) comes back with
 * the cells renumbered and the nets in a stable order of their smallest new cell (pin_cell_out, pin_off_out,
 * net_start_out), with the pins of every cell: cell_start i32 [num_mov + 1] and cell_pin i32 [num_pins] (new pin
 * indices in order, cell by cell; the fixed pins' entries after cell_start[num_mov]). Synchronises the stream. */
AP_API int ap_level_renumber(const float* pos, int num_nodes, int num_mov, int nbx, int nby, double bx, double by,
                             const int* pin_cell, const float* pin_off, const int* net_start, int num_nets,
                             int num_pins, int* perm, int* pin_cell_out, float* pin_off_out, int* net_start_out,
                             int* cell_start, int* cell_pin, void* stream);
/* No counterpart in the reference. The room the density binning gives each tile's bucket in a pass: percent % of the
 * entries the bucket had in the pass before, plus extra (defaults 150 and 64; percent 0..150, extra 0..64). Entries
 * past a bucket's room take the slower spill path (global integer atomics), so a smaller room changes the time, never
 * the density map; a room of 0 % + 0 sends every entry down that path, which is what it is for (tests). Drops the
 * level's captured graphs. */
AP_API int ap_level_set_bucket_room(ap_level* level, int percent, int extra);
/* No counterpart in the reference. Entries that took the spill path since the level was created, u64[2]: force grid,
 * overflow grid (0 without one). Synchronizes the device. */
AP_API int ap_level_bucket_spills(ap_level* level, unsigned long long* spills);
/* No counterpart in the reference. A density tile whose bucket had, in the pass before, 2 x slice entries or more and
 * factor times the grid's mean per tile or more is summed by up to 16 workgroups, a slice of its entries each
 * (defaults: slice 4096, factor 4; slice 0: never split, else 64 .. 2^24; factor 0 .. 1024): the level's start can
 * hold most nodes in a few tiles, which one workgroup each would sum alone while the rest of the GPU idles. Integer
 * sums, so the map does not change; a small slice with factor 0 splits every tile, which is what it is for (tests).
 * Drops the level's captured graphs. tile_splits: u64[2], the split tiles summed since the level was created (force
 * grid, overflow grid). Synchronizes the device. */
AP_API int ap_level_set_tile_split(ap_level* level, int slice, int factor);
AP_API int ap_level_tile_splits(ap_level* level, unsigned long long* splits);

/* ---------------------------------------------------------------- batch: many independent levels per launch
 * No counterpart in the reference. count levels (blocks) described as for ap_level_create (each with its own
 * buffers and control state S / kp / c; the FFT Poisson method and nets of at most 64 pins) run their passes
 * together: every stage of a pass is one launch over all blocks' workgroups, and a run of passes replays one HIP
 * graph. A block that has stopped is skipped, so each block's result is the one ap_level_run gives it. */
typedef struct ap_batch ap_batch;
AP_API int ap_batch_create(const ap_level_desc* blocks, int count, ap_batch** batch);
AP_API int ap_batch_destroy(ap_batch* batch);
AP_API int ap_batch_prepare(ap_batch* batch, int passes);
AP_API int ap_batch_run(ap_batch* batch, int passes, int use_graph, void* stream);
/* ms between the first and last kernel of the last ap_batch_run */
AP_API int ap_batch_last_ms(ap_batch* batch, float* ms);
/* ap_level_autotune for every block (optional layout i32[2 count]: side and group of block b's grid at 2b, 2b + 1) */
AP_API int ap_batch_autotune(ap_batch* batch, int reps, int* layout);
/* the tile layout of every block's grid: layout i32[2 count], side and group of block b at 2b, 2b + 1 */
AP_API int ap_batch_tile_layout(ap_batch* batch, int* layout);

/* ---------------------------------------------------------------- part 4: learned schedule controller
 * No counterpart in the reference: a 12 -> 64 -> 64 -> 2 ReLU MLP (FP32, parameters f32[ap_mlp_param_count()] laid
 * out W1[64][12], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]) whose outputs, clipped to +-0.2, are the logs of
 * factors on the density-weight growth and on the smoothing of every pass of the engine; the reference's rule
 * (gp_cuda_kernel.cu:198-209) is what they scale. Features (12, each clamped to [-10, 10]): overflow,
 * 100 x its change, log(HPWL / HPWL at the controller's start), 10 x relative HPWL change, WA model error
 * 1 - WA / HPWL, log(gamma / base gamma), density share lambda A / (pins + lambda A), log(step / first step),
 * 0.2 x log(|wirelength gradient| / |weighted density gradient|), iteration / 1000, log(1 + motion), phase
 * (1 coarse level, 0 cell level). */

AP_API int ap_mlp_param_count(void);
/* Workspace (floats) ap_mlp_grad needs for n samples. */
AP_API long long ap_mlp_work_floats(int n);
/* out f32[n][2] = the MLP (unclipped) on x f32[n][12]. */
AP_API int ap_mlp_forward(const float* params, const float* x, int n, float* out, void* stream);
/* Weighted squared loss sum_s weight[s] |mlp(x_s) - target_s|^2 (target f32[n][2]) into loss f64[1] and its
 * gradient into grad f32[ap_mlp_param_count()]; deterministic. work: f32[ap_mlp_work_floats(n)]. */
AP_API int ap_mlp_grad(const float* params, const float* x, const float* target, const float* weight, int n,
                       float* grad, double* loss, float* work, void* stream);
/* One Adam step (bias-corrected, step >= 1) on params with moments m, v. */
AP_API int ap_mlp_adam(float* params, float* m, float* v, const float* grad, int step, float lr, float beta1,
                       float beta2, float eps, void* stream);
/* Attach the controller to a level's passes (params null: the plain rule). sigma: Gaussian exploration noise added
 * to the outputs before the clip (counter-based draws from seed, a new draw every `hold` learned passes); phase: 1
 * for a coarse level; log: null or
 * f32[log_cap][16] receiving, per learned pass, the 12 features, the 2 applied outputs, overflow and HPWL; window:
 * the plain rule takes over for the rest of the level after `window` passes without a 0.1 % overflow improvement,
 * counted once the density share lambda A / (pins + lambda A) is 0.95 or more (where the rule's stall test counts).
 * Resets the controller's state and drops the level's captured graphs. */
AP_API int ap_level_set_controller(ap_level* level, const float* params, float sigma, unsigned long long seed,
                                   int hold, float phase, float* log, int log_cap, int window);
/* The controller's state f64[8]: started, previous overflow, first HPWL, first step, best overflow, pass of the last
 * improvement, fallback (1: the rule took over), learned passes so far. Waits for the level's stream. */
AP_API int ap_level_controller_state(ap_level* level, double* state);

/* ---------------------------------------------------------------- part 5: problem preparation on the device
 * Clustering of the two-level placement: replaces cluster_cells (src/cluster.py:38-85), which the reference runs as
 * torch operators inside the GP clock. The caller drives the rounds (cluster.py:64-85): for each, it draws the heads
 * (torch's CPU generator, as the reference) and passes them as u8 [num] on the device. */
typedef struct ap_cluster ap_cluster;
/* The weighted cell graph of the nets of 2 .. max_degree pins (cluster.py:44-58): pin_cell i32 [P] grouped by net
 * (num_cells = fixed pin), net_start i32 [num_nets + 1], area i64 [num_cells]. Synchronises the stream. */
AP_API int ap_cluster_create(int num_cells, const int* pin_cell, const int* net_start, int num_nets,
                             const long long* area, int max_degree, void* stream, ap_cluster** out);
/* One round of star matching (cluster.py:66-84) with heads u8 [num] on the device, the merged-area cap (float32, as
 * the reference's tensor), the big-cell limit and per_head (1 .. 4). *num_out: the clusters after the round (equal to
 * the clusters before it when nothing merged; the state is then unchanged). Synchronises the stream. */
AP_API int ap_cluster_round(ap_cluster* c, const unsigned char* head, float cap, long long big, int per_head,
                            int* num_out, void* stream);
/* The heads drawn by the library instead (the same draws): ap_cluster_seed_heads starts the stream of
 * torch.manual_seed(seed) (MT19937 with its standard seeding; a head is a float32 draw below 0.5, i.e. bit 23 of the
 * 32-bit output clear), and every ap_cluster_round_seeded takes the next `num` draws of it, `num` the clusters before
 * the round, as ap_cluster_round with those heads would. The host draws the next round's heads while the device runs
 * a round. Synchronises the stream. */
AP_API int ap_cluster_seed_heads(ap_cluster* c, unsigned int seed);
AP_API int ap_cluster_round_seeded(ap_cluster* c, float cap, long long big, int per_head, int* num_out,
                                   void* stream);
/* Host only: the first `count` head flags of the stream of `seed` (1: head), u8 [count] in host memory. */
AP_API int ap_head_draws(unsigned int seed, int count, unsigned char* out);
AP_API int ap_cluster_num(const ap_cluster* c);
AP_API int ap_cluster_edges(const ap_cluster* c);
/* cid i64 [num_cells]: the cluster of every cell; area i64 [num]: the clusters' areas. */
AP_API int ap_cluster_result(ap_cluster* c, long long* cid, long long* area, void* stream);
AP_API int ap_cluster_destroy(ap_cluster* c);

/* Free-area fraction of a grid (src/database.py:103-134) on the device: xs i64 [nx], ys i64 [ny] the edges of every
 * row / blocker box clipped to the region, usable u8 [(nx - 1) x (ny - 1)] the edge-grid cells inside a row and
 * under no blocker; out f64 [nbx x nby]. Equal to the host form to rounding (its sums run in another order). */
AP_API int ap_free_fraction(const long long* xs, int nx, const long long* ys, int ny, const unsigned char* usable,
                            double x0, double y0, double x1, double y1, int nbx, int nby, double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
