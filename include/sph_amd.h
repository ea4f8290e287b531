/*
 * sph_amd.h -- C ABI of the MI355X (gfx950) 2D SPH step library (libsph_amd.so).
 *
 * One opaque context owns every device buffer of one particle system on one GPU
 * and one HIP stream. Every call returns SPH_OK (0) or a negative status; the
 * message of the last failure is returned by sph_last_error(). Calls are
 * asynchronous on the context stream unless they return host data. Plain
 * pointers and sizes only: host arrays are float / int32 / uint32 / uint64.
 *
 * Each entry point replaces one role of the reference's device "acceleration"
 * API (jose-pinzon0407/sph-simulations, core/acceleration/); the replaced
 * declaration is cited beside it. Differences that are deliberate:
 *   - the context owns the buffers (the reference passes raw device pointers
 *     held in globals, app/cu_simulation_configuration.h:15-20);
 *   - errors are reported (the reference checks nothing);
 *   - the order is an enum, not a std::string (cuda_callbacks.cuh:26);
 *   - the CSR list capacity grows inside the context instead of being freed and
 *     reallocated by the callee every step (cuNeighborLists.cu:107-119).
 */
#ifndef SPH_AMD_H
#define SPH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_OK 0
#define SPH_ERR_INVALID (-1)  /* bad argument / unsupported configuration */
#define SPH_ERR_HIP (-2)      /* a HIP runtime call failed */
#define SPH_ERR_CAPACITY (-3) /* more particles than the context was created for */
#define SPH_ERR_STATE (-4)    /* call order violated (e.g. lists requested but not built) */

enum sph_order { SPH_ORDER_MORTON = 0, SPH_ORDER_ROWMAJOR = 1 };
enum sph_physics { SPH_PHYS_REFERENCE = 0, SPH_PHYS_WCSPH = 1 };
enum sph_boundary { SPH_BC_WALLS = 0, SPH_BC_PERIODIC = 1 };
enum sph_compute { SPH_COMPUTE_SORTED = 0, SPH_COMPUTE_LISTS = 1 };
/* Physics kernels of the fused step (sph_step); all give bit-identical results.
 * Block variants need Morton order (row-major always uses PARTICLE). */
enum sph_step_kernel {
  SPH_STEP_SPLIT_BLOCK = 0, /* density kernel + force kernel over LDS-staged Morton blocks */
  SPH_STEP_FUSED_BLOCK = 1, /* one kernel, two-cell halo, densities never leave LDS */
  SPH_STEP_PARTICLE = 2     /* one thread per particle, fixed-width neighbor list */
};
/* Re-sort inside sph_step. FULL: radix sort of every key, then the permutation.
 * INCREMENTAL: only particles whose cell key changed are radix-sorted, then merged
 * with the unchanged ones (already in order); the result is the same stable
 * order by (key, previous index). */
enum sph_resort { SPH_RESORT_FULL = 0, SPH_RESORT_INCREMENTAL = 1 };

/* Simulation parameters. Reference physics (SPH_PHYS_REFERENCE) reproduces the
 * macros of core/ConfigVariables.hpp:9-15 when filled with h=1, eos_k=30,
 * rho0=25, mu=0.01, gravity_y=-10, mass=1. */
typedef struct sph_params {
  double lo[2];      /* Domain::l_corner (core/Domain.h:7) */
  double hi[2];      /* Domain::h_corner (core/Domain.h:8) */
  double h;          /* kernel support radius (H) */
  double cell_size;  /* <= 0 -> h; must be >= h */
  double dt;
  double mass;       /* uniform particle mass */
  double rho0;       /* rest density */
  double gravity_y;
  double eos_k;      /* reference physics: p = eos_k (rho - rho0) */
  double mu;         /* reference physics: viscosity */
  double c0;         /* WCSPH: numerical sound speed */
  double alpha;      /* WCSPH: Monaghan artificial viscosity */
  double eps_visc;   /* WCSPH: mu_ij regulariser, times h^2 */
  int32_t gamma;     /* WCSPH: integer Tait exponent */
  int32_t physics;   /* enum sph_physics */
  int32_t boundary;  /* enum sph_boundary */
  int32_t order;     /* enum sph_order */
  int32_t ell_width; /* neighbor slots per particle in the fused step (0 -> 32) */
  int32_t reserved;
} sph_params;

/* Uniform cell grid derived from the parameters (cuGridInformation,
 * core/acceleration/cuParticles.cuh:31-38, plus MortonOrder padding). */
typedef struct sph_grid_info {
  float cs_x, cs_y;
  float lo_x, lo_y;
  float len_x, len_y;
  int32_t nx, ny;   /* logical cells per axis */
  int32_t nxp, nyp; /* padded extents of the key space */
  int64_t n_keys;   /* cell-table size = largest key + 1 */
  int32_t key_bits; /* significant key bits (radix passes = ceil(bits / 8)) */
  int32_t order;
} sph_grid_info;

/* Device view of the current (sorted) state, for consumers that read it in
 * place on the same stream (all arrays hold n elements, float unless noted).
 * The pid buffer alternates from step to step: take the view again after a
 * step call. */
typedef struct sph_device_view {
  int64_t n;
  float *rx, *ry, *vx, *vy, *ax, *ay, *rho, *pr; /* pr = p / rho^2 */
  uint32_t* pid;
  int32_t* ell_count; /* neighbors of each particle (self excluded) */
  int32_t* ell_index; /* n * ell_width slots, stencil order */
  int32_t ell_width;
  int32_t pad;
  const int32_t* cell_range; /* int2 (start, end) per key */
  void* stream;              /* hipStream_t of the context */
} sph_device_view;

/* Stage indices of sph_stage_times(), named after the reference timers
 * (app/simulation_configuration.h:72-91) where one exists. Each launch is
 * bracketed by hipEvents on the context stream when timing is on. */
enum sph_stage {
  SPH_T_INTEGRATION = 0,   /* standalone integration */
  SPH_T_NEIGHBOR_SEARCH,   /* cell keys; permutation (full sort) + cell ranges */
  SPH_T_NEIGHBOR_LIST,     /* CSR build (count, scan, fill) */
  SPH_T_DENSITY,           /* density + EOS (+ fixed-width neighbor list) */
  SPH_T_FORCE,             /* forces (+ integration and next keys inside sph_step) */
  SPH_T_SORT_HIST,         /* radix sort: digit histogram kernel */
  SPH_T_SORT_PASS,         /* radix sort: one onesweep pass (several per sort) */
  SPH_T_BLOCK_STEP,        /* fused density + forces + integration + keys (sph_step) */
  SPH_T_MEASURE,
  SPH_T_MOVERS,            /* incremental re-sort: mover report (when the step kernel
                              writes none), tile prefixes, movers sorted */
  SPH_T_MERGE,             /* incremental re-sort: merge movers and stayers (permutation) */
  SPH_T_COUNT
};

/* ---- lifetime (cuDeviceProperties / cuCreateContainer / cuCreateCellBuffers /
 *      cuSetupNeighborLists, core/acceleration/cudaLaunchers.cuh:7-15) ---- */
int sph_create(const sph_params* params, int64_t n_max, int device, void** out_ctx);
int sph_destroy(void* ctx);
const char* sph_last_error(const void* ctx); /* ctx may be NULL (creation errors) */
int sph_get_grid(const void* ctx, sph_grid_info* out);
/* Derived float constants in the order of the oracle's orc_consts struct. */
int sph_get_constants(const void* ctx, float* out, int32_t count);

/* ---- data movement (cuReadContainer / cuWriteContainer, cudaLaunchers.cuh:9-10) ---- */
int sph_upload(void* ctx, int64_t n, const float* rx, const float* ry, const float* vx,
               const float* vy, const uint32_t* pid /* NULL -> 0..n-1 */);
int sph_download(void* ctx, float* rx, float* ry, float* vx, float* vy, uint32_t* pid);
/* Any pointer may be NULL. pres is p (not p / rho^2). */
int sph_download_fields(void* ctx, float* ax, float* ay, float* rho, float* pres);
int64_t sph_size(const void* ctx);

/* ---- hot-path stages ---- */
/* cuCallbackFindNeighbors (cuda_callbacks.cuh:19-26): cell keys, stable radix sort
 * by key, permutation of r, v, a and particle ids, cell start/end ranges. */
int sph_find_neighbors(void* ctx);
/* cuCallbackNeighborLists (cuda_callbacks.cuh:48-57) / MultipassNeighborLists
 * (core/algorithmic/MultipassNeighborLists.h:14-18): CSR lists in stencil order.
 * include_self = 1 gives the CPU convention, 0 the GPU convention. */
int sph_build_neighbor_lists(void* ctx, int32_t include_self, int64_t* total);
/* Caller-built lists for SPH_COMPUTE_LISTS: the neighbor_list / offset_per_particle
 * pair that ListsForceComputation and MPListsForceComputation are constructed with
 * (core/compute/ListsForceComputation.h:5-66) and evaluate as given, without a
 * distance test (ListsForceComputation.cpp:3-145, MPListsForceComputation.cpp:
 * 3-147). Host CSR over the context's current sorted particles: offsets[n + 1]
 * (offsets[0] = 0, non-decreasing) and indices in [0, n). include_self = 1: every
 * row lists its own particle exactly once (the CPU convention; that entry is
 * dropped and the kernels add the self term first); 0: no row lists itself.
 * Violations return SPH_ERR_INVALID and leave the previous lists in place.
 * Replaces the lists of sph_build_neighbor_lists until the next one. Synchronises. */
int sph_set_neighbor_lists(void* ctx, const int64_t* offsets, const int32_t* indices,
                           int32_t include_self);
/* cuCallbackApproximateDensitySorted + cuInvertBuffer + cuKernelComputePressureSorted
 * (cuda_callbacks.cuh:40-46; cudaLaunchers.cuh:17-53), or the list variant
 * cuKernelApproximateDensityWithNeighborLists for SPH_COMPUTE_LISTS. */
int sph_density_pressure(void* ctx, int32_t compute);
/* cuCallbackComputeForcesSorted (cuda_callbacks.cuh:28-38) or
 * cuKernelComputeForcesWithNeighborLists (cudaLaunchers.cuh:23-33). */
int sph_forces(void* ctx, int32_t compute);
/* cuCallbackIntegrate (core/acceleration/integrator/cuIntegratorLauncher.cuh:23-28). */
int sph_integrate(void* ctx);
/* cuKernelUpdateVelocities / cuKernelUpdatePositions (cuIntegratorLauncher.cuh:5-21):
 * the two halves of sph_integrate, v += dt a, then r += dt v with the boundary. */
int sph_update_velocities(void* ctx);
int sph_update_positions(void* ctx);
/* Time step of the integration stages (the dt argument of cuCallbackIntegrate). */
int sph_set_dt(void* ctx, double dt);
/* cuKernelMeasurementsPerStep (core/acceleration/utils/cuUtilsLauncher.cuh:5-9) in
 * FP64: out[0] kinetic energy, out[1] max |v|^2, out[2] particles inside the
 * domain, out[3] potential energy, out[4] total mass. Synchronises. */
int sph_measure(void* ctx, double* out5);

/* Allocation-free sph_measure over caller device arrays. It has no
 * counterpart in the reference; it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). No context: rx, ry, vx, vy are n device floats, out5 is 5
 * device doubles (the values of sph_measure, bit for bit), params supplies
 * mass, gravity_y and the domain. Runs on the calling thread's current
 * device on `stream` (a hipStream_t; NULL = the null stream), asynchronously,
 * and calls no hipMalloc/hipFree: the partial sums live in a static 1 MiB
 * device array per GPU, so n <= 8,388,608 (larger n: SPH_ERR_CAPACITY).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * race. Errors: sph_last_error(NULL). */
int sph_measure_noalloc(const sph_params* params, int64_t n, const float* rx, const float* ry,
                        const float* vx, const float* vy, double* out5, void* stream);

/* Time step from the CFL and force criteria with FP64 maxima over the current
 * state (velocities and the accelerations of the last force evaluation):
 *   dt = min(cfl * h / (c0 + max|v|), 0.25 * sqrt(h / max|a|))
 * (the first term is the fixed dt = cfl h / c0 of a fluid at rest).
 * out3 = {dt, max|v|, max|a|}. Does not apply it: pass dt to sph_set_dt. The
 * reference runs with a fixed dt; this backs adaptive stepping. Synchronises. */
int sph_cfl(void* ctx, double cfl, double* out3);

/* ---- binary snapshots (replace positions_step_<k>.txt of
 *      core/utils/CheckPointWriter.cpp:12-27, which keeps positions only) ----
 * Little-endian file: "SPHSNAP1", uint32 version (1), uint32 header bytes (192),
 * int64 n, int64 step, sph_params (144 bytes), uint32 flags (1: arrays in the
 * context's sorted order), uint32 reserved, uint64 FNV-1a of the arrays; then
 * float32 rx[n], ry[n], vx[n], vy[n], uint32 pid[n]. Saving the sorted order
 * with the ids makes a resumed run continue bit-identically (sph_prepare after
 * the load re-sorts an already-sorted state). Single-scene contexts only. */
int sph_snapshot_info(const char* path, sph_params* params, int64_t* n, int64_t* step);
int sph_snapshot_save(void* ctx, const char* path, int64_t step);
/* Needs a context created with the file's parameters (compared byte for byte,
 * except dt, which the load applies) and n <= its capacity; leaves it as after
 * sph_upload. */
int sph_snapshot_load(void* ctx, const char* path, int64_t* step);

/* ---- fused pipeline (no reference counterpart; same arithmetic as the stages) ----
 * sph_prepare: find neighbors + density with the fixed-width neighbor list.
 * sph_step: n iterations of [density + forces + integration + keys] -> radix
 *   sort -> [permutation + cell ranges]; density and forces are left for the
 *   state before the last integration (written by the last iteration only;
 *   sph_density_pressure + sph_forces refresh).
 * Reference-equivalent run of S iterations (simulation_configuration.h:297-387):
 *   sph_prepare; sph_step(S-1); sph_density_pressure(SORTED); sph_forces(SORTED).
 * By default (SPH_STEP_PARTICLE, SPH_RESORT_INCREMENTAL) one iteration is four
 * launches: the density kernel (density + EOS + fixed-width list), the force
 * kernel (forces + integration + next keys + the movers' flags, each mover put
 * into the bucket of its merge tile), the merge of the incremental re-sort
 * (each tile's bucket ranked in LDS and merged with the tile's other
 * particles) and the cell-range kernel. A multi-step call writes the separate
 * state arrays back when a later call first needs them, unless a device view
 * was handed out. */
int sph_prepare(void* ctx);
int sph_step(void* ctx, int32_t n_steps);
/* Capture fused iterations in a hipGraph and replay it in sph_step: enable = 0
 * off, k >= 1 (at most 64) a graph of 2k iterations; the even part of the steps
 * left over replays one shorter graph, an odd last step runs without. Graphs
 * are kept per buffer assignment (the pid and key buffers alternate every
 * step) and iteration count, up to four at once. */
int sph_set_graph(void* ctx, int32_t enable);
/* Capture, without running anything, the graphs that sph_step(ctx, n_steps)
 * would replay from the current state, so that call spends no time capturing
 * (no reference counterpart: for callers that time or schedule sph_step). */
int sph_step_graphs(void* ctx, int32_t n_steps);
/* Number of particles whose neighbor count exceeded the fixed list width in the
 * last density pass (they are handled by a stencil fallback). Synchronises. */
int sph_ell_overflow(void* ctx, int64_t* count);
/* Choose the physics kernels of sph_step (enum sph_step_kernel). */
int sph_set_step_kernel(void* ctx, int32_t variant);
/* How sph_step re-sorts (enum sph_resort); both give the same permutation. */
int sph_set_resort(void* ctx, int32_t mode);
/* Particles whose cell key changed in the last incremental re-sort. Synchronises. */
int sph_movers(void* ctx, int64_t* count);
/* Which way the last incremental re-sort ordered its movers: 1 through the
 * per-tile buckets (every tile's movers within its bucket capacity), 0 through
 * the sorted-movers fallback. Synchronises. No counterpart in the reference
 * (its sort is a full std::sort / thrust sort every step). */
int sph_resort_path(void* ctx, int32_t* path);
/* Re-evaluate the block kernels on the current state with counters on:
 * out[0..2] density (or fused) kernel: occupied blocks, blocks over the LDS
 * capacity, staged particles; out[3..5] the same for the force kernel. */
int sph_block_stats(void* ctx, int64_t* out6);

/* ---- batched scenes (no reference counterpart: the reference runs one scene
 * per process) ----
 * A batch context holds n_scenes independent particle systems that share
 * physics, boundary, order and ell_width but each have their own box, h, dt,
 * mass and derived constants. The scenes' keys live in one key space (scene s
 * owns the key range [base_s, base_s + its cell-table size), base_s the sum of
 * the earlier scenes' table sizes each rounded up to even), so one radix sort
 * per step serves all scenes and each scene keeps a fixed, contiguous segment
 * [first_s, first_s + count_s) of the sorted order; every workgroup of the
 * physics kernels covers one tile of one segment. Upload and download take the
 * scenes' particles concatenated in scene order (sph_upload with n = the total
 * count); particle ids are global. On a batch context these calls work:
 * sph_upload, sph_download(_fields), sph_size, sph_get_grid (the shared key
 * space), sph_find_neighbors, sph_prepare, sph_step, sph_set_graph,
 * sph_set_resort, sph_density_pressure and sph_forces (both through the
 * fixed-width list, same arithmetic as SORTED), sph_get_sorted_keys,
 * sph_get_permutation, sph_get_cell_ranges, sph_get_ell, sph_movers, timing and
 * sph_synchronize; the others return SPH_ERR_STATE. Every scene's trajectory is
 * bit-identical to a single-scene context with the same parameters and
 * cell_size = cell_factor * h. */
typedef struct sph_batch_tuning {
  double cell_factor; /* cell size = cell_factor * h of every scene; >= 1 */
  int32_t tile;       /* particles per workgroup: 64, 128, 256 or 512 */
  int32_t resort;     /* enum sph_resort used by sph_step */
} sph_batch_tuning;

/* scenes[0..n_scenes) with counts[s] >= 1 particles each; the scenes' cell_size
 * fields are ignored (the tuning sets the cell size). Starts with tuning
 * {1.0, 256, SPH_RESORT_INCREMENTAL}. */
int sph_batch_create(const sph_params* scenes, int32_t n_scenes, const int64_t* counts,
                     int device, void** out_ctx);
int32_t sph_batch_scenes(const void* ctx); /* 0 for a single-scene context */
/* Segment of scene s in the concatenated / sorted order and its grid. */
int sph_batch_scene_info(const void* ctx, int32_t scene, int64_t* first, int64_t* count,
                         sph_grid_info* grid);
/* Apply a tuning: rebuilds the key space and tiles; the state is kept but must be
 * re-sorted, so call sph_prepare before sph_step. */
int sph_batch_set_tuning(void* ctx, const sph_batch_tuning* tuning);
int sph_batch_get_tuning(const void* ctx, sph_batch_tuning* tuning);
/* The built-in candidate set of the autotuner (cell_factor x tile x resort);
 * returns the number of candidates (writes at most cap). */
int32_t sph_batch_default_candidates(sph_batch_tuning* out, int32_t cap);
/* Autotuner: for each candidate, from a device copy of the current state,
 * applies it, prepares, runs 2 warm-up steps and times trial_steps steps of
 * graph-replayed sph_step with hipEvents; ms_per_step[i] receives the time (-1
 * for a candidate that cannot be applied). Applies the fastest, restores the
 * state it started from, and leaves the context as after sph_upload (call
 * sph_prepare next). *best = index of the chosen candidate. */
int sph_batch_autotune(void* ctx, const sph_batch_tuning* candidates, int32_t n_candidates,
                       int32_t trial_steps, double* ms_per_step, int32_t* best);
/* sph_measure per scene: out[5 * s + k] as out5 of sph_measure. Synchronises. */
int sph_batch_measure(void* ctx, double* out);

/* ---- device-resident state exchange (no reference counterpart) ----
 * For data pipelines that keep snapshots in device memory: export writes the
 * current state into caller device memory out[4 n] = rx | ry | vx | vy in
 * particle-id order; upload_device replaces the state from caller device arrays
 * (ids 0..n-1), like sph_upload. Both synchronise the context stream. */
int sph_export_state(void* ctx, float* d_out);
/* Device memory for such pipelines (hipMalloc / hipFree / hipMemcpy on a device;
 * kind: 0 host->device, 1 device->host, 2 device->device). */
int sph_device_malloc(int device, int64_t bytes, void** out_ptr);
int sph_device_free(void* ptr);
int sph_device_copy(void* dst, const void* src, int64_t bytes, int32_t kind);
int sph_upload_device(void* ctx, int64_t n, const float* d_rx, const float* d_ry,
                      const float* d_vx, const float* d_vy);

/* ---- learned surrogate (no reference counterpart) ----
 * A message-passing network over the fixed-width neighbor lists predicts each
 * particle's acceleration from invariant features (distances, relative
 * velocities, densities and p / rho^2 of up to ell_width neighbors) with an
 * equivariant readout (separation vectors times learned edge scalars). Node
 * linear layers run as BF16 MFMA GEMMs with FP32 accumulation, edge layers in
 * FP32, training with Adam on FP32 master weights. A network handle is tied to
 * no context; every compute call takes the context whose sorted state (after
 * sph_prepare / sph_density_pressure) it reads and whose stream it uses. */
typedef struct sph_nn_config {
  int32_t width;     /* hidden width; must be 128 */
  int32_t layers;    /* message-passing layers, 1..8 */
  int32_t train;     /* 1: keep per-layer activations for backpropagation */
  int32_t periodic;  /* minimum-image separations in a periodic box */
  int64_t max_nodes; /* largest particle count of a call */
  double h;          /* smoothing length of the scenes it is used on */
  double rho0, c0;   /* density and pressure feature scales */
  double v_scale;    /* relative-velocity feature scale */
  double a_scale;    /* acceleration scale of the readout */
  double gravity_x, gravity_y; /* added to the readout */
  double len_x, len_y;         /* periodic box lengths (periodic = 1) */
} sph_nn_config;

int sph_nn_create(const sph_nn_config* cfg, int device, void** out_nn);
int sph_nn_destroy(void* nn);
const char* sph_nn_last_error(const void* nn); /* nn may be NULL (creation errors) */
int64_t sph_nn_num_params(const void* nn);
/* FP32 parameters in the order of the network's parameter layout. */
int sph_nn_set_params(void* nn, const float* host, int64_t count);
int sph_nn_get_params(void* nn, float* host, int64_t count);
int sph_nn_get_grad(void* nn, float* host, int64_t count); /* last sph_nn_loss_grad */
int sph_nn_reset_optimizer(void* nn);
/* Global gradient-norm limit applied by sph_nn_adam before the update (the
 * gradient is scaled to norm max_norm when it is larger; 0, the default, turns
 * it off). The norm is computed in FP64 on the device. */
int sph_nn_set_clip(void* nn, double max_norm);
/* Predicted accelerations of the context's sorted particles; ax, ay (host, n
 * each, sorted order) may be NULL. Needs the fixed-width list (sph_prepare or
 * sph_density_pressure on a batch / after sph_prepare). */
int sph_nn_forward(void* nn, void* ctx, float* ax, float* ay);
/* Forward, loss against the context's accelerations t (call sph_forces first),
 * backward into the gradient. Loss = sum |a - t|^2 / sum |t|^2 (FP64 sums).
 * out3 = {loss, RMS relative error = sqrt(loss), RMS |t|}.
 * Needs symmetric lists: fails if a list overflowed ell_width. Synchronises. */
int sph_nn_loss_grad(void* nn, void* ctx, double* out3);
int sph_nn_adam(void* nn, void* ctx, double lr, double beta1, double beta2, double eps);
/* loss_grad + adam(lr, 0.9, 0.999, 1e-8). */
int sph_nn_train_step(void* nn, void* ctx, double lr, double* out3);

/* Surrogate in the loop: n_steps steps of the context where the network supplies
 * the accelerations; every check_every steps the plain SPH forces are computed
 * as well and the RMS relative error sqrt(sum |a_nn - a|^2 / sum |a|^2) (FP64)
 * is compared with threshold. mode 0 (step): a failing check step integrates
 * with the plain forces. mode 1 (window): a failing check restores the state at
 * the start of the window and advances the window with plain SPH. Needs
 * sph_prepare first; leaves the context prepared for sph_step. */
typedef struct sph_surrogate_stats {
  int64_t steps, checks, failed_checks, nn_steps, plain_steps;
  double last_error, max_error, mean_error;
} sph_surrogate_stats;
int sph_surrogate_run(void* ctx, void* nn, int32_t n_steps, int32_t check_every,
                      double threshold, int32_t mode, sph_surrogate_stats* stats);

/* ---- inspection (host copies; synchronise) ---- */
int sph_get_sorted_keys(void* ctx, uint64_t* keys);  /* n keys of the last sort */
int sph_get_permutation(void* ctx, uint32_t* perm);  /* sorted slot -> previous index */
int sph_get_cell_ranges(void* ctx, int32_t* start, int32_t* end); /* n_keys each */
int sph_get_neighbor_lists(void* ctx, int64_t* offsets, int32_t* indices); /* CSR */
int sph_get_ell(void* ctx, int32_t* counts, int32_t* slots); /* n, n*ell_width */
int sph_get_device_view(void* ctx, sph_device_view* out);

/* ---- timing (hipEvents on the context stream) ---- */
int sph_set_timing(void* ctx, int32_t enable);
/* Accumulated milliseconds per enum sph_stage and launch counts; resets them. */
int sph_stage_times(void* ctx, double* ms, int64_t* launches, int32_t count);
/* Waits for the context stream and reads the device status word (through a
 * pinned host word): SPH_ERR_HIP when a bounded wait on the device gave up
 * (radix-sort look-back, re-sort merge) since the context was created. About
 * 12 us on an idle stream; a caller timing only the steps can wait with
 * hipDeviceSynchronize and call this afterwards. */
int sph_synchronize(void* ctx);

#ifdef __cplusplus
}
#endif
#endif /* SPH_AMD_H */
