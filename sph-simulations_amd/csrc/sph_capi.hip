// C-ABI of libsph_amd.so: context lifetime, buffers, stage orchestration,
// hipGraph capture of the fused iteration, and hipEvent stage timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sph_amd.h"
#include "sph_internal.hpp"
#include "sph_nn.hpp"

using namespace sphamd;

namespace {

thread_local std::string g_create_error;
// message of the last failed context-free call that must not allocate (static text)
thread_local const char* g_static_error = nullptr;

struct PendingTiming {
  int stage;
  hipEvent_t a, b;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  sph_params prm{};
  sph_grid_info grid{};
  Launch L{};
  int64_t n_max = 0;
  DevState s{};
  SortBuffers sb{};
  std::vector<void*> allocs;
  uint32_t* host_status = nullptr;  // pinned: the device status word's copy
  int64_t csr_cap = 0, csr_total = -1;
  int csr_self = -1;
  int64_t* scan_scratch = nullptr;
  double* meas_part = nullptr;
  double* meas_out = nullptr;
  int64_t* ovf = nullptr;
  std::string err;
  // state flags
  bool sorted = false;         // S is in the order of the last sort, ranges valid
  bool density_ell = false;    // rho/pr/ELL valid for S
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<PendingTiming> pending;
  double stage_ms[SPH_T_COUNT] = {};
  int64_t launches[SPH_T_COUNT] = {};
  // inside sph_step on the per-particle kernels: the re-sort writes only the
  // packed copy of the state, the separate arrays are restored when it ends
  bool soa_lazy = false;
  // the packed copy (and, with packed_ok_narrow, the narrow keys) still holds
  // the sorted state a lazy sph_step call left: the next call needs no k_pack.
  // Every entry point but the step calls and a few that only read clears it
  // (as_ctx); never set once a device view was handed out (the caller may then
  // write the state between calls without an API call)
  bool packed_ok = false;
  bool packed_ok_narrow = false;
  bool view_out = false;
  // after a lazy sph_step call (no device view out) the separate state
  // arrays, and with soa_stale_narrow the u64 sorted keys, are written back
  // from the packed copy only when a later call needs them (settle_soa, from
  // as_ctx / as_ctx_read, or the next step call of another kind)
  bool soa_stale = false;
  bool soa_stale_narrow = false;
  // the fused step writes the accelerations, the separate density and
  // pressure arrays and the incremental re-sort's permutation (off for the
  // steps of a sph_step call but its last: only the last step's are
  // observable; the steps read the packed copies)
  bool fields_out = true;
  // ax, ay, rho, pr and p hold sph_step's last iteration in the order before its
  // final re-sort; settle_fields moves them into the current order
  bool fields_prev_order = false;
  // graph
  bool use_graph = false;
  int graph_pairs = 16;  // iteration pairs a captured graph holds
  // captured graphs, each for one buffer assignment and iteration count (the
  // pid and key buffers swap every iteration, so a call that starts after an
  // odd number of steps needs the other assignment's graph)
  struct GraphEntry {
    const void* sig[10];  // buffer assignment the graph was captured with
    int iters;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;
  int graph_n = -1;
  int block_grid = 0;             // persistent workgroups of the block kernels
  int step_kernel = SPH_STEP_PARTICLE;
  int resort = SPH_RESORT_INCREMENTAL;  // how sph_step re-sorts
  ResortBuffers rb{};
  unsigned long long* block_stats = nullptr;
  // batched scenes (n_scenes == 0: single-scene context)
  int n_scenes = 0;
  std::vector<sph_params> scene_prm;
  std::vector<int64_t> scene_first, scene_count;
  std::vector<sph_grid_info> scene_grid;
  std::vector<SceneDev> scene_dev;
  sph_batch_tuning tuning{1.0, 256, SPH_RESORT_INCREMENTAL};
  SceneDev* d_scenes = nullptr;
  TileDev* d_tiles = nullptr;
  int64_t tiles_cap = 0;
  int64_t cells_cap = 0;
  int passes_cap = 0;
  BatchLaunch B{};
  float* backup = nullptr;  // autotuner's copy of rx, ry, vx, vy, ax, ay (6 n_max)
  uint32_t* backup_pid = nullptr;
};

// Buffer assignment the fused iteration reads and writes; the stages swap
// these pointers, so a captured graph is only valid for the assignment it saw.
void graph_signature(const Ctx* c, const void** sig) {
  const DevState& s = c->s;
  const void* v[10] = {s.rx, s.tx, s.vx, s.tvx, s.ax, s.tax, s.pid, s.pid_alt,
                       s.narrow ? (const void*)s.nkeys : (const void*)s.keys_sorted,
                       reinterpret_cast<const void*>((intptr_t)s.n * 2 + (s.narrow ? 1 : 0))};
  for (int i = 0; i < 10; ++i) sig[i] = v[i];
}

// The cached graph of `iters` iterations for the current buffer assignment, or
// null.
hipGraphExec_t find_graph(const Ctx* c, int iters) {
  const void* sig[10];
  graph_signature(c, sig);
  for (const Ctx::GraphEntry& g : c->graphs) {
    if (g.iters != iters) continue;
    bool same = true;
    for (int i = 0; i < 10; ++i) same = same && sig[i] == g.sig[i];
    if (same) return g.exec;
  }
  return nullptr;
}

int fail(Ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(Ctx* c, hipError_t e, const char* what) {
  return fail(c, SPH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(ctx, expr)                                  \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
  } while (0)

int next_pow2(int n) {
  int p = 1;
  while (p < n) p *= 2;
  return p;
}

uint64_t morton_host(uint32_t x, uint32_t y) {
  uint64_t z = 0;
  for (int i = 0; i < 32; ++i) {
    z |= (uint64_t)((x >> i) & 1u) << (2 * i);
    z |= (uint64_t)((y >> i) & 1u) << (2 * i + 1);
  }
  return z;
}

// Grid of MortonOrder::Initialize / RowMajorOrder::Initialize
// (core/alg/orders/MortonOrder.h:15-47, RowMajorOrder.h:13-25): ceil(L / cs) cells
// per axis in float, Morton extents padded to powers of two; the table is sized
// by the largest Morton code. Periodic boxes use floor(L / h) cells of L / n.
int derive_grid(const sph_params& p, sph_grid_info& g, std::string& why) {
  float lo_x = (float)p.lo[0], lo_y = (float)p.lo[1];
  float len_x = (float)p.hi[0] - lo_x, len_y = (float)p.hi[1] - lo_y;
  float cs = (float)(p.cell_size > 0 ? p.cell_size : p.h);
  if (!(p.h > 0) || !(cs > 0) || !(len_x > 0) || !(len_y > 0)) {
    why = "h, cell size and box lengths must be positive";
    return -1;
  }
  if (p.cell_size > 0 && p.cell_size < p.h) {
    why = "cell_size must be >= h (the 3x3 stencil must cover the support)";
    return -1;
  }
  int nx, ny;
  float csx = cs, csy = cs;
  if (p.boundary == SPH_BC_PERIODIC) {
    nx = (int)std::floor(len_x / cs);
    ny = (int)std::floor(len_y / cs);
    if (nx < 3 || ny < 3) {
      why = "periodic boxes need at least 3 cells per axis";
      return -1;
    }
    csx = len_x / (float)nx;
    csy = len_y / (float)ny;
  } else {
    nx = (int)std::ceil(len_x / cs);
    ny = (int)std::ceil(len_y / cs);
  }
  if (nx < 1 || ny < 1 || nx > (1 << 30) || ny > (1 << 30)) {
    why = "cell grid out of range";
    return -1;
  }
  g.cs_x = csx; g.cs_y = csy;
  g.lo_x = lo_x; g.lo_y = lo_y;
  g.len_x = len_x; g.len_y = len_y;
  g.nx = nx; g.ny = ny;
  g.order = p.order;
  if (p.order == SPH_ORDER_MORTON) {
    g.nxp = next_pow2(nx);
    g.nyp = next_pow2(ny);
    g.n_keys = (int64_t)morton_host((uint32_t)(g.nxp - 1), (uint32_t)(g.nyp - 1)) + 1;
  } else {
    g.nxp = nx; g.nyp = ny;
    g.n_keys = (int64_t)nx * ny;
  }
  if (g.n_keys > (int64_t)1 << 31) {
    why = "cell table larger than 2^31 entries";
    return -1;
  }
  int bits = 0;
  while (bits < 63 && ((uint64_t)1 << bits) < (uint64_t)g.n_keys) ++bits;
  g.key_bits = std::max(bits, 1);
  return 0;
}

// Derived constants; the same double expressions as the oracle, rounded once.
void derive_phys(const sph_params& p, const sph_grid_info& g, DevPhys& c, DevGrid& dg) {
  const double PI = 3.14159265358979323846;
  double h = p.h, h2 = h * h, h4 = h2 * h2, h5 = h4 * h, h6 = h4 * h2, h8 = h4 * h4,
         h9 = h8 * h;
  c.h = (float)h;
  c.h2 = c.h * c.h;
  c.tol = 1.0e-12f;
  c.physics = p.physics;
  c.boundary = p.boundary;
  c.gamma = p.gamma;
  c.gx = 0.0f;
  c.gy = (float)p.gravity_y;
  c.dt = (float)p.dt;
  c.rho0 = (float)p.rho0;
  c.eos_k = (float)p.eos_k;
  if (p.physics == SPH_PHYS_REFERENCE) {
    double c_rho = 315.0 / (64.0 * PI * h9);  // poly6, core/Kernel.cpp:8-10
    double cp = 15.0 / (PI * h6);             // spiky, core/Kernel.cpp:38-40
    c.mc_rho = (float)(p.mass * c_rho);
    c.c_pres = (float)(1.5 * cp * p.mass);
    c.c_visc = (float)(45.0 * p.mu / (PI * h6));  // cuParticles.cuh:51
    c.inv_rho0 = 0.0f;
    c.tait_b = 0.0f;
  } else {
    double c_rho = 4.0 / (PI * h8);  // poly6 normalised in 2D
    double cg = 30.0 / (PI * h5);    // spiky gradient normalised in 2D
    c.mc_rho = (float)(p.mass * c_rho);
    c.c_pres = (float)(p.mass * cg);
    c.c_visc = 0.0f;
    c.inv_rho0 = (float)(1.0 / p.rho0);
    c.tait_b = (float)(p.rho0 * p.c0 * p.c0 / (double)p.gamma);
  }
  c.neg_alpha_c0 = (float)(-p.alpha * p.c0);
  c.eps_h2 = (float)(p.eps_visc * h2);
  c.lo_x = (float)p.lo[0];
  c.lo_y = (float)p.lo[1];
  c.hi_x = (float)p.hi[0];
  c.hi_y = (float)p.hi[1];
  c.len_x = c.hi_x - c.lo_x;
  c.len_y = c.hi_y - c.lo_y;
  c.half_x = c.len_x * 0.5f;
  c.half_y = c.len_y * 0.5f;
  dg.lo_x = g.lo_x;
  dg.lo_y = g.lo_y;
  dg.cs_x = g.cs_x;
  dg.cs_y = g.cs_y;
  dg.nx = g.nx;
  dg.ny = g.ny;
  dg.order = g.order;
  dg.periodic = p.boundary == SPH_BC_PERIODIC ? 1 : 0;
}

template <class T>
int dev_alloc(Ctx* c, T** p, int64_t count) {
  void* q = nullptr;
  size_t bytes = (size_t)std::max<int64_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return hip_fail(c, e, "hipMalloc");
  c->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return SPH_OK;
}

int settle_soa(Ctx* c);

// The context of an entry point that may read or change the state: the
// separate arrays are brought up to date first, and the packed copy no longer
// counts as current. Null when bringing them up to date failed (the reason is
// in sph_last_error).
Ctx* as_ctx(void* p) {
  Ctx* c = static_cast<Ctx*>(p);
  if (!c) return nullptr;
  c->packed_ok = false;
  return settle_soa(c) == SPH_OK ? c : nullptr;
}
// The context of a call that reads the state and changes nothing.
Ctx* as_ctx_read(void* p) {
  Ctx* c = static_cast<Ctx*>(p);
  if (!c) return nullptr;
  return settle_soa(c) == SPH_OK ? c : nullptr;
}
// The context of the step calls and of calls that touch no state array.
Ctx* as_ctx_keep(void* p) { return static_cast<Ctx*>(p); }

// ---- timing helpers
hipEvent_t take_event(Ctx* c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct StageTimer {
  Ctx* c;
  int stage;
  hipEvent_t a = nullptr;
  StageTimer(Ctx* ctx, int st) : c(ctx), stage(st) {
    if (c->timing) {
      a = take_event(c);
      if (a) (void)hipEventRecord(a, c->stream);
    }
  }
  ~StageTimer() {
    if (!a) return;
    hipEvent_t b = take_event(c);
    if (!b) return;
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back({stage, a, b});
  }
};

int collect_timings(Ctx* c) {
  if (c->pending.empty()) return SPH_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto& t : c->pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      c->stage_ms[t.stage] += ms;
      c->launches[t.stage] += 1;
    }
    c->ev_pool.push_back(t.a);
    c->ev_pool.push_back(t.b);
  }
  c->pending.clear();
  return SPH_OK;
}

// ---- pipeline pieces
int sort_keys(Ctx* c) {
  SortBuffers b = c->sb;
  b.clear_n = c->s.n_prev;
  {
    StageTimer t(c, SPH_T_SORT_HIST);
    HIPCHK(c, radix_sort_hist(b, c->s.n, c->grid.key_bits, c->stream));
  }
  for (int p = 0; p < sort_passes(c->grid.key_bits); ++p) {
    StageTimer t(c, SPH_T_SORT_PASS);
    HIPCHK(c, radix_sort_pass(b, c->s.n, c->grid.key_bits, p, c->stream));
  }
  return SPH_OK;
}

int hist_words(Ctx* c) { return sort_passes(c->grid.key_bits) * 256; }

int settle_fields(Ctx* c);

// Standalone find-neighbors: keys of S, sort, permute S (r, v, a, pid) into T,
// ranges, then swap S and T.
int find_neighbors(Ctx* c) {
  DevState& s = c->s;
  int rc0 = settle_fields(c);  // the gather permutes a and uses the T arrays
  if (rc0) return rc0;
  {
    StageTimer t(c, SPH_T_NEIGHBOR_SEARCH);
    if (c->n_scenes)
      HIPCHK(c, launch_batch_keys(c->L, c->B, s, c->sb.keys_in));
    else
      HIPCHK(c, launch_keys(c->L, s, c->sb.keys_in));
  }
  int rc = sort_keys(c);
  if (rc) return rc;
  {
    StageTimer t(c, SPH_T_NEIGHBOR_SEARCH);
    HIPCHK(c, launch_gather(c->L, s, s.rx, s.ry, s.vx, s.vy, s.pid, s.ax, s.ay, s.tx, s.ty,
                            s.tvx, s.tvy, s.pid_alt, s.tax, s.tay, hist_words(c)));
  }
  std::swap(s.rx, s.tx);
  std::swap(s.ry, s.ty);
  std::swap(s.vx, s.tvx);
  std::swap(s.vy, s.tvy);
  std::swap(s.ax, s.tax);
  std::swap(s.ay, s.tay);
  std::swap(s.pid, s.pid_alt);
  s.n_prev = s.n;
  c->sorted = true;
  c->density_ell = false;
  c->csr_total = -1;
  return SPH_OK;
}

// Re-sort of the integrated state T (keys in sb.keys_in) against the previous
// sorted order: mover flags -> sorted movers (one launch) -> merge/scatter
// T -> S -> ranges. Same permutation as sort_keys + gather.
// flags_ready: the producer of the new keys wrote the mover flags.
int incremental_resort(Ctx* c, bool flags_ready, bool bucketed) {
  DevState& s = c->s;
  ResortBuffers& r = c->rb;
  if (!bucketed || !kMergePrefix) {
    StageTimer t(c, SPH_T_MOVERS);
    HIPCHK(c,
           launch_movers(c->L, s, r, c->sb.keys_in, c->grid.key_bits, flags_ready, bucketed));
  }
  {
    // after a bucketing producer the merge also forms the chunk counters and
    // runs the sorted path (no launch between the force kernel and the merge)
    StageTimer t(c, SPH_T_MERGE);
    const bool pub = bucketed && kMergePrefix;
    HIPCHK(c, launch_merge(c->L, s, r, r.keys_alt, !c->soa_lazy, c->fields_out,
                           pub ? c->sb.keys_in : nullptr, c->grid.key_bits));
  }
  if (s.narrow) {
    std::swap(s.nkeys, s.nkeys_alt);
  } else {
    std::swap(s.keys_sorted, r.keys_alt);
    c->sb.keys_out = s.keys_sorted;
  }
  std::swap(s.pid, s.pid_alt);
  {
    StageTimer t(c, SPH_T_NEIGHBOR_SEARCH);
    HIPCHK(c, launch_ranges(c->L, s, r, hist_words(c)));
  }
  s.n_prev = s.n;
  return SPH_OK;
}

int resort_integrated(Ctx* c, bool flags_ready = false, bool bucketed = false);

// One fused iteration: forces(ELL) + integrate + keys -> sort -> gather T->S
// + ranges -> density + EOS + ELL.
int fused_iteration(Ctx* c) {
  DevState& s = c->s;
  if (c->n_scenes) {
    // batched scenes: per-tile scene constants, fixed-width list
    DevState fs = s;
    if (!c->fields_out) fs.ax = fs.ay = fs.rho = fs.pr = nullptr;
    {
      StageTimer t(c, SPH_T_DENSITY);
      HIPCHK(c, launch_batch_density(c->L, c->B, fs));
    }
    // the force kernel also flags the movers for the incremental re-sort and
    // puts them into their owner tiles' buckets
    const bool inc = c->resort == SPH_RESORT_INCREMENTAL;
    const bool bucketed = inc && kFusedBucket && s.n <= kResortMaxParticles;
    {
      StageTimer t(c, SPH_T_FORCE);
      HIPCHK(c, launch_batch_forces(c->L, c->B, fs, c->sb.keys_in,
                                    !inc       ? MoverSink{}
                                    : bucketed ? bucket_sink(c->rb, s, (int)resort_tiles(s.n))
                                               : mover_sink(c->rb)));
    }
    return resort_integrated(c, inc, bucketed);
  } else if (s.block_bits && c->step_kernel == SPH_STEP_FUSED_BLOCK) {
    // Morton order, one kernel: density, forces, integration and keys
    StageTimer t(c, SPH_T_BLOCK_STEP);
    HIPCHK(c, launch_block_step(c->L, s, c->sb.keys_in, c->block_grid, c->step_kernel, nullptr));
  } else if (s.block_bits && c->step_kernel == SPH_STEP_SPLIT_BLOCK) {
    // Morton order: block density kernel, then block force + integration + keys
    {
      StageTimer t(c, SPH_T_DENSITY);
      HIPCHK(c, launch_block_step(c->L, s, c->sb.keys_in, c->block_grid, c->step_kernel,
                                  nullptr, 1));
    }
    StageTimer t(c, SPH_T_FORCE);
    HIPCHK(c, launch_block_step(c->L, s, c->sb.keys_in, c->block_grid, c->step_kernel, nullptr,
                                2));
  } else {
    // per-particle kernels: density with the fixed-width list, then forces from
    // it; neighbor data gathered from the packed copy s.xyuv (kept by the
    // re-sorts, made by sph_step on entry) and s.rp (written by the density)
    DevState fs = s;
    if (!c->fields_out) fs.ax = fs.ay = fs.rho = fs.pr = nullptr;
    {
      StageTimer t(c, SPH_T_DENSITY);
      HIPCHK(c, launch_density(c->L, fs, NB_ELL, false, true));
    }
    // the force kernel also flags the movers for the incremental re-sort and
    // puts them into their owner tiles' buckets
    const bool inc = c->resort == SPH_RESORT_INCREMENTAL;
    const bool bucketed = inc && kFusedBucket && s.n <= kResortMaxParticles;
    {
      StageTimer t(c, SPH_T_FORCE);
      HIPCHK(c, launch_forces(c->L, fs, NB_ELL, true, c->sb.keys_in,
                              !inc       ? MoverSink{}
                              : bucketed ? bucket_sink(c->rb, s, (int)resort_tiles(s.n))
                                         : mover_sink(c->rb),
                              true));
    }
    return resort_integrated(c, inc, bucketed);
  }
  return resort_integrated(c);
}

// Re-sort the integrated twin T (keys in sb.keys_in) into S with new ranges.
int resort_integrated(Ctx* c, bool flags_ready, bool bucketed) {
  DevState& s = c->s;
  if (c->resort == SPH_RESORT_INCREMENTAL && s.n <= kResortMaxParticles)
    return incremental_resort(c, flags_ready, bucketed);
  int rc = sort_keys(c);
  if (rc) return rc;
  {
    StageTimer t(c, SPH_T_NEIGHBOR_SEARCH);
    HIPCHK(c, launch_gather(c->L, s, s.tx, s.ty, s.tvx, s.tvy, s.pid, nullptr, nullptr, s.rx,
                            s.ry, s.vx, s.vy, s.pid_alt, nullptr, nullptr, hist_words(c)));
  }
  std::swap(s.pid, s.pid_alt);
  s.n_prev = s.n;
  return SPH_OK;
}

void destroy_graph(Ctx* c) {
  for (Ctx::GraphEntry& g : c->graphs) {
    (void)hipGraphExecDestroy(g.exec);
    (void)hipGraphDestroy(g.graph);
  }
  c->graphs.clear();
  c->graph_n = -1;
}

// Graphs kept at once: both buffer assignments of the full-length graph and of
// a shorter one for a call's remaining steps.
constexpr size_t kGraphCache = 4;

// A graph of `iters` fused iterations with fixed buffer pointers. The pid and
// key buffers swap once per iteration, so `iters` is even: an even number of
// swaps returns every pointer to where it started, and capturing runs nothing.
int capture_graph(Ctx* c, int iters, hipGraphExec_t* out) {
  if (c->graph_n != c->s.n || c->graphs.size() >= kGraphCache) destroy_graph(c);
  bool was_timing = c->timing;
  c->timing = false;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  // only the last iteration writes the observable fields
  int rc = SPH_OK;
  for (int it = 0; it < iters && rc == SPH_OK; ++it) {
    c->fields_out = it + 1 == iters;
    rc = fused_iteration(c);
  }
  c->fields_out = true;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  c->timing = was_timing;
  if (rc != SPH_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return hip_fail(c, e, "hipGraphInstantiate");
  }
  Ctx::GraphEntry ent{};
  graph_signature(c, ent.sig);
  ent.iters = iters;
  ent.graph = g;
  ent.exec = x;
  c->graphs.push_back(ent);
  c->graph_n = c->s.n;
  *out = x;
  return SPH_OK;
}

// The graph of `iters` iterations for the current buffer assignment, captured
// if it is not cached.
int get_graph(Ctx* c, int iters, hipGraphExec_t* out) {
  if ((*out = find_graph(c, iters))) return SPH_OK;
  return capture_graph(c, iters, out);
}

// Synchronise and report a radix-sort look-back that gave up (see sph_sort.hip).
int sync_and_check(Ctx* c) {
  *c->host_status = 0u;
  HIPCHK(c, hipMemcpyAsync(c->host_status, c->sb.status, sizeof(uint32_t),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const uint32_t st = *c->host_status;
  if (st != 0)
    return fail(c, SPH_ERR_HIP,
                std::string(st & 8u ? "incremental re-sort merge" : "radix sort look-back") +
                    " gave up waiting (device status " + std::to_string(st) + ")");
  return SPH_OK;
}

// The first call after sph_step that reads or changes the fields (or the T
// arrays, which are free between calls) puts them into the current order,
// so that index i of every array is the same particle. Costs one gather of
// five arrays, once; the steps themselves never pay for it.
// The separate state arrays (and the u64 sorted keys after narrow steps)
// from the packed copy a lazy sph_step call left: one pass, once.
int settle_soa(Ctx* c) {
  if (!c->soa_stale) return SPH_OK;
  c->soa_stale = false;
  DevState s = c->s;
  s.narrow = c->soa_stale_narrow;
  const hipError_t e = launch_unpack(c->L, s);
  return e == hipSuccess ? SPH_OK : hip_fail(c, e, "unpack");
}

int settle_fields(Ctx* c) {
  if (!c->fields_prev_order) return SPH_OK;
  c->fields_prev_order = false;
  DevState& s = c->s;
  float* tmp[5] = {s.tx, s.ty, s.tvx, s.tvy, s.tax};
  HIPCHK(c, launch_settle_fields(c->L, s, c->n_scenes != 0, tmp));
  float* dst[5] = {s.ax, s.ay, s.rho, s.pr, s.pres};
  const size_t fb = (size_t)s.n * sizeof(float);
  for (int k = 0; k < 5; ++k)
    HIPCHK(c, hipMemcpyAsync(dst[k], tmp[k], fb, hipMemcpyDeviceToDevice, c->stream));
  return SPH_OK;
}

int single_only(const Ctx* c) {
  return fail(const_cast<Ctx*>(c), SPH_ERR_STATE,
              "not available on a batch context (single-scene call)");
}

int check_n(Ctx* c) {
  if (c->s.n <= 0) return fail(c, SPH_ERR_STATE, "no particles uploaded");
  return SPH_OK;
}

// Stream and every device buffer of a context for n_max particles and a cell
// table of n_keys entries (keys of at most key_bits bits); block_bits 0 -> no
// Morton block table.
int alloc_state(Ctx* c, int64_t n_max, int64_t n_keys, int key_bits, int block_bits,
                int64_t meas_words) {
  int rc = SPH_OK;
  auto bail = [&](int code) { return code; };
  const sph_params& p = c->prm;
  const int device = c->device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    hip_fail(c, e, "hipSetDevice");
    return bail(SPH_ERR_HIP);
  }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    hip_fail(c, e, "hipStreamCreate");
    return bail(SPH_ERR_HIP);
  }
  c->L.stream = c->stream;
  // a pinned word for sync_and_check's read of the device status: a copy
  // into pageable memory costs ~15 us more per synchronisation
  e = hipHostMalloc(reinterpret_cast<void**>(&c->host_status), sizeof(uint32_t),
                    hipHostMallocDefault);
  if (e != hipSuccess) {
    c->host_status = nullptr;
    hip_fail(c, e, "hipHostMalloc");
    return bail(SPH_ERR_HIP);
  }
  DevState& s = c->s;
  s.ell_width = p.ell_width > 0 ? p.ell_width : 32;
  s.n_keys = n_keys;
  const int64_t N = n_max;
  float** farrs[] = {&s.rx, &s.ry, &s.vx, &s.vy, &s.tx, &s.ty, &s.tvx, &s.tvy, &s.ax,
                     &s.ay, &s.tax, &s.tay, &s.rho, &s.pr, &s.pres};
  for (float** a : farrs)
    if ((rc = dev_alloc(c, a, N))) return bail(rc);
  if ((rc = dev_alloc(c, &s.pid, N)) || (rc = dev_alloc(c, &s.pid_alt, N)) ||
      (rc = dev_alloc(c, &s.xyuv, N)) || (rc = dev_alloc(c, &s.rp, N)) ||
      (rc = dev_alloc(c, &s.cells, n_keys)) ||
      (rc = dev_alloc(c, &s.ell_count, N)) ||
      (rc = dev_alloc(c, &s.ell_idx, N * s.ell_width)) ||
      (rc = dev_alloc(c, &s.csr_count, N)) || (rc = dev_alloc(c, &s.csr_off, N + 1)))
    return bail(rc);
  const int tiles = (int)((N + sort_tile_size() - 1) / sort_tile_size());
  SortBuffers& b = c->sb;
  if ((rc = dev_alloc(c, &b.keys_in, N)) || (rc = dev_alloc(c, &b.keys_tmp, N)) ||
      (rc = dev_alloc(c, &b.keys_out, N)) || (rc = dev_alloc(c, &b.vals_a, N)) ||
      (rc = dev_alloc(c, &b.vals_b, N)) || (rc = dev_alloc(c, &b.vals_out, N)) ||
      (rc = dev_alloc(c, &b.hist, kMaxPasses * 256)) ||
      (rc = dev_alloc(c, &b.tickets, kMaxPasses)) || (rc = dev_alloc(c, &b.status, 1)) ||
      (rc = dev_alloc(c, &b.lookback, (int64_t)sort_passes(key_bits) * tiles * 256)))
    return bail(rc);
  b.max_tiles = tiles;
  s.block_bits = block_bits;
  s.n_blocks = s.block_bits ? ((n_keys + (1 << (2 * s.block_bits)) - 1) >> (2 * s.block_bits))
                            : 0;
  if (s.block_bits && (rc = dev_alloc(c, &s.blocks, s.n_blocks))) return bail(rc);
  if ((rc = dev_alloc(c, &c->block_stats, 6))) return bail(rc);
  {
    ResortBuffers& r = c->rb;
    if ((rc = dev_alloc(c, &r.mkey, N)) || (rc = dev_alloc(c, &r.midx, N)) ||
        (rc = dev_alloc(c, &r.skey, N)) || (rc = dev_alloc(c, &r.sidx, N)) ||
        (rc = dev_alloc(c, &r.ktmp, N)) || (rc = dev_alloc(c, &r.keys_alt, N)) ||
        (rc = dev_alloc(c, &r.flags, resort_flag_words((int)N))) ||
        (rc = dev_alloc(c, &r.tile_base, resort_tiles((int)N))) ||
        (rc = dev_alloc(c, &s.nkeys, N)) || (rc = dev_alloc(c, &s.nkeys_in, N)) ||
        (rc = dev_alloc(c, &s.nkeys_alt, N)) || (rc = dev_alloc(c, &s.t4, N)) ||
        (rc = dev_alloc(c, &s.tile_first, resort_tiles((int)N))) ||
        (rc = dev_alloc(c, &s.tile_count, resort_count_words((int)N))) ||
        (rc = dev_alloc(c, &r.vtmp, N)) || (rc = dev_alloc(c, &r.mode, kResortStateWords)) ||
        (rc = dev_alloc(c, &r.bucket, resort_tiles((int)N) * kBucketCap * 2)))
      return bail(rc);
    r.n_movers = reinterpret_cast<int*>(r.mode + 1);
    r.claim = r.mode + 2;
    r.n_last = reinterpret_cast<int*>(r.mode + 3);
    r.mode_last = r.mode + 4;
    r.ready = r.mode + 5;
    r.status = b.status;
  }
  b.blocks = s.blocks;
  b.block_bits = s.block_bits;
  int cu = 256;
  (void)hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device);
  b.hist_blocks = std::max(cu, 1);
  c->block_grid = block_step_grid(s.n_blocks, cu);
  b.cells = s.cells;
  b.n_keys = (uint64_t)n_keys;
  s.keys_sorted = b.keys_out;
  s.perm = b.vals_out;
  s.hist = b.hist;
  if ((rc = dev_alloc(c, &c->scan_scratch, scan_scratch_words((int)N))) ||
      (rc = dev_alloc(c, &c->meas_part, measure_partial_words((int)N))) ||
      (rc = dev_alloc(c, &c->meas_out, meas_words)) || (rc = dev_alloc(c, &c->ovf, 1)))
    return bail(rc);
  e = hipMemsetAsync(b.hist, 0, kMaxPasses * 256 * sizeof(uint32_t), c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(b.status, 0, sizeof(uint32_t), c->stream);
  // the re-sort's step words: count 0, path buckets, claim free
  if (e == hipSuccess)
    e = hipMemsetAsync(c->rb.mode, 0, kResortStateWords * sizeof(uint32_t), c->stream);
  if (e == hipSuccess) e = hipMemsetD32Async(c->rb.mode, 1, 1, c->stream);
  // producers that OR their mover bits in expect zero flag words
  if (e == hipSuccess)
    e = hipMemsetAsync(c->rb.flags, 0, resort_flag_words((int)N) * sizeof(unsigned long long),
                       c->stream);
  if (e == hipSuccess) e = launch_fill_cells(c->L, s.cells, n_keys);
  if (e == hipSuccess && s.block_bits) e = launch_fill_cells(c->L, s.blocks, s.n_blocks);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    hip_fail(c, e, "initialise buffers");
    return bail(SPH_ERR_HIP);
  }
  c->cells_cap = n_keys;
  c->passes_cap = sort_passes(key_bits);
  return SPH_OK;
}


// Grids, constants and key bases of every scene of a batch for one tuning, and
// the per-scene totals of the shared key space; no device work.
int batch_layout(const Ctx* c, const sph_batch_tuning& t, std::vector<sph_grid_info>& grids,
                 std::vector<SceneDev>& dev, int64_t& total, std::string& why) {
  if (!(t.cell_factor >= 1.0) || !std::isfinite(t.cell_factor)) {
    why = "cell_factor must be >= 1";
    return -1;
  }
  if (t.tile < kWave || t.tile > kMaxTile || t.tile % kWave != 0 || (t.tile & (t.tile - 1))) {
    why = "tile must be 64, 128, 256 or 512";
    return -1;
  }
  if (t.resort != SPH_RESORT_FULL && t.resort != SPH_RESORT_INCREMENTAL) {
    why = "unknown re-sort mode";
    return -1;
  }
  const int S = c->n_scenes;
  grids.assign(S, sph_grid_info{});
  dev.assign(S, SceneDev{});
  total = 0;
  for (int i = 0; i < S; ++i) {
    sph_params p = c->scene_prm[i];
    p.cell_size = t.cell_factor * p.h;
    if (derive_grid(p, grids[i], why) != 0) {
      why = "scene " + std::to_string(i) + ": " + why;
      return -1;
    }
    derive_phys(p, grids[i], dev[i].c, dev[i].g);
    dev[i].base = (uint64_t)total;
    // every scene's table starts at an even entry (16-byte aligned): the
    // density walk loads pairs of ranges
    total += (grids[i].n_keys + 1) & ~(int64_t)1;
    if (total > ((int64_t)1 << 31)) {
      why = "batch cell table larger than 2^31 entries";
      return -1;
    }
  }
  return 0;
}

int key_bits_of(int64_t n_keys) {
  int bits = 0;
  while (bits < 63 && ((uint64_t)1 << bits) < (uint64_t)n_keys) ++bits;
  return std::max(bits, 1);
}

// Apply a tuning to a batch context: scene constants and key bases, tile table,
// a cleared cell table; the state must be re-sorted afterwards.
int batch_configure(Ctx* c, const sph_batch_tuning& t) {
  std::vector<sph_grid_info> grids;
  std::vector<SceneDev> dev;
  int64_t total = 0;
  std::string why;
  if (batch_layout(c, t, grids, dev, total, why) != 0) return fail(c, SPH_ERR_INVALID, why);
  const int bits = key_bits_of(total);
  if (total > c->cells_cap || sort_passes(bits) > c->passes_cap)
    return fail(c, SPH_ERR_INVALID, "tuning needs a larger cell table than the context holds");
  std::vector<TileDev> tiles;
  for (int i = 0; i < c->n_scenes; ++i) {
    const int64_t b = c->scene_first[i], e = b + c->scene_count[i];
    for (int64_t q = b; q < e; q += t.tile)
      tiles.push_back(TileDev{i, (int32_t)q, (int32_t)std::min<int64_t>(q + t.tile, e), 0});
  }
  if ((int64_t)tiles.size() > c->tiles_cap) return fail(c, SPH_ERR_INVALID, "tile table overflow");
  destroy_graph(c);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_scenes, dev.data(), dev.size() * sizeof(SceneDev),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_tiles, tiles.data(), tiles.size() * sizeof(TileDev),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, launch_fill_cells(c->L, c->s.cells, total));
  HIPCHK(c, hipStreamSynchronize(c->stream));  // host vectors go out of scope
  c->scene_grid = grids;
  c->scene_dev = dev;
  c->tuning = t;
  c->resort = t.resort;
  c->B.scenes = c->d_scenes;
  c->B.tiles = c->d_tiles;
  c->B.n_tiles = (int)tiles.size();
  c->B.tile = t.tile;
  sph_grid_info& g = c->grid;
  g = sph_grid_info{};
  g.n_keys = total;
  g.key_bits = bits;
  g.order = c->prm.order;
  c->s.n_keys = total;
  c->sb.n_keys = (uint64_t)total;
  c->s.n_prev = 0;  // the whole table was just cleared
  c->sorted = false;
  c->density_ell = false;
  c->fields_prev_order = false;
  c->csr_total = -1;
  return SPH_OK;
}


// ---- learned surrogate handles
thread_local std::string g_nn_create_error;

struct NnH {
  Net net;
};

NnH* as_nn(void* p) { return static_cast<NnH*>(p); }

int nn_fail(NnH* h, int code, const std::string& msg) {
  if (h) h->net.err = msg;
  return code;
}

#define NNCHK(h, expr)                                                               \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return nn_fail(h, SPH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Shared checks of the calls that run the network on a context's sorted state.
int nn_ctx_check(NnH* h, Ctx* c, bool need_train) {
  if (!h || !c) return SPH_ERR_INVALID;
  if (c->s.n <= 0) return nn_fail(h, SPH_ERR_STATE, "context holds no particles");
  if (!c->sorted || !c->density_ell)
    return nn_fail(h, SPH_ERR_STATE, "context needs a fixed-width neighbor list (sph_prepare)");
  if (c->s.n > h->net.n_cap) return nn_fail(h, SPH_ERR_CAPACITY, "more particles than max_nodes");
  if (need_train && !h->net.train) return nn_fail(h, SPH_ERR_STATE, "network created with train = 0");
  return SPH_OK;
}

}  // namespace

extern "C" {

int sph_create(const sph_params* params, int64_t n_max, int device, void** out_ctx) {
  g_create_error.clear();
  g_static_error = nullptr;
  if (!params || !out_ctx || n_max <= 0 || n_max >= ((int64_t)1 << 30)) {
    g_create_error = "sph_create: invalid arguments (n_max must be in [1, 2^30))";
    return SPH_ERR_INVALID;
  }
  const sph_params& p = *params;
  if (p.physics != SPH_PHYS_REFERENCE && p.physics != SPH_PHYS_WCSPH) {
    g_create_error = "sph_create: unknown physics";
    return SPH_ERR_INVALID;
  }
  if (p.boundary != SPH_BC_WALLS && p.boundary != SPH_BC_PERIODIC) {
    g_create_error = "sph_create: unknown boundary";
    return SPH_ERR_INVALID;
  }
  if (p.order != SPH_ORDER_MORTON && p.order != SPH_ORDER_ROWMAJOR) {
    g_create_error = "sph_create: unknown order";
    return SPH_ERR_INVALID;
  }
  if (p.physics == SPH_PHYS_WCSPH && (p.gamma < 1 || p.gamma > 16 || !(p.rho0 > 0))) {
    g_create_error = "sph_create: WCSPH needs an integer gamma in [1,16] and rho0 > 0";
    return SPH_ERR_INVALID;
  }
  Ctx* c = new Ctx();
  c->prm = p;
  c->n_max = n_max;
  c->device = device;
  std::string why;
  if (derive_grid(p, c->grid, why) != 0) {
    g_create_error = "sph_create: " + why;
    delete c;
    return SPH_ERR_INVALID;
  }
  derive_phys(p, c->grid, c->L.phys, c->L.grid);
  int rc = alloc_state(c, n_max, c->grid.n_keys, c->grid.key_bits,
                       p.order == SPH_ORDER_MORTON ? 3 : 0, 8);
  if (rc != SPH_OK) {
    g_create_error = c->err.empty() ? std::string("sph_create failed") : c->err;
    sph_destroy(c);
    return rc;
  }
  *out_ctx = c;
  return SPH_OK;
}

int sph_destroy(void* ctx) {
  Ctx* c = as_ctx_keep(ctx);
  if (!c) return SPH_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  destroy_graph(c);
  for (auto& t : c->pending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->host_status) (void)hipHostFree(c->host_status);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return SPH_OK;
}

const char* sph_last_error(const void* ctx) {
  const Ctx* c = static_cast<const Ctx*>(ctx);
  if (!c) return g_static_error ? g_static_error : g_create_error.c_str();
  return c->err.c_str();
}

int sph_get_grid(const void* ctx, sph_grid_info* out) {
  const Ctx* c = static_cast<const Ctx*>(ctx);
  if (!c || !out) return SPH_ERR_INVALID;
  *out = c->grid;
  return SPH_OK;
}

int sph_get_constants(const void* ctx, float* out, int32_t count) {
  const Ctx* c = static_cast<const Ctx*>(ctx);
  if (!c || !out) return SPH_ERR_INVALID;
  if (c->n_scenes) return single_only(c);
  const DevPhys& p = c->L.phys;
  // order of orc_consts: h, h2, tol, mc_rho, c_pres, c_visc, gx, gy, eos_k, rho0,
  // inv_rho0, tait_b, neg_alpha_c0, eps_h2, dt, lo_x, lo_y, hi_x, hi_y, len_x, len_y
  const float v[] = {p.h,      p.h2,     p.tol,    p.mc_rho, p.c_pres, p.c_visc, p.gx,
                     p.gy,     p.eos_k,  p.rho0,   p.inv_rho0, p.tait_b, p.neg_alpha_c0,
                     p.eps_h2, p.dt,     p.lo_x,   p.lo_y,   p.hi_x,   p.hi_y,   p.len_x,
                     p.len_y};
  int m = std::min<int>(count, (int)(sizeof(v) / sizeof(v[0])));
  for (int i = 0; i < m; ++i) out[i] = v[i];
  return m;
}

int sph_upload(void* ctx, int64_t n, const float* rx, const float* ry, const float* vx,
               const float* vy, const uint32_t* pid) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (n < 0 || n > c->n_max) return fail(c, SPH_ERR_CAPACITY, "n exceeds the context capacity");
  if (c->n_scenes && n != c->n_max)
    return fail(c, SPH_ERR_INVALID, "a batch upload holds every scene's particles");
  if (n > 0 && (!rx || !ry || !vx || !vy))
    return fail(c, SPH_ERR_INVALID, "null position or velocity array");
  DevState& s = c->s;
  s.n = (int)n;
  size_t fb = (size_t)n * sizeof(float);
  HIPCHK(c, hipMemcpyAsync(s.rx, rx, fb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.ry, ry, fb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.vx, vx, fb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.vy, vy, fb, hipMemcpyHostToDevice, c->stream));
  std::vector<uint32_t> ids;
  const uint32_t* src = pid;
  if (!src) {
    ids.resize(n);
    for (int64_t i = 0; i < n; ++i) ids[i] = (uint32_t)i;
    src = ids.data();
  }
  HIPCHK(c, hipMemcpyAsync(s.pid, src, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(s.ax, 0, fb, c->stream));
  HIPCHK(c, hipMemsetAsync(s.ay, 0, fb, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->sorted = false;
  c->density_ell = false;
  c->fields_prev_order = false;
  c->csr_total = -1;
  if (c->graph_n != s.n) destroy_graph(c);
  return SPH_OK;
}

int sph_download(void* ctx, float* rx, float* ry, float* vx, float* vy, uint32_t* pid) {
  Ctx* c = as_ctx_read(ctx);
  if (!c) return SPH_ERR_INVALID;
  DevState& s = c->s;
  size_t fb = (size_t)s.n * sizeof(float);
  if (rx) HIPCHK(c, hipMemcpyAsync(rx, s.rx, fb, hipMemcpyDeviceToHost, c->stream));
  if (ry) HIPCHK(c, hipMemcpyAsync(ry, s.ry, fb, hipMemcpyDeviceToHost, c->stream));
  if (vx) HIPCHK(c, hipMemcpyAsync(vx, s.vx, fb, hipMemcpyDeviceToHost, c->stream));
  if (vy) HIPCHK(c, hipMemcpyAsync(vy, s.vy, fb, hipMemcpyDeviceToHost, c->stream));
  if (pid) HIPCHK(c, hipMemcpyAsync(pid, s.pid, (size_t)s.n * 4, hipMemcpyDeviceToHost, c->stream));
  return sync_and_check(c);
}

int sph_download_fields(void* ctx, float* ax, float* ay, float* rho, float* pres) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  int rc = settle_fields(c);
  if (rc) return rc;
  DevState& s = c->s;
  size_t fb = (size_t)s.n * sizeof(float);
  if (ax) HIPCHK(c, hipMemcpyAsync(ax, s.ax, fb, hipMemcpyDeviceToHost, c->stream));
  if (ay) HIPCHK(c, hipMemcpyAsync(ay, s.ay, fb, hipMemcpyDeviceToHost, c->stream));
  if (rho) HIPCHK(c, hipMemcpyAsync(rho, s.rho, fb, hipMemcpyDeviceToHost, c->stream));
  if (pres) HIPCHK(c, hipMemcpyAsync(pres, s.pres, fb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPH_OK;
}

int64_t sph_size(const void* ctx) {
  const Ctx* c = static_cast<const Ctx*>(ctx);
  return c ? c->s.n : -1;
}

int sph_find_neighbors(void* ctx) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  int rc = check_n(c);
  if (rc) return rc;
  return find_neighbors(c);
}

int sph_build_neighbor_lists(void* ctx, int32_t include_self, int64_t* total) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (c->n_scenes) return single_only(c);
  int rc = check_n(c);
  if (rc) return rc;
  if (!c->sorted) return fail(c, SPH_ERR_STATE, "call sph_find_neighbors first");
  DevState& s = c->s;
  StageTimer t(c, SPH_T_NEIGHBOR_LIST);
  HIPCHK(c, launch_csr_count(c->L, s, include_self != 0));
  HIPCHK(c, launch_scan(c->L, s.csr_count, s.csr_off, s.n, c->scan_scratch));
  int64_t tot = 0;
  HIPCHK(c, hipMemcpyAsync(&tot, s.csr_off + s.n, sizeof(int64_t), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (tot > c->csr_cap) {
    int64_t cap = std::max<int64_t>(tot + tot / 4, 1024);
    if (s.csr_idx) {
      auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)s.csr_idx);
      if (it != c->allocs.end()) c->allocs.erase(it);
      HIPCHK(c, hipFree(s.csr_idx));
      s.csr_idx = nullptr;
    }
    if ((rc = dev_alloc(c, &s.csr_idx, cap))) return rc;
    c->csr_cap = cap;
  }
  HIPCHK(c, launch_csr_fill(c->L, s, include_self != 0));
  c->csr_total = tot;
  c->csr_self = include_self != 0;
  if (total) *total = tot;
  return SPH_OK;
}

int sph_set_neighbor_lists(void* ctx, const int64_t* offsets, const int32_t* indices,
                           int32_t include_self) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (c->n_scenes) return single_only(c);
  int rc = check_n(c);
  if (rc) return rc;
  if (!c->sorted) return fail(c, SPH_ERR_STATE, "call sph_find_neighbors first");
  if (!offsets) return fail(c, SPH_ERR_INVALID, "offsets is null");
  const int64_t n = c->s.n;
  if (offsets[0] != 0) return fail(c, SPH_ERR_INVALID, "offsets[0] must be 0");
  for (int64_t i = 0; i < n; ++i)
    if (offsets[i + 1] < offsets[i])
      return fail(c, SPH_ERR_INVALID, "offsets decrease at particle " + std::to_string(i));
  const int64_t total = offsets[n];
  if (total > 0 && !indices) return fail(c, SPH_ERR_INVALID, "indices is null");
  if (total > (int64_t)INT32_MAX) return fail(c, SPH_ERR_INVALID, "more than 2^31-1 entries");
  // every index is checked here: the kernels gather through them unchecked
  std::vector<int64_t> off(n + 1);
  std::vector<int32_t> idx;
  idx.reserve((size_t)total);
  off[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    int self = 0;
    for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) {
      const int32_t j = indices[k];
      if (j < 0 || (int64_t)j >= n)
        return fail(c, SPH_ERR_INVALID, "index " + std::to_string(j) + " of particle " +
                                            std::to_string(i) + " outside [0, n)");
      if (j == i) {
        ++self;
        if (include_self) continue;
        return fail(c, SPH_ERR_INVALID, "particle " + std::to_string(i) +
                                            " lists itself (pass include_self = 1)");
      }
      idx.push_back(j);
    }
    if (include_self && self != 1)
      return fail(c, SPH_ERR_INVALID, "particle " + std::to_string(i) + " lists itself " +
                                          std::to_string(self) + " times (expected once)");
    off[i + 1] = (int64_t)idx.size();
  }
  DevState& s = c->s;
  const int64_t tot = (int64_t)idx.size();
  if (tot > c->csr_cap) {
    int64_t cap = std::max<int64_t>(tot + tot / 4, 1024);
    if (s.csr_idx) {
      auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)s.csr_idx);
      if (it != c->allocs.end()) c->allocs.erase(it);
      HIPCHK(c, hipFree(s.csr_idx));
      s.csr_idx = nullptr;
    }
    if ((rc = dev_alloc(c, &s.csr_idx, cap))) return rc;
    c->csr_cap = cap;
  }
  HIPCHK(c, hipMemcpyAsync(s.csr_off, off.data(), (size_t)(n + 1) * 8, hipMemcpyHostToDevice,
                           c->stream));
  if (tot > 0)
    HIPCHK(c, hipMemcpyAsync(s.csr_idx, idx.data(), (size_t)tot * 4, hipMemcpyHostToDevice,
                             c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));  // the host vectors go out of scope
  c->csr_total = tot;
  c->csr_self = 0;
  c->density_ell = false;  // SPH_COMPUTE_LISTS now means these lists
  return SPH_OK;
}

int sph_density_pressure(void* ctx, int32_t compute) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  int rc = check_n(c);
  if (rc || (rc = settle_fields(c))) return rc;
  if (!c->sorted) return fail(c, SPH_ERR_STATE, "call sph_find_neighbors first");
  if (c->n_scenes) {
    if (compute != SPH_COMPUTE_SORTED && compute != SPH_COMPUTE_LISTS)
      return fail(c, SPH_ERR_INVALID, "unknown compute type");
    StageTimer t(c, SPH_T_DENSITY);
    HIPCHK(c, launch_pack(c->L, c->s));  // the state may have changed since the sort
    HIPCHK(c, launch_batch_density(c->L, c->B, c->s));
    c->density_ell = true;
    return SPH_OK;
  }
  int nb = NB_STENCIL;
  if (compute == SPH_COMPUTE_LISTS) {
    if (c->csr_total < 0 || c->csr_self != 0)
      return fail(c, SPH_ERR_STATE, "lists compute needs sph_build_neighbor_lists(self=0)");
    nb = NB_CSR;
  } else if (compute != SPH_COMPUTE_SORTED) {
    return fail(c, SPH_ERR_INVALID, "unknown compute type");
  }
  StageTimer t(c, SPH_T_DENSITY);
  HIPCHK(c, launch_density(c->L, c->s, nb, true));
  c->density_ell = false;
  return SPH_OK;
}

int sph_forces(void* ctx, int32_t compute) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  int rc = check_n(c);
  if (rc || (rc = settle_fields(c))) return rc;
  if (!c->sorted) return fail(c, SPH_ERR_STATE, "call sph_find_neighbors first");
  if (c->n_scenes) {
    if (compute != SPH_COMPUTE_SORTED && compute != SPH_COMPUTE_LISTS)
      return fail(c, SPH_ERR_INVALID, "unknown compute type");
    if (!c->density_ell) return fail(c, SPH_ERR_STATE, "call sph_density_pressure first");
    // the fused kernel also writes the integrated twin and keys into scratch
    StageTimer t(c, SPH_T_FORCE);
    HIPCHK(c, launch_pack(c->L, c->s));  // the state may have changed since the sort
    HIPCHK(c, launch_batch_forces(c->L, c->B, c->s, c->sb.keys_in));
    return SPH_OK;
  }
  int nb = NB_STENCIL;
  if (compute == SPH_COMPUTE_LISTS) {
    if (c->density_ell) {
      nb = NB_ELL;
    } else if (c->csr_total >= 0 && c->csr_self == 0) {
      nb = NB_CSR;
    } else {
      return fail(c, SPH_ERR_STATE, "lists compute needs neighbor lists (self excluded)");
    }
  } else if (compute != SPH_COMPUTE_SORTED) {
    return fail(c, SPH_ERR_INVALID, "unknown compute type");
  }
  StageTimer t(c, SPH_T_FORCE);
  HIPCHK(c, launch_forces(c->L, c->s, nb, false, nullptr, MoverSink{}));
  return SPH_OK;
}

int sph_integrate(void* ctx) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (c->n_scenes) return single_only(c);
  int rc = check_n(c);
  if (rc || (rc = settle_fields(c))) return rc;
  StageTimer t(c, SPH_T_INTEGRATION);
  HIPCHK(c, launch_integrate(c->L, c->s));
  c->density_ell = false;  // positions moved: neighbor data is stale
  c->csr_total = -1;
  return SPH_OK;
}

int sph_set_dt(void* ctx, double dt) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (c->n_scenes) return single_only(c);
  if (!(dt > 0) || !std::isfinite(dt)) return fail(c, SPH_ERR_INVALID, "dt must be positive");
  c->prm.dt = dt;
  c->L.phys.dt = (float)dt;
  destroy_graph(c);  // dt is baked into the captured launches
  return SPH_OK;
}

int sph_update_velocities(void* ctx) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (c->n_scenes) return single_only(c);
  int rc = check_n(c);
  if (rc || (rc = settle_fields(c))) return rc;
  StageTimer t(c, SPH_T_INTEGRATION);
  HIPCHK(c, launch_update_velocities(c->L, c->s));
  c->density_ell = false;
  return SPH_OK;
}

int sph_update_positions(void* ctx) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (c->n_scenes) return single_only(c);
  int rc = check_n(c);
  if (rc || (rc = settle_fields(c))) return rc;
  StageTimer t(c, SPH_T_INTEGRATION);
  HIPCHK(c, launch_update_positions(c->L, c->s));
  c->density_ell = false;
  c->csr_total = -1;
  return SPH_OK;
}

int sph_measure(void* ctx, double* out5) {
  Ctx* c = as_ctx(ctx);
  if (!c || !out5) return SPH_ERR_INVALID;
  if (c->n_scenes) return single_only(c);
  {
    StageTimer t(c, SPH_T_MEASURE);
    HIPCHK(c, launch_measure(c->L, c->s, c->prm.mass, c->prm.gravity_y, c->meas_part, c->meas_out));
  }
  HIPCHK(c, hipMemcpyAsync(out5, c->meas_out, 5 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPH_OK;
}

int sph_prepare(void* ctx) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  int rc = check_n(c);
  if (rc) return rc;
  if ((rc = find_neighbors(c))) return rc;
  {
    StageTimer t(c, SPH_T_DENSITY);
    if (c->n_scenes)
      HIPCHK(c, launch_batch_density(c->L, c->B, c->s));
    else
      HIPCHK(c, launch_density(c->L, c->s, NB_ELL, true));
  }
  c->density_ell = true;
  return SPH_OK;
}

// sph_step, or with graphs_only the capture of the graphs such a call replays
// (nothing launched, no state changed).
int step_impl(Ctx* c, int32_t n_steps, bool graphs_only) {
  int rc = check_n(c);
  if (rc) return rc;
  if (!c->sorted) return fail(c, SPH_ERR_STATE, "call sph_prepare first");
  // per-particle kernels (one scene or batched), two steps or more: the steps
  // keep only the packed copy current and the separate arrays are written back
  // once at the end (a single step has the merge write them, which costs less
  // than the pass over the packed copy)
  const bool lazy =
      n_steps >= 2 &&
      (c->n_scenes || !(c->s.block_bits && (c->step_kernel == SPH_STEP_FUSED_BLOCK ||
                                            c->step_kernel == SPH_STEP_SPLIT_BLOCK)));
  // and with cell keys below 2^32 and the incremental re-sort, the steps keep
  // the sorted keys as u32 (k_pack narrows them, k_unpack widens them back)
  c->s.narrow = lazy && c->resort == SPH_RESORT_INCREMENTAL && c->grid.key_bits <= 32 &&
                c->s.n <= kResortMaxParticles;
  if (graphs_only) {
    c->soa_lazy = lazy;
    rc = SPH_OK;
    if (c->use_graph && n_steps > 0) {
      const int per = 2 * c->graph_pairs, rest = (n_steps % per) & ~1;
      hipGraphExec_t x = nullptr;
      if (n_steps >= per) rc = get_graph(c, per, &x);
      if (rc == SPH_OK && rest >= 2) rc = get_graph(c, rest, &x);
    }
    c->soa_lazy = false;
    c->s.narrow = false;
    return rc;
  }
  // separate arrays a lazy call left stale are written back unless this call
  // is lazy in the same way (it reads only the packed copy and the keys it
  // left)
  if (c->soa_stale && !(lazy && c->s.narrow == c->soa_stale_narrow) && (rc = settle_soa(c))) {
    c->s.narrow = false;
    return rc;
  }
  // the packed copy of the sorted state the step kernels gather from (other
  // calls may have changed the state since the last re-sort wrote it; not
  // after a lazy step call with nothing in between)
  const bool pack = !(c->packed_ok && (!c->s.narrow || c->packed_ok_narrow));
  c->packed_ok = false;
  if (pack) {
    const hipError_t e = launch_pack(c->L, c->s);
    if (e != hipSuccess) {
      c->s.narrow = false;
      return hip_fail(c, e, "pack");
    }
  }
  if (n_steps <= 0) {
    c->s.narrow = false;
    c->csr_total = -1;
    return SPH_OK;
  }
  c->soa_lazy = lazy;
  auto run = [&]() -> int {
    int done = 0, rc2 = SPH_OK;
    if (c->use_graph && !c->timing) {
      // whole graphs, then the even part of what is left in one shorter graph
      // (an odd last step runs without)
      const int per = 2 * c->graph_pairs;
      hipGraphExec_t x = nullptr;
      for (; done + per <= n_steps; done += per) {
        if (!x && (rc2 = get_graph(c, per, &x))) return rc2;
        HIPCHK(c, hipGraphLaunch(x, c->stream));
      }
      const int rest = (n_steps - done) & ~1;
      if (rest >= 2) {
        if ((rc2 = get_graph(c, rest, &x))) return rc2;
        HIPCHK(c, hipGraphLaunch(x, c->stream));
        done += rest;
      }
    }
    for (; done < n_steps; ++done) {
      c->fields_out = done + 1 == n_steps;
      if ((rc2 = fused_iteration(c))) return rc2;
    }
    return SPH_OK;
  };
  c->fields_prev_order = false;  // every field is rewritten by the last iteration
  rc = run();
  c->fields_out = true;
  c->soa_lazy = false;
  const bool narrow = c->s.narrow;
  if (lazy) {
    if (rc == SPH_OK && !c->view_out) {
      // the separate arrays are written back when something needs them
      c->soa_stale = true;
      c->soa_stale_narrow = narrow;
    } else {
      const hipError_t e = launch_unpack(c->L, c->s);
      if (rc == SPH_OK && e != hipSuccess) rc = hip_fail(c, e, "unpack");
    }
  }
  c->s.narrow = false;
  if (rc) return rc;
  c->packed_ok = lazy && !c->view_out;
  c->packed_ok_narrow = narrow;
  c->csr_total = -1;
  c->density_ell = false;
  c->fields_prev_order = true;  // computed before the final re-sort
  return SPH_OK;
}

int sph_step(void* ctx, int32_t n_steps) {
  Ctx* c = as_ctx_keep(ctx);
  if (!c) return SPH_ERR_INVALID;
  return step_impl(c, n_steps, false);
}

int sph_step_graphs(void* ctx, int32_t n_steps) {
  Ctx* c = as_ctx_keep(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (n_steps < 0) return fail(c, SPH_ERR_INVALID, "negative step count");
  return step_impl(c, n_steps, true);
}

int sph_set_graph(void* ctx, int32_t enable) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (enable < 0 || enable > 64) return fail(c, SPH_ERR_INVALID, "graph: 0 to 64 iteration pairs");
  c->use_graph = enable != 0;
  if (enable != 0 && enable != c->graph_pairs) {
    c->graph_pairs = enable;
    destroy_graph(c);
  }
  if (!c->use_graph) destroy_graph(c);
  return SPH_OK;
}

int sph_ell_overflow(void* ctx, int64_t* count) {
  Ctx* c = as_ctx(ctx);
  if (!c || !count) return SPH_ERR_INVALID;
  if (!c->density_ell) return fail(c, SPH_ERR_STATE, "no fixed-width list built");
  HIPCHK(c, hipMemsetAsync(c->ovf, 0, sizeof(int64_t), c->stream));
  HIPCHK(c, launch_ell_overflow(c->L, c->s, c->ovf));
  HIPCHK(c, hipMemcpyAsync(count, c->ovf, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPH_OK;
}

int sph_set_step_kernel(void* ctx, int32_t variant) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (c->n_scenes) return single_only(c);
  if (variant < SPH_STEP_SPLIT_BLOCK || variant > SPH_STEP_PARTICLE)
    return fail(c, SPH_ERR_INVALID, "unknown step kernel");
  c->step_kernel = variant;
  destroy_graph(c);
  return SPH_OK;
}

int sph_set_resort(void* ctx, int32_t mode) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (mode != SPH_RESORT_FULL && mode != SPH_RESORT_INCREMENTAL)
    return fail(c, SPH_ERR_INVALID, "unknown re-sort mode");
  c->resort = mode;
  c->tuning.resort = mode;
  destroy_graph(c);
  return SPH_OK;
}

int sph_movers(void* ctx, int64_t* count) {
  Ctx* c = as_ctx_keep(ctx);
  if (!c || !count) return SPH_ERR_INVALID;
  int m = 0;
  HIPCHK(c, hipMemcpyAsync(&m, c->rb.n_last, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *count = m;
  return SPH_OK;
}

int sph_resort_path(void* ctx, int32_t* path) {
  Ctx* c = as_ctx_keep(ctx);
  if (!c || !path) return SPH_ERR_INVALID;
  uint32_t m = 0;
  HIPCHK(c, hipMemcpyAsync(&m, c->rb.mode_last, sizeof(uint32_t), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *path = m != 0u ? 1 : 0;
  return SPH_OK;
}

int sph_block_stats(void* ctx, int64_t* out6) {
  Ctx* c = as_ctx(ctx);
  if (!c || !out6) return SPH_ERR_INVALID;
  if (c->n_scenes) return single_only(c);
  if (!c->sorted || !c->s.block_bits)
    return fail(c, SPH_ERR_STATE, "needs a sorted Morton-order state");
  HIPCHK(c, hipMemsetAsync(c->block_stats, 0, 6 * sizeof(unsigned long long), c->stream));
  int variant = c->step_kernel == SPH_STEP_PARTICLE ? SPH_STEP_SPLIT_BLOCK : c->step_kernel;
  // writes only the unsorted twin, keys_in and the per-particle fields
  HIPCHK(c, launch_block_step(c->L, c->s, c->sb.keys_in, c->block_grid, variant,
                              c->block_stats));
  unsigned long long v[6];
  HIPCHK(c, hipMemcpyAsync(v, c->block_stats, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < 6; ++i) out6[i] = (int64_t)v[i];
  return SPH_OK;
}

int sph_get_sorted_keys(void* ctx, uint64_t* keys) {
  Ctx* c = as_ctx(ctx);
  if (!c || !keys) return SPH_ERR_INVALID;
  if (!c->sorted) return fail(c, SPH_ERR_STATE, "no sort has run");
  HIPCHK(c, hipMemcpyAsync(keys, c->s.keys_sorted, (size_t)c->s.n * 8, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPH_OK;
}

int sph_get_permutation(void* ctx, uint32_t* perm) {
  Ctx* c = as_ctx(ctx);
  if (!c || !perm) return SPH_ERR_INVALID;
  if (!c->sorted) return fail(c, SPH_ERR_STATE, "no sort has run");
  HIPCHK(c, hipMemcpyAsync(perm, c->s.perm, (size_t)c->s.n * 4, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPH_OK;
}

int sph_get_cell_ranges(void* ctx, int32_t* start, int32_t* end) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (!c->sorted) return fail(c, SPH_ERR_STATE, "no sort has run");
  std::vector<int2> cells(c->grid.n_keys);
  HIPCHK(c, hipMemcpyAsync(cells.data(), c->s.cells, cells.size() * sizeof(int2),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (size_t k = 0; k < cells.size(); ++k) {
    if (start) start[k] = cells[k].x;
    if (end) end[k] = cells[k].y;
  }
  return SPH_OK;
}

int sph_get_neighbor_lists(void* ctx, int64_t* offsets, int32_t* indices) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (c->n_scenes) return single_only(c);
  if (c->csr_total < 0) return fail(c, SPH_ERR_STATE, "lists not built");
  if (offsets)
    HIPCHK(c, hipMemcpyAsync(offsets, c->s.csr_off, (size_t)(c->s.n + 1) * 8,
                             hipMemcpyDeviceToHost, c->stream));
  if (indices && c->csr_total > 0)
    HIPCHK(c, hipMemcpyAsync(indices, c->s.csr_idx, (size_t)c->csr_total * 4,
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPH_OK;
}

int sph_get_ell(void* ctx, int32_t* counts, int32_t* slots) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (!c->density_ell) return fail(c, SPH_ERR_STATE, "no fixed-width list built");
  const int n = c->s.n, w = c->s.ell_width;
  std::vector<int32_t> cm(slots ? (size_t)n * w : 0);
  if (counts)
    HIPCHK(c, hipMemcpyAsync(counts, c->s.ell_count, (size_t)n * 4, hipMemcpyDeviceToHost,
                             c->stream));
  if (slots)
    HIPCHK(c, hipMemcpyAsync(cm.data(), c->s.ell_idx, cm.size() * 4, hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (slots)
    for (int k = 0; k < w; ++k)
      for (int i = 0; i < n; ++i) slots[(size_t)i * w + k] = cm[(size_t)k * n + i];
  return SPH_OK;
}

int sph_get_device_view(void* ctx, sph_device_view* out) {
  Ctx* c = as_ctx(ctx);
  if (!c || !out) return SPH_ERR_INVALID;
  c->view_out = true;  // from now on every sph_step call packs the state afresh
  int rc = settle_fields(c);  // the view's fields follow the current order
  if (rc) return rc;
  DevState& s = c->s;
  out->n = s.n;
  out->rx = s.rx; out->ry = s.ry; out->vx = s.vx; out->vy = s.vy;
  out->ax = s.ax; out->ay = s.ay; out->rho = s.rho; out->pr = s.pr;
  out->pid = s.pid;
  out->ell_count = s.ell_count;
  out->ell_index = s.ell_idx;
  out->ell_width = s.ell_width;
  out->pad = 0;
  out->cell_range = reinterpret_cast<const int32_t*>(s.cells);
  out->stream = c->stream;
  return SPH_OK;
}

int sph_set_timing(void* ctx, int32_t enable) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  int rc = collect_timings(c);
  c->timing = enable != 0;
  return rc;
}

int sph_stage_times(void* ctx, double* ms, int64_t* launches, int32_t count) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  int rc = collect_timings(c);
  if (rc) return rc;
  for (int i = 0; i < count && i < SPH_T_COUNT; ++i) {
    if (ms) ms[i] = c->stage_ms[i];
    if (launches) launches[i] = c->launches[i];
    c->stage_ms[i] = 0;
    c->launches[i] = 0;
  }
  return SPH_OK;
}

int sph_synchronize(void* ctx) {
  Ctx* c = as_ctx_keep(ctx);
  if (!c) return SPH_ERR_INVALID;
  return sync_and_check(c);
}

int sph_batch_create(const sph_params* scenes, int32_t n_scenes, const int64_t* counts,
                     int device, void** out_ctx) {
  g_create_error.clear();
  g_static_error = nullptr;
  if (!scenes || !counts || !out_ctx || n_scenes < 1 || n_scenes > (1 << 20)) {
    g_create_error = "sph_batch_create: invalid arguments";
    return SPH_ERR_INVALID;
  }
  int64_t total = 0;
  for (int i = 0; i < n_scenes; ++i) {
    if (counts[i] < 1) {
      g_create_error = "sph_batch_create: every scene needs at least one particle";
      return SPH_ERR_INVALID;
    }
    total += counts[i];
  }
  if (total >= ((int64_t)1 << 30)) {
    g_create_error = "sph_batch_create: more than 2^30 particles";
    return SPH_ERR_INVALID;
  }
  const sph_params& p0 = scenes[0];
  for (int i = 0; i < n_scenes; ++i) {
    const sph_params& p = scenes[i];
    if (p.physics != p0.physics || p.boundary != p0.boundary || p.order != p0.order ||
        p.ell_width != p0.ell_width) {
      g_create_error = "sph_batch_create: scenes must share physics, boundary, order, ell_width";
      return SPH_ERR_INVALID;
    }
    if (p.physics == SPH_PHYS_WCSPH && (p.gamma < 1 || p.gamma > 16 || !(p.rho0 > 0))) {
      g_create_error = "sph_batch_create: WCSPH needs an integer gamma in [1,16] and rho0 > 0";
      return SPH_ERR_INVALID;
    }
  }
  if ((p0.physics != SPH_PHYS_REFERENCE && p0.physics != SPH_PHYS_WCSPH) ||
      (p0.boundary != SPH_BC_WALLS && p0.boundary != SPH_BC_PERIODIC) ||
      (p0.order != SPH_ORDER_MORTON && p0.order != SPH_ORDER_ROWMAJOR)) {
    g_create_error = "sph_batch_create: unknown physics, boundary or order";
    return SPH_ERR_INVALID;
  }
  Ctx* c = new Ctx();
  c->prm = p0;
  c->n_max = total;
  c->device = device;
  c->n_scenes = n_scenes;
  c->scene_prm.assign(scenes, scenes + n_scenes);
  c->scene_count.assign(counts, counts + n_scenes);
  c->scene_first.resize(n_scenes);
  int64_t off = 0;
  for (int i = 0; i < n_scenes; ++i) {
    c->scene_first[i] = off;
    off += counts[i];
  }
  c->step_kernel = SPH_STEP_PARTICLE;
  // the cell table is sized for cell_factor 1, the largest any tuning needs
  std::vector<sph_grid_info> grids;
  std::vector<SceneDev> dev;
  int64_t n_keys = 0;
  std::string why;
  sph_batch_tuning t0{1.0, 256, SPH_RESORT_INCREMENTAL};
  if (batch_layout(c, t0, grids, dev, n_keys, why) != 0) {
    g_create_error = "sph_batch_create: " + why;
    delete c;
    return SPH_ERR_INVALID;
  }
  derive_phys(p0, grids[0], c->L.phys, c->L.grid);  // physics / boundary selectors
  int rc = alloc_state(c, total, n_keys, key_bits_of(n_keys), 0,
                       std::max<int64_t>(8, 5 * (int64_t)n_scenes));
  if (rc == SPH_OK) {
    c->tiles_cap = total / kWave + n_scenes + 1;
    rc = dev_alloc(c, &c->d_scenes, n_scenes);
    if (rc == SPH_OK) rc = dev_alloc(c, &c->d_tiles, c->tiles_cap);
  }
  if (rc == SPH_OK) rc = batch_configure(c, t0);
  if (rc != SPH_OK) {
    g_create_error = c->err.empty() ? std::string("sph_batch_create failed") : c->err;
    sph_destroy(c);
    return rc;
  }
  *out_ctx = c;
  return SPH_OK;
}

int32_t sph_batch_scenes(const void* ctx) {
  const Ctx* c = static_cast<const Ctx*>(ctx);
  return c ? c->n_scenes : SPH_ERR_INVALID;
}

int sph_batch_scene_info(const void* ctx, int32_t scene, int64_t* first, int64_t* count,
                         sph_grid_info* grid) {
  const Ctx* c = static_cast<const Ctx*>(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (scene < 0 || scene >= c->n_scenes)
    return fail(const_cast<Ctx*>(c), SPH_ERR_INVALID, "scene index out of range");
  if (first) *first = c->scene_first[scene];
  if (count) *count = c->scene_count[scene];
  if (grid) *grid = c->scene_grid[scene];
  return SPH_OK;
}

int sph_batch_set_tuning(void* ctx, const sph_batch_tuning* tuning) {
  Ctx* c = as_ctx(ctx);
  if (!c || !tuning) return SPH_ERR_INVALID;
  if (!c->n_scenes) return fail(c, SPH_ERR_STATE, "not a batch context");
  return batch_configure(c, *tuning);
}

int sph_batch_get_tuning(const void* ctx, sph_batch_tuning* tuning) {
  const Ctx* c = static_cast<const Ctx*>(ctx);
  if (!c || !tuning) return SPH_ERR_INVALID;
  if (!c->n_scenes) return fail(const_cast<Ctx*>(c), SPH_ERR_STATE, "not a batch context");
  *tuning = c->tuning;
  return SPH_OK;
}

int32_t sph_batch_default_candidates(sph_batch_tuning* out, int32_t cap) {
  const double factors[] = {1.0, 1.25, 1.5};
  const int32_t tiles[] = {64, 128, 256, 512};
  const int32_t resorts[] = {SPH_RESORT_INCREMENTAL, SPH_RESORT_FULL};
  int32_t n = 0;
  for (double f : factors)
    for (int32_t t : tiles)
      for (int32_t r : resorts) {
        if (out && n < cap) out[n] = sph_batch_tuning{f, t, r};
        ++n;
      }
  return n;
}

int sph_batch_autotune(void* ctx, const sph_batch_tuning* candidates, int32_t n_candidates,
                       int32_t trial_steps, double* ms_per_step, int32_t* best) {
  Ctx* c = as_ctx(ctx);
  if (!c || !candidates || n_candidates < 1 || trial_steps < 1) return SPH_ERR_INVALID;
  if (!c->n_scenes) return fail(c, SPH_ERR_STATE, "not a batch context");
  int rc = check_n(c);
  if (rc) return rc;
  DevState& s = c->s;
  const int64_t N = c->n_max;
  if (!c->backup) {
    if ((rc = dev_alloc(c, &c->backup, 6 * N)) || (rc = dev_alloc(c, &c->backup_pid, N)))
      return rc;
  }
  const size_t fb = (size_t)s.n * sizeof(float);
  auto copy_state = [&](bool save) -> hipError_t {
    float* arrs[] = {s.rx, s.ry, s.vx, s.vy, s.ax, s.ay};
    for (int k = 0; k < 6; ++k) {
      float* bk = c->backup + (size_t)k * N;
      hipError_t e = hipMemcpyAsync(save ? bk : arrs[k], save ? arrs[k] : bk, fb,
                                    hipMemcpyDeviceToDevice, c->stream);
      if (e != hipSuccess) return e;
    }
    return hipMemcpyAsync(save ? c->backup_pid : s.pid, save ? s.pid : c->backup_pid,
                          (size_t)s.n * 4, hipMemcpyDeviceToDevice, c->stream);
  };
  HIPCHK(c, copy_state(true));
  const sph_batch_tuning start = c->tuning;
  const bool was_graph = c->use_graph, was_timing = c->timing;
  if ((rc = collect_timings(c))) return rc;
  c->use_graph = true;
  c->timing = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  int best_i = -1;
  double best_ms = 0;
  for (int i = 0; i < n_candidates && rc == SPH_OK; ++i) {
    if (ms_per_step) ms_per_step[i] = -1.0;
    if (batch_configure(c, candidates[i]) != SPH_OK) continue;  // not applicable
    c->soa_stale = false;  // the trial's state is replaced
    hipError_t e = copy_state(false);
    if (e != hipSuccess) { rc = hip_fail(c, e, "restore state"); break; }
    if ((rc = sph_prepare(c)) || (rc = sph_step(c, 2))) break;
    if ((e = hipEventRecord(e0, c->stream)) != hipSuccess) { rc = hip_fail(c, e, "event"); break; }
    if ((rc = sph_step(c, trial_steps))) break;
    if ((e = hipEventRecord(e1, c->stream)) != hipSuccess) { rc = hip_fail(c, e, "event"); break; }
    if ((rc = sync_and_check(c))) break;
    float ms = 0;
    if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) { rc = hip_fail(c, e, "timing"); break; }
    const double per = (double)ms / trial_steps;
    if (ms_per_step) ms_per_step[i] = per;
    if (best_i < 0 || per < best_ms) {
      best_i = i;
      best_ms = per;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->use_graph = was_graph;
  c->packed_ok = false;  // the trials' state is replaced below
  c->timing = was_timing;
  if (rc == SPH_OK && best_i < 0) rc = fail(c, SPH_ERR_INVALID, "no candidate could be applied");
  const std::string err = c->err;
  // leave the chosen (or, after a failure, the starting) tuning and the saved state
  int rc2 = batch_configure(c, rc == SPH_OK ? candidates[best_i] : start);
  c->soa_stale = false;  // the trials' state is replaced
  if (rc2 == SPH_OK) {
    hipError_t e = copy_state(false);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc2 = hip_fail(c, e, "restore state");
  }
  if (rc != SPH_OK) {
    c->err = err;
    return rc;
  }
  if (rc2 != SPH_OK) return rc2;
  if (best) *best = best_i;
  return SPH_OK;
}

int sph_batch_measure(void* ctx, double* out) {
  Ctx* c = as_ctx(ctx);
  if (!c || !out) return SPH_ERR_INVALID;
  if (!c->n_scenes) return fail(c, SPH_ERR_STATE, "not a batch context");
  int rc = check_n(c);
  if (rc) return rc;
  {
    StageTimer t(c, SPH_T_MEASURE);
    for (int i = 0; i < c->n_scenes; ++i) {
      // scene i occupies [first, first + count) of every particle array
      DevState v = c->s;
      const int64_t f = c->scene_first[i];
      v.rx += f; v.ry += f; v.vx += f; v.vy += f;
      v.n = (int)c->scene_count[i];
      Launch L = c->L;
      L.phys = c->scene_dev[i].c;
      HIPCHK(c, launch_measure(L, v, c->scene_prm[i].mass, c->scene_prm[i].gravity_y,
                               c->meas_part, c->meas_out + 5 * (int64_t)i));
    }
  }
  HIPCHK(c, hipMemcpyAsync(out, c->meas_out, 5 * sizeof(double) * c->n_scenes,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPH_OK;
}

int sph_export_state(void* ctx, float* d_out) {
  Ctx* c = as_ctx(ctx);
  if (!c || !d_out) return SPH_ERR_INVALID;
  HIPCHK(c, launch_export_state(c->L, c->s, d_out));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SPH_OK;
}

int sph_device_malloc(int device, int64_t bytes, void** out_ptr) {
  if (!out_ptr || bytes <= 0) return SPH_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return SPH_ERR_HIP;
  return hipMalloc(out_ptr, (size_t)bytes) == hipSuccess ? SPH_OK : SPH_ERR_HIP;
}

int sph_device_free(void* ptr) {
  if (!ptr) return SPH_OK;
  return hipFree(ptr) == hipSuccess ? SPH_OK : SPH_ERR_HIP;
}

int sph_device_copy(void* dst, const void* src, int64_t bytes, int32_t kind) {
  if (!dst || !src || bytes < 0 || kind < 0 || kind > 2) return SPH_ERR_INVALID;
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                                    : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  return hipMemcpy(dst, src, (size_t)bytes, k) == hipSuccess ? SPH_OK : SPH_ERR_HIP;
}

int sph_upload_device(void* ctx, int64_t n, const float* d_rx, const float* d_ry,
                      const float* d_vx, const float* d_vy) {
  Ctx* c = as_ctx(ctx);
  if (!c) return SPH_ERR_INVALID;
  if (n < 0 || n > c->n_max) return fail(c, SPH_ERR_CAPACITY, "n exceeds the context capacity");
  if (c->n_scenes && n != c->n_max)
    return fail(c, SPH_ERR_INVALID, "a batch upload holds every scene's particles");
  if (n > 0 && (!d_rx || !d_ry || !d_vx || !d_vy))
    return fail(c, SPH_ERR_INVALID, "null position or velocity array");
  DevState& s = c->s;
  s.n = (int)n;
  const size_t fb = (size_t)n * sizeof(float);
  HIPCHK(c, hipMemcpyAsync(s.rx, d_rx, fb, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.ry, d_ry, fb, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.vx, d_vx, fb, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(s.vy, d_vy, fb, hipMemcpyDeviceToDevice, c->stream));
  std::vector<uint32_t> ids(n);
  for (int64_t i = 0; i < n; ++i) ids[i] = (uint32_t)i;
  HIPCHK(c, hipMemcpyAsync(s.pid, ids.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(s.ax, 0, fb, c->stream));
  HIPCHK(c, hipMemsetAsync(s.ay, 0, fb, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->sorted = false;
  c->density_ell = false;
  c->fields_prev_order = false;
  c->csr_total = -1;
  if (c->graph_n != s.n) destroy_graph(c);
  return SPH_OK;
}

int sph_nn_create(const sph_nn_config* cfg, int device, void** out_nn) {
  g_nn_create_error.clear();
  if (!cfg || !out_nn) {
    g_nn_create_error = "sph_nn_create: invalid arguments";
    return SPH_ERR_INVALID;
  }
  const sph_nn_config& q = *cfg;
  if (q.width != kNnWidth || q.layers < 1 || q.layers > 8 || q.max_nodes < 1 ||
      q.max_nodes >= ((int64_t)1 << 30) || !(q.h > 0) || !(q.rho0 > 0) || !(q.c0 > 0) ||
      !(q.v_scale > 0) || !(q.a_scale > 0) ||
      (q.periodic && (!(q.len_x > 0) || !(q.len_y > 0)))) {
    g_nn_create_error = "sph_nn_create: width must be 128, layers 1..8, positive scales";
    return SPH_ERR_INVALID;
  }
  NnConsts k{};
  k.h = (float)q.h;
  k.inv_h = (float)(1.0 / q.h);
  k.rho0 = (float)q.rho0;
  k.inv_rho0 = (float)(1.0 / q.rho0);
  k.rs = 50.0f;
  k.ps = (float)(50.0 * q.rho0 / (q.c0 * q.c0));
  k.inv_vs = (float)(1.0 / q.v_scale);
  k.a_s = (float)q.a_scale;
  k.gx = (float)q.gravity_x;
  k.gy = (float)q.gravity_y;
  k.len_x = (float)q.len_x;
  k.len_y = (float)q.len_y;
  k.half_x = k.len_x * 0.5f;
  k.half_y = k.len_y * 0.5f;
  k.periodic = q.periodic ? 1 : 0;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) {
    g_nn_create_error = std::string("hipSetDevice: ") + hipGetErrorString(e);
    return SPH_ERR_HIP;
  }
  NnH* h = new NnH();
  h->net.device = device;
  if (nn_create(&h->net, q.layers, q.train != 0, q.max_nodes, k) != 0) {
    g_nn_create_error = h->net.err;
    nn_free(&h->net);
    delete h;
    return SPH_ERR_HIP;
  }
  *out_nn = h;
  return SPH_OK;
}

int sph_nn_destroy(void* nn) {
  NnH* h = as_nn(nn);
  if (!h) return SPH_OK;
  (void)hipDeviceSynchronize();
  nn_free(&h->net);
  delete h;
  return SPH_OK;
}

const char* sph_nn_last_error(const void* nn) {
  const NnH* h = static_cast<const NnH*>(nn);
  return h ? h->net.err.c_str() : g_nn_create_error.c_str();
}

int64_t sph_nn_num_params(const void* nn) {
  const NnH* h = static_cast<const NnH*>(nn);
  return h ? h->net.n_params : SPH_ERR_INVALID;
}

int sph_nn_set_params(void* nn, const float* host, int64_t count) {
  NnH* h = as_nn(nn);
  if (!h || !host) return SPH_ERR_INVALID;
  if (count != h->net.n_params) return nn_fail(h, SPH_ERR_INVALID, "parameter count mismatch");
  NNCHK(h, hipDeviceSynchronize());
  NNCHK(h, hipMemcpy(h->net.theta, host, count * sizeof(float), hipMemcpyHostToDevice));
  NNCHK(h, nn_sync_weights(&h->net, nullptr));
  NNCHK(h, hipDeviceSynchronize());
  return SPH_OK;
}

int sph_nn_get_params(void* nn, float* host, int64_t count) {
  NnH* h = as_nn(nn);
  if (!h || !host) return SPH_ERR_INVALID;
  if (count != h->net.n_params) return nn_fail(h, SPH_ERR_INVALID, "parameter count mismatch");
  NNCHK(h, hipDeviceSynchronize());
  NNCHK(h, hipMemcpy(host, h->net.theta, count * sizeof(float), hipMemcpyDeviceToHost));
  return SPH_OK;
}

int sph_nn_get_grad(void* nn, float* host, int64_t count) {
  NnH* h = as_nn(nn);
  if (!h || !host) return SPH_ERR_INVALID;
  if (count != h->net.n_params) return nn_fail(h, SPH_ERR_INVALID, "parameter count mismatch");
  NNCHK(h, hipDeviceSynchronize());
  NNCHK(h, hipMemcpy(host, h->net.grad, count * sizeof(float), hipMemcpyDeviceToHost));
  return SPH_OK;
}

int sph_nn_reset_optimizer(void* nn) {
  NnH* h = as_nn(nn);
  if (!h) return SPH_ERR_INVALID;
  NNCHK(h, hipDeviceSynchronize());
  NNCHK(h, hipMemset(h->net.m, 0, h->net.n_params * sizeof(float)));
  NNCHK(h, hipMemset(h->net.v, 0, h->net.n_params * sizeof(float)));
  h->net.adam_t = 0;
  return SPH_OK;
}

int sph_nn_set_clip(void* nn, double max_norm) {
  NnH* h = as_nn(nn);
  if (!h) return SPH_ERR_INVALID;
  if (!(max_norm >= 0) || !std::isfinite(max_norm))
    return nn_fail(h, SPH_ERR_INVALID, "gradient-norm limit must be finite and >= 0");
  h->net.clip = (float)max_norm;
  return SPH_OK;
}

int sph_nn_forward(void* nn, void* ctx, float* ax, float* ay) {
  NnH* h = as_nn(nn);
  Ctx* c = as_ctx(ctx);
  int rc = nn_ctx_check(h, c, false);
  if (rc) return rc;
  Net& net = h->net;
  float* dax = net.a_out;
  float* day = net.a_out + net.n_cap;
  NNCHK(h, nn_forward(&net, c->s, c->stream, dax, day));
  const size_t fb = (size_t)c->s.n * sizeof(float);
  if (ax) NNCHK(h, hipMemcpyAsync(ax, dax, fb, hipMemcpyDeviceToHost, c->stream));
  if (ay) NNCHK(h, hipMemcpyAsync(ay, day, fb, hipMemcpyDeviceToHost, c->stream));
  NNCHK(h, hipStreamSynchronize(c->stream));
  return SPH_OK;
}

int sph_nn_loss_grad(void* nn, void* ctx, double* out3) {
  NnH* h = as_nn(nn);
  Ctx* c = as_ctx(ctx);
  int rc = nn_ctx_check(h, c, true);
  if (rc) return rc;
  Net& net = h->net;
  // the sender half of the edge backward walks each list backwards: the lists
  // must be complete (no particle over ell_width neighbors)
  int64_t over = 0;
  NNCHK(h, hipMemsetAsync(c->ovf, 0, sizeof(int64_t), c->stream));
  DevState capped = c->s;  // the network reads at most 32 slots per particle
  capped.ell_width = std::min(capped.ell_width, 32);
  NNCHK(h, launch_ell_overflow(c->L, capped, c->ovf));
  NNCHK(h, hipMemcpyAsync(&over, c->ovf, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  NNCHK(h, hipStreamSynchronize(c->stream));
  if (over != 0) return nn_fail(h, SPH_ERR_STATE, "a neighbor list overflowed ell_width");
  NNCHK(h, nn_loss_backward(&net, c->s, c->stream, c->s.ax, c->s.ay));
  double sums[2] = {0, 0};
  NNCHK(h, hipMemcpyAsync(sums, net.red_out, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
  NNCHK(h, hipStreamSynchronize(c->stream));
  if (out3) {
    const double n = c->s.n;
    out3[0] = sums[1] > 0 ? sums[0] / sums[1] : 0.0;
    out3[1] = std::sqrt(out3[0]);
    out3[2] = std::sqrt(sums[1] / n);
  }
  return SPH_OK;
}

int sph_nn_adam(void* nn, void* ctx, double lr, double beta1, double beta2, double eps) {
  NnH* h = as_nn(nn);
  Ctx* c = as_ctx(ctx);
  if (!h || !c) return SPH_ERR_INVALID;
  if (!h->net.train) return nn_fail(h, SPH_ERR_STATE, "network created with train = 0");
  if (!(lr > 0) || !(beta1 >= 0 && beta1 < 1) || !(beta2 >= 0 && beta2 < 1) || !(eps > 0))
    return nn_fail(h, SPH_ERR_INVALID, "invalid Adam hyper-parameters");
  NNCHK(h, nn_adam(&h->net, c->stream, (float)lr, (float)beta1, (float)beta2, (float)eps));
  return SPH_OK;
}

int sph_nn_train_step(void* nn, void* ctx, double lr, double* out3) {
  int rc = sph_nn_loss_grad(nn, ctx, out3);
  if (rc) return rc;
  return sph_nn_adam(nn, ctx, lr, 0.9, 0.999, 1e-8);
}

int sph_surrogate_run(void* ctx, void* nn, int32_t n_steps, int32_t check_every,
                      double threshold, int32_t mode, sph_surrogate_stats* stats) {
  Ctx* c = as_ctx(ctx);
  NnH* h = as_nn(nn);
  if (!c || !h || n_steps < 0 || check_every < 1 || !(threshold > 0) || (mode != 0 && mode != 1))
    return SPH_ERR_INVALID;
  if (c->n_scenes) return single_only(c);
  if (c->s.n <= 0) return fail(c, SPH_ERR_STATE, "no particles uploaded");
  if (!c->sorted) return fail(c, SPH_ERR_STATE, "call sph_prepare first");
  if (c->s.n > h->net.n_cap) return fail(c, SPH_ERR_CAPACITY, "more particles than max_nodes");
  DevState& s = c->s;
  const int64_t N = c->n_max;
  int rc = settle_fields(c);
  if (rc) return rc;
  if (mode == 1 && !c->backup) {
    if ((rc = dev_alloc(c, &c->backup, 6 * N)) || (rc = dev_alloc(c, &c->backup_pid, N)))
      return rc;
  }
  sph_surrogate_stats st{};
  double err_sum = 0;
  const size_t fb = (size_t)s.n * sizeof(float);
  auto snapshot = [&](bool save) -> hipError_t {
    float* arrs[] = {s.rx, s.ry, s.vx, s.vy};
    for (int k = 0; k < 4; ++k) {
      float* bk = c->backup + (size_t)k * N;
      hipError_t e = hipMemcpyAsync(save ? bk : arrs[k], save ? arrs[k] : bk, fb,
                                    hipMemcpyDeviceToDevice, c->stream);
      if (e != hipSuccess) return e;
    }
    return hipMemcpyAsync(save ? c->backup_pid : s.pid, save ? s.pid : c->backup_pid,
                          (size_t)s.n * 4, hipMemcpyDeviceToDevice, c->stream);
  };
  double* d_sums = h->net.red_out + 2;
  for (int t = 0; t < n_steps;) {
    const int k = std::min(check_every, n_steps - t);
    if (mode == 1) HIPCHK(c, snapshot(true));
    bool failed = false;
    int nn_done = 0, plain_done = 0;
    for (int q = 0; q < k; ++q) {
      {
        StageTimer tm(c, SPH_T_DENSITY);
        HIPCHK(c, launch_density(c->L, s, NB_ELL, false));
      }
      {
        StageTimer tm(c, SPH_T_FORCE);
        HIPCHK(c, nn_forward(&h->net, s, c->stream, s.tax, s.tay));
      }
      const float* ax = s.tax;
      const float* ay = s.tay;
      if (q == k - 1) {
        {
          StageTimer tm(c, SPH_T_FORCE);
          HIPCHK(c, launch_forces(c->L, s, NB_ELL, false, nullptr, MoverSink{}));
        }
        double sums[2] = {0, 0};
        HIPCHK(c, nn_error_sums(&h->net, s.n, s.tax, s.tay, s.ax, s.ay, c->stream, d_sums));
        HIPCHK(c, hipMemcpyAsync(sums, d_sums, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const double err = sums[1] > 0 ? std::sqrt(sums[0] / sums[1]) : 0.0;
        st.checks += 1;
        st.last_error = err;
        st.max_error = std::max(st.max_error, err);
        err_sum += err;
        if (err > threshold) {
          st.failed_checks += 1;
          failed = true;
          if (mode == 1) break;
          ax = s.ax;  // this step integrates with the plain forces
          ay = s.ay;
        }
      }
      {
        StageTimer tm(c, SPH_T_INTEGRATION);
        HIPCHK(c, launch_integrate_keys(c->L, s, ax, ay, c->sb.keys_in));
      }
      if ((rc = resort_integrated(c))) return rc;
      if (ax == s.ax) ++plain_done; else ++nn_done;
    }
    if (failed && mode == 1) {
      // discard the window: back to its start, then plain SPH for its k steps
      HIPCHK(c, snapshot(false));
      if ((rc = find_neighbors(c))) return rc;
      for (int q = 0; q < k; ++q)
        if ((rc = fused_iteration(c))) return rc;
      nn_done = 0;
      plain_done = k;
    }
    st.nn_steps += nn_done;
    st.plain_steps += plain_done;
    st.steps += k;
    t += k;
  }
  if ((rc = sync_and_check(c))) return rc;
  st.mean_error = st.checks ? err_sum / (double)st.checks : 0.0;
  c->density_ell = false;
  c->csr_total = -1;
  if (stats) *stats = st;
  return SPH_OK;
}

int sph_cfl(void* ctx, double cfl, double* out3) {
  Ctx* c = as_ctx(ctx);
  if (!c || !out3) return SPH_ERR_INVALID;
  if (!(cfl > 0) || !std::isfinite(cfl)) return fail(c, SPH_ERR_INVALID, "cfl must be positive");
  if (c->n_scenes) return single_only(c);
  int rc = check_n(c);
  if (rc || (rc = settle_fields(c))) return rc;
  double m[2] = {0, 0};
  HIPCHK(c, launch_max_va(c->L, c->s, c->meas_part, c->meas_out));
  HIPCHK(c, hipMemcpyAsync(m, c->meas_out, sizeof(m), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const double vmax = std::sqrt(m[0]), amax = std::sqrt(m[1]), h = c->prm.h;
  const double c0 = c->prm.physics == SPH_PHYS_WCSPH ? c->prm.c0 : 0.0;
  double dt = HUGE_VAL;
  if (c0 + vmax > 0) dt = std::min(dt, cfl * h / (c0 + vmax));
  if (amax > 0) dt = std::min(dt, 0.25 * std::sqrt(h / amax));
  if (!std::isfinite(dt)) dt = c->prm.dt;
  out3[0] = dt;
  out3[1] = vmax;
  out3[2] = amax;
  return SPH_OK;
}

}  // extern "C"

namespace {

#pragma pack(push, 1)
struct SnapHeader {
  char magic[8];
  uint32_t version;
  uint32_t header_bytes;
  int64_t n;
  int64_t step;
  sph_params params;
  uint32_t flags;
  uint32_t reserved;
  uint64_t checksum;
};
#pragma pack(pop)
static_assert(sizeof(sph_params) == 144, "sph_params layout");
static_assert(sizeof(SnapHeader) == 192, "snapshot header layout");

constexpr char kSnapMagic[8] = {'S', 'P', 'H', 'S', 'N', 'A', 'P', '1'};

uint64_t fnv1a(uint64_t h, const void* p, size_t bytes) {
  const unsigned char* c = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < bytes; ++i) {
    h ^= c[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

// Reads and validates a header; msg names the problem.
bool read_header(FILE* f, SnapHeader& hd, std::string& msg) {
  if (std::fread(&hd, sizeof(hd), 1, f) != 1) {
    msg = "snapshot: short header";
    return false;
  }
  if (std::memcmp(hd.magic, kSnapMagic, 8) != 0) {
    msg = "snapshot: not an SPHSNAP1 file";
    return false;
  }
  if (hd.version != 1 || hd.header_bytes != sizeof(SnapHeader)) {
    msg = "snapshot: unsupported version";
    return false;
  }
  if (hd.n < 0 || hd.n >= ((int64_t)1 << 30)) {
    msg = "snapshot: particle count out of range";
    return false;
  }
  return true;
}

bool same_params_but_dt(sph_params a, sph_params b) {
  a.dt = 0;
  b.dt = 0;
  return std::memcmp(&a, &b, sizeof(a)) == 0;
}

}  // namespace

extern "C" {

int sph_snapshot_info(const char* path, sph_params* params, int64_t* n, int64_t* step) {
  g_create_error.clear();
  g_static_error = nullptr;
  if (!path) return SPH_ERR_INVALID;
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    g_create_error = std::string("snapshot: cannot open ") + path;
    return SPH_ERR_INVALID;
  }
  SnapHeader hd{};
  std::string msg;
  const bool ok = read_header(f, hd, msg);
  std::fclose(f);
  if (!ok) {
    g_create_error = msg;
    return SPH_ERR_INVALID;
  }
  if (params) *params = hd.params;
  if (n) *n = hd.n;
  if (step) *step = hd.step;
  return SPH_OK;
}

int sph_snapshot_save(void* ctx, const char* path, int64_t step) {
  Ctx* c = as_ctx(ctx);
  if (!c || !path) return SPH_ERR_INVALID;
  if (c->n_scenes) return single_only(c);
  const int64_t n = c->s.n;
  std::vector<float> r(4 * (size_t)n);
  std::vector<uint32_t> pid((size_t)n);
  const size_t fb = (size_t)n * sizeof(float);
  DevState& s = c->s;
  const float* src[4] = {s.rx, s.ry, s.vx, s.vy};
  for (int k = 0; k < 4; ++k)
    HIPCHK(c, hipMemcpyAsync(r.data() + (size_t)k * n, src[k], fb, hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHK(c, hipMemcpyAsync(pid.data(), s.pid, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  SnapHeader hd{};
  std::memcpy(hd.magic, kSnapMagic, 8);
  hd.version = 1;
  hd.header_bytes = sizeof(SnapHeader);
  hd.n = n;
  hd.step = step;
  hd.params = c->prm;
  hd.flags = 1;
  uint64_t h = 0xcbf29ce484222325ull;
  h = fnv1a(h, r.data(), r.size() * sizeof(float));
  h = fnv1a(h, pid.data(), pid.size() * sizeof(uint32_t));
  hd.checksum = h;
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(c, SPH_ERR_INVALID, std::string("snapshot: cannot create ") + path);
  bool ok = std::fwrite(&hd, sizeof(hd), 1, f) == 1;
  ok = ok && (n == 0 || std::fwrite(r.data(), sizeof(float), r.size(), f) == r.size());
  ok = ok && (n == 0 || std::fwrite(pid.data(), sizeof(uint32_t), pid.size(), f) == pid.size());
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(c, SPH_ERR_INVALID, std::string("snapshot: write failed: ") + path);
  return SPH_OK;
}

int sph_snapshot_load(void* ctx, const char* path, int64_t* step) {
  Ctx* c = as_ctx(ctx);
  if (!c || !path) return SPH_ERR_INVALID;
  if (c->n_scenes) return single_only(c);
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(c, SPH_ERR_INVALID, std::string("snapshot: cannot open ") + path);
  SnapHeader hd{};
  std::string msg;
  if (!read_header(f, hd, msg)) {
    std::fclose(f);
    return fail(c, SPH_ERR_INVALID, msg);
  }
  if (hd.n > c->n_max) {
    std::fclose(f);
    return fail(c, SPH_ERR_CAPACITY, "snapshot: more particles than the context holds");
  }
  if (!same_params_but_dt(hd.params, c->prm)) {
    std::fclose(f);
    return fail(c, SPH_ERR_INVALID, "snapshot: parameters differ from the context's");
  }
  const int64_t n = hd.n;
  std::vector<float> r(4 * (size_t)n);
  std::vector<uint32_t> pid((size_t)n);
  bool ok = n == 0 || std::fread(r.data(), sizeof(float), r.size(), f) == r.size();
  ok = ok && (n == 0 || std::fread(pid.data(), sizeof(uint32_t), pid.size(), f) == pid.size());
  std::fclose(f);
  if (!ok) return fail(c, SPH_ERR_INVALID, "snapshot: truncated arrays");
  uint64_t h = 0xcbf29ce484222325ull;
  h = fnv1a(h, r.data(), r.size() * sizeof(float));
  h = fnv1a(h, pid.data(), pid.size() * sizeof(uint32_t));
  if (h != hd.checksum) return fail(c, SPH_ERR_INVALID, "snapshot: checksum mismatch");
  int rc = sph_upload(c, n, r.data(), r.data() + n, r.data() + 2 * n, r.data() + 3 * n,
                      pid.data());
  if (rc) return rc;
  if (hd.params.dt != c->prm.dt && (rc = sph_set_dt(c, hd.params.dt))) return rc;
  if (step) *step = hd.step;
  return SPH_OK;
}

}  // extern "C"

// ---- the allocation-free measurement (sph_measure_noalloc) ----
namespace {

// Partial sums of the FP64 measurement kernels for callers that cannot
// allocate: 4 doubles per 256-particle block, 1 MiB per GPU (held until the
// process ends), so at most kNoAllocMaxN particles per call. Host code never
// passes `scratch` itself to a kernel (its host-side name is no device
// address): it uses the address hipGetSymbolAddress returns for the current
// device, looked up once per device.
constexpr int64_t kNoAllocWords = 131072;
constexpr int64_t kNoAllocMaxN = kNoAllocWords / 4 * 256;  // 8,388,608
constexpr int kNoAllocDevices = 64;
__device__ __align__(256) double scratch[kNoAllocWords];

std::mutex g_scratch_mu;
double* g_scratch_ptr[kNoAllocDevices] = {};

// Address of this device's copy of `scratch` (set up on the first call on a
// device, with that device current; no allocation).
hipError_t scratch_for_current_device(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kNoAllocDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

int noalloc_fail(int code, const char* msg) {
  g_static_error = msg;
  return code;
}

}  // namespace

extern "C" int sph_measure_noalloc(const sph_params* params, int64_t n, const float* rx,
                                   const float* ry, const float* vx, const float* vy,
                                   double* out5, void* stream) {
  if (!params || !out5) return noalloc_fail(SPH_ERR_INVALID, "sph_measure_noalloc: null argument");
  if (n < 0 || n > kNoAllocMaxN)
    return noalloc_fail(SPH_ERR_CAPACITY,
                        "sph_measure_noalloc: n must be in [0, 8388608] (the static scratch "
                        "holds the partial sums of 8,388,608 particles)");
  if (n > 0 && (!rx || !ry || !vx || !vy))
    return noalloc_fail(SPH_ERR_INVALID, "sph_measure_noalloc: null particle array");
  double* part = nullptr;
  if (scratch_for_current_device(&part) != hipSuccess)
    return noalloc_fail(SPH_ERR_HIP, "sph_measure_noalloc: hipGetSymbolAddress of the scratch "
                                     "array failed on the current device");
  if (measure_partial_words((int)n) > kNoAllocWords)
    return noalloc_fail(SPH_ERR_CAPACITY, "sph_measure_noalloc: scratch too small");
  // the same kernels and partial layout as sph_measure (bit-identical results);
  // of the physics constants only the domain bounds are read
  Launch L{};
  L.stream = static_cast<hipStream_t>(stream);
  L.phys.lo_x = (float)params->lo[0];
  L.phys.lo_y = (float)params->lo[1];
  L.phys.hi_x = (float)params->hi[0];
  L.phys.hi_y = (float)params->hi[1];
  DevState s{};
  s.n = (int)n;
  s.rx = const_cast<float*>(rx);
  s.ry = const_cast<float*>(ry);
  s.vx = const_cast<float*>(vx);
  s.vy = const_cast<float*>(vy);
  if (launch_measure(L, s, params->mass, params->gravity_y, part, out5) != hipSuccess)
    return noalloc_fail(SPH_ERR_HIP, "sph_measure_noalloc: kernel launch failed");
  g_static_error = nullptr;
  return SPH_OK;
}

