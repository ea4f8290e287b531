// gfx950 kernels of the 2D SPH step other than the radix sort: cell keys,
// permutation + cell ranges, neighbor lists, density + EOS, forces (optionally
// fused with integration and the next step's keys), integration, FP64
// diagnostics. One thread per particle in Morton order, so the neighbor gathers
// of a wavefront hit a compact patch of cells and coalesce in L2.
#include <hip/hip_runtime.h>

#include "sph_internal.hpp"

namespace sphamd {

namespace {

#ifndef SPH_BLK
#define SPH_BLK 256  // 128, 256 and 512 measured at 1M: 256 fastest
#endif
constexpr int BLK = SPH_BLK;

inline int blocks_for(int64_t n, int blk = BLK) { return (int)((n + blk - 1) / blk); }

__global__ __launch_bounds__(BLK) void k_fill_cells(int2* __restrict__ cells, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * BLK)
    cells[i] = make_int2(-1, -1);
}

// Cell key of every particle of the sorted state (GenericNeighborSearch.h:200-232).
__global__ __launch_bounds__(BLK) void k_keys(int n, DevGrid g, const float* __restrict__ rx,
                                              const float* __restrict__ ry,
                                              uint64_t* __restrict__ keys) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  keys[i] = cell_key(g, cell_coord(rx[i], g.lo_x, g.cs_x, g.nx),
                     cell_coord(ry[i], g.lo_y, g.cs_y, g.ny));
}

// Cell start/end of sorted index i (cuSetCellRanges, cuFindNeighbors.cu:103-122):
// run starts are found with one wave64 ballot of "key differs from the previous
// key"; a run ends where the next one starts, so the end mask is the start mask
// shifted by one lane (the last lane compares with the next wave's first key).
// The same is done one level up for Morton blocks. The three key loads are
// unconditional (clamped into [0, n), n >= 1) so that they issue together.
// Also keeps the key of every kResortTile-th slot, and clears the tile's counter,
// for the incremental re-sort.
// Returns i < n.
__device__ __forceinline__ bool write_ranges(int i, int n, const uint64_t* __restrict__ keys,
                                             int2* __restrict__ cells, int64_t n_keys,
                                             int2* __restrict__ blocks, int block_bits,
                                             uint64_t* __restrict__ tile_first,
                                             unsigned long long* __restrict__ tile_count) {
  const bool live = i < n;
  const int ic = min(i, n - 1);
  const uint64_t kc = keys[ic], kpc = keys[max(ic - 1, 0)], knc = keys[min(ic + 1, n - 1)];
  uint64_t k = live ? kc : ~0ull;
  if (live && i % kResortTile == 0) {
    tile_first[i / kResortTile] = k;
    tile_count[i / kResortTile] = 0;
    if (i % (kResortTile * kChunkTiles) == 0)
      tile_count[(n + kResortTile - 1) / kResortTile + i / (kResortTile * kChunkTiles)] = 0;
  }
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t prev = __shfl_up(k, 1);
  if (lane == 0) prev = (i > 0 && live) ? kpc : ~k;
  const bool is_start = live && (i == 0 || k != prev);
  const uint64_t starts = __ballot(is_start);
  bool is_end;
  if (lane < kWave - 1) {
    is_end = live && (i == n - 1 || ((starts >> (lane + 1)) & 1ull));
  } else {
    is_end = live && (i == n - 1 || k != knc);
  }
  if (block_bits > 0) {
    // the same boundary detection one level up: Morton blocks of 4^block_bits cells
    const int sh = 2 * block_bits;
    const bool b_start = live && (i == 0 || (k >> sh) != (prev >> sh));
    const uint64_t bstarts = __ballot(b_start);
    bool b_end;
    if (lane < kWave - 1)
      b_end = live && (i == n - 1 || ((bstarts >> (lane + 1)) & 1ull));
    else
      b_end = live && (i == n - 1 || (k >> sh) != (knc >> sh));
    if (live && k < (uint64_t)n_keys) {
      const int bk = (int)(k >> sh);
      if (b_start) blocks[bk].x = i;
      if (b_end) blocks[bk].y = i + 1;
    }
  }
  if (k < (uint64_t)n_keys) {
    if (is_start) cells[k].x = i;
    if (is_end) cells[k].y = i + 1;
  }
  return live;
}

// Permutation through the sort result plus cell ranges. Also clears the sort's
// digit histogram.
__global__ __launch_bounds__(BLK) void k_gather(
    int n, const uint32_t* __restrict__ perm, const uint64_t* __restrict__ keys,
    const float* __restrict__ srx, const float* __restrict__ sry,
    const float* __restrict__ svx, const float* __restrict__ svy,
    const uint32_t* __restrict__ spid, const float* __restrict__ sax,
    const float* __restrict__ say, float* __restrict__ drx, float* __restrict__ dry,
    float* __restrict__ dvx, float* __restrict__ dvy, uint32_t* __restrict__ dpid,
    float* __restrict__ dax, float* __restrict__ day, int2* __restrict__ cells,
    int64_t n_keys, uint32_t* __restrict__ hist, int hist_words, int2* __restrict__ blocks,
    int block_bits, uint64_t* __restrict__ tile_first, unsigned long long* __restrict__ tile_count,
    float4* __restrict__ xyuv) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (blockIdx.x == 0)
    for (int w = threadIdx.x; w < hist_words; w += BLK) hist[w] = 0;
  const bool live =
      write_ranges(i, n, keys, cells, n_keys, blocks, block_bits, tile_first, tile_count);
  if (!live) return;
  const uint32_t j = min(perm[i], (uint32_t)(n - 1));  // in bounds even if perm were bad
  const float x = srx[j], y = sry[j], u = svx[j], v = svy[j];
  drx[i] = x;
  dry[i] = y;
  dvx[i] = u;
  dvy[i] = v;
  xyuv[i] = make_float4(x, y, u, v);
  dpid[i] = spid[j];
  if (sax) {
    dax[i] = sax[j];
    day[i] = say[j];
  }
}

// Cell and block ranges of an order produced by the incremental re-sort; clears
// the digit histogram its mover sort used. Each thread takes kRangesPer slots a
// grid stride apart, every key it needs (its own, the one before and the one
// after) loaded in one round: few enough waves that they all run at once.
#ifndef SPH_RANGES_PER
#define SPH_RANGES_PER 2  // 1, 2, 4 and 8 measured at 1M: 2 fastest
#endif
constexpr int kRangesPer = SPH_RANGES_PER;

template <typename K>
__global__ __launch_bounds__(BLK) void k_ranges(int n, const K* __restrict__ keys,
                                                int2* __restrict__ cells, int64_t n_keys,
                                                int2* __restrict__ blocks, int block_bits,
                                                uint32_t* __restrict__ hist, int hist_words,
                                                uint64_t* __restrict__ tile_first,
                                                unsigned long long* __restrict__ tile_count,
                                                ResortBuffers r) {
  const int stride = gridDim.x * BLK;
  const int i0 = blockIdx.x * BLK + threadIdx.x;
  if (blockIdx.x == 0)
    for (int w = threadIdx.x; w < hist_words; w += BLK) hist[w] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // the re-sort that just ended: its count and path for the host; then the
    // next step's start values (buckets unless a bucket overflows)
    *r.n_last = *r.n_movers;
    *r.mode_last = *r.mode;
    *r.n_movers = 0;
    *r.mode = 1u;
    *r.claim = 0u;
    *r.ready = 0u;
  }
  // each slot's key, and one neighbor key per lane (the one before for lane
  // 0, after for the others) loaded together; the neighbors inside the wave
  // come from the adjacent lanes (slot q of the wave's lanes is 64
  // consecutive indices)
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t kc[kRangesPer], kp[kRangesPer], kn[kRangesPer];
#pragma unroll
  for (int q = 0; q < kRangesPer; ++q) {
    const int ic = min(i0 + q * stride, n - 1);
    kc[q] = keys[ic];
    kn[q] = keys[lane == 0 ? max(ic - 1, 0) : min(ic + 1, n - 1)];
  }
#pragma unroll
  for (int q = 0; q < kRangesPer; ++q) {
    const uint64_t edge = kn[q];
    kp[q] = __shfl_up(kc[q], 1);
    kn[q] = __shfl_down(kc[q], 1);
    if (lane == 0) kp[q] = edge;
    if (lane == kWave - 1) kn[q] = edge;
  }
  const int sh = 2 * block_bits;
#pragma unroll
  for (int q = 0; q < kRangesPer; ++q) {
    const int i = i0 + q * stride;
    if (i >= n) break;
    const uint64_t k = kc[q];
    if (i % kResortTile == 0) {
      tile_first[i / kResortTile] = k;
      tile_count[i / kResortTile] = 0;
      if (i % (kResortTile * kChunkTiles) == 0)
        tile_count[(n + kResortTile - 1) / kResortTile + i / (kResortTile * kChunkTiles)] = 0;
    }
    if (k >= (uint64_t)n_keys) continue;
    if (i == 0 || k != kp[q]) cells[k].x = i;
    if (i == n - 1 || k != kn[q]) cells[k].y = i + 1;
    if (block_bits > 0) {
      const int bk = (int)(k >> sh);
      if (i == 0 || (k >> sh) != (kp[q] >> sh)) blocks[bk].x = i;
      if (i == n - 1 || (k >> sh) != (kn[q] >> sh)) blocks[bk].y = i + 1;
    }
  }
}

// Summation density + EOS. NB: stencil traversal, CSR rows, or stencil with the
// fixed-width (column-major) neighbor list written on the way.
template <int PHYS, int BC, int NB>
__global__ __launch_bounds__(BLK) void k_density(
    int n, DevGrid g, DevPhys c, const float* __restrict__ rx, const float* __restrict__ ry,
    const int2* __restrict__ cells, const int64_t* __restrict__ csr_off,
    const int32_t* __restrict__ csr_idx, int32_t* __restrict__ ell_count,
    int32_t* __restrict__ ell_idx, int ell_w, float* __restrict__ rho_out,
    float* __restrict__ pr_out, float* __restrict__ p_out, const float4* __restrict__ xyuv,
    float2* __restrict__ rp_out) {
  const int i = xcd_tile(blockIdx.x, gridDim.x) * BLK + threadIdx.x;
  if (i >= n) return;
  float xi, yi;
  if (NB == NB_ELL && xyuv) {
    const float2 q = ld8_of16(xyuv, i);  // the packed copy is the current state
    xi = q.x;
    yi = q.y;
  } else {
    xi = rx[i];
    yi = ry[i];
  }
  float acc = density_term(c.h2, 0.0f);
  if (NB == NB_CSR) {
    for (int64_t k = csr_off[i]; k < csr_off[i + 1]; ++k) {
      int j = csr_idx[k];
      float dx, dy;
      separation<BC>(c, xi, yi, rx[j], ry[j], dx, dy);
      acc += density_term(c.h2, dx * dx + dy * dy);
    }
  } else if (NB == NB_STENCIL) {
    for_each_neighbor<BC>(g, c, rx, ry, cells, n, i, xi, yi,
                          [&](int, float, float, float r2) { acc += density_term(c.h2, r2); });
  } else {
    __shared__ int s_list[kEllLds * BLK];
    if (xyuv)
      acc = density_ell<BC, true>(g, c, rx, ry, cells, n, i, xi, yi, s_list + threadIdx.x, BLK,
                                  ell_count, ell_idx, ell_w, xyuv);
    else
      acc = density_ell<BC>(g, c, rx, ry, cells, n, i, xi, yi, s_list + threadIdx.x, BLK,
                            ell_count, ell_idx, ell_w);
  }
  float rho, p, pr;
  eos<PHYS>(c, acc, rho, p, pr);
  if (rho_out) {  // (null: a step whose separate fields nobody reads)
    rho_out[i] = rho;
    pr_out[i] = pr;
  }
  if (p_out) p_out[i] = p;
  if (rp_out) rp_out[i] = make_float2(rho, pr);
}

// The separate arrays of the sorted state from its packed copy.
__global__ __launch_bounds__(BLK) void k_unpack(int n, const float4* __restrict__ xyuv,
                                                float* __restrict__ rx, float* __restrict__ ry,
                                                float* __restrict__ vx, float* __restrict__ vy,
                                                const uint32_t* __restrict__ k32,
                                                uint64_t* __restrict__ k64) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const float4 q = xyuv[i];
  rx[i] = q.x;
  ry[i] = q.y;
  vx[i] = q.z;
  vy[i] = q.w;
  if (k32) k64[i] = k32[i];  // the narrow keys back (null: none)
}

// The sorted state packed per particle (x, y, vx, vy) for the neighbor gathers.
__global__ __launch_bounds__(BLK) void k_pack(int n, const float* __restrict__ rx,
                                              const float* __restrict__ ry,
                                              const float* __restrict__ vx,
                                              const float* __restrict__ vy,
                                              float4* __restrict__ xyuv,
                                              const uint64_t* __restrict__ k64,
                                              uint32_t* __restrict__ k32) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  xyuv[i] = make_float4(rx[i], ry[i], vx[i], vy[i]);
  if (k32) k32[i] = (uint32_t)k64[i];  // narrow copy of the sorted keys (null: none)
}

// Accelerations. FUSED additionally integrates (symplectic Euler) into the T
// arrays and writes the cell key of the new position for the next sort (keys
// of type K: u32 for narrow keys).
template <int PHYS, int BC, int NB, bool FUSED, typename K = uint64_t>
__global__ __launch_bounds__(BLK) void k_forces(
    int n, DevGrid g, DevPhys c, const float* __restrict__ rx, const float* __restrict__ ry,
    const float* __restrict__ vx, const float* __restrict__ vy,
    const float* __restrict__ rho, const float* __restrict__ pr,
    const int2* __restrict__ cells, const int64_t* __restrict__ csr_off,
    const int32_t* __restrict__ csr_idx, const int32_t* __restrict__ ell_count,
    const int32_t* __restrict__ ell_idx, int ell_w, float* __restrict__ ax_out,
    float* __restrict__ ay_out, float* __restrict__ tx, float* __restrict__ ty,
    float* __restrict__ tvx, float* __restrict__ tvy, K* __restrict__ keys_out,
    const K* __restrict__ kold, MoverSink movers, const float4* __restrict__ xyuv,
    const float2* __restrict__ rp, float4* __restrict__ t4) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int i = tile * BLK + threadIdx.x;
  if (i >= n) return;
  // the bucketing's bounds of the wave's re-sort tile, read first (scalar
  // loads when a workgroup is one tile)
  int rtile = tile;
  if constexpr (BLK != kResortTile)
    rtile = (tile * BLK + (int)(threadIdx.x & ~(unsigned)(kWave - 1))) / kResortTile;
  const TileKeys tk = (FUSED && movers.bucket) ? tile_keys(movers, rtile) : TileKeys{};
  float xi, yi, vxi, vyi, rhoi, pri;
  if (NB == NB_ELL && xyuv) {
    // the particle's own state from the packed copies: two loads instead of six
    const float4 q = ld16(xyuv, i);
    const float2 r = ld8(rp, i);
    xi = q.x; yi = q.y; vxi = q.z; vyi = q.w; rhoi = r.x; pri = r.y;
  } else {
    xi = rx[i]; yi = ry[i]; vxi = vx[i]; vyi = vy[i]; rhoi = rho[i]; pri = pr[i];
  }
  const uint64_t ko = (FUSED && movers.flags) ? (uint64_t)kold[i] : 0ull;
  Acc a{0.f, 0.f, 0.f, 0.f};
  auto pair = [&](int j, float dx, float dy, float r2) {
    force_pair<PHYS>(c, dx, dy, r2, rhoi, rho[j], pri, pr[j], vxi, vyi, vx[j], vy[j], a);
  };
  auto from_list = [&](int j) {
    float dx, dy;
    separation<BC>(c, xi, yi, rx[j], ry[j], dx, dy);
    pair(j, dx, dy, dx * dx + dy * dy);
  };
  if (NB == NB_CSR) {
    for (int64_t k = csr_off[i]; k < csr_off[i + 1]; ++k) from_list(csr_idx[k]);
  } else if (NB == NB_ELL) {
    if (xyuv)
      a = forces_ell<PHYS, BC, true>(g, c, rx, ry, vx, vy, rho, pr, cells, ell_count, ell_idx,
                                     ell_w, n, i, xi, yi, vxi, vyi, rhoi, pri, xyuv, rp);
    else
      a = forces_ell<PHYS, BC>(g, c, rx, ry, vx, vy, rho, pr, cells, ell_count, ell_idx, ell_w,
                               n, i, xi, yi, vxi, vyi, rhoi, pri);
  } else {
    for_each_neighbor<BC>(g, c, rx, ry, cells, n, i, xi, yi, pair);
  }
  float axi, ayi;
  force_finish<PHYS>(c, a, axi, ayi);
  if (ax_out) {  // (null: a step whose accelerations nobody reads)
    ax_out[i] = axi;
    ay_out[i] = ayi;
  }
  if (FUSED) {
    float x = xi, y = yi, u = vxi, v = vyi;
    integrate_one<BC>(c, x, y, u, v, axi, ayi);
    if constexpr (sizeof(K) == 4) {
      t4[i] = make_float4(x, y, u, v);  // a narrow step keeps T packed
    } else {
      tx[i] = x;
      ty[i] = y;
      tvx[i] = u;
      tvy[i] = v;
    }
    const uint64_t kn = cell_key(g, cell_coord(x, g.lo_x, g.cs_x, g.nx),
                                 cell_coord(y, g.lo_y, g.cs_y, g.ny));
    keys_out[i] = (K)kn;
    // particles whose key changes (tiles are BLK-aligned, so wave w of the
    // tile covers flag word i / 64; waves past n write none)
    if (movers.flags) {
      record_movers(movers, i, kn != ko);
      if (movers.bucket) bucket_movers(movers, i, kn, kn != ko, make_float4(x, y, u, v), tk);
    }
  }
}

template <int BC>
__global__ __launch_bounds__(BLK) void k_integrate(int n, DevPhys c, float* __restrict__ rx,
                                                   float* __restrict__ ry,
                                                   float* __restrict__ vx,
                                                   float* __restrict__ vy,
                                                   const float* __restrict__ ax,
                                                   const float* __restrict__ ay) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  float x = rx[i], y = ry[i], u = vx[i], v = vy[i];
  integrate_one<BC>(c, x, y, u, v, ax[i], ay[i]);
  rx[i] = x;
  ry[i] = y;
  vx[i] = u;
  vy[i] = v;
}

// The two halves of integrate_one, for callers that drive them separately
// (cuUpdateVelocities / cuUpdatePositions, cuIntegrator.cu:60-121).
__global__ __launch_bounds__(BLK) void k_update_velocities(int n, DevPhys c,
                                                           float* __restrict__ vx,
                                                           float* __restrict__ vy,
                                                           const float* __restrict__ ax,
                                                           const float* __restrict__ ay) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  vx[i] = vx[i] + c.dt * ax[i];
  vy[i] = vy[i] + c.dt * ay[i];
}

template <int BC>
__global__ __launch_bounds__(BLK) void k_update_positions(int n, DevPhys c,
                                                          float* __restrict__ rx,
                                                          float* __restrict__ ry,
                                                          float* __restrict__ vx,
                                                          float* __restrict__ vy) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  float x = rx[i], y = ry[i], u = vx[i], v = vy[i];
  x = x + c.dt * u;
  y = y + c.dt * v;
  if (BC == 1) {
    if (x >= c.hi_x) x = x - c.len_x; else if (x < c.lo_x) x = x + c.len_x;
    if (y >= c.hi_y) y = y - c.len_y; else if (y < c.lo_y) y = y + c.len_y;
  } else {
    if (x > c.hi_x) { x = c.hi_x; u = -u; }
    if (x < c.lo_x) { x = c.lo_x; u = -u; }
    if (y > c.hi_y) { y = c.hi_y; v = -v; }
    if (y < c.lo_y) { y = c.lo_y; v = -v; }
  }
  rx[i] = x;
  ry[i] = y;
  vx[i] = u;
  vy[i] = v;
}

// Multipass CSR lists (cuCountNeighborsPerParticle / cuAssignIndices,
// cuNeighborLists.cuh:6-136).
template <int BC, bool SELF>
__global__ __launch_bounds__(BLK) void k_csr_count(int n, DevGrid g, DevPhys c,
                                                   const float* __restrict__ rx,
                                                   const float* __restrict__ ry,
                                                   const int2* __restrict__ cells,
                                                   int32_t* __restrict__ count) {
  const int i = xcd_tile(blockIdx.x, gridDim.x) * BLK + threadIdx.x;
  if (i >= n) return;
  int k = 0;
  for_each_neighbor<BC, SELF>(g, c, rx, ry, cells, n, i, rx[i], ry[i],
                              [&](int, float, float, float) { ++k; });
  count[i] = k;
}

template <int BC, bool SELF>
__global__ __launch_bounds__(BLK) void k_csr_fill(int n, DevGrid g, DevPhys c,
                                                  const float* __restrict__ rx,
                                                  const float* __restrict__ ry,
                                                  const int2* __restrict__ cells,
                                                  const int64_t* __restrict__ off,
                                                  int32_t* __restrict__ idx) {
  const int i = xcd_tile(blockIdx.x, gridDim.x) * BLK + threadIdx.x;
  if (i >= n) return;
  int64_t o = off[i];
  for_each_neighbor<BC, SELF>(g, c, rx, ry, cells, n, i, rx[i], ry[i],
                              [&](int j, float, float, float) { idx[o++] = j; });
}

// ---------------------------------------------------------------- scan
constexpr int SCAN_BLK = 1024;

__device__ __forceinline__ int64_t block_incl_scan64(int64_t v, int64_t* s_w) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  int64_t x = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    int64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) s_w[w] = x;
  __syncthreads();
  int64_t off = 0;
  for (int k = 0; k < w; ++k) off += s_w[k];
  __syncthreads();
  return x + off;
}

__global__ __launch_bounds__(SCAN_BLK) void k_scan_reduce(const int32_t* __restrict__ in, int n,
                                                          int64_t* __restrict__ part) {
  __shared__ int64_t s_w[SCAN_BLK / kWave];
  const int i = blockIdx.x * SCAN_BLK + threadIdx.x;
  int64_t v = i < n ? in[i] : 0;
  int64_t incl = block_incl_scan64(v, s_w);
  if (threadIdx.x == SCAN_BLK - 1) part[blockIdx.x] = incl;
}

__global__ __launch_bounds__(SCAN_BLK) void k_scan_top(int64_t* __restrict__ part, int nparts) {
  __shared__ int64_t s_w[SCAN_BLK / kWave];
  __shared__ int64_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int base = 0; base < nparts; base += SCAN_BLK) {
    int i = base + threadIdx.x;
    int64_t v = i < nparts ? part[i] : 0;
    int64_t incl = block_incl_scan64(v, s_w);
    int64_t carry = s_carry;
    if (i < nparts) part[i] = carry + incl - v;  // exclusive
    __syncthreads();
    if (threadIdx.x == SCAN_BLK - 1) s_carry = carry + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) part[nparts] = s_carry;
}

__global__ __launch_bounds__(SCAN_BLK) void k_scan_down(const int32_t* __restrict__ in, int n,
                                                        const int64_t* __restrict__ part,
                                                        int64_t* __restrict__ out) {
  __shared__ int64_t s_w[SCAN_BLK / kWave];
  const int i = blockIdx.x * SCAN_BLK + threadIdx.x;
  int64_t v = i < n ? in[i] : 0;
  int64_t incl = block_incl_scan64(v, s_w);
  if (i < n) out[i] = part[blockIdx.x] + incl - v;
  if (i == n - 1) out[n] = part[blockIdx.x] + incl;
}

// ----------------------------------------------------------- diagnostics
// Per-block FP64 partials of sum 1/2 m |v|^2, sum m (-g) (y - lo), max |v|^2 and
// the count inside the domain (MeasuringFunctions.h:5-62), then one block that
// adds the partials in a fixed order.
__global__ __launch_bounds__(BLK) void k_measure_partial(int n, DevPhys c, double mass, double g,
                                                         const float* __restrict__ rx,
                                                         const float* __restrict__ ry,
                                                         const float* __restrict__ vx,
                                                         const float* __restrict__ vy,
                                                         double* __restrict__ part) {
  __shared__ double s_ke[BLK], s_pe[BLK], s_vm[BLK], s_in[BLK];
  const int i = blockIdx.x * BLK + threadIdx.x;
  double ke = 0, pe = 0, vm = 0, in = 0;
  if (i < n) {
    double u = vx[i], v = vy[i];
    double v2 = u * u + v * v;
    ke = 0.5 * mass * v2;
    pe = mass * g * ((double)ry[i] - (double)c.lo_y);
    vm = v2;
    float x = rx[i], y = ry[i];
    bool inside = isfinite(x) && isfinite(y) && !(x < c.lo_x) && !(x > c.hi_x) &&
                  !(y < c.lo_y) && !(y > c.hi_y);
    in = inside ? 1.0 : 0.0;
  }
  s_ke[threadIdx.x] = ke;
  s_pe[threadIdx.x] = pe;
  s_vm[threadIdx.x] = vm;
  s_in[threadIdx.x] = in;
  __syncthreads();
  for (int s = BLK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      s_ke[threadIdx.x] += s_ke[threadIdx.x + s];
      s_pe[threadIdx.x] += s_pe[threadIdx.x + s];
      s_vm[threadIdx.x] = fmax(s_vm[threadIdx.x], s_vm[threadIdx.x + s]);
      s_in[threadIdx.x] += s_in[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x + 0] = s_ke[0];
    part[4 * blockIdx.x + 1] = s_pe[0];
    part[4 * blockIdx.x + 2] = s_vm[0];
    part[4 * blockIdx.x + 3] = s_in[0];
  }
}

__global__ __launch_bounds__(BLK) void k_measure_final(int nparts, int n, double mass,
                                                       const double* __restrict__ part,
                                                       double* __restrict__ out) {
  __shared__ double s_ke[BLK], s_pe[BLK], s_vm[BLK], s_in[BLK];
  double ke = 0, pe = 0, vm = 0, in = 0;
  for (int b = threadIdx.x; b < nparts; b += BLK) {
    ke += part[4 * b + 0];
    pe += part[4 * b + 1];
    vm = fmax(vm, part[4 * b + 2]);
    in += part[4 * b + 3];
  }
  s_ke[threadIdx.x] = ke;
  s_pe[threadIdx.x] = pe;
  s_vm[threadIdx.x] = vm;
  s_in[threadIdx.x] = in;
  __syncthreads();
  for (int s = BLK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      s_ke[threadIdx.x] += s_ke[threadIdx.x + s];
      s_pe[threadIdx.x] += s_pe[threadIdx.x + s];
      s_vm[threadIdx.x] = fmax(s_vm[threadIdx.x], s_vm[threadIdx.x + s]);
      s_in[threadIdx.x] += s_in[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s_ke[0];
    out[1] = s_vm[0];
    out[2] = s_in[0];
    out[3] = s_pe[0];
    out[4] = mass * (double)n;
  }
}

// FP64 maxima of |v|^2 and |a|^2 for the time-step criteria: per-block maxima,
// then one block over the partials.
__global__ __launch_bounds__(BLK) void k_max_va_partial(int n, const float* __restrict__ vx,
                                                        const float* __restrict__ vy,
                                                        const float* __restrict__ ax,
                                                        const float* __restrict__ ay,
                                                        double* __restrict__ part) {
  __shared__ double s_v[BLK], s_a[BLK];
  const int i = blockIdx.x * BLK + threadIdx.x;
  double v2 = 0, a2 = 0;
  if (i < n) {
    const double u = vx[i], v = vy[i], p = ax[i], q = ay[i];
    v2 = u * u + v * v;
    a2 = p * p + q * q;
  }
  s_v[threadIdx.x] = v2;
  s_a[threadIdx.x] = a2;
  __syncthreads();
  for (int o = BLK / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s_v[threadIdx.x] = fmax(s_v[threadIdx.x], s_v[threadIdx.x + o]);
      s_a[threadIdx.x] = fmax(s_a[threadIdx.x], s_a[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s_v[0];
    part[2 * blockIdx.x + 1] = s_a[0];
  }
}

__global__ __launch_bounds__(BLK) void k_max_va_final(int nparts, const double* __restrict__ part,
                                                      double* __restrict__ out) {
  __shared__ double s_v[BLK], s_a[BLK];
  double v = 0, a = 0;
  for (int b = threadIdx.x; b < nparts; b += BLK) {
    v = fmax(v, part[2 * b]);
    a = fmax(a, part[2 * b + 1]);
  }
  s_v[threadIdx.x] = v;
  s_a[threadIdx.x] = a;
  __syncthreads();
  for (int o = BLK / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s_v[threadIdx.x] = fmax(s_v[threadIdx.x], s_v[threadIdx.x + o]);
      s_a[threadIdx.x] = fmax(s_a[threadIdx.x], s_a[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s_v[0];
    out[1] = s_a[0];
  }
}

__global__ __launch_bounds__(BLK) void k_ell_overflow(int n, int w,
                                                      const int32_t* __restrict__ cnt,
                                                      unsigned long long* __restrict__ out) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  bool over = i < n && cnt[i] > w;
  uint64_t m = __ballot(over);
  if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicAdd(out, (unsigned long long)__popcll(m));
}

}  // namespace

// ------------------------------------------------------------ launchers
hipError_t launch_fill_cells(const Launch& L, int2* cells, int64_t n_keys) {
  int b = (int)std::min<int64_t>(blocks_for(n_keys), 8192);
  k_fill_cells<<<std::max(b, 1), BLK, 0, L.stream>>>(cells, n_keys);
  return hipGetLastError();
}

hipError_t launch_keys(const Launch& L, const DevState& s, uint64_t* keys_out) {
  if (s.n > 0) k_keys<<<blocks_for(s.n), BLK, 0, L.stream>>>(s.n, L.grid, s.rx, s.ry, keys_out);
  return hipGetLastError();
}

hipError_t launch_gather(const Launch& L, const DevState& s, const float* src_rx,
                         const float* src_ry, const float* src_vx, const float* src_vy,
                         const uint32_t* src_pid, const float* src_ax, const float* src_ay,
                         float* dst_rx, float* dst_ry, float* dst_vx, float* dst_vy,
                         uint32_t* dst_pid, float* dst_ax, float* dst_ay, int hist_words) {
  if (s.n > 0)
    k_gather<<<blocks_for(s.n), BLK, 0, L.stream>>>(
        s.n, s.perm, s.keys_sorted, src_rx, src_ry, src_vx, src_vy, src_pid, src_ax, src_ay,
        dst_rx, dst_ry, dst_vx, dst_vy, dst_pid, dst_ax, dst_ay, s.cells, s.n_keys, s.hist,
        hist_words, s.blocks, s.block_bits, s.tile_first, s.tile_count, s.xyuv);
  return hipGetLastError();
}

hipError_t launch_ranges(const Launch& L, const DevState& s, const ResortBuffers& r,
                         int hist_words) {
  if (s.n > 0) {
    const int grid = (blocks_for(s.n) + kRangesPer - 1) / kRangesPer;
    if (s.narrow)
      k_ranges<uint32_t><<<grid, BLK, 0, L.stream>>>(s.n, s.nkeys, s.cells, s.n_keys, s.blocks,
                                                     s.block_bits, s.hist, hist_words,
                                                     s.tile_first, s.tile_count, r);
    else
      k_ranges<uint64_t><<<grid, BLK, 0, L.stream>>>(s.n, s.keys_sorted, s.cells, s.n_keys,
                                                     s.blocks, s.block_bits, s.hist, hist_words,
                                                     s.tile_first, s.tile_count, r);
  }
  return hipGetLastError();
}

template <int PHYS, int BC>
static hipError_t density_dispatch(const Launch& L, const DevState& s, int nb, bool write_p,
                                   bool packed) {
  float* p = write_p ? s.pres : nullptr;
  const int nb_ = s.n > 0 ? blocks_for(s.n) : 0;
  if (nb_ == 0) return hipSuccess;
  if (nb == NB_CSR)
    k_density<PHYS, BC, NB_CSR><<<nb_, BLK, 0, L.stream>>>(
        s.n, L.grid, L.phys, s.rx, s.ry, s.cells, s.csr_off, s.csr_idx, s.ell_count, s.ell_idx,
        s.ell_width, s.rho, s.pr, p, nullptr, nullptr);
  else if (nb == NB_ELL)
    k_density<PHYS, BC, NB_ELL><<<nb_, BLK, 0, L.stream>>>(
        s.n, L.grid, L.phys, s.rx, s.ry, s.cells, s.csr_off, s.csr_idx, s.ell_count, s.ell_idx,
        s.ell_width, s.rho, s.pr, p, packed ? s.xyuv : nullptr, packed ? s.rp : nullptr);
  else
    k_density<PHYS, BC, NB_STENCIL><<<nb_, BLK, 0, L.stream>>>(
        s.n, L.grid, L.phys, s.rx, s.ry, s.cells, s.csr_off, s.csr_idx, s.ell_count, s.ell_idx,
        s.ell_width, s.rho, s.pr, p, nullptr, nullptr);
  return hipGetLastError();
}

hipError_t launch_density(const Launch& L, const DevState& s, int nb, bool write_p,
                          bool packed) {
  const int ph = L.phys.physics, bc = L.phys.boundary;
  if (ph == 0 && bc == 0) return density_dispatch<0, 0>(L, s, nb, write_p, packed);
  if (ph == 0 && bc == 1) return density_dispatch<0, 1>(L, s, nb, write_p, packed);
  if (ph == 1 && bc == 0) return density_dispatch<1, 0>(L, s, nb, write_p, packed);
  return density_dispatch<1, 1>(L, s, nb, write_p, packed);
}

// The fields of the last iteration of sph_step (computed before its final
// re-sort) moved into the current order: out[i] = field[perm[i]]. p is
// recomputed from rho with the density kernel's EOS (same function, same
// input: the same bits) unless the step wrote it (batched scenes, whose
// constants differ per scene), in which case it is gathered.
template <int PHYS>
__global__ __launch_bounds__(BLK) void k_settle_fields(
    int n, DevPhys c, const uint32_t* __restrict__ perm, const float* __restrict__ ax,
    const float* __restrict__ ay, const float* __restrict__ rho, const float* __restrict__ pr,
    const float* __restrict__ pres, float* __restrict__ oax, float* __restrict__ oay,
    float* __restrict__ orho, float* __restrict__ opr, float* __restrict__ opres) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = min(perm[i], (uint32_t)(n - 1));  // in bounds even if perm were bad
  const float r = rho[j];
  oax[i] = ax[j];
  oay[i] = ay[j];
  orho[i] = r;
  opr[i] = pr[j];
  opres[i] = pres ? pres[j] : eos_pressure<PHYS>(c, r);
}

hipError_t launch_settle_fields(const Launch& L, const DevState& s, bool gather_p,
                                float* const* out5) {
  if (s.n <= 0) return hipSuccess;
  const float* pres = gather_p ? s.pres : nullptr;
  if (L.phys.physics == 0)
    k_settle_fields<0><<<blocks_for(s.n), BLK, 0, L.stream>>>(
        s.n, L.phys, s.perm, s.ax, s.ay, s.rho, s.pr, pres, out5[0], out5[1], out5[2], out5[3],
        out5[4]);
  else
    k_settle_fields<1><<<blocks_for(s.n), BLK, 0, L.stream>>>(
        s.n, L.phys, s.perm, s.ax, s.ay, s.rho, s.pr, pres, out5[0], out5[1], out5[2], out5[3],
        out5[4]);
  return hipGetLastError();
}

hipError_t launch_unpack(const Launch& L, const DevState& s) {
  if (s.n > 0)
    k_unpack<<<blocks_for(s.n), BLK, 0, L.stream>>>(s.n, s.xyuv, s.rx, s.ry, s.vx, s.vy,
                                                    s.narrow ? s.nkeys : nullptr, s.keys_sorted);
  return hipGetLastError();
}

hipError_t launch_pack(const Launch& L, const DevState& s) {
  if (s.n > 0)
    k_pack<<<blocks_for(s.n), BLK, 0, L.stream>>>(s.n, s.rx, s.ry, s.vx, s.vy, s.xyuv,
                                                  s.keys_sorted, s.narrow ? s.nkeys : nullptr);
  return hipGetLastError();
}

template <int PHYS, int BC, int NB, bool FUSED>
static hipError_t forces_go(const Launch& L, const DevState& s, uint64_t* keys_out,
                            MoverSink movers, bool packed) {
  if constexpr (NB == NB_ELL && FUSED) {
    if (s.narrow) {  // the step's narrow keys: read nkeys, write nkeys_in
      k_forces<PHYS, BC, NB, FUSED, uint32_t><<<blocks_for(s.n), BLK, 0, L.stream>>>(
          s.n, L.grid, L.phys, s.rx, s.ry, s.vx, s.vy, s.rho, s.pr, s.cells, s.csr_off,
          s.csr_idx, s.ell_count, s.ell_idx, s.ell_width, s.ax, s.ay, s.tx, s.ty, s.tvx, s.tvy,
          s.nkeys_in, s.nkeys, movers, packed ? s.xyuv : nullptr, packed ? s.rp : nullptr,
          s.t4);
      return hipGetLastError();
    }
  }
  k_forces<PHYS, BC, NB, FUSED><<<blocks_for(s.n), BLK, 0, L.stream>>>(
      s.n, L.grid, L.phys, s.rx, s.ry, s.vx, s.vy, s.rho, s.pr, s.cells, s.csr_off, s.csr_idx,
      s.ell_count, s.ell_idx, s.ell_width, s.ax, s.ay, s.tx, s.ty, s.tvx, s.tvy, keys_out,
      s.keys_sorted, FUSED ? movers : MoverSink{}, packed && NB == NB_ELL ? s.xyuv : nullptr,
      packed && NB == NB_ELL ? s.rp : nullptr, nullptr);
  return hipGetLastError();
}

template <int PHYS, int BC>
static hipError_t forces_dispatch(const Launch& L, const DevState& s, int nb, bool fused,
                                  uint64_t* keys_out, MoverSink movers, bool packed) {
  if (s.n <= 0) return hipSuccess;
  if (fused) {
    if (nb == NB_ELL) return forces_go<PHYS, BC, NB_ELL, true>(L, s, keys_out, movers, packed);
    if (nb == NB_CSR) return forces_go<PHYS, BC, NB_CSR, true>(L, s, keys_out, movers, false);
    return forces_go<PHYS, BC, NB_STENCIL, true>(L, s, keys_out, movers, false);
  }
  if (nb == NB_ELL)
    return forces_go<PHYS, BC, NB_ELL, false>(L, s, keys_out, MoverSink{}, packed);
  if (nb == NB_CSR)
    return forces_go<PHYS, BC, NB_CSR, false>(L, s, keys_out, MoverSink{}, false);
  return forces_go<PHYS, BC, NB_STENCIL, false>(L, s, keys_out, MoverSink{}, false);
}

hipError_t launch_forces(const Launch& L, const DevState& s, int nb, bool fused,
                         uint64_t* keys_out, MoverSink movers, bool packed) {
  const int ph = L.phys.physics, bc = L.phys.boundary;
  if (ph == 0 && bc == 0)
    return forces_dispatch<0, 0>(L, s, nb, fused, keys_out, movers, packed);
  if (ph == 0 && bc == 1)
    return forces_dispatch<0, 1>(L, s, nb, fused, keys_out, movers, packed);
  if (ph == 1 && bc == 0)
    return forces_dispatch<1, 0>(L, s, nb, fused, keys_out, movers, packed);
  return forces_dispatch<1, 1>(L, s, nb, fused, keys_out, movers, packed);
}

hipError_t launch_integrate(const Launch& L, const DevState& s) {
  if (s.n <= 0) return hipSuccess;
  if (L.phys.boundary == 1)
    k_integrate<1><<<blocks_for(s.n), BLK, 0, L.stream>>>(s.n, L.phys, s.rx, s.ry, s.vx, s.vy,
                                                          s.ax, s.ay);
  else
    k_integrate<0><<<blocks_for(s.n), BLK, 0, L.stream>>>(s.n, L.phys, s.rx, s.ry, s.vx, s.vy,
                                                          s.ax, s.ay);
  return hipGetLastError();
}

hipError_t launch_update_velocities(const Launch& L, const DevState& s) {
  if (s.n <= 0) return hipSuccess;
  k_update_velocities<<<blocks_for(s.n), BLK, 0, L.stream>>>(s.n, L.phys, s.vx, s.vy, s.ax,
                                                              s.ay);
  return hipGetLastError();
}

hipError_t launch_update_positions(const Launch& L, const DevState& s) {
  if (s.n <= 0) return hipSuccess;
  if (L.phys.boundary == 1)
    k_update_positions<1><<<blocks_for(s.n), BLK, 0, L.stream>>>(s.n, L.phys, s.rx, s.ry, s.vx,
                                                                 s.vy);
  else
    k_update_positions<0><<<blocks_for(s.n), BLK, 0, L.stream>>>(s.n, L.phys, s.rx, s.ry, s.vx,
                                                                 s.vy);
  return hipGetLastError();
}

template <int BC, bool SELF>
static void csr_count_go(const Launch& L, const DevState& s) {
  k_csr_count<BC, SELF><<<blocks_for(s.n), BLK, 0, L.stream>>>(s.n, L.grid, L.phys, s.rx, s.ry,
                                                               s.cells, s.csr_count);
}
template <int BC, bool SELF>
static void csr_fill_go(const Launch& L, const DevState& s) {
  k_csr_fill<BC, SELF><<<blocks_for(s.n), BLK, 0, L.stream>>>(s.n, L.grid, L.phys, s.rx, s.ry,
                                                              s.cells, s.csr_off, s.csr_idx);
}

hipError_t launch_csr_count(const Launch& L, const DevState& s, bool self) {
  if (s.n <= 0) return hipSuccess;
  const int bc = L.phys.boundary;
  if (bc == 0) self ? csr_count_go<0, true>(L, s) : csr_count_go<0, false>(L, s);
  else self ? csr_count_go<1, true>(L, s) : csr_count_go<1, false>(L, s);
  return hipGetLastError();
}

hipError_t launch_csr_fill(const Launch& L, const DevState& s, bool self) {
  if (s.n <= 0) return hipSuccess;
  const int bc = L.phys.boundary;
  if (bc == 0) self ? csr_fill_go<0, true>(L, s) : csr_fill_go<0, false>(L, s);
  else self ? csr_fill_go<1, true>(L, s) : csr_fill_go<1, false>(L, s);
  return hipGetLastError();
}

int64_t scan_scratch_words(int n) { return (int64_t)blocks_for(n, SCAN_BLK) + 2; }

hipError_t launch_scan(const Launch& L, const int32_t* in, int64_t* out, int n,
                       int64_t* scratch) {
  if (n <= 0) return hipSuccess;
  const int nb = blocks_for(n, SCAN_BLK);
  k_scan_reduce<<<nb, SCAN_BLK, 0, L.stream>>>(in, n, scratch);
  k_scan_top<<<1, SCAN_BLK, 0, L.stream>>>(scratch, nb);
  k_scan_down<<<nb, SCAN_BLK, 0, L.stream>>>(in, n, scratch, out);
  return hipGetLastError();
}

int64_t measure_partial_words(int n) { return 4 * (int64_t)std::max(blocks_for(n), 1); }

hipError_t launch_max_va(const Launch& L, const DevState& s, double* partials, double* out2) {
  const int nb = std::max(blocks_for(s.n), 1);
  k_max_va_partial<<<nb, BLK, 0, L.stream>>>(s.n, s.vx, s.vy, s.ax, s.ay, partials);
  k_max_va_final<<<1, BLK, 0, L.stream>>>(nb, partials, out2);
  return hipGetLastError();
}

hipError_t launch_measure(const Launch& L, const DevState& s, double mass, double gravity_y,
                          double* partials, double* out5) {
  const int nb = std::max(blocks_for(s.n), 1);
  k_measure_partial<<<nb, BLK, 0, L.stream>>>(s.n, L.phys, mass, -gravity_y, s.rx, s.ry, s.vx,
                                              s.vy, partials);
  k_measure_final<<<1, BLK, 0, L.stream>>>(nb, s.n, mass, partials, out5);
  return hipGetLastError();
}

hipError_t launch_ell_overflow(const Launch& L, const DevState& s, int64_t* out_count) {
  if (s.n <= 0) return hipSuccess;
  k_ell_overflow<<<blocks_for(s.n), BLK, 0, L.stream>>>(
      s.n, s.ell_width, s.ell_count, reinterpret_cast<unsigned long long*>(out_count));
  return hipGetLastError();
}

}  // namespace sphamd
