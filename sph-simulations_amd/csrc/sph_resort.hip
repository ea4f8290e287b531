// Incremental re-sort of an already sorted particle set (gfx950).
//
// Between two steps a particle moves a small fraction of a cell, so almost every
// particle keeps its cell key. The particles whose key changed ("movers") are
// compacted and radix-sorted on their own; the rest ("stayers") are already in
// key order, because they are a subsequence of the previous sorted order with
// unchanged keys. The new order is the stable merge of the two sorted sequences
// by (key, previous index) -- exactly the order a stable sort of all keys gives.
//
//   k_mover_flags    flag movers per 256-particle tile, count them, clear the
//                    cell / block ranges of the previous order
//   k_tile_scan      exclusive scan of the tile counts (one workgroup)
//   k_mover_compact  movers in previous-index order
//   k_small_sort     up to kSmallSortCap movers sorted in LDS (one workgroup);
//                    above that the radix sort runs on the device-side count
//   k_merge   final slot of every particle: stayer j -> (stayers before j) +
//             (movers ordered before it); mover r -> r + (stayers ordered before
//             it); scatter r, v, pid, key and the permutation
//   k_ranges  cell start/end and Morton block ranges of the new order
#include <hip/hip_runtime.h>

#include "sph_internal.hpp"

namespace sphamd {

namespace {

constexpr int TW = 256;  // particles per tile / threads per workgroup
constexpr uint32_t FLAG_AGG = 1u << 30;
constexpr uint32_t FLAG_PREFIX = 2u << 30;
constexpr uint32_t COUNT_MASK = (1u << 30) - 1;
constexpr uint32_t MAX_POLLS = 1u << 18;
constexpr int WIN = 512;  // movers a merge tile may hold in LDS

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// (ka, ia) < (kb, ib) lexicographically
__device__ __forceinline__ bool pair_less(uint64_t ka, uint32_t ia, uint64_t kb, uint32_t ib) {
  return ka < kb || (ka == kb && ia < ib);
}

// Movers of each tile: one ballot word per wavefront and the tile's count (into
// tile_base, scanned in place by k_tile_scan). Also clears the cell / block
// ranges of the previous order (k_ranges writes the new ones).
__global__ __launch_bounds__(TW) void k_mover_flags(
    int n, const uint64_t* __restrict__ knew, const uint64_t* __restrict__ kold,
    int2* __restrict__ cells, int2* __restrict__ blocks, int block_bits, uint64_t n_keys,
    unsigned long long* __restrict__ flags, uint32_t* __restrict__ tile_count) {
  __shared__ uint32_t s_wcnt[TW / kWave];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int tile = blockIdx.x;
  const int i = tile * TW + tid;
  const bool live = i < n;
  const uint64_t kn = live ? knew[i] : 0, ko = live ? kold[i] : 0;
  if (live) {
    const uint64_t kp = i > 0 ? kold[i - 1] : ~ko;
    if (ko != kp && ko < n_keys) cells[ko] = make_int2(-1, -1);
    if (block_bits > 0 && (ko >> (2 * block_bits)) != (kp >> (2 * block_bits)) && ko < n_keys)
      blocks[ko >> (2 * block_bits)] = make_int2(-1, -1);
  }
  const unsigned long long bal = __ballot(live && kn != ko);
  if (lane == 0) {
    s_wcnt[wave] = (uint32_t)__popcll(bal);
    flags[(size_t)tile * (TW / kWave) + wave] = bal;
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t c = 0;
    for (int w = 0; w < TW / kWave; ++w) c += s_wcnt[w];
    tile_count[tile] = c;
  }
}

// In-place exclusive scan of the per-tile mover counts (one workgroup, fixed
// order) and the total.
constexpr int SCAN_TW = 1024;
__global__ __launch_bounds__(SCAN_TW) void k_tile_scan(int tiles, uint32_t* __restrict__ tile_base,
                                                       int* __restrict__ n_movers) {
  __shared__ uint32_t s_w[SCAN_TW / kWave];
  __shared__ uint32_t s_carry;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  if (tid == 0) s_carry = 0;
  __syncthreads();
  for (int base = 0; base < tiles; base += SCAN_TW) {
    const int t = base + tid;
    const uint32_t v = t < tiles ? tile_base[t] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == kWave - 1) s_w[wave] = x;
    __syncthreads();
    uint32_t off = s_carry;
    for (int w = 0; w < wave; ++w) off += s_w[w];
    if (t < tiles) tile_base[t] = off + x - v;
    __syncthreads();
    if (tid == SCAN_TW - 1) s_carry = off + x;
    __syncthreads();
  }
  if (tid == 0) *n_movers = (int)s_carry;
}

// Movers in previous-index order: (new key, previous index).
__global__ __launch_bounds__(TW) void k_mover_compact(
    int n, const uint64_t* __restrict__ knew, const unsigned long long* __restrict__ flags,
    const uint32_t* __restrict__ tile_base, uint64_t* __restrict__ mkey,
    uint32_t* __restrict__ midx) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int tile = blockIdx.x;
  const int i = tile * TW + tid;
  const unsigned long long fw = flags[(size_t)tile * (TW / kWave) + wave];
  if (i >= n || !((fw >> lane) & 1ull)) return;
  uint32_t r = tile_base[tile];
  for (int w = 0; w < wave; ++w) r += (uint32_t)__popcll(flags[(size_t)tile * (TW / kWave) + w]);
  r += (uint32_t)__popcll(fw & ((1ull << lane) - 1ull));
  mkey[r] = knew[i];
  midx[r] = (uint32_t)i;
}

// Movers sorted by (key, previous index) in one workgroup: a bitonic network in
// LDS over the next power of two (sentinels sort last). Does nothing when the
// count exceeds SMALL_CAP: the radix sort handles those.
constexpr int SMALL_TW = 1024;
__global__ __launch_bounds__(SMALL_TW) void k_small_sort(const int* __restrict__ n_dev,
                                                         const uint64_t* __restrict__ mkey,
                                                         const uint32_t* __restrict__ midx,
                                                         uint64_t* __restrict__ skey,
                                                         uint32_t* __restrict__ sidx) {
  __shared__ uint64_t s_k[kSmallSortCap];
  __shared__ uint32_t s_i[kSmallSortCap];
  const int n = *n_dev;
  if (n <= 0 || n > kSmallSortCap) return;
  int m = 1;
  while (m < n) m <<= 1;
  for (int t = threadIdx.x; t < m; t += SMALL_TW) {
    s_k[t] = t < n ? mkey[t] : ~0ull;
    s_i[t] = t < n ? midx[t] : ~0u;
  }
  __syncthreads();
  for (int size = 2; size <= m; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < (m >> 1); t += SMALL_TW) {
        const int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
        const bool up = (lo & size) == 0;
        const uint64_t ka = s_k[lo], kb = s_k[hi];
        const uint32_t ia = s_i[lo], ib = s_i[hi];
        if (pair_less(kb, ib, ka, ia) == up) {
          s_k[lo] = kb;
          s_k[hi] = ka;
          s_i[lo] = ib;
          s_i[hi] = ia;
        }
      }
      __syncthreads();
    }
  }
  for (int t = threadIdx.x; t < n; t += SMALL_TW) {
    skey[t] = s_k[t];
    sidx[t] = s_i[t];
  }
}

// Movers among the previous-order indices [0, p) (p <= n).
__device__ __forceinline__ uint32_t movers_before(int p, int n, int n_mov,
                                                  const unsigned long long* flags,
                                                  const uint32_t* tile_base) {
  if (p >= n) return (uint32_t)n_mov;
  const int tile = p / TW, within = p % TW;
  uint32_t r = tile_base[tile];
  const int w = within / kWave, l = within % kWave;
  for (int k = 0; k < w; ++k) r += (uint32_t)__popcll(flags[(size_t)tile * (TW / kWave) + k]);
  r += (uint32_t)__popcll(flags[(size_t)tile * (TW / kWave) + w] & ((1ull << l) - 1ull));
  return r;
}

// Number of previous-order particles j with (kold[j], j) < (k, idx): the
// sequence (kold[j], j) is strictly increasing in j.
__device__ __forceinline__ int rank_in_previous(const uint64_t* kold, int n, uint64_t k,
                                                uint32_t idx) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (pair_less(kold[mid], (uint32_t)mid, k, idx)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Number of sorted movers r in [lo, hi) with (skey[r], sidx[r]) < (k, idx).
__device__ __forceinline__ int rank_in_movers(const uint64_t* skey, const uint32_t* sidx, int lo,
                                              int hi, uint64_t k, uint32_t idx) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (pair_less(skey[mid], sidx[mid], k, idx)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void scatter(int pos, int src, uint64_t key, const float* tx,
                                        const float* ty, const float* tvx, const float* tvy,
                                        const uint32_t* pid_src, float* rx, float* ry,
                                        float* vx, float* vy, uint32_t* pid_dst,
                                        uint64_t* keys_dst, uint32_t* perm) {
  rx[pos] = tx[src];
  ry[pos] = ty[src];
  vx[pos] = tvx[src];
  vy[pos] = tvy[src];
  pid_dst[pos] = pid_src[src];
  keys_dst[pos] = key;
  perm[pos] = (uint32_t)src;
}

// Grid: ceil(n / TW) workgroups for the previous order (stayers), then as many
// again for the sorted movers (those past the device-side mover count exit).
__global__ __launch_bounds__(TW) void k_merge(
    int n, int tiles, const uint64_t* __restrict__ kold,
    const unsigned long long* __restrict__ flags, const uint32_t* __restrict__ tile_base,
    const uint64_t* __restrict__ skey, const uint32_t* __restrict__ sidx,
    const int* __restrict__ n_movers_p, const float* __restrict__ tx,
    const float* __restrict__ ty, const float* __restrict__ tvx,
    const float* __restrict__ tvy, const uint32_t* __restrict__ pid_src,
    float* __restrict__ rx, float* __restrict__ ry, float* __restrict__ vx,
    float* __restrict__ vy, uint32_t* __restrict__ pid_dst, uint64_t* __restrict__ keys_dst,
    uint32_t* __restrict__ perm) {
  __shared__ uint64_t s_k[WIN];
  __shared__ uint32_t s_i[WIN];
  __shared__ int s_lo, s_hi;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int n_mov = *n_movers_p;
  if ((int)blockIdx.x < tiles) {
    const int tile = blockIdx.x;
    const int j = tile * TW + tid;
    const bool live = j < n;
    const unsigned long long fw = flags[(size_t)tile * (TW / kWave) + wave];
    const bool mv = live && ((fw >> lane) & 1ull);
    // window of sorted movers that can fall between this tile's first and last
    // previous-order elements
    const int j0 = tile * TW, j1 = min(n, j0 + TW) - 1;
    if (tid == 0) s_lo = rank_in_movers(skey, sidx, 0, n_mov, kold[j0], (uint32_t)j0);
    if (tid == 1) s_hi = rank_in_movers(skey, sidx, 0, n_mov, kold[j1], (uint32_t)j1);
    __syncthreads();
    const int lo = s_lo, hi = s_hi;
    const bool windowed = hi - lo <= WIN;
    if (windowed)
      for (int r = lo + tid; r < hi; r += TW) {
        s_k[r - lo] = skey[r];
        s_i[r - lo] = sidx[r];
      }
    __syncthreads();
    if (live && !mv) {
      const uint64_t k = kold[j];
      uint32_t before = tile_base[tile];
      for (int w = 0; w < wave; ++w)
        before += (uint32_t)__popcll(flags[(size_t)tile * (TW / kWave) + w]);
      before += (uint32_t)__popcll(fw & ((1ull << lane) - 1ull));
      int b;
      if (windowed) {
        int a = 0, z = hi - lo;
        while (a < z) {
          const int mid = (a + z) >> 1;
          if (pair_less(s_k[mid], s_i[mid], k, (uint32_t)j)) a = mid + 1;
          else z = mid;
        }
        b = lo + a;
      } else {
        b = rank_in_movers(skey, sidx, lo, hi, k, (uint32_t)j);
      }
      const int pos = (j - (int)before) + b;
      scatter(pos, j, k, tx, ty, tvx, tvy, pid_src, rx, ry, vx, vy, pid_dst, keys_dst, perm);
    }
  } else {
    const int r = ((int)blockIdx.x - tiles) * TW + tid;
    if (r >= n_mov) return;
    const uint64_t k = skey[r];
    const uint32_t src = sidx[r];
    const int p = rank_in_previous(kold, n, k, src);
    const int stayers_before = p - (int)movers_before(p, n, n_mov, flags, tile_base);
    scatter(r + stayers_before, (int)src, k, tx, ty, tvx, tvy, pid_src, rx, ry, vx, vy, pid_dst,
            keys_dst, perm);
  }
}

}  // namespace

int64_t resort_flag_words(int n) { return ((int64_t)n + TW - 1) / TW * (TW / kWave); }
int64_t resort_tiles(int n) { return ((int64_t)n + TW - 1) / TW; }

hipError_t launch_movers(const Launch& L, const DevState& s, const ResortBuffers& r,
                         const uint64_t* knew) {
  if (s.n <= 0) return hipSuccess;
  const int tiles = (int)resort_tiles(s.n);
  k_mover_flags<<<tiles, TW, 0, L.stream>>>(s.n, knew, s.keys_sorted, s.cells, s.blocks,
                                            s.block_bits, (uint64_t)s.n_keys, r.flags,
                                            r.tile_base);
  k_tile_scan<<<1, SCAN_TW, 0, L.stream>>>(tiles, r.tile_base, r.n_movers);
  k_mover_compact<<<tiles, TW, 0, L.stream>>>(s.n, knew, r.flags, r.tile_base, r.mkey, r.midx);
  return hipGetLastError();
}

hipError_t launch_small_sort(const Launch& L, const ResortBuffers& r) {
  k_small_sort<<<1, SMALL_TW, 0, L.stream>>>(r.n_movers, r.mkey, r.midx, r.skey, r.sidx);
  return hipGetLastError();
}

hipError_t launch_merge(const Launch& L, const DevState& s, const ResortBuffers& r,
                        uint64_t* keys_dst) {
  if (s.n <= 0) return hipSuccess;
  const int tiles = (int)resort_tiles(s.n);
  k_merge<<<2 * tiles, TW, 0, L.stream>>>(
      s.n, tiles, s.keys_sorted, r.flags, r.tile_base, r.skey, r.sidx, r.n_movers, s.tx, s.ty,
      s.tvx, s.tvy, s.pid, s.rx, s.ry, s.vx, s.vy, s.pid_alt, keys_dst, s.perm);
  return hipGetLastError();
}

}  // namespace sphamd
