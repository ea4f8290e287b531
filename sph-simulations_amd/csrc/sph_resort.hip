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
//   k_mover_sort     movers sorted by (key, previous index): one workgroup in
//                    LDS for up to kSmallSortCap movers, a grid-wide radix
//                    sort above that (one launch either way)
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
    unsigned long long* __restrict__ flags, uint32_t* __restrict__ tile_count,
    uint32_t* __restrict__ bar) {
  __shared__ uint32_t s_wcnt[TW / kWave];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int tile = blockIdx.x;
  if (tile == 0 && tid == 0) *bar = 0;  // the mover sort's grid barrier
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

// Movers in previous-index order: (new key, previous index), one thread per
// particle.
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

// ------------------------------------------------------------ mover sort
// The compacted movers sorted by (new key, previous index) into skey / sidx, in
// one launch of kMoverGrid workgroups: with at most kSmallSortCap movers
// workgroup 0 sorts them in LDS with a bitonic network (the others return at
// once); above that all workgroups run an LSD radix sort of 8-bit digits (per
// pass: digit counts per workgroup chunk, then a stable scatter in
// 1024-element sub-tiles with wave64 ballot ranking, a grid barrier after each
// phase). The grid is small enough to be resident at once; every barrier wait
// is bounded and reports to the status word if it gives up.
constexpr int MS_TW = 1024;  // threads per mover-sort workgroup
constexpr int MS_WAVES = MS_TW / kWave;
// the grid radix path reads every workgroup's 256 digit counts into the LDS of
// the small-sort key buffers
static_assert(kMoverGrid * 256 <= 4 * kSmallSortCap, "digit counts exceed the LDS arrays");
constexpr uint32_t MS_POLLS = 1u << 20;

// Grid-wide barrier on a monotonic counter (zeroed by k_mover_flags every step).
__device__ __forceinline__ bool grid_barrier(uint32_t* bar, uint32_t& gen, uint32_t* status,
                                             int* s_ok) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    atomicAdd(bar, 1u);
    const uint32_t target = (gen + 1) * gridDim.x;
    uint32_t polls = 0;
    int ok = 1;
    while (ld_agent(bar) < target) {
      if (++polls >= MS_POLLS) {
        atomicOr(status, 4u);  // never expected: the grid was not resident
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __threadfence();
    *s_ok = ok;
  }
  __syncthreads();
  ++gen;
  return *s_ok != 0;
}

// In-place exclusive scan of 256 digit counts by the first four wavefronts
// (scratch: 4 words); ends with a workgroup barrier.
__device__ __forceinline__ void digit_excl_scan(uint32_t* v, uint32_t* scratch) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  uint32_t x = 0, c = 0;
  if (tid < 256) {
    c = v[tid];
    x = c;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == kWave - 1) scratch[wave] = x;
  }
  __syncthreads();
  if (tid < 256) {
    uint32_t off = 0;
    for (int w = 0; w < wave; ++w) off += scratch[w];
    v[tid] = off + x - c;
  }
  __syncthreads();
}

// Stable rank of this lane's digit d among the valid lanes of its wave with a
// lower lane id and the same digit; *same gets the mask of those lanes' group.
__device__ __forceinline__ uint32_t wave_digit_rank(uint32_t d, bool valid,
                                                    unsigned long long* same) {
  const int lane = threadIdx.x & (kWave - 1);
  unsigned long long m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const unsigned long long bb = __ballot((d >> b) & 1u);
    m &= ((d >> b) & 1u) ? bb : ~bb;
  }
  *same = m;
  return (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

__global__ __launch_bounds__(MS_TW) void k_mover_sort(
    int passes, const int* __restrict__ n_movers, uint64_t* __restrict__ mkey,
    uint32_t* __restrict__ midx, uint64_t* __restrict__ ktmp, uint32_t* __restrict__ vtmp,
    uint64_t* __restrict__ skey, uint32_t* __restrict__ sidx, uint32_t* __restrict__ counts,
    uint32_t* __restrict__ bar, uint32_t* __restrict__ status) {
  // small case: two key/value buffers (ping-pong) of kSmallSortCap each
  __shared__ uint64_t s_k[2 * kSmallSortCap];
  __shared__ uint32_t s_i[2 * kSmallSortCap];
  __shared__ uint32_t s_wc[MS_WAVES * 256];  // per-wave digit counts -> offsets
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_tot[256];
  __shared__ int s_ok;
  const int tid = threadIdx.x, wave = tid / kWave;
  const uint32_t total = (uint32_t)*n_movers;
  if (total <= (uint32_t)kSmallSortCap) {
    if (blockIdx.x != 0) return;
    // LSD radix sort in LDS: per 8-bit pass, digit counts of all elements give
    // each digit's start; elements are then placed in order, 1024 at a time,
    // at start + (same digit earlier in this round) + (same digit earlier in
    // the wave) -- stable, so ties keep their (key, previous index) order
    const int m = (int)total;
    for (int t = tid; t < m; t += MS_TW) {
      s_k[t] = mkey[t];
      s_i[t] = midx[t];
    }
    int cur = 0;
    for (int p = 0; p < passes; ++p) {
      const int shift = 8 * p;
      uint64_t* kin = s_k + cur * kSmallSortCap;
      uint32_t* vin = s_i + cur * kSmallSortCap;
      uint64_t* kout = s_k + (cur ^ 1) * kSmallSortCap;
      uint32_t* vout = s_i + (cur ^ 1) * kSmallSortCap;
      if (tid < 256) s_base[tid] = 0;
      __syncthreads();
      for (int t = tid; t < m; t += MS_TW) atomicAdd(&s_base[(uint32_t)(kin[t] >> shift) & 255u], 1u);
      __syncthreads();
      digit_excl_scan(s_base, s_tot);

      for (int base = 0; base < m; base += MS_TW) {
        for (int q = tid; q < MS_WAVES * 256; q += MS_TW) s_wc[q] = 0;
        __syncthreads();
        const int e = base + tid;
        const bool valid = e < m;
        const uint64_t key = valid ? kin[e] : 0ull;
        const uint32_t val = valid ? vin[e] : 0u;
        const uint32_t d = (uint32_t)(key >> shift) & 255u;
        unsigned long long same;
        const uint32_t wrank = wave_digit_rank(d, valid, &same);
        if (valid && wrank == 0) s_wc[wave * 256 + d] = (uint32_t)__popcll(same);
        __syncthreads();
        if (tid < 256) {
          uint32_t run = 0;
          for (int w = 0; w < MS_WAVES; ++w) {
            const uint32_t c = s_wc[w * 256 + tid];
            s_wc[w * 256 + tid] = run;
            run += c;
          }
          s_tot[tid] = run;
        }
        __syncthreads();
        if (valid) {
          const uint32_t pos = s_base[d] + s_wc[wave * 256 + d] + wrank;
          kout[pos] = key;
          vout[pos] = val;
        }
        __syncthreads();
        if (tid < 256) s_base[tid] += s_tot[tid];
        __syncthreads();
      }
      cur ^= 1;
    }
    for (int t = tid; t < m; t += MS_TW) {
      skey[t] = s_k[cur * kSmallSortCap + t];
      sidx[t] = s_i[cur * kSmallSortCap + t];
    }
    return;
  }
  // ---- more movers than one workgroup sorts: LSD radix sort over the grid
  uint32_t gen = 0;
  uint32_t* s_cnt = s_wc;  // [MS_WAVES][256] per-wave digit counts
  uint32_t* s_off = s_base;  // [256] running digit offsets
  const int nb = gridDim.x;
  const int chunk = (int)((total + (uint32_t)(nb * MS_TW) - 1) / (uint32_t)(nb * MS_TW)) * MS_TW;
  const int c0 = min((int)total, (int)blockIdx.x * chunk);
  const int c1 = min((int)total, c0 + chunk);
  const int lane = tid & (kWave - 1);
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  for (int p = 0; p < passes; ++p) {
    const uint64_t* kin = p == 0 ? mkey : ((p & 1) ? ktmp : mkey);
    const uint32_t* vin = p == 0 ? midx : ((p & 1) ? vtmp : midx);
    uint64_t* kout = p == passes - 1 ? skey : ((p & 1) ? mkey : ktmp);
    uint32_t* vout = p == passes - 1 ? sidx : ((p & 1) ? midx : vtmp);
    const int shift = 8 * p;
    // A: digit counts of this workgroup's chunk
    if (tid < 256) s_off[tid] = 0;
    __syncthreads();
    for (int e = c0 + tid; e < c1; e += MS_TW)
      atomicAdd(&s_off[(uint32_t)(kin[e] >> shift) & 255u], 1u);
    __syncthreads();
    if (tid < 256) counts[(size_t)blockIdx.x * 256 + tid] = s_off[tid];
    if (!grid_barrier(bar, gen, status, &s_ok)) return;
    // B: where each digit of this chunk starts: every smaller digit, then the
    // same digit in lower chunks. The nb x 256 counts come into LDS with one
    // round of independent loads, then thread d sums its column.
    uint32_t* s_all = reinterpret_cast<uint32_t*>(s_k);  // [nb][256]
    for (int q = tid; q < nb * 256; q += MS_TW) s_all[q] = counts[q];
    __syncthreads();
    uint32_t col = 0, before = 0;
    if (tid < 256) {
      for (int w = 0; w < nb; ++w) {
        const uint32_t c = s_all[w * 256 + tid];
        if (w < (int)blockIdx.x) before += c;
        col += c;
      }
      s_cnt[tid] = col;
      s_cnt[256 + tid] = before;
    }
    __syncthreads();
    if (tid < 256) s_off[tid] = s_cnt[tid];
    __syncthreads();
    digit_excl_scan(s_off, s_tot);
    if (tid < 256) s_off[tid] += s_cnt[256 + tid];
    __syncthreads();
    // C: stable scatter in sub-tiles of MS_TW elements
    for (int base = c0; base < c1; base += MS_TW) {
      for (int q = tid; q < MS_WAVES * 256; q += MS_TW) s_cnt[q] = 0;
      __syncthreads();
      const int e = base + tid;
      const bool valid = e < c1;
      const uint64_t key = valid ? kin[e] : 0ull;
      const uint32_t val = valid ? vin[e] : 0u;
      const uint32_t d = (uint32_t)(key >> shift) & 255u;
      unsigned long long same = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const unsigned long long bb = __ballot((d >> b) & 1u);
        same &= ((d >> b) & 1u) ? bb : ~bb;
      }
      const uint32_t wrank = (uint32_t)__popcll(same & lt_mask);
      if (valid && wrank == 0) s_cnt[wave * 256 + d] = (uint32_t)__popcll(same);
      __syncthreads();
      // per digit: exclusive prefix over the waves, and the sub-tile total
      if (tid < 256) {
        uint32_t run = 0;
        for (int w = 0; w < MS_WAVES; ++w) {
          const uint32_t c = s_cnt[w * 256 + tid];
          s_cnt[w * 256 + tid] = run;
          run += c;
        }
        s_tot[tid] = run;
      }
      __syncthreads();
      if (valid) {
        const uint32_t pos = s_off[d] + s_cnt[wave * 256 + d] + wrank;
        kout[pos] = key;
        vout[pos] = val;
      }
      __syncthreads();
      if (tid < 256) s_off[tid] += s_tot[tid];
      __syncthreads();
    }
    if (!grid_barrier(bar, gen, status, &s_ok)) return;
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
                         const uint64_t* knew, int key_bits) {
  if (s.n <= 0) return hipSuccess;
  const int tiles = (int)resort_tiles(s.n);
  const int passes = (key_bits + 7) / 8;
  if (passes < 1 || passes > kMaxPasses) return hipErrorInvalidValue;
  k_mover_flags<<<tiles, TW, 0, L.stream>>>(s.n, knew, s.keys_sorted, s.cells, s.blocks,
                                            s.block_bits, (uint64_t)s.n_keys, r.flags,
                                            r.tile_base, r.bar);
  k_tile_scan<<<1, SCAN_TW, 0, L.stream>>>(tiles, r.tile_base, r.n_movers);
  k_mover_compact<<<tiles, TW, 0, L.stream>>>(s.n, knew, r.flags, r.tile_base, r.mkey, r.midx);
  k_mover_sort<<<kMoverGrid, MS_TW, 0, L.stream>>>(passes, r.n_movers, r.mkey, r.midx, r.ktmp,
                                                   r.vtmp, r.skey, r.sidx, r.counts, r.bar,
                                                   r.status);
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
