// Stable LSD radix sort of (u64 key, u32 value) pairs for gfx950: one digit
// histogram pass over all digits, then one "onesweep" kernel per 8-bit digit.
//
// Onesweep pass (decoupled look-back): each workgroup takes the tile of its id
// (workgroups are dispatched in id order on each XCD, so every tile it waits on
// is resident or done; a pool sorting a device-side count takes tiles from an
// atomic ticket instead), ranks its 2048 keys stably in LDS with wave64
// ballots (rank = items of the same digit in earlier waves / slots / lanes),
// publishes its per-digit count, walks back over predecessor tiles until it
// meets an inclusive prefix, publishes its own inclusive prefix, reorders the
// tile by digit in LDS and writes it out in runs of consecutive addresses.
// Only the significant key bits are sorted (ceil(key_bits / 8) passes).
#include <hip/hip_runtime.h>

#include "sph_internal.hpp"

namespace sphamd {

namespace {

constexpr int RB = 8;
constexpr int RADIX = 1 << RB;
constexpr int WG = 256;
constexpr int WAVES = WG / kWave;
constexpr int IPT = 8;
constexpr int TILE = WG * IPT;
constexpr int HIST_IPT = 16;
constexpr uint32_t FLAG_AGG = 1u << 30;
constexpr uint32_t FLAG_PREFIX = 2u << 30;
constexpr uint32_t COUNT_MASK = (1u << 30) - 1;
// Bound on look-back polls of one predecessor word (each poll sleeps ~64 clk).
// Reaching it means a predecessor never published: the pass records an error in
// the status word instead of spinning forever.
constexpr uint32_t MAX_POLLS = 1u << 18;

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of one value per thread over a 256-thread workgroup.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wsum,
                                                    uint32_t& total) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) s_wsum[w] = x;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < WAVES; ++k) {
    uint32_t s = s_wsum[k];
    off += (k < w) ? s : 0u;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return off + x - v;
}

// Digit histogram of every pass at once; also clears the look-back words of the
// passes about to run, the tile tickets, and the cell ranges of the previous
// sort (read from its sorted keys before the last pass overwrites them). The
// global histogram must be zero on entry (the permutation kernel clears it).
__global__ __launch_bounds__(WG) void k_hist(const uint64_t* __restrict__ keys, int n,
                                             const int* __restrict__ n_dev, int passes,
                                             uint32_t* __restrict__ hist,
                                             uint32_t* __restrict__ lookback,
                                             uint32_t lookback_stride,
                                             uint32_t* __restrict__ tickets,
                                             const uint64_t* __restrict__ old_sorted,
                                             int old_n, int2* __restrict__ cells,
                                             uint64_t n_keys, int2* __restrict__ blocks,
                                             int block_bits) {
  __shared__ uint32_t s_h[kMaxPasses * RADIX];
  if (n_dev) n = *n_dev;  // count known on the device only
  for (int i = threadIdx.x; i < passes * RADIX; i += WG) s_h[i] = 0;
  __syncthreads();
  const int gtid = blockIdx.x * WG + threadIdx.x, gsize = gridDim.x * WG;
  for (int chunk = gtid; (int64_t)chunk * HIST_IPT < n; chunk += gsize) {
    int base = chunk * HIST_IPT;
    int cnt = min(HIST_IPT, n - base);
    uint64_t k[HIST_IPT];
#pragma unroll
    for (int t = 0; t < HIST_IPT; ++t) k[t] = t < cnt ? keys[base + t] : 0ull;
    for (int p = 0; p < passes; ++p) {
      // run-length accumulation: neighbouring keys mostly share their high digits
      uint32_t cur = (uint32_t)(k[0] >> (p * RB)) & (RADIX - 1), run = 1;
#pragma unroll
      for (int t = 1; t < HIST_IPT; ++t) {
        if (t >= cnt) break;
        uint32_t d = (uint32_t)(k[t] >> (p * RB)) & (RADIX - 1);
        if (d == cur) {
          ++run;
        } else {
          atomicAdd(&s_h[p * RADIX + cur], run);
          cur = d;
          run = 1;
        }
      }
      atomicAdd(&s_h[p * RADIX + cur], run);
    }
  }
  for (int i = gtid; i < old_n; i += gsize) {
    uint64_t k = old_sorted[i];
    uint64_t kp = i > 0 ? old_sorted[i - 1] : ~k;
    if (k != kp && k < n_keys) cells[k] = make_int2(-1, -1);
    if (block_bits > 0 && (k >> (2 * block_bits)) != (kp >> (2 * block_bits)) && k < n_keys)
      blocks[k >> (2 * block_bits)] = make_int2(-1, -1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < passes * RADIX; i += WG)
    if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
  const uint32_t used = (uint32_t)((n + TILE - 1) / TILE) * RADIX;  // words per pass in use
  for (int p = 0; p < passes; ++p) {
    uint4* lb = reinterpret_cast<uint4*>(lookback + (size_t)p * lookback_stride);
    for (uint32_t i = gtid; i < used / 4; i += gsize) lb[i] = make_uint4(0, 0, 0, 0);
  }
  if (blockIdx.x == 0 && threadIdx.x < kMaxPasses) tickets[threadIdx.x] = 0;
}

__global__ __launch_bounds__(WG) void k_onesweep(int n, const int* __restrict__ n_dev,
                                                 int shift, int pass,
                                                 const uint64_t* __restrict__ kin,
                                                 const uint32_t* __restrict__ vin,
                                                 uint64_t* __restrict__ kout,
                                                 uint32_t* __restrict__ vout,
                                                 const uint32_t* __restrict__ hist,
                                                 uint32_t* __restrict__ lookback,
                                                 uint32_t* __restrict__ tickets,
                                                 uint32_t* __restrict__ status) {
  __shared__ uint64_t s_keys[TILE];
  __shared__ uint32_t s_vals[TILE];
  __shared__ uint32_t s_whist[WAVES * RADIX];
  __shared__ uint32_t s_dstart[RADIX];
  __shared__ uint32_t s_gbase[RADIX];
  __shared__ uint32_t s_wsum[WAVES];
  __shared__ uint32_t s_tile;

  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  if (n_dev) n = *n_dev;
  // With a device-side count the grid is a small persistent pool that keeps taking
  // tickets until every tile is done; otherwise one tile per workgroup.
  for (;;) {
  // one tile per workgroup in id order (dispatch order per XCD keeps every
  // predecessor resident); the persistent pool takes tickets instead
  if (tid == 0) s_tile = n_dev ? atomicAdd(&tickets[pass], 1u) : blockIdx.x;
  for (int i = tid; i < WAVES * RADIX; i += WG) s_whist[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  if ((int64_t)tile * TILE >= n) return;
  const int base = (int)tile * TILE + wave * (kWave * IPT);

  uint64_t key[IPT];
  uint32_t val[IPT], rank[IPT];
#pragma unroll
  for (int s = 0; s < IPT; ++s) {
    int idx = base + s * kWave + lane;
    if (idx < n) {
      key[s] = kin[idx];
      val[s] = vin ? vin[idx] : (uint32_t)idx;
    } else {
      key[s] = 0;
      val[s] = 0;
    }
  }

  // Stable in-wave ranking: lanes holding the same digit are found with RB
  // ballots; the lowest such lane bumps the wave's running count.
  const uint64_t lt_mask = (1ull << lane) - 1ull;
#pragma unroll
  for (int s = 0; s < IPT; ++s) {
    const bool valid = base + s * kWave + lane < n;
    const uint32_t d = (uint32_t)(key[s] >> shift) & (RADIX - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      uint64_t bb = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    if (valid) {
      uint32_t below = (uint32_t)__popcll(peers & lt_mask);
      uint32_t c = s_whist[wave * RADIX + d];
      rank[s] = c + below;
      if (below == 0) s_whist[wave * RADIX + d] = c + (uint32_t)__popcll(peers);
    }
  }
  __syncthreads();

  // Per digit (one per thread): tile count and exclusive offsets across waves.
  const int d = tid;
  uint32_t cnt = 0;
#pragma unroll
  for (int w = 0; w < WAVES; ++w) {
    uint32_t c = s_whist[w * RADIX + d];
    s_whist[w * RADIX + d] = cnt;
    cnt += c;
  }

  uint32_t* lb = lookback + (size_t)tile * RADIX;
  if (tile == 0) {
    uint32_t tot;
    uint32_t excl = block_excl_scan(hist[pass * RADIX + d], s_wsum, tot);
    st_agent(&lb[d], FLAG_PREFIX | (excl + cnt));
    s_gbase[d] = excl;
  } else {
    st_agent(&lb[d], FLAG_AGG | cnt);
    uint32_t excl = 0;
    int t = (int)tile - 1;
    uint32_t polls = 0;
    while (t >= 0) {
      uint32_t w = ld_agent(&lookback[(size_t)t * RADIX + d]);
      uint32_t flag = w & ~COUNT_MASK;
      if (flag == 0) {
        if (++polls < MAX_POLLS) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        atomicOr(status, 1u);  // give up on this predecessor: result is wrong, flagged
        polls = 0;
        --t;
        continue;
      }
      polls = 0;
      excl += w & COUNT_MASK;
      if (flag == FLAG_PREFIX) break;
      --t;
    }
    st_agent(&lb[d], FLAG_PREFIX | (excl + cnt));
    s_gbase[d] = excl;
  }
  uint32_t tile_total;
  s_dstart[d] = block_excl_scan(cnt, s_wsum, tile_total);
  __syncthreads();

#pragma unroll
  for (int s = 0; s < IPT; ++s) {
    if (base + s * kWave + lane < n) {
      uint32_t dd = (uint32_t)(key[s] >> shift) & (RADIX - 1);
      uint32_t pos = s_dstart[dd] + s_whist[wave * RADIX + dd] + rank[s];
      s_keys[pos] = key[s];
      s_vals[pos] = val[s];
    }
  }
  __syncthreads();

  for (int k = tid; k < (int)tile_total; k += WG) {
    uint64_t kk = s_keys[k];
    uint32_t dd = (uint32_t)(kk >> shift) & (RADIX - 1);
    uint32_t g = s_gbase[dd] + ((uint32_t)k - s_dstart[dd]);
    if (g < (uint32_t)n) {
      kout[g] = kk;
      vout[g] = s_vals[k];
    }
  }
  if (!n_dev) return;
  __syncthreads();  // LDS reused by the next tile
  }
}

}  // namespace

int sort_tile_size() { return TILE; }

int sort_passes(int key_bits) { return (key_bits + RB - 1) / RB; }

static uint32_t lookback_words_per_pass(int n) {
  return (uint32_t)((size_t)((n + TILE - 1) / TILE) * RADIX);
}

hipError_t radix_sort_hist(const SortBuffers& b, int n, int key_bits, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const int passes = sort_passes(key_bits);
  const int tiles = (n + TILE - 1) / TILE;
  if (passes > kMaxPasses || tiles > b.max_tiles) return hipErrorInvalidValue;
  const int hist_chunks = (n + HIST_IPT - 1) / HIST_IPT;
  const int hist_blocks =
      max(1, min(b.n_dev ? b.pool : b.hist_blocks, (hist_chunks + WG - 1) / WG));
  // `passes` look-back arrays laid out back to back with stride lookback_words_per_pass(n)
  k_hist<<<hist_blocks, WG, 0, stream>>>(b.keys_in, n, b.n_dev, passes, b.hist, b.lookback,
                                         lookback_words_per_pass(n), b.tickets, b.keys_out,
                                         b.clear_n, b.cells, b.n_keys, b.blocks, b.block_bits);
  return hipGetLastError();
}

hipError_t radix_sort_pass(const SortBuffers& b, int n, int key_bits, int p,
                           hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  const int passes = sort_passes(key_bits);
  const int tiles = (n + TILE - 1) / TILE;
  const uint32_t lb_words = lookback_words_per_pass(n);
  const uint64_t* kin = p == 0 ? b.keys_in : ((p & 1) ? b.keys_tmp : b.keys_in);
  uint64_t* kout = p == passes - 1 ? b.keys_out : ((p & 1) ? b.keys_in : b.keys_tmp);
  const uint32_t* vin = p == 0 ? b.vals_in : ((p & 1) ? b.vals_a : b.vals_b);
  uint32_t* vout = p == passes - 1 ? b.vals_out : ((p & 1) ? b.vals_b : b.vals_a);
  const int grid = b.n_dev ? min(tiles, b.pool) : tiles;
  k_onesweep<<<grid, WG, 0, stream>>>(n, b.n_dev, p * RB, p, kin, vin, kout, vout, b.hist,
                                       b.lookback + (size_t)p * lb_words, b.tickets,
                                       b.status);
  return hipGetLastError();
}

hipError_t radix_sort_pairs(const SortBuffers& b, int n, int key_bits, hipStream_t stream) {
  hipError_t e = radix_sort_hist(b, n, key_bits, stream);
  for (int p = 0; e == hipSuccess && p < sort_passes(key_bits); ++p)
    e = radix_sort_pass(b, n, key_bits, p, stream);
  return e;
}

}  // namespace sphamd
