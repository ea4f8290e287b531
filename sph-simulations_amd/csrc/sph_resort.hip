// Incremental re-sort of an already sorted particle set (gfx950).
//
// Between two steps a particle moves a small fraction of a cell, so almost every
// particle keeps its cell key. The particles whose key changed ("movers") are
// compacted and radix-sorted on their own; the rest ("stayers") are already in
// key order, because they are a subsequence of the previous sorted order with
// unchanged keys. The new order is the stable merge of the two sorted sequences
// by (key, previous index) -- exactly the order a stable sort of all keys gives.
//
// The movers' report (record_movers): one ballot word per 64 particles of the
// previous order, and lists of the movers' (previous index, new key) in eight
// segments, each filled in any order through one atomic counter per wave with
// movers; written by the force kernel of the step (or by k_mover_flags for other
// producers); k_merge zeroes the counters.
//
//   k_bucket  every listed mover (all of them listed, at most kMidSortCap) into
//             the bucket of its owner tile -- the merge tile of the last
//             previous-order particle ordered before it, found from the
//             previous cell range of its new cell (a search of the tile-first
//             table when that cell was empty); packed per-tile counters count
//             the owners (slot in the bucket) and the previous-index tiles
//   k_movers  bucket path: workgroup 0 scans the packed counters into the
//             256-particle tile prefixes and the window starts. Radix path
//             (lists past capacity, or a bucket past kBucketCap): workgroup 0
//             compacts the movers from the flag words in previous-index order
//             and the grid runs an LSD radix sort with bounded grid barriers,
//             then the window starts by binary search. Either way the merge
//             gets the tile prefixes and, per tile, the movers ordered before
//             its first particle.
//   k_merge   one workgroup per tile of the previous order: clears the tile's
//             old cell / block ranges, takes its window (its bucket, sorted in
//             LDS by counting, or a range of the sorted movers), places its
//             stayers (rank among the window's movers, searched in LDS) and the
//             window's movers (rank among the tile's keys, searched in LDS);
//             scatters r, v, pid, key and the permutation
//   k_ranges  cell start/end and Morton block ranges of the new order
#include <hip/hip_runtime.h>

#include "sph_internal.hpp"

namespace sphamd {

namespace {

constexpr int TW = kResortTile;  // particles per tile (four flag words)
constexpr int WIN = 512;  // sorted movers a merge tile holds in LDS at once

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// (ka, ia) < (kb, ib) lexicographically
__device__ __forceinline__ bool pair_less(uint64_t ka, uint32_t ia, uint64_t kb, uint32_t ib) {
  return ka < kb || (ka == kb && ia < ib);
}

// Loads that a lane predicate would guard are issued unconditionally at a
// clamped index inside the buffer and masked afterwards: a load under a branch
// is waited for before the branch closes, which serialises a run of them.

// Flag word w of the previous order (zero past the last particle; n >= 1).
__device__ __forceinline__ unsigned long long flag_word(const unsigned long long* flags, int w,
                                                        int n) {
  const unsigned long long v = flags[min(w, (n - 1) / kWave)];
  return (int64_t)w * kWave < n ? v : 0ull;
}

// The movers' flag words for producers of the new keys that do not write them.
constexpr int MF_TW = 1024;
__global__ __launch_bounds__(MF_TW) void k_mover_flags(int n, int words,
                                                       const uint64_t* __restrict__ knew,
                                                       const uint64_t* __restrict__ kold,
                                                       MoverSink m) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int i = blockIdx.x * MF_TW + tid;
  const int ic = min(i, n - 1);
  const uint64_t kn = knew[ic], ko = kold[ic];
  const unsigned long long bal = __ballot(i < n && kn != ko);
  if (lane == 0 && i / kWave < words) m.flags[i / kWave] = bal;
}

// ------------------------------------------------------------------ movers
constexpr int MS_TW = 1024;  // threads per k_movers workgroup
constexpr int MS_WAVES = MS_TW / kWave;
// the grid radix path reads every workgroup's 256 digit counts into its LDS key
// buffers
static_assert(kMoverGrid * 256 <= 4 * kMoverLds, "digit counts exceed the LDS arrays");
constexpr uint32_t MS_POLLS = 1u << 20;

// Grid-wide barrier on a monotonic counter (zero at kernel start: k_merge
// resets it every step).
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

// Inclusive prefix sum over the 64 lanes of a wave with DPP moves (no LDS
// round trips): shifts by 1, 2, 4 and 8 lanes inside each row of 16, then row 1
// and 3 add the last lane of the row before, then rows 2 and 3 add lane 31.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  int x = (int)v;
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return (uint32_t)x;
}

// Exclusive scan of one value per thread over the workgroup; *total gets the
// sum. scratch: MS_WAVES words. Ends with a workgroup barrier.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* scratch,
                                                    uint32_t* total) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint32_t x = wave_incl_scan(v);
  if (lane == kWave - 1) scratch[wave] = x;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < MS_WAVES; ++w) {
    const uint32_t c = scratch[w];
    off += w < wave ? c : 0u;
    tot += c;
  }
  *total = tot;
  __syncthreads();
  return off + x - v;
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

// Number of sorted movers r in [lo, hi) with (keys[r], idx[r]) < (k, i).
__device__ __forceinline__ int rank_in_sorted(const uint64_t* keys, const uint32_t* idx, int lo,
                                              int hi, uint64_t k, uint32_t i) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (pair_less(keys[mid], idx[mid], k, i)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}


// Flag words per thread and per batch of the ordered compaction.
constexpr int MS_WPT = 12;
constexpr int MS_BATCH = MS_WPT * MS_TW;  // 12288 words: 786,432 particles

// Exclusive prefix over the waves of each digit's per-wave counts (threads
// 0..255, one digit each; the 16 counts are loaded before any is replaced) and
// the digit totals into tot.
__device__ __forceinline__ void wave_digit_prefix(uint32_t* wc, uint32_t* tot) {
  const int tid = threadIdx.x;
  if (tid < 256) {
    uint32_t c[MS_WAVES];
#pragma unroll
    for (int w = 0; w < MS_WAVES; ++w) c[w] = wc[w * 256 + tid];
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < MS_WAVES; ++w) {
      wc[w * 256 + tid] = run;
      run += c[w];
    }
    tot[tid] = run;
  }
}

using lds_u32 = __attribute__((address_space(3))) uint32_t;

// Workgroup 0 of the radix path: tile prefixes and the movers' previous
// indices in order, into midx. Per batch each wave takes MS_WPT * 64
// consecutive flag words (lane l: words q * 64 + l, coalesced); the prefixes
// come from wave scans plus the waves' offsets. s_pre: MS_BATCH words; s_w:
// MS_WAVES words (both LDS).
__device__ __forceinline__ void compact_movers(int n, int tiles,
                                            const unsigned long long* __restrict__ flags,
                                            uint32_t* __restrict__ tile_base,
                                            uint32_t* __restrict__ tile_lo,
                                            int* __restrict__ n_movers,
                                            uint32_t* __restrict__ midx, lds_u32* s_pre,
                                            lds_u32* s_w) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int words = 4 * tiles;
  uint32_t done = 0;
  for (int b0 = 0; b0 < words; b0 += MS_BATCH) {
    const int wbase = b0 + wave * (MS_WPT * kWave);
    lds_u32* pre = s_pre + wave * (MS_WPT * kWave) + lane;  // pre[q * kWave]
    unsigned long long wd[MS_WPT];
#pragma unroll
    for (int q = 0; q < MS_WPT; ++q) wd[q] = flag_word(flags, wbase + q * kWave + lane, n);
    uint32_t wtot = 0;
#pragma unroll
    for (int q = 0; q < MS_WPT; ++q) {
      const uint32_t c = (uint32_t)__popcll(wd[q]);
      const uint32_t x = wave_incl_scan(c);
      pre[q * kWave] = wtot + x - c;
      wtot += (uint32_t)__builtin_amdgcn_readlane((int)x, kWave - 1);
    }
    if (lane == 0) s_w[wave] = wtot;
    __syncthreads();
    uint32_t woff = done, btot = 0;
#pragma unroll
    for (int w = 0; w < MS_WAVES; ++w) {
      const uint32_t c = s_w[w];
      woff += w < wave ? c : 0u;
      btot += c;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MS_WPT; ++q) {
      const int w = wbase + q * kWave + lane;
      uint32_t pos = woff + pre[q * kWave];
      if ((w & 3) == 0 && (w >> 2) < tiles) tile_base[w >> 2] = pos;
      unsigned long long bits = wd[q];
      while (bits) {
        midx[pos] = (uint32_t)w * kWave + (uint32_t)(__ffsll((long long)bits) - 1);
        ++pos;
        bits &= bits - 1ull;
      }
    }
    done += btot;
  }
  if (tid == 0) {
    *n_movers = (int)done;
    tile_lo[0] = 0;
    tile_lo[tiles] = done;
  }
}

// (key, previous index) order of the listed movers: one word key << 32 | index
// when the keys have at most 32 bits (one compare), else the pair.
template <bool WIDE>
struct Ord;
template <>
struct Ord<false> {
  uint64_t w;
  __device__ static Ord make(uint64_t k, uint32_t i) { return Ord{(k << 32) | i}; }
  __device__ bool operator<(const Ord& o) const { return w < o.w; }
};
template <>
struct Ord<true> {
  uint64_t k;
  uint32_t i;
  __device__ static Ord make(uint64_t k, uint32_t i) { return Ord{k, i}; }
  __device__ bool operator<(const Ord& o) const { return pair_less(k, i, o.k, o.i); }
};

// ---- bucket path: every mover goes into the bucket of the merge tile that
// owns it -- the tile of the last previous-order particle ordered before it
// by (key, previous index). When its new cell k held particles [s, e) in the
// previous order, that particle is s - 1 if its previous index i < s (its
// old key was smaller) and e - 1 otherwise (i lies outside [s, e): its old
// key differs), one load of the cell table; movers into a cell that was
// empty search the tile-first table in LDS. Each workgroup takes KB_TW flag
// words (one per thread) and lists their set bits in LDS; the owner's packed
// counter (tile_count, high half) gives the slot in the bucket, the
// previous-index tile's counter (low half) counts the movers before each
// tile. The count word (bovf) sums the movers; a bucket past kBucketCap, a
// workgroup past kBucketList movers or 2^16 movers in all send the step to
// the radix path.
constexpr int KB_TW = 256;
constexpr int kBucketList = 2048;
constexpr int kBucketTab = 1024;  // sampled tile-first keys in LDS (every S-th tile)
static_assert(kBucketTab % KB_TW == 0, "tile-first table loads");
static_assert(KB_TW % 4 == 0, "four flag words per merge tile");

template <bool WIDE>
__global__ __launch_bounds__(KB_TW) void k_bucket(
    int n, int tiles, const uint64_t* __restrict__ knew,
    const unsigned long long* __restrict__ flags, const uint64_t* __restrict__ tfirst,
    uint32_t* __restrict__ tcount, const int2* __restrict__ cells, uint64_t n_keys,
    uint4* __restrict__ bucket, uint32_t* __restrict__ bovf) {
  using O = Ord<WIDE>;
  __shared__ uint32_t s_list[kBucketList];
  __shared__ uint64_t s_tab[kBucketTab];
  __shared__ uint32_t s_w[KB_TW / kWave];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int words = (n + kWave - 1) / kWave;
  const int w = (int)blockIdx.x * KB_TW + tid;
  const unsigned long long fw = flags[min(w, words - 1)];
  // the sampled tile-first table for owner searches, loaded with the flags
  const int S = (tiles + kBucketTab - 1) / kBucketTab;
  const int ne = (tiles + S - 1) / S;
  uint64_t tv[kBucketTab / KB_TW];
#pragma unroll
  for (int q = 0; q < kBucketTab / KB_TW; ++q)
    tv[q] = tfirst[(size_t)min(q * KB_TW + tid, ne - 1) * S];
  const unsigned long long bits = w < words ? fw : 0ull;
  const uint32_t c = (uint32_t)__popcll(bits);
  // movers whose previous index lies in each merge tile (four words per tile,
  // four adjacent lanes)
  uint32_t hc = c + (uint32_t)__shfl_xor((int)c, 1);
  hc += (uint32_t)__shfl_xor((int)hc, 2);
  if ((lane & 3) == 0 && hc != 0u) atomicAdd(&tcount[w >> 2], hc);
  // the workgroup's movers listed in LDS (exclusive scan of the word counts)
  const uint32_t x = wave_incl_scan(c);
  if (lane == kWave - 1) s_w[wave] = x;
  __syncthreads();
  uint32_t off = x - c, tot = 0;
#pragma unroll
  for (int v = 0; v < KB_TW / kWave; ++v) {
    const uint32_t t = s_w[v];
    off += v < wave ? t : 0u;
    tot += t;
  }
  if (tot == 0u) return;
  if (tot > (uint32_t)kBucketList) {
    if (tid == 0) atomicOr(bovf, kBucketOverflow);
    return;
  }
  if (tid == 0) atomicAdd(bovf, tot);
  for (unsigned long long b = bits; b != 0ull; b &= b - 1ull)
    s_list[off++] = (uint32_t)w * kWave + (uint32_t)(__ffsll((long long)b) - 1);
#pragma unroll
  for (int q = 0; q < kBucketTab / KB_TW; ++q) {
    const int e2 = q * KB_TW + tid;
    if (e2 < ne) {
      if constexpr (WIDE) s_tab[e2] = tv[q];
      else s_tab[e2] = O::make(tv[q], (uint32_t)(e2 * S * TW)).w;
    }
  }
  __syncthreads();
  for (uint32_t r0 = 0; r0 < tot; r0 += KB_TW) {
    const uint32_t r = r0 + (uint32_t)tid;
    if (r >= tot) break;
    const uint32_t mi = s_list[r];
    const uint64_t mk = knew[mi];
    const int2 cr = cells[min(mk, n_keys - 1)];
    const bool need = cr.x < 0;
    const O mo = O::make(mk, mi);
    uint32_t own;
    if (!need) {
      const int p = (int)mi < cr.x ? cr.x - 1 : cr.y - 1;
      own = p < 0 ? 0u : (uint32_t)p / (uint32_t)TW;
    } else {
      auto tab = [&](int e2) {
        if constexpr (WIDE) return O::make(s_tab[e2], (uint32_t)(e2 * S * TW));
        else return O{s_tab[e2]};
      };
      // the last sampled entry before the mover, then the tiles between two samples
      int base = 0;
      for (int len = ne; len > 1;) {
        const int half = len >> 1;
        if (tab(base + half - 1) < mo) base += half;
        len -= half;
      }
      if (tab(base) < mo) ++base;
      const int cc = base - 1;
      int o = max(cc, 0) * S;
      if (cc >= 0)
        for (int k = 1; k < S; ++k) {
          const int t = cc * S + k;
          if (t < tiles && O::make(tfirst[t], (uint32_t)t * TW) < mo) o = t;
        }
      own = (uint32_t)o;
    }
    const uint32_t slot = atomicAdd(&tcount[own], 1u << 16) >> 16;
    if (slot < (uint32_t)kBucketCap)
      bucket[(size_t)own * kBucketCap + slot] =
          make_uint4((uint32_t)mk, (uint32_t)(mk >> 32), mi, 0u);
    else
      atomicOr(bovf, kBucketOverflow);
  }
}

// ---- tile prefixes and, when the buckets cannot be used, the movers sorted.
// Bucket path (listed movers in their buckets, none past its capacity):
// workgroup 0 scans the packed tile counters into the tile prefixes (movers
// before each tile's first particle by previous index) and the window starts
// (movers ordered before each tile's first particle); the merge sorts each
// window from its bucket. Otherwise (lists past their capacity or a bucket
// overflowed) workgroup 0 compacts the movers from the flag words in
// previous-index order and the grid runs an LSD radix sort of all of them,
// and the merge reads its window from the sorted movers. *mode tells the
// merge which: 1 buckets, 0 sorted.
template <bool WIDE>
__global__ __launch_bounds__(MS_TW) void k_movers(
    int n, int tiles, int passes, const uint64_t* __restrict__ knew,
    const uint64_t* __restrict__ tfirst, const uint32_t* __restrict__ tcount,
    const unsigned long long* __restrict__ flags, const uint32_t* __restrict__ bovf,
    uint32_t* __restrict__ mode,
    uint32_t* __restrict__ tile_base, uint32_t* __restrict__ tile_lo, int* __restrict__ n_movers,
    uint64_t* __restrict__ mkey, uint32_t* __restrict__ midx, uint64_t* __restrict__ ktmp,
    uint32_t* __restrict__ vtmp, uint64_t* __restrict__ skey, uint32_t* __restrict__ sidx,
    uint32_t* __restrict__ counts, uint32_t* __restrict__ bar, uint32_t* __restrict__ status) {
  // key buffers: the word prefixes of the compaction and the digit counts of
  // the grid sort
  __shared__ uint64_t s_k[2 * kMoverLds];
  // digit counts: u32 [MS_WAVES][256] per wave in the grid sort
  __shared__ uint32_t s_wc[2 * MS_WAVES * 256];
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_tot[256];
  __shared__ uint32_t s_w[MS_WAVES];
  __shared__ int s_ok;
  const int tid = threadIdx.x, wave = tid / kWave;
  const uint32_t bucketed_movers = *bovf;
  // workgroup 0 loads the first tile counters with them (used on the bucket path)
  uint32_t c0v[4] = {0u, 0u, 0u, 0u};
  if (blockIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) c0v[k] = tcount[min(4 * tid + k, tiles - 1)];
  }
  // no overflow and fewer than 2^16 movers (the packed counters' prefixes fit)
  const bool bucketed = bucketed_movers < (1u << 16);
  if (bucketed) {
    if (blockIdx.x != 0) return;
    // exclusive scan of the packed counters (previous-index tile low half,
    // owner tile high half; both prefixes stay below 2^16), four tiles per
    // thread per round
    uint32_t carry = 0;
    for (int t0 = 0; t0 < tiles; t0 += 4 * MS_TW) {
      uint32_t c[4], sum = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int t = t0 + 4 * tid + k;
        const uint32_t v = t0 == 0 ? c0v[k] : tcount[min(t, tiles - 1)];
        c[k] = t < tiles ? v : 0u;
        sum += c[k];
      }
      uint32_t rtot;
      uint32_t run = carry + block_excl_scan(sum, s_w, &rtot);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int t = t0 + 4 * tid + k;
        if (t < tiles) {
          tile_base[t] = run & 0xffffu;
          tile_lo[t] = run >> 16;
        }
        run += c[k];
      }
      carry += rtot;
    }
    if (tid == 0) {
      *n_movers = (int)bucketed_movers;
      tile_lo[tiles] = bucketed_movers;
      *mode = 1u;
    }
    return;
  }
  if (blockIdx.x == 0 && tid == 0) *mode = 0u;
  uint32_t gen = 0;

  // ---- more movers than one workgroup groups: workgroup 0 writes the tile
  // prefixes and compacts the movers in previous-index order into midx
  static_assert(MS_BATCH * 4 <= 2 * kMoverLds * 8, "word prefixes exceed the key buffers");
  if (blockIdx.x == 0)
    compact_movers(n, tiles, flags, tile_base, tile_lo, n_movers, midx,
                   (lds_u32*)(s_k), (lds_u32*)(s_w));

  // ---- more movers than one workgroup sorts: LSD radix sort over the grid
  // (the indices in midx from workgroup 0; the first pass reads the keys
  // through them)
  if (!grid_barrier(bar, gen, status, &s_ok)) return;
  // the exact number of movers, counted from the flag words by workgroup 0 (the
  // list counters stop short of it when the producer listed nothing)
  const uint32_t total = ld_agent(reinterpret_cast<const uint32_t*>(n_movers));
  uint32_t* s_cnt = s_wc;    // [MS_WAVES][256] per-wave digit counts
  uint32_t* s_off = s_base;  // [256] running digit offsets
  const int nb = gridDim.x;
  const int chunk = (int)((total + (uint32_t)(nb * MS_TW) - 1) / (uint32_t)(nb * MS_TW)) * MS_TW;
  const int c0 = min((int)total, (int)blockIdx.x * chunk);
  const int c1 = min((int)total, c0 + chunk);
  const int lane = tid & (kWave - 1);
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  for (int p = 0; p < passes; ++p) {
    // pass 0 reads the keys through midx; then ktmp / mkey alternate
    const uint64_t* kin = (p & 1) ? ktmp : mkey;
    const uint32_t* vin = (p == 0) ? midx : ((p & 1) ? vtmp : midx);
    uint64_t* kout = p == passes - 1 ? skey : ((p & 1) ? mkey : ktmp);
    uint32_t* vout = p == passes - 1 ? sidx : ((p & 1) ? midx : vtmp);
    auto key_of = [&](int e) { return p == 0 ? knew[vin[e]] : kin[e]; };
    const int shift = 8 * p;
    // A: digit counts of this workgroup's chunk
    if (tid < 256) s_off[tid] = 0;
    __syncthreads();
    for (int e = c0 + tid; e < c1; e += MS_TW)
      atomicAdd(&s_off[(uint32_t)(key_of(e) >> shift) & 255u], 1u);
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
      const uint64_t key = valid ? key_of(e) : 0ull;
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
      wave_digit_prefix(s_cnt, s_tot);
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
  // window starts of the merge tiles (binary searches in the sorted movers)
  for (int t = 1 + (int)blockIdx.x * MS_TW + tid; t < tiles; t += nb * MS_TW)
    tile_lo[t] = (uint32_t)rank_in_sorted(skey, sidx, 0, (int)total, tfirst[t],
                                          (uint32_t)t * TW);
}

// ------------------------------------------------------------------- merge
// One workgroup per tile of the previous order [j0, j0 + TW). Its window of
// sorted movers [lo, hi) holds the movers ordered after the tile's first
// particle and before the next tile's first particle (tile 0: from the first
// mover), so every mover belongs to exactly one window.
//   stayer j:  slot = (j - movers before j) + lo + (window movers before it)
//   mover r:   slot = r + (stayers before it), where the previous-order
//              particles before it are j0 + (tile particles before it)
__global__ __launch_bounds__(TW) void k_merge(
    int n, const uint64_t* __restrict__ kold, const unsigned long long* __restrict__ flags,
    const uint32_t* __restrict__ tile_base, const uint32_t* __restrict__ tile_lo,
    const uint64_t* __restrict__ skey, const uint32_t* __restrict__ sidx,
    const float* __restrict__ tx, const float* __restrict__ ty, const float* __restrict__ tvx,
    const float* __restrict__ tvy, const uint32_t* __restrict__ pid_src, float* __restrict__ rx,
    float* __restrict__ ry, float* __restrict__ vx, float* __restrict__ vy,
    uint32_t* __restrict__ pid_dst, uint64_t* __restrict__ keys_dst, uint32_t* __restrict__ perm,
    int2* __restrict__ cells, int2* __restrict__ blocks, int block_bits, uint64_t n_keys,
    uint32_t* __restrict__ bar, float4* __restrict__ xyuv,
    const uint32_t* __restrict__ status, const uint32_t* __restrict__ mode,
    const uint4* __restrict__ bucket, uint32_t* __restrict__ bovf) {
  // A mover kernel whose grid barrier gave up left its windows partly written:
  // merge nothing (the host reports the status word at the next check).
  if (*status != 0u) return;
  __shared__ uint64_t s_tk[TW];  // the tile's previous keys (all ones past n)
  __shared__ uint64_t s_mk[WIN];
  __shared__ uint32_t s_mi[WIN];
  __shared__ uint32_t s_wpre[TW / kWave + 1];  // movers of the tile before each wave
  __shared__ unsigned long long s_fw[TW / kWave];  // the tile's flag words
  __shared__ uint64_t s_bk_k[kBucketCap];          // the tile's bucket, arrival order
  __shared__ uint32_t s_bk_i[kBucketCap];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int tile = blockIdx.x;
  const int j0 = tile * TW, j = j0 + tid;
  const bool live = j < n;
  // every load that depends only on j, issued together
  const unsigned long long fw = flag_word(flags, j0 / kWave + wave, n);
  const uint32_t tb = tile_base[tile];
  // windows lie in [0, n) (clamped, so that a bad table cannot read past skey)
  const int lo = min((int)tile_lo[tile], n), hi = max(lo, min((int)tile_lo[tile + 1], n));
  const int jc = min(j, n - 1);
  const uint64_t kc = kold[jc], kpc = kold[max(jc - 1, 0)];
  const uint64_t k = live ? kc : ~0ull;
  const uint64_t kp = (live && j > 0) ? kpc : ~k;
  const float x = tx[jc], y = ty[jc], u = tvx[jc], v = tvy[jc];
  const uint32_t id = pid_src[jc];
  if (tile == 0 && tid == 0) *bar = 0;  // the next k_movers' grid barrier
  // the next step's bucket count word (k_bucket and k_movers have ended)
  if (tile == 0 && tid == 0) *bovf = 0u;
  const bool from_bucket = *mode != 0u;
  // the previous order's ranges (k_ranges writes the new ones)
  if (live && k != kp && k < n_keys) {
    cells[k] = make_int2(-1, -1);
    if (block_bits > 0 && (k >> (2 * block_bits)) != (kp >> (2 * block_bits)))
      blocks[k >> (2 * block_bits)] = make_int2(-1, -1);
  }
  s_tk[tid] = k;
  if (lane == 0) {
    s_wpre[wave + 1] = (uint32_t)__popcll(fw);
    s_fw[wave] = fw;
  }
  __syncthreads();
  if (tid == 0) {
    s_wpre[0] = 0;
    for (int w = 1; w <= TW / kWave; ++w) s_wpre[w] += s_wpre[w - 1];
  }
  __syncthreads();
  const bool mv = live && ((fw >> lane) & 1ull);
  int before_movers = 0;  // window movers ordered before stayer j
  for (int c0 = lo; c0 < hi; c0 += WIN) {
    const int len = min(WIN, hi - c0);
    if (from_bucket) {
      // the window is the tile's bucket (at most kBucketCap movers, one
      // chunk): sorted here by (key, previous index), rank by counting
      const int blen = min(len, kBucketCap);
      for (int r = tid; r < blen; r += TW) {
        const uint4 b = bucket[(size_t)tile * kBucketCap + r];
        s_bk_k[r] = (uint64_t)b.x | ((uint64_t)b.y << 32);
        s_bk_i[r] = b.z;
      }
      __syncthreads();
      for (int r = tid; r < blen; r += TW) {
        const uint64_t km = s_bk_k[r];
        const uint32_t im = s_bk_i[r];
        int rank = 0;
        for (int u = 0; u < blen; ++u) rank += pair_less(s_bk_k[u], s_bk_i[u], km, im) ? 1 : 0;
        s_mk[rank] = km;
        s_mi[rank] = im;
      }
    } else {
      for (int r = tid; r < len; r += TW) {
        s_mk[r] = skey[c0 + r];
        s_mi[r] = sidx[c0 + r];
      }
    }
    __syncthreads();
    if (live && !mv) {
      int a = 0, z = len;
      while (a < z) {
        const int mid = (a + z) >> 1;
        if (pair_less(s_mk[mid], s_mi[mid], k, (uint32_t)j)) a = mid + 1;
        else z = mid;
      }
      before_movers += a;
    }
    for (int r = tid; r < len; r += TW) {
      const uint64_t km = s_mk[r];
      const uint32_t src = s_mi[r];
      // the mover's state, in flight during the search
      const uint32_t sc = min(src, (uint32_t)(n - 1));
      const float mx = tx[sc], my = ty[sc], mu = tvx[sc], mv = tvy[sc];
      const uint32_t mid_ = pid_src[sc];
      // tile particles ordered before the mover (keys past n compare greater)
      int a = 0, z = TW;
      while (a < z) {
        const int mid = (a + z) >> 1;
        if (pair_less(s_tk[mid], (uint32_t)(j0 + mid), km, src)) a = mid + 1;
        else z = mid;
      }
      // movers among the tile particles [j0, j0 + a)
      const int w = a / kWave, l = a % kWave;
      uint32_t mb = tb + s_wpre[w];
      if (w < TW / kWave) mb += (uint32_t)__popcll(s_fw[w] & ((1ull << l) - 1ull));
      const int pos = (c0 + r) + (j0 + a - (int)mb);
      if ((unsigned)pos >= (unsigned)n) continue;  // only a corrupt table gets here
      if (rx) {
        rx[pos] = mx;
        ry[pos] = my;
        vx[pos] = mu;
        vy[pos] = mv;
      }
      xyuv[pos] = make_float4(mx, my, mu, mv);
      pid_dst[pos] = mid_;
      keys_dst[pos] = km;
      if (perm) perm[pos] = src;
    }
    __syncthreads();
  }
  if (live && !mv) {
    const uint32_t mb = tb + s_wpre[wave] + (uint32_t)__popcll(fw & ((1ull << lane) - 1ull));
    const int pos = (j - (int)mb) + lo + before_movers;
    if ((unsigned)pos >= (unsigned)n) return;  // only a corrupt table gets here
    if (rx) {
      rx[pos] = x;
      ry[pos] = y;
      vx[pos] = u;
      vy[pos] = v;
    }
    xyuv[pos] = make_float4(x, y, u, v);
    pid_dst[pos] = id;
    keys_dst[pos] = k;
    if (perm) perm[pos] = (uint32_t)j;
  }
}

}  // namespace

int64_t resort_flag_words(int n) { return ((int64_t)n + TW - 1) / TW * (TW / kWave); }
int64_t resort_tiles(int n) { return ((int64_t)n + TW - 1) / TW; }

hipError_t launch_movers(const Launch& L, const DevState& s, const ResortBuffers& r,
                         const uint64_t* knew, int key_bits, bool flags_ready) {
  if (s.n <= 0) return hipSuccess;
  const int tiles = (int)resort_tiles(s.n);
  const int passes = (key_bits + 7) / 8;
  if (passes < 1 || passes > kMaxPasses) return hipErrorInvalidValue;
  if (!flags_ready)
    k_mover_flags<<<(s.n + MF_TW - 1) / MF_TW, MF_TW, 0, L.stream>>>(
        s.n, (int)resort_flag_words(s.n), knew, s.keys_sorted, mover_sink(r));
  // keys of at most 32 bits: (key, index) in one word
  auto bk = passes <= 4 ? k_bucket<false> : k_bucket<true>;
  const int words = (int)((s.n + kWave - 1) / kWave);
  bk<<<(words + KB_TW - 1) / KB_TW, KB_TW, 0, L.stream>>>(
      s.n, tiles, knew, r.flags, s.tile_first, s.tile_count, s.cells, (uint64_t)s.n_keys,
      r.bucket, r.bovf);
  auto go = passes <= 4 ? k_movers<false> : k_movers<true>;
  go<<<kMoverGrid, MS_TW, 0, L.stream>>>(s.n, tiles, passes, knew, s.tile_first, s.tile_count,
                                         r.flags, r.bovf, r.mode, r.tile_base,
                                         r.tile_lo, r.n_movers, r.mkey, r.midx, r.ktmp, r.vtmp,
                                         r.skey, r.sidx, r.counts, r.bar, r.status);
  return hipGetLastError();
}

hipError_t launch_merge(const Launch& L, const DevState& s, const ResortBuffers& r,
                        uint64_t* keys_dst, bool soa, bool perm) {
  if (s.n <= 0) return hipSuccess;
  const int tiles = (int)resort_tiles(s.n);
  k_merge<<<tiles, TW, 0, L.stream>>>(s.n, s.keys_sorted, r.flags, r.tile_base, r.tile_lo,
                                      r.skey, r.sidx, s.tx, s.ty, s.tvx, s.tvy, s.pid,
                                      soa ? s.rx : nullptr, s.ry, s.vx, s.vy, s.pid_alt,
                                      keys_dst, perm ? s.perm : nullptr, s.cells, s.blocks,
                                      s.block_bits, (uint64_t)s.n_keys, r.bar, s.xyuv,
                                      r.status, r.mode, r.bucket, r.bovf);
  return hipGetLastError();
}

}  // namespace sphamd
