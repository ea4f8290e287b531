// Incremental re-sort of an already sorted particle set (gfx950).
//
// Between two steps a particle moves a small fraction of a cell, so almost every
// particle keeps its cell key. The particles whose key changed ("movers") are
// ordered on their own; the rest ("stayers") are already in key order, because
// they are a subsequence of the previous sorted order with unchanged keys. The
// new order is the stable merge of the two sequences by (key, previous index)
// -- exactly the order a stable sort of all keys gives.
//
// The movers' report (record_movers): one ballot word per 64 particles of the
// previous order, bit set when the particle's new key differs from its old one;
// written by the force kernel of the step (or by k_mover_flags for other
// producers).
//
//   k_bucket  one flag word per thread: each workgroup lists its movers in LDS
//             and puts every mover, with its state, into the bucket of its
//             owner tile -- the merge tile of the last previous-order particle
//             ordered before it, found from the previous cell range of its new
//             cell (a search of the sampled tile-first table when that cell was
//             empty). Packed 64-bit counters per tile and per chunk of
//             kChunkTiles tiles count the movers by previous index (low half)
//             and by owner (high half). When a bucket overflows, the workgroup
//             whose mover first finds it full (and first claims the step)
//             orders every mover itself (ordered compaction, then a stable
//             radix sort) for the sorted path.
//   k_chunks  in place of k_bucket when the force kernel bucketed the movers
//             itself (bucket_movers, adding to the tile counters only): the
//             chunk counters as sums of the tile counters, and the sorted path
//             when a bucket overflowed
//   k_merge   one workgroup per MT tiles of the previous order (four by
//             default; a thread per particle of each): clears the old cell /
//             block ranges its movers left; bucket path: its prefixes from the
//             packed counters, its window (the movers it owns) from its bucket,
//             sorted in LDS by counting; sorted path: the window from the sorted
//             movers. It places its stayers (rank among the window's movers,
//             searched in LDS) and the window's movers (rank among the tile's
//             keys, searched in LDS); scatters r, v, pid, key and the
//             permutation
//   k_ranges  cell start/end and Morton block ranges of the new order, the
//             tile-first keys; zeroes the counters and resets the step's
//             mover count, path and claim word for the next step (keeping the
//             count and path for the host)
#include <hip/hip_runtime.h>

#include "sph_internal.hpp"

namespace sphamd {

namespace {

constexpr int TW = kResortTile;  // particles per tile (four flag words)
constexpr int WIN = 512;  // sorted movers a merge tile holds in LDS at once (sorted path)
constexpr unsigned long long kOwn = 1ull << 32;  // one owned mover in a packed counter
static_assert((kBucketFast & (kBucketFast - 1)) == 0 && kBucketFast <= kWave &&
                  kBucketCap % kBucketFast == 0 && kBucketCap <= TW,
              "bucket loads: a power of two of at most one wave first, the rest by thread");

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

// Sum over the 64 lanes of a wave (every lane gets it).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(v), kWave - 1);
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

// (key, previous index) order of the movers: one word key << 32 | index
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

// ------------------------------------------------ sorted path (one workgroup)
constexpr int KB_TW = 256;  // threads of a k_bucket workgroup
constexpr int KB_WAVES = KB_TW / kWave;
constexpr int FB_WPT = 12;  // flag words per thread and batch of the ordered compaction
static_assert(KB_TW == 256, "the digit scans take one digit per thread");

// Movers in previous-index order into midx, and the movers before each tile
// (tile_base); returns their number. Per batch each wave takes WPT * 64
// consecutive flag words (lane l: words q * 64 + l, coalesced); the prefixes
// come from wave scans plus the waves' offsets. s_pre: WPT * KB_TW words,
// s_w: KB_WAVES words (LDS). (The merge's copy takes fewer words per batch to
// keep its register count.)
template <int WPT = FB_WPT>
__device__ __forceinline__ uint32_t compact_movers(int n, int tiles,
                                                   const unsigned long long* __restrict__ flags,
                                                   uint32_t* __restrict__ tile_base,
                                                   uint32_t* __restrict__ midx, uint32_t* s_pre,
                                                   uint32_t* s_w) {
  constexpr int FB_WPT = WPT;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int words = 4 * tiles;
  uint32_t done = 0;
  for (int b0 = 0; b0 < words; b0 += FB_WPT * KB_TW) {
    const int wbase = b0 + wave * (FB_WPT * kWave);
    uint32_t* pre = s_pre + wave * (FB_WPT * kWave) + lane;  // pre[q * kWave]
    unsigned long long wd[FB_WPT];
#pragma unroll
    for (int q = 0; q < FB_WPT; ++q) wd[q] = flag_word(flags, wbase + q * kWave + lane, n);
    uint32_t wtot = 0;
#pragma unroll
    for (int q = 0; q < FB_WPT; ++q) {
      const uint32_t c = (uint32_t)__popcll(wd[q]);
      const uint32_t x = wave_incl_scan(c);
      pre[q * kWave] = wtot + x - c;
      wtot += (uint32_t)__builtin_amdgcn_readlane((int)x, kWave - 1);
    }
    if (lane == 0) s_w[wave] = wtot;
    __syncthreads();
    uint32_t woff = done, btot = 0;
#pragma unroll
    for (int w = 0; w < KB_WAVES; ++w) {
      const uint32_t c = s_w[w];
      woff += w < wave ? c : 0u;
      btot += c;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < FB_WPT; ++q) {
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
  return done;
}

// Stable LSD radix sort (8-bit digits, `passes` passes) of the m movers whose
// previous indices midx holds in previous-index order, by their new keys (read
// through midx in the first pass); result in skey / sidx. One workgroup; s:
// KB_WAVES * 256 + 512 words of LDS.
template <typename K, int HB = 8>  // HB: keys per thread in flight while counting
__device__ __forceinline__ void sort_movers(int m, int passes, const K* __restrict__ knew,
                                            const ResortBuffers& r, uint32_t* s) {
  uint32_t* s_cnt = s;                    // [KB_WAVES][256] per-wave digit counts
  uint32_t* s_off = s + KB_WAVES * 256;  // [256] running digit offsets
  uint32_t* s_tot = s_off + 256;          // [256] digit totals of a chunk
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  for (int p = 0; p < passes; ++p) {
    const uint64_t* kin = (p & 1) ? r.ktmp : r.mkey;
    const uint32_t* vin = p == 0 ? r.midx : ((p & 1) ? r.vtmp : r.midx);
    uint64_t* kout = p == passes - 1 ? r.skey : ((p & 1) ? r.mkey : r.ktmp);
    uint32_t* vout = p == passes - 1 ? r.sidx : ((p & 1) ? r.midx : r.vtmp);
    auto key_of = [&](int e) { return p == 0 ? (uint64_t)knew[vin[e]] : kin[e]; };
    const int shift = 8 * p;
    s_off[tid] = 0u;
    __syncthreads();
    for (int e0 = 0; e0 < m; e0 += HB * KB_TW) {
      uint64_t kk[HB];
#pragma unroll
      for (int q = 0; q < HB; ++q) kk[q] = key_of(min(e0 + q * KB_TW + tid, m - 1));
#pragma unroll
      for (int q = 0; q < HB; ++q)
        if (e0 + q * KB_TW + tid < m) atomicAdd(&s_off[(uint32_t)(kk[q] >> shift) & 255u], 1u);
    }
    __syncthreads();
    digit_excl_scan(s_off, s_tot);
    // stable scatter, KB_TW keys at a time: rank inside the wave by the digit's
    // ballot mask, the waves before by their digit counts
    for (int base = 0; base < m; base += KB_TW) {
#pragma unroll
      for (int q = 0; q < KB_WAVES; ++q) s_cnt[q * KB_TW + tid] = 0u;
      const int e = base + tid;
      const bool valid = e < m;
      const int ec = min(e, m - 1);
      const uint64_t key = key_of(ec);
      const uint32_t val = vin[ec];
      __syncthreads();
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
      {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < KB_WAVES; ++w) {
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
      s_off[tid] += s_tot[tid];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ buckets
// Every mover goes into the bucket of the merge tile that owns it -- the tile
// of the last previous-order particle ordered before it by (key, previous
// index). When its new cell k held particles [s, e) in the previous order,
// that particle is s - 1 if its previous index i < s (its old key was smaller)
// and e - 1 otherwise (i lies outside [s, e): its old key differs), one load
// of the cell table; movers into a cell that was empty search the tile-first
// table in LDS. Each workgroup takes KB_TW flag words (one per thread, the
// kChunkTiles tiles of one counter chunk) and lists their set bits in LDS,
// kBucketList at a time; the owner's packed counter (high half) gives the
// slot in the bucket, the previous-index tile's counter (low half) counts the
// movers before each tile, and the chunk counters (tcount[tiles + c]) sum both
// over their tiles, so the merge finds its prefixes from at most kChunkTiles +
// chunks words. A bucket entry carries the mover's key, previous index and
// state (pid, x, y, vx, vy from T), so the merge reads it in one round. Each
// workgroup adds its movers to the step's count without waiting for the
// result. The path stays the buckets (*r.mode = 1, set by k_ranges) unless a
// bucket overflows: the mover that takes the first slot past a bucket's end
// claims the sort (one compare-and-swap on the claim word, so one workgroup
// of the step wins), and its workgroup then orders every mover itself --
// compaction and sort read only the flags and keys of the previous kernel --
// and sets the sorted path. No workgroup reads what another one wrote, so
// none needs a device-scope fence, and none waits for an atomic's result
// unless it overflowed a bucket.
constexpr int kBucketList = 2048;
constexpr int kBucketTab = 1024;  // sampled tile-first keys in LDS (every S-th tile)
static_assert(kBucketTab % KB_TW == 0, "tile-first table loads");
static_assert(KB_TW == 4 * kChunkTiles, "one k_bucket workgroup per counter chunk");
// LDS words of the list and the table, reused by the sorted path
constexpr int KB_LDS = kBucketList + 2 * kBucketTab;
static_assert(FB_WPT * KB_TW <= KB_LDS && KB_WAVES * 256 + 512 <= KB_LDS,
              "the sorted path's LDS exceeds the list and table");

template <bool WIDE, typename K>
__global__ __launch_bounds__(KB_TW) void k_bucket(
    int n, int tiles, int passes, const K* __restrict__ knew,
    const unsigned long long* __restrict__ flags, const uint64_t* __restrict__ tfirst,
    unsigned long long* __restrict__ tcount, const int2* __restrict__ cells, uint64_t n_keys,
    const float* __restrict__ tx, const float* __restrict__ ty, const float* __restrict__ tvx,
    const float* __restrict__ tvy, const uint32_t* __restrict__ pid_src,
    const float4* __restrict__ t4, ResortBuffers r) {
  using O = Ord<WIDE>;
  constexpr bool NARROW = sizeof(K) == 4;  // a narrow step: T packed in t4
  __shared__ uint64_t s_mem64[KB_LDS / 2];
  __shared__ uint32_t s_w[KB_WAVES];
  __shared__ uint32_t s_cc[3];  // owners in the chunks before, of and after this one
  __shared__ uint32_t s_sort;  // this workgroup won the claim: it orders the movers
  uint32_t* s_mem = reinterpret_cast<uint32_t*>(s_mem64);
  uint32_t* s_list = s_mem;
  uint64_t* s_tab = s_mem64 + kBucketList / 2;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int words = (n + kWave - 1) / kWave;
  const int w = (int)blockIdx.x * KB_TW + tid;
  const int ch = (int)blockIdx.x;
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
  if ((lane & 3) == 0 && hc != 0u) atomicAdd(&tcount[w >> 2], (unsigned long long)hc);
  // the workgroup's movers listed in LDS (exclusive scan of the word counts)
  const uint32_t x = wave_incl_scan(c);
  if (lane == kWave - 1) s_w[wave] = x;
  if (tid < 3) s_cc[tid] = 0u;
  if (tid == 0) s_sort = 0u;
  __syncthreads();
  uint32_t off = x - c, tot = 0;
#pragma unroll
  for (int v = 0; v < KB_WAVES; ++v) {
    const uint32_t t = s_w[v];
    off += v < wave ? t : 0u;
    tot += t;
  }
  if (tot != 0u) {  // uniform
#pragma unroll
    for (int q = 0; q < kBucketTab / KB_TW; ++q) {
      const int e2 = q * KB_TW + tid;
      if (e2 < ne) {
        if constexpr (WIDE) s_tab[e2] = tv[q];
        else s_tab[e2] = O::make(tv[q], (uint32_t)(e2 * S * TW)).w;
      }
    }
    auto tab = [&](int e2) {
      if constexpr (WIDE) return O::make(s_tab[e2], (uint32_t)(e2 * S * TW));
      else return O{s_tab[e2]};
    };
    for (uint32_t base = 0; base < tot; base += kBucketList) {
      // this round's movers: workgroup ranks [base, base + kBucketList)
      uint32_t o = off;
      for (unsigned long long b = bits; b != 0ull; b &= b - 1ull, ++o)
        if (o >= base && o < base + kBucketList)
          s_list[o - base] = (uint32_t)w * kWave + (uint32_t)(__ffsll((long long)b) - 1);
      __syncthreads();
      const uint32_t cnt = min(tot - base, (uint32_t)kBucketList);
      for (uint32_t r0 = 0; r0 < cnt; r0 += KB_TW) {
        const uint32_t rr = r0 + (uint32_t)tid;
        if (rr >= cnt) break;
        const uint32_t mi = s_list[rr];
        const uint64_t mk = (uint64_t)knew[mi];
        float mx, my, mu, mv;
        if constexpr (NARROW) {
          const float4 tq = t4[mi];
          mx = tq.x;
          my = tq.y;
          mu = tq.z;
          mv = tq.w;
        } else {
          mx = tx[mi];
          my = ty[mi];
          mu = tvx[mi];
          mv = tvy[mi];
        }
        const uint32_t mp = pid_src[mi];
        const int2 cr = cells[min(mk, n_keys - 1)];
        const O mo = O::make(mk, mi);
        uint32_t own;
        if (cr.x >= 0) {
          const int p = (int)mi < cr.x ? cr.x - 1 : cr.y - 1;
          own = p < 0 ? 0u : (uint32_t)p / (uint32_t)TW;
        } else {
          // the last sampled entry before the mover, then the tiles between
          // two samples
          int sb = 0;
          for (int len = ne; len > 1;) {
            const int half = len >> 1;
            if (tab(sb + half - 1) < mo) sb += half;
            len -= half;
          }
          if (tab(sb) < mo) ++sb;
          const int cc = sb - 1;
          int ot = max(cc, 0) * S;
          if (cc >= 0)
            for (int k = 1; k < S; ++k) {
              const int t = cc * S + k;
              if (t < tiles && O::make(tfirst[t], (uint32_t)t * TW) < mo) ot = t;
            }
          own = (uint32_t)ot;
        }
        const int oc = (int)(own / kChunkTiles) - ch + 1;
        if ((unsigned)oc < 3u) atomicAdd(&s_cc[oc], 1u);
        else atomicAdd(&tcount[tiles + own / kChunkTiles], kOwn);
        const uint32_t slot = (uint32_t)(atomicAdd(&tcount[own], kOwn) >> 32);
        if (slot < (uint32_t)kBucketCap) {
          uint4* e = r.bucket + 2 * ((size_t)own * kBucketCap + slot);
          e[0] = make_uint4((uint32_t)mk, (uint32_t)(mk >> 32), mi, mp);
          e[1] = make_uint4(__float_as_uint(mx), __float_as_uint(my), __float_as_uint(mu),
                            __float_as_uint(mv));
        } else if (slot == (uint32_t)kBucketCap && atomicCAS(r.claim, 0u, 1u) == 0u) {
          s_sort = 1u;  // the first overflow of a bucket, and the first claim of the step
        }
      }
      __syncthreads();
    }
    if (tid < 3) {
      const int oc = ch - 1 + tid;
      const unsigned long long v =
          (unsigned long long)s_cc[tid] * kOwn + (tid == 1 ? (unsigned long long)tot : 0ull);
      if (v != 0ull && oc >= 0 && oc * kChunkTiles < tiles) atomicAdd(&tcount[tiles + oc], v);
    }
  }
  // the step's mover count (read by the merge of the sorted path and, through
  // k_ranges, by the host); s_sort is complete after the last round's barrier
  if (tid == 0 && tot != 0u) atomicAdd(r.n_movers, (int)tot);
  if (s_sort == 0u) return;
  // sorted path: every mover in previous-index order, then sorted by (key,
  // previous index); the merge searches the tile-first keys in them
  const uint32_t m = compact_movers(n, tiles, flags, r.tile_base, r.midx, s_mem, s_w);
  __syncthreads();
  sort_movers((int)m, passes, knew, r, s_mem);
  if (tid == 0) *r.mode = 0u;
}

// After a producer that bucketed the movers itself (bucket_movers), which
// adds only to the tile counters: each wave sums one chunk's kChunkTiles tile
// counters (both halves) into the chunk counter the merge reads; and when a
// bucket overflowed (the claim word), workgroup 0 then orders every mover for
// the sorted path, with the same ordered compaction and stable radix sort as
// k_bucket's.
template <typename K>
__global__ __launch_bounds__(KB_TW) void k_chunks(int n, int tiles, int passes,
                                                  const K* __restrict__ knew,
                                                  const unsigned long long* __restrict__ flags,
                                                  unsigned long long* __restrict__ tcount,
                                                  ResortBuffers r) {
  __shared__ uint64_t s_mem64[KB_LDS / 2];
  __shared__ uint32_t s_w[KB_WAVES];
  static_assert(kChunkTiles == kWave, "one wave per chunk, one lane per tile");
  const int lane = threadIdx.x & (kWave - 1);
  const int chunks = (tiles + kChunkTiles - 1) / kChunkTiles;
  const int ch = (int)blockIdx.x * KB_WAVES + (int)threadIdx.x / kWave;
  const int t = ch * kChunkTiles + lane;
  const unsigned long long v = tcount[min(t, tiles - 1)];
  const bool in = ch < chunks && t < tiles;
  const uint32_t lo = wave_sum(in ? (uint32_t)v : 0u), hi = wave_sum(in ? (uint32_t)(v >> 32) : 0u);
  if (ch < chunks && lane == 0) tcount[tiles + ch] = ((unsigned long long)hi << 32) | lo;
  if (blockIdx.x != 0 || *r.claim == 0u) return;
  uint32_t* s_mem = reinterpret_cast<uint32_t*>(s_mem64);
  const uint32_t m = compact_movers(n, tiles, flags, r.tile_base, r.midx, s_mem, s_w);
  __syncthreads();
  sort_movers((int)m, passes, knew, r, s_mem);
  if (threadIdx.x == 0) {
    *r.n_movers = (int)m;
    *r.mode = 0u;
  }
}

// ------------------------------------------------------------------- merge
// Per tile of the previous order [j0, j0 + TW): its window of sorted movers
// [lo, hi) holds the movers ordered after the tile's first particle and
// before the next tile's first particle (tile 0: from the first mover), so
// every mover belongs to exactly one window.
//   stayer j:  slot = (j - movers before j) + lo + (window movers before it)
//   mover r:   slot = r + (stayers before it), where the previous-order
//              particles before it are j0 + (tile particles before it)
// Bucket path: the movers before the tile by previous index and the window
// start are the two halves of the sum of the packed counters of the chunk's
// earlier tiles and of the earlier chunks; the window is the tile's bucket.
// Sorted path: the movers before the tile from tile_base, the window's ends by
// binary search of the tile-first keys in the sorted movers.
// A workgroup takes MT consecutive tiles, each thread one particle of each:
// every load of a particle's state is issued in one round for all of them, so
// at 1M particles the whole grid is resident at once.
#ifndef SPH_MERGE_TILES
#define SPH_MERGE_TILES 4  // 1, 2, 4 (buckets of 64) and 8 (of 32) measured at 1M: 4 fastest
#endif
constexpr int MT = SPH_MERGE_TILES;  // tiles per merge workgroup
constexpr int BH = TW / MT;    // threads that hold one tile's bucket entries
static_assert(kBucketCap <= BH && MT * kBucketCap <= WIN,
              "a tile's bucket maps to its share of the threads and of the window buffer");

// The merge's LDS, also the scratch of the sorted path that workgroup 0 of a
// claimed step runs first when the merge forms the chunk counters itself.
struct MergeShared {
  uint64_t tk[MT][TW];
  uint64_t mk[WIN];
  uint32_t mi[WIN];
  uint64_t ak[MT][kBucketCap];
  uint32_t ai[MT][kBucketCap];
  alignas(16) uint32_t ms[MT][kBucketCap][4];
  uint32_t mp[MT][kBucketCap];
  uint32_t wpre[MT][TW / kWave + 1];
  unsigned long long fw[MT][TW / kWave];
  uint32_t plo[TW / kWave], phi[TW / kWave];
  uint32_t win[MT + 1];
};
union MergeLds {
  MergeShared m;
  struct {
    uint32_t mem[KB_LDS];
    uint32_t w[KB_WAVES];
  } sort;
};

// Published chunk counters (bit 63 set by the publishing workgroup).
constexpr unsigned long long kPublished = 1ull << 63;
constexpr uint32_t kMaxChunkPolls = 1u << 20;   // ~60 ms of polling at ~128 clk each
constexpr uint32_t kMaxReadyPolls = 1u << 24;   // ~1 s: the sorted path's one-workgroup sort
constexpr uint32_t kStatusMergeWait = 8u;       // a merge wait gave up (status word)

__device__ __forceinline__ void st_agent64(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The counter a workgroup of lower id publishes, without its flag bit; gives
// up (status bit, value 0) if it never appears.
__device__ __forceinline__ unsigned long long wait_published(unsigned long long* p,
                                                             uint32_t* status) {
  unsigned long long v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t polls = 0;
  while (!(v & kPublished)) {
    if (++polls >= kMaxChunkPolls) {
      atomicOr(status, kStatusMergeWait);
      return 0ull;
    }
    __builtin_amdgcn_s_sleep(2);
    v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return v & ~kPublished;
}


template <typename K, bool PUB>
__global__ __launch_bounds__(TW) void k_merge(
    int n, int passes, const K* __restrict__ knew, const K* __restrict__ kold,
    unsigned long long* __restrict__ flags,
    const uint64_t* __restrict__ tfirst, unsigned long long* __restrict__ tcount,
    const float* __restrict__ tx, const float* __restrict__ ty, const float* __restrict__ tvx,
    const float* __restrict__ tvy, const uint32_t* __restrict__ pid_src, float* __restrict__ rx,
    float* __restrict__ ry, float* __restrict__ vx, float* __restrict__ vy,
    uint32_t* __restrict__ pid_dst, K* __restrict__ keys_dst, uint32_t* __restrict__ perm,
    int2* __restrict__ cells, int2* __restrict__ blocks, int block_bits, uint64_t n_keys,
    float4* __restrict__ xyuv, const float4* __restrict__ t4, ResortBuffers r) {
  constexpr bool NARROW = sizeof(K) == 4;  // a narrow step: T packed in t4
  // (workgroup 0 of a claimed PUB step first orders the movers in the same LDS)
  __shared__ MergeLds lds;
  auto& s_tk = lds.m.tk;  // the tiles' previous keys (all ones past n)
  // window keys and previous indices: bucket path tile q at q * kBucketCap,
  // sorted path one tile's chunk at a time
  auto& s_mk = lds.m.mk;
  auto& s_mi = lds.m.mi;
  auto& s_ak = lds.m.ak;  // bucket path: keys, arrival order
  auto& s_ai = lds.m.ai;
  auto& s_ms = *reinterpret_cast<float4 (*)[MT][kBucketCap]>(&lds.m.ms);  // bucket path:
  auto& s_mp = lds.m.mp;                                                  // sorted order
  auto& s_wpre = lds.m.wpre;  // movers of a tile before each wave
  auto& s_fw = lds.m.fw;      // the tiles' flag words
  auto& s_plo = lds.m.plo;    // counter sums per wave
  auto& s_phi = lds.m.phi;
  auto& s_win = lds.m.win;  // sorted path: the windows' ends
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int tiles = (n + TW - 1) / TW;
  const int t0 = (int)blockIdx.x * MT;
  uint32_t claimed = 0u;
  if constexpr (PUB) {
    // chunk blockIdx.x's counter from its tile counters, published with a
    // flag bit before anything else (a workgroup waits only for chunks of
    // lower-id workgroups, so every chunk it waits for is published by a
    // workgroup that is resident or done)
    const int cb = (int)blockIdx.x;
    if (wave == 0 && cb * kChunkTiles < tiles) {
      const int t = cb * kChunkTiles + lane;
      const unsigned long long cv = tcount[min(t, tiles - 1)];
      const bool in = t < tiles;
      const uint32_t lo = wave_sum(in ? (uint32_t)cv : 0u);
      const uint32_t hi = wave_sum(in ? (uint32_t)(cv >> 32) : 0u);
      if (lane == 0) st_agent64(&tcount[tiles + cb], kPublished | ((uint64_t)hi << 32) | lo);
    }
    claimed = *r.claim;
    if (claimed != 0u && blockIdx.x == 0) {
      // a bucket overflowed: order every mover here first (ordered compaction
      // and a stable radix sort, as k_bucket's sorted path), then release the
      // other workgroups, which wait for it before they place their windows
      const uint32_t m =
          compact_movers<4>(n, tiles, flags, r.tile_base, r.midx, lds.sort.mem, lds.sort.w);
      __syncthreads();
      sort_movers<K, 2>((int)m, passes, knew, r, lds.sort.mem);
      __syncthreads();
      if (tid == 0) {
        *r.n_movers = (int)m;
        *r.mode = 0u;
        __hip_atomic_store(r.ready, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
      __syncthreads();
    }
  }
  // every load that depends only on the particles, issued together
  unsigned long long fw[MT];
  uint64_t k[MT];
  float x[MT], y[MT], u[MT], v[MT];
  uint32_t id[MT];
  bool live[MT];
#pragma unroll
  for (int q = 0; q < MT; ++q) {
    const int j0 = (t0 + q) * TW, j = j0 + tid;
    live[q] = j < n;
    fw[q] = flag_word(flags, j0 / kWave + wave, n);
    const int jc = min(j, n - 1);
    const uint64_t kc = (uint64_t)kold[jc];
    k[q] = live[q] ? kc : ~0ull;
    if constexpr (NARROW) {
      const float4 tq = t4[jc];
      x[q] = tq.x;
      y[q] = tq.y;
      u[q] = tq.z;
      v[q] = tq.w;
    } else {
      x[q] = tx[jc];
      y[q] = ty[jc];
      u[q] = tvx[jc];
      v[q] = tvy[jc];
    }
    id[q] = pid_src[jc];
  }
  // the packed counters: the chunk's tiles before t0 (lanes of wave 0), the
  // chunks before it, and the workgroup's own tiles
  const int ch = t0 / kChunkTiles, tin = t0 % kChunkTiles;
  const unsigned long long tcv =
      tcount[min(ch * kChunkTiles + (tid & (kChunkTiles - 1)), tiles - 1)];
  unsigned long long own[MT];
#pragma unroll
  for (int q = 0; q < MT; ++q) own[q] = tcount[min(t0 + q, tiles - 1)];
  unsigned long long pc = tid < tin ? tcv : 0ull;
  // the first entries of each tile's bucket (thread qb * BH + rb: entry rb of
  // tile t0 + qb), loaded whichever path the step took (entries past the
  // window's length are stale and not used)
  const int qb = tid / BH, rb = tid % BH;
  const int tq = min(t0 + qb, tiles - 1);
  const uint4* be = r.bucket + 2 * ((size_t)tq * kBucketCap + (rb & (kBucketFast - 1)));
  uint4 b0 = be[0], b1 = be[1];
  bool from_bucket;
  if constexpr (PUB) {
    // the earlier chunks' counters, waited for (see the start of the kernel)
    for (int q = tid; q < ch; q += TW) pc += wait_published(&tcount[tiles + q], r.status);
    from_bucket = claimed == 0u;
  } else {
    for (int q = tid; q < ch; q += TW) pc += tcount[tiles + q];
    from_bucket = *r.mode != 0u;
  }
  // k_ranges writes the range of every cell and block that holds particles in
  // the new order, so of the previous ranges only those of cells and blocks
  // that may have become empty are cleared: the ones a mover left
#pragma unroll
  for (int q = 0; q < MT; ++q)
    if (live[q] && ((fw[q] >> lane) & 1ull) && k[q] < n_keys) {
      cells[k[q]] = make_int2(-1, -1);
      if (block_bits > 0) blocks[k[q] >> (2 * block_bits)] = make_int2(-1, -1);
    }
#pragma unroll
  for (int q = 0; q < MT; ++q) s_tk[q][tid] = k[q];
  {
    const uint32_t plo = wave_sum((uint32_t)pc), phi = wave_sum((uint32_t)(pc >> 32));
    if (lane == 0) {
      s_plo[wave] = plo;
      s_phi[wave] = phi;
#pragma unroll
      for (int q = 0; q < MT; ++q) {
        s_wpre[q][wave + 1] = (uint32_t)__popcll(fw[q]);
        s_fw[q][wave] = fw[q];
      }
    }
  }
  if constexpr (PUB) {
    if (!from_bucket) {
      // the sorted movers of workgroup 0
      if (tid == 0) {
        uint32_t polls = 0;
        while (__hip_atomic_load(r.ready, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
          if (++polls >= kMaxReadyPolls) {
            atomicOr(r.status, kStatusMergeWait);
            break;
          }
          __builtin_amdgcn_s_sleep(2);
        }
      }
      __syncthreads();
    }
  }
  if (!from_bucket && tid <= MT) {
    // sorted path: movers ordered before the first particle of tile t0 + tid
    // (all of them past the last tile)
    const int m = *r.n_movers;
    const int t = t0 + tid;
    s_win[tid] = t == 0 ? 0u
                 : t >= tiles
                     ? (uint32_t)m
                     : (uint32_t)rank_in_sorted(r.skey, r.sidx, 0, m, tfirst[t], (uint32_t)t * TW);
  }
  __syncthreads();
  if (tid < MT) {
    s_wpre[tid][0] = 0;
    for (int w = 1; w <= TW / kWave; ++w) s_wpre[tid][w] += s_wpre[tid][w - 1];
  }
  // the tiles' flag words are read (s_fw): zero them for producers that OR in
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < MT; ++q)
      if (t0 + q < tiles) flags[(t0 + q) * (TW / kWave) + wave] = 0ull;
  uint32_t tb[MT];
  int lo[MT], hi[MT];
  {
    uint32_t wlo[MT], whi[MT];
    if (from_bucket) {
      uint32_t slo = 0, shi = 0;
#pragma unroll
      for (int w = 0; w < TW / kWave; ++w) {
        slo += s_plo[w];
        shi += s_phi[w];
      }
#pragma unroll
      for (int q = 0; q < MT; ++q) {
        const unsigned long long o = t0 + q < tiles ? own[q] : 0ull;
        tb[q] = slo;
        wlo[q] = shi;
        whi[q] = shi + (uint32_t)(o >> 32);
        slo += (uint32_t)o;
        shi = whi[q];
      }
      // the last workgroup's prefix is the step's mover count (producers that
      // bucket the movers themselves do not add it up)
      if (t0 + MT >= tiles && tid == 0) *r.n_movers = (int)slo;
    } else {
#pragma unroll
      for (int q = 0; q < MT; ++q) {
        tb[q] = r.tile_base[min(t0 + q, tiles - 1)];
        wlo[q] = s_win[q];
        whi[q] = s_win[q + 1];
      }
    }
    // windows lie in [0, n) (clamped, so that a bad table cannot read past skey)
#pragma unroll
    for (int q = 0; q < MT; ++q) {
      lo[q] = min((int)wlo[q], n);
      hi[q] = max(lo[q], min((int)whi[q], n));
    }
  }
  __syncthreads();
  bool mv[MT];
  int before[MT];  // window movers ordered before a stayer
#pragma unroll
  for (int q = 0; q < MT; ++q) {
    mv[q] = live[q] && ((fw[q] >> lane) & 1ull);
    before[q] = 0;
  }
  // a mover with key km, previous index src and its state goes to window slot
  // c of tile q: rank among the tile's keys, minus the tile's movers before it
  auto place = [&](int q, int c, uint64_t km, uint32_t src, float mx, float my, float mu,
                   float mw, uint32_t mid_) {
    const int j0 = (t0 + q) * TW;
    int a = 0, z = TW;
    while (a < z) {
      const int mid = (a + z) >> 1;
      if (pair_less(s_tk[q][mid], (uint32_t)(j0 + mid), km, src)) a = mid + 1;
      else z = mid;
    }
    const int w = a / kWave, l = a % kWave;
    uint32_t mb = tb[q] + s_wpre[q][w];
    if (w < TW / kWave) mb += (uint32_t)__popcll(s_fw[q][w] & ((1ull << l) - 1ull));
    const int pos = c + (j0 + a - (int)mb);
    if ((unsigned)pos >= (unsigned)n) return;  // only a corrupt table gets here
    if (rx) {
      rx[pos] = mx;
      ry[pos] = my;
      vx[pos] = mu;
      vy[pos] = mw;
    }
    xyuv[pos] = make_float4(mx, my, mu, mw);
    pid_dst[pos] = mid_;
    keys_dst[pos] = (K)km;
    if (perm) perm[pos] = src;
  };
  auto stayer_rank = [&](const uint64_t* mk, const uint32_t* mi, int len, uint64_t kk,
                         uint32_t j) {
    int a = 0, z = len;
    while (a < z) {
      const int mid = (a + z) >> 1;
      if (pair_less(mk[mid], mi[mid], kk, j)) a = mid + 1;
      else z = mid;
    }
    return a;
  };
  if (from_bucket) {
    // each tile's window is its bucket (at most kBucketCap movers): keys in
    // arrival order into LDS, ranked by counting, then placed
    const int blen = min(hi[qb] - lo[qb], kBucketCap);
    if (rb < blen) {
      if (rb >= kBucketFast) {
        const uint4* bx = r.bucket + 2 * ((size_t)tq * kBucketCap + rb);
        b0 = bx[0];
        b1 = bx[1];
      }
      s_ak[qb][rb] = (uint64_t)b0.x | ((uint64_t)b0.y << 32);
      s_ai[qb][rb] = b0.z;
    }
    __syncthreads();
    if (rb < blen) {
      const uint64_t km = (uint64_t)b0.x | ((uint64_t)b0.y << 32);
      const uint32_t im = b0.z;
      int rank = 0;
      for (int e = 0; e < blen; ++e) rank += pair_less(s_ak[qb][e], s_ai[qb][e], km, im) ? 1 : 0;
      s_mk[qb * kBucketCap + rank] = km;
      s_mi[qb * kBucketCap + rank] = im;
      s_ms[qb][rank] = make_float4(__uint_as_float(b1.x), __uint_as_float(b1.y),
                                   __uint_as_float(b1.z), __uint_as_float(b1.w));
      s_mp[qb][rank] = b0.w;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MT; ++q)
      if (live[q] && !mv[q])
        before[q] = stayer_rank(s_mk + q * kBucketCap, s_mi + q * kBucketCap,
                                min(hi[q] - lo[q], kBucketCap), k[q],
                                (uint32_t)((t0 + q) * TW + tid));
    if (rb < blen) {
      const float4 st = s_ms[qb][rb];
      place(qb, lo[qb] + rb, s_mk[qb * kBucketCap + rb], s_mi[qb * kBucketCap + rb], st.x, st.y,
            st.z, st.w, s_mp[qb][rb]);
    }
  } else {
    // sorted path: each tile's window in chunks of WIN sorted movers, the
    // movers' state gathered from T
    for (int q = 0; q < MT; ++q) {
      for (int c0 = lo[q]; c0 < hi[q]; c0 += WIN) {
        const int len = min(WIN, hi[q] - c0);
        for (int e = tid; e < len; e += TW) {
          s_mk[e] = r.skey[c0 + e];
          s_mi[e] = r.sidx[c0 + e];
        }
        __syncthreads();
        if (live[q] && !mv[q])
          before[q] += stayer_rank(s_mk, s_mi, len, k[q], (uint32_t)((t0 + q) * TW + tid));
        for (int e = tid; e < len; e += TW) {
          const uint32_t src = s_mi[e];
          const uint32_t sc = min(src, (uint32_t)(n - 1));
          if constexpr (NARROW) {
            const float4 tq = t4[sc];
            place(q, c0 + e, s_mk[e], src, tq.x, tq.y, tq.z, tq.w, pid_src[sc]);
          } else {
            place(q, c0 + e, s_mk[e], src, tx[sc], ty[sc], tvx[sc], tvy[sc], pid_src[sc]);
          }
        }
        __syncthreads();
      }
    }
  }
#pragma unroll
  for (int q = 0; q < MT; ++q) {
    if (!live[q] || mv[q]) continue;
    const int j = (t0 + q) * TW + tid;
    const uint32_t mb = tb[q] + s_wpre[q][wave] + (uint32_t)__popcll(fw[q] & ((1ull << lane) - 1ull));
    const int pos = (j - (int)mb) + lo[q] + before[q];
    if ((unsigned)pos >= (unsigned)n) continue;  // only a corrupt table gets here
    if (rx) {
      rx[pos] = x[q];
      ry[pos] = y[q];
      vx[pos] = u[q];
      vy[pos] = v[q];
    }
    xyuv[pos] = make_float4(x[q], y[q], u[q], v[q]);
    pid_dst[pos] = id[q];
    keys_dst[pos] = (K)k[q];
    if (perm) perm[pos] = (uint32_t)j;
  }
}

}  // namespace

int64_t resort_flag_words(int n) { return ((int64_t)n + TW - 1) / TW * (TW / kWave); }
int64_t resort_tiles(int n) { return ((int64_t)n + TW - 1) / TW; }
int64_t resort_count_words(int n) {
  const int64_t t = resort_tiles(n);
  return t + (t + kChunkTiles - 1) / kChunkTiles;
}

hipError_t launch_movers(const Launch& L, const DevState& s, const ResortBuffers& r,
                         const uint64_t* knew, int key_bits, bool flags_ready, bool bucketed) {
  if (s.n <= 0) return hipSuccess;
  const int tiles = (int)resort_tiles(s.n);
  const int passes = (key_bits + 7) / 8;
  if (passes < 1 || passes > kMaxPasses) return hipErrorInvalidValue;
  if (bucketed) {
    // the producer bucketed the movers: the chunk counters, and the sorted
    // path when claimed
    if (!flags_ready) return hipErrorInvalidValue;
    const int chunks = (tiles + kChunkTiles - 1) / kChunkTiles;
    const int grid = (chunks + KB_WAVES - 1) / KB_WAVES;
    if (s.narrow)
      k_chunks<uint32_t><<<grid, KB_TW, 0, L.stream>>>(s.n, tiles, passes, s.nkeys_in, r.flags,
                                                       s.tile_count, r);
    else
      k_chunks<uint64_t><<<grid, KB_TW, 0, L.stream>>>(s.n, tiles, passes, knew, r.flags,
                                                       s.tile_count, r);
    return hipGetLastError();
  }
  if (!flags_ready) {
    // producers without their own report write u64 keys
    if (s.narrow) return hipErrorInvalidValue;
    k_mover_flags<<<(s.n + MF_TW - 1) / MF_TW, MF_TW, 0, L.stream>>>(
        s.n, (int)resort_flag_words(s.n), knew, s.keys_sorted, mover_sink(r));
  }
  const int words = (int)((s.n + kWave - 1) / kWave);
  const int grid = (words + KB_TW - 1) / KB_TW;
  // keys of at most 32 bits: (key, index) in one word
  if (s.narrow)
    k_bucket<false, uint32_t><<<grid, KB_TW, 0, L.stream>>>(
        s.n, tiles, passes, s.nkeys_in, r.flags, s.tile_first, s.tile_count, s.cells,
        (uint64_t)s.n_keys, s.tx, s.ty, s.tvx, s.tvy, s.pid, s.t4, r);
  else if (passes <= 4)
    k_bucket<false, uint64_t><<<grid, KB_TW, 0, L.stream>>>(
        s.n, tiles, passes, knew, r.flags, s.tile_first, s.tile_count, s.cells,
        (uint64_t)s.n_keys, s.tx, s.ty, s.tvx, s.tvy, s.pid, nullptr, r);
  else
    k_bucket<true, uint64_t><<<grid, KB_TW, 0, L.stream>>>(
        s.n, tiles, passes, knew, r.flags, s.tile_first, s.tile_count, s.cells,
        (uint64_t)s.n_keys, s.tx, s.ty, s.tvx, s.tvy, s.pid, nullptr, r);
  return hipGetLastError();
}

template <bool PUB>
static hipError_t merge_go(const Launch& L, const DevState& s, const ResortBuffers& r,
                           const uint64_t* knew, int passes, uint64_t* keys_dst, bool soa,
                           bool perm) {
  const int tiles = (int)resort_tiles(s.n);
  const int grid = (tiles + MT - 1) / MT;
  if (s.narrow)  // the step's narrow keys: nkeys into nkeys_alt
    k_merge<uint32_t, PUB><<<grid, TW, 0, L.stream>>>(
        s.n, passes, s.nkeys_in, s.nkeys, r.flags, s.tile_first, s.tile_count, s.tx, s.ty,
        s.tvx, s.tvy, s.pid, soa ? s.rx : nullptr, s.ry, s.vx, s.vy, s.pid_alt, s.nkeys_alt,
        perm ? s.perm : nullptr, s.cells, s.blocks, s.block_bits, (uint64_t)s.n_keys, s.xyuv,
        s.t4, r);
  else
    k_merge<uint64_t, PUB><<<grid, TW, 0, L.stream>>>(
        s.n, passes, knew, s.keys_sorted, r.flags, s.tile_first, s.tile_count, s.tx, s.ty,
        s.tvx, s.tvy, s.pid, soa ? s.rx : nullptr, s.ry, s.vx, s.vy, s.pid_alt, keys_dst,
        perm ? s.perm : nullptr, s.cells, s.blocks, s.block_bits, (uint64_t)s.n_keys, s.xyuv,
        nullptr, r);
  return hipGetLastError();
}

hipError_t launch_merge(const Launch& L, const DevState& s, const ResortBuffers& r,
                        uint64_t* keys_dst, bool soa, bool perm, const uint64_t* knew,
                        int key_bits) {
  if (s.n <= 0) return hipSuccess;
  if (knew) {
    const int passes = (key_bits + 7) / 8;
    if (passes < 1 || passes > kMaxPasses) return hipErrorInvalidValue;
    return merge_go<true>(L, s, r, knew, passes, keys_dst, soa, perm);
  }
  return merge_go<false>(L, s, r, nullptr, 0, keys_dst, soa, perm);
}

}  // namespace sphamd
