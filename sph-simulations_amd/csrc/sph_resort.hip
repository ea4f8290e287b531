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
//             and by owner (high half). The workgroup that completes the tally
//             word (one atomic per workgroup) chooses the path: when a bucket
//             overflowed it orders the movers itself (ordered compaction, then
//             a stable radix sort) for the sorted path.
//   k_merge   one workgroup per tile of the previous order: clears the tile's
//             old cell / block ranges; bucket path: its prefixes from the
//             packed counters, its window (the movers it owns) from its bucket,
//             sorted in LDS by counting; sorted path: the window from the sorted
//             movers. It places its stayers (rank among the window's movers,
//             searched in LDS) and the window's movers (rank among the tile's
//             keys, searched in LDS); scatters r, v, pid, key and the
//             permutation
//   k_ranges  cell start/end and Morton block ranges of the new order, the
//             tile-first keys, and zeroes the counters for the next step
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
// (tile_base); returns their number. Per batch each wave takes FB_WPT * 64
// consecutive flag words (lane l: words q * 64 + l, coalesced); the prefixes
// come from wave scans plus the waves' offsets. s_pre: FB_WPT * KB_TW words,
// s_w: KB_WAVES words (LDS).
__device__ uint32_t compact_movers(int n, int tiles, const unsigned long long* __restrict__ flags,
                                   uint32_t* __restrict__ tile_base,
                                   uint32_t* __restrict__ midx, uint32_t* s_pre, uint32_t* s_w) {
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
__device__ void sort_movers(int m, int passes, const uint64_t* __restrict__ knew,
                            const ResortBuffers& r, uint32_t* s) {
  uint32_t* s_cnt = s;                    // [KB_WAVES][256] per-wave digit counts
  uint32_t* s_off = s + KB_WAVES * 256;  // [256] running digit offsets
  uint32_t* s_tot = s_off + 256;          // [256] digit totals of a chunk
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  constexpr int HB = 8;  // keys per thread in flight while counting
  for (int p = 0; p < passes; ++p) {
    const uint64_t* kin = (p & 1) ? r.ktmp : r.mkey;
    const uint32_t* vin = p == 0 ? r.midx : ((p & 1) ? r.vtmp : r.midx);
    uint64_t* kout = p == passes - 1 ? r.skey : ((p & 1) ? r.mkey : r.ktmp);
    uint32_t* vout = p == passes - 1 ? r.sidx : ((p & 1) ? r.midx : r.vtmp);
    auto key_of = [&](int e) { return p == 0 ? knew[vin[e]] : kin[e]; };
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
// workgroup adds its movers, itself and whether one of its movers found a full
// bucket to the tally word in one atomic; the one that completes the tally
// tells the merge which path (*r.mode: 1 buckets, 0 sorted) and, for the
// sorted path, orders the movers itself. It reads nothing other workgroups
// wrote except the tally (flags and keys come from the previous kernel), so
// no workgroup needs a device-scope fence.
constexpr int kBucketList = 2048;
constexpr int kBucketTab = 1024;  // sampled tile-first keys in LDS (every S-th tile)
static_assert(kBucketTab % KB_TW == 0, "tile-first table loads");
static_assert(KB_TW == 4 * kChunkTiles, "one k_bucket workgroup per counter chunk");
// LDS words of the list and the table, reused by the sorted path
constexpr int KB_LDS = kBucketList + 2 * kBucketTab;
static_assert(FB_WPT * KB_TW <= KB_LDS && KB_WAVES * 256 + 512 <= KB_LDS,
              "the sorted path's LDS exceeds the list and table");

template <bool WIDE>
__global__ __launch_bounds__(KB_TW) void k_bucket(
    int n, int tiles, int passes, const uint64_t* __restrict__ knew,
    const unsigned long long* __restrict__ flags, const uint64_t* __restrict__ tfirst,
    unsigned long long* __restrict__ tcount, const int2* __restrict__ cells, uint64_t n_keys,
    const float* __restrict__ tx, const float* __restrict__ ty, const float* __restrict__ tvx,
    const float* __restrict__ tvy, const uint32_t* __restrict__ pid_src, ResortBuffers r) {
  using O = Ord<WIDE>;
  __shared__ uint64_t s_mem64[KB_LDS / 2];
  __shared__ uint32_t s_w[KB_WAVES];
  __shared__ uint32_t s_cc[3];  // owners in the chunks before, of and after this one
  __shared__ uint32_t s_full;   // a mover of this workgroup found its bucket full
  __shared__ unsigned long long s_tally;  // the tally with this workgroup added, last one only
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
  if (tid == 0) s_full = 0u;
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
        const uint64_t mk = knew[mi];
        const float mx = tx[mi], my = ty[mi], mu = tvx[mi], mv = tvy[mi];
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
        } else {
          s_full = 1u;
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
  // the tally: this workgroup's movers, itself, and whether it found a full
  // bucket (s_full is complete after the last round's barrier)
  if (tid == 0) {
    const unsigned long long add =
        (unsigned long long)tot + (1ull << 32) + (s_full ? (1ull << 48) : 0ull);
    const unsigned long long t = atomicAdd(r.tally, add) + add;
    s_tally = ((t >> 32) & 0xffffull) == gridDim.x ? t : 0ull;
  }
  __syncthreads();
  const unsigned long long tally = s_tally;
  if (tally == 0ull) return;  // not the last workgroup
  if ((tally >> 48) == 0ull) {
    if (tid == 0) {
      *r.n_movers = (int)(uint32_t)tally;
      *r.mode = 1u;
    }
    return;
  }
  // sorted path: every mover in previous-index order, then sorted by (key,
  // previous index); the merge searches the tile-first keys in them
  const uint32_t m = compact_movers(n, tiles, flags, r.tile_base, r.midx, s_mem, s_w);
  __syncthreads();
  sort_movers((int)m, passes, knew, r, s_mem);
  if (tid == 0) {
    *r.n_movers = (int)m;
    *r.mode = 0u;
  }
}

// ------------------------------------------------------------------- merge
// One workgroup per tile of the previous order [j0, j0 + TW). Its window of
// sorted movers [lo, hi) holds the movers ordered after the tile's first
// particle and before the next tile's first particle (tile 0: from the first
// mover), so every mover belongs to exactly one window.
//   stayer j:  slot = (j - movers before j) + lo + (window movers before it)
//   mover r:   slot = r + (stayers before it), where the previous-order
//              particles before it are j0 + (tile particles before it)
// Bucket path: the movers before the tile by previous index and the window
// start are the two halves of the sum of the packed counters of the chunk's
// earlier tiles and of the earlier chunks; the window is the tile's bucket.
// Sorted path: the movers before the tile from tile_base, the window's ends by
// binary search of the tile-first keys in the sorted movers.
__global__ __launch_bounds__(TW) void k_merge(
    int n, const uint64_t* __restrict__ kold, unsigned long long* __restrict__ flags,
    const uint64_t* __restrict__ tfirst, const unsigned long long* __restrict__ tcount,
    const float* __restrict__ tx, const float* __restrict__ ty, const float* __restrict__ tvx,
    const float* __restrict__ tvy, const uint32_t* __restrict__ pid_src, float* __restrict__ rx,
    float* __restrict__ ry, float* __restrict__ vx, float* __restrict__ vy,
    uint32_t* __restrict__ pid_dst, uint64_t* __restrict__ keys_dst, uint32_t* __restrict__ perm,
    int2* __restrict__ cells, int2* __restrict__ blocks, int block_bits, uint64_t n_keys,
    float4* __restrict__ xyuv, ResortBuffers r) {
  __shared__ uint64_t s_tk[TW];  // the tile's previous keys (all ones past n)
  __shared__ uint64_t s_mk[WIN];
  __shared__ uint32_t s_mi[WIN];
  __shared__ uint32_t s_wpre[TW / kWave + 1];  // movers of the tile before each wave
  __shared__ unsigned long long s_fw[TW / kWave];  // the tile's flag words
  __shared__ uint32_t s_plo[TW / kWave], s_phi[TW / kWave];  // counter sums per wave
  __shared__ uint32_t s_win[2];                    // sorted path: the window's ends
  __shared__ uint64_t s_bk_k[kBucketCap];          // the tile's bucket, arrival order
  __shared__ uint32_t s_bk_i[kBucketCap];
  __shared__ uint4 s_bk_s[kBucketCap];             // (x, y, vx, vy) bits, arrival order
  __shared__ uint32_t s_bk_p[kBucketCap];          // pid, arrival order
  __shared__ float4 s_ms[kBucketCap];              // the window's state, sorted order
  __shared__ uint32_t s_mp[kBucketCap];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int tile = blockIdx.x;
  const int tiles = (n + TW - 1) / TW;
  const int j0 = tile * TW, j = j0 + tid;
  const bool live = j < n;
  // every load that depends only on j, issued together
  const unsigned long long fw = flag_word(flags, j0 / kWave + wave, n);
  // the packed counters: the chunk's tiles before this one (lanes of wave 0)
  // and the chunks before this one
  const int ch = tile / kChunkTiles, tin = tile % kChunkTiles;
  const unsigned long long tcv =
      tcount[min(ch * kChunkTiles + (tid & (kChunkTiles - 1)), tiles - 1)];
  const unsigned long long own = tcount[tile];
  unsigned long long pc = tid < tin ? tcv : 0ull;
  for (int q = tid; q < ch; q += TW) pc += tcount[tiles + q];
  const int jc = min(j, n - 1);
  const uint64_t kc = kold[jc], kpc = kold[max(jc - 1, 0)];
  const uint64_t k = live ? kc : ~0ull;
  const uint64_t kp = (live && j > 0) ? kpc : ~k;
  const float x = tx[jc], y = ty[jc], u = tvx[jc], v = tvy[jc];
  const uint32_t id = pid_src[jc];
  // the first entries of the tile's bucket, loaded whichever path the step
  // took (entries past the window's length are stale and not used)
  const uint4* be = r.bucket + 2 * ((size_t)tile * kBucketCap + (tid & (kBucketFast - 1)));
  uint4 b0 = be[0], b1 = be[1];
  const bool from_bucket = *r.mode != 0u;
  if (tile == 0 && tid == 0) *r.tally = 0ull;  // for the next k_bucket (this one has ended)
  // the previous order's ranges (k_ranges writes the new ones)
  if (live && k != kp && k < n_keys) {
    cells[k] = make_int2(-1, -1);
    if (block_bits > 0 && (k >> (2 * block_bits)) != (kp >> (2 * block_bits)))
      blocks[k >> (2 * block_bits)] = make_int2(-1, -1);
  }
  s_tk[tid] = k;
  {
    const uint32_t plo = wave_sum((uint32_t)pc), phi = wave_sum((uint32_t)(pc >> 32));
    if (lane == 0) {
      s_plo[wave] = plo;
      s_phi[wave] = phi;
      s_wpre[wave + 1] = (uint32_t)__popcll(fw);
      s_fw[wave] = fw;
    }
  }
  if (!from_bucket && tid < 2) {
    // sorted path: movers ordered before this tile's (tid 0) and the next
    // tile's (tid 1) first particle
    const int m = *r.n_movers;
    const int t = tile + tid;
    s_win[tid] = t == 0 ? 0u
                 : t >= tiles
                     ? (uint32_t)m
                     : (uint32_t)rank_in_sorted(r.skey, r.sidx, 0, m, tfirst[t], (uint32_t)t * TW);
  }
  __syncthreads();
  if (tid == 0) {
    s_wpre[0] = 0;
    for (int w = 1; w <= TW / kWave; ++w) s_wpre[w] += s_wpre[w - 1];
  }
  // the tile's flag words are read (s_fw): zero them for producers that OR in
  if (lane == 0) flags[j0 / kWave + wave] = 0ull;
  uint32_t tb, wlo, whi;
  if (from_bucket) {
    uint32_t slo = 0, shi = 0;
#pragma unroll
    for (int w = 0; w < TW / kWave; ++w) {
      slo += s_plo[w];
      shi += s_phi[w];
    }
    tb = slo;
    wlo = shi;
    whi = shi + (uint32_t)(own >> 32);
  } else {
    tb = r.tile_base[tile];
    wlo = s_win[0];
    whi = s_win[1];
  }
  // windows lie in [0, n) (clamped, so that a bad table cannot read past skey)
  const int lo = min((int)wlo, n), hi = max(lo, min((int)whi, n));
  __syncthreads();
  const bool mv = live && ((fw >> lane) & 1ull);
  int before_movers = 0;  // window movers ordered before stayer j
  for (int c0 = lo; c0 < hi; c0 += WIN) {
    const int len = min(WIN, hi - c0);
    if (from_bucket) {
      // the window is the tile's bucket (at most kBucketCap movers, one
      // chunk): sorted here by (key, previous index), rank by counting
      const int blen = min(len, kBucketCap);
      if (blen > kBucketFast && tid >= kBucketFast && tid < blen) {
        const uint4* bx = r.bucket + 2 * ((size_t)tile * kBucketCap + tid);
        b0 = bx[0];
        b1 = bx[1];
      }
      if (tid < blen) {
        s_bk_k[tid] = (uint64_t)b0.x | ((uint64_t)b0.y << 32);
        s_bk_i[tid] = b0.z;
        s_bk_p[tid] = b0.w;
        s_bk_s[tid] = b1;
      }
      __syncthreads();
      for (int q = tid; q < blen; q += TW) {
        const uint64_t km = s_bk_k[q];
        const uint32_t im = s_bk_i[q];
        int rank = 0;
        for (int e = 0; e < blen; ++e) rank += pair_less(s_bk_k[e], s_bk_i[e], km, im) ? 1 : 0;
        s_mk[rank] = km;
        s_mi[rank] = im;
        const uint4 st = s_bk_s[q];
        s_ms[rank] = make_float4(__uint_as_float(st.x), __uint_as_float(st.y),
                                 __uint_as_float(st.z), __uint_as_float(st.w));
        s_mp[rank] = s_bk_p[q];
      }
    } else {
      for (int q = tid; q < len; q += TW) {
        s_mk[q] = r.skey[c0 + q];
        s_mi[q] = r.sidx[c0 + q];
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
    for (int q = tid; q < len; q += TW) {
      const uint64_t km = s_mk[q];
      const uint32_t src = s_mi[q];
      // the mover's state: from its bucket entry, or gathered from T (in
      // flight during the search)
      float mx, my, mu, mw;
      uint32_t mid_;
      if (from_bucket) {
        const float4 st = s_ms[q];
        mx = st.x;
        my = st.y;
        mu = st.z;
        mw = st.w;
        mid_ = s_mp[q];
      } else {
        const uint32_t sc = min(src, (uint32_t)(n - 1));
        mx = tx[sc];
        my = ty[sc];
        mu = tvx[sc];
        mw = tvy[sc];
        mid_ = pid_src[sc];
      }
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
      const int pos = (c0 + q) + (j0 + a - (int)mb);
      if ((unsigned)pos >= (unsigned)n) continue;  // only a corrupt table gets here
      if (rx) {
        rx[pos] = mx;
        ry[pos] = my;
        vx[pos] = mu;
        vy[pos] = mw;
      }
      xyuv[pos] = make_float4(mx, my, mu, mw);
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
int64_t resort_count_words(int n) {
  const int64_t t = resort_tiles(n);
  return t + (t + kChunkTiles - 1) / kChunkTiles;
}

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
      s.n, tiles, passes, knew, r.flags, s.tile_first, s.tile_count, s.cells,
      (uint64_t)s.n_keys, s.tx, s.ty, s.tvx, s.tvy, s.pid, r);
  return hipGetLastError();
}

hipError_t launch_merge(const Launch& L, const DevState& s, const ResortBuffers& r,
                        uint64_t* keys_dst, bool soa, bool perm) {
  if (s.n <= 0) return hipSuccess;
  const int tiles = (int)resort_tiles(s.n);
  k_merge<<<tiles, TW, 0, L.stream>>>(s.n, s.keys_sorted, r.flags, s.tile_first, s.tile_count,
                                      s.tx, s.ty, s.tvx, s.tvy, s.pid, soa ? s.rx : nullptr,
                                      s.ry, s.vx, s.vy, s.pid_alt, keys_dst,
                                      perm ? s.perm : nullptr, s.cells, s.blocks, s.block_bits,
                                      (uint64_t)s.n_keys, s.xyuv, r);
  return hipGetLastError();
}

}  // namespace sphamd
