// Internal interfaces between the C-ABI layer (sph_capi.hip) and the kernel
// translation units (sph_kernels.hip, sph_sort.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sph_amd.h"
#include "sph_device.hpp"

namespace sphamd {

constexpr int kMaxPasses = 8;  // 64-bit keys, 8-bit digits
constexpr int kResortTile = 256;     // particles per tile of the incremental re-sort
#ifndef SPH_BUCKET_CAP
#define SPH_BUCKET_CAP 64  // at most the merge's threads per tile (256 / tiles per workgroup)
#endif
constexpr int kBucketCap = SPH_BUCKET_CAP;  // movers a merge tile's bucket holds
constexpr int kBucketFast = 8;       // bucket entries the merge loads with its first round
constexpr int kChunkTiles = 64;      // merge tiles per counter chunk (one k_bucket workgroup)
// k_bucket's tally word counts its workgroups (one per chunk) in 16 bits:
// larger sets re-sort in full
constexpr int64_t kResortMaxParticles = 65535ll * kResortTile * kChunkTiles;
// the single-scene force kernel buckets the movers itself (bucket_movers)
// instead of a k_bucket launch
#ifndef SPH_FUSED_BUCKET
#define SPH_FUSED_BUCKET 1
#endif
constexpr bool kFusedBucket = SPH_FUSED_BUCKET != 0;
// after it, the merge forms the chunk counters and runs the sorted path
// itself (no k_chunks launch)
#ifndef SPH_MERGE_PREFIX
#define SPH_MERGE_PREFIX 1
#endif
constexpr bool kMergePrefix = SPH_MERGE_PREFIX != 0;

struct SortBuffers {
  // n passed to the launchers sizes the grids; when n_dev is set the kernels
  // sort the first *n_dev elements (a count produced on the device)
  const int* n_dev = nullptr;
  const uint32_t* vals_in = nullptr;  // first-pass values; null -> 0..n-1
  int pool = 64;                      // persistent workgroups when n_dev is set
  uint64_t* keys_in;   // unsorted keys (clobbered when passes >= 3)
  uint64_t* keys_tmp;
  uint64_t* keys_out;  // sorted keys
  uint32_t* vals_a;
  uint32_t* vals_b;
  uint32_t* vals_out;  // permutation: sorted slot -> input index
  uint32_t* hist;      // [kMaxPasses][256], zero on entry
  uint32_t* tickets;   // [kMaxPasses]
  uint32_t* lookback;  // [kMaxPasses][max_tiles][256]
  int max_tiles;
  int hist_blocks;
  // cell ranges written by the previous sort: the histogram kernel clears them
  // (cells of keys_out[0..clear_n) boundaries) before keys_out is overwritten
  int2* cells;
  int clear_n;
  uint64_t n_keys;
  uint32_t* status;  // nonzero after a look-back gave up (never expected)
  // Morton blocks of 2^block_bits x 2^block_bits cells (0: no block table)
  int2* blocks;
  int block_bits;
};

int sort_tile_size();
int sort_passes(int key_bits);
hipError_t radix_sort_pairs(const SortBuffers& b, int n, int key_bits, hipStream_t stream);
// the two launch kinds of radix_sort_pairs, for per-kernel timing
hipError_t radix_sort_hist(const SortBuffers& b, int n, int key_bits, hipStream_t stream);
hipError_t radix_sort_pass(const SortBuffers& b, int n, int key_bits, int pass,
                           hipStream_t stream);

// Particle arrays of one context. "S" is the sorted state the stages operate on;
// "T" holds an unsorted copy (integrated positions in the fused step, or the
// permutation target of the standalone find-neighbors stage).
struct DevState {
  int n;
  int n_prev;  // particles covered by keys_sorted / the cell table (0: none)
  float *rx, *ry, *vx, *vy;
  float *tx, *ty, *tvx, *tvy;
  float *ax, *ay, *tax, *tay;
  float *rho, *pr, *pres;
  float4* xyuv;  // sorted state packed per particle (x, y, vx, vy) for neighbor gathers
  float2* rp;    // (rho, pr) per particle, for neighbor gathers
  uint32_t *pid, *pid_alt;
  uint64_t* keys_sorted;
  uint32_t* perm;
  int2* cells;
  int64_t n_keys;
  int32_t *ell_count, *ell_idx;  // ELL: slot k of particle i at ell_idx[k * n + i]
  int ell_width;
  int32_t* csr_count;
  int64_t* csr_off;
  int32_t* csr_idx;
  uint32_t* hist;  // the sort's histogram, cleared by the permutation kernel
  uint64_t* tile_first;  // key of every kResortTile-th sorted slot (with the ranges)
  // re-sort counters per kResortTile tile, then per chunk of kChunkTiles tiles
  // (at tiles + c), zeroed with the ranges: movers by previous index (low 32
  // bits) and by owner tile (high 32 bits)
  unsigned long long* tile_count;
  // Morton blocks (square groups of 2^block_bits cells per side, a contiguous key
  // range): particle range per block, -1 when empty
  int block_bits;   // 0 when the order has no block structure (row-major)
  int2* blocks;
  int64_t n_blocks;
  // Narrow steps (cell keys below 2^32, inside a sph_step call): the sorted
  // keys, the next keys and the merge's second buffer as u32, and the
  // integrated state T packed per particle (x, y, vx, vy). When narrow is set
  // the step's kernels (forces, re-sort, ranges) use these in place of
  // keys_sorted / the new keys / the merge's u64 buffer and of tx, ty, tvx,
  // tvy; k_pack fills nkeys from keys_sorted on entry and k_unpack writes it
  // back on exit.
  bool narrow;
  uint32_t *nkeys, *nkeys_in, *nkeys_alt;
  float4* t4;
};

enum { NB_STENCIL = 0, NB_CSR = 1, NB_ELL = 2 };

constexpr int kResortStateWords = 8;

// Buffers of the incremental re-sort (sph_resort.hip).
struct ResortBuffers {
  unsigned long long* flags;       // mover bit per particle, one word per wave
  // [tiles][kBucketCap] pairs: (key lo, key hi, previous index, pid), (x, y, vx, vy)
  uint4* bucket;
  // Per step (k_ranges copies the first two to the last two and resets them
  // for the next step): the movers, one add per k_bucket workgroup; the path
  // (1: the merge windows come from the buckets, 0: sorted -- set by the
  // workgroup that wins the claim word when a bucket overflows); the claim
  // word; the last step's mover count and path, for the host. One allocation
  // of kResortStateWords words.
  int* n_movers;
  uint32_t* mode;
  uint32_t* claim;
  int* n_last;
  uint32_t* mode_last;
  uint32_t* ready;   // the merge's sorted movers are complete (reset by k_ranges)
  uint32_t* status;  // device status word: bits for waits that gave up
  uint64_t* keys_alt;              // second sorted-key buffer (ping-pong)
  // sorted path (a bucket overflowed): movers in previous-index order, then
  // sorted by (key, previous index), and the movers before each 256-particle tile
  uint64_t* mkey;
  uint32_t* midx;
  uint64_t* skey;
  uint32_t* sidx;
  uint64_t* ktmp;                  // sort scratch
  uint32_t* vtmp;
  uint32_t* tile_base;
};

// Where the producer of the new keys reports the movers: one ballot word per
// 64 particles of the previous order (flags == nullptr: report nothing). With
// bucket set the producer also puts every mover into its owner tile's bucket
// (bucket_movers), the work of k_bucket, so the re-sort needs no bucketing
// launch; the other members are then the re-sort's tables of the previous
// order.
struct MoverSink {
  unsigned long long* flags = nullptr;
  uint4* bucket = nullptr;
  unsigned long long* tcount = nullptr;  // packed tile counters, then chunk counters
  const uint64_t* tfirst = nullptr;      // first key of each tile
  const int2* cells = nullptr;           // cell ranges of the previous order
  const uint32_t* pid = nullptr;         // ids in the previous order
  uint32_t* claim = nullptr;             // set when a bucket overflows
  uint64_t n_keys = 0;
  int tiles = 0;
};

inline MoverSink mover_sink(const ResortBuffers& r) {
  MoverSink m;
  m.flags = r.flags;
  return m;
}

// The report and the bucketing in the producer of the new keys.
inline MoverSink bucket_sink(const ResortBuffers& r, const DevState& s, int tiles) {
  MoverSink m = mover_sink(r);
  m.bucket = r.bucket;
  m.tcount = s.tile_count;
  m.tfirst = s.tile_first;
  m.cells = s.cells;
  m.pid = s.pid;
  m.claim = r.claim;
  m.n_keys = (uint64_t)s.n_keys;
  m.tiles = tiles;
  return m;
}

// One wave's part of the report, called by all its active lanes (the wave's
// particles are the 64 indices from a multiple of 64, its first lane active).
__device__ __forceinline__ void record_movers(const MoverSink& m, int i, bool moved) {
  const unsigned long long bal = __ballot(moved);
  if ((threadIdx.x & (kWave - 1)) == 0) m.flags[i / kWave] = bal;
}

// (key, index) < (tile t's first key, its first index)
__device__ __forceinline__ bool tile_before(const MoverSink& m, int t, uint64_t k, uint32_t i) {
  const uint64_t tk = m.tfirst[t];
  return tk < k || (tk == k && (uint32_t)t * (uint32_t)kResortTile < i);
}

// Owner tile of a mover (key k, previous index i) whose new cell held no
// particle: the last tile whose first (key, index) is ordered before it, or
// tile 0. With all 64 lanes active the wave searches together: a sample of
// every S-th tile's first key (S = tiles / 256 rounded up; four per lane, one
// round of loads, counted by ballots), then the tiles between two samples
// (64 per round). Otherwise the calling lane alone bisects the whole table.
__device__ __forceinline__ uint32_t owner_of_empty(const MoverSink& m, bool full, uint64_t k,
                                                   uint32_t i) {
  const int lane = threadIdx.x & (kWave - 1);
  if (!full) {
    int lo = 0, hi = m.tiles;  // tiles [0, lo) are before the mover
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (tile_before(m, mid, k, i)) lo = mid + 1;
      else hi = mid;
    }
    return lo > 0 ? (uint32_t)(lo - 1) : 0u;
  }
  const int S = (m.tiles + 4 * kWave - 1) / (4 * kWave);
  bool lt[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int t = (q * kWave + lane) * S;
    lt[q] = t < m.tiles && tile_before(m, min(t, m.tiles - 1), k, i);
  }
  int c1 = 0;  // samples before the mover (a prefix of the sampled tiles)
#pragma unroll
  for (int q = 0; q < 4; ++q) c1 += __popcll(__ballot(lt[q]));
  if (c1 == 0) return 0u;
  const int base = (c1 - 1) * S, end = min(base + S, m.tiles);
  int c2 = 0;  // tiles after the sample that are before the mover
  for (int t0 = base + 1; t0 < end; t0 += kWave) {
    const int t = t0 + lane;
    const bool b = t < end && tile_before(m, min(t, m.tiles - 1), k, i);
    c2 += __popcll(__ballot(b));
  }
  return (uint32_t)(base + c2);
}

// First keys of a tile of the previous order and of the next tile (all ones
// past the last tile). A mover of that tile whose (key, index) lies between
// the two is owned by its own tile. Wave-uniform: the producer reads them with
// scalar loads at its start, so they are at hand when its movers are known.
struct TileKeys {
  uint64_t first = 0, next = 0;
};

__device__ __forceinline__ TileKeys tile_keys(const MoverSink& m, int t) {
  TileKeys b;
  b.first = m.tfirst[t];
  b.next = t + 1 < m.tiles ? m.tfirst[t + 1] : ~0ull;
  return b;
}

// The bucketing of k_bucket done by the wave that found its movers: the count
// of the wave's movers by previous index (low half of its tile's counter),
// then each mover's owner tile -- the
// tile of the last previous-order particle ordered before it by (key,
// previous index): from the previous range [s, e) of its new cell k, s - 1 if
// its index i < s (its old key was smaller) and e - 1 otherwise, or by
// owner_of_empty when that cell was empty -- one 64-bit atomic on the owner's
// counter (high half) for the mover's slot in the owner's bucket, and the
// entry (key, previous index, pid, integrated state). A bucket that is full
// sets the claim word: the step then orders the movers in k_chunks. Only the
// tile counters are added to here: the chunk counters and the step's mover
// count are sums that k_chunks and the merge form afterwards. (Adds to the
// same word from every wave serialise at the memory: the 64 chunk counters
// of a 1M set took 20-25 us per step that way.)
// Called by all active lanes of a wave; st is the lane's integrated state.
// ALIGNED: the wave's particles are the 64 indices from a multiple of 64 (one
// tile, whose first keys are tk), else consecutive indices from anywhere (the
// batched scenes' segments: each mover adds itself to its tile). In the
// aligned form a mover between tk.first and tk.next (about nine in ten: most
// movers step to a cell the tile already spans) is owned by its own tile,
// known without its cell range: lane 0 adds the wave's count and takes the
// slots of all such movers in one returning atomic, so a wave whose movers
// are all of this kind waits for one memory round (that atomic, with the
// pids) instead of two (the cell ranges, then the atomics).
template <bool ALIGNED = true>
__device__ __forceinline__ void bucket_movers(const MoverSink& m, int i, uint64_t k,
                                              bool moved, float4 st, TileKeys tk = TileKeys{}) {
  const int lane = threadIdx.x & (kWave - 1);
  const unsigned long long bal = __ballot(moved);
  if (bal == 0ull) return;
  const unsigned long long active = __ballot(true);
  const int tile = i / kResortTile;
  // (tk.first, first index of the tile) < (k, i) and (k, i) < (tk.next, first
  // index of the next tile); i lies in the tile, so the second is k <= tk.next
  const bool home = ALIGNED && moved &&
                    (tk.first < k || (tk.first == k && (uint32_t)tile * kResortTile < (uint32_t)i)) &&
                    k <= tk.next;
  const unsigned long long hbal = __ballot(home);
  int2 cr = make_int2(0, 0);
  uint32_t pid = 0u;
  if (moved) {
    pid = m.pid[i];
    if (!home) cr = m.cells[min(k, m.n_keys - 1)];
  }
  uint32_t hbase = 0u;
  if (ALIGNED) {
    const unsigned long long add =
        (unsigned long long)__popcll(bal) | ((unsigned long long)__popcll(hbal) << 32);
    if (hbal != 0ull) {  // wave-uniform
      unsigned long long old = 0ull;
      if (lane == 0) old = atomicAdd(&m.tcount[tile], add);
      hbase = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(old >> 32), 0);
    } else if (lane == 0) {
      atomicAdd(&m.tcount[tile], add);
    }
  } else if (moved) {
    atomicAdd(&m.tcount[tile], 1ull);
  }
  uint32_t own = (uint32_t)tile;
  if (moved && !home && cr.x >= 0) {
    const int p = i < cr.x ? cr.x - 1 : cr.y - 1;
    own = p < 0 ? 0u : (uint32_t)p / (uint32_t)kResortTile;
  }
  unsigned long long empty = __ballot(moved && !home && cr.x < 0);
  while (empty != 0ull) {  // wave-uniform
    const int l = __ffsll((long long)empty) - 1;
    empty &= empty - 1ull;
    const uint64_t kl = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(k >> 32), l)
                         << 32) |
                        (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)k, l);
    const uint32_t il = (uint32_t)__builtin_amdgcn_readlane(i, l);
    if (active == ~0ull) {
      const uint32_t o = owner_of_empty(m, true, kl, il);
      if (lane == l) own = o;
    } else if (lane == l) {
      own = owner_of_empty(m, false, kl, il);
    }
  }
  if (!moved) return;
  constexpr unsigned long long kOwnOne = 1ull << 32;
  const uint32_t slot = home ? hbase + (uint32_t)__popcll(hbal & ((1ull << lane) - 1ull))
                             : (uint32_t)(atomicAdd(&m.tcount[own], kOwnOne) >> 32);
  if (slot < (uint32_t)kBucketCap) {
    uint4* e = m.bucket + 2 * ((size_t)own * kBucketCap + slot);
    e[0] = make_uint4((uint32_t)k, (uint32_t)(k >> 32), (uint32_t)i, pid);
    e[1] = make_uint4(__float_as_uint(st.x), __float_as_uint(st.y), __float_as_uint(st.z),
                      __float_as_uint(st.w));
  } else {
    *m.claim = 1u;
  }
}

struct Launch {
  hipStream_t stream;
  DevGrid grid;
  DevPhys phys;
};

hipError_t launch_fill_cells(const Launch& L, int2* cells, int64_t n_keys);
// keys of S (standalone stage); clears the ranges left by the previous sort
hipError_t launch_keys(const Launch& L, const DevState& s, uint64_t* keys_out);
// gather src -> dst through perm, write cell ranges, clear the sort histogram
hipError_t launch_gather(const Launch& L, const DevState& s, const float* src_rx,
                         const float* src_ry, const float* src_vx, const float* src_vy,
                         const uint32_t* src_pid, const float* src_ax, const float* src_ay,
                         float* dst_rx, float* dst_ry, float* dst_vx, float* dst_vy,
                         uint32_t* dst_pid, float* dst_ax, float* dst_ay, int hist_words);
// packed (NB_ELL only): neighbor positions from s.xyuv, and (rho, pr) written
// to s.rp as well
hipError_t launch_density(const Launch& L, const DevState& s, int nb_mode, bool write_p,
                          bool packed = false);
// ax, ay, rho, pr, p of the previous order into out5[0..4] through s.perm (p
// gathered from s.pres when gather_p, else recomputed from rho)
hipError_t launch_settle_fields(const Launch& L, const DevState& s, bool gather_p,
                                float* const* out5);
// s.xyuv from the sorted state, and back
hipError_t launch_pack(const Launch& L, const DevState& s);
hipError_t launch_unpack(const Launch& L, const DevState& s);
// fused: also integrate into T and write the next keys; movers (fused only,
// flags may be null): the particles whose next key differs from s.keys_sorted
// (the input of the incremental re-sort), as flag words
hipError_t launch_forces(const Launch& L, const DevState& s, int nb_mode, bool fused,
                         uint64_t* keys_out, MoverSink movers = MoverSink{},
                         bool packed = false);
hipError_t launch_integrate(const Launch& L, const DevState& s);
// density + forces + integration + next keys over LDS-staged Morton blocks
// (variant: enum sph_step_kernel); stats (may be null): 3 counters per kernel;
// part selects the kernels of the split variant: 0 both, 1 density, 2 force
hipError_t launch_block_step(const Launch& L, const DevState& s, uint64_t* keys_out, int grid,
                             int variant, unsigned long long* stats, int part = 0);
int block_step_grid(int64_t n_blocks, int cus);
// incremental re-sort: flags -> buckets (or sorted movers) -> merge -> ranges
int64_t resort_flag_words(int n);
int64_t resort_tiles(int n);
// words of DevState::tile_count for n particles: the tiles, then the chunks
int64_t resort_count_words(int n);
// movers into their owner tiles' buckets with the tile and chunk counters
// (k_bucket; when a bucket overflowed the workgroup that claimed the step sorts
// the movers into skey / sidx instead); k_mover_flags first when the producer
// of knew wrote no flag words. bucketed: the producer bucketed the movers
// itself (bucket_sink) and added only to the tile counters; k_chunks sums them
// into the chunk counters and runs the sorted path when a bucket overflowed
hipError_t launch_movers(const Launch& L, const DevState& s, const ResortBuffers& r,
                         const uint64_t* knew, int key_bits, bool flags_ready,
                         bool bucketed = false);
// soa: also write the separate arrays rx, ry, vx, vy (else only s.xyuv);
// perm: also write the permutation s.perm (read only through the API).
// knew (the new keys, key_bits significant) given: the producer bucketed the
// movers and added only to the tile counters (bucket_movers), and no
// k_bucket / k_chunks ran: the merge forms the chunk counters itself (each
// published by one workgroup, waited for by the later ones) and, when a
// bucket overflowed, its workgroup 0 orders the movers for the sorted path
// while the others wait
hipError_t launch_merge(const Launch& L, const DevState& s, const ResortBuffers& r,
                        uint64_t* keys_dst, bool soa = true, bool perm = true,
                        const uint64_t* knew = nullptr, int key_bits = 0);
// cell and block ranges of the sorted keys s.keys_sorted; clears the histogram
// and resets the re-sort's step words in r for the next step
hipError_t launch_ranges(const Launch& L, const DevState& s, const ResortBuffers& r,
                         int hist_words);
hipError_t launch_update_velocities(const Launch& L, const DevState& s);
hipError_t launch_update_positions(const Launch& L, const DevState& s);
hipError_t launch_csr_count(const Launch& L, const DevState& s, bool self);
hipError_t launch_csr_fill(const Launch& L, const DevState& s, bool self);
// exclusive scan of csr_count into csr_off[0..n]; scratch >= scan_scratch_words(n)
int64_t scan_scratch_words(int n);
hipError_t launch_scan(const Launch& L, const int32_t* in, int64_t* out, int n,
                       int64_t* scratch);
// FP64 diagnostics into out5 (device), partials >= measure_partial_words(n)
int64_t measure_partial_words(int n);
hipError_t launch_measure(const Launch& L, const DevState& s, double mass, double gravity_y,
                          double* partials, double* out5);
hipError_t launch_ell_overflow(const Launch& L, const DevState& s, int64_t* out_count);
// FP64 max |v|^2 and max |a|^2 of S into out2 (device); partials >= measure_partial_words(n)
hipError_t launch_max_va(const Launch& L, const DevState& s, double* partials, double* out2);

// ------------------------------------------------------------ batched scenes
// Per-scene constants of a batch: the scene's grid and physics, and the offset
// of its cell table in the shared key space (keys of scene s lie in
// [base, base + its table size), so one global sort keeps every scene's
// particles in one contiguous, fixed segment of the sorted order).
struct SceneDev {
  DevGrid g;
  DevPhys c;
  uint64_t base;
};

// One workgroup's particles: [begin, end) of one scene's segment.
struct TileDev {
  int32_t scene, begin, end, pad;
};

struct BatchLaunch {
  const SceneDev* scenes;
  const TileDev* tiles;
  int n_tiles;
  int tile;  // workgroup size: 64, 128, 256 or 512
};

constexpr int kMaxTile = 512;

// keys of S (scene base + cell key in the scene's grid)
hipError_t launch_batch_keys(const Launch& L, const BatchLaunch& B, const DevState& s,
                             uint64_t* keys_out);
// density + EOS + fixed-width list of every scene
hipError_t launch_batch_density(const Launch& L, const BatchLaunch& B, const DevState& s);
// forces + integration into T + next keys of every scene; movers (flags may be
// null): the particles whose next key differs from s.keys_sorted, ORed into
// flag words that are zero on entry (the merge zeroes them after reading)
hipError_t launch_batch_forces(const Launch& L, const BatchLaunch& B, const DevState& s,
                               uint64_t* keys_out, MoverSink movers = MoverSink{});

}  // namespace sphamd
