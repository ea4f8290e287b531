// Device-side helpers shared by the gfx950 kernels: grid/physics constants,
// 64-bit Morton keys, cell coordinates and the per-pair SPH arithmetic.
//
// Arithmetic contract: every kernel is compiled with -ffp-contract=off and IEEE
// float division / sqrt, and evaluates each expression in the order written
// here, so integer outputs (keys, neighbor sets) are bit-exact with the CPU
// definition and float outputs follow one fixed operation order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace sphamd {

constexpr int kWave = 64;
constexpr int kXcds = 8;  // accelerator dies of the MI355X, each with its own L2

// Workgroups are dealt to the XCDs round-robin (workgroup b runs on die b % 8).
// Renumbering so that die x processes a contiguous range of tiles keeps tiles
// that read each other's particles (neighbors in Morton order) on one L2. A
// bijection of [0, nb); not used by kernels whose tiles wait on predecessors.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int x = b % kXcds, k = b / kXcds;
  const int q = nb / kXcds, r = nb % kXcds;
  return x * q + min(x, r) + k;
}

// Element j of a 4-byte array whose base is wave-uniform (a kernel argument),
// addressed by a 32-bit byte offset so that the load takes the base in scalar
// registers and only the offset in a vector register. Contexts hold fewer than
// 2^30 particles, so 4 j fits in 32 bits.
template <class T>
__device__ __forceinline__ T ld4(const T* __restrict__ base, int j) {
  static_assert(sizeof(T) == 4, "4-byte elements");
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                     ((uint32_t)j << 2));
}

// 8 and 16 bytes of element j of a packed array (32-bit byte offsets, as ld4).
__device__ __forceinline__ float2 ld8(const float2* __restrict__ base, int j) {
  return *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(base) +
                                          ((uint32_t)j << 3));
}
__device__ __forceinline__ float2 ld8_of16(const float4* __restrict__ base, int j) {
  return *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(base) +
                                          ((uint32_t)j << 4));
}
__device__ __forceinline__ float4 ld16(const float4* __restrict__ base, int j) {
  return *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(base) +
                                          ((uint32_t)j << 4));
}

// Pins loaded values at this point. A value used only under a lane predicate
// would otherwise have its load moved into that branch, where it is waited for
// before the branch closes -- one load in flight at a time. After a batch of
// loads, pinning them keeps the whole batch in flight together.
__device__ __forceinline__ void pin(float v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void pin(int v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void pin(unsigned v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void pin(unsigned long long v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void pin(unsigned long v) { asm volatile("" ::"v"(v)); }

struct DevGrid {
  float lo_x, lo_y, cs_x, cs_y;
  int32_t nx, ny;
  int32_t order;     // 0 Morton, 1 row-major
  int32_t periodic;  // 1 -> wrapped stencil
};

struct DevPhys {
  float h, h2, tol, mc_rho, c_pres, c_visc, gx, gy;
  float eos_k, rho0, inv_rho0, tait_b, neg_alpha_c0, eps_h2, dt;
  float lo_x, lo_y, hi_x, hi_y, len_x, len_y, half_x, half_y;
  int32_t physics, boundary, gamma;
};

// Spread the low 32 bits of v to the even bit positions of a 64-bit word
// (Morton dilation; x -> even bits, y -> odd bits as in MortonOrder.h:117-126).
// Bits 0..15 of v to the even bits of a 32-bit word.
__device__ __forceinline__ uint32_t dilate16(uint32_t v) {
  uint32_t x = v & 0xffffu;
  x = (x | (x << 8)) & 0x00ff00ffu;
  x = (x | (x << 4)) & 0x0f0f0f0fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}
__device__ __forceinline__ uint64_t dilate(uint32_t v) {
  return (uint64_t)dilate16(v) | ((uint64_t)dilate16(v >> 16) << 32);
}

__device__ __forceinline__ uint64_t cell_key(const DevGrid& g, int cx, int cy) {
  if (g.order == 0) return dilate((uint32_t)cx) | (dilate((uint32_t)cy) << 1);
  return (uint64_t)cx + (uint64_t)cy * (uint64_t)g.nx;
}

// floor((x - lo) / cs) clamped to [0, n-1]; NaN -> 0.
__device__ __forceinline__ int cell_coord(float x, float lo, float cs, int n) {
  float f = floorf((x - lo) / cs);
  if (f >= (float)n) return n - 1;
  if (f >= 0.0f) return (int)f;
  return 0;
}

__device__ __forceinline__ float min_image(float d, float len, float half) {
  if (d > half) return d - len;
  if (d < -half) return d + len;
  return d;
}

template <int BC>
__device__ __forceinline__ void separation(const DevPhys& c, float xi, float yi, float xj,
                                           float yj, float& dx, float& dy) {
  dx = xi - xj;
  dy = yi - yj;
  if (BC == 1) {
    dx = min_image(dx, c.len_x, c.half_x);
    dy = min_image(dy, c.len_y, c.half_y);
  }
}

__device__ __forceinline__ float density_term(float h2, float r2) {
  float t = h2 - r2;
  return (t * t) * t;
}

__device__ __forceinline__ float ipowf(float x, int e) {
  float r = 1.0f, b = x;
  while (e > 0) {
    if (e & 1) r = r * b;
    e >>= 1;
    if (e) b = b * b;
  }
  return r;
}

// p from the EOS: linear (reference) or Tait.
template <int PHYS>
__device__ __forceinline__ float eos_pressure(const DevPhys& c, float rho) {
  if (PHYS == 0) return c.eos_k * (rho - c.rho0);
  const float x = rho * c.inv_rho0;
  return c.tait_b * (ipowf(x, c.gamma) - 1.0f);
}

// rho from the kernel sum; p from the EOS; pr = p / rho^2.
template <int PHYS>
__device__ __forceinline__ void eos(const DevPhys& c, float acc, float& rho, float& p,
                                   float& pr) {
  rho = c.mc_rho * acc;
  p = eos_pressure<PHYS>(c, rho);
  pr = p / (rho * rho);
}

struct Acc {
  float fpx, fpy, fvx, fvy;
};

// Contribution of neighbor j to the acceleration of i (d = r_i - r_j).
template <int PHYS>
__device__ __forceinline__ void force_pair(const DevPhys& c, float dx, float dy, float r2,
                                           float rho_i, float rho_j, float pr_i, float pr_j,
                                           float vxi, float vyi, float vxj, float vyj, Acc& a) {
  float r = sqrtf(r2);
  float s = c.h - r;
  if (PHYS == 0) {
    if (r > c.tol) {
      float coef = ((c.c_pres * (s * s)) / r) * (pr_i + pr_j);
      a.fpx += coef * dx;
      a.fpy += coef * dy;
    }
    float cv = (c.c_visc * s) / rho_j;
    a.fvx += cv * (vxj - vxi);
    a.fvy += cv * (vyj - vyi);
  } else {
    if (r > c.tol) {
      float dvx = vxi - vxj, dvy = vyi - vyj;
      float vr = dvx * dx + dvy * dy;
      float pij = 0.0f;
      if (vr < 0.0f) {
        // Monaghan: -alpha c0 mu_ij / rho_ij with mu_ij = h vr / (r^2 + eps h^2),
        // rho_ij = (rho_i + rho_j) / 2, as one quotient
        pij = (c.neg_alpha_c0 * (c.h * vr)) / ((r2 + c.eps_h2) * ((rho_i + rho_j) * 0.5f));
      }
      float coef = ((c.c_pres * (s * s)) / r) * ((pr_i + pr_j) + pij);
      a.fpx += coef * dx;
      a.fpy += coef * dy;
    }
  }
}

template <int PHYS>
__device__ __forceinline__ void force_finish(const DevPhys& c, const Acc& a, float& ax,
                                             float& ay) {
  ax = c.gx + a.fpx;
  ay = c.gy + a.fpy;
  if (PHYS == 0) {
    ax = ax + a.fvx;
    ay = ay + a.fvy;
  }
}

// Symplectic Euler and the boundary treatment.
template <int BC>
__device__ __forceinline__ void integrate_one(const DevPhys& c, float& x, float& y, float& vx,
                                              float& vy, float ax, float ay) {
  vx = vx + c.dt * ax;
  vy = vy + c.dt * ay;
  x = x + c.dt * vx;
  y = y + c.dt * vy;
  if (BC == 1) {
    if (x >= c.hi_x) x = x - c.len_x; else if (x < c.lo_x) x = x + c.len_x;
    if (y >= c.hi_y) y = y - c.len_y; else if (y < c.lo_y) y = y + c.len_y;
  } else {
    if (x > c.hi_x) { x = c.hi_x; vx = -vx; }
    if (x < c.lo_x) { x = c.lo_x; vx = -vx; }
    if (y > c.hi_y) { y = c.hi_y; vy = -vy; }
    if (y < c.lo_y) { y = c.lo_y; vy = -vy; }
  }
}

// Visit the neighbors of sorted particle i in stencil order (x offset outer,
// y offset inner, ascending index in a cell), self excluded unless SELF. cells
// holds one int2 (start, end) per key, -1 for empty cells.
// pos(j) gives particle j's position (float2).
template <int BC, bool SELF = false, class P, class F>
__device__ __forceinline__ void for_each_neighbor_p(const DevGrid& g, const DevPhys& c, P&& pos,
                                                    const int2* __restrict__ cells, int n, int i,
                                                    float xi, float yi, F&& fn) {
  int cx = cell_coord(xi, g.lo_x, g.cs_x, g.nx);
  int cy = cell_coord(yi, g.lo_y, g.cs_y, g.ny);
#pragma unroll
  for (int ox = -1; ox <= 1; ++ox) {
    int cxj = cx + ox;
    if (BC == 1) {
      cxj = cxj < 0 ? cxj + g.nx : (cxj >= g.nx ? cxj - g.nx : cxj);
    } else if (cxj < 0 || cxj >= g.nx) {
      continue;
    }
#pragma unroll
    for (int oy = -1; oy <= 1; ++oy) {
      int cyj = cy + oy;
      if (BC == 1) {
        cyj = cyj < 0 ? cyj + g.ny : (cyj >= g.ny ? cyj - g.ny : cyj);
      } else if (cyj < 0 || cyj >= g.ny) {
        continue;
      }
      int2 se = cells[cell_key(g, cxj, cyj)];
      if (se.x < 0 || se.y < 0) continue;
      const int je = min(se.y, n);
      for (int j = se.x; j < je; ++j) {
        if (j == i) {
          if (SELF) fn(j, 0.0f, 0.0f, 0.0f);
          continue;
        }
        float dx, dy;
        const float2 pj = pos(j);
        separation<BC>(c, xi, yi, pj.x, pj.y, dx, dy);
        float r2 = dx * dx + dy * dy;
        if (r2 > c.h2) continue;
        fn(j, dx, dy, r2);
      }
    }
  }
}

template <int BC, bool SELF = false, class F>
__device__ __forceinline__ void for_each_neighbor(const DevGrid& g, const DevPhys& c,
                                                  const float* __restrict__ rx,
                                                  const float* __restrict__ ry,
                                                  const int2* __restrict__ cells, int n, int i,
                                                  float xi, float yi, F&& fn) {
  for_each_neighbor_p<BC, SELF>(
      g, c, [&](int j) { return make_float2(rx[j], ry[j]); }, cells, n, i, xi, yi, fn);
}

constexpr int kEllStage = 8;   // list slots per thread staged in LDS before the store
constexpr int kEllLds = kEllStage + 9;  // LDS words per thread: staged slots + 9 cell shifts
#ifndef SPH_CAND_BATCH
#define SPH_CAND_BATCH 4
#endif
#ifndef SPH_LIST_BATCH
#define SPH_LIST_BATCH 7
#endif
constexpr int kCandBatch = SPH_CAND_BATCH;  // stencil candidates loaded together (density)
#ifndef SPH_PAIR_LOADS
#define SPH_PAIR_LOADS 1  // 20-step calls 0.0668 -> 0.0664 ms, 400-step calls unchanged
#endif
constexpr bool kPairLoads = SPH_PAIR_LOADS;  // stencil rows as pair + single range loads
constexpr int kListBatch = SPH_LIST_BATCH;  // list slots loaded together (forces)
#ifndef SPH_LIST_BATCH2
#define SPH_LIST_BATCH2 4
#endif
// 2D lists at h = 1.3 dx hold about five neighbors and rarely more than
// seven: the first batch loads seven slots with the count, so one round of
// loads serves almost every wave; longer lists continue four at a time
// (first / later batch of 4/4, 4/3, 6/1, 7/1, 7/4, 8/1 slots measured on the
// 1M dam break: 7/1 and 7/4 fastest, 0.8-1.4 us per step ahead of 4/4)
constexpr int kListBatch2 = SPH_LIST_BATCH2;
static_assert(kListBatch2 <= kListBatch, "later batches reuse the first batch's registers");

// Cell ranges of the 3x3 stencil of position (x, y) in visiting order (x offset
// outer, y offset inner); cells outside a walled box give (0, 0) and empty
// cells (-1, -1) or (0, 0): a length max(end - start, 0) of zero either way.
// Issuing the nine loads together keeps them in flight at once.
template <int BC>
__device__ __forceinline__ void stencil_ranges(const DevGrid& g, const int2* __restrict__ cells,
                                               float x, float y, int2* rng) {
  const int cx = cell_coord(x, g.lo_x, g.cs_x, g.nx);
  const int cy = cell_coord(y, g.lo_y, g.cs_y, g.ny);
  if (g.order == 0 && g.nx <= 16384 && g.ny <= 16384) {
    // Morton keys below 2^28: the cell's coordinates are dilated once and the
    // neighbors' parts stepped in dilated form (x on the even bits: x + 1 is
    // ((d | odd bits) + 1) & even bits, x - 1 is (d - 1) & even bits; y the
    // same on the odd bits), and the nine loads take 32-bit byte offsets from
    // the table base. Empty cells keep (-1, -1), which the callers read as no
    // candidates; cells outside a walled box are loaded at the cell's own key
    // and masked to (0, 0).
    constexpr uint32_t kEven = 0x55555555u, kOdd = 0xAAAAAAAAu;
    const uint32_t dx = dilate16((uint32_t)cx), dy = dilate16((uint32_t)cy) << 1;
    const uint32_t dxm = (dx - 1u) & kEven, dxp = ((dx | kOdd) + 1u) & kEven;
    const uint32_t dym = (dy - 2u) & kOdd, dyp = ((dy | kEven) + 2u) & kOdd;
    const bool xlo = cx > 0, xhi = cx + 1 < g.nx, ylo = cy > 0, yhi = cy + 1 < g.ny;
    uint32_t kx[3], ky[3];
    kx[1] = dx;
    ky[1] = dy;
    if (BC == 1) {
      kx[0] = xlo ? dxm : dilate16((uint32_t)(g.nx - 1));
      kx[2] = xhi ? dxp : 0u;
      ky[0] = ylo ? dym : dilate16((uint32_t)(g.ny - 1)) << 1;
      ky[2] = yhi ? dyp : 0u;
    } else {
      kx[0] = xlo ? dxm : dx;
      kx[2] = xhi ? dxp : dx;
      ky[0] = ylo ? dym : dy;
      ky[2] = yhi ? dyp : dy;
    }
    const char* base = reinterpret_cast<const char*>(cells);
    // Cells x and x + 1 with x even are consecutive keys: each row of the
    // stencil is one 16-byte load of the even-odd pair it holds and one
    // 8-byte load of its third cell (six loads, not nine). Needs the pair's
    // odd member to be the row's neighbor: always in a walled box (a pair
    // past the last column reads a padded cell, masked below) and in a
    // periodic one with an even number of columns. Tables start 16-byte
    // aligned (a batched scene's table at an even entry).
    if (kPairLoads && g.nx >= 2 && (BC == 0 || (g.nx & 1) == 0)) {
      const bool odd = (cx & 1) != 0;
      const uint32_t pxd = odd ? dxm : dx;        // the pair's even member
      const uint32_t sxd = odd ? kx[2] : kx[0];   // the row's third cell
#pragma unroll
      for (int oy = 0; oy < 3; ++oy) {
        const int4 a = *reinterpret_cast<const int4*>(base + ((pxd + ky[oy]) << 3));
        const int2 b = *reinterpret_cast<const int2*>(base + ((sxd + ky[oy]) << 3));
        const int2 lo = make_int2(a.x, a.y), hi = make_int2(a.z, a.w);
        rng[oy] = odd ? lo : b;
        rng[3 + oy] = odd ? hi : lo;
        rng[6 + oy] = odd ? b : hi;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 9; ++q)
        rng[q] = *reinterpret_cast<const int2*>(base + ((kx[q / 3] + ky[q % 3]) << 3));
    }
    if (BC == 0) {
      const bool okx[3] = {xlo, true, xhi}, oky[3] = {ylo, true, yhi};
#pragma unroll
      for (int q = 0; q < 9; ++q)
        if (!(okx[q / 3] && oky[q % 3])) rng[q] = make_int2(0, 0);
    }
    return;
  }
  // Key = (x part) + (y part) in both orders (the Morton parts have disjoint
  // bits). Cells outside a walled box are loaded at a clamped, valid key and
  // masked afterwards, so the nine loads carry no branch between them.
  uint64_t kx[3], ky[3];
  bool okx[3], oky[3];
#pragma unroll
  for (int o = 0; o < 3; ++o) {
    int cxj = cx + o - 1, cyj = cy + o - 1;
    if (BC == 1) {
      cxj = cxj < 0 ? cxj + g.nx : (cxj >= g.nx ? cxj - g.nx : cxj);
      cyj = cyj < 0 ? cyj + g.ny : (cyj >= g.ny ? cyj - g.ny : cyj);
      okx[o] = oky[o] = true;
    } else {
      okx[o] = cxj >= 0 && cxj < g.nx;
      oky[o] = cyj >= 0 && cyj < g.ny;
      cxj = min(max(cxj, 0), g.nx - 1);
      cyj = min(max(cyj, 0), g.ny - 1);
    }
    if (g.order != 0) {
      kx[o] = (uint64_t)cxj;
      ky[o] = (uint64_t)cyj * (uint64_t)g.nx;
    } else if (g.nx <= 65536 && g.ny <= 65536) {  // coordinates fit 16 bits
      kx[o] = dilate16((uint32_t)cxj);
      ky[o] = (uint64_t)dilate16((uint32_t)cyj) << 1;
    } else {
      kx[o] = dilate((uint32_t)cxj);
      ky[o] = dilate((uint32_t)cyj) << 1;
    }
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) rng[q] = cells[kx[q / 3] + ky[q % 3]];
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const int2 se = rng[q];
    const bool ok = okx[q / 3] & oky[q % 3] & (se.x >= 0) & (se.y >= 0);
    rng[q] = ok ? se : make_int2(0, 0);
  }
}

// Kernel sum of the summation density of sorted particle i (self term first,
// then the neighbors in stencil order), writing the neighbor indices (self
// excluded) to the column-major fixed-width list: slot k of i at ell_idx[k*n+i].
// The first kEllStage slots go through stage[k * stride] (this thread's LDS
// column of kEllLds words) and are stored slot by slot at the end, so a
// wavefront's stores of one slot are consecutive addresses. Returns the sum;
// ell_count[i] gets the count (which may exceed ell_w: the force pass then walks
// the stencil instead).
template <int BC, bool PACKED = false>
__device__ __forceinline__ float density_ell(const DevGrid& g, const DevPhys& c,
                                             const float* __restrict__ rx,
                                             const float* __restrict__ ry,
                                             const int2* __restrict__ cells, int n, int i,
                                             float xi, float yi, int* stage, int stride,
                                             int32_t* __restrict__ ell_count,
                                             int32_t* __restrict__ ell_idx, int ell_w,
                                             const float4* __restrict__ xyuv = nullptr) {
  float acc = density_term(c.h2, 0.0f);
  int2 rng[9];
  stencil_ranges<BC>(g, cells, xi, yi, rng);
  // The candidates of the nine cells form one sequence in stencil order: slot t
  // of it is index t + shift[q] for the last cell q with first[q] <= t. Walking
  // it kCandBatch slots at a time lets every slot's position loads be in flight
  // together (slots past the end load particle i itself, a valid address).
  int first[9], shift[9];
  int total = 0;
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const int len = max(min(rng[q].y, n) - rng[q].x, 0);
    first[q] = total;
    shift[q] = rng[q].x - total;
    total += len;
  }
  // Cell of slot t: the number q of cells 1..8 with first[q] <= t; the slot's
  // index is t + shift[q], the shifts in this thread's LDS column. With fewer
  // than 128 candidates in every lane of the wave, q comes from a byte-wise
  // compare of t with first[1..8] packed four to a word (bit 7 of each byte
  // difference) and two popcounts; otherwise from eight compares. Slot indices
  // are computed for every slot of a batch and masked afterwards.
  int* shift_col = stage + kEllStage * stride;
#pragma unroll
  for (int q = 0; q < 9; ++q) shift_col[q * stride] = shift[q];
  const bool packed = __all(total < 128);
  const uint32_t f_lo = (uint32_t)(first[1] & 255) | ((uint32_t)(first[2] & 255) << 8) |
                        ((uint32_t)(first[3] & 255) << 16) | ((uint32_t)(first[4] & 255) << 24);
  const uint32_t f_hi = (uint32_t)(first[5] & 255) | ((uint32_t)(first[6] & 255) << 8) |
                        ((uint32_t)(first[7] & 255) << 16) | ((uint32_t)(first[8] & 255) << 24);
  const int f1 = first[1], f2 = first[2], f3 = first[3], f4 = first[4], f5 = first[5],
            f6 = first[6], f7 = first[7], f8 = first[8];
  int cnt = 0;
  auto walk = [&](auto cell_of) {
    for (int t0 = 0; t0 < total; t0 += kCandBatch) {
      int jb[kCandBatch];
      float xb[kCandBatch], yb[kCandBatch];
      int sh[kCandBatch];
#pragma unroll
      for (int u = 0; u < kCandBatch; ++u) sh[u] = shift_col[cell_of(t0 + u) * stride];
#pragma unroll
      for (int u = 0; u < kCandBatch; ++u) pin(sh[u]);
#pragma unroll
      for (int u = 0; u < kCandBatch; ++u) jb[u] = t0 + u < total ? t0 + u + sh[u] : i;
#pragma unroll
      for (int u = 0; u < kCandBatch; ++u) {
        if constexpr (PACKED) {
          const float2 q = ld8_of16(xyuv, jb[u]);
          xb[u] = q.x;
          yb[u] = q.y;
        } else {
          xb[u] = ld4(rx, jb[u]);
          yb[u] = ld4(ry, jb[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < kCandBatch; ++u) {
        pin(xb[u]);
        pin(yb[u]);
      }
#pragma unroll
      for (int u = 0; u < kCandBatch; ++u) {
        const int j = jb[u];
        if (t0 + u >= total || j == i) continue;
        float dx, dy;
        separation<BC>(c, xi, yi, xb[u], yb[u], dx, dy);
        const float r2 = dx * dx + dy * dy;
        if (r2 > c.h2) continue;
        acc += density_term(c.h2, r2);
        if (cnt < kEllStage) stage[cnt * stride] = j;
        else if (cnt < ell_w) ell_idx[(int64_t)cnt * n + i] = j;
        ++cnt;
      }
    }
  };
  if (packed) {
    walk([&](int t) {
      const uint32_t tb = ((uint32_t)t | 0x80u) * 0x01010101u;
      return __builtin_popcount((tb - f_lo) & 0x80808080u) +
             __builtin_popcount((tb - f_hi) & 0x80808080u);
    });
  } else {
    walk([&](int t) {
      return (t >= f1) + (t >= f2) + (t >= f3) + (t >= f4) + (t >= f5) + (t >= f6) + (t >= f7) +
             (t >= f8);
    });
  }
  ell_count[i] = cnt;
  const int staged = min(cnt, min(kEllStage, ell_w));
  for (int k = 0; k < staged; ++k) ell_idx[(int64_t)k * n + i] = stage[k * stride];
  return acc;
}

// Pressure (+ viscous) sum of sorted particle i over its fixed-width list, or
// over the stencil when the list overflowed.
template <int PHYS, int BC, bool PACKED = false>
__device__ __forceinline__ Acc forces_ell(const DevGrid& g, const DevPhys& c,
                                          const float* __restrict__ rx,
                                          const float* __restrict__ ry,
                                          const float* __restrict__ vx,
                                          const float* __restrict__ vy,
                                          const float* __restrict__ rho,
                                          const float* __restrict__ pr,
                                          const int2* __restrict__ cells,
                                          const int32_t* __restrict__ ell_count,
                                          const int32_t* __restrict__ ell_idx, int ell_w, int n,
                                          int i, float xi, float yi, float vxi, float vyi,
                                          float rhoi, float pri,
                                          const float4* __restrict__ xyuv = nullptr,
                                          const float2* __restrict__ rp = nullptr) {
  Acc a{0.f, 0.f, 0.f, 0.f};
  auto pair = [&](int j, float dx, float dy, float r2) {
    if constexpr (PACKED) {
      const float4 q = xyuv[j];
      const float2 r = rp[j];
      force_pair<PHYS>(c, dx, dy, r2, rhoi, r.x, pri, r.y, vxi, vyi, q.z, q.w, a);
    } else {
      force_pair<PHYS>(c, dx, dy, r2, rhoi, rho[j], pri, pr[j], vxi, vyi, vx[j], vy[j], a);
    }
  };
  // Slot loads are unconditional (clamped to the list width, so always inside
  // the buffer; slots past the count are replaced by i before any use): the
  // first batch's slots load together with the count. Then kListBatch slots at
  // a time, every neighbor load of the batch issued before the first use
  // (particle i itself for unused slots); the sums still run in slot order.
  const int cnt = ell_count[i];
  int jl[kListBatch];
#pragma unroll
  for (int u = 0; u < kListBatch; ++u) jl[u] = ell_idx[(int64_t)min(u, ell_w - 1) * n + i];
  auto batch = [&](int k0, auto size) {
    constexpr int B = decltype(size)::value;
    float xb[B], yb[B], ub[B], wb[B], db[B], pb[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int j = k0 + u < cnt ? jl[u] : i;
      if constexpr (PACKED) {
        const float4 q = ld16(xyuv, j);
        const float2 r = ld8(rp, j);
        xb[u] = q.x;
        yb[u] = q.y;
        ub[u] = q.z;
        wb[u] = q.w;
        db[u] = r.x;
        pb[u] = r.y;
      } else {
        xb[u] = ld4(rx, j);
        yb[u] = ld4(ry, j);
        ub[u] = ld4(vx, j);
        wb[u] = ld4(vy, j);
        db[u] = ld4(rho, j);
        pb[u] = ld4(pr, j);
      }
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      pin(xb[u]);
      pin(yb[u]);
      pin(ub[u]);
      pin(wb[u]);
      pin(db[u]);
      pin(pb[u]);
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      if (k0 + u >= cnt) continue;
      float dx, dy;
      separation<BC>(c, xi, yi, xb[u], yb[u], dx, dy);
      force_pair<PHYS>(c, dx, dy, dx * dx + dy * dy, rhoi, db[u], pri, pb[u], vxi, vyi, ub[u],
                       wb[u], a);
    }
  };
  if (cnt <= ell_w) {
    batch(0, std::integral_constant<int, kListBatch>{});
    for (int k0 = kListBatch; k0 < cnt; k0 += kListBatch2) {
#pragma unroll
      for (int u = 0; u < kListBatch2; ++u)
        jl[u] = ell_idx[(int64_t)min(k0 + u, ell_w - 1) * n + i];
      batch(k0, std::integral_constant<int, kListBatch2>{});
    }
  } else if constexpr (PACKED) {
    // the list overflowed: the stencil, state from the packed copies
    for_each_neighbor_p<BC>(
        g, c, [&](int j) { return make_float2(xyuv[j].x, xyuv[j].y); }, cells, n, i, xi, yi,
        pair);
  } else {
    for_each_neighbor<BC>(g, c, rx, ry, cells, n, i, xi, yi, pair);
  }
  return a;
}

}  // namespace sphamd
