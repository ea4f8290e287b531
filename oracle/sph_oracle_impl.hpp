// CPU restatement of the reference's 2D SPH hot path -- TEST INFRASTRUCTURE ONLY.
//
// Shared by the strict checker build (sph_oracle.cpp, FP32 + FP64, no contraction,
// no fast-math) and the timed CPU baseline (cpu_baseline.cpp, the reference's
// -O3 -ffast-math flags). Every function cites the reference code it restates;
// paths are relative to the reference repository root.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "sph_oracle.h"

namespace orc {

// ---------------------------------------------------------------- grid + keys
// MortonOrder::MortonCode2d (core/alg/orders/MortonOrder.h:117-126): x bits go to
// even positions, y bits to odd positions. Widened to 32 bits per axis in a
// 64-bit key; for codes below 2^32 this equals the reference's int code.
inline uint64_t morton_loop(uint32_t x, uint32_t y) {
  uint64_t z = 0;
  for (int i = 0; i < 32; ++i) {
    z |= (uint64_t)((x >> i) & 1u) << (2 * i);
    z |= (uint64_t)((y >> i) & 1u) << (2 * i + 1);
  }
  return z;
}

// Magic-bits form of cuMortonOrder::Partition1By1 (core/acceleration/order/
// cuMortonOrder.cuh:44-53), widened to 64 bits.
inline uint64_t part1by1(uint64_t x) {
  x &= 0xffffffffull;
  x = (x | (x << 16)) & 0x0000ffff0000ffffull;
  x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
  x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}
inline uint64_t compact1by1(uint64_t x) {  // cuMortonOrder.cuh:55-63
  x &= 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0f0f0f0f0f0f0f0full;
  x = (x | (x >> 4)) & 0x00ff00ff00ff00ffull;
  x = (x | (x >> 8)) & 0x0000ffff0000ffffull;
  x = (x | (x >> 16)) & 0x00000000ffffffffull;
  return x;
}
inline uint64_t morton_magic(uint32_t x, uint32_t y) { return part1by1(x) | (part1by1(y) << 1); }

inline int next_pow2(int n) {  // MortonOrder::NexPowerOfTwo (MortonOrder.h:82-90)
  int p = 1;
  while (p < n) p *= 2;
  return p;
}

// Grid set-up: MortonOrder::Initialize / AddCellPadding (MortonOrder.h:15-47,
// 92-103) and RowMajorOrder::Initialize (RowMajorOrder.h:13-25). The cell count
// per axis is ceil(L / cs) evaluated in Real (float) as the reference does. The
// periodic variant (not in the reference) uses floor(L / h) cells of edge L / n.
inline int make_grid(const orc_params* p, orc_grid* g) {
  float lo_x = (float)p->lo[0], lo_y = (float)p->lo[1];
  float len_x = (float)p->hi[0] - lo_x, len_y = (float)p->hi[1] - lo_y;
  float cs = (float)(p->cell_size > 0 ? p->cell_size : p->h);
  if (!(cs > 0) || !(len_x > 0) || !(len_y > 0)) return -1;
  if (p->cell_size > 0 && p->cell_size < p->h) return -4;  // stencil must cover h
  int nx, ny;
  float csx = cs, csy = cs;
  if (p->boundary == 1) {
    nx = (int)std::floor(len_x / cs);
    ny = (int)std::floor(len_y / cs);
    if (nx < 3 || ny < 3) return -2;
    csx = len_x / (float)nx;
    csy = len_y / (float)ny;
  } else {
    nx = (int)std::ceil(len_x / cs);
    ny = (int)std::ceil(len_y / cs);
  }
  if (nx < 1 || ny < 1) return -3;
  g->cs_x = csx; g->cs_y = csy;
  g->lo_x = lo_x; g->lo_y = lo_y;
  g->len_x = len_x; g->len_y = len_y;
  g->nx = nx; g->ny = ny;
  g->order = p->order;
  if (p->order == 0) {
    g->nxp = next_pow2(nx);
    g->nyp = next_pow2(ny);
    // Table sized by the largest code (the reference sizes it nxp*nyp, which
    // overflows once the padded aspect ratio exceeds 2).
    g->n_keys = (int64_t)morton_loop((uint32_t)(g->nxp - 1), (uint32_t)(g->nyp - 1)) + 1;
  } else {
    g->nxp = nx; g->nyp = ny;
    g->n_keys = (int64_t)nx * (int64_t)ny;
  }
  int bits = 0;
  while (bits < 64 && ((uint64_t)1 << bits) < (uint64_t)g->n_keys) ++bits;
  g->key_bits = bits < 1 ? 1 : bits;
  return 0;
}

// Cell coordinate of GenericNeighborSearch::SetParticleCellIds
// (core/alg/GenericNeighborSearch.h:212-222): floor(x / cs) clamped above at
// n-1. Canonical additions: measured from the lower corner, clamped below at 0,
// NaN -> 0. IEEE division, no contraction.
template <class R>
inline int cell_coord(R x, R lo, R cs, int n) {
  R f = std::floor((x - lo) / cs);
  if (f >= (R)n) return n - 1;
  if (f >= (R)0) return (int)f;
  return 0;
}

inline uint64_t make_key(const orc_grid& g, int cx, int cy) {
  if (g.order == 0) return morton_loop((uint32_t)cx, (uint32_t)cy);
  return (uint64_t)cx + (uint64_t)cy * (uint64_t)g.nx;  // RowMajorOrder.h:54-56
}

// --------------------------------------------------------------- constants
template <class R>
struct Consts {
  R h, h2, tol, mc_rho, c_pres, c_visc, gx, gy, eos_k, rho0, inv_rho0, tait_b;
  R neg_alpha_c0, eps_h2, dt, lo_x, lo_y, hi_x, hi_y, len_x, len_y, half_x, half_y;
  R cs_x, cs_y;
  int physics, boundary, gamma;
};

// Kernel constants. Reference physics: poly6 315/(64 pi h^9), spiky 15/(pi h^6),
// viscosity 45 mu/(pi h^6) (core/Kernel.cpp:8-10,38-40; cuParticles.cuh:48-51;
// GenericForceComputation.tpp:36-38). WCSPH physics: the same kernel shapes
// normalised in 2D (poly6 4/(pi h^8), spiky gradient 30/(pi h^5)), Tait EOS.
// Powers of h are formed by products in double, then rounded once.
template <class R>
inline Consts<R> make_consts(const orc_params* p, const orc_grid& g) {
  Consts<R> c{};
  const double PI = 3.14159265358979323846;
  double h = p->h, h2 = h * h, h4 = h2 * h2, h5 = h4 * h, h6 = h4 * h2, h8 = h4 * h4,
         h9 = h8 * h;
  c.h = (R)h;
  c.h2 = c.h * c.h;
  c.tol = (R)1.0e-12;  // TOL, core/ConfigVariables.hpp:12
  c.physics = p->physics;
  c.boundary = p->boundary;
  c.gamma = p->gamma;
  c.gx = (R)0;
  c.gy = (R)p->gravity_y;
  c.dt = (R)p->dt;
  c.rho0 = (R)p->rho0;
  c.eos_k = (R)p->eos_k;
  if (p->physics == 0) {
    double c_rho = 315.0 / (64.0 * PI * h9);
    double cp = 15.0 / (PI * h6);
    c.mc_rho = (R)(p->mass * c_rho);
    // fp += -3 cp (h-r)^2/r d * (-1/2)(P_i+P_j) m  (GenericForceComputation.tpp:246-252)
    c.c_pres = (R)(1.5 * cp * p->mass);
    c.c_visc = (R)(45.0 * p->mu / (PI * h6));  // cuParticles.cuh:51
    c.inv_rho0 = (R)0;
    c.tait_b = (R)0;
  } else {
    double c_rho = 4.0 / (PI * h8);
    double cg = 30.0 / (PI * h5);
    c.mc_rho = (R)(p->mass * c_rho);
    c.c_pres = (R)(p->mass * cg);
    c.c_visc = (R)0;
    c.inv_rho0 = (R)(1.0 / p->rho0);
    c.tait_b = (R)(p->rho0 * p->c0 * p->c0 / (double)p->gamma);
  }
  c.neg_alpha_c0 = (R)(-p->alpha * p->c0);
  c.eps_h2 = (R)(p->eps_visc * h2);
  c.lo_x = (R)p->lo[0];
  c.lo_y = (R)p->lo[1];
  c.hi_x = (R)p->hi[0];
  c.hi_y = (R)p->hi[1];
  c.len_x = c.hi_x - c.lo_x;
  c.len_y = c.hi_y - c.lo_y;
  c.half_x = c.len_x * (R)0.5;
  c.half_y = c.len_y * (R)0.5;
  if (sizeof(R) == sizeof(float)) {
    c.cs_x = (R)g.cs_x;
    c.cs_y = (R)g.cs_y;
  } else {
    double cs = p->cell_size > 0 ? p->cell_size : p->h;
    c.cs_x = p->boundary == 1 ? (R)(c.len_x / (R)g.nx) : (R)cs;
    c.cs_y = p->boundary == 1 ? (R)(c.len_y / (R)g.ny) : (R)cs;
  }
  return c;
}

// Integer power by binary exponentiation (fixed evaluation order).
template <class R>
inline R ipow(R x, int e) {
  R r = (R)1, b = x;
  while (e > 0) {
    if (e & 1) r = r * b;
    e >>= 1;
    if (e) b = b * b;
  }
  return r;
}

// Test hook: a nonzero ORC_MUTANT plants one known mistake in the restatement so
// that tests/test_oracle.py can show its closed-form checks catch it (1: Monaghan
// gate flipped, 2: mean density without the 1/2, 3: +alpha instead of -alpha,
// 4: minimum image wraps one side only, 5: minimum image wraps the wrong way).
// The checker and the CPU baseline are always built with 0.
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

// Minimum-image separation for periodic boxes (no counterpart in the reference).
template <class R>
inline R min_image(R d, R len, R half) {
#if ORC_MUTANT == 4
  if (d > half) return d - len;
  return d;
#elif ORC_MUTANT == 5
  if (d > half) return d + len;
  if (d < -half) return d - len;
  return d;
#else
  if (d > half) return d - len;
  if (d < -half) return d + len;
  return d;
#endif
}

template <class R>
inline void separation(const Consts<R>& c, R xi, R yi, R xj, R yj, R& dx, R& dy) {
  dx = xi - xj;
  dy = yi - yj;
  if (c.boundary == 1) {
    dx = min_image(dx, c.len_x, c.half_x);
    dy = min_image(dy, c.len_y, c.half_y);
  }
}

// ------------------------------------------------------------ neighbor lists
// MultipassNeighborLists::CountNeighborsPerParticle / AssignIndices
// (core/algorithmic/MultipassNeighborLists.h:63-181) with the GPU convention of
// cuCountNeighborsPerParticle / cuAssignIndices (cuNeighborLists.cuh:50-58,
// 115-127): stencil order x-offset outer, y-offset inner, ascending index within a
// cell, keep r^2 <= h^2, self excluded unless include_self.
template <class R, class F>
inline void for_each_neighbor(const orc_grid& g, const Consts<R>& c, const R* rx, const R* ry,
                              const int32_t* start, const int32_t* end, int64_t i,
                              bool include_self, F&& fn) {
  R xi = rx[i], yi = ry[i];
  int cx = cell_coord<R>(xi, c.lo_x, c.cs_x, g.nx);
  int cy = cell_coord<R>(yi, c.lo_y, c.cs_y, g.ny);
  for (int ox = -1; ox <= 1; ++ox) {
    int cxj = cx + ox;
    if (c.boundary == 1) {
      cxj = cxj < 0 ? cxj + g.nx : (cxj >= g.nx ? cxj - g.nx : cxj);
    } else if (cxj < 0 || cxj >= g.nx) {
      continue;
    }
    for (int oy = -1; oy <= 1; ++oy) {
      int cyj = cy + oy;
      if (c.boundary == 1) {
        cyj = cyj < 0 ? cyj + g.ny : (cyj >= g.ny ? cyj - g.ny : cyj);
      } else if (cyj < 0 || cyj >= g.ny) {
        continue;
      }
      uint64_t key = make_key(g, cxj, cyj);
      int32_t s = start[key], e = end[key];
      if (s < 0 || e < 0) continue;
      for (int32_t j = s; j < e; ++j) {
        if (j == i) {
          if (include_self) fn((int32_t)j, (R)0, (R)0, (R)0);
          continue;
        }
        R dx, dy;
        separation(c, xi, yi, rx[j], ry[j], dx, dy);
        R r2 = dx * dx + dy * dy;
        if (r2 > c.h2) continue;
        fn((int32_t)j, dx, dy, r2);
      }
    }
  }
}

template <class R>
inline int64_t neighbor_csr(const orc_params* p, int64_t n, const R* rx, const R* ry,
                            const int32_t* start, const int32_t* end, int include_self,
                            int32_t* counts, int64_t* offsets, int32_t* indices) {
  orc_grid g;
  if (make_grid(p, &g) != 0) return -1;
  Consts<R> c = make_consts<R>(p, g);
  for (int64_t i = 0; i < n; ++i) {
    int32_t k = 0;
    for_each_neighbor<R>(g, c, rx, ry, start, end, i, include_self != 0,
                         [&](int32_t, R, R, R) { ++k; });
    counts[i] = k;
  }
  offsets[0] = 0;
  for (int64_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + counts[i];
  if (indices) {
    for (int64_t i = 0; i < n; ++i) {
      int64_t o = offsets[i];
      for_each_neighbor<R>(g, c, rx, ry, start, end, i, include_self != 0,
                           [&](int32_t j, R, R, R) { indices[o++] = j; });
    }
  }
  return offsets[n];
}

// ------------------------------------------------------------------ physics
// Summation density over self + list (MPListsForceComputation.cpp:3-38; the
// self term first as cuApproximateDensityWithLists adds it,
// cuComputationWithLists.cu:118-122), then the EOS (reference: p = k(rho-rho0),
// GenericForceComputation.tpp:22-33; WCSPH: Tait) and pr = p / rho^2.
template <class R>
inline void density_one(const Consts<R>& c, R acc, R& rho, R& pres, R& pr) {
  rho = c.mc_rho * acc;
  if (c.physics == 0) {
    pres = c.eos_k * (rho - c.rho0);
  } else {
    R x = rho * c.inv_rho0;
    pres = c.tait_b * (ipow<R>(x, c.gamma) - (R)1);
  }
  pr = pres / (rho * rho);
}

template <class R>
inline R density_term(const Consts<R>& c, R r2) {
  R t = c.h2 - r2;
  return (t * t) * t;
}

template <class R>
inline void density_lists(const orc_params* p, int64_t n, const R* rx, const R* ry,
                          const int64_t* off, const int32_t* idx, R* rho, R* pres, R* pr) {
  orc_grid g;
  make_grid(p, &g);
  Consts<R> c = make_consts<R>(p, g);
  for (int64_t i = 0; i < n; ++i) {
    R acc = density_term<R>(c, (R)0);
    for (int64_t k = off[i]; k < off[i + 1]; ++k) {
      int32_t j = idx[k];
      R dx, dy;
      separation(c, rx[i], ry[i], rx[j], ry[j], dx, dy);
      R r2 = dx * dx + dy * dy;
      acc += density_term<R>(c, r2);
    }
    density_one<R>(c, acc, rho[i], pres[i], pr[i]);
  }
}

// Pairwise acceleration contribution of j on i.
// Reference physics (GenericForceComputation.tpp:229-265; MPListsForceComputation.cpp:382-418):
//   pressure  (r > TOL): +1.5 cp m (h-r)^2 / r (P_i + P_j) d,   d = r_i - r_j
//   viscosity          : 45 mu/(pi h^6) (h-r) / rho_j (v_j - v_i)
// WCSPH: -m (P_i + P_j + Pi_ij) grad W with the 2D spiky gradient and Monaghan's
// Pi_ij = -alpha c0 mu_ij / rho_bar for approaching pairs.
template <class R>
struct Accum {
  R fpx = 0, fpy = 0, fvx = 0, fvy = 0;
};

template <class R>
inline void force_pair(const Consts<R>& c, R dx, R dy, R r2, R rho_i, R rho_j, R pr_i, R pr_j,
                       R vxi, R vyi, R vxj, R vyj, Accum<R>& a) {
  R r = std::sqrt(r2);
  R s = c.h - r;
  if (c.physics == 0) {
    if (r > c.tol) {
      R coef = ((c.c_pres * (s * s)) / r) * (pr_i + pr_j);
      a.fpx += coef * dx;
      a.fpy += coef * dy;
    }
    R cv = (c.c_visc * s) / rho_j;
    a.fvx += cv * (vxj - vxi);
    a.fvy += cv * (vyj - vyi);
  } else {
    if (r > c.tol) {
      R dvx = vxi - vxj, dvy = vyi - vyj;
      R vr = dvx * dx + dvy * dy;
      R pij = (R)0;
#if ORC_MUTANT == 1
      if (vr > (R)0) {
#else
      if (vr < (R)0) {
#endif
        // Monaghan: -alpha c0 mu_ij / rho_ij with mu_ij = h vr / (r^2 + eps h^2),
        // rho_ij = (rho_i + rho_j) / 2, as one quotient
#if ORC_MUTANT == 2
        pij = (c.neg_alpha_c0 * (c.h * vr)) / ((r2 + c.eps_h2) * (rho_i + rho_j));
#elif ORC_MUTANT == 3
        pij = (-c.neg_alpha_c0 * (c.h * vr)) / ((r2 + c.eps_h2) * ((rho_i + rho_j) * (R)0.5));
#else
        pij = (c.neg_alpha_c0 * (c.h * vr)) / ((r2 + c.eps_h2) * ((rho_i + rho_j) * (R)0.5));
#endif
      }
      R coef = ((c.c_pres * (s * s)) / r) * ((pr_i + pr_j) + pij);
      a.fpx += coef * dx;
      a.fpy += coef * dy;
    }
  }
}

template <class R>
inline void force_finish(const Consts<R>& c, const Accum<R>& a, R& ax, R& ay) {
  ax = c.gx + a.fpx;  // f = g; f += fp; f += fv  (GenericForceComputation.tpp:193-279)
  ay = c.gy + a.fpy;
  if (c.physics == 0) {
    ax = ax + a.fvx;
    ay = ay + a.fvy;
  }
}

template <class R>
inline void forces_lists(const orc_params* p, int64_t n, const R* rx, const R* ry, const R* vx,
                         const R* vy, const R* rho, const R* pr, const int64_t* off,
                         const int32_t* idx, R* ax, R* ay) {
  orc_grid g;
  make_grid(p, &g);
  Consts<R> c = make_consts<R>(p, g);
  for (int64_t i = 0; i < n; ++i) {
    Accum<R> a;
    for (int64_t k = off[i]; k < off[i + 1]; ++k) {
      int32_t j = idx[k];
      R dx, dy;
      separation(c, rx[i], ry[i], rx[j], ry[j], dx, dy);
      R r2 = dx * dx + dy * dy;
      force_pair<R>(c, dx, dy, r2, rho[i], rho[j], pr[i], pr[j], vx[i], vy[i], vx[j], vy[j], a);
    }
    force_finish<R>(c, a, ax[i], ay[i]);
  }
}

// Symplectic Euler + boundary (Integrator.cpp:7-47; SoAContainer.cpp:109-132):
// v += dt a; r += dt v; per axis: r > hi -> r = hi, v = -v; r < lo -> r = lo, v = -v.
// Periodic boxes wrap instead (no counterpart in the reference).
template <class R>
inline void integrate_one(const Consts<R>& c, R& x, R& y, R& vx, R& vy, R ax, R ay) {
  vx = vx + c.dt * ax;
  vy = vy + c.dt * ay;
  x = x + c.dt * vx;
  y = y + c.dt * vy;
  if (c.boundary == 1) {
    if (x >= c.hi_x) x = x - c.len_x; else if (x < c.lo_x) x = x + c.len_x;
    if (y >= c.hi_y) y = y - c.len_y; else if (y < c.lo_y) y = y + c.len_y;
  } else {
    if (x > c.hi_x) { x = c.hi_x; vx = -vx; }
    if (x < c.lo_x) { x = c.lo_x; vx = -vx; }
    if (y > c.hi_y) { y = c.hi_y; vy = -vy; }
    if (y < c.lo_y) { y = c.lo_y; vy = -vy; }
  }
}

template <class R>
inline void integrate_all(const orc_params* p, int64_t n, R* rx, R* ry, R* vx, R* vy,
                          const R* ax, const R* ay) {
  orc_grid g;
  make_grid(p, &g);
  Consts<R> c = make_consts<R>(p, g);
  for (int64_t i = 0; i < n; ++i) integrate_one<R>(c, rx[i], ry[i], vx[i], vy[i], ax[i], ay[i]);
}

// ------------------------------------------------------------ sort + ranges
// Canonical permutation: stable sort by (key, current index). The reference sorts
// particle ids with an unstable std::sort by cell id
// (GenericNeighborSearch.h:69-76) and thrust::sort_by_key on the GPU
// (cuFindNeighbors.cu:82-100); stability fixes the intra-cell order.
inline void stable_sort_keys(int64_t n, const uint64_t* keys, uint32_t* perm,
                             uint64_t* sorted) {
  std::vector<uint32_t> id(n);
  std::iota(id.begin(), id.end(), 0u);
  std::stable_sort(id.begin(), id.end(),
                   [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
  for (int64_t i = 0; i < n; ++i) {
    perm[i] = id[i];
    sorted[i] = keys[id[i]];
  }
}

// GenericNeighborSearch::SetCellRanges (GenericNeighborSearch.h:92-108) and
// cuSetCellRanges (cuFindNeighbors.cu:103-122): start = first sorted index of a
// key, end = last + 1, -1 for empty cells.
inline void cell_ranges(int64_t n, const uint64_t* sk, int64_t n_keys, int32_t* start,
                        int32_t* end) {
  std::fill(start, start + n_keys, -1);
  std::fill(end, end + n_keys, -1);
  for (int64_t i = 0; i < n; ++i) {
    uint64_t k = sk[i];
    if (i == 0 || k != sk[i - 1]) start[k] = (int32_t)i;
    if (i == n - 1 || k != sk[i + 1]) end[k] = (int32_t)(i + 1);
  }
}

// -------------------------------------------------------------- pipeline
template <class R>
inline int run(const orc_params* p, int64_t n, R* rx, R* ry, R* vx, R* vy, R* ax, R* ay,
               R* rho_out, uint32_t* pid, int32_t steps) {
  orc_grid g;
  if (make_grid(p, &g) != 0) return -1;
  Consts<R> c = make_consts<R>(p, g);
  std::vector<uint64_t> keys(n), skeys(n);
  std::vector<uint32_t> perm(n);
  std::vector<int32_t> start(g.n_keys), end(g.n_keys), counts(n);
  std::vector<int64_t> off(n + 1);
  std::vector<int32_t> idx;
  std::vector<R> tmp(n), rho(n), pres(n), pr(n);
  std::vector<uint32_t> tpid(n);
  for (int it = 0; it < steps; ++it) {
    if (it > 0) integrate_all<R>(p, n, rx, ry, vx, vy, ax, ay);
    for (int64_t i = 0; i < n; ++i)
      keys[i] = make_key(g, cell_coord<R>(rx[i], c.lo_x, c.cs_x, g.nx),
                         cell_coord<R>(ry[i], c.lo_y, c.cs_y, g.ny));
    stable_sort_keys(n, keys.data(), perm.data(), skeys.data());
    R* arrs[6] = {rx, ry, vx, vy, ax, ay};
    for (R* a : arrs) {
      for (int64_t i = 0; i < n; ++i) tmp[i] = a[perm[i]];
      std::copy(tmp.begin(), tmp.end(), a);
    }
    for (int64_t i = 0; i < n; ++i) tpid[i] = pid[perm[i]];
    std::copy(tpid.begin(), tpid.end(), pid);
    cell_ranges(n, skeys.data(), g.n_keys, start.data(), end.data());
    int64_t tot = neighbor_csr<R>(p, n, rx, ry, start.data(), end.data(), 0, counts.data(),
                                  off.data(), nullptr);
    idx.resize(tot > 0 ? tot : 1);
    neighbor_csr<R>(p, n, rx, ry, start.data(), end.data(), 0, counts.data(), off.data(),
                    idx.data());
    density_lists<R>(p, n, rx, ry, off.data(), idx.data(), rho.data(), pres.data(), pr.data());
    forces_lists<R>(p, n, rx, ry, vx, vy, rho.data(), pr.data(), off.data(), idx.data(), ax, ay);
  }
  if (rho_out) std::copy(rho.begin(), rho.end(), rho_out);
  return 0;
}

}  // namespace orc
