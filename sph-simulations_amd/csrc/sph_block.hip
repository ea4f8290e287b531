// SPH physics over Morton blocks staged in LDS (gfx950).
//
// A Morton block is a square of 2^B x 2^B cells whose keys form one contiguous
// range, so its particles are one contiguous run of the sorted arrays. A
// workgroup takes a block and stages in LDS the particles of the block (one
// coalesced copy) plus a halo of H cells around it (cell by cell), then
// evaluates its particles from LDS only and writes results with coalesced
// stores. Neighbor visits keep the stencil order of the global kernels (x offset
// outer, y inner, ascending index), so every float is produced by the same
// operation sequence as the oracle and the per-particle kernels.
//
// Kernel modes:
//   SPLIT_DENSITY (H = 1): density + EOS of the block's particles -> rho, pr.
//   SPLIT_FORCE   (H = 1): accelerations from staged r, v, rho, pr; symplectic
//                          Euler into the unsorted twin; cell key of the new
//                          position for the next sort.
//   FUSED         (H = 2): both in one launch; the halo's ring 2 completes the
//                          density of ring-1 particles, so no rho round trip.
// A block whose staged particles exceed the LDS capacity is evaluated from
// global memory with the same arithmetic (FUSED recomputes neighbor densities;
// the split kernels read them).
//
// Work distribution: the dense range of blocks is cut into chunks of CHUNK
// consecutive blocks (spatial neighbours share halo cells in the L2 of the CU's
// XCD); chunks are dealt round robin to the persistent workgroups.
#include <hip/hip_runtime.h>

#include "sph_internal.hpp"

namespace sphamd {

namespace {

constexpr int BWG = 128;
constexpr int CHUNK = 4;

enum { FUSED = 0, SPLIT_DENSITY = 1, SPLIT_FORCE = 2 };

__device__ __forceinline__ uint32_t compact_even(uint64_t x) {
  x &= 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0f0f0f0f0f0f0f0full;
  x = (x | (x >> 4)) & 0x00ff00ff00ff00ffull;
  x = (x | (x >> 8)) & 0x0000ffff0000ffffull;
  x = (x | (x >> 16)) & 0x00000000ffffffffull;
  return (uint32_t)x;
}

__device__ __forceinline__ int wrap_mod(int v, int n) {
  v %= n;
  return v < 0 ? v + n : v;
}

// Density + EOS of sorted particle j from global memory.
template <int PHYS, int BC>
__device__ __forceinline__ void density_global(const DevGrid& g, const DevPhys& c,
                                               const float* rx, const float* ry,
                                               const int2* cells, int n, int j, float& rho,
                                               float& pr) {
  float acc = density_term(c.h2, 0.0f);
  for_each_neighbor<BC>(g, c, rx, ry, cells, n, j, rx[j], ry[j],
                        [&](int, float, float, float r2) { acc += density_term(c.h2, r2); });
  float p;
  eos<PHYS>(c, acc, rho, p, pr);
}

template <int BC>
__device__ __forceinline__ void write_integrated(const DevGrid& g, const DevPhys& c, int i,
                                                 float x, float y, float u, float v, float axi,
                                                 float ayi, float* ax_out, float* ay_out,
                                                 float* tx, float* ty, float* tvx, float* tvy,
                                                 uint64_t* keys_out) {
  ax_out[i] = axi;
  ay_out[i] = ayi;
  integrate_one<BC>(c, x, y, u, v, axi, ayi);
  tx[i] = x;
  ty[i] = y;
  tvx[i] = u;
  tvy[i] = v;
  keys_out[i] = cell_key(g, cell_coord(x, g.lo_x, g.cs_x, g.nx),
                         cell_coord(y, g.lo_y, g.cs_y, g.ny));
}

template <int PHYS, int BC, int B, int H, int MODE, int MAXP>
__global__ __launch_bounds__(BWG) void k_blocks(
    int n, int n_blocks, DevGrid g, DevPhys c, const int2* __restrict__ blocks,
    const int2* __restrict__ cells, const float* __restrict__ rx,
    const float* __restrict__ ry, const float* __restrict__ vx, const float* __restrict__ vy,
    float* __restrict__ rho_g, float* __restrict__ pr_g, float* __restrict__ ax_out,
    float* __restrict__ ay_out, float* __restrict__ tx, float* __restrict__ ty,
    float* __restrict__ tvx, float* __restrict__ tvy, uint64_t* __restrict__ keys_out,
    unsigned long long* __restrict__ stats) {
  constexpr int BS = 1 << B;
  constexpr int BOX = BS + 2 * H;
  constexpr int NBOX = BOX * BOX;
  constexpr int QPT = (NBOX + BWG - 1) / BWG;
  constexpr int WAVES = BWG / kWave;
  constexpr bool NEED_V = MODE != SPLIT_DENSITY;
  constexpr bool NEED_RHO = MODE != SPLIT_DENSITY;  // rho/pr arrays in LDS
  __shared__ float s_x[MAXP], s_y[MAXP];
  __shared__ float s_vx[NEED_V ? MAXP : 1], s_vy[NEED_V ? MAXP : 1];
  __shared__ float s_rho[NEED_RHO ? MAXP : 1], s_pr[NEED_RHO ? MAXP : 1];
  __shared__ int s_gi[MAXP];
  __shared__ short s_q[MAXP];
  __shared__ int s_off[NBOX], s_cnt[NBOX];
  __shared__ int s_wsum[WAVES];

  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  unsigned long long st_blocks = 0, st_fallback = 0, st_staged = 0;
  const int n_chunks = (n_blocks + CHUNK - 1) / CHUNK;
  for (int ch = blockIdx.x; ch < n_chunks; ch += gridDim.x) {
    const int bk_end = min(n_blocks, (ch + 1) * CHUNK);
    for (int bk = ch * CHUNK; bk < bk_end; ++bk) {
      const int2 br = blocks[bk];
      if (br.x < 0 || br.y <= br.x) continue;  // empty block (uniform)
      const int bs = br.x, nb = br.y - br.x;
      const uint64_t key0 = (uint64_t)bk << (2 * B);
      const int cx0 = (int)compact_even(key0), cy0 = (int)compact_even(key0 >> 1);

      // box cells: counts and global starts; own cells map into [bs, bs + nb)
      int qcnt[QPT], qgs[QPT];
      bool qown[QPT];
      int halo = 0;
#pragma unroll
      for (int u = 0; u < QPT; ++u) {
        const int q = tid * QPT + u;
        int cnt = 0, gs = 0;
        bool own = false;
        if (q < NBOX) {
          const int lx = q % BOX, ly = q / BOX;
          int gx = cx0 - H + lx, gy = cy0 - H + ly;
          own = gx >= cx0 && gx < cx0 + BS && gy >= cy0 && gy < cy0 + BS && gx < g.nx &&
                gy < g.ny;
          bool ok = true;
          if (BC == 1) {
            if (!own) {
              gx = wrap_mod(gx, g.nx);
              gy = wrap_mod(gy, g.ny);
            }
          } else {
            ok = gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny;
          }
          if (ok) {
            const int2 se = cells[cell_key(g, gx, gy)];
            if (se.x >= 0 && se.y > se.x) {
              cnt = se.y - se.x;
              gs = se.x;
            }
          }
        }
        qcnt[u] = cnt;
        qgs[u] = gs;
        qown[u] = own;
        halo += own ? 0 : cnt;
      }
      int incl = halo;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
      }
      if (lane == kWave - 1) s_wsum[wave] = incl;
      __syncthreads();
      int woff = 0, total = 0;
#pragma unroll
      for (int w = 0; w < WAVES; ++w) {
        woff += w < wave ? s_wsum[w] : 0;
        total += s_wsum[w];
      }
      int run = nb + woff + incl - halo;
#pragma unroll
      for (int u = 0; u < QPT; ++u) {
        const int q = tid * QPT + u;
        if (q < NBOX) {
          s_cnt[q] = qcnt[u];
          if (qown[u]) {
            s_off[q] = qgs[u] - bs;
          } else {
            s_off[q] = run;
            run += qcnt[u];
          }
        }
      }
      const int P = nb + total;
      ++st_blocks;
      st_staged += (unsigned long long)P;

      if (P <= MAXP) {
        // ---- stage: the block's run, then the halo cells
        for (int k = tid; k < nb; k += BWG) {
          const int i = bs + k;
          s_x[k] = rx[i];
          s_y[k] = ry[i];
          if (NEED_V) {
            s_vx[k] = vx[i];
            s_vy[k] = vy[i];
          }
          if (MODE == SPLIT_FORCE) {
            s_rho[k] = rho_g[i];
            s_pr[k] = pr_g[i];
          }
          s_gi[k] = i;
        }
#pragma unroll
        for (int u = 0; u < QPT; ++u) {
          const int q = tid * QPT + u;
          if (q < NBOX && qcnt[u] > 0) {
            const int o = s_off[q];
            for (int m = 0; m < qcnt[u]; ++m) {
              s_q[o + m] = (short)q;
              if (!qown[u]) {
                const int j = qgs[u] + m;
                s_x[o + m] = rx[j];
                s_y[o + m] = ry[j];
                if (NEED_V) {
                  s_vx[o + m] = vx[j];
                  s_vy[o + m] = vy[j];
                }
                if (MODE == SPLIT_FORCE) {
                  s_rho[o + m] = rho_g[j];
                  s_pr[o + m] = pr_g[j];
                }
                s_gi[o + m] = j;
              }
            }
          }
        }
        __syncthreads();

        if (MODE != SPLIT_FORCE) {
          // ---- density: own particles (split) or all within one cell (fused)
          const int dcount = MODE == SPLIT_DENSITY ? nb : P;
          for (int d = tid; d < dcount; d += BWG) {
            const int q = s_q[d];
            const int lx = q % BOX, ly = q / BOX;
            if (MODE == FUSED && (lx < 1 || lx > BOX - 2 || ly < 1 || ly > BOX - 2)) continue;
            const float xi = s_x[d], yi = s_y[d];
            const int gi = s_gi[d];
            float acc = density_term(c.h2, 0.0f);
#pragma unroll
            for (int ox = -1; ox <= 1; ++ox) {
#pragma unroll
              for (int oy = -1; oy <= 1; ++oy) {
                const int qq = (lx + ox) + (ly + oy) * BOX;
                const int o = s_off[qq], m = s_cnt[qq];
                for (int e = o; e < o + m; ++e) {
                  if (s_gi[e] == gi) continue;
                  float dx, dy;
                  separation<BC>(c, xi, yi, s_x[e], s_y[e], dx, dy);
                  const float r2 = dx * dx + dy * dy;
                  if (r2 > c.h2) continue;
                  acc += density_term(c.h2, r2);
                }
              }
            }
            float rho, p, pr;
            eos<PHYS>(c, acc, rho, p, pr);
            if (MODE == SPLIT_DENSITY) {
              rho_g[gi] = rho;
              pr_g[gi] = pr;
            } else {
              s_rho[d] = rho;
              s_pr[d] = pr;
            }
          }
        }
        if (MODE != SPLIT_DENSITY) {
          if (MODE == FUSED) __syncthreads();
          // ---- forces, integration and next keys of the block's particles
          for (int d = tid; d < nb; d += BWG) {
            const int q = s_q[d];
            const int lx = q % BOX, ly = q / BOX;
            const float xi = s_x[d], yi = s_y[d], vxi = s_vx[d], vyi = s_vy[d];
            const float rhoi = s_rho[d], pri = s_pr[d];
            const int gi = s_gi[d];
            Acc a{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ox = -1; ox <= 1; ++ox) {
#pragma unroll
              for (int oy = -1; oy <= 1; ++oy) {
                const int qq = (lx + ox) + (ly + oy) * BOX;
                const int o = s_off[qq], m = s_cnt[qq];
                for (int e = o; e < o + m; ++e) {
                  if (s_gi[e] == gi) continue;
                  float dx, dy;
                  separation<BC>(c, xi, yi, s_x[e], s_y[e], dx, dy);
                  const float r2 = dx * dx + dy * dy;
                  if (r2 > c.h2) continue;
                  force_pair<PHYS>(c, dx, dy, r2, rhoi, s_rho[e], pri, s_pr[e], vxi, vyi,
                                   s_vx[e], s_vy[e], a);
                }
              }
            }
            float axi, ayi;
            force_finish<PHYS>(c, a, axi, ayi);
            if (MODE == FUSED) {
              rho_g[gi] = rhoi;
              pr_g[gi] = pri;
            }
            write_integrated<BC>(g, c, gi, xi, yi, vxi, vyi, axi, ayi, ax_out, ay_out, tx, ty,
                                 tvx, tvy, keys_out);
          }
        }
      } else {
        // ---- over the LDS capacity: the same arithmetic from global memory
        ++st_fallback;
        if (MODE == SPLIT_DENSITY) {
          for (int d = tid; d < nb; d += BWG) {
            float rho, pr;
            density_global<PHYS, BC>(g, c, rx, ry, cells, n, bs + d, rho, pr);
            rho_g[bs + d] = rho;
            pr_g[bs + d] = pr;
          }
        } else {
          for (int d = tid; d < nb; d += BWG) {
            const int i = bs + d;
            const float xi = rx[i], yi = ry[i], vxi = vx[i], vyi = vy[i];
            float rhoi, pri;
            if (MODE == FUSED) {
              density_global<PHYS, BC>(g, c, rx, ry, cells, n, i, rhoi, pri);
            } else {
              rhoi = rho_g[i];
              pri = pr_g[i];
            }
            Acc a{0.f, 0.f, 0.f, 0.f};
            for_each_neighbor<BC>(g, c, rx, ry, cells, n, i, xi, yi,
                                  [&](int j, float dx, float dy, float r2) {
                                    float rhoj, prj;
                                    if (MODE == FUSED) {
                                      density_global<PHYS, BC>(g, c, rx, ry, cells, n, j, rhoj,
                                                               prj);
                                    } else {
                                      rhoj = rho_g[j];
                                      prj = pr_g[j];
                                    }
                                    force_pair<PHYS>(c, dx, dy, r2, rhoi, rhoj, pri, prj, vxi,
                                                     vyi, vx[j], vy[j], a);
                                  });
            float axi, ayi;
            force_finish<PHYS>(c, a, axi, ayi);
            if (MODE == FUSED) {
              rho_g[i] = rhoi;
              pr_g[i] = pri;
            }
            write_integrated<BC>(g, c, i, xi, yi, vxi, vyi, axi, ayi, ax_out, ay_out, tx, ty,
                                 tvx, tvy, keys_out);
          }
        }
      }
      __syncthreads();  // LDS reused by the next block
    }
  }
  if (stats && tid == 0) {
    atomicAdd(&stats[0], st_blocks);
    atomicAdd(&stats[1], st_fallback);
    atomicAdd(&stats[2], st_staged);
  }
}

constexpr int kMaxPFused = 768;
constexpr int kMaxPSplit = 1024;

template <int PHYS, int BC>
hipError_t blocks_go(const Launch& L, const DevState& s, uint64_t* keys_out, int grid,
                     int variant, unsigned long long* stats, int part) {
  if (s.block_bits != 3 || grid <= 0) return hipErrorInvalidValue;
  const int nbk = (int)s.n_blocks;
  if (variant == SPH_STEP_FUSED_BLOCK) {
    k_blocks<PHYS, BC, 3, 2, FUSED, kMaxPFused><<<grid, BWG, 0, L.stream>>>(
        s.n, nbk, L.grid, L.phys, s.blocks, s.cells, s.rx, s.ry, s.vx, s.vy, s.rho, s.pr, s.ax,
        s.ay, s.tx, s.ty, s.tvx, s.tvy, keys_out, stats);
  } else {
    if (part != 2) {
      k_blocks<PHYS, BC, 3, 1, SPLIT_DENSITY, kMaxPSplit><<<grid, BWG, 0, L.stream>>>(
          s.n, nbk, L.grid, L.phys, s.blocks, s.cells, s.rx, s.ry, s.vx, s.vy, s.rho, s.pr,
          s.ax, s.ay, s.tx, s.ty, s.tvx, s.tvy, keys_out, stats);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    if (part != 1)
      k_blocks<PHYS, BC, 3, 1, SPLIT_FORCE, kMaxPSplit><<<grid, BWG, 0, L.stream>>>(
          s.n, nbk, L.grid, L.phys, s.blocks, s.cells, s.rx, s.ry, s.vx, s.vy, s.rho, s.pr,
          s.ax, s.ay, s.tx, s.ty, s.tvx, s.tvy, keys_out, stats ? stats + 3 : nullptr);
  }
  return hipGetLastError();
}

}  // namespace

int block_step_grid(int64_t n_blocks, int cus) {
  int64_t chunks = (n_blocks + CHUNK - 1) / CHUNK;
  return (int)std::max<int64_t>(1, std::min<int64_t>(chunks, (int64_t)cus * 8));
}

hipError_t launch_block_step(const Launch& L, const DevState& s, uint64_t* keys_out, int grid,
                             int variant, unsigned long long* stats, int part) {
  if (s.n <= 0) return hipSuccess;
  const int ph = L.phys.physics, bc = L.phys.boundary;
  if (ph == 0 && bc == 0) return blocks_go<0, 0>(L, s, keys_out, grid, variant, stats, part);
  if (ph == 0 && bc == 1) return blocks_go<0, 1>(L, s, keys_out, grid, variant, stats, part);
  if (ph == 1 && bc == 0) return blocks_go<1, 0>(L, s, keys_out, grid, variant, stats, part);
  return blocks_go<1, 1>(L, s, keys_out, grid, variant, stats, part);
}

}  // namespace sphamd
