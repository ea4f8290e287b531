// Fused SPH step over occupied Morton blocks (gfx950).
//
// A Morton block is a square of 2^B x 2^B cells whose keys form one contiguous
// range, so its particles are one contiguous run of the sorted arrays. One
// workgroup takes a block and stages in LDS the particles of the block plus a
// two-cell halo ((2^B + 4)^2 cells): the block's own particles with one
// coalesced copy, the halo cell by cell. It then computes, from LDS only,
//   1. density + EOS of every staged particle within one cell of the block
//      (the halo ring 2 completes their stencils),
//   2. accelerations of the block's particles (their neighbors' densities are
//      all in LDS),
//   3. symplectic Euler and the cell key of the new position for the next sort,
// and writes the results with coalesced stores. Neighbor visits keep the stencil
// order of the global kernels (x offset outer, y inner, ascending index), so
// every float is produced by the same operation sequence as the oracle.
// A block whose staged particles exceed the LDS capacity is computed from global
// memory instead (densities of neighbors recomputed), with the same arithmetic.
#include <hip/hip_runtime.h>

#include "sph_internal.hpp"

namespace sphamd {

namespace {

constexpr int BWG = 128;

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

// Density + EOS of sorted particle j from global memory (fallback path).
template <int PHYS, int BC>
__device__ void density_global(const DevGrid& g, const DevPhys& c, const float* rx,
                               const float* ry, const int2* cells, int n, int j, float& rho,
                               float& pr) {
  float acc = density_term(c.h2, 0.0f);
  for_each_neighbor<BC>(g, c, rx, ry, cells, n, j, rx[j], ry[j],
                        [&](int, float, float, float r2) { acc += density_term(c.h2, r2); });
  float p;
  eos<PHYS>(c, acc, rho, p, pr);
}

template <int PHYS, int BC, int B, int MAXP>
__global__ __launch_bounds__(BWG) void k_block_step(
    int n, DevGrid g, DevPhys c, const int* __restrict__ active,
    const int* __restrict__ n_active_p, const int2* __restrict__ blocks,
    const int2* __restrict__ cells, const float* __restrict__ rx,
    const float* __restrict__ ry, const float* __restrict__ vx, const float* __restrict__ vy,
    float* __restrict__ rho_out, float* __restrict__ pr_out, float* __restrict__ ax_out,
    float* __restrict__ ay_out, float* __restrict__ tx, float* __restrict__ ty,
    float* __restrict__ tvx, float* __restrict__ tvy, uint64_t* __restrict__ keys_out) {
  constexpr int BS = 1 << B;
  constexpr int BOX = BS + 4;
  constexpr int NBOX = BOX * BOX;
  constexpr int QPT = (NBOX + BWG - 1) / BWG;
  constexpr int WAVES = BWG / kWave;
  __shared__ float s_x[MAXP], s_y[MAXP], s_vx[MAXP], s_vy[MAXP], s_rho[MAXP], s_pr[MAXP];
  __shared__ int s_gi[MAXP];
  __shared__ short s_q[MAXP];
  __shared__ int s_off[NBOX], s_cnt[NBOX];
  __shared__ int s_wsum[WAVES];

  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int na = *n_active_p;
  // XCD-aware order: workgroups are dealt to the 8 XCDs round robin, so each
  // XCD takes a contiguous slice of the occupied-block list (the grid is a
  // multiple of 8, so every slice has gridDim.x / 8 workgroups).
  constexpr int nxcd = 8;
  const int per_xcd = gridDim.x / nxcd;
  const int slice = (na + nxcd - 1) / nxcd;
  const int lo_t = (blockIdx.x % nxcd) * slice, hi_t = min(na, lo_t + slice);
  for (int t = lo_t + blockIdx.x / nxcd; t < hi_t; t += per_xcd) {
    const int bk = active[t];
    const int2 br = blocks[bk];
    const int bs = br.x, nb = br.y - br.x;
    const uint64_t key0 = (uint64_t)bk << (2 * B);
    const int cx0 = (int)compact_even(key0), cy0 = (int)compact_even(key0 >> 1);

    // box cells: counts, global starts, own/halo; exclusive scan of halo counts
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
        int gx = cx0 - 2 + lx, gy = cy0 - 2 + ly;
        own = gx >= cx0 && gx < cx0 + BS && gy >= cy0 && gy < cy0 + BS && gx < g.nx &&
              gy < g.ny;
        bool ok;
        if (BC == 1) {
          ok = own || true;
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
    // workgroup exclusive scan of the halo counts
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
    __syncthreads();

    if (P <= MAXP) {
      // ---- stage: own particles (one coalesced run), then halo cells
      for (int k = tid; k < nb; k += BWG) {
        const int i = bs + k;
        s_x[k] = rx[i];
        s_y[k] = ry[i];
        s_vx[k] = vx[i];
        s_vy[k] = vy[i];
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
              s_vx[o + m] = vx[j];
              s_vy[o + m] = vy[j];
              s_gi[o + m] = j;
            }
          }
        }
      }
      __syncthreads();

      // ---- density of every staged particle within one cell of the block
      for (int d = tid; d < P; d += BWG) {
        const int q = s_q[d];
        const int lx = q % BOX, ly = q / BOX;
        if (lx < 1 || lx > BOX - 2 || ly < 1 || ly > BOX - 2) continue;
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
        s_rho[d] = rho;
        s_pr[d] = pr;
      }
      __syncthreads();

      // ---- forces, integration and next keys of the block's own particles
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
              force_pair<PHYS>(c, dx, dy, r2, rhoi, s_rho[e], pri, s_pr[e], vxi, vyi, s_vx[e],
                               s_vy[e], a);
            }
          }
        }
        float axi, ayi;
        force_finish<PHYS>(c, a, axi, ayi);
        const int i = bs + d;
        rho_out[i] = rhoi;
        pr_out[i] = pri;
        ax_out[i] = axi;
        ay_out[i] = ayi;
        float x = xi, y = yi, u = vxi, v = vyi;
        integrate_one<BC>(c, x, y, u, v, axi, ayi);
        tx[i] = x;
        ty[i] = y;
        tvx[i] = u;
        tvy[i] = v;
        keys_out[i] = cell_key(g, cell_coord(x, g.lo_x, g.cs_x, g.nx),
                               cell_coord(y, g.lo_y, g.cs_y, g.ny));
      }
    } else {
      // ---- over capacity: the same arithmetic from global memory
      for (int d = tid; d < nb; d += BWG) {
        const int i = bs + d;
        const float xi = rx[i], yi = ry[i], vxi = vx[i], vyi = vy[i];
        float rhoi, pri;
        density_global<PHYS, BC>(g, c, rx, ry, cells, n, i, rhoi, pri);
        Acc a{0.f, 0.f, 0.f, 0.f};
        for_each_neighbor<BC>(g, c, rx, ry, cells, n, i, xi, yi,
                              [&](int j, float dx, float dy, float r2) {
                                float rhoj, prj;
                                density_global<PHYS, BC>(g, c, rx, ry, cells, n, j, rhoj, prj);
                                force_pair<PHYS>(c, dx, dy, r2, rhoi, rhoj, pri, prj, vxi, vyi,
                                                 vx[j], vy[j], a);
                              });
        float axi, ayi;
        force_finish<PHYS>(c, a, axi, ayi);
        rho_out[i] = rhoi;
        pr_out[i] = pri;
        ax_out[i] = axi;
        ay_out[i] = ayi;
        float x = xi, y = yi, u = vxi, v = vyi;
        integrate_one<BC>(c, x, y, u, v, axi, ayi);
        tx[i] = x;
        ty[i] = y;
        tvx[i] = u;
        tvy[i] = v;
        keys_out[i] = cell_key(g, cell_coord(x, g.lo_x, g.cs_x, g.nx),
                               cell_coord(y, g.lo_y, g.cs_y, g.ny));
      }
    }
    __syncthreads();
  }
}

constexpr int kMaxP = 768;

template <int PHYS, int BC>
hipError_t block_step_go(const Launch& L, const DevState& s, uint64_t* keys_out, int grid) {
  if (s.block_bits != 3 || grid <= 0 || grid % 8 != 0) return hipErrorInvalidValue;
  k_block_step<PHYS, BC, 3, kMaxP><<<grid, BWG, 0, L.stream>>>(
      s.n, L.grid, L.phys, s.active, s.n_active, s.blocks, s.cells, s.rx, s.ry, s.vx, s.vy,
      s.rho, s.pr, s.ax, s.ay, s.tx, s.ty, s.tvx, s.tvy, keys_out);
  return hipGetLastError();
}

}  // namespace

int block_step_max_particles(int) { return kMaxP; }

hipError_t launch_block_step(const Launch& L, const DevState& s, uint64_t* keys_out, int grid) {
  if (s.n <= 0) return hipSuccess;
  const int ph = L.phys.physics, bc = L.phys.boundary;
  if (ph == 0 && bc == 0) return block_step_go<0, 0>(L, s, keys_out, grid);
  if (ph == 0 && bc == 1) return block_step_go<0, 1>(L, s, keys_out, grid);
  if (ph == 1 && bc == 0) return block_step_go<1, 0>(L, s, keys_out, grid);
  return block_step_go<1, 1>(L, s, keys_out, grid);
}

}  // namespace sphamd
