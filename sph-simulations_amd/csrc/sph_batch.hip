// gfx950 kernels of the batched step: many independent scenes advance in one
// launch per stage. Every workgroup covers one tile of one scene's contiguous
// segment of the sorted order (the tile table is static because scene segments
// never move), so the scene's constants are workgroup-uniform and come in
// through scalar loads; the per-particle arithmetic is the single-scene code,
// neighbor data gathered from the packed copies (x, y, vx, vy) and (rho, pr).
#include <hip/hip_runtime.h>

#include "sph_internal.hpp"

namespace sphamd {

namespace {

template <int BC>
__global__ __launch_bounds__(kMaxTile) void k_bkeys(const TileDev* __restrict__ tiles,
                                                    const SceneDev* __restrict__ scenes,
                                                    const float* __restrict__ rx,
                                                    const float* __restrict__ ry,
                                                    uint64_t* __restrict__ keys) {
  const TileDev t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const int i = t.begin + (int)threadIdx.x;
  if (i >= t.end) return;
  const DevGrid g = scenes[t.scene].g;
  const uint64_t base = scenes[t.scene].base;
  keys[i] = base + cell_key(g, cell_coord(rx[i], g.lo_x, g.cs_x, g.nx),
                            cell_coord(ry[i], g.lo_y, g.cs_y, g.ny));
}

template <int PHYS, int BC>
__global__ __launch_bounds__(kMaxTile) void k_bdensity(
    int n, const TileDev* __restrict__ tiles, const SceneDev* __restrict__ scenes,
    const float* __restrict__ rx, const float* __restrict__ ry, const int2* __restrict__ cells,
    int32_t* __restrict__ ell_count, int32_t* __restrict__ ell_idx, int ell_w,
    float* __restrict__ rho_out, float* __restrict__ pr_out, float* __restrict__ p_out,
    const float4* __restrict__ xyuv, float2* __restrict__ rp_out) {
  extern __shared__ int s_list[];  // kEllLds x blockDim.x
  const TileDev t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const int i = t.begin + (int)threadIdx.x;
  if (i >= t.end) return;
  const SceneDev& S = scenes[t.scene];
  const DevGrid g = S.g;
  const DevPhys c = S.c;
  const float2 qi = ld8_of16(xyuv, i);  // the packed copy is the current state
  const float xi = qi.x, yi = qi.y;
  const float acc = density_ell<BC, true>(g, c, rx, ry, cells + S.base, n, i, xi, yi,
                                          s_list + threadIdx.x, (int)blockDim.x, ell_count,
                                          ell_idx, ell_w, xyuv);
  float rho, p, pr;
  eos<PHYS>(c, acc, rho, p, pr);
  if (rho_out) {  // (null: a step whose separate fields nobody reads)
    rho_out[i] = rho;
    pr_out[i] = pr;
  }
  if (p_out) p_out[i] = p;
  rp_out[i] = make_float2(rho, pr);
}

template <int PHYS, int BC, typename K>
__global__ __launch_bounds__(kMaxTile) void k_bforces(
    int n, const TileDev* __restrict__ tiles, const SceneDev* __restrict__ scenes,
    const float* __restrict__ rx, const float* __restrict__ ry, const float* __restrict__ vx,
    const float* __restrict__ vy, const float* __restrict__ rho, const float* __restrict__ pr,
    const int2* __restrict__ cells, const int32_t* __restrict__ ell_count,
    const int32_t* __restrict__ ell_idx, int ell_w, float* __restrict__ ax_out,
    float* __restrict__ ay_out, float* __restrict__ tx, float* __restrict__ ty,
    float* __restrict__ tvx, float* __restrict__ tvy, K* __restrict__ keys_out,
    const float4* __restrict__ xyuv, const float2* __restrict__ rp,
    const K* __restrict__ kold, MoverSink movers, float4* __restrict__ t4) {
  unsigned long long* __restrict__ flags = movers.flags;
  const TileDev t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
  const int i = t.begin + (int)threadIdx.x;
  if (i >= t.end) return;
  // the previous key, for the re-sort's mover report (flags null: none)
  const uint64_t ko = flags ? (uint64_t)kold[i] : 0ull;
  const SceneDev& S = scenes[t.scene];
  const DevGrid g = S.g;
  const DevPhys c = S.c;
  const uint64_t base = S.base;
  const float4 qi = ld16(xyuv, i);  // the packed copies are the current state
  const float2 ri = ld8(rp, i);
  const float xi = qi.x, yi = qi.y, vxi = qi.z, vyi = qi.w, rhoi = ri.x, pri = ri.y;
  const Acc a = forces_ell<PHYS, BC, true>(g, c, rx, ry, vx, vy, rho, pr, cells + base,
                                           ell_count, ell_idx, ell_w, n, i, xi, yi, vxi, vyi,
                                           rhoi, pri, xyuv, rp);
  float axi, ayi;
  force_finish<PHYS>(c, a, axi, ayi);
  if (ax_out) {  // (null: a step whose accelerations nobody reads)
    ax_out[i] = axi;
    ay_out[i] = ayi;
  }
  float x = xi, y = yi, u = vxi, v = vyi;
  integrate_one<BC>(c, x, y, u, v, axi, ayi);
  if constexpr (sizeof(K) == 4) {
    t4[i] = make_float4(x, y, u, v);  // a narrow step keeps T packed
  } else {
    tx[i] = x;
    ty[i] = y;
    tvx[i] = u;
    tvy[i] = v;
  }
  const uint64_t kn = base + cell_key(g, cell_coord(x, g.lo_x, g.cs_x, g.nx),
                                      cell_coord(y, g.lo_y, g.cs_y, g.ny));
  keys_out[i] = (K)kn;
  if (flags) {
    // The mover flags of the incremental re-sort: a workgroup's particles start
    // anywhere in the order, so a wave's 64 bits straddle at most two flag
    // words; OR them in (the merge zeroes the words after reading them).
    const unsigned long long bal = __ballot(kn != ko);
    if ((threadIdx.x & (kWave - 1)) == 0 && bal != 0ull) {
      const int w0 = i / kWave, sh = i % kWave;
      atomicOr(&flags[w0], bal << sh);
      if (sh != 0 && (bal >> (kWave - sh)) != 0ull) atomicOr(&flags[w0 + 1], bal >> (kWave - sh));
    }
    // and each mover into its owner tile's bucket (bucket_sink)
    if (movers.bucket) bucket_movers<false>(movers, i, kn, kn != ko, make_float4(x, y, u, v));
  }
}

bool tile_ok(const BatchLaunch& B) {
  return B.n_tiles > 0 && B.tile >= kWave && B.tile <= kMaxTile && B.tile % kWave == 0;
}

}  // namespace

hipError_t launch_batch_keys(const Launch& L, const BatchLaunch& B, const DevState& s,
                             uint64_t* keys_out) {
  if (s.n <= 0) return hipSuccess;
  if (!tile_ok(B)) return hipErrorInvalidValue;
  if (L.phys.boundary == 1)
    k_bkeys<1><<<B.n_tiles, B.tile, 0, L.stream>>>(B.tiles, B.scenes, s.rx, s.ry, keys_out);
  else
    k_bkeys<0><<<B.n_tiles, B.tile, 0, L.stream>>>(B.tiles, B.scenes, s.rx, s.ry, keys_out);
  return hipGetLastError();
}

template <int PHYS, int BC>
static void bdensity_go(const Launch& L, const BatchLaunch& B, const DevState& s) {
  const size_t lds = (size_t)kEllLds * B.tile * sizeof(int);
  k_bdensity<PHYS, BC><<<B.n_tiles, B.tile, lds, L.stream>>>(
      s.n, B.tiles, B.scenes, s.rx, s.ry, s.cells, s.ell_count, s.ell_idx, s.ell_width, s.rho,
      s.pr, s.pres, s.xyuv, s.rp);
}

hipError_t launch_batch_density(const Launch& L, const BatchLaunch& B, const DevState& s) {
  if (s.n <= 0) return hipSuccess;
  if (!tile_ok(B)) return hipErrorInvalidValue;
  const int ph = L.phys.physics, bc = L.phys.boundary;
  if (ph == 0 && bc == 0) bdensity_go<0, 0>(L, B, s);
  else if (ph == 0 && bc == 1) bdensity_go<0, 1>(L, B, s);
  else if (ph == 1 && bc == 0) bdensity_go<1, 0>(L, B, s);
  else bdensity_go<1, 1>(L, B, s);
  return hipGetLastError();
}

template <int PHYS, int BC>
static void bforces_go(const Launch& L, const BatchLaunch& B, const DevState& s,
                       uint64_t* keys_out, const MoverSink& flags) {
  if (s.narrow)  // the step's narrow keys: read nkeys, write nkeys_in
    k_bforces<PHYS, BC, uint32_t><<<B.n_tiles, B.tile, 0, L.stream>>>(
        s.n, B.tiles, B.scenes, s.rx, s.ry, s.vx, s.vy, s.rho, s.pr, s.cells, s.ell_count,
        s.ell_idx, s.ell_width, s.ax, s.ay, s.tx, s.ty, s.tvx, s.tvy, s.nkeys_in, s.xyuv, s.rp,
        s.nkeys, flags, s.t4);
  else
    k_bforces<PHYS, BC, uint64_t><<<B.n_tiles, B.tile, 0, L.stream>>>(
        s.n, B.tiles, B.scenes, s.rx, s.ry, s.vx, s.vy, s.rho, s.pr, s.cells, s.ell_count,
        s.ell_idx, s.ell_width, s.ax, s.ay, s.tx, s.ty, s.tvx, s.tvy, keys_out, s.xyuv, s.rp,
        s.keys_sorted, flags, nullptr);
}

hipError_t launch_batch_forces(const Launch& L, const BatchLaunch& B, const DevState& s,
                               uint64_t* keys_out, MoverSink movers) {
  if (s.n <= 0) return hipSuccess;
  if (!tile_ok(B)) return hipErrorInvalidValue;
  const int ph = L.phys.physics, bc = L.phys.boundary;
  const MoverSink& f = movers;
  if (ph == 0 && bc == 0) bforces_go<0, 0>(L, B, s, keys_out, f);
  else if (ph == 0 && bc == 1) bforces_go<0, 1>(L, B, s, keys_out, f);
  else if (ph == 1 && bc == 0) bforces_go<1, 0>(L, B, s, keys_out, f);
  else bforces_go<1, 1>(L, B, s, keys_out, f);
  return hipGetLastError();
}

}  // namespace sphamd
