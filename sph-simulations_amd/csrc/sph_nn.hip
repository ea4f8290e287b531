// Surrogate message-passing network on gfx950: BF16 MFMA GEMMs with FP32
// accumulation for every node-level linear layer (forward, input gradients and
// weight gradients), FP32 edge kernels over the fixed-width neighbor lists, and
// Adam on FP32 master weights. The network (features, layers, readout) is
// written out next to the CPU oracle that checks it; the parameter order here
// is the one the oracle's param_layout() lists.
//
// Rounding contract: GEMM operands (weights and activations) are rounded to
// bfloat16 (nearest even) and accumulated in FP32; edge kernels use the
// bf16-rounded weight matrices in FP32 arithmetic; biases stay FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "sph_nn.hpp"

namespace sphamd {

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;

constexpr int W = kNnWidth;
constexpr int kNodesPerWave = 16;         // edge kernels: nodes per wavefront
constexpr int kEdgeNodesPerBlock = 4 * kNodesPerWave;
constexpr int kEdgePartial = W * kEdgeF + W + W + 1 + kEdgeF;  // dC, dc, dw, db, du
constexpr int kRowsPerChunk = 256;        // weight-gradient row chunks (one partial each)
constexpr int kColRows = 64;              // bias-gradient (column sum) row chunks
constexpr int kMaxNnNeighbors = 32;       // neighbors per node the network reads

__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
// SiLU and its derivative from the hardware exp2 / reciprocal approximations
// (about 1 ulp each; the network is checked against the oracle with tolerances,
// not bit for bit).
__device__ __forceinline__ float sigmoid_f(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504f * z));
}
__device__ __forceinline__ float silu_f(float z) { return z * sigmoid_f(z); }
__device__ __forceinline__ float dsilu_f(float z) {
  const float s = sigmoid_f(z);
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ float fast_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// ------------------------------------------------------------------ GEMM
// Y[n x Mo] = [X1 | X2] * B^T with B [Mo x K] bf16 row-major (K = K1 + K2, a
// compile-time 32, 128 or 256), computed as Y^T = B X^T so that each lane's
// accumulators are four consecutive output columns of one row (16-byte stores).
// A workgroup stages its 128-column block of B in LDS once (rows padded by 16
// bytes against bank conflicts) and walks 128-row tiles; wave w owns rows
// 32w..32w+31 of a tile and all 128 columns as four 32x32 MFMA tiles
// (v_mfma_f32_32x32x16_bf16). Operand layout of that instruction: lane l holds
// row (l % 32) of the A operand (here B's columns) and column (l % 32) of the B
// operand (here X's rows) at k = 8 (l / 32) .. +8; accumulator register q of
// lane l is element (8 (q / 4) + 4 (l / 32) + q % 4, l % 32) -- output column
// and row respectively.
struct GemmArgs {
  int n, K1, K2, Mo;
  const uint16_t* X1;
  const uint16_t* X2;
  const uint16_t* B;
  const float* bias;      // [Mo] or null
  const float* rowscale;  // [n] or null: + rowscale[row] * bias2[col]
  const float* bias2;
  const float* R;         // residual [n x Mo] or null (may alias Yf)
  float* Yf;              // value before the activation, or null
  uint16_t* Yb;           // bf16 of act(value), or null
  int act;                // 0 identity, 1 SiLU (applies to Yb only)
};

template <int K>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs a) {
  constexpr int LDB = K + 8;  // padded LDS row (elements)
  __shared__ __attribute__((aligned(16))) uint16_t sB[128 * LDB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col0 = blockIdx.y * 128;
  for (int idx = threadIdx.x; idx < 128 * (K / 8); idx += 256) {
    const int r = idx / (K / 8), c8 = idx % (K / 8);
    *reinterpret_cast<u32x4*>(sB + r * LDB + c8 * 8) =
        *reinterpret_cast<const u32x4*>(a.B + (int64_t)(col0 + r) * K + c8 * 8);
  }
  __syncthreads();
  const int kh = (lane >> 5) * 8;
  const int tiles = (a.n + 127) / 128;
  for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int row = tile * 128 + wave * 32 + (lane & 31);  // this lane's output row
    f32x16 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[t][q] = 0.0f;
    u32x4 xv[K / 16];
#pragma unroll
    for (int ks = 0; ks < K / 16; ++ks) {
      const int k = ks * 16;
      xv[ks] = u32x4{0u, 0u, 0u, 0u};
      if (row < a.n)
        xv[ks] = *reinterpret_cast<const u32x4*>(
            k < a.K1 ? a.X1 + (int64_t)row * a.K1 + k + kh
                     : a.X2 + (int64_t)row * a.K2 + (k - a.K1) + kh);
    }
#pragma unroll
    for (int ks = 0; ks < K / 16; ++ks) {
      const bf16x8 X = as_bf16x8(xv[ks]);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const u32x4 bv =
            *reinterpret_cast<const u32x4*>(sB + (t * 32 + (lane & 31)) * LDB + ks * 16 + kh);
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(bv), X, acc[t], 0, 0, 0);
      }
    }
    if (row >= a.n) continue;
    const float rs = a.rowscale ? a.rowscale[row] : 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int col = col0 + t * 32 + g * 8 + (lane >> 5) * 4;
        float4 v = make_float4(acc[t][4 * g], acc[t][4 * g + 1], acc[t][4 * g + 2],
                               acc[t][4 * g + 3]);
        if (a.bias) {
          const float4 bc = *reinterpret_cast<const float4*>(a.bias + col);
          v.x += bc.x; v.y += bc.y; v.z += bc.z; v.w += bc.w;
        }
        if (a.rowscale) {
          const float4 b2 = *reinterpret_cast<const float4*>(a.bias2 + col);
          v.x += rs * b2.x; v.y += rs * b2.y; v.z += rs * b2.z; v.w += rs * b2.w;
        }
        const int64_t o = (int64_t)row * a.Mo + col;
        if (a.R) {
          const float4 rr = *reinterpret_cast<const float4*>(a.R + o);
          v.x += rr.x; v.y += rr.y; v.z += rr.z; v.w += rr.w;
        }
        if (a.Yf) *reinterpret_cast<float4*>(a.Yf + o) = v;
        if (a.Yb) {
          float4 y = v;
          if (a.act) {
            y.x = silu_f(y.x); y.y = silu_f(y.y); y.z = silu_f(y.z); y.w = silu_f(y.w);
          }
          const uint32_t lo = (uint32_t)f2bf(y.x) | ((uint32_t)f2bf(y.y) << 16);
          const uint32_t hi = (uint32_t)f2bf(y.z) | ((uint32_t)f2bf(y.w) << 16);
          *reinterpret_cast<uint2*>(a.Yb + o) = make_uint2(lo, hi);
        }
      }
    }
  }
}

// Weight gradient partials: partial[chunk][m][k] = sum over the chunk's rows r
// of dY[r][m] * X[r][k], X = [X1 | X2]. Grid (chunks, Mo / 128, ceil(K / 128));
// wave w owns m rows 32w..+31 and up to four 32-column k tiles; the reduction
// index (rows) is the MFMA's k dimension.
struct WgradArgs {
  int n, Mo, K1, K2;
  const uint16_t* dY;
  const uint16_t* X1;
  const uint16_t* X2;
  float* partial;
};

__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int K = a.K1 + a.K2;
  const int m = blockIdx.y * 128 + wave * 32 + (lane & 31);
  const int kc0 = blockIdx.z * 128;
  const int ntk = min(4, (K - kc0) / 32);
  const int kh = (lane >> 5) * 8;
  const int r_begin = blockIdx.x * kRowsPerChunk;
  const int r_end = min(a.n, r_begin + kRowsPerChunk);
  f32x16 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[t][q] = 0.0f;
  for (int r = r_begin; r < r_end; r += 16) {
    uint16_t av[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int row = r + kh + q;
      av[q] = row < r_end ? a.dY[(int64_t)row * a.Mo + m] : (uint16_t)0;
    }
    const bf16x8 A = __builtin_bit_cast(bf16x8, av);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (t >= ntk) break;
      const int k = kc0 + t * 32 + (lane & 31);
      uint16_t bv[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int row = r + kh + q;
        uint16_t v = 0;
        if (row < r_end)
          v = k < a.K1 ? a.X1[(int64_t)row * a.K1 + k] : a.X2[(int64_t)row * a.K2 + (k - a.K1)];
        bv[q] = v;
      }
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, __builtin_bit_cast(bf16x8, bv), acc[t],
                                                        0, 0, 0);
    }
  }
  float* out = a.partial + (int64_t)blockIdx.x * a.Mo * K;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    if (t >= ntk) break;
    const int k = kc0 + t * 32 + (lane & 31);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int mm = blockIdx.y * 128 + wave * 32 + (q >> 2) * 8 + (lane >> 5) * 4 + (q & 3);
      out[(int64_t)mm * K + k] = acc[t][q];
    }
  }
}

// out[m * kr + k] = sum_p partial[p][m][k] (k < kr; partial rows have kp
// entries). G threads per output sum the parts p = g, g + G, ... and a tree
// over g in LDS combines them: a fixed order, so the result is reproducible.
template <int G>
__global__ __launch_bounds__(256) void k_reduce(int parts, int rows, int kp, int kr,
                                                const float* __restrict__ partial,
                                                float* __restrict__ out) {
  constexpr int OB = 256 / G;  // outputs per workgroup
  __shared__ float s_acc[256];
  const int ol = threadIdx.x % OB, g = threadIdx.x / OB;
  const int idx = blockIdx.x * OB + ol;
  const bool live = idx < rows * kr;
  float acc = 0.0f;
  if (live) {
    const int mr = idx / kr, kk = idx % kr;
    const int64_t stride = (int64_t)rows * kp;
    const float* src = partial + (int64_t)mr * kp + kk;
#pragma unroll 4
    for (int p = g; p < parts; p += G) acc += src[p * stride];
  }
  s_acc[threadIdx.x] = acc;  // [g][ol]
  __syncthreads();
#pragma unroll
  for (int h = G / 2; h > 0; h >>= 1) {
    if (g < h) s_acc[threadIdx.x] += s_acc[threadIdx.x + h * OB];
    __syncthreads();
  }
  if (g == 0 && live) out[idx] = s_acc[ol];
}

// Column sums of Y [n x Mo] fp32 (optionally weighted by rowscale) per chunk of
// kColRows rows: thread t takes column t % Mo and every (256 / Mo)-th row of
// the chunk (Mo <= 256), a fixed-order tree in LDS combines the row groups.
__global__ __launch_bounds__(256) void k_colsum(int n, int Mo, const float* __restrict__ Y,
                                                const float* __restrict__ rowscale,
                                                float* __restrict__ partial) {
  __shared__ float s_acc[256];
  const int r_begin = blockIdx.x * kColRows;
  const int r_end = min(n, r_begin + kColRows);
  const int groups = 256 / Mo;
  const int c = threadIdx.x % Mo, rg = threadIdx.x / Mo;
  float acc = 0.0f;
  if (rg < groups) {
#pragma unroll 4
    for (int r = r_begin + rg; r < r_end; r += groups) {
      const float v = Y[(int64_t)r * Mo + c];
      acc += rowscale ? rowscale[r] * v : v;
    }
  }
  s_acc[threadIdx.x] = acc;
  __syncthreads();
  for (int h = groups / 2; h > 0; h >>= 1) {
    if (rg < h) s_acc[threadIdx.x] += s_acc[threadIdx.x + h * Mo];
    __syncthreads();
  }
  if (rg == 0) partial[(int64_t)blockIdx.x * Mo + c] = s_acc[threadIdx.x];
}

// Threads per output of k_reduce for a number of parts: about eight parts each.
int reduce_group(int parts) {
  int g = 1;
  while (g < 256 && g * 8 < parts) g <<= 1;
  return g;
}

void launch_reduce(hipStream_t st, int parts, int rows, int kp, int kr, const float* partial,
                   float* out) {
  const int g = reduce_group(parts);
  const int64_t outs = (int64_t)rows * kr;
  const int blocks = (int)((outs * g + 255) / 256);
  switch (g) {
#define SPH_REDUCE_CASE(G) \
  case G: k_reduce<G><<<blocks, 256, 0, st>>>(parts, rows, kp, kr, partial, out); break;
    SPH_REDUCE_CASE(1)
    SPH_REDUCE_CASE(2)
    SPH_REDUCE_CASE(4)
    SPH_REDUCE_CASE(8)
    SPH_REDUCE_CASE(16)
    SPH_REDUCE_CASE(32)
    SPH_REDUCE_CASE(64)
    SPH_REDUCE_CASE(128)
    SPH_REDUCE_CASE(256)
#undef SPH_REDUCE_CASE
    default: break;
  }
}

// ------------------------------------------------------------ elementwise
__global__ __launch_bounds__(256) void k_silu_bwd(int64_t count, const float* __restrict__ g,
                                                  const float* __restrict__ pre,
                                                  float* __restrict__ out,
                                                  uint16_t* __restrict__ outb) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  const float v = g[i] * dsilu_f(pre[i]);
  out[i] = v;
  outb[i] = f2bf(v);
}

// bf16 copies of a weight matrix [rows x kr]: row-major with K padded to kp, and
// its transpose [kp x rows].
__global__ __launch_bounds__(256) void k_wcopy(int rows, int kr, int kp, const float* __restrict__ w,
                                               uint16_t* __restrict__ wb,
                                               uint16_t* __restrict__ wbt) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= rows * kp) return;
  const int mr = idx / kp, kk = idx % kp;
  const uint16_t b = kk < kr ? f2bf(w[mr * kr + kk]) : (uint16_t)0;
  wb[idx] = b;
  wbt[(int64_t)kk * rows + mr] = b;
}

// Sum of squares of the gradient in FP64: per-workgroup partials (grid
// stride), then one workgroup adds them in a fixed order.
constexpr int kNormBlocks = 256;
__global__ __launch_bounds__(256) void k_sq_partial(int64_t count, const float* __restrict__ g,
                                                    double* __restrict__ part) {
  __shared__ double s[256];
  double a = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * 256) {
    const double x = g[i];
    a += x * x;
  }
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_sq_final(int parts, const double* __restrict__ part,
                                                  double* __restrict__ out) {
  __shared__ double s[256];
  double a = 0.0;
  for (int p = threadIdx.x; p < parts; p += 256) a += part[p];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// Adam on FP32 master weights. With g2 (the gradient's squared norm, on the
// device) the gradient is first scaled to a global norm of at most clip.
__global__ __launch_bounds__(256) void k_adam(int64_t count, float* __restrict__ w,
                                              const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v,
                                              float lr, float b1, float b2, float eps,
                                              float c1, float c2, const double* __restrict__ g2,
                                              float clip) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  float scale = 1.0f;
  if (g2) {
    const double nrm = sqrt(*g2);
    if (nrm > (double)clip) scale = (float)((double)clip / nrm);
  }
  const float gi = g[i] * scale;
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / c1, vh = vi / c2;
  w[i] = w[i] - lr * mh / (sqrtf(vh) + eps);
}

// ------------------------------------------------------------- features
struct NodeIn {
  const float *rx, *ry, *vx, *vy, *rho, *pr;
  const int32_t* cnt;
  const int32_t* idx;  // column-major: slot e of particle i at idx[e * n + i]
  int n, w;
};

__global__ __launch_bounds__(256) void k_node_features(NodeIn in, NnConsts k,
                                                       uint16_t* __restrict__ x,
                                                       float* __restrict__ deg) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= in.n) return;
  const float d = (float)min(in.cnt[i], min(in.w, kMaxNnNeighbors));
  deg[i] = d;
  float f[kNodeF] = {k.rs * (in.rho[i] * k.inv_rho0 - 1.0f), k.ps * in.pr[i], d * 0.125f, 1.0f,
                     0.0f, 0.0f, 0.0f, 0.0f};
  uint16_t* xr = x + (int64_t)i * kNodeK;
#pragma unroll
  for (int c = 0; c < kNodeK; ++c) xr[c] = c < kNodeF ? f2bf(f[c]) : (uint16_t)0;
}

// Invariant features of edge (receiver i, sender j); (dx, dy) = r_i - r_j;
// pri is p / rho^2 of the receiver.
template <int BC>
__device__ __forceinline__ void edge_feat(const NnConsts& k, float xi, float yi, float vxi,
                                          float vyi, float pri, float xj, float yj, float vxj,
                                          float vyj, float rhoj, float prj, float& dx, float& dy,
                                          float* f) {
  dx = xi - xj;
  dy = yi - yj;
  if (BC == 1) {
    dx = dx > k.half_x ? dx - k.len_x : (dx < -k.half_x ? dx + k.len_x : dx);
    dy = dy > k.half_y ? dy - k.len_y : (dy < -k.half_y ? dy + k.len_y : dy);
  }
  const float r = fast_sqrt(dx * dx + dy * dy);
  const float q = r * k.inv_h;
  const float w = (1.0f - q) * (1.0f - q);
  const float dvx = vxi - vxj, dvy = vyi - vyj;
  f[0] = q;
  f[1] = w;
  f[2] = fast_div(w, fmaxf(q, 0.05f));
  f[3] = fast_div((dvx * dx + dvy * dy) * k.inv_vs, fmaxf(r, 0.05f * k.h));
  f[4] = fast_sqrt(dvx * dvx + dvy * dvy) * k.inv_vs;
  f[5] = k.rs * (rhoj * k.inv_rho0 - 1.0f);
  f[6] = k.ps * prj;
  f[7] = f[2] * (k.ps * pri + k.ps * prj);  // the form of the SPH pressure term
}

// Edge kernels: one wavefront per node, 16 nodes per wavefront in turn. For a
// node, lane e (< its neighbor count, at most kMaxNnNeighbors) loads neighbor e
// and computes that edge's invariant features in parallel; the hidden-column
// phase (lane l owns columns 2l and 2l + 1) then broadcasts each edge's values
// with readlane and fetches the neighbor rows of 8 edges at a time, so the loads
// of a node are in flight together instead of one edge after another.

__device__ __forceinline__ float bcast(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// per-lane data of edge `lane` of a node: sender j, features, separation
struct EdgeLane {
  int j;
  float f[kEdgeF];
  float dx, dy;
};

template <int BC>
__device__ __forceinline__ void load_edge_lane(const NodeIn& in, const NnConsts& k, int i, int deg,
                                               float xi, float yi, float vxi, float vyi,
                                               float pri, EdgeLane& el) {
  // Every load is unconditional (the slot clamped into the list, the node
  // itself standing in for a missing neighbor), so that they issue together;
  // lanes past the degree are zeroed afterwards.
  const int lane = threadIdx.x & 63;
  const bool ok = lane < deg;
  const int jl = in.idx[(int64_t)min(lane, in.w - 1) * in.n + i];
  const int j = ok ? jl : i;
  el.j = j;
  const float xj = in.rx[j], yj = in.ry[j], vxj = in.vx[j], vyj = in.vy[j], rhoj = in.rho[j],
              prj = in.pr[j];
  edge_feat<BC>(k, xi, yi, vxi, vyi, pri, xj, yj, vxj, vyj, rhoj, prj, el.dx, el.dy, el.f);
  if (!ok) {
    el.dx = el.dy = 0.0f;
#pragma unroll
    for (int q = 0; q < kEdgeF; ++q) el.f[q] = 0.0f;
  }
}

// --------------------------------------------------------- edge forward
// MODE 0: S_i = sum_j silu(P_i + Q_j + C f_ij + c).
// MODE 1: a_i = g + (a_s / h) sum_j (r_i - r_j) (w . silu(...) + b + u . f_ij).
template <int MODE, int BC>
__global__ __launch_bounds__(256) void k_edge_fwd(NodeIn in, NnConsts k, const float* __restrict__ P,
                                                  const float* __restrict__ Q,
                                                  const uint16_t* __restrict__ Cb,
                                                  const float* __restrict__ cb,
                                                  const uint16_t* __restrict__ wvec,
                                                  const float* __restrict__ bo,
                                                  const float* __restrict__ uvec,
                                                  float* __restrict__ S,
                                                  uint16_t* __restrict__ Sb,
                                                  float* __restrict__ ax, float* __restrict__ ay) {
  const int lane = threadIdx.x & 63, wave = wave_id();
  const int c0 = 2 * lane;
  float C0[kEdgeF], C1[kEdgeF], u[kEdgeF];
#pragma unroll
  for (int e = 0; e < kEdgeF; ++e) {
    C0[e] = bf2f(Cb[c0 * kEdgeF + e]);
    C1[e] = bf2f(Cb[(c0 + 1) * kEdgeF + e]);
    u[e] = MODE == 1 ? uvec[e] : 0.0f;
  }
  const float cb0 = cb[c0], cb1 = cb[c0 + 1];
  float w0 = 0.0f, w1 = 0.0f, b = 0.0f;
  if (MODE == 1) {
    w0 = bf2f(wvec[c0]);
    w1 = bf2f(wvec[c0 + 1]);
    b = bo[0];
  }
  const float sc = k.a_s * k.inv_h;
  const int base = (xcd_tile(blockIdx.x, gridDim.x) * 4 + wave) * kNodesPerWave;
  for (int t = 0; t < kNodesPerWave; ++t) {
    const int i = base + t;
    if (i >= in.n) break;
    const float xi = in.rx[i], yi = in.ry[i], vxi = in.vx[i], vyi = in.vy[i], pri = in.pr[i];
    const int deg = min(in.cnt[i], min(in.w, kMaxNnNeighbors));
    EdgeLane el;
    load_edge_lane<BC>(in, k, i, deg, xi, yi, vxi, vyi, pri, el);
    // the linear readout path is per edge: lane-parallel, reduced below
    float lx = 0.0f, ly = 0.0f;
    if (MODE == 1 && lane < deg) {
      float uf = 0.0f;
#pragma unroll
      for (int q = 0; q < kEdgeF; ++q) uf += u[q] * el.f[q];
      lx = el.dx * uf;
      ly = el.dy * uf;
    }
    const float2 p = *reinterpret_cast<const float2*>(P + (int64_t)i * W + c0);
    float s0 = 0.0f, s1 = 0.0f, tx0 = 0.0f, ty0 = 0.0f, tx1 = 0.0f, ty1 = 0.0f;
    for (int e0 = 0; e0 < deg; e0 += 8) {
      // rows of 8 edges, loaded unconditionally (a missing edge's sender is the
      // node itself) and pinned so that they are in flight together
      float2 qv[8];
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int j = __builtin_amdgcn_readlane(el.j, min(e0 + v, 63));
        qv[v] = *reinterpret_cast<const float2*>(Q + (int64_t)j * W + c0);
      }
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        pin(qv[v].x);
        pin(qv[v].y);
      }
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int e = e0 + v;
        if (e >= deg) break;
        float z0 = p.x + qv[v].x + cb0, z1 = p.y + qv[v].y + cb1;
#pragma unroll
        for (int q = 0; q < kEdgeF; ++q) {
          const float fq = bcast(el.f[q], e);
          z0 += C0[q] * fq;
          z1 += C1[q] * fq;
        }
        const float a0 = silu_f(z0), a1 = silu_f(z1);
        if (MODE == 0) {
          s0 += a0;
          s1 += a1;
        } else {
          const float dx = bcast(el.dx, e), dy = bcast(el.dy, e);
          tx0 += dx * a0;
          ty0 += dy * a0;
          tx1 += dx * a1;
          ty1 += dy * a1;
        }
      }
    }
    if (MODE == 0) {
      const int64_t o = (int64_t)i * W + c0;
      if (S) *reinterpret_cast<float2*>(S + o) = make_float2(s0, s1);
      *reinterpret_cast<uint32_t*>(Sb + o) = (uint32_t)f2bf(s0) | ((uint32_t)f2bf(s1) << 16);
    } else {
      // sum_j d_ij (w . s_ij + b + u . f_ij): w-part per column, b and u parts per edge lane
      const float bd = lane < deg ? b : 0.0f;
      float px = w0 * tx0 + w1 * tx1 + bd * el.dx + lx;
      float py = w0 * ty0 + w1 * ty1 + bd * el.dy + ly;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        px += __shfl_xor(px, o);
        py += __shfl_xor(py, o);
      }
      if (lane == 0) {
        ax[i] = k.gx + sc * px;
        ay[i] = k.gy + sc * py;
      }
    }
  }
}

// --------------------------------------------------------- edge backward
// One wavefront per node k. Receiver part: dz_kj = g_kj * silu'(z_kj) over k's
// list gives dP_k and the C, c (and readout w, b, u) gradient partials; sender
// part: dQ_k = sum over receivers i with k in i's list of dz_ik, visited through
// k's own list (the lists are symmetric).
// MODE 0: g_kj = dS_k. MODE 1: g_kj = (a_s / h) (da_k . (r_k - r_j)) w.
template <int MODE, int BC>
__global__ __launch_bounds__(256) void k_edge_bwd(
    NodeIn in, NnConsts k, const float* __restrict__ P, const float* __restrict__ Q,
    const uint16_t* __restrict__ Cb, const float* __restrict__ cb,
    const uint16_t* __restrict__ wvec, const float* __restrict__ G, float* __restrict__ dP,
    float* __restrict__ dQ, uint16_t* __restrict__ dPb, uint16_t* __restrict__ dQb,
    float* __restrict__ partial) {
  __shared__ float s_red[4][kEdgePartial];
  const int lane = threadIdx.x & 63, wave = wave_id();
  const int c0 = 2 * lane;
  float C0[kEdgeF], C1[kEdgeF], gC0[kEdgeF], gC1[kEdgeF], gu[kEdgeF];
#pragma unroll
  for (int e = 0; e < kEdgeF; ++e) {
    C0[e] = bf2f(Cb[c0 * kEdgeF + e]);
    C1[e] = bf2f(Cb[(c0 + 1) * kEdgeF + e]);
    gC0[e] = 0.0f;
    gC1[e] = 0.0f;
    gu[e] = 0.0f;
  }
  const float cb0 = cb[c0], cb1 = cb[c0 + 1];
  float w0 = 0.0f, w1 = 0.0f;
  if (MODE == 1) {
    w0 = bf2f(wvec[c0]);
    w1 = bf2f(wvec[c0 + 1]);
  }
  float gc0 = 0.0f, gc1 = 0.0f, gw0 = 0.0f, gw1 = 0.0f, gb = 0.0f;
  const float sc = k.a_s * k.inv_h;
  const int base = (xcd_tile(blockIdx.x, gridDim.x) * 4 + wave) * kNodesPerWave;
  for (int t = 0; t < kNodesPerWave; ++t) {
    const int n0 = base + t;
    if (n0 >= in.n) break;
    const float xk = in.rx[n0], yk = in.ry[n0], vxk = in.vx[n0], vyk = in.vy[n0],
                rhok = in.rho[n0], prk = in.pr[n0];
    const int deg = min(in.cnt[n0], min(in.w, kMaxNnNeighbors));
    // lane e: edge k <- j_e (receiver k) and the reverse edge j_e <- k
    EdgeLane el;
    load_edge_lane<BC>(in, k, n0, deg, xk, yk, vxk, vyk, prk, el);
    float rf[kEdgeF], rdx = 0.0f, rdy = 0.0f, dphi_k = 0.0f, dphi_r = 0.0f;
#pragma unroll
    for (int q = 0; q < kEdgeF; ++q) rf[q] = 0.0f;
    {
      // unconditional loads (el.j is the node itself past the degree), masked
      const bool ok = lane < deg;
      const int j = el.j;
      const float xj = in.rx[j], yj = in.ry[j], vxj = in.vx[j], vyj = in.vy[j], prj = in.pr[j];
      float gjx = 0.0f, gjy = 0.0f, gkx = 0.0f, gky = 0.0f;
      if (MODE == 1) {
        gkx = G[2 * (int64_t)n0];
        gky = G[2 * (int64_t)n0 + 1];
        gjx = G[2 * (int64_t)j];
        gjy = G[2 * (int64_t)j + 1];
      }
      edge_feat<BC>(k, xj, yj, vxj, vyj, prj, xk, yk, vxk, vyk, rhok, prk, rdx, rdy, rf);
      if (!ok) {
        rdx = rdy = 0.0f;
#pragma unroll
        for (int q = 0; q < kEdgeF; ++q) rf[q] = 0.0f;
      }
      if (MODE == 1 && ok) {
        dphi_k = sc * (gkx * el.dx + gky * el.dy);
        dphi_r = sc * (gjx * rdx + gjy * rdy);
#pragma unroll
        for (int q = 0; q < kEdgeF; ++q) gu[q] += dphi_k * el.f[q];
        gb += dphi_k;
      }
    }
    const float2 pk = *reinterpret_cast<const float2*>(P + (int64_t)n0 * W + c0);
    const float2 qk = *reinterpret_cast<const float2*>(Q + (int64_t)n0 * W + c0);
    float2 gk = make_float2(0.0f, 0.0f);
    if (MODE == 0) gk = *reinterpret_cast<const float2*>(G + (int64_t)n0 * W + c0);
    float dp0 = 0.0f, dp1 = 0.0f, dq0 = 0.0f, dq1 = 0.0f;
    for (int e0 = 0; e0 < deg; e0 += 8) {
      float2 qj[8], pj[8], gj[8];
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int j = __builtin_amdgcn_readlane(el.j, min(e0 + v, 63));
        const int64_t o = (int64_t)j * W + c0;
        qj[v] = *reinterpret_cast<const float2*>(Q + o);
        pj[v] = *reinterpret_cast<const float2*>(P + o);
        gj[v] = MODE == 0 ? *reinterpret_cast<const float2*>(G + o) : make_float2(0.0f, 0.0f);
      }
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        pin(qj[v].x);
        pin(qj[v].y);
        pin(pj[v].x);
        pin(pj[v].y);
        if (MODE == 0) {
          pin(gj[v].x);
          pin(gj[v].y);
        }
      }
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int e = e0 + v;
        if (e >= deg) break;
        // k receives from j
        {
          float z0 = pk.x + qj[v].x + cb0, z1 = pk.y + qj[v].y + cb1;
          float fq[kEdgeF];
#pragma unroll
          for (int q = 0; q < kEdgeF; ++q) {
            fq[q] = bcast(el.f[q], e);
            z0 += C0[q] * fq[q];
            z1 += C1[q] * fq[q];
          }
          float g0, g1;
          if (MODE == 0) {
            g0 = gk.x;
            g1 = gk.y;
          } else {
            const float dphi = bcast(dphi_k, e);
            gw0 += dphi * silu_f(z0);
            gw1 += dphi * silu_f(z1);
            g0 = dphi * w0;
            g1 = dphi * w1;
          }
          const float dz0 = g0 * dsilu_f(z0), dz1 = g1 * dsilu_f(z1);
          dp0 += dz0;
          dp1 += dz1;
          gc0 += dz0;
          gc1 += dz1;
#pragma unroll
          for (int q = 0; q < kEdgeF; ++q) {
            gC0[q] += dz0 * fq[q];
            gC1[q] += dz1 * fq[q];
          }
        }
        // j receives from k
        {
          float z0 = pj[v].x + qk.x + cb0, z1 = pj[v].y + qk.y + cb1;
#pragma unroll
          for (int q = 0; q < kEdgeF; ++q) {
            const float fq = bcast(rf[q], e);
            z0 += C0[q] * fq;
            z1 += C1[q] * fq;
          }
          float g0, g1;
          if (MODE == 0) {
            g0 = gj[v].x;
            g1 = gj[v].y;
          } else {
            const float dphi = bcast(dphi_r, e);
            g0 = dphi * w0;
            g1 = dphi * w1;
          }
          dq0 += g0 * dsilu_f(z0);
          dq1 += g1 * dsilu_f(z1);
        }
      }
    }
    const int64_t o = (int64_t)n0 * W + c0;
    *reinterpret_cast<float2*>(dP + o) = make_float2(dp0, dp1);
    *reinterpret_cast<float2*>(dQ + o) = make_float2(dq0, dq1);
    *reinterpret_cast<uint32_t*>(dPb + o) = (uint32_t)f2bf(dp0) | ((uint32_t)f2bf(dp1) << 16);
    *reinterpret_cast<uint32_t*>(dQb + o) = (uint32_t)f2bf(dq0) | ((uint32_t)f2bf(dq1) << 16);
  }
  // workgroup partials of dC, dc, dw, db, du (fixed order over lanes and waves)
  if (MODE == 1) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      gb += __shfl_xor(gb, o);
#pragma unroll
      for (int q = 0; q < kEdgeF; ++q) gu[q] += __shfl_xor(gu[q], o);
    }
  }
  float* r = s_red[wave];
#pragma unroll
  for (int q = 0; q < kEdgeF; ++q) {
    r[c0 * kEdgeF + q] = gC0[q];
    r[(c0 + 1) * kEdgeF + q] = gC1[q];
  }
  r[W * kEdgeF + c0] = gc0;
  r[W * kEdgeF + c0 + 1] = gc1;
  r[W * kEdgeF + W + c0] = gw0;
  r[W * kEdgeF + W + c0 + 1] = gw1;
  if (lane == 0) {
    r[W * kEdgeF + 2 * W] = gb;
#pragma unroll
    for (int q = 0; q < kEdgeF; ++q) r[W * kEdgeF + 2 * W + 1 + q] = gu[q];
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < kEdgePartial; idx += 256)
    partial[(int64_t)blockIdx.x * kEdgePartial + idx] =
        ((s_red[0][idx] + s_red[1][idx]) + s_red[2][idx]) + s_red[3][idx];
}

// ------------------------------------------------------------ loss / error
// Per-block FP64 sums of |a - t|^2 and |t|^2.
__global__ __launch_bounds__(256) void k_loss(int n, const float* __restrict__ ax,
                                              const float* __restrict__ ay,
                                              const float* __restrict__ tx,
                                              const float* __restrict__ ty,
                                              double* __restrict__ part) {
  __shared__ double s0[256], s1[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double e2 = 0.0, t2 = 0.0;
  if (i < n) {
    const float ex = ax[i] - tx[i], ey = ay[i] - ty[i];
    e2 = (double)ex * ex + (double)ey * ey;
    t2 = (double)tx[i] * tx[i] + (double)ty[i] * ty[i];
  }
  s0[threadIdx.x] = e2;
  s1[threadIdx.x] = t2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s0[threadIdx.x] += s0[threadIdx.x + o];
      s1[threadIdx.x] += s1[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0[0];
    part[2 * blockIdx.x + 1] = s1[0];
  }
}

// da = 2 (a - t) / sum |t|^2 (the gradient of the normalised loss), interleaved.
__global__ __launch_bounds__(256) void k_loss_grad(int n, const float* __restrict__ ax,
                                                   const float* __restrict__ ay,
                                                   const float* __restrict__ tx,
                                                   const float* __restrict__ ty,
                                                   const double* __restrict__ sums,
                                                   float* __restrict__ da) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float inv = sums[1] > 0.0 ? (float)(2.0 / sums[1]) : 0.0f;
  da[2 * (int64_t)i] = (ax[i] - tx[i]) * inv;
  da[2 * (int64_t)i + 1] = (ay[i] - ty[i]) * inv;
}

__global__ __launch_bounds__(256) void k_sum2(int parts, const double* __restrict__ part,
                                              double* __restrict__ out) {
  __shared__ double s0[256], s1[256];
  double a = 0.0, b = 0.0;
  for (int p = threadIdx.x; p < parts; p += 256) {
    a += part[2 * p];
    b += part[2 * p + 1];
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s0[threadIdx.x] += s0[threadIdx.x + o];
      s1[threadIdx.x] += s1[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s0[0];
    out[1] = s1[0];
  }
}

// ------------------------------------------------------- SPH-side helpers
template <int BC>
__global__ __launch_bounds__(256) void k_integrate_keys(
    int n, DevGrid g, DevPhys c, const float* __restrict__ rx, const float* __restrict__ ry,
    const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ ax,
    const float* __restrict__ ay, float* __restrict__ tx, float* __restrict__ ty,
    float* __restrict__ tvx, float* __restrict__ tvy, uint64_t* __restrict__ keys) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float x = rx[i], y = ry[i], u = vx[i], v = vy[i];
  integrate_one<BC>(c, x, y, u, v, ax[i], ay[i]);
  tx[i] = x;
  ty[i] = y;
  tvx[i] = u;
  tvy[i] = v;
  keys[i] = cell_key(g, cell_coord(x, g.lo_x, g.cs_x, g.nx), cell_coord(y, g.lo_y, g.cs_y, g.ny));
}

__global__ __launch_bounds__(256) void k_export(int n, const uint32_t* __restrict__ pid,
                                                const float* __restrict__ rx,
                                                const float* __restrict__ ry,
                                                const float* __restrict__ vx,
                                                const float* __restrict__ vy,
                                                float* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = min(pid[i], (uint32_t)(n - 1));
  out[p] = rx[i];
  out[(int64_t)n + p] = ry[i];
  out[2 * (int64_t)n + p] = vx[i];
  out[3 * (int64_t)n + p] = vy[i];
}

// ------------------------------------------------------------- host side
inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

template <class T>
hipError_t nn_alloc(Net* net, T** p, int64_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, (size_t)std::max<int64_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return e;
  net->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return hipSuccess;
}

enum { T_ENC_W1 = 0, T_ENC_B1, T_ENC_W2, T_ENC_B2, T_LAYER0 };
enum { L_A = 0, L_B, L_C, L_c, L_D, L_d, L_E, L_e, L_F, L_f, L_COUNT };
enum { R_A = 0, R_B, R_C, R_c, R_w, R_b, R_u };

const NnTensor& ten(const Net* n, int idx) { return n->t[idx]; }
const NnTensor& lt(const Net* n, int l, int which) { return n->t[T_LAYER0 + l * L_COUNT + which]; }
const NnTensor& rt(const Net* n, int which) { return n->t[T_LAYER0 + n->L * L_COUNT + which]; }
const float* P_(const Net* n, const NnTensor& t) { return n->theta + t.off; }
float* G_(const Net* n, const NnTensor& t) { return n->grad + t.off; }
const uint16_t* WB(const Net* n, const NnTensor& t) { return n->wb + t.boff; }
const uint16_t* WBT(const Net* n, const NnTensor& t) {
  return n->wb + t.boff + (int64_t)t.rows * t.kp;
}

void gemm(hipStream_t st, int n, const uint16_t* X1, int K1, const uint16_t* X2, int K2,
          const uint16_t* B, int Mo, const float* bias, const float* rowscale, const float* bias2,
          const float* R, float* Yf, uint16_t* Yb, int act) {
  GemmArgs a{n, K1, K2, Mo, X1, X2, B, bias, rowscale, bias2, R, Yf, Yb, act};
  // up to 8 workgroups per CU, each walking a few row tiles
  const dim3 grid(std::min(cdiv(n, 128), 2048), Mo / 128);
  switch (K1 + K2) {
    case kNodeK: k_gemm<kNodeK><<<grid, 256, 0, st>>>(a); break;
    case W: k_gemm<W><<<grid, 256, 0, st>>>(a); break;
    case 2 * W: k_gemm<2 * W><<<grid, 256, 0, st>>>(a); break;
    default: break;  // the network uses no other K
  }
}

// weight gradient of tensor t (rows Mo, padded K kp = K1 + K2) into grad
void wgrad(Net* net, hipStream_t st, int n, const uint16_t* dY, int Mo, const uint16_t* X1, int K1,
           const uint16_t* X2, int K2, const NnTensor& t) {
  const int chunks = cdiv(n, kRowsPerChunk);
  const int K = K1 + K2;
  WgradArgs a{n, Mo, K1, K2, dY, X1, X2, net->partial};
  k_wgrad<<<dim3(chunks, Mo / 128, cdiv(K, 128)), 256, 0, st>>>(a);
  launch_reduce(st, chunks, Mo, K, t.cols, net->partial, G_(net, t));
}

void colsum(Net* net, hipStream_t st, int n, int Mo, const float* Y, const float* rowscale,
            const NnTensor& t) {
  const int chunks = cdiv(n, kColRows);
  k_colsum<<<chunks, 256, 0, st>>>(n, Mo, Y, rowscale, net->partial);
  launch_reduce(st, chunks, 1, Mo, Mo, net->partial, G_(net, t));
}

NodeIn node_in(const DevState& s) {
  return NodeIn{s.rx, s.ry, s.vx, s.vy, s.rho, s.pr, s.ell_count, s.ell_idx, s.n, s.ell_width};
}

template <int MODE>
void edge_fwd(const Net* net, const DevState& s, hipStream_t st, const float* P, const float* Q,
              const NnTensor& C, const NnTensor& c, const NnTensor* w, const NnTensor* b,
              const NnTensor* u, float* S, uint16_t* Sb, float* ax, float* ay) {
  const NodeIn in = node_in(s);
  const int blocks = cdiv(s.n, kEdgeNodesPerBlock);
  const uint16_t* wv = w ? WB(net, *w) : nullptr;
  const float* bv = b ? P_(net, *b) : nullptr;
  const float* uv = u ? P_(net, *u) : nullptr;
  if (net->k.periodic)
    k_edge_fwd<MODE, 1><<<blocks, 256, 0, st>>>(in, net->k, P, Q, WB(net, C), P_(net, c), wv, bv,
                                                uv, S, Sb, ax, ay);
  else
    k_edge_fwd<MODE, 0><<<blocks, 256, 0, st>>>(in, net->k, P, Q, WB(net, C), P_(net, c), wv, bv,
                                                uv, S, Sb, ax, ay);
}

// edge backward + reduction of its partials into the C, c (, w, b) gradients
template <int MODE>
void edge_bwd(Net* net, const DevState& s, hipStream_t st, const float* P, const float* Q,
              const NnTensor& C, const NnTensor& c, const NnTensor* w, const NnTensor* b,
              const NnTensor* u, const float* G) {
  const NodeIn in = node_in(s);
  const int blocks = cdiv(s.n, kEdgeNodesPerBlock);
  const uint16_t* wv = w ? WB(net, *w) : nullptr;
  if (net->k.periodic)
    k_edge_bwd<MODE, 1><<<blocks, 256, 0, st>>>(in, net->k, P, Q, WB(net, C), P_(net, c), wv, G,
                                                net->dP, net->dQ, net->dPb, net->dQb,
                                                net->partial);
  else
    k_edge_bwd<MODE, 0><<<blocks, 256, 0, st>>>(in, net->k, P, Q, WB(net, C), P_(net, c), wv, G,
                                                net->dP, net->dQ, net->dPb, net->dQb,
                                                net->partial);
  // partial rows: [dC (W x 8) | dc (W) | dw (W) | db | du (8)]
  const int stride = kEdgePartial;
  launch_reduce(st, blocks, 1, stride, W * kEdgeF, net->partial, G_(net, C));
  launch_reduce(st, blocks, 1, stride, W, net->partial + W * kEdgeF, G_(net, c));
  if (w) {
    launch_reduce(st, blocks, 1, stride, W, net->partial + W * kEdgeF + W, G_(net, *w));
    launch_reduce(st, blocks, 1, stride, 1, net->partial + W * kEdgeF + 2 * W, G_(net, *b));
    launch_reduce(st, blocks, 1, stride, kEdgeF, net->partial + W * kEdgeF + 2 * W + 1,
                  G_(net, *u));
  }
}

}  // namespace

int nn_create(Net* net, int layers, bool train, int64_t n_cap, const NnConsts& k) {
  net->L = layers;
  net->train = train;
  net->n_cap = n_cap;
  net->k = k;
  net->t.clear();
  int64_t off = 0, boff = 0;
  auto add = [&](int rows, int cols, int kp) {
    NnTensor t{off, rows, cols, kp, kp ? boff : -1};
    off += (int64_t)rows * cols;
    if (kp) boff += 2 * (int64_t)rows * kp;
    net->t.push_back(t);
  };
  add(W, kNodeF, kNodeK);  // enc_w1
  add(W, 1, 0);
  add(W, W, W);            // enc_w2
  add(W, 1, 0);
  for (int l = 0; l < layers; ++l) {
    add(W, W, W);            // A
    add(W, W, W);            // B
    add(W, kEdgeF, kEdgeF);  // C (bf16 copy read by the edge kernels)
    add(W, 1, 0);            // c
    add(W, W, W);            // D
    add(W, 1, 0);            // d
    add(W, 2 * W, 2 * W);    // E
    add(W, 1, 0);            // e
    add(W, W, W);            // F
    add(W, 1, 0);            // f
  }
  add(W, W, W);            // ro_A
  add(W, W, W);            // ro_B
  add(W, kEdgeF, kEdgeF);  // ro_C
  add(W, 1, 0);            // ro_c
  add(1, W, W);            // ro_w
  add(1, 1, 0);            // ro_b
  add(kEdgeF, 1, 0);       // ro_u (FP32: the linear path of the readout)
  net->n_params = off;
  net->n_wb = boff;
  const int64_t N = n_cap, NW = N * W;
  const int nl = train ? layers + 1 : 1;
  hipError_t e = hipSuccess;
  auto A = [&](auto** p, int64_t count) {
    if (e == hipSuccess) e = nn_alloc(net, p, count);
  };
  A(&net->theta, off);
  A(&net->grad, off);
  A(&net->m, off);
  A(&net->v, off);
  A(&net->wb, boff);
  A(&net->x, N * kNodeK);
  A(&net->deg, N);
  A(&net->pre1, NW);
  A(&net->a1b, NW);
  net->lb.assign(nl, NnLayerBufs{});
  for (auto& b : net->lb) {
    A(&b.h, NW); A(&b.P, NW); A(&b.Q, NW);
    if (train) A(&b.preU, NW);
    A(&b.hb, NW); A(&b.Sb, NW); A(&b.Mb, NW); A(&b.Ub, NW);
  }
  A(&net->a_out, 2 * N);
  if (train) {
    A(&net->dh, NW); A(&net->dP, NW); A(&net->dQ, NW); A(&net->dS, NW);
    A(&net->dpre, NW);
    A(&net->dhb, NW); A(&net->dPb, NW); A(&net->dQb, NW); A(&net->dMb, NW); A(&net->dpreb, NW);
    A(&net->da, 2 * N);
    const int64_t chunks = cdiv(N, kRowsPerChunk);
    const int64_t eb = cdiv(N, kEdgeNodesPerBlock);
    net->partial_cap = std::max({(int64_t)chunks * W * 2 * W, (int64_t)eb * kEdgePartial,
                                 (int64_t)cdiv(N, kColRows) * 2 * W});
    A(&net->partial, net->partial_cap);
  }
  net->red_cap = 2 * (int64_t)cdiv(N, 256);
  A(&net->red, net->red_cap);
  A(&net->red_out, 4);
  if (train) A(&net->gnorm, kNormBlocks + 1);
  if (e == hipSuccess) e = hipMemset(net->m, 0, off * sizeof(float));
  if (e == hipSuccess) e = hipMemset(net->v, 0, off * sizeof(float));
  if (e == hipSuccess) e = hipMemset(net->grad, 0, off * sizeof(float));
  if (e == hipSuccess) e = hipMemset(net->theta, 0, off * sizeof(float));
  if (e != hipSuccess) {
    net->err = std::string("network allocation: ") + hipGetErrorString(e);
    return -1;
  }
  return 0;
}

void nn_free(Net* net) {
  for (void* p : net->allocs) (void)hipFree(p);
  net->allocs.clear();
}

hipError_t nn_sync_weights(Net* net, hipStream_t st) {
  for (const NnTensor& t : net->t) {
    if (!t.kp) continue;
    k_wcopy<<<cdiv((int64_t)t.rows * t.kp, 256), 256, 0, st>>>(
        t.rows, t.cols, t.kp, net->theta + t.off, net->wb + t.boff,
        net->wb + t.boff + (int64_t)t.rows * t.kp);
  }
  return hipGetLastError();
}

hipError_t nn_forward(Net* net, const DevState& s, hipStream_t st, float* ax, float* ay) {
  const int n = s.n;
  if (n <= 0) return hipSuccess;
  const NodeIn in = node_in(s);
  k_node_features<<<cdiv(n, 256), 256, 0, st>>>(in, net->k, net->x, net->deg);
  gemm(st, n, net->x, kNodeK, nullptr, 0, WB(net, ten(net, T_ENC_W1)), W,
       P_(net, ten(net, T_ENC_B1)), nullptr, nullptr, nullptr, net->train ? net->pre1 : nullptr,
       net->a1b, 1);
  NnLayerBufs& b0 = net->lb[0];
  gemm(st, n, net->a1b, W, nullptr, 0, WB(net, ten(net, T_ENC_W2)), W,
       P_(net, ten(net, T_ENC_B2)), nullptr, nullptr, nullptr, b0.h, b0.hb, 0);
  for (int l = 0; l < net->L; ++l) {
    NnLayerBufs& b = net->lb[net->train ? l : 0];
    NnLayerBufs& nb = net->lb[net->train ? l + 1 : 0];
    gemm(st, n, b.hb, W, nullptr, 0, WB(net, lt(net, l, L_A)), W, nullptr, nullptr, nullptr,
         nullptr, b.P, nullptr, 0);
    gemm(st, n, b.hb, W, nullptr, 0, WB(net, lt(net, l, L_B)), W, nullptr, nullptr, nullptr,
         nullptr, b.Q, nullptr, 0);
    // only the bf16 copies of S and M are read later (forward and backward)
    edge_fwd<0>(net, s, st, b.P, b.Q, lt(net, l, L_C), lt(net, l, L_c), nullptr, nullptr,
                nullptr, nullptr, b.Sb, nullptr, nullptr);
    gemm(st, n, b.Sb, W, nullptr, 0, WB(net, lt(net, l, L_D)), W, nullptr, net->deg,
         P_(net, lt(net, l, L_d)), nullptr, nullptr, b.Mb, 0);
    gemm(st, n, b.hb, W, b.Mb, W, WB(net, lt(net, l, L_E)), W, P_(net, lt(net, l, L_e)), nullptr,
         nullptr, nullptr, net->train ? b.preU : nullptr, b.Ub, 1);
    // h_{l+1} = h_l + F U + f (in place when activations are not kept; the
    // readout reads only the bf16 copy, so inference skips the last FP32 h)
    const bool keep_h = net->train || l + 1 < net->L;
    gemm(st, n, b.Ub, W, nullptr, 0, WB(net, lt(net, l, L_F)), W, P_(net, lt(net, l, L_f)),
         nullptr, nullptr, b.h, keep_h ? nb.h : nullptr, nb.hb, 0);
  }
  NnLayerBufs& f = net->lb[net->train ? net->L : 0];
  gemm(st, n, f.hb, W, nullptr, 0, WB(net, rt(net, R_A)), W, nullptr, nullptr, nullptr, nullptr,
       f.P, nullptr, 0);
  gemm(st, n, f.hb, W, nullptr, 0, WB(net, rt(net, R_B)), W, nullptr, nullptr, nullptr, nullptr,
       f.Q, nullptr, 0);
  const NnTensor& rw = rt(net, R_w);
  const NnTensor& rb = rt(net, R_b);
  const NnTensor& ru = rt(net, R_u);
  edge_fwd<1>(net, s, st, f.P, f.Q, rt(net, R_C), rt(net, R_c), &rw, &rb, &ru, nullptr, nullptr,
              ax, ay);
  return hipGetLastError();
}

hipError_t nn_error_sums(Net* net, int n, const float* ax, const float* ay, const float* tx,
                         const float* ty, hipStream_t st, double* out2) {
  const int blocks = cdiv(n, 256);
  k_loss<<<blocks, 256, 0, st>>>(n, ax, ay, tx, ty, net->red);
  k_sum2<<<1, 256, 0, st>>>(blocks, net->red, out2);
  return hipGetLastError();
}

hipError_t nn_loss_backward(Net* net, const DevState& s, hipStream_t st, const float* tax,
                            const float* tay) {
  const int n = s.n;
  if (!net->train) return hipErrorInvalidValue;
  hipError_t e = nn_forward(net, s, st, net->a_out, net->a_out + net->n_cap);
  if (e != hipSuccess) return e;
  const float* ax = net->a_out;
  const float* ay = net->a_out + net->n_cap;
  const int blocks = cdiv(n, 256);
  k_loss<<<blocks, 256, 0, st>>>(n, ax, ay, tax, tay, net->red);
  k_sum2<<<1, 256, 0, st>>>(blocks, net->red, net->red_out);
  k_loss_grad<<<blocks, 256, 0, st>>>(n, ax, ay, tax, tay, net->red_out, net->da);
  const int L = net->L;
  // readout
  NnLayerBufs& f = net->lb[L];
  const NnTensor& rw = rt(net, R_w);
  const NnTensor& rb = rt(net, R_b);
  const NnTensor& ru = rt(net, R_u);
  edge_bwd<1>(net, s, st, f.P, f.Q, rt(net, R_C), rt(net, R_c), &rw, &rb, &ru, net->da);
  wgrad(net, st, n, net->dPb, W, f.hb, W, nullptr, 0, rt(net, R_A));
  wgrad(net, st, n, net->dQb, W, f.hb, W, nullptr, 0, rt(net, R_B));
  gemm(st, n, net->dPb, W, nullptr, 0, WBT(net, rt(net, R_A)), W, nullptr, nullptr, nullptr,
       nullptr, net->dh, nullptr, 0);
  gemm(st, n, net->dQb, W, nullptr, 0, WBT(net, rt(net, R_B)), W, nullptr, nullptr, nullptr,
       net->dh, net->dh, net->dhb, 0);
  for (int l = L - 1; l >= 0; --l) {
    NnLayerBufs& b = net->lb[l];
    // h_{l+1} = h_l + F U + f
    wgrad(net, st, n, net->dhb, W, b.Ub, W, nullptr, 0, lt(net, l, L_F));
    colsum(net, st, n, W, net->dh, nullptr, lt(net, l, L_f));
    gemm(st, n, net->dhb, W, nullptr, 0, WBT(net, lt(net, l, L_F)), W, nullptr, nullptr, nullptr,
         nullptr, net->dpre, nullptr, 0);
    k_silu_bwd<<<cdiv((int64_t)n * W, 256), 256, 0, st>>>((int64_t)n * W, net->dpre, b.preU,
                                                          net->dpre, net->dpreb);
    // U = silu(E [h, M] + e)
    wgrad(net, st, n, net->dpreb, W, b.hb, W, b.Mb, W, lt(net, l, L_E));
    colsum(net, st, n, W, net->dpre, nullptr, lt(net, l, L_e));
    const uint16_t* Et = WBT(net, lt(net, l, L_E));  // [2W x W]: rows 0..W-1 -> h, W.. -> M
    gemm(st, n, net->dpreb, W, nullptr, 0, Et, W, nullptr, nullptr, nullptr, net->dh, net->dh,
         nullptr, 0);
    gemm(st, n, net->dpreb, W, nullptr, 0, Et + (int64_t)W * W, W, nullptr, nullptr, nullptr,
         nullptr, net->dS, net->dMb, 0);
    // M = D S + deg d  (dS holds dM here)
    wgrad(net, st, n, net->dMb, W, b.Sb, W, nullptr, 0, lt(net, l, L_D));
    colsum(net, st, n, W, net->dS, net->deg, lt(net, l, L_d));
    gemm(st, n, net->dMb, W, nullptr, 0, WBT(net, lt(net, l, L_D)), W, nullptr, nullptr, nullptr,
         nullptr, net->dS, nullptr, 0);
    // S = sum silu(P_i + Q_j + C f + c)
    edge_bwd<0>(net, s, st, b.P, b.Q, lt(net, l, L_C), lt(net, l, L_c), nullptr, nullptr,
                nullptr, net->dS);
    wgrad(net, st, n, net->dPb, W, b.hb, W, nullptr, 0, lt(net, l, L_A));
    wgrad(net, st, n, net->dQb, W, b.hb, W, nullptr, 0, lt(net, l, L_B));
    gemm(st, n, net->dPb, W, nullptr, 0, WBT(net, lt(net, l, L_A)), W, nullptr, nullptr, nullptr,
         net->dh, net->dh, nullptr, 0);
    gemm(st, n, net->dQb, W, nullptr, 0, WBT(net, lt(net, l, L_B)), W, nullptr, nullptr, nullptr,
         net->dh, net->dh, net->dhb, 0);
  }
  // encoder
  wgrad(net, st, n, net->dhb, W, net->a1b, W, nullptr, 0, ten(net, T_ENC_W2));
  colsum(net, st, n, W, net->dh, nullptr, ten(net, T_ENC_B2));
  gemm(st, n, net->dhb, W, nullptr, 0, WBT(net, ten(net, T_ENC_W2)), W, nullptr, nullptr, nullptr,
       nullptr, net->dpre, nullptr, 0);
  k_silu_bwd<<<cdiv((int64_t)n * W, 256), 256, 0, st>>>((int64_t)n * W, net->dpre, net->pre1,
                                                        net->dpre, net->dpreb);
  wgrad(net, st, n, net->dpreb, W, net->x, kNodeK, nullptr, 0, ten(net, T_ENC_W1));
  colsum(net, st, n, W, net->dpre, nullptr, ten(net, T_ENC_B1));
  return hipGetLastError();
}

hipError_t nn_adam(Net* net, hipStream_t st, float lr, float b1, float b2, float eps) {
  net->adam_t += 1;
  const float c1 = 1.0f - (float)std::pow((double)b1, net->adam_t);
  const float c2 = 1.0f - (float)std::pow((double)b2, net->adam_t);
  const double* g2 = nullptr;
  if (net->clip > 0.0f) {
    k_sq_partial<<<kNormBlocks, 256, 0, st>>>(net->n_params, net->grad, net->gnorm);
    k_sq_final<<<1, 256, 0, st>>>(kNormBlocks, net->gnorm, net->gnorm + kNormBlocks);
    g2 = net->gnorm + kNormBlocks;
  }
  k_adam<<<cdiv(net->n_params, 256), 256, 0, st>>>(net->n_params, net->theta, net->grad, net->m,
                                                   net->v, lr, b1, b2, eps, c1, c2, g2,
                                                   net->clip);
  return nn_sync_weights(net, st);
}

hipError_t launch_integrate_keys(const Launch& L, const DevState& s, const float* ax,
                                 const float* ay, uint64_t* keys_out) {
  if (s.n <= 0) return hipSuccess;
  if (L.phys.boundary == 1)
    k_integrate_keys<1><<<cdiv(s.n, 256), 256, 0, L.stream>>>(
        s.n, L.grid, L.phys, s.rx, s.ry, s.vx, s.vy, ax, ay, s.tx, s.ty, s.tvx, s.tvy, keys_out);
  else
    k_integrate_keys<0><<<cdiv(s.n, 256), 256, 0, L.stream>>>(
        s.n, L.grid, L.phys, s.rx, s.ry, s.vx, s.vy, ax, ay, s.tx, s.ty, s.tvx, s.tvy, keys_out);
  return hipGetLastError();
}

hipError_t launch_export_state(const Launch& L, const DevState& s, float* out) {
  if (s.n <= 0) return hipSuccess;
  k_export<<<cdiv(s.n, 256), 256, 0, L.stream>>>(s.n, s.pid, s.rx, s.ry, s.vx, s.vy, out);
  return hipGetLastError();
}

}  // namespace sphamd
