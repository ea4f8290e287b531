// Internal interface of the surrogate network (sph_nn.hip) used by the C-ABI
// layer: parameter storage, activations and the forward / backward / Adam
// pipelines over the fixed-width neighbor lists of a DevState.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "sph_internal.hpp"

namespace sphamd {

constexpr int kNnWidth = 128;  // hidden width (the kernels are written for 128)
constexpr int kNodeF = 8;      // node features (padded to kNodeK for the MFMA)
constexpr int kNodeK = 32;
constexpr int kEdgeF = 8;      // edge features

// Normalisation and physics constants of the features and the readout.
struct NnConsts {
  float h, inv_h, rho0, inv_rho0, rs, ps, inv_vs, a_s, gx, gy;
  float len_x, len_y, half_x, half_y;
  int periodic;
};

// One parameter tensor in the flat FP32 vector; matrices also have a bf16 copy
// (row-major, K padded to kp) and its transpose.
struct NnTensor {
  int64_t off;      // into theta / grad
  int rows, cols;   // cols = 1 for a vector
  int kp;           // padded K of the bf16 copies (0: vector, no copy)
  int64_t boff;     // into wb: rows x kp, then kp x rows (transpose)
};

struct NnLayerBufs {
  float *h, *P, *Q, *preU;      // fp32 [n x W] (preU only when training)
  uint16_t *hb, *Sb, *Mb, *Ub;  // bf16 activations
};

struct Net {
  int L = 3;
  bool train = true;
  int64_t n_cap = 0;
  NnConsts k{};
  std::vector<NnTensor> t;  // enc(4), L x 10, ro(6)
  int64_t n_params = 0, n_wb = 0;
  float *theta = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  uint16_t* wb = nullptr;
  int adam_t = 0;
  // activations
  uint16_t* x = nullptr;   // [n x kNodeK]
  float* deg = nullptr;    // [n]
  float* pre1 = nullptr;   // [n x W]
  uint16_t* a1b = nullptr;
  std::vector<NnLayerBufs> lb;  // L + 1 entries (the last holds the final h, Po, Qo)
  float* a_out = nullptr;       // [n x 2] (ax, then ay)
  // backward
  float *dh = nullptr, *dP = nullptr, *dQ = nullptr, *dS = nullptr, *dHM = nullptr,
        *dpre = nullptr;
  uint16_t *dhb = nullptr, *dPb = nullptr, *dQb = nullptr, *dMb = nullptr, *dpreb = nullptr;
  float* da = nullptr;          // [n x 2]
  float* partial = nullptr;     // reduction partials
  int64_t partial_cap = 0;
  double* red = nullptr;        // FP64 reduction scratch
  int64_t red_cap = 0;
  double* red_out = nullptr;    // 4 doubles
  double* gnorm = nullptr;      // gradient-norm partials + the squared norm (train only)
  float clip = 0.0f;            // global gradient-norm limit of Adam (0: none)
  std::vector<void*> allocs;
  std::string err;
  int device = 0;
};

int nn_create(Net* net, int layers, bool train, int64_t n_cap, const NnConsts& k);
void nn_free(Net* net);
// refresh the bf16 copies of every matrix from theta
hipError_t nn_sync_weights(Net* net, hipStream_t s);
// forward on the sorted state of s (fixed-width list valid); predicted
// accelerations into ax, ay (device, n each)
hipError_t nn_forward(Net* net, const DevState& s, hipStream_t st, float* ax, float* ay);
// loss against (tax, tay) and backward into net->grad; FP64 sums
// (sum |a - t|^2, sum |t|^2) into net->red_out[0..1]
hipError_t nn_loss_backward(Net* net, const DevState& s, hipStream_t st, const float* tax,
                            const float* tay);
hipError_t nn_adam(Net* net, hipStream_t st, float lr, float b1, float b2, float eps);
// FP64 sums (sum |a - t|^2, sum |t|^2) of two acceleration fields into out2 (device)
hipError_t nn_error_sums(Net* net, int n, const float* ax, const float* ay, const float* tx,
                         const float* ty, hipStream_t st, double* out2);
// integration from given accelerations into T plus the next cell keys
hipError_t launch_integrate_keys(const Launch& L, const DevState& s, const float* ax,
                                 const float* ay, uint64_t* keys_out);
// device-to-device scatter of S into particle-id order: out = [rx | ry | vx | vy]
hipError_t launch_export_state(const Launch& L, const DevState& s, float* out);

}  // namespace sphamd
