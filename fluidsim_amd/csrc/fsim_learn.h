// fsim_learn.h -- the learned Poisson preconditioner: internal interface
// between the solver (fsim_api.hip) and the network (fsim_learn.hip).
//
// The network is a 4-level convolutional U-Net (16/32/48/64 channels, 3x3
// kernels, no biases, ReLU) with BF16 weights and activations and FP32
// accumulation on the MFMA units.  It is the coarse solver of the multigrid
// preconditioner inside flexible CG: the levels down to the first one of at
// most 64 x 64 cells are smoothed and restricted as in the multigrid cycle;
// the network maps that level's residual (plus its operator's coupling ratios)
// to its correction, which is prolongated and post-smoothed on the way up.
// Activations are NHWC with the channel count a multiple of 16.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fsim_internal.h"

namespace fsl {

constexpr int kLayers = 14;      // 3x3 convolutions
constexpr int kInCh = 5;         // r/(s d), cx_left/d, cx_right/d, cy_below/d, cy_above/d
constexpr int kPadIn = 16;       // input channels as stored
constexpr int kOutCh = 16;       // channels feeding the 1x1 output convolution
constexpr int kAlign = 8;        // network grid sides are multiples of 8 (three poolings)

enum InMode { kPlain = 0, kPool = 1, kUpCat = 2 };

// Layer table: input mode, input channels (a, b for the concatenation), output
// channels, resolution level.
struct LayerDesc {
  int mode, ca, cb, cout, level;
};
extern const LayerDesc kNet[kLayers];

// Activations and gradients of one batch at one resolution.
struct Acts {
  int N = 0, H = 0, W = 0;          // level-0 size of the network grid
  uint16_t* y[kLayers] = {};        // layer outputs (post-ReLU, bf16)
  uint16_t* dpre[kLayers] = {};     // gradients at the layer pre-activations (bf16)
  float* dx = nullptr;              // fp32 scratch for input gradients (largest layer)
  float* dx2 = nullptr;             // second scratch (pooling branch)
  uint16_t* in = nullptr;           // network input [N][H][W][16]
  float* out = nullptr;             // e_hat [N][H][W]
  float* dout = nullptr;            // dL/de_hat
  float* rhat = nullptr;            // r / s [N][H][W]
  float* op = nullptr;              // d, cx, cy [N][3][H][W] (loss operator)
  float* scale = nullptr;           // s per sample
  size_t bytes = 0;
  uint64_t id = 0;                  // allocation number, unique in the process (0: none)
};

struct Net {
  // parameters: fp32 master copy, bf16 copies for the forward and (tap-flipped,
  // channel-transposed) for the input-gradient convolution, Adam moments
  float* w = nullptr;
  float* g = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  uint16_t* wb = nullptr;
  uint16_t* wt = nullptr;
  size_t off[kLayers + 1] = {};     // parameter offsets; [kLayers] = the 1x1 output weights
  size_t nparam = 0;
  float* wpart = nullptr;           // weight-gradient partial sums
  size_t wpart_len = 0;
  int adam_t = 0;
  // training set: samples of (r, d, cx, cy) on an H x W coarse grid, fp32
  float* data = nullptr;
  int cap = 0, count = 0, H = 0, W = 0;
  Acts train, infer;
  float* lpart = nullptr;           // loss partials
  hipStream_t st = nullptr;
  uint64_t rng = 0;
};

// network grid side for a coarse level side
inline int net_side(int n) { return (n + kAlign - 1) / kAlign * kAlign; }

int net_init(Net* net, int cap, int H, int W, uint64_t seed, hipStream_t st);
void net_free(Net* net);
int acts_alloc(Acts& a, int N, int H, int W);
void acts_free(Acts& a);

// Inference inside the solver: from the coarse level's residual, operator and
// the partial sums of r^2/d written by the finest-level down kernel (the
// scale), write the coarse correction into L.e.
int net_apply(Net* net, const MgLevel& L, const double* part, int nb, DevScalars* sc,
              hipStream_t st);
// Append the coarse level's current residual and operator as a training sample.
int net_store_sample(Net* net, const MgLevel& L, hipStream_t st);
// Regenerate the BF16 weight copies after the FP32 masters changed.
void net_refresh(Net* net);
// One Adam step on the samples idx[0..B) of the training set; mean loss.
int train_step(Net* net, const int* idx, int B, float lr, double* loss);
// Loss and gradient of a host batch (see fs_learn_loss_grad).
int loss_grad(Net* net, int B, int H, int W, const float* in5, const float* rhat, const float* op,
              double* loss, float* grad, float* out);

}  // namespace fsl

// The C-ABI handle of the learned preconditioner.
struct fs_learn {
  fsl::Net net;
};
