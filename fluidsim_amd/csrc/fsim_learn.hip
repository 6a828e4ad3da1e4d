// fsim_learn.hip -- the learned coarse solver: U-Net forward and backward on
// the MFMA units (BF16 operands, FP32 accumulation), Adam, the training set
// and the solver-side inference.
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x16_bf16: one wave
// computes 16 pixels x (16 * NT) output channels; the K loop runs over the 9
// taps and the input channels in blocks of 16 (all channel counts are
// multiples of 16).  Operand layouts of the instruction (checked on the
// device): lane l holds A[l % 16][4 (l / 16) + e] and B[4 (l / 16) + e][l % 16],
// e = 0..3, and receives D[4 (l / 16) + i][l % 16], i = 0..3.
//
// Rounding points (the CPU reference network follows them): the input, every
// layer output (after ReLU), every pooled value and every pre-activation
// gradient are rounded to BF16 (round to nearest even); products and sums are
// FP32; weights are FP32 masters with BF16 copies.
#include <hip/hip_runtime.h>

#include <atomic>
#include <math.h>
#include <string.h>

#include <vector>

#include "fsim.h"
#include "fsim_learn.h"

namespace fsl {

// 16 pixels per wave, 4 waves per block
const LayerDesc kNet[kLayers] = {
    {kPlain, kPadIn, 0, 16, 0},  // e0a
    {kPlain, 16, 0, 16, 0},      // e0b
    {kPool, 16, 0, 32, 1},       // e1a  (2x2 average of e0b)
    {kPlain, 32, 0, 32, 1},      // e1b
    {kPool, 32, 0, 48, 2},       // e2a
    {kPlain, 48, 0, 48, 2},      // e2b
    {kPool, 48, 0, 64, 3},       // e3a
    {kPlain, 64, 0, 64, 3},      // e3b
    {kUpCat, 64, 48, 48, 2},     // d2a  (upsampled e3b ++ e2b)
    {kPlain, 48, 0, 48, 2},      // d2b
    {kUpCat, 48, 32, 32, 1},     // d1a  (upsampled d2b ++ e1b)
    {kPlain, 32, 0, 32, 1},      // d1b
    {kUpCat, 32, 16, 16, 0},     // d0a  (upsampled d1b ++ e0b)
    {kPlain, 16, 0, 16, 0},      // d0b
};

namespace {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kLossSlots = 1 << 20;  // floats: loss partials, then the batch indices

int layer_cin(int l) { return kNet[l].mode == kUpCat ? kNet[l].ca + kNet[l].cb : kNet[l].ca; }
// the skip input of the decoder layers
int skip_src(int l) { return l == 8 ? 5 : l == 10 ? 3 : l == 12 ? 1 : -1; }

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

struct ConvArgs {
  const uint16_t* xa;    // plain input; fine input (pool); half-resolution input (up-concat)
  const uint16_t* xb;    // skip input (up-concat)
  int ca, cb, mode;
  const uint16_t* w;     // [9][cout][cin] bf16
  int cout;
  uint16_t* y;           // bf16 output [N][H][W][cout] (may be null)
  float* yf;             // fp32 output (may be null)
  const uint16_t* mask;  // output *= (mask > 0), mask with cout channels (may be null)
  int relu;
  int N, H, W;           // output resolution
};

// Four input channels c0..c0+3 of pixel (n, hh, ww) at the output resolution.
template <int MODE>
__device__ __forceinline__ bf4 conv_in4(const ConvArgs& A, int n, int hh, int ww, int c0) {
  if (MODE == kPlain)
    return *(const bf4*)(A.xa + (((size_t)n * A.H + hh) * A.W + ww) * A.ca + c0);
  if (MODE == kPool) {
    const size_t W2 = 2 * (size_t)A.W;
    const size_t b = (((size_t)n * 2 * A.H + 2 * hh) * W2 + 2 * ww) * A.ca + c0;
    const bf4 q0 = *(const bf4*)(A.xa + b), q1 = *(const bf4*)(A.xa + b + A.ca);
    const bf4 q2 = *(const bf4*)(A.xa + b + W2 * A.ca), q3 = *(const bf4*)(A.xa + b + (W2 + 1) * A.ca);
    bf4 r;
#pragma unroll
    for (int e = 0; e < 4; e++)
      r[e] = (short)f2bf((bf2f((uint16_t)q0[e]) + bf2f((uint16_t)q1[e]) + bf2f((uint16_t)q2[e]) +
                          bf2f((uint16_t)q3[e])) *
                         0.25f);
    return r;
  }
  if (c0 < A.ca)
    return *(const bf4*)(A.xa + (((size_t)n * (A.H / 2) + hh / 2) * (A.W / 2) + ww / 2) * A.ca + c0);
  return *(const bf4*)(A.xb + (((size_t)n * A.H + hh) * A.W + ww) * A.cb + (c0 - A.ca));
}

// One input channel c of pixel p shifted by tap t (zero outside the grid).
__device__ __forceinline__ float conv_in1(const ConvArgs& A, int p, int t, int c) {
  const int HW = A.H * A.W, n = p / HW, rem = p % HW;
  const int hh = rem / A.W + t / 3 - 1, ww = rem % A.W + t % 3 - 1;
  if (hh < 0 || hh >= A.H || ww < 0 || ww >= A.W) return 0.0f;
  if (A.mode == kPlain) return bf2f(A.xa[(((size_t)n * A.H + hh) * A.W + ww) * A.ca + c]);
  if (A.mode == kPool) {
    const size_t W2 = 2 * (size_t)A.W;
    const size_t b = (((size_t)n * 2 * A.H + 2 * hh) * W2 + 2 * ww) * A.ca + c;
    return bf2f(f2bf((bf2f(A.xa[b]) + bf2f(A.xa[b + A.ca]) + bf2f(A.xa[b + W2 * A.ca]) +
                      bf2f(A.xa[b + (W2 + 1) * A.ca])) *
                     0.25f));
  }
  if (c < A.ca) return bf2f(A.xa[(((size_t)n * (A.H / 2) + hh / 2) * (A.W / 2) + ww / 2) * A.ca + c]);
  return bf2f(A.xb[(((size_t)n * A.H + hh) * A.W + ww) * A.cb + (c - A.ca)]);
}

// Forward (and input-gradient) 3x3 convolution: each wave computes NT output
// tiles of 16 channels (from tile NT * blockIdx.y on; cout = 16 * NT *
// gridDim.y) for 16 pixels, over CB = cin / 16 input-channel blocks, fully
// unrolled so that the operand loads of all taps are issued ahead of the MFMA
// chain.  The network application (one sample) launches one tile per wave
// and the output tiles along y, so that the deep layers' few pixels still
// spread over many waves with short MFMA chains.
template <int MODE, int NT, int CB>
__global__ void __launch_bounds__(256) k_conv(ConvArgs A) {
  constexpr int cin = 16 * CB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int co0 = (int)blockIdx.y * NT * 16;
  const int HW = A.H * A.W, total = A.N * HW;
  const int p0 = (blockIdx.x * 4 + wave) * 16;
  if (p0 >= total) return;  // whole wave; the kernel has no barriers
  const int kg = lane >> 4, col = lane & 15;
  const int pa = p0 + col;
  const bool pin = pa < total;
  const int n = pin ? pa / HW : 0, rem = pin ? pa % HW : 0;
  const int h = rem / A.W, w = rem % A.W;
  f4 acc[NT];
#pragma unroll
  for (int nt = 0; nt < NT; nt++) acc[nt] = (f4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < 9; t++) {
    const int hh = h + t / 3 - 1, ww = w + t % 3 - 1;
    const bool ok = pin && hh >= 0 && hh < A.H && ww >= 0 && ww < A.W;
    const uint16_t* wt = A.w + ((size_t)t * A.cout + co0 + col) * cin + 4 * kg;
    bf4 a[CB], b[CB][NT];
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
      a[cb] = ok ? conv_in4<MODE>(A, n, hh, ww, 16 * cb + 4 * kg) : (bf4){0, 0, 0, 0};
#pragma unroll
      for (int nt = 0; nt < NT; nt++) b[cb][nt] = *(const bf4*)(wt + (size_t)nt * 16 * cin + 16 * cb);
    }
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
      for (int nt = 0; nt < NT; nt++)
        acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[cb], b[cb][nt], acc[nt], 0, 0, 0);
  }
#pragma unroll
  for (int nt = 0; nt < NT; nt++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int p = p0 + 4 * kg + i;
      if (p >= total) continue;
      const int co = co0 + nt * 16 + col;
      const size_t o = (size_t)p * A.cout + co;
      float v = acc[nt][i];
      if (A.relu) v = fmaxf(v, 0.0f);
      if (A.mask && !(bf2f(A.mask[o]) > 0.0f)) v = 0.0f;
      if (A.y) A.y[o] = f2bf(v);
      if (A.yf) A.yf[o] = v;
    }
}

// Partial sums of a weight gradient: at most this many pixel ranges.
constexpr int kWgradParts = 256;

// Weight gradient dW[t][o][i] = sum_p x(p + tap t)[i] dpre(p)[o]: GEMM with
// M = output channels, N = input channels, K = pixels.  Block (gx, gy): pixel
// chunks of 64 in the range of gx, 16 (tap, o-tile, i-tile) combinations of gy
// (4 per wave).  Partials per gx, reduced by k_wgrad_reduce.
struct WgradArgs {
  ConvArgs X;            // the layer's input (mode, channels, resolution)
  const uint16_t* dp;    // dpre [N][H][W][cout]
  int cout;
  float* part;           // [gridDim.x][9][cout][cin]
};

__global__ void __launch_bounds__(256) k_wgrad(WgradArgs A) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kg = lane >> 4, col = lane & 15;
  const ConvArgs& X = A.X;
  const int cin = X.mode == kUpCat ? X.ca + X.cb : X.ca;
  const int nm = A.cout / 16, nn = cin / 16, ncombo = 9 * nm * nn;
  const int cbase = (blockIdx.y * 4 + wave) * 4;
  if (cbase >= ncombo) return;  // whole wave
  int tap[4], m0[4], n0[4];
  bool act[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int c = cbase + j;
    act[j] = c < ncombo;
    const int cc = act[j] ? c : 0;
    tap[j] = cc / (nm * nn);
    m0[j] = (cc / nn) % nm * 16;
    n0[j] = cc % nn * 16;
  }
  f4 acc[4];
#pragma unroll
  for (int j = 0; j < 4; j++) acc[j] = (f4){0.f, 0.f, 0.f, 0.f};
  const int total = X.N * X.H * X.W, chunks = (total + 63) / 64;
  const int per = (chunks + gridDim.x - 1) / gridDim.x;
  const int ch0 = blockIdx.x * per, ch1 = min(chunks, ch0 + per);
  for (int ch = ch0; ch < ch1; ch++)
    for (int ks = 0; ks < 4; ks++) {
      const int k0 = ch * 64 + ks * 16 + 4 * kg;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (!act[j]) continue;
        bf4 a, b;
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int p = k0 + e;
          const bool in = p < total;
          a[e] = in ? (short)A.dp[(size_t)p * A.cout + m0[j] + col] : (short)0;
          b[e] = in ? (short)f2bf(conv_in1(X, p, tap[j], n0[j] + col)) : (short)0;
        }
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
      }
    }
  float* out = A.part + (size_t)blockIdx.x * 9 * A.cout * cin;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (!act[j]) continue;
#pragma unroll
    for (int i = 0; i < 4; i++)
      out[((size_t)tap[j] * A.cout + m0[j] + 4 * kg + i) * cin + n0[j] + col] = acc[j][i];
  }
}

__global__ void k_wgrad_reduce(const float* part, int nparts, int len, float* g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float s = 0.0f;
  for (int k = 0; k < nparts; k++) s += part[(size_t)k * len + i];
  g[i] = s;
}

// Adam on the FP32 masters.
__global__ void k_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1,
                       float b2, float c1, float c2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

struct WOff {
  int off[kLayers + 1];
  int cin[kLayers], cout[kLayers];
};

// BF16 copies of the weights: wb as stored, wt with the taps flipped and the
// channel axes swapped (the input-gradient convolution).
__global__ void k_weights_bf16(const float* w, uint16_t* wb, uint16_t* wt, WOff o, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  wb[i] = f2bf(w[i]);
  int l = 0;
  while (l < kLayers && i >= o.off[l + 1]) l++;
  if (l == kLayers) return;  // 1x1 output weights
  const int k = i - o.off[l], ci = o.cin[l], co = o.cout[l];
  // wt[t'][i][o] = w[8 - t'][o][i]
  const int tp = k / (ci * co), rem = k % (ci * co), ii = rem / co, oo = rem % co;
  wt[i] = f2bf(w[o.off[l] + ((8 - tp) * co + oo) * ci + ii]);
}

// Training batch: sample scale s = sqrt(sum r^2/d / HW), inputs, r/s and the
// loss operator.  One block per sample.
__global__ void __launch_bounds__(1024) k_gather(const float* data, const int* idx, int H, int W,
                                                 uint16_t* in, float* rhat, float* op,
                                                 float* scale) {
  __shared__ float s_red[32];
  const int b = blockIdx.x, HW = H * W;
  const float* r = data + (size_t)idx[b] * 4 * HW;
  const float *d = r + HW, *cx = r + 2 * HW, *cy = r + 3 * HW;
  float acc = 0.0f;
  for (int p = threadIdx.x; p < HW; p += blockDim.x) acc += r[p] * r[p] / d[p];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.0f;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += s_red[k];
    t = sqrtf(t / HW);
    s_red[31] = t > 0.0f ? t : 1.0f;
  }
  __syncthreads();
  const float s = s_red[31];
  if (threadIdx.x == 0) scale[b] = s;
  for (int p = threadIdx.x; p < HW; p += blockDim.x) {
    const int h = p / W, w = p % W;
    const float dd = d[p];
    uint16_t* o = in + ((size_t)b * HW + p) * kPadIn;
    o[0] = f2bf(r[p] / (s * dd));
    o[1] = f2bf((w > 0 ? cx[p - 1] : 0.0f) / dd);
    o[2] = f2bf(cx[p] / dd);
    o[3] = f2bf((h > 0 ? cy[p - W] : 0.0f) / dd);
    o[4] = f2bf(cy[p] / dd);
    for (int c = kInCh; c < kPadIn; c++) o[c] = 0;
    rhat[(size_t)b * HW + p] = r[p] / s;
    float* ob = op + (size_t)b * 3 * HW;
    ob[p] = dd;
    ob[HW + p] = cx[p];
    ob[2 * HW + p] = cy[p];
  }
}

// e_hat = y13 . w_out (FP32).
__global__ void k_out_fwd(const uint16_t* y, const float* wout, float* out, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float s = 0.0f;
  for (int c = 0; c < kOutCh; c++) s += bf2f(y[(size_t)p * kOutCh + c]) * bf2f(f2bf(wout[c]));
  out[p] = s;
}

// Energy loss per sample, L = (e.A e - 2 e.r_hat) / HW (the A-norm error of
// e up to a constant), and dL/de = 2 (A e - r_hat) / (HW N).  Partials per block.
__global__ void k_loss(const float* e, const float* rhat, const float* op, int N, int H, int W,
                       float* dout, float* lpart) {
  __shared__ float s_red[4];
  const int HW = H * W, total = N * HW;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (p < total) {
    const int n = p / HW, q = p % HW, h = q / W, w = q % W;
    const float* ob = op + (size_t)n * 3 * HW;
    const float *d = ob, *cx = ob + HW, *cy = ob + 2 * HW;
    const float* en = e + (size_t)n * HW;
    float ae = d[q] * en[q];
    if (w > 0) ae -= cx[q - 1] * en[q - 1];
    if (w + 1 < W) ae -= cx[q] * en[q + 1];
    if (h > 0) ae -= cy[q - W] * en[q - W];
    if (h + 1 < H) ae -= cy[q] * en[q + W];
    l = (en[q] * ae - 2.0f * en[q] * rhat[p]) / HW;
    dout[p] = 2.0f * (ae - rhat[p]) / ((float)HW * N);
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_down(l, o);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) lpart[blockIdx.x] = s_red[0] + s_red[1] + s_red[2] + s_red[3];
}

// Output layer backward: dpre13 = (y13 > 0) dout w_out (BF16) and the
// partials of dw_out = sum_p dout y13.
__global__ void k_out_bwd(const uint16_t* y, const float* wout, const float* dout, int n,
                          uint16_t* dpre, float* wpart) {
  __shared__ float s_w[4][kOutCh];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  float gw[kOutCh];
  for (int c = 0; c < kOutCh; c++) gw[c] = 0.0f;
  if (p < n) {
    const float g = dout[p];
    for (int c = 0; c < kOutCh; c++) {
      const float yv = bf2f(y[(size_t)p * kOutCh + c]);
      gw[c] = g * yv;
      dpre[(size_t)p * kOutCh + c] = yv > 0.0f ? f2bf(g * bf2f(f2bf(wout[c]))) : (uint16_t)0;
    }
  }
  for (int c = 0; c < kOutCh; c++) {
    float v = gw[c];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < kOutCh) {
    const int c = threadIdx.x;
    wpart[(size_t)blockIdx.x * kOutCh + c] = s_w[0][c] + s_w[1][c] + s_w[2][c] + s_w[3][c];
  }
}

// Gradient into the half-resolution input of an up-concat layer:
// dpre_src = (y_src > 0) * sum over the 2x2 children of dx_cat[:, c], c < ca.
__global__ void k_up_bwd(const float* dxcat, int ccat, const uint16_t* ysrc, int csrc, int N,
                         int Hc, int Wc, uint16_t* dpre) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = N * Hc * Wc * csrc;
  if (i >= total) return;
  const int c = i % csrc, p = i / csrc, n = p / (Hc * Wc), q = p % (Hc * Wc);
  const int h = q / Wc, w = q % Wc, Wf = 2 * Wc;
  const size_t f = ((size_t)n * 2 * Hc + 2 * h) * Wf + 2 * w;
  const float s = dxcat[f * ccat + c] + dxcat[(f + 1) * ccat + c] +
                  dxcat[(f + Wf) * ccat + c] + dxcat[(f + Wf + 1) * ccat + c];
  dpre[i] = bf2f(ysrc[i]) > 0.0f ? f2bf(s) : (uint16_t)0;
}

// Gradient into an encoder output that feeds both a skip connection and a
// pooling: dpre = (y > 0) (dx_cat[:, ca + c] + dx_pool(parent)[c] / 4).
__global__ void k_skip_bwd(const float* dxcat, int ccat, int ca, const float* dxpool,
                           const uint16_t* y, int c, int N, int H, int W, uint16_t* dpre) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = N * H * W * c;
  if (i >= total) return;
  const int ch = i % c, p = i / c, n = p / (H * W), q = p % (H * W);
  const int h = q / W, w = q % W;
  const size_t pc = ((size_t)n * (H / 2) + h / 2) * (W / 2) + w / 2;
  const float s = dxcat[(size_t)p * ccat + ca + ch] + 0.25f * dxpool[pc * c + ch];
  dpre[i] = bf2f(y[i]) > 0.0f ? f2bf(s) : (uint16_t)0;
}

// Solver side: network input from the coarse level (FP64) with the scale from
// the finest down kernel's partials; padding cells are zero.
__global__ void k_nn_input(MgLevel L, const double* part, int nb, DevScalars* sc, int H, int W,
                           uint16_t* in) {
  __shared__ double s_red[16];
  if (sc->done) return;
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v += part[3 * b];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += s_red[k];
    t = sqrt(t / ((double)H * W));
    s_red[15] = t > 0.0 ? t : 1.0;
  }
  __syncthreads();
  const double s = s_red[15];
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->nn_scale = s;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= H * W) return;
  const int h = p / W, w = p % W;
  uint16_t* o = in + (size_t)p * kPadIn;
  for (int c = 0; c < kPadIn; c++) o[c] = 0;
  if (h >= L.ny || w >= L.nx) return;
  const size_t m = (size_t)(h + 1) * L.P + (w + 1);
  const double d = L.d[m];
  o[0] = f2bf((float)(L.r[m] / (s * d)));
  o[1] = f2bf((float)(L.cx[m - 1] / d));
  o[2] = f2bf((float)(L.cx[m] / d));
  o[3] = f2bf((float)(L.cy[m - L.P] / d));
  o[4] = f2bf((float)(L.cy[m] / d));
}

__global__ void k_nn_output(const uint16_t* y, const float* wout, MgLevel L, DevScalars* sc,
                            int W) {
  if (sc->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.nx * L.ny) return;
  const int w = i % L.nx, h = i / L.nx;
  const size_t p = (size_t)h * W + w;
  float s = 0.0f;
  for (int c = 0; c < kOutCh; c++) s += bf2f(y[p * kOutCh + c]) * bf2f(f2bf(wout[c]));
  L.e[(size_t)(h + 1) * L.P + (w + 1)] = sc->nn_scale * (double)s;
}

__global__ void k_nn_store(MgLevel L, int H, int W, float* out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= H * W) return;
  const int h = p / W, w = p % W, HW = H * W;
  float r = 0.0f, d = 1.0f, cx = 0.0f, cy = 0.0f;
  if (h < L.ny && w < L.nx) {
    const size_t m = (size_t)(h + 1) * L.P + (w + 1);
    r = (float)L.r[m];
    d = (float)L.d[m];
    cx = (float)L.cx[m];
    cy = (float)L.cy[m];
  }
  out[p] = r;
  out[HW + p] = d;
  out[2 * HW + p] = cx;
  out[3 * HW + p] = cy;
}

#define HIPC(x)                         \
  do {                                  \
    if ((x) != hipSuccess) return -1;   \
  } while (0)

// The (mode, cout/16, cin/16) combinations of the forward and input-gradient
// convolutions of the network.
#define FSL_CONV_CASES(X)                                                              \
  X(kPlain, 1, 1) X(kPlain, 2, 2) X(kPlain, 3, 3) X(kPlain, 4, 4) X(kPlain, 1, 2)      \
  X(kPlain, 2, 3) X(kPlain, 3, 4) X(kPlain, 7, 3) X(kPlain, 5, 2) X(kPlain, 3, 1)      \
  X(kPool, 2, 1) X(kPool, 3, 2) X(kPool, 4, 3) X(kUpCat, 3, 7) X(kUpCat, 2, 5)         \
  X(kUpCat, 1, 3) X(kPlain, 1, 3) X(kPlain, 1, 4) X(kPool, 1, 1) X(kPool, 1, 2)        \
  X(kPool, 1, 3) X(kUpCat, 1, 5) X(kUpCat, 1, 7)

// split: one output tile per wave, the tiles along the grid's y (the
// network application); otherwise every tile in each wave (training batches).
int launch_conv(const ConvArgs& a, hipStream_t st, bool split = false) {
  const int total = a.N * a.H * a.W;
  const int ntile = a.cout / 16, cb = (a.mode == kUpCat ? a.ca + a.cb : a.ca) / 16;
  const int nt = split ? 1 : ntile;
  const dim3 g((total + 63) / 64, split ? ntile : 1), b(256);
#define FSL_CONV_LAUNCH(M, NT, CB)                                        \
  if (a.mode == M && nt == NT && cb == CB) {                              \
    hipLaunchKernelGGL((k_conv<M, NT, CB>), g, b, 0, st, a);              \
    return 0;                                                             \
  }
  FSL_CONV_CASES(FSL_CONV_LAUNCH)
#undef FSL_CONV_LAUNCH
  return -1;
}

// The forward convolution of layer l over activations a.
ConvArgs fwd_args(const Net* net, const Acts& a, int l) {
  const LayerDesc& L = kNet[l];
  ConvArgs c;
  memset(&c, 0, sizeof(c));
  c.xa = l == 0 ? a.in : a.y[l - 1];
  c.xb = L.mode == kUpCat ? a.y[skip_src(l)] : nullptr;
  c.ca = L.ca;
  c.cb = L.cb;
  c.mode = L.mode;
  c.w = net->wb + net->off[l];
  c.cout = L.cout;
  c.y = a.y[l];
  c.relu = 1;
  c.N = a.N;
  c.H = a.H >> L.level;
  c.W = a.W >> L.level;
  return c;
}

int forward(const Net* net, const Acts& a, hipStream_t st, bool split = false) {
  for (int l = 0; l < kLayers; l++)
    if (launch_conv(fwd_args(net, a, l), st, split)) return -1;
  return 0;
}

int ceil_div(size_t a, size_t b) { return (int)((a + b - 1) / b); }

// dW of layer l into net->g, from a.dpre[l].
void wgrad(Net* net, const Acts& a, int l, hipStream_t st) {
  WgradArgs w;
  w.X = fwd_args(net, a, l);
  w.dp = a.dpre[l];
  w.cout = kNet[l].cout;
  w.part = net->wpart;
  const int cin = layer_cin(l), len = 9 * w.cout * cin;
  const int ncombo = 9 * (w.cout / 16) * (cin / 16);
  const int total = a.N * (a.H >> kNet[l].level) * (a.W >> kNet[l].level);
  // pixel chunks of 64, at least 4 per block, at most kWgradParts blocks along x
  int gx = ceil_div(total, 64 * 4);
  const int cap = (int)(net->wpart_len / (size_t)len);
  gx = gx < 1 ? 1 : gx > kWgradParts ? kWgradParts : gx;
  if (gx > cap) gx = cap;
  hipLaunchKernelGGL(k_wgrad, dim3(gx, ceil_div(ncombo, 16)), dim3(256), 0, st, w);
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(ceil_div(len, 256)), dim3(256), 0, st, net->wpart, gx,
                     len, net->g + net->off[l]);
}

// Input gradient of layer l (l > 0): masked BF16 into dpre[l - 1] for plain
// layers, FP32 into `dx` for pooling and up-concat layers.
int dgrad(Net* net, const Acts& a, int l, float* dx, hipStream_t st) {
  const LayerDesc& L = kNet[l];
  ConvArgs c;
  memset(&c, 0, sizeof(c));
  c.xa = a.dpre[l];
  c.ca = L.cout;
  c.mode = kPlain;
  c.w = net->wt + net->off[l];
  c.cout = layer_cin(l);
  c.N = a.N;
  c.H = a.H >> L.level;
  c.W = a.W >> L.level;
  if (L.mode == kPlain) {
    c.y = a.dpre[l - 1];
    c.mask = a.y[l - 1];
  } else if (L.mode == kPool) {
    c.yf = dx;  // gradient w.r.t. the pooled input, at this layer's resolution
  } else {
    c.yf = dx;
  }
  return launch_conv(c, st);
}

WOff woff(const Net* net) {
  WOff o;
  for (int l = 0; l <= kLayers; l++) o.off[l] = (int)net->off[l];
  for (int l = 0; l < kLayers; l++) {
    o.cin[l] = layer_cin(l);
    o.cout[l] = kNet[l].cout;
  }
  return o;
}

void refresh_bf16(Net* net, hipStream_t st) {
  hipLaunchKernelGGL(k_weights_bf16, dim3(ceil_div(net->nparam, 256)), dim3(256), 0, st, net->w,
                     net->wb, net->wt, woff(net), (int)net->nparam);
}

// Full backward from a.dout; parameter gradients into net->g.
int backward(Net* net, const Acts& a, hipStream_t st) {
  const int n0 = a.N * a.H * a.W;
  const float* wout = net->w + net->off[kLayers];
  hipLaunchKernelGGL(k_out_bwd, dim3(ceil_div(n0, 256)), dim3(256), 0, st, a.y[13], wout, a.dout,
                     n0, a.dpre[13], net->wpart);
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(1), dim3(64), 0, st, net->wpart, ceil_div(n0, 256),
                     kOutCh, net->g + net->off[kLayers]);
  // per-level scratch for the concatenated-input gradients, kept until the
  // matching skip is consumed
  float* dxcat[3];
  size_t offc = 0;
  for (int lv = 0; lv < 3; lv++) {
    dxcat[lv] = a.dx + offc;
    const int l = 12 - 2 * lv;
    offc += (size_t)a.N * (a.H >> lv) * (a.W >> lv) * layer_cin(l);
  }
  for (int l = kLayers - 1; l >= 0; l--) {
    wgrad(net, a, l, st);
    if (l == 0) break;
    const LayerDesc& L = kNet[l];
    if (L.mode == kPlain) {
      if (dgrad(net, a, l, nullptr, st)) return -1;
    } else if (L.mode == kUpCat) {
      const int lv = L.level;
      if (dgrad(net, a, l, dxcat[lv], st)) return -1;
      const int Hc = a.H >> (lv + 1), Wc = a.W >> (lv + 1);
      const int total = a.N * Hc * Wc * L.ca;
      hipLaunchKernelGGL(k_up_bwd, dim3(ceil_div(total, 256)), dim3(256), 0, st, dxcat[lv],
                         L.ca + L.cb, a.y[l - 1], L.ca, a.N, Hc, Wc, a.dpre[l - 1]);
    } else {  // pooling layer at level lv + 1 reading encoder output l - 1 at level lv
      const int lv = L.level - 1;
      if (dgrad(net, a, l, a.dx2, st)) return -1;
      const int c = L.ca, H = a.H >> lv, W = a.W >> lv;
      const int dl = 12 - 2 * lv;  // the decoder layer holding the skip
      const int total = a.N * H * W * c;
      hipLaunchKernelGGL(k_skip_bwd, dim3(ceil_div(total, 256)), dim3(256), 0, st, dxcat[lv],
                         layer_cin(dl), kNet[dl].ca, a.dx2, a.y[l - 1], c, a.N, H, W,
                         a.dpre[l - 1]);
    }
  }
  return 0;
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
double unif(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

int acts_alloc(Acts& a, int N, int H, int W) {
  acts_free(a);
  const size_t P0 = (size_t)N * H * W;
  std::vector<size_t> sz;
  for (int l = 0; l < kLayers; l++) sz.push_back((P0 >> (2 * kNet[l].level)) * kNet[l].cout * 2);
  for (int l = 0; l < kLayers; l++) sz.push_back((P0 >> (2 * kNet[l].level)) * kNet[l].cout * 2);
  size_t dxbytes = 0;
  for (int lv = 0; lv < 3; lv++) dxbytes += (P0 >> (2 * lv)) * layer_cin(12 - 2 * lv) * 4;
  sz.push_back(dxbytes);
  size_t dx2 = 0;  // largest pooled-input gradient (layers 2, 4, 6)
  for (int l = 2; l <= 6; l += 2) {
    const size_t b = (P0 >> (2 * kNet[l].level)) * kNet[l].ca * 4;
    dx2 = b > dx2 ? b : dx2;
  }
  sz.push_back(dx2);
  sz.push_back(P0 * kPadIn * 2);  // in
  sz.push_back(P0 * 4);           // out
  sz.push_back(P0 * 4);           // dout
  sz.push_back(P0 * 4);           // rhat
  sz.push_back((size_t)N * 4);    // scale
  sz.push_back(P0 * 3 * 4);       // op
  size_t total = 0;
  for (size_t b : sz) total += (b + 255) / 256 * 256;
  char* base = nullptr;
  if (hipMalloc(&base, total) != hipSuccess) return -1;
  // graphs captured over these buffers are keyed by this number
  static std::atomic<uint64_t> next_id{1};
  a.id = next_id.fetch_add(1);
  a.N = N;
  a.H = H;
  a.W = W;
  a.bytes = total;
  size_t o = 0, k = 0;
  auto take = [&]() {
    char* p = base + o;
    o += (sz[k++] + 255) / 256 * 256;
    return p;
  };
  for (int l = 0; l < kLayers; l++) a.y[l] = (uint16_t*)take();
  for (int l = 0; l < kLayers; l++) a.dpre[l] = (uint16_t*)take();
  a.dx = (float*)take();
  a.dx2 = (float*)take();
  a.in = (uint16_t*)take();
  a.out = (float*)take();
  a.dout = (float*)take();
  a.rhat = (float*)take();
  a.scale = (float*)take();
  a.op = (float*)take();
  return 0;
}

void acts_free(Acts& a) {
  if (a.y[0]) (void)hipFree(a.y[0]);
  a = Acts();
}

int net_init(Net* net, int cap, int H, int W, uint64_t seed, hipStream_t st) {
  net->st = st;
  net->H = H;
  net->W = W;
  net->cap = cap;
  net->count = 0;
  net->rng = seed;
  size_t o = 0;
  for (int l = 0; l < kLayers; l++) {
    net->off[l] = o;
    o += (size_t)9 * kNet[l].cout * layer_cin(l);
  }
  net->off[kLayers] = o;
  o += kOutCh;
  net->nparam = o;
  size_t maxlen = 0;
  for (int l = 0; l < kLayers; l++) {
    const size_t len = (size_t)9 * kNet[l].cout * layer_cin(l);
    maxlen = len > maxlen ? len : maxlen;
  }
  net->wpart_len = kWgradParts * maxlen;
  HIPC(hipMalloc(&net->w, o * 4));
  HIPC(hipMalloc(&net->g, o * 4));
  HIPC(hipMalloc(&net->m, o * 4));
  HIPC(hipMalloc(&net->v, o * 4));
  HIPC(hipMalloc(&net->wb, o * 2));
  HIPC(hipMalloc(&net->wt, o * 2));
  HIPC(hipMalloc(&net->wpart, net->wpart_len * 4));
  HIPC(hipMalloc(&net->lpart, (size_t)kLossSlots * 8));
  if (cap > 0) HIPC(hipMalloc(&net->data, (size_t)cap * 4 * H * W * 4));
  // He initialisation (input channels beyond the five used stay zero)
  std::vector<float> w(o, 0.0f);
  uint64_t s = seed;
  for (int l = 0; l < kLayers; l++) {
    const int ci = layer_cin(l), co = kNet[l].cout;
    const int fan = 9 * (l == 0 ? kInCh : ci);
    const double sd = sqrt(2.0 / fan);
    for (int t = 0; t < 9; t++)
      for (int oo = 0; oo < co; oo++)
        for (int ii = 0; ii < ci; ii++) {
          const double u1 = unif(s), u2 = unif(s);
          const double g = sqrt(-2.0 * log(u1 > 1e-300 ? u1 : 1e-300)) * cos(6.283185307179586 * u2);
          w[net->off[l] + ((size_t)t * co + oo) * ci + ii] =
              (l == 0 && ii >= kInCh) ? 0.0f : (float)(sd * g);
        }
  }
  for (int c = 0; c < kOutCh; c++) w[net->off[kLayers] + c] = (float)(0.25 * (unif(s) - 0.5));
  HIPC(hipMemcpy(net->w, w.data(), o * 4, hipMemcpyHostToDevice));
  HIPC(hipMemset(net->m, 0, o * 4));
  HIPC(hipMemset(net->v, 0, o * 4));
  HIPC(hipMemset(net->g, 0, o * 4));
  net->adam_t = 0;
  refresh_bf16(net, st);
  HIPC(hipStreamSynchronize(st));
  return 0;
}

void net_free(Net* net) {
  (void)hipFree(net->w);
  (void)hipFree(net->g);
  (void)hipFree(net->m);
  (void)hipFree(net->v);
  (void)hipFree(net->wb);
  (void)hipFree(net->wt);
  (void)hipFree(net->wpart);
  (void)hipFree(net->lpart);
  (void)hipFree(net->data);
  acts_free(net->train);
  acts_free(net->infer);
}

int net_apply(Net* net, const MgLevel& L, const double* part, int nb, DevScalars* sc,
              hipStream_t st) {
  Acts& a = net->infer;
  const int H = net_side(L.ny), W = net_side(L.nx);
  if (a.N != 1 || a.H != H || a.W != W) {
    if (acts_alloc(a, 1, H, W)) return -1;
  }
  hipLaunchKernelGGL(k_nn_input, dim3(ceil_div((size_t)H * W, 256)), dim3(256), 0, st, L, part,
                     nb, sc, H, W, a.in);
  if (forward(net, a, st, true)) return -1;
  hipLaunchKernelGGL(k_nn_output, dim3(ceil_div((size_t)L.nx * L.ny, 256)), dim3(256), 0, st,
                     a.y[13], net->w + net->off[kLayers], L, sc, W);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

void net_refresh(Net* net) { refresh_bf16(net, net->st); }

int net_store_sample(Net* net, const MgLevel& L, hipStream_t st) {
  if (net->count >= net->cap || L.nx > net->W || L.ny > net->H) return 1;
  float* out = net->data + (size_t)net->count * 4 * net->H * net->W;
  hipLaunchKernelGGL(k_nn_store, dim3(ceil_div((size_t)net->H * net->W, 256)), dim3(256), 0, st,
                     L, net->H, net->W, out);
  net->count++;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// One training step on `B` samples (indices given), returns the mean loss.
int train_step(Net* net, const int* idx_host, int B, float lr, double* loss) {
  Acts& a = net->train;
  hipStream_t st = net->st;
  if (a.N != B || a.H != net->H || a.W != net->W) {
    if (acts_alloc(a, B, net->H, net->W)) return -1;
  }
  int* d_idx = (int*)(net->lpart + kLossSlots / 2);
  HIPC(hipMemcpyAsync(d_idx, idx_host, B * sizeof(int), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_gather, dim3(B), dim3(1024), 0, st, net->data, d_idx, net->H, net->W,
                     a.in, a.rhat, a.op, a.scale);
  if (forward(net, a, st)) return -1;
  const int n0 = B * net->H * net->W;
  hipLaunchKernelGGL(k_out_fwd, dim3(ceil_div(n0, 256)), dim3(256), 0, st, a.y[13],
                     net->w + net->off[kLayers], a.out, n0);
  const int nbl = ceil_div(n0, 256);
  if (nbl > kLossSlots / 2 || B > kLossSlots) return -1;
  hipLaunchKernelGGL(k_loss, dim3(nbl), dim3(256), 0, st, a.out, a.rhat, a.op, B, net->H, net->W,
                     a.dout, net->lpart);
  if (backward(net, a, st)) return -1;
  net->adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)net->adam_t), c2 = 1.0f - powf(b2, (float)net->adam_t);
  hipLaunchKernelGGL(k_adam, dim3(ceil_div(net->nparam, 256)), dim3(256), 0, st, net->w, net->g,
                     net->m, net->v, (int)net->nparam, lr, b1, b2, c1, c2, 1e-8f);
  refresh_bf16(net, st);
  std::vector<float> lp(nbl);
  HIPC(hipMemcpyAsync(lp.data(), net->lpart, nbl * sizeof(float), hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  double s = 0.0;
  for (float v : lp) s += v;
  *loss = s / B;
  return 0;
}

// Loss and parameter gradient of a host batch (for checks against the CPU
// reference): in [B][H][W][5] (r/(s d) and the coupling ratios), rhat, op.
int loss_grad(Net* net, int B, int H, int W, const float* in5, const float* rhat, const float* op,
              double* loss, float* grad, float* out) {
  Acts& a = net->train;
  hipStream_t st = net->st;
  if (acts_alloc(a, B, H, W)) return -1;
  const size_t n0 = (size_t)B * H * W;
  std::vector<uint16_t> in(n0 * kPadIn, 0);
  for (size_t p = 0; p < n0; p++)
    for (int c = 0; c < kInCh; c++) {
      uint32_t u;
      memcpy(&u, &in5[p * kInCh + c], 4);
      u += 0x7fffu + ((u >> 16) & 1u);
      in[p * kPadIn + c] = (uint16_t)(u >> 16);
    }
  HIPC(hipMemcpyAsync(a.in, in.data(), in.size() * 2, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(a.rhat, rhat, n0 * 4, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(a.op, op, n0 * 3 * 4, hipMemcpyHostToDevice, st));
  // a single sample runs the forward in the network application's launch
  // shape (one output tile per wave), a batch in the training shape
  if (forward(net, a, st, B == 1)) return -1;
  hipLaunchKernelGGL(k_out_fwd, dim3(ceil_div(n0, 256)), dim3(256), 0, st, a.y[13],
                     net->w + net->off[kLayers], a.out, (int)n0);
  const int nbl = ceil_div(n0, 256);
  if (nbl > kLossSlots / 2) return -1;
  hipLaunchKernelGGL(k_loss, dim3(nbl), dim3(256), 0, st, a.out, a.rhat, a.op, B, H, W, a.dout,
                     net->lpart);
  if (backward(net, a, st)) return -1;
  std::vector<float> lp(nbl);
  HIPC(hipMemcpyAsync(lp.data(), net->lpart, nbl * 4, hipMemcpyDeviceToHost, st));
  if (grad) HIPC(hipMemcpyAsync(grad, net->g, net->nparam * 4, hipMemcpyDeviceToHost, st));
  if (out) HIPC(hipMemcpyAsync(out, a.out, n0 * 4, hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  double s = 0.0;
  for (float v : lp) s += v;
  *loss = s / B;
  return 0;
}

}  // namespace fsl
