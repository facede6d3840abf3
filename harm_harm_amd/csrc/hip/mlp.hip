// mlp.hip -- training and inference kernels of the learned inversion guess,
// and the batched inversion seeded by it.
//
// Training runs in one persistent workgroup (4 wavefronts) that loops over
// minibatches of 128 pairs: forward (three BF16 matrix products with FP32
// accumulation, tanh hidden layers), mean-squared error on the standardised
// targets, backward (the weight gradients as transposed products X^T G, the
// input gradients as G W^T), and Adam on an FP32 master copy.  Every gradient
// element is produced by exactly one lane with a fixed summation order, so a
// training run is bit-reproducible.
//
// The seeded inversion evaluates the network per wavefront on 64 zones (the
// matrix instructions need the whole wave, so every lane joins in, active or
// not), starts the FP64 Newton solve from the predicted W', and falls back to
// the plain guess for any zone whose seeded solve does not succeed.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <vector>
#include <algorithm>
#include "mlp.h"
#include "hhost.h"
#include "gkern.h"
#include "dinvert.h"

using hm::v4f;

struct harm_mlp {
  float* w32 = nullptr;     // kNParam master parameters
  float* adam_m = nullptr;  // first moments
  float* adam_v = nullptr;  // second moments
  uint16_t* wbf = nullptr;  // kNBf BF16 weight images
  float* norm = nullptr;    // kNNorm standardisation constants
  float* loss = nullptr;    // per-launch mean loss (device scalar)
  int64_t t = 0;            // Adam steps taken
  std::vector<float> hnorm; // host copy of norm
};

namespace {

constexpr int kBlock = 256;
constexpr int kBatch = 128;     // training minibatch
constexpr int kWaves = kBlock / 64;

// ---------------------------------------------------------------------------
// shared pieces
// ---------------------------------------------------------------------------

// BF16 images of the weights from the FP32 parameters (zero padding).
__global__ void k_mlp_pack(const float* __restrict__ w, uint16_t* __restrict__ wbf)
{
  for (int e = threadIdx.x + blockIdx.x * blockDim.x; e < hm::kNBf; e += blockDim.x * gridDim.x) {
    float v = 0.0f;
    if (e < hm::kBfW2) {
      const int f = e / hm::kHid, h = e % hm::kHid;
      if (f < hm::kIn) v = w[hm::kOffW1 + f * hm::kHid + h];
    } else if (e < hm::kBfW3) {
      v = w[hm::kOffW2 + (e - hm::kBfW2)];
    } else {
      const int x = e - hm::kBfW3, h = x / hm::kOutPad, o = x % hm::kOutPad;
      if (o < hm::kOut) v = w[hm::kOffW3 + h * hm::kOut + o];
    }
    wbf[e] = hm::f2bf(v);
  }
}

// Weights and biases into LDS for the forward products.
__device__ void stage_weights(const uint16_t* __restrict__ wbf, const float* __restrict__ w32, uint16_t* sW,
                              float* sBias)
{
  for (int e = threadIdx.x; e < hm::kNBf; e += blockDim.x) sW[e] = wbf[e];
  for (int e = threadIdx.x; e < hm::kHid; e += blockDim.x) {
    sBias[e] = w32[hm::kOffB1 + e];
    sBias[hm::kHid + e] = w32[hm::kOffB2 + e];
  }
  if (threadIdx.x < hm::kOut) sBias[2 * hm::kHid + threadIdx.x] = w32[hm::kOffB3 + threadIdx.x];
}

// Forward pass of one wavefront's 64 samples.  X (standardised, BF16) is in
// bufA with row stride 16; the hidden layers go to bufB then bufA (row stride
// 64); the two standardised outputs of sample m land in y[2*m + o].  Every
// thread of the block must call it (it synchronises the block between
// layers); the caller must make X visible first.
__device__ void wave_forward(uint16_t* bufA, uint16_t* bufB, const uint16_t* sW, const float* sBias, float* y)
{
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  // layer 1: [64 x 16] x [16 x 64]
  for (int t = 0; t < 16; t++) {
    const int mt = t >> 2, nt = t & 3;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    acc = hm::mma16(bufA + mt * 16 * hm::kInPad, hm::kInPad, 1, sW + hm::kBfW1 + nt * 16, hm::kHid, 1, hm::kInPad, acc);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int m = mt * 16 + 4 * q + i, n = nt * 16 + r;
      bufB[m * hm::kHid + n] = hm::f2bf(tanhf(acc[i] + sBias[n]));
    }
  }
  __syncthreads();
  // layer 2: [64 x 64] x [64 x 64]
  for (int t = 0; t < 16; t++) {
    const int mt = t >> 2, nt = t & 3;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    acc = hm::mma16(bufB + mt * 16 * hm::kHid, hm::kHid, 1, sW + hm::kBfW2 + nt * 16, hm::kHid, 1, hm::kHid, acc);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int m = mt * 16 + 4 * q + i, n = nt * 16 + r;
      bufA[m * hm::kHid + n] = hm::f2bf(tanhf(acc[i] + sBias[hm::kHid + n]));
    }
  }
  __syncthreads();
  // layer 3: [64 x 64] x [64 x 16 (2 used)]
  for (int mt = 0; mt < 4; mt++) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    acc = hm::mma16(bufA + mt * 16 * hm::kHid, hm::kHid, 1, sW + hm::kBfW3, hm::kOutPad, 1, hm::kHid, acc);
    if (r < hm::kOut) {
#pragma unroll
      for (int i = 0; i < 4; i++) y[(mt * 16 + 4 * q + i) * 2 + r] = acc[i] + sBias[2 * hm::kHid + r];
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// training pairs
// ---------------------------------------------------------------------------

// Features and targets of n inversion problems from their conserved state U
// (UEVOLVE, SoA), geometry (SoA) and converged primitives p (SoA).  X[z][8]
// raw features, T[z][2] = (log(W'/D), log(u~^2)), ok[z] = 1 when usable.
__global__ __launch_bounds__(kBlock) void k_mlp_pairs(const double* __restrict__ U, const double* __restrict__ geom,
                                                      const double* __restrict__ p, long n, double gam,
                                                      double gammafail, float* __restrict__ X,
                                                      float* __restrict__ T, int32_t* __restrict__ ok)
{
  const long z = (long)blockIdx.x * kBlock + threadIdx.x;
  if (z >= n) return;
  hd::Geom g;
  hd::load_geom(geom, z, n, g);
  const double ig = (g.gdet >= 0.0 ? 1.0 : -1.0) / (fabs(g.gdet) + D_SMALL);
  double Ut[HARM_NPR], pz[HARM_NPR];
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) {
    Ut[pl] = g.alpha * (U[pl * n + z] * ig);
    pz[pl] = p[pl * n + z];
  }
  hd::Inv s;
  double Qtcon[4], Bcon[4];
  hd::inv_setup(Ut, g, gam, gammafail, s, Qtcon, Bcon);
  float f[hm::kIn];
  hm::features(s, f);
  double Wp, utsq;
  hm::wp_of_prim(pz, g, s.D, gam, Wp, utsq);
  bool good = s.D > 0.0 && Wp > 0.0 && utsq >= 0.0 && isfinite(Wp) && isfinite(utsq);
  for (int k = 0; k < hm::kIn; k++) {
    good = good && isfinite(f[k]);
    X[z * hm::kIn + k] = f[k];
  }
  T[z * 2 + 0] = (float)log(Wp / s.D + hm::kTinyWp);
  T[z * 2 + 1] = (float)log(utsq + hm::kTinyUt);
  ok[z] = good ? 1 : 0;
}

// ---------------------------------------------------------------------------
// training (one persistent workgroup)
// ---------------------------------------------------------------------------
struct TrainArgs {
  float* w32;
  float* am;
  float* av;
  uint16_t* wbf;
  const float* norm;
  const float* X;
  const float* T;
  const int32_t* idx;  // sample order
  long nidx;
  long nx;             // rows of X and T
  long first;          // position in idx of this launch's first sample
  int nsteps;
  long t0;             // Adam steps before this launch
  float lr, beta1, beta2, eps;
  int update;          // 0: write gradients of one batch to grad_out instead of stepping
  float* grad_out;
  float* loss_out;     // mean loss over the launch's steps
};

__global__ __launch_bounds__(kBlock) void k_mlp_train(TrainArgs a)
{
  __shared__ uint16_t sX[kBatch * hm::kInPad];
  __shared__ uint16_t sH1[kBatch * hm::kHid];
  __shared__ uint16_t sH2[kBatch * hm::kHid];
  __shared__ uint16_t sG1[kBatch * hm::kHid];
  __shared__ uint16_t sG2[kBatch * hm::kHid];
  __shared__ uint16_t sDY[kBatch * hm::kOutPad];
  __shared__ float sT[kBatch * hm::kOut];
  __shared__ float sErr[kBatch * hm::kOut];
  __shared__ float sDYf[kBatch * hm::kOut];
  __shared__ uint16_t sW[hm::kNBf];
  __shared__ float sBias[2 * hm::kHid + hm::kOut];

  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, r = l & 15, q = l >> 4;
  stage_weights(a.wbf, a.w32, sW, sBias);
  for (int e = tid; e < kBatch * hm::kOutPad; e += kBlock) sDY[e] = 0;
  float loss_acc = 0.0f;
  __syncthreads();

  for (int step = 0; step < a.nsteps; step++) {
    // ---- batch ----
    if (tid < kBatch) {
      const long pos = (a.first + (long)step * kBatch + tid) % a.nidx;
      long sidx = a.idx[pos];
      if (sidx < 0 || sidx >= a.nx) sidx = 0;  // out-of-range indices are rejected on the host
#pragma unroll
      for (int k = 0; k < hm::kIn; k++)
        sX[tid * hm::kInPad + k] = hm::f2bf((a.X[sidx * hm::kIn + k] - a.norm[k]) * a.norm[hm::kIn + k]);
#pragma unroll
      for (int k = hm::kIn; k < hm::kInPad; k++) sX[tid * hm::kInPad + k] = 0;
#pragma unroll
      for (int o = 0; o < hm::kOut; o++)
        sT[tid * 2 + o] = (a.T[sidx * 2 + o] - a.norm[2 * hm::kIn + o]) / a.norm[2 * hm::kIn + hm::kOut + o];
    }
    __syncthreads();
    // ---- layer 1 ----
    for (int t = w * 8; t < w * 8 + 8; t++) {
      const int mt = t >> 2, nt = t & 3;
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      acc = hm::mma16(sX + mt * 16 * hm::kInPad, hm::kInPad, 1, sW + hm::kBfW1 + nt * 16, hm::kHid, 1, hm::kInPad,
                      acc);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int m = mt * 16 + 4 * q + i, n = nt * 16 + r;
        sH1[m * hm::kHid + n] = hm::f2bf(tanhf(acc[i] + sBias[n]));
      }
    }
    __syncthreads();
    // ---- layer 2 ----
    for (int t = w * 8; t < w * 8 + 8; t++) {
      const int mt = t >> 2, nt = t & 3;
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      acc = hm::mma16(sH1 + mt * 16 * hm::kHid, hm::kHid, 1, sW + hm::kBfW2 + nt * 16, hm::kHid, 1, hm::kHid, acc);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int m = mt * 16 + 4 * q + i, n = nt * 16 + r;
        sH2[m * hm::kHid + n] = hm::f2bf(tanhf(acc[i] + sBias[hm::kHid + n]));
      }
    }
    __syncthreads();
    // ---- layer 3 and the loss gradient ----
    for (int mt = 2 * w; mt < 2 * w + 2; mt++) {
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      acc = hm::mma16(sH2 + mt * 16 * hm::kHid, hm::kHid, 1, sW + hm::kBfW3, hm::kOutPad, 1, hm::kHid, acc);
      if (r < hm::kOut) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int m = mt * 16 + 4 * q + i;
          const float d = acc[i] + sBias[2 * hm::kHid + r] - sT[m * 2 + r];
          const float dy = d * (2.0f / (kBatch * hm::kOut));
          sErr[m * 2 + r] = d * d;
          sDYf[m * 2 + r] = dy;
          sDY[m * hm::kOutPad + r] = hm::f2bf(dy);
        }
      }
    }
    __syncthreads();
    // ---- dW3 = H2^T dY (one 16x16 tile per wave); dH2 = dY W3^T ----
    v4f g3 = {0.f, 0.f, 0.f, 0.f};
    g3 = hm::mma16(sH2 + w * 16, 1, hm::kHid, sDY, hm::kOutPad, 1, kBatch, g3);
    for (int t = w * 8; t < w * 8 + 8; t++) {
      const int mt = t >> 2, nt = t & 3;
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      acc = hm::mma16(sDY + mt * 16 * hm::kOutPad, hm::kOutPad, 1, sW + hm::kBfW3 + nt * 16 * hm::kOutPad, 1,
                      hm::kOutPad, hm::kOutPad, acc);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int m = mt * 16 + 4 * q + i, n = nt * 16 + r;
        const float h = hm::bf2f(sH2[m * hm::kHid + n]);
        sG2[m * hm::kHid + n] = hm::f2bf(acc[i] * (1.0f - h * h));
      }
    }
    __syncthreads();
    // ---- dW2 = H1^T dZ2 (four tiles per wave); dH1 = dZ2 W2^T ----
    v4f g2[4];
    for (int u = 0; u < 4; u++) {
      const int t = 4 * w + u, mt = t >> 2, nt = t & 3;
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      g2[u] = hm::mma16(sH1 + mt * 16, 1, hm::kHid, sG2 + nt * 16, hm::kHid, 1, kBatch, acc);
    }
    for (int t = w * 8; t < w * 8 + 8; t++) {
      const int mt = t >> 2, nt = t & 3;
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      acc = hm::mma16(sG2 + mt * 16 * hm::kHid, hm::kHid, 1, sW + hm::kBfW2 + nt * 16 * hm::kHid, 1, hm::kHid, hm::kHid,
                      acc);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int m = mt * 16 + 4 * q + i, n = nt * 16 + r;
        const float h = hm::bf2f(sH1[m * hm::kHid + n]);
        sG1[m * hm::kHid + n] = hm::f2bf(acc[i] * (1.0f - h * h));
      }
    }
    __syncthreads();
    // ---- dW1 = X^T dZ1 (one tile per wave) and the bias gradients ----
    v4f g1 = {0.f, 0.f, 0.f, 0.f};
    g1 = hm::mma16(sX, 1, hm::kInPad, sG1 + w * 16, hm::kHid, 1, kBatch, g1);
    float gb = 0.0f;  // bias gradient owned by this thread
    if (tid < hm::kHid) {
      for (int m = 0; m < kBatch; m++) gb += hm::bf2f(sG1[m * hm::kHid + tid]);
    } else if (tid < 2 * hm::kHid) {
      for (int m = 0; m < kBatch; m++) gb += hm::bf2f(sG2[m * hm::kHid + tid - hm::kHid]);
    } else if (tid < 2 * hm::kHid + hm::kOut) {
      for (int m = 0; m < kBatch; m++) gb += sDYf[m * 2 + tid - 2 * hm::kHid];
    }
    if (tid == 0) {
      float sl = 0.0f;
      for (int e = 0; e < kBatch * hm::kOut; e++) sl += sErr[e];
      loss_acc += sl / (kBatch * hm::kOut);
    }
    __syncthreads();  // every read of the current weights is done
    // ---- parameter step ----
    const long tstep = a.t0 + step + 1;
    const float bc1 = 1.0f - powf(a.beta1, (float)tstep), bc2 = 1.0f - powf(a.beta2, (float)tstep);
    auto apply = [&](int pidx, float g) {
      if (!a.update) { a.grad_out[pidx] = g; return a.w32[pidx]; }
      const float m = a.beta1 * a.am[pidx] + (1.0f - a.beta1) * g;
      const float v = a.beta2 * a.av[pidx] + (1.0f - a.beta2) * g * g;
      a.am[pidx] = m;
      a.av[pidx] = v;
      const float wn = a.w32[pidx] - a.lr * (m / bc1) / (sqrtf(v / bc2) + a.eps);
      a.w32[pidx] = wn;
      return wn;
    };
    if (r < hm::kOut) {  // W3 tile of wave w: rows h = 16w + 4q + i
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int h = w * 16 + 4 * q + i;
        const float wn = apply(hm::kOffW3 + h * hm::kOut + r, g3[i]);
        sW[hm::kBfW3 + h * hm::kOutPad + r] = hm::f2bf(wn);
      }
    }
    for (int u = 0; u < 4; u++) {
      const int t = 4 * w + u, mt = t >> 2, nt = t & 3;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int hin = mt * 16 + 4 * q + i, hout = nt * 16 + r;
        const float wn = apply(hm::kOffW2 + hin * hm::kHid + hout, g2[u][i]);
        sW[hm::kBfW2 + hin * hm::kHid + hout] = hm::f2bf(wn);
      }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int f = 4 * q + i, h = w * 16 + r;
      if (f < hm::kIn) {
        const float wn = apply(hm::kOffW1 + f * hm::kHid + h, g1[i]);
        sW[hm::kBfW1 + f * hm::kHid + h] = hm::f2bf(wn);
      }
    }
    if (tid < hm::kHid) sBias[tid] = apply(hm::kOffB1 + tid, gb);
    else if (tid < 2 * hm::kHid) sBias[tid] = apply(hm::kOffB2 + tid - hm::kHid, gb);
    else if (tid < 2 * hm::kHid + hm::kOut) sBias[tid] = apply(hm::kOffB3 + tid - 2 * hm::kHid, gb);
    __syncthreads();
  }
  // the BF16 images of the final weights
  for (int e = tid; e < hm::kNBf; e += kBlock) a.wbf[e] = sW[e];
  if (tid == 0) a.loss_out[0] = loss_acc / (a.nsteps > 0 ? a.nsteps : 1);
}

// ---------------------------------------------------------------------------
// inference
// ---------------------------------------------------------------------------
struct Fwd {
  uint16_t sW[hm::kNBf];
  float sBias[2 * hm::kHid + hm::kOut];
  uint16_t bufA[kWaves][64 * hm::kHid];
  uint16_t bufB[kWaves][64 * hm::kHid];
  float y[kWaves][64 * 2];
};

// X[n][8] raw features -> Y[n][2] raw predictions (log W'/D, log u~^2)
__global__ __launch_bounds__(kBlock) void k_mlp_forward(const uint16_t* __restrict__ wbf,
                                                        const float* __restrict__ w32, const float* __restrict__ norm,
                                                        const float* __restrict__ X, float* __restrict__ Y, long n)
{
  __shared__ Fwd S;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  stage_weights(wbf, w32, S.sW, S.sBias);
  const long z = (long)blockIdx.x * kBlock + threadIdx.x;
#pragma unroll
  for (int k = 0; k < hm::kInPad; k++) {
    float v = 0.0f;
    if (z < n && k < hm::kIn) v = (X[z * hm::kIn + k] - norm[k]) * norm[hm::kIn + k];
    S.bufA[w][l * hm::kInPad + k] = hm::f2bf(v);
  }
  __syncthreads();
  wave_forward(S.bufA[w], S.bufB[w], S.sW, S.sBias, S.y[w]);
  if (z < n) {
#pragma unroll
    for (int o = 0; o < hm::kOut; o++)
      Y[z * 2 + o] = S.y[w][l * 2 + o] * norm[2 * hm::kIn + hm::kOut + o] + norm[2 * hm::kIn + o];
  }
}

// Batched inversion started from the network's W' (harm_invert_batch's
// arguments, plus the count of zones that fell back to the plain guess).
__global__ __launch_bounds__(kBlock) void k_invert_batch_mlp(const uint16_t* __restrict__ wbf,
                                                             const float* __restrict__ w32,
                                                             const float* __restrict__ norm,
                                                             const double* __restrict__ U,
                                                             const double* __restrict__ geom, double* __restrict__ p,
                                                             int32_t* __restrict__ pflag, int32_t* __restrict__ niter,
                                                             long n, double gam, double gammafail,
                                                             unsigned long long* __restrict__ nfallback)
{
  __shared__ Fwd S;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  stage_weights(wbf, w32, S.sW, S.sBias);
  const long z = (long)blockIdx.x * kBlock + threadIdx.x;
  const bool act = z < n;
  double D = 1.0;
  {
    float f[hm::kIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (act) {
      hd::Geom g;
      hd::load_geom(geom, z, n, g);
      const double ig = (g.gdet >= 0.0 ? 1.0 : -1.0) / (fabs(g.gdet) + D_SMALL);
      double Ut[HARM_NPR];
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) Ut[pl] = g.alpha * (U[pl * n + z] * ig);
      hd::Inv s;
      double Qtcon[4], Bcon[4];
      hd::inv_setup(Ut, g, gam, gammafail, s, Qtcon, Bcon);
      hm::features(s, f);
      D = s.D;
    }
#pragma unroll
    for (int k = 0; k < hm::kInPad; k++) {
      float v = 0.0f;
      if (act && k < hm::kIn) v = (f[k] - norm[k]) * norm[hm::kIn + k];
      S.bufA[w][l * hm::kInPad + k] = hm::f2bf(v);
    }
  }
  __syncthreads();
  wave_forward(S.bufA[w], S.bufB[w], S.sW, S.sBias, S.y[w]);
  int fb = 0;
  if (act) {
    const float y0 = S.y[w][l * 2] * norm[2 * hm::kIn + hm::kOut] + norm[2 * hm::kIn];
    const double seed = D * exp((double)fminf(fmaxf(y0, -60.0f), 60.0f));
    hd::Geom g;
    hd::load_geom(geom, z, n, g);
    double Uz[HARM_NPR], pz[HARM_NPR], pg[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) {
      Uz[pl] = U[pl * n + z];
      pg[pl] = p[pl * n + z];
      pz[pl] = pg[pl];
    }
    int ns = 0;
    int fl = hd::utoprimgen(Uz, g, gam, gammafail, pz, ns, seed);
    if (fl != 0) {  // the plain guess decides
      fb = 1;
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) pz[pl] = pg[pl];
      fl = hd::utoprimgen(Uz, g, gam, gammafail, pz, ns, __builtin_nan(""));
    }
    if (hd::needs_after_hot(fl, true, false)) {  // the cold retry, as harm_invert_batch
      double Ugf[HARM_NPR];
      hd::geometry_free(Uz, g, Ugf);
      fl = hd::after_hot(fl, Ugf, g, gam, gammafail, pg, pz, ns, true, false);
    }
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) p[pl * n + z] = pz[pl];
    pflag[z] = fl;
    if (niter) niter[z] = ns;
  }
  if (nfallback) {
    __shared__ unsigned int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    if (fb) atomicAdd(&cnt, 1u);
    __syncthreads();
    if (threadIdx.x == 0 && cnt) atomicAdd(nfallback, (unsigned long long)cnt);
  }
}

// deterministic host initialisation: Glorot-uniform weights, zero biases
void init_params(std::vector<float>& w, uint64_t seed)
{
  uint64_t s = seed * 0x9E3779B97F4A7C15ull + 0x1234567ull;
  auto uni = [&]() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t x = s;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
  };
  w.assign(hm::kNParam, 0.0f);
  auto fill = [&](int off, int fin, int fout) {
    const double a = sqrt(6.0 / (fin + fout));
    for (int e = 0; e < fin * fout; e++) w[off + e] = (float)((2.0 * uni() - 1.0) * a);
  };
  fill(hm::kOffW1, hm::kIn, hm::kHid);
  fill(hm::kOffW2, hm::kHid, hm::kHid);
  fill(hm::kOffW3, hm::kHid, hm::kOut);
}

harm_status pack(harm_mlp* m, hipStream_t s)
{
  hipLaunchKernelGGL(k_mlp_pack, dim3(8), dim3(kBlock), 0, s, m->w32, m->wbf);
  return hh::check_launch("k_mlp_pack");
}

// The step's inversion with the learned guess (harm_set_inversion_guess):
// the zones of k_invert in the same order; per wave the features of 64 zones
// go through the network on the matrix cores, each zone's Newton solve then
// starts from the predicted W' (falling back to the previous primitives when
// it fails), and the final substep's floors follow in the same kernel
// (floors_zone, as the plain inversion's floors pass applies them; DIAG:
// keeping the states the fixes started from).
template <bool DIAG>
__global__ __launch_bounds__(kBlock) void k_invert_grid_mlp(const uint16_t* __restrict__ wbf,
                                                            const float* __restrict__ w32,
                                                            const float* __restrict__ norm, hd::Dims d, hk::Params P,
                                                            const double* __restrict__ U,
                                                            const double* __restrict__ pg, double* __restrict__ pf,
                                                            int32_t* __restrict__ pflag,
                                                            const double* __restrict__ gcent,
                                                            const double* __restrict__ floors, int finalstep,
                                                            unsigned long long* __restrict__ di,
                                                            double* __restrict__ fx_p0, double* __restrict__ fx_pfl,
                                                            uint8_t* __restrict__ fx_flag)
{
  __shared__ Fwd S;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  stage_weights(wbf, w32, S.sW, S.sBias);
  const int lb = hk::xcd_tile(blockIdx.x, gridDim.x);
  const int t = lb * kBlock + threadIdx.x;
  const bool act = t < d.n1 * d.n2 * d.n3;
  long o = 0, pp = 0;
  double D = 1.0;
  {
    float f[hm::kIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (act) {
      int i, j, k;
      hk::zone_ijk(d, t, i, j, k);
      o = hd::cidx(d, i, j, k);
      pp = hd::pidx(d, i, j);
      hd::Geom g;
      hd::load_geom(gcent, pp, d.nplane, g);
      const double ig = (g.gdet >= 0.0 ? 1.0 : -1.0) / (fabs(g.gdet) + D_SMALL);
      double Ut[HARM_NPR];
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) Ut[pl] = g.alpha * (U[pl * d.ncell + o] * ig);
      hd::Inv s;
      double Qtcon[4], Bcon[4];
      hd::inv_setup(Ut, g, P.gam, P.gammafail, s, Qtcon, Bcon);
      hm::features(s, f);
      D = s.D;
    }
#pragma unroll
    for (int k = 0; k < hm::kInPad; k++) {
      float v = 0.0f;
      if (act && k < hm::kIn) v = (f[k] - norm[k]) * norm[hm::kIn + k];
      S.bufA[w][l * hm::kInPad + k] = hm::f2bf(v);
    }
  }
  __syncthreads();
  wave_forward(S.bufA[w], S.bufB[w], S.sW, S.sBias, S.y[w]);
  int ns = 0, nf = 0, fb = 0;
  if (act) {
    const float y0 = S.y[w][l * 2] * norm[2 * hm::kIn + hm::kOut] + norm[2 * hm::kIn];
    const double seed = D * exp((double)fminf(fmaxf(y0, -60.0f), 60.0f));
    double Uz[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) Uz[pl] = U[pl * d.ncell + o];
    nf = hk::invert_zone<true, DIAG>(d, P, Uz, U, pg, pf, pflag, gcent, floors, finalstep, o, pp, ns, seed, &fb,
                                     fx_p0, fx_pfl, fx_flag) > 0;
  }
  hk::newton_counts_to(ns, nf, di);
  __shared__ unsigned int nfb;
  if (threadIdx.x == 0) nfb = 0;
  __syncthreads();
  if (fb) atomicAdd(&nfb, 1u);
  __syncthreads();
  if (threadIdx.x == 0 && nfb) atomicAdd(&di[hk::DI_NFALLBACK], (unsigned long long)nfb);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" harm_status harm_mlp_create(uint64_t seed, harm_mlp** out)
{
  if (!out) return hh::set_error(HARM_EINVAL, "harm_mlp_create: null argument");
  harm_mlp* m = new harm_mlp();
  auto fail = [&](harm_status st) {
    harm_mlp_destroy(m);
    return st;
  };
  harm_status st;
  if ((st = hh::check_hip(hipMalloc((void**)&m->w32, sizeof(float) * hm::kNParam), "hipMalloc")) != HARM_OK ||
      (st = hh::check_hip(hipMalloc((void**)&m->adam_m, sizeof(float) * hm::kNParam), "hipMalloc")) != HARM_OK ||
      (st = hh::check_hip(hipMalloc((void**)&m->adam_v, sizeof(float) * hm::kNParam), "hipMalloc")) != HARM_OK ||
      (st = hh::check_hip(hipMalloc((void**)&m->wbf, sizeof(uint16_t) * hm::kNBf), "hipMalloc")) != HARM_OK ||
      (st = hh::check_hip(hipMalloc((void**)&m->norm, sizeof(float) * hm::kNNorm), "hipMalloc")) != HARM_OK ||
      (st = hh::check_hip(hipMalloc((void**)&m->loss, sizeof(float) * 4), "hipMalloc")) != HARM_OK)
    return fail(st);
  std::vector<float> w;
  init_params(w, seed);
  std::vector<float> z(hm::kNParam, 0.0f);
  m->hnorm.assign(hm::kNNorm, 0.0f);
  for (int k = 0; k < hm::kIn; k++) m->hnorm[hm::kIn + k] = 1.0f;
  for (int o = 0; o < hm::kOut; o++) m->hnorm[2 * hm::kIn + hm::kOut + o] = 1.0f;
  if ((st = hh::check_hip(hipMemcpy(m->w32, w.data(), sizeof(float) * hm::kNParam, hipMemcpyHostToDevice), "init")) !=
          HARM_OK ||
      (st = hh::check_hip(hipMemcpy(m->adam_m, z.data(), sizeof(float) * hm::kNParam, hipMemcpyHostToDevice),
                          "init")) != HARM_OK ||
      (st = hh::check_hip(hipMemcpy(m->adam_v, z.data(), sizeof(float) * hm::kNParam, hipMemcpyHostToDevice),
                          "init")) != HARM_OK ||
      (st = hh::check_hip(hipMemcpy(m->norm, m->hnorm.data(), sizeof(float) * hm::kNNorm, hipMemcpyHostToDevice),
                          "init")) != HARM_OK ||
      (st = pack(m, nullptr)) != HARM_OK)
    return fail(st);
  if ((st = hh::check_hip(hipDeviceSynchronize(), "harm_mlp_create")) != HARM_OK) return fail(st);
  *out = m;
  return HARM_OK;
}

extern "C" harm_status harm_mlp_destroy(harm_mlp* m)
{
  if (!m) return HARM_OK;
  void* bufs[] = {m->w32, m->adam_m, m->adam_v, m->wbf, m->norm, m->loss};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete m;
  return HARM_OK;
}

extern "C" harm_status harm_mlp_get_params(harm_mlp* m, float* params, float* norm)
{
  if (!m) return hh::set_error(HARM_EINVAL, "harm_mlp_get_params: null handle");
  if (params) HH_HIP(hipMemcpy(params, m->w32, sizeof(float) * hm::kNParam, hipMemcpyDeviceToHost));
  if (norm) memcpy(norm, m->hnorm.data(), sizeof(float) * hm::kNNorm);
  return HARM_OK;
}

extern "C" harm_status harm_mlp_set_params(harm_mlp* m, const float* params, const float* norm)
{
  if (!m) return hh::set_error(HARM_EINVAL, "harm_mlp_set_params: null handle");
  if (params) {
    HH_HIP(hipMemcpy(m->w32, params, sizeof(float) * hm::kNParam, hipMemcpyHostToDevice));
    HH_TRY(pack(m, nullptr));
  }
  if (norm) {
    for (int k = 0; k < hm::kIn; k++)
      if (!(norm[hm::kIn + k] > 0.0f)) return hh::set_error(HARM_EINVAL, "harm_mlp_set_params: feature scale <= 0");
    for (int o = 0; o < hm::kOut; o++)
      if (!(norm[2 * hm::kIn + hm::kOut + o] > 0.0f))
        return hh::set_error(HARM_EINVAL, "harm_mlp_set_params: target spread <= 0");
    m->hnorm.assign(norm, norm + hm::kNNorm);
    HH_HIP(hipMemcpy(m->norm, norm, sizeof(float) * hm::kNNorm, hipMemcpyHostToDevice));
  }
  return hh::check_hip(hipDeviceSynchronize(), "harm_mlp_set_params");
}

extern "C" harm_status harm_mlp_pairs(const double* U, const double* geom, const double* p, int64_t n, double gam,
                                      double gammafail, float* X, float* T, int32_t* ok, void* stream)
{
  if (n <= 0 || !U || !geom || !p || !X || !T || !ok) return hh::set_error(HARM_EINVAL, "harm_mlp_pairs: bad arguments");
  const long nb = (long)((n + kBlock - 1) / kBlock);
  if (nb > 0x7fffffffL) return hh::set_error(HARM_EINVAL, "harm_mlp_pairs: n too large");
  hipLaunchKernelGGL(k_mlp_pairs, dim3((unsigned)nb), dim3(kBlock), 0, (hipStream_t)stream, U, geom, p, (long)n, gam,
                     gammafail, X, T, ok);
  return hh::check_launch("k_mlp_pairs");
}

extern "C" harm_status harm_mlp_fit_norm(harm_mlp* m, const float* X, const float* T, const int32_t* ok, int64_t n)
{
  if (!m || !X || !T || !ok || n <= 0) return hh::set_error(HARM_EINVAL, "harm_mlp_fit_norm: bad arguments");
  std::vector<float> hx((size_t)n * hm::kIn), ht((size_t)n * hm::kOut);
  std::vector<int32_t> hok(n);
  HH_HIP(hipMemcpy(hx.data(), X, hx.size() * sizeof(float), hipMemcpyDeviceToHost));
  HH_HIP(hipMemcpy(ht.data(), T, ht.size() * sizeof(float), hipMemcpyDeviceToHost));
  HH_HIP(hipMemcpy(hok.data(), ok, hok.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  double s1[hm::kIn + hm::kOut] = {0}, s2[hm::kIn + hm::kOut] = {0};
  long cnt = 0;
  for (int64_t z = 0; z < n; z++) {
    if (!hok[z]) continue;
    cnt++;
    for (int k = 0; k < hm::kIn; k++) { s1[k] += hx[z * hm::kIn + k]; s2[k] += (double)hx[z * hm::kIn + k] * hx[z * hm::kIn + k]; }
    for (int o = 0; o < hm::kOut; o++) {
      s1[hm::kIn + o] += ht[z * 2 + o];
      s2[hm::kIn + o] += (double)ht[z * 2 + o] * ht[z * 2 + o];
    }
  }
  if (cnt < 2) return hh::set_error(HARM_EINVAL, "harm_mlp_fit_norm: fewer than two usable pairs");
  std::vector<float> nm(hm::kNNorm);
  for (int k = 0; k < hm::kIn + hm::kOut; k++) {
    const double mean = s1[k] / cnt, var = std::max(s2[k] / cnt - mean * mean, 1e-12);
    const double sd = sqrt(var);
    if (k < hm::kIn) { nm[k] = (float)mean; nm[hm::kIn + k] = (float)(1.0 / sd); }
    else { nm[2 * hm::kIn + (k - hm::kIn)] = (float)mean; nm[2 * hm::kIn + hm::kOut + (k - hm::kIn)] = (float)sd; }
  }
  return harm_mlp_set_params(m, nullptr, nm.data());
}

extern "C" harm_status harm_mlp_train(harm_mlp* m, const float* X, const float* T, int64_t nx, const int32_t* order,
                                      int64_t norder, int64_t first, int32_t nsteps, float lr, float* loss,
                                      void* stream)
{
  if (!m || !X || !T || nx <= 0 || !order || norder <= 0 || nsteps <= 0 || !(lr > 0.0f) || first < 0)
    return hh::set_error(HARM_EINVAL, "harm_mlp_train: bad arguments");
  TrainArgs a{};
  a.w32 = m->w32; a.am = m->adam_m; a.av = m->adam_v; a.wbf = m->wbf; a.norm = m->norm;
  a.X = X; a.T = T; a.nx = (long)nx; a.idx = order; a.nidx = (long)norder; a.first = (long)first; a.nsteps = nsteps;
  a.t0 = (long)m->t;
  a.lr = lr; a.beta1 = 0.9f; a.beta2 = 0.999f; a.eps = 1e-8f; a.update = 1; a.grad_out = nullptr; a.loss_out = m->loss;
  hipLaunchKernelGGL(k_mlp_train, dim3(1), dim3(kBlock), 0, (hipStream_t)stream, a);
  HH_TRY(hh::check_launch("k_mlp_train"));
  m->t += nsteps;
  if (loss) {
    HH_HIP(hipStreamSynchronize((hipStream_t)stream));
    HH_HIP(hipMemcpy(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost));
  }
  return HARM_OK;
}

extern "C" harm_status harm_mlp_grad(harm_mlp* m, const float* X, const float* T, int64_t nx, const int32_t* order,
                                     float* grad, float* loss)
{
  if (!m || !X || !T || nx <= 0 || !order || !grad) return hh::set_error(HARM_EINVAL, "harm_mlp_grad: bad arguments");
  TrainArgs a{};
  a.w32 = m->w32; a.am = m->adam_m; a.av = m->adam_v; a.wbf = m->wbf; a.norm = m->norm;
  a.X = X; a.T = T; a.nx = (long)nx; a.idx = order; a.nidx = kBatch; a.first = 0; a.nsteps = 1; a.t0 = 0;
  a.lr = 0.0f; a.beta1 = 0.9f; a.beta2 = 0.999f; a.eps = 1e-8f; a.update = 0; a.grad_out = grad; a.loss_out = m->loss;
  hipLaunchKernelGGL(k_mlp_train, dim3(1), dim3(kBlock), 0, nullptr, a);
  HH_TRY(hh::check_launch("k_mlp_train"));
  HH_HIP(hipDeviceSynchronize());
  if (loss) HH_HIP(hipMemcpy(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost));
  return HARM_OK;
}

extern "C" harm_status harm_mlp_forward(harm_mlp* m, const float* X, float* Y, int64_t n, void* stream)
{
  if (!m || !X || !Y || n <= 0) return hh::set_error(HARM_EINVAL, "harm_mlp_forward: bad arguments");
  const long nb = (long)((n + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)nb), dim3(kBlock), 0, (hipStream_t)stream, m->wbf, m->w32, m->norm,
                     X, Y, (long)n);
  return hh::check_launch("k_mlp_forward");
}

extern "C" harm_status harm_invert_batch_mlp(harm_mlp* m, const double* U, const double* geom, double* p_inout,
                                             int32_t* pflag, int32_t* niter, int64_t n, double gam, double gammafail,
                                             unsigned long long* nfallback, void* stream)
{
  if (!m || n <= 0 || !U || !geom || !p_inout || !pflag)
    return hh::set_error(HARM_EINVAL, "harm_invert_batch_mlp: bad arguments");
  const long nb = (long)((n + kBlock - 1) / kBlock);
  if (nb > 0x7fffffffL) return hh::set_error(HARM_EINVAL, "harm_invert_batch_mlp: n too large");
  hipLaunchKernelGGL(k_invert_batch_mlp, dim3((unsigned)nb), dim3(kBlock), 0, (hipStream_t)stream, m->wbf, m->w32,
                     m->norm, U, geom, p_inout, pflag, niter, (long)n, gam, gammafail, nfallback);
  return hh::check_launch("k_invert_batch_mlp");
}

harm_status hk::launch_invert_grid_mlp(const harm_mlp* m, const hd::Dims& d, const hk::Params& P, const double* U,
                                       const double* pg, double* pf, int32_t* pflag, const double* gcent,
                                       const double* floors, int finalstep, unsigned long long* di, hipStream_t s,
                                       double* fx_p0, double* fx_pfl, uint8_t* fx_flag)
{
  const long nb = ((long)d.n1 * d.n2 * d.n3 + kBlock - 1) / kBlock;
  if (fx_p0)
    hipLaunchKernelGGL(k_invert_grid_mlp<true>, dim3((unsigned)nb), dim3(kBlock), 0, s, m->wbf, m->w32, m->norm, d, P,
                       U, pg, pf, pflag, gcent, floors, finalstep, di, fx_p0, fx_pfl, fx_flag);
  else
    hipLaunchKernelGGL(k_invert_grid_mlp<false>, dim3((unsigned)nb), dim3(kBlock), 0, s, m->wbf, m->w32, m->norm, d,
                       P, U, pg, pf, pflag, gcent, floors, finalstep, di, nullptr, nullptr, nullptr);
  return hh::check_launch("k_invert_grid_mlp");
}
