// mlp.h -- the learned initial guess for the conserved-to-primitive inversion:
// an 8 -> 64 -> 64 -> 2 perceptron (tanh hidden layers) evaluated with BF16
// operands on the matrix cores (v_mfma_f32_16x16x16_bf16, FP32 accumulation).
//
// Inputs are eight scale-free functions of the invariants the 1-D W' Newton
// solve starts from (D, B^2, Q~^2, (Q.B)^2, Q.n; inv_setup in dphys.h); the
// two outputs are log(W'/D) and log(u~^2) of the solution, both standardised
// with the training set's mean and spread.  The Newton solve in FP64 stays the
// judge: the network only replaces the starting W'.
//
// The matrix helpers work on one wavefront: every lane of the wave must call
// them (the matrix instruction reads all 64 lanes), operands are BF16 in LDS
// or global memory with arbitrary element strides, so transposed products
// (X^T G, G W^T) need no copies.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "dphys.h"

namespace hm {

constexpr int kIn = 8;     // input features
constexpr int kInPad = 16; // K of the first product, padded to the MFMA depth
constexpr int kHid = 64;   // hidden width
constexpr int kOut = 2;    // outputs
constexpr int kOutPad = 16;

// Parameter block (FP32 master copy), row-major [in][out]:
// W1[8][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2]
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kIn * kHid;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kOut;
constexpr int kNParam = kOffB3 + kOut;  // 4866

// BF16 weight images as the products read them (zero padded):
// W1 [16][64], W2 [64][64], W3 [64][16]
constexpr int kBfW1 = 0;
constexpr int kBfW2 = kBfW1 + kInPad * kHid;
constexpr int kBfW3 = kBfW2 + kHid * kHid;
constexpr int kNBf = kBfW3 + kHid * kOutPad;  // 6144 bf16 values

// standardisation: feature mean[8], feature scale[8] (1/std), target mean[2], target std[2]
constexpr int kNNorm = 2 * kIn + 2 * kOut;

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (NaN kept quiet)
__host__ __device__ inline uint16_t f2bf(float f)
{
  union { float f; uint32_t u; } v;
  v.f = f;
  if ((v.u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((v.u >> 16) | 0x40u);
  const uint32_t r = ((v.u >> 16) & 1u) + 0x7fffu;
  return (uint16_t)((v.u + r) >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b)
{
  union { float f; uint32_t u; } v;
  v.u = (uint32_t)b << 16;
  return v.f;
}

// acc += A (16 x K) * B (K x 16) for one wavefront, K a multiple of 16.
// A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn].
// Lane l holds A(l%16, k0 + 4*(l/16) + i) and B(k0 + 4*(l/16) + i, l%16) for
// i = 0..3, and returns acc[i] = D(4*(l/16) + i, l%16).
__device__ __forceinline__ v4f mma16(const uint16_t* A, int sam, int sak, const uint16_t* B, int sbk, int sbn, int K,
                                     v4f acc)
{
  const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
  for (int k0 = 0; k0 < K; k0 += 16) {
    v4s a, b;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = k0 + 4 * q + i;
      a[i] = (short)A[r * sam + k * sak];
      b[i] = (short)B[k * sbk + r * sbn];
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  return acc;
}

__host__ __device__ inline double slog(double x) { return x >= 0.0 ? log1p(x) : -log1p(-x); }

// The eight raw features of an inversion problem (before standardisation).
// All are ratios to powers of D, so they do not depend on the density scale.
__device__ __forceinline__ void features(const hd::Inv& s, float* f)
{
  const double D = s.D > 0.0 ? s.D : D_SMALL;
  const double e = s.Qdotnp / D, q2 = s.Qtsq / (D * D), b2 = s.Bsq / D;
  const double qb2 = s.QdotBsq / (D * D * D);
  const double q = sqrt(fmax(q2, 0.0));
  double wcold = 0.0;  // the cold-hydro W'/D, a closed-form first estimate
  {
    const double det = 1.0 + q2;
    wcold = q2 / (1.0 + sqrt(det));
  }
  f[0] = (float)slog(e);
  f[1] = (float)slog(q2);
  f[2] = (float)slog(b2);
  f[3] = (float)slog(qb2);
  f[4] = (float)(q / (fabs(e) + 1.0 + q));
  f[5] = (float)(b2 / (b2 + fabs(e) + 1.0));
  f[6] = (float)(s.QdotBsq / (s.Qtsq * s.Bsq + D_SMALL));
  f[7] = (float)slog(wcold);
}

// W' of a primitive state in the frame of the solve (set_guess_Wp,
// utoprim_jon.c:1346): D u~^2/(1+gamma) + (u + p) gamma^2.
__device__ __forceinline__ void wp_of_prim(const double* prim, const hd::Geom& g, double D, double gam, double& Wp,
                                           double& utsq)
{
  utsq = 0.;
#pragma unroll
  for (int i = 1; i < 4; i++)
#pragma unroll
    for (int j = 1; j < 4; j++) utsq += hd::gco(g, i, j) * prim[HARM_U1 + i - 1] * prim[HARM_U1 + j - 1];
  const double gammasq = 1. + utsq;
  const double gamma = sqrt(gammasq);
  const double u = prim[HARM_UU];
  const double p = (gam - 1.0) * u;
  Wp = D * utsq / (1.0 + gamma) + (u + p) * gammasq;
}

// Floors of the target transforms (log of non-negative quantities).
constexpr double kTinyWp = 1e-14, kTinyUt = 1e-14;

}  // namespace hm
