// dphys.h -- per-zone GRMHD physics as gfx950 device functions (FP64).
//
// Everything here runs in registers for one zone or one face state; the
// kernels that call it own the memory layout.  Operation order follows the
// reference expressions so that a build with -ffp-contract=off reproduces the
// CPU results bit for bit.
//
// Reference (HARM): get_state phys.tools.c:1513, primtoflux :30,
// mhd_calc_norestmass_ma :1071, mhd_calc_primfield_em :1133, vchar vchar.c:34,
// simplefast :260, source_conn phys.c:132, Utoprim_jon utoprim_jon.c:136-4520.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <float.h>
#include "harmstep.h"

#define D_SMALL 1.e-300
#define D_KINDASMALL 1.e-100
#define D_VERYBIG 1.e150
#define D_NUMEPSILON DBL_EPSILON
#define D_SQRTMINNUMREPRESENT 1.4916681462400413e-154

namespace hd {

// Geometry of one grid point, unpacked from the 27-field record.
struct Geom {
  double gcov[10];  // packed symmetric
  double gcon[10];
  double gpert_tt, gdet, alpha, bsqoa2, beta1, beta2, beta3;
};

__host__ __device__ constexpr int sym(int a, int b)
{
  return (a <= b) ? ((a == 0) ? b : (a == 1) ? 3 + b : (a == 2) ? 5 + b : 9)
                  : ((b == 0) ? a : (b == 1) ? 3 + a : (b == 2) ? 5 + a : 9);
}

// A copy of a global-memory pointer that the compiler cannot see through (an
// empty asm).  Loads through it are not merged with earlier loads of the same
// addresses, so a value read before a register-hungry stretch of code is read
// again afterwards (an L1/L2 hit) instead of being held in registers across
// it.  The copy stays in the global address space, so the loads remain
// global_load (a generic pointer would give flat_load).  The pointer must be
// the same in every lane of the wave (a base address); it is taken from the
// first lane into scalar registers.
__device__ __forceinline__ const double* fresh(const double* p)
{
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  auto g = (const __attribute__((address_space(1))) double*)(((unsigned long long)hi << 32) | lo);
  asm volatile("" : "+s"(g));
  return (const double*)g;
}

// 1/x for |x| in the normal range and far from overflow: the hardware
// reciprocal estimate refined by two Newton steps (within an ulp of the IEEE
// quotient, without the scaling and special-value steps of the IEEE division
// sequence, which such an x never needs).  Used only by the face-flux
// kernels (the FAST paths below), whose results are compared with a
// tolerance; everything bit-exact divides.
__device__ __forceinline__ double recip_nr(double x)
{
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  r = fma(fma(-x, r, 1.0), r, r);
  return r;
}

// Load a record stored SoA with field stride `fs` starting at `base`.
__device__ __forceinline__ void load_geom(const double* __restrict__ tab, long base, long fs, Geom& g)
{
#pragma unroll
  for (int f = 0; f < 10; f++) g.gcov[f] = tab[base + f * fs];
#pragma unroll
  for (int f = 0; f < 10; f++) g.gcon[f] = tab[base + (10 + f) * fs];
  g.gpert_tt = tab[base + HARM_G_GCOVPERT_TT * fs];
  g.gdet = tab[base + HARM_G_GDET * fs];
  g.alpha = tab[base + HARM_G_ALPHA * fs];
  g.bsqoa2 = tab[base + HARM_G_BETASQOALPHASQ * fs];
  g.beta1 = tab[base + HARM_G_BETA1 * fs];
  g.beta2 = tab[base + HARM_G_BETA2 * fs];
  g.beta3 = tab[base + HARM_G_BETA3 * fs];
}

__device__ __forceinline__ double gco(const Geom& g, int a, int b) { return g.gcov[sym(a, b)]; }
__device__ __forceinline__ double gcn(const Geom& g, int a, int b) { return g.gcon[sym(a, b)]; }

__device__ __forceinline__ void lower(const double* v, const Geom& g, double* w)
{
#pragma unroll
  for (int a = 0; a < 4; a++) w[a] = gco(g, a, 0) * v[0] + gco(g, a, 1) * v[1] + gco(g, a, 2) * v[2] + gco(g, a, 3) * v[3];
}
__device__ __forceinline__ void raise(const double* v, const Geom& g, double* w)
{
#pragma unroll
  for (int a = 0; a < 4; a++) w[a] = gcn(g, a, 0) * v[0] + gcn(g, a, 1) * v[1] + gcn(g, a, 2) * v[2] + gcn(g, a, 3) * v[3];
}

struct State {
  double ucon[4], ucov[4];
  double gamma, qsq, pressure, bsq, p1ud0;
  double iu0;  // 1/u^0, set by get_state<true> only
};

// The functions below take FAST = true only in the face-flux kernels: their
// divisions by quantities that are positive and normal by construction
// (alpha, 1 + gamma alpha, u^0, the characteristic quadratic's leading
// coefficient, the fast-speed denominator >= SMALL) become multiplications by
// recip_nr, and 1/u^0 is formed once per state.  FAST = false is the
// reference's operation order, bit for bit.

// get_state for VELREL4 primitives (ucon_calc_rel4vel, compute_1plusud0_rel4vel,
// bsq_calc_fromq_rel4vel, ideal-gas pressure).
template <bool FAST = false>
__device__ __forceinline__ void get_state(const double* pr, const Geom& g, double gam, State& q)
{
  const double u1 = pr[HARM_U1], u2 = pr[HARM_U2], u3 = pr[HARM_U3];
  double qsq = gco(g, 1, 1) * u1 * u1 + gco(g, 2, 2) * u2 * u2 + gco(g, 3, 3) * u3 * u3
               + 2. * (gco(g, 1, 2) * u1 * u2 + gco(g, 1, 3) * u1 * u3);
  qsq = qsq * (qsq >= 0.0);
  const double gamma = sqrt(1. + qsq);
  q.qsq = qsq;
  q.gamma = gamma;
  q.ucon[0] = FAST ? gamma * recip_nr(g.alpha) : gamma / g.alpha;
  q.ucon[1] = u1 - q.ucon[0] * g.beta1;
  q.ucon[2] = u2 - q.ucon[0] * g.beta2;
  q.ucon[3] = u3 - q.ucon[0] * g.beta3;
  lower(q.ucon, g, q.ucov);
  double ud0tilde = 0.0;
  ud0tilde += u1 * gco(g, 0, 1);
  ud0tilde += u2 * gco(g, 0, 2);
  ud0tilde += u3 * gco(g, 0, 3);
  const double alpha = g.alpha, alphasq = alpha * alpha;
  q.p1ud0 = FAST ? ud0tilde + (g.gpert_tt - alphasq * (g.bsqoa2 + qsq)) * recip_nr(1.0 + gamma * alpha)
                 : ud0tilde + (g.gpert_tt - alphasq * (g.bsqoa2 + qsq)) / (1.0 + gamma * alpha);
  const double Bcon1 = pr[HARM_B1], Bcon2 = pr[HARM_B2], Bcon3 = pr[HARM_B3];
  double Bcov[4];
#pragma unroll
  for (int jj = 0; jj < 4; jj++) {
    Bcov[jj] = 0.0;
    Bcov[jj] += Bcon1 * gco(g, jj, 1);
    Bcov[jj] += Bcon2 * gco(g, jj, 2);
    Bcov[jj] += Bcon3 * gco(g, jj, 3);
  }
  double Bsq = 0.0, udotB = 0.0;
  Bsq += Bcon1 * Bcov[1]; Bsq += Bcon2 * Bcov[2]; Bsq += Bcon3 * Bcov[3];
  udotB += Bcon1 * q.ucov[1]; udotB += Bcon2 * q.ucov[2]; udotB += Bcon3 * q.ucov[3];
  double bsq = (Bsq + udotB * udotB);
  if (FAST) {
    q.iu0 = recip_nr(q.ucon[0]);
    bsq *= q.iu0 * q.iu0;
  } else {
    bsq *= 1.0 / (q.ucon[0] * q.ucon[0]);
  }
  q.bsq = bsq;
  q.pressure = (gam - 1.0) * pr[HARM_UU];
}

// One row T^dir_mu of the rest-mass-removed stress tensor: matter + EM parts
// kept separate so callers can add them in the reference order.
template <bool FAST = false>
__device__ __forceinline__ void mhd_parts(const double* pr, int dir, const Geom& g, const State& q, double* ma, double* em)
{
  const double rho = pr[HARM_RHO], u = pr[HARM_UU], P = q.pressure;
  const double w = P + rho + u;
  const double ud = q.ucon[dir];
  ma[0] = (P + u) * ud * q.ucov[0] + rho * ud * q.p1ud0;
#pragma unroll
  for (int j = 1; j < 4; j++) ma[j] = w * ud * q.ucov[j];
  ma[dir] += P;
  const double Bc[4] = {0.0, pr[HARM_B1], pr[HARM_B2], pr[HARM_B3]};
  double Bcov[4];
#pragma unroll
  for (int jj = 0; jj < 4; jj++) {
    Bcov[jj] = 0.0;
    Bcov[jj] += Bc[1] * gco(g, jj, 1);
    Bcov[jj] += Bc[2] * gco(g, jj, 2);
    Bcov[jj] += Bc[3] * gco(g, jj, 3);
  }
  double Bsq = 0.0, udotB = 0.0;
  Bsq += Bc[1] * Bcov[1]; Bsq += Bc[2] * Bcov[2]; Bsq += Bc[3] * Bcov[3];
  udotB += Bc[1] * q.ucov[1]; udotB += Bc[2] * q.ucov[2]; udotB += Bc[3] * q.ucov[3];
  const double Bd = Bc[dir];
#pragma unroll
  for (int mu = 0; mu < 4; mu++)
    em[mu] = (Bsq * ud * (q.ucov[mu]) - Bd * Bcov[mu] - udotB * (Bd * (q.ucov[mu]) + Bcov[mu] * ud));
  em[dir] += 0.5 * (Bsq + udotB * udotB);
  const double ogs = FAST ? q.iu0 * q.iu0 : 1.0 / ((q.ucon[0]) * (q.ucon[0]));
#pragma unroll
  for (int mu = 0; mu < 4; mu++) em[mu] *= ogs;
}

// primtoflux(UEVOLVE): flux (dir>0) or conserved state (dir=0), x gdet.
template <bool FAST = false>
__device__ __forceinline__ void primtoflux(const double* pr, const State& q, int dir, const Geom& g, double* F)
{
  double ma[4], em[4], fl[HARM_NPR];
  fl[HARM_RHO] = pr[HARM_RHO] * q.ucon[dir];
  mhd_parts<FAST>(pr, dir, g, q, ma, em);
#pragma unroll
  for (int j = 0; j < 4; j++) fl[HARM_UU + j] = ma[j] + em[j];
  if (dir > 0) {
    const double Bd = pr[HARM_B1 + dir - 1];
    if (FAST) {
      fl[HARM_B1] = (pr[HARM_B1] * q.ucon[dir] - Bd * q.ucon[1]) * q.iu0;
      fl[HARM_B2] = (pr[HARM_B2] * q.ucon[dir] - Bd * q.ucon[2]) * q.iu0;
      fl[HARM_B3] = (pr[HARM_B3] * q.ucon[dir] - Bd * q.ucon[3]) * q.iu0;
    } else {
      fl[HARM_B1] = (pr[HARM_B1] * q.ucon[dir] - Bd * q.ucon[1]) / q.ucon[0];
      fl[HARM_B2] = (pr[HARM_B2] * q.ucon[dir] - Bd * q.ucon[2]) / q.ucon[0];
      fl[HARM_B3] = (pr[HARM_B3] * q.ucon[dir] - Bd * q.ucon[3]) / q.ucon[0];
    }
  } else {
    fl[HARM_B1] = pr[HARM_B1];
    fl[HARM_B2] = pr[HARM_B2];
    fl[HARM_B3] = pr[HARM_B3];
  }
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) F[pl] = fl[pl] * g.gdet;
}

// vchar + simplefast: fast magnetosonic speeds along dir.
template <bool FAST = false>
__device__ __forceinline__ void vchar(const double* pr, const State& q, int dir, const Geom& g, double gam,
                                      double& vmax, double& vmin)
{
  const double bsq = q.bsq;
  double rho = pr[HARM_RHO], u = pr[HARM_UU];
  if (u < 0) u = 0;
  if (rho < 0) rho = D_SMALL;
  const double P = q.pressure;
  double WW = rho + u + P, cs2;
  if (WW > 0.0) {
    const double pressure = (gam - 1.0) * u;
    const double h = rho + u + pressure;  // > 0 here: h = WW when u >= 0, else h = rho > -P > 0
    cs2 = FAST ? gam * pressure * recip_nr(h) : gam * pressure / h;
  } else {
    WW = D_SMALL;
    cs2 = D_SMALL;
  }
  if (cs2 >= 1.0) cs2 = 1.0 - D_NUMEPSILON * 10.0;
  else if (cs2 <= 0.0) cs2 = D_SMALL;
  const double EF = D_SMALL + fabs(bsq + WW);
  const double va2 = FAST ? bsq * recip_nr(EF) : bsq / EF;
  double cms2 = cs2 + va2 - cs2 * va2;
  if (cms2 < 0.0) cms2 = 0.0;
  if (cms2 > 1.0) cms2 = 1.0;
  const double Asq = gcn(g, dir, dir);
  const double Bsq = gcn(g, 0, 0);
  const double Au = q.ucon[dir];
  const double Bu = q.ucon[0];
  const double AB = gcn(g, dir, 0);
  const double Au2 = Au * Au, Bu2 = Bu * Bu, AuBu = Au * Bu;
  const double B = 2. * (AuBu * (1.0 - cms2) - AB * cms2);
  const double A = Bu2 * (1.0 - cms2) - Bsq * cms2;
  double discr = 4.0 * cms2 * ((AB * AB - Asq * Bsq) * cms2 + (2.0 * AB * Au * Bu - Asq * Bu2 - Bsq * Au2) * (cms2 - 1.0));
  if (discr < 0.0) discr = 0.0;
  discr = sqrt(discr);
  double vm, vp;
  if (FAST) {  // A >= min(u0^2, -g^00) > 0
    const double i2A = recip_nr(2. * A);
    vm = -(-B + discr) * i2A;
    vp = -(-B - discr) * i2A;
  } else {
    vm = -(-B + discr) / (2. * A);
    vp = -(-B - discr) / (2. * A);
  }
  if (vp < vm) { vmax = vm; vmin = vp; }
  else { vmax = vp; vmin = vm; }
}

// source_conn: dU^{UU+l} = gdet (sum_{j,k} T^j_k Gamma^k_{lj} - rho u^j Gamma^t_{lj}).
// conn(c) returns packed Gamma^a_{bc} entry c = a*10 + sym(b,c).
template <class ConnF>
__device__ __forceinline__ void source(const double* pr, const Geom& g, const State& q, ConnF conn, double* dU)
{
  double mhd[4][4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    double ma[4], em[4];
    mhd_parts(pr, j, g, q, ma, em);
#pragma unroll
    for (int k = 0; k < 4; k++) mhd[j][k] = (ma[k] + 0.0) + em[k];
  }
  double dUc[4];
#pragma unroll
  for (int l = 0; l < 4; l++) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int k = 0; k < 4; k++) s += mhd[j][k] * conn(k * 10 + sym(l, j));
    dUc[l] = s;
  }
#pragma unroll
  for (int l = 0; l < 4; l++) {
    double s = dUc[l];
#pragma unroll
    for (int j = 0; j < 4; j++) s += -pr[HARM_RHO] * q.ucon[j] * conn(0 * 10 + sym(l, j));
    dUc[l] = s;
  }
  dU[HARM_RHO] = 0.0 + g.gdet * 0.0;
#pragma unroll
  for (int l = 0; l < 4; l++) dU[HARM_UU + l] = 0.0 + g.gdet * dUc[l];
  dU[HARM_B1] = 0.0 + g.gdet * 0.0;
  dU[HARM_B2] = 0.0 + g.gdet * 0.0;
  dU[HARM_B3] = 0.0 + g.gdet * 0.0;
}

// The same connection source, evaluated one stress-tensor row at a time so
// that only four T^j_k values are live: for each l the terms are added in the
// same (j, k) order as in source(), so the sums are identical.  Returns the
// four energy-momentum components (without the gdet factor applied to the
// zero rest-mass and field entries) times gdet, and T^0_0 (rest mass removed,
// the UU entry of U before the gdet factor).
template <class ConnF>
__device__ __forceinline__ void source_rows(const double* pr, const Geom& g, const State& q, ConnF conn,
                                            double* dUl, double& T00)
{
  double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < 4; j++) {
    double ma[4], em[4], row[4];
    mhd_parts(pr, j, g, q, ma, em);
#pragma unroll
    for (int k = 0; k < 4; k++) row[k] = (ma[k] + 0.0) + em[k];
    if (j == 0) T00 = ma[0] + em[0];
#pragma unroll
    for (int l = 0; l < 4; l++)
#pragma unroll
      for (int k = 0; k < 4; k++) s[l] += row[k] * conn(k * 10 + sym(l, j));
  }
#pragma unroll
  for (int l = 0; l < 4; l++) {
    double v = s[l];
#pragma unroll
    for (int j = 0; j < 4; j++) v += -pr[HARM_RHO] * q.ucon[j] * conn(0 * 10 + sym(l, j));
    dUl[l] = 0.0 + g.gdet * v;
  }
}

// ---------------------------------------------------------------------------
// Inversion: Jon's 1-D W' Newton (WHICHHOTINVERTER 3), no line search.
// ---------------------------------------------------------------------------
struct Inv {
  double Bsq, QdotB, QdotBsq, Qtsq, Qdotnp, D;
  double w0, w1, w2;
  double gam, gsq_too_big, utsq_too_big;
};

__device__ __forceinline__ double utsq_calc(double W, const Inv& s)
{
  const double S = s.QdotB, SoW = S / W;
  const double top = s.Qtsq + (s.Bsq + 2.0 * W) * SoW * SoW;
  const double bot = (s.Bsq * s.Bsq - s.Qtsq) + 2.0 * s.Bsq * W + W * W - SoW * SoW * (s.Bsq + 2.0 * W);
  return top / bot;
}

__device__ __forceinline__ bool vsqgtr1(double W, const Inv& s)
{
  return (W * W * W * (W + 2. * s.Bsq) - s.QdotBsq * (2. * W + s.Bsq) <= W * W * (s.Qtsq - s.Bsq * s.Bsq));
}

__device__ __forceinline__ int coldgrhd(double Qtsq, double D, double& Wp)
{
  const double det = D * D + Qtsq;
  if (det < 0.0) return 1;
  if (D > 0.0) Wp = Qtsq / (D + sqrt(det));
  else Wp = Qtsq / (D - sqrt(det));
  return 0;
}

// Residual E'(W') and its derivative (func_Eprime_opt).
__device__ __forceinline__ void func_eprime(double Wp, const Inv& s, double& dx, double& resid, double& jac, double& f,
                                            double& df)
{
  const double D = s.D, Bsq = s.Bsq;
  const double W = Wp + D;
  const double X = Bsq + W + D_SMALL;
  const double X2 = X * X, X3 = X2 * X;
  const double utsq = utsq_calc(W, s);
  const double gammasq = 1.0 + utsq;
  const double gamma = sqrt(gammasq);
  const double wmrho0 = (Wp - D * utsq / (1.0 + gamma)) / gammasq;
  if (utsq < -0.9 * s.utsq_too_big || utsq != utsq) {
    const double Ep = -D_VERYBIG;
    resid = Ep;
    jac = 0.0;
    dx = D_VERYBIG;
    f = 0.5 * (Ep * Ep);
    df = -2. * f;
    return;
  }
  const double igr = (s.gam - 1.0) / s.gam;
  const double pofchi = igr * wmrho0, dpdrho = 0.0, dpdchi = (s.gam - 1.0) / s.gam;
  double Ep = s.Qdotnp + Wp - pofchi + 0.5 * Bsq + 0.5 * (Bsq * s.Qtsq - s.QdotBsq) / X2;
  if (pofchi != pofchi || Ep != Ep) Ep = -D_VERYBIG;
  const double Xd = Bsq + W, Wsq = W * W, W3 = Wsq * W, Xd3 = Xd * Xd * Xd;
  const double dvsq = -2.0 / Xd3 * (s.Qtsq + s.QdotBsq * (3.0 * W * Xd + Bsq * Bsq) / W3);
  const double dwmrho0dW = 1.0 / gammasq;
  const double drho0dW = 0.0;
  const double dwmrho0dvsq = (D * (gamma * 0.5 - 1.0) - Wp);
  const double drho0dvsq = -D * gamma * 0.5;
  const double dpdW = drho0dW * dpdrho + dwmrho0dW * dpdchi;
  const double dpdvsq = drho0dvsq * dpdrho + dwmrho0dvsq * dpdchi;
  const double dpdWp = dpdW + dpdvsq * dvsq;
  double dE = (1.0 - dpdWp) - (Bsq * s.Qtsq - s.QdotBsq) / X3;
  if (dvsq != dvsq || dpdW != dpdW || dpdvsq != dpdvsq || dpdWp != dpdWp || dE != dE) dE = 0.0;
  resid = Ep;
  jac = dE;
  dx = -Ep / dE;
  f = 0.5 * (Ep * Ep);
  df = -2. * f;
}

// Cold GRMHD residual W^2 P^2 of the momentum equation and its W' derivative
// (func_Psq, utoprim_jon.c:3103, with Psq_Wp :1923 and dPsqdWp_Wp :1944).
__device__ __forceinline__ void func_psq(double Wp, const Inv& s, double& dx, double& resid, double& jac, double& f,
                                         double& df)
{
  const double D = s.D, Bsq = s.Bsq, Qtsq = s.Qtsq;
  const double W = Wp + D;
  const double Y = Bsq + 2.0 * W;
  const double W2 = W * W;
  const double S = s.QdotB;
  const double r = -Qtsq * W2 + Wp * (D + W) * (W2 + Bsq * Y) - Y * S * S;
  const double drdW = 2.0 * (D * ((Bsq + D) * (Bsq + D) - Qtsq) - (S * S) + ((Bsq + D) * (Bsq + 5.0 * D) - Qtsq) * Wp +
                             3.0 * (Bsq + 2.0 * D) * (Wp * Wp) + 2.0 * (Wp * Wp * Wp));
  resid = r;
  jac = drdW;
  dx = -r / drdW;
  f = 0.5 * r * r;
  df = -2. * f;
}

template <bool COLD>
__device__ __forceinline__ void residual(double Wp, const Inv& s, double& dx, double& resid, double& jac, double& f,
                                         double& df)
{
  if (COLD) func_psq(Wp, s, dx, resid, jac, f, df);
  else func_eprime(Wp, s, dx, resid, jac, f, df);
}

__device__ __forceinline__ bool ultrarel(double x0, const Inv& s) { return (fabs(s.w1) > s.w2) || (fabs(x0) > s.w2); }
__device__ __forceinline__ bool errcheck(double errx, double x0, const Inv& s)
{
  return ultrarel(x0, s) ? (fabs(errx) <= D_NUMEPSILON) : (fabs(errx) <= 1E3 * D_NUMEPSILON);
}
__device__ __forceinline__ int extracheck(double x0, const Inv& s) { return ultrarel(x0, s) ? 4 : 2; }

__device__ __forceinline__ void validate_wp(double Wpold, double& Wpnew, const Inv& s)
{
  if (Wpnew < -s.D && s.D > 0.0) {
    double Wp = 0.0;
    coldgrhd(s.Qtsq, s.D, Wp);
    Wpnew = Wp * 10.0;
  }
  Wpnew = (fabs(Wpnew / s.w1) > s.gsq_too_big) ? 0.5 * (Wpold + Wpnew) : Wpnew;
}

// general_newton_raphson for n=1 without line search (COLD: on the cold
// residual).  Returns retval.
template <bool COLD = false>
__device__ __forceinline__ int newton(int& pflag, double& xio, const Inv& s, int& nstroke)
{
  double x = xio, x_old = x, x_older = x, dx = 0.0, dx_old = D_VERYBIG, resid, jac, f = 1., df = 1.;
  double errx = 1., damp = 1.0;
  int n_iter = 0, i_extra = 0, lntries = 0;
  bool doing_extra = false, keep = true, diddamp = false;
  while (keep) {
    nstroke++;
    lntries++;
    if (lntries > 10 && !diddamp) damp = 0.5;
    residual<COLD>(x, s, dx, resid, jac, f, df);
    x_older = x_old;
    x_old = x;
    diddamp = false;
    if (n_iter == 0) {
      x_older = x;
      x_old = x;
      dx_old = dx;
    } else {
      if (resid <= -D_VERYBIG || jac == 0.0) {
        damp = 0.5 * damp;
        x = x_older;
        dx = dx_old;
        diddamp = true;
        residual<COLD>(x, s, dx, resid, jac, f, df);
        if (resid <= -D_VERYBIG || jac == 0.0) { xio = x; return 1; }
        x_older = x;
        x_old = x;
        dx_old = dx;
      }
    }
    x += damp * dx;
    errx = (s.w0 == 0.) ? fabs(dx_old) : fabs(dx_old / fmax(fabs(s.w0), fabs(x)));
    validate_wp(x_old, x, s);
    dx_old = dx;
    const bool ec = errcheck(errx, x, s);
    const int xc = extracheck(x, s);
    if (ec && !doing_extra && xc > 0) doing_extra = true;
    if (doing_extra) i_extra++;
    if ((ec && !doing_extra) || (i_extra > xc) || (n_iter >= 199)) keep = false;
    n_iter++;
  }
  if (!isfinite(x) || !isfinite(f) || !isfinite(df)) {
    pflag = 8;
    xio = x;
    return 1;
  }
  x = x_old;
  xio = x;
  if (errcheck(errx, x, s)) return 0;
  if (fabs(errx) > 1E-6) {
    pflag = 54;
    return 1;
  }
  return 0;
}

// Utoprim_new_body on alpha-scaled U (Ut) and guess (pt).  If wp_seed is
// finite it replaces the W' guess (MLP-seeded form); Newton stays the judge.
// The frame-independent invariants of the conserved state that the 1-D W'
// solve works with (the set-up part of Utoprim_new_body, utoprim_jon.c:716):
// D, B^2, the spatial momentum Q~^2, Q.B and Q.n (rest mass removed), plus
// Q~^mu and B^mu for the final primitive reconstruction.
__device__ __forceinline__ void inv_setup(const double* Ut, const Geom& g, double gam, double gammafail, Inv& s,
                                          double* Qtcon, double* Bcon)
{
  s.gam = gam;
  s.gsq_too_big = gammafail * gammafail;
  s.utsq_too_big = (gammafail - 1.0) * (gammafail - 1.0);
  s.D = Ut[HARM_RHO];
  Bcon[0] = 0.;
  Bcon[1] = Ut[HARM_B1];
  Bcon[2] = Ut[HARM_B2];
  Bcon[3] = Ut[HARM_B3];
  double Bcov[4];
  lower(Bcon, g, Bcov);
  s.Bsq = 0.;
  s.Bsq += Bcon[1] * Bcov[1]; s.Bsq += Bcon[2] * Bcov[2]; s.Bsq += Bcon[3] * Bcov[3];
  const double ncov[4] = {-g.alpha, 0., 0., 0.};
  double ncon[4];
  raise(ncov, g, ncon);
  const double Qcovp[4] = {Ut[HARM_UU], Ut[HARM_U1], Ut[HARM_U2], Ut[HARM_U3]};
  double Qconp[4];
  raise(Qcovp, g, Qconp);
  Qtcon[0] = 0.0; Qtcon[1] = 0.0; Qtcon[2] = 0.0; Qtcon[3] = 0.0;
#pragma unroll
  for (int j = 1; j < 4; j++)
#pragma unroll
    for (int i = 1; i < 4; i++) Qtcon[j] += Qcovp[i] * (gcn(g, i, j) + ncon[i] * ncon[j]);
  double Qtcov[4];
  lower(Qtcon, g, Qtcov);
  s.Qtsq = 0.;
#pragma unroll
  for (int i = 0; i < 4; i++) s.Qtsq += Qtcov[i] * Qtcon[i];
  const double Qcov[4] = {Qcovp[0] - s.D, Qcovp[1], Qcovp[2], Qcovp[3]};
  double Qcon[4];
  raise(Qcov, g, Qcon);
  s.QdotB = 0.;
  s.QdotB += Qcov[1] * Bcon[1]; s.QdotB += Qcov[2] * Bcon[2]; s.QdotB += Qcov[3] * Bcon[3];
  s.QdotBsq = s.QdotB * s.QdotB;
  {
    const double alpha = g.alpha, alphasq = alpha * alpha;
    const double Dfactor = (-g.gpert_tt + alphasq * (g.bsqoa2)) / (alphasq + alpha);
    s.Qdotnp = Qconp[0] * ncov[0] + (s.D) * (Dfactor);
  }
}

// COLD: the cold GRMHD solve (eomtype EOMCOLDGRMHD): the cold EOS (pressure 0,
// pickeos_eomtype eos.c:259, coldeos.c) in the guess and the result, u = 0.
template <bool COLD = false>
__device__ __forceinline__ int new_body(int& pflag, const double* Ut, const Geom& g, double gam, double gammafail,
                                        double* prim, int& nstroke, double wp_seed)
{
  Inv s;
  pflag = 0;
  prim[HARM_B1] = Ut[HARM_B1];
  prim[HARM_B2] = Ut[HARM_B2];
  prim[HARM_B3] = Ut[HARM_B3];
  double Qtcon[4], Bcon[4];
  inv_setup(Ut, g, gam, gammafail, s, Qtcon, Bcon);

  // set_guess_Wp + verify_Wlast
  double Wp_last, W_last;
  {
    double utsq = 0.;
#pragma unroll
    for (int i = 1; i < 4; i++)
#pragma unroll
      for (int j = 1; j < 4; j++) utsq += gco(g, i, j) * prim[HARM_U1 + i - 1] * prim[HARM_U1 + j - 1];
    if ((utsq < 0.) && (fabs(utsq) < 0.9 * s.utsq_too_big)) utsq = 0.0;
    if (!(utsq >= 0. && utsq < s.utsq_too_big)) utsq = 0.0;
    const double gammasq = 1. + utsq;
    const double gamma = sqrt(gammasq);
    const double rho0 = (s.D) / gamma;
    const double u = prim[HARM_UU];
    const double p = COLD ? 0.0 : (gam - 1.0) * u;
    const double BB[4] = {0.0, prim[HARM_B1], prim[HARM_B2], prim[HARM_B3]};
    double BBc[4];
    lower(BB, g, BBc);
    double bsq = 0.0;
    bsq += BB[1] * BBc[1]; bsq += BB[2] * BBc[2]; bsq += BB[3] * BBc[3];
    bsq *= ((g.alpha * g.alpha)) / gammasq;
    Wp_last = (s.D) * utsq / (1.0 + gamma) + (u + p) * gammasq;
    W_last = Wp_last + (s.D);
    const double Wpabs = fabs((s.D) * utsq) / (1.0 + fabs(gamma)) + (fabs(u) + fabs(p)) * fabs(gammasq) + D_SQRTMINNUMREPRESENT;
    const double nuabs = fabs(u) + fabs(p) + fabs(bsq);
    const double etaabs = fabs(rho0) + nuabs;
    s.w2 = 1e2 * fabs(rho0);
    if (gammasq * etaabs >= (s.w2)) {
      double Wpcoldhd = 0.0;
      if (coldgrhd(s.Qtsq, s.D, Wpcoldhd)) { pflag = 3; Wpcoldhd = Wp_last; }
      s.w0 = fmax(fmax(Wpcoldhd, Wp_last), Wpabs);
      Wp_last = fmax(2.0 * Wpcoldhd, Wp_last);
    } else {
      s.w0 = Wpabs;
    }
    s.w1 = fabs(rho0) + s.w0 + D_SQRTMINNUMREPRESENT;
    if (isfinite(wp_seed) && wp_seed > 0.0) {
      Wp_last = wp_seed;
      W_last = Wp_last + s.D;
    }
    if (W_last < -(s.D) || Wp_last < 0) { W_last = -(s.D); Wp_last = 0; }
    int i_increase = 0;
    while (vsqgtr1(W_last, s) && (i_increase < 10)) {
      Wp_last *= 10.;
      W_last = Wp_last + (s.D);
      i_increase++;
    }
    int nattempt = 0;
    while (true) {
      const double ut2 = utsq_calc(W_last, s);
      if (ut2 >= 0.0 && ut2 == ut2) break;
      double coldWp = 0.0;
      coldgrhd(s.Qtsq, s.D, coldWp);
      if (Wp_last < coldWp) Wp_last = coldWp;
      if (nattempt > 0) Wp_last = fmax(fmax(fabs(Wp_last) * 10.0, D_NUMEPSILON * 100.0 * fabs(s.D)), D_KINDASMALL);
      W_last = Wp_last + (s.D);
      if (nattempt > 1000) { pflag = 4; return 1; }
      nattempt++;
    }
  }

  double Wp = Wp_last;
  const int rv = newton<COLD>(pflag, Wp, s, nstroke);
  if (rv != 0) return rv;

  if (Wp == 1.e100) pflag = 51;
  else if (Wp <= -(s.D)) pflag = 51;
  else if (Wp / s.w1 > s.gsq_too_big) pflag = 2;

  // Wp2prim
  double pr0[HARM_NPR];
  prim[HARM_B1] = Ut[HARM_B1];
  prim[HARM_B2] = Ut[HARM_B2];
  prim[HARM_B3] = Ut[HARM_B3];
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) pr0[pl] = prim[pl];
  const double W = Wp + s.D;
  const double utsq = utsq_calc(W, s);
  const double gammasq = 1.0 + utsq;
  const double gamma = sqrt(gammasq);
  const double wmrho0 = (Wp - s.D * utsq / (1.0 + gamma)) / gammasq;
  const double rho0 = s.D / gamma;
  const double p = COLD ? 0.0 : ((gam - 1.0) / gam) * wmrho0;
  const double u = wmrho0 - p;
  if ((rho0 <= 0.) || ((p < 0.) && !COLD)) {
    if ((rho0 <= 0.) && (p >= 0.)) pflag = 9;
    if ((rho0 > 0.) && (p < 0.)) pflag = 10;
    if ((rho0 <= 0.) && (p < 0.)) pflag = 11;
  }
  prim[HARM_RHO] = rho0;
  prim[HARM_UU] = COLD ? 0.0 : u;
#pragma unroll
  for (int i = 1; i < 4; i++) prim[HARM_U1 + i - 1] = gamma / (W + s.Bsq) * (Qtcon[i] + s.QdotB * Bcon[i] / W);
  bool foundnan = false;
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++)
    if (!isfinite(prim[pl])) foundnan = true;
  if (foundnan) {
    pflag = 8;
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) prim[pl] = pr0[pl];
    return 1;
  }
  return 0;
}

// Utoprim_jon_nonrelcompat_inputnorestmass on geometry-free Ugf.
template <bool COLD = false>
__device__ __forceinline__ int jon(const double* Ugf, const Geom& g, double gam, double gammafail, double* pr,
                                   int& nstroke, double wp_seed)
{
  double Ut[HARM_NPR], pt[HARM_NPR];
  pr[HARM_B1] = Ugf[HARM_B1];
  pr[HARM_B2] = Ugf[HARM_B2];
  pr[HARM_B3] = Ugf[HARM_B3];
  const double alpha = g.alpha;
#pragma unroll
  for (int i = 0; i < HARM_NPR; i++) Ut[i] = alpha * Ugf[i];
#pragma unroll
  for (int i = 0; i < HARM_NPR; i++) pt[i] = pr[i];
  pt[HARM_B1] = alpha * pr[HARM_B1];
  pt[HARM_B2] = alpha * pr[HARM_B2];
  pt[HARM_B3] = alpha * pr[HARM_B3];
  int pflag = 0;
  const int ret = new_body<COLD>(pflag, Ut, g, gam, gammafail, pt, nstroke, wp_seed);
  if (ret == 0) {
#pragma unroll
    for (int i = 0; i <= HARM_U3; i++) pr[i] = pt[i];
  } else if (pflag <= 0) {
    pflag = ret + 51;
  }
  return pflag;
}

// UtoU(UEVOLVE -> UNOTHING): U without its gdet factor.
__device__ __forceinline__ void geometry_free(const double* U, const Geom& g, double* Ugf)
{
  const double ig = (g.gdet >= 0.0 ? 1.0 : -1.0) / (fabs(g.gdet) + D_SMALL);
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) Ugf[pl] = U[pl] * ig;
}

// The hot solve of Utoprimgen(UEVOLVE): strip gdet then invert.
__device__ __forceinline__ int utoprimgen(const double* U, const Geom& g, double gam, double gammafail, double* pr,
                                          int& nstroke, double wp_seed)
{
  double Ugf[HARM_NPR];
  geometry_free(U, g, Ugf);
  return jon(Ugf, g, gam, gammafail, pr, nstroke, wp_seed);
}

// check_on_inversion with FAILIFBADCHECK (utoprimgen.c:697-850, :18) after the
// hot solve: a converged zone (flag 0) whose U(p) differs from the inverted
// geometry-free Ugf by more than 10% in rho, u, a momentum or a field
// component, each against its own norm (the momenta made comparable through
// u^i/u^t, REMOVERESTMASSFROMUU 2), gets UTOPRIMFAILCONVBADINVERTCOMPARE.  The
// momentum norm sums the U2 terms twice and leaves U1 out, as the reference
// does.  The reference forms a third of those sums with a long double 1/3;
// the double quotient by 3 below is the same correctly rounded value (the
// quotient of a double by 3 is never within that rounding's error of a
// halfway point), and the field ratio's long double division can differ from
// this double one in the last bit only, which matters for a ratio equal to
// 0.1 to 16 digits.
__device__ __forceinline__ int check_inversion(int flag, const double* Ugf, const double* pr, const Geom& g, double gam)
{
  if (flag != 0) return flag;
  Geom G1 = g;
  G1.gdet = 1.0;  // primtoU(UNOTHING)
  State q;
  double Un[HARM_NPR];
  get_state(pr, g, gam, q);
  primtoflux(pr, q, 0, G1, Un);
  double on[HARM_NPR], nn[HARM_NPR];
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) { on[pl] = Ugf[pl]; nn[pl] = Un[pl]; }
#pragma unroll
  for (int jj = 1; jj <= 3; jj++) {
    on[HARM_UU + jj] = Ugf[HARM_UU + jj] * (q.ucon[jj] / q.ucon[0]);
    nn[HARM_UU + jj] = Un[HARM_UU + jj] * (q.ucon[jj] / q.ucon[0]);
  }
  int bad = 0;
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) {
    double fdiff;
    if (pl == HARM_RHO || pl == HARM_UU) {
      fdiff = fabs(nn[pl] - on[pl]) / (fabs(nn[pl]) + fabs(on[pl]) + D_SMALL);
    } else if (pl <= HARM_U3) {
      double en = (fabs(nn[HARM_U2]) + fabs(on[HARM_U2]) + fabs(nn[HARM_U2]) + fabs(on[HARM_U2]) + fabs(nn[HARM_U3]) +
                   fabs(on[HARM_U3])) / 3.0;
      const double eu = 0.5 * (fabs(on[HARM_UU]) + fabs(nn[HARM_UU]));
      en = en > eu ? en : eu;  // MAX, global.nondepnmemonics.h:761
      fdiff = fabs(nn[pl] - on[pl]) / (en + D_SMALL);
    } else {
      fdiff = fabs(nn[pl] - on[pl]) / ((fabs(nn[HARM_B1]) + fabs(on[HARM_B1]) + fabs(nn[HARM_B2]) + fabs(on[HARM_B2]) +
                                        fabs(nn[HARM_B3]) + fabs(on[HARM_B3])) / 3.0 + D_SMALL);
    }
    if (fdiff > 1E-1 && fabs(on[pl]) > D_SMALL && fabs(nn[pl]) > D_SMALL) bad++;
  }
  return bad ? HARM_UTOPRIMFAILCONVBADINVERTCOMPARE : flag;
}

// The rest of Utoprimgen's EOMGRMHD branch after the hot solve
// (utoprimgen.c:215-263): check_on_inversion when checkinv, then for a failed
// hot solve (any failure: USECOLDIFHOT* = 1, :587-591) the cold solve from
// the same guess pg when hot2cold (trycoldinversion, :597-690).  A converged
// cold solve keeps its density and velocities, takes u back from the guess
// and marks the zone UTOPRIMFAILU2AVG1, so that failure averaging recomputes
// u alone (:655-660); otherwise the hot result and flag stand.  Ugf: the
// geometry-free conserved state; the cold solve's strokes add to ns.
__device__ __forceinline__ int after_hot(int fl, const double* Ugf, const Geom& g, double gam, double gammafail,
                                         const double* pg, double* pr, int& ns, bool hot2cold, bool checkinv)
{
  if (checkinv) fl = check_inversion(fl, Ugf, pr, g, gam);
  if (hot2cold && fl > 0) {
    double pc[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) pc[pl] = pg[pl];
    const int cf = jon<true>(Ugf, g, gam, gammafail, pc, ns, __builtin_nan(""));
    if (cf >= HARM_UTOPRIMFAILFIXEDCOLD && cf <= 0) {  // IFUTOPRIMNOFAILORFIXED
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) pr[pl] = pc[pl];
      pr[HARM_UU] = pg[HARM_UU];
      fl = HARM_UTOPRIMFAILU2AVG1;
    }
  }
  return fl;
}

// whether after_hot has anything to do for a hot-solve flag
__device__ __forceinline__ bool needs_after_hot(int fl, bool hot2cold, bool checkinv)
{
  return (checkinv && fl == 0) || (hot2cold && fl > 0);
}

}  // namespace hd
