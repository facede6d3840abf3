// dgrid.h -- grid-level device helpers: storage indexing, reconstruction,
// face states and the small per-zone fixes (limit_gamma, inflow check).
//
// Reference (HARM): interppoint.para.c:568-1200 (para4), flux.c:1952-2110 (getplpr,
// check_plpr), fixup.c:2723 (limit_gamma), fixup.c:3255 (inflow_check_rel4vel).
#pragma once
#include "dphys.h"

namespace hd {

// Active sizes n, ghost widths ng, storage box ns = n + 2 ng, and the device
// row pitch ps1 >= ns1 (a multiple of 16 doubles, so every storage row
// starts on a 128-byte line).  ncell and nplane count pitched positions.
struct Dims {
  int n1, n2, n3, ng1, ng2, ng3, ns1, ns2, ns3, ps1;
  long ncell, nplane;
};

__device__ __forceinline__ long cidx(const Dims& d, int i, int j, int k)
{
  return ((long)(k + d.ng3) * d.ns2 + (j + d.ng2)) * d.ps1 + (i + d.ng1);
}
__device__ __forceinline__ long pidx(const Dims& d, int i, int j) { return (long)(j + d.ng2) * d.ps1 + (i + d.ng1); }

__device__ __forceinline__ double minmod(double a, double b)
{
  return ((a) * (b) <= 0) ? 0.0 : ((fabs(a) < fabs(b)) ? (a) : (b));
}
__device__ __forceinline__ double igdetnosing(double gdet) { return (gdet >= 0.0 ? 1.0 : -1.0) / (fabs(gdet) + D_SMALL); }

__device__ __forceinline__ double gamma_of(const double* pr, const Geom& G)
{
  const double u1 = pr[HARM_U1], u2 = pr[HARM_U2], u3 = pr[HARM_U3];
  double qsq = gco(G, 1, 1) * u1 * u1 + gco(G, 2, 2) * u2 * u2 + gco(G, 3, 3) * u3 * u3
               + 2. * (gco(G, 1, 2) * u1 * u2 + gco(G, 1, 3) * u1 * u3);
  qsq = qsq * (qsq >= 0.0);
  return sqrt(1. + qsq);
}

// returns whether the velocity was rescaled
__device__ __forceinline__ bool limit_gamma(double gammamax, double* pr, const Geom& G)
{
  if (gammamax <= 1.0) return false;
  const double gamma = gamma_of(pr, G);
  if ((gamma > gammamax && (gamma != 1.0))) {
    const double pref = (gammamax * gammamax - 1.) / (gamma * gamma - 1.);
    const double f = sqrt(pref);
    pr[HARM_U1] *= f;
    pr[HARM_U2] *= f;
    pr[HARM_U3] *= f;
    return true;
  }
  return false;
}

// para4 with MC-limited slopes and the CW84 monotonicity conditions.
__device__ __forceinline__ void para4(const double* y, double& lout, double& rout)
{
  double dq1[4], dq2[3], dq[3];
#pragma unroll
  for (int m = -1; m <= 1; m++) {
    dq1[m + 1] = (y[m] - y[m - 1]);
    dq2[m + 1] = 0.5 * (y[m + 1] - y[m - 1]);
  }
  dq1[3] = (y[2] - y[1]);
#pragma unroll
  for (int m = -1; m <= 1; m++) {
    const double Dqm = 2.0 * dq1[m + 1];
    const double Dqp = 2.0 * dq1[m + 2];
    const double Dqc = dq2[m + 1];
    dq[m + 1] = minmod(Dqc, minmod(Dqm, Dqp));
  }
  const double ddql = (dq[1] - dq[0]);
  const double ddqr = (dq[2] - dq[1]);
  double l = 0.5 * (y[0] + y[-1]) - ddql * (1.0 / 8.0);
  double r = 0.5 * (y[1] + y[0]) - ddqr * (1.0 / 8.0);
  const double qa = (r - y[0]) * (y[0] - l);
  const double qd = (r - l);
  const double qe = 4.0 * (y[0] - 0.5 * (l + r));
  if (qa <= 0.0) {
    l = y[0];
    r = y[0];
  } else {
    if (qd * (qd - qe) < 0.0) l = (-(2.0 + 4.0) * r + 2.0 * 4.0 * y[0]) / (4.0 - 2.0);
    else if (qd * (qd + qe) < 0.0) r = (-(2.0 + 4.0) * l + 2.0 * 4.0 * y[0]) / (4.0 - 2.0);
  }
  lout = l;
  rout = r;
}

// Single-zone reconstruction for the schemes that need only the zone's
// five-point stencil (PARA, donor).  WENO5BND reconstructs whole lines
// (weno.hip).
template <int RECON>
__device__ __forceinline__ void recon_cell(const double* y, double& l, double& r)
{
  if (RECON == HARM_RECON_PARA) para4(y, l, r);
  else { l = y[0]; r = y[0]; }
}

// inflow_check_rel4vel for dir 1 on the outer radial boundary.
__device__ __forceinline__ void inflow_check_x1up(double* pr, const Geom& G, double gam, double gammamax)
{
  State q;
  get_state(pr, G, gam, q);
  if (!(q.ucon[1] < 0.)) return;
  const double gamma = gamma_of(pr, G);
  pr[HARM_U1] /= gamma;
  pr[HARM_U2] /= gamma;
  pr[HARM_U3] /= gamma;
  const double alpha = 1. / sqrt(-gcn(G, 0, 0));
  const double betacon = gcn(G, 0, 1) * alpha * alpha;
  pr[HARM_U1] = betacon / alpha;
  double vsq = 0.;
#pragma unroll
  for (int j = 1; j < 4; j++)
#pragma unroll
    for (int k = 1; k < 4; k++) vsq += gco(G, j, k) * pr[HARM_U1 + j - 1] * pr[HARM_U1 + k - 1];
  if (vsq < 0.0) vsq = 0.0;
  if (vsq >= 1.0) vsq = 1.0 - 1.0 / (gammamax * gammamax);
  const double gnew = 1. / sqrt(1. - vsq);
  pr[HARM_U1] *= gnew;
  pr[HARM_U2] *= gnew;
  pr[HARM_U3] *= gnew;
}

// ---------------------------------------------------------------------------
// Bondi accretion boundaries (initboundcode/bounds.sasha.c:896-1160)
// ---------------------------------------------------------------------------
// The exact integrals of the Bondi problem's solution, used by the outer
// boundary (bounds.sasha.c:61-62: C2 and K of the Gamma = 4/3, r_c = 8 flow).
constexpr double kBondiC2x = 1.373125, kBondiKx = 1.0;

// bondi_compute_integrals (bounds.sasha.c:896): C1 = T^n u^r r^2,
// C2 = (1 + (1+n) T)^2 (1 - 2/r + (u^r)^2), K = p / rho^Gamma of a zone at
// radius r, with u^r the Kerr-Schild radial 4-velocity (spin 0) and
// T = p/rho, n = 1/(Gamma-1).
__device__ __forceinline__ void bondi_integrals(const double* pr, const Geom& G, double r, double R0, double gam,
                                                double& C1, double& C2, double& K)
{
  State q;
  get_state(pr, G, gam, q);
  const double ur = (r - R0) * q.ucon[1];  // d r / d X1 = r - R0
  const double n = 1 / (gam - 1);
  const double p = (gam - 1) * pr[HARM_UU];
  const double rho = pr[HARM_RHO];
  const double T = p / rho;
  C1 = pow(T, n) * ur * r * r;
  C2 = pow(1 + (1 + n) * T, 2.) * (1 - 2. / r + ur * ur);
  K = p / pow(rho, gam);
}

// bondi_invert_integrals (bounds.sasha.c:1101): the temperature at radius r
// from the integrals, then rho, u and u^r.  The temperature is the root of
// h(T) = T - f(T), f(T) = (sqrt(C2/(1 - 2/r + u^2)) - 1)/(1 + n),
// u = C1/(r^2 T^n).  For r > 9 the reference's plain iteration T <- f(T) is
// used as written.  For 0.35 < r < 2.5 the root is a repelling fixed point of
// f, which the reference reaches by the damped iteration
// T <- T + 0.02 (T - f(T)), converging linearly (100-170 steps at the inner
// boundary); here that iteration runs until its steps fall below 1e-4 of T
// and Newton steps on h finish it (the same root, to 1.3e-14 of the
// reference's over randomized integrals; about 30 steps).  A Newton
// step moving T by more than 1 % hands back to the damped iteration to the
// reference's own tolerance.  Returns false outside those radii or when the
// outer iteration does not decrease monotonically (the reference stops the
// run there).
__device__ __forceinline__ bool bondi_invert(double n, double C1, double C2, double K, double r, double& rho,
                                             double& uu, double& ur)
{
  const double relative_accuracy = 1.0e-15;
  const int max_iter = 250;
  double T0 = 0.2, T1 = 0.1, u1;
  int iter;
  if (r > 0.35 && r < 2.5) {
    for (iter = 0; iter < max_iter && fabs((T0 - T1) / T1) > 1.0e-4; iter++) {
      T0 = T1;
      u1 = C1 / (r * r * pow(T0, n));
      T1 = T0 + 0.02 * (T0 - (sqrt(C2 / (1 - 2 / r + u1 * u1)) - 1) / (1 + n));
    }
    bool polished = false;
    double T = T1;
    for (int k = 0; k < 8; k++) {
      const double u = C1 / (r * r * pow(T, n));
      const double w = 1 - 2 / r + u * u;
      const double f = (sqrt(C2 / w) - 1) / (1 + n);
      const double fp = n * sqrt(C2) * u * u / ((1 + n) * T * w * sqrt(w));
      const double Tn = T - (T - f) / (1 - fp);
      if (!(fabs(Tn - T) <= 1e-2 * fabs(T))) break;  // leaving the basin (or NaN)
      // a step below 1e-12 of T leaves an error of order its square: T is
      // converged to the last bits (a tighter test could chase the rounding
      // noise of pow and sqrt)
      if (fabs(Tn - T) <= 1e-12 * fabs(Tn)) {
        T = Tn;
        polished = true;
        break;
      }
      T = Tn;
    }
    if (polished) T1 = T;
    else
      for (; iter < max_iter && fabs((T0 - T1) / T1) > relative_accuracy; iter++) {
        T0 = T1;
        u1 = C1 / (r * r * pow(T0, n));
        T1 = T0 + 0.02 * (T0 - (sqrt(C2 / (1 - 2 / r + u1 * u1)) - 1) / (1 + n));
      }
  } else if (r > 9) {
    for (iter = 0; iter < max_iter && (T0 - T1) / T1 > relative_accuracy; iter++) {
      T0 = T1;
      u1 = C1 / (r * r * pow(T0, n));
      T1 = (sqrt(C2 / (1 - 2 / r + u1 * u1)) - 1) / (1 + n);
    }
    if (T1 > T0) return false;
  } else {
    return false;
  }
  rho = pow(T1 / K, n);
  uu = rho * T1 * n;
  ur = C1 / (pow(T1, n) * r * r);
  return true;
}

// A radial Kerr-Schild 4-velocity u^r (u^theta = u^phi = 0) at a zone of the
// (spin 0) MKS grid as primitive relative 4-velocities: u^X1 = u^r/(r - R0),
// u^t from the normalisation with the zone's metric (the root with u^t > 0,
// inside the horizon too), then u~^i = u^i + u^t beta^i (pr2ucon VEL4 and
// ucon2pr, phys.tools.c).
__device__ __forceinline__ void radial_ks_to_rel4(double ur, double r, double R0, const Geom& G, double* pr)
{
  const double u1 = ur / (r - R0);
  const double AA = gco(G, 0, 0), BB = 2. * gco(G, 0, 1) * u1, CC = 1. + gco(G, 1, 1) * u1 * u1;
  const double discr = BB * BB - 4. * AA * CC;
  const double ut = (-BB - sqrt(discr)) / (2. * AA);
  pr[HARM_U1] = u1 + ut * G.beta1;
  pr[HARM_U2] = ut * G.beta2;
  pr[HARM_U3] = ut * G.beta3;
}

}  // namespace hd
