// dgrid.h -- grid-level device helpers: storage indexing, reconstruction,
// face states and the small per-zone fixes (limit_gamma, inflow check).
//
// Reference (HARM): interppoint.para.c:568-1200 (para4), reconstructeno.c
// :1693-2030, 2211-2330, 3015 (WENO weights), flux.c:1952-2110 (getplpr,
// check_plpr), fixup.c:2723 (limit_gamma), fixup.c:3255 (inflow_check_rel4vel).
#pragma once
#include "dphys.h"

namespace hd {

struct Dims {
  int n1, n2, n3, ng1, ng2, ng3, ns1, ns2, ns3;
  long ncell, nplane;
};

__device__ __forceinline__ long cidx(const Dims& d, int i, int j, int k)
{
  return ((long)(k + d.ng3) * d.ns2 + (j + d.ng2)) * d.ns1 + (i + d.ng1);
}
__device__ __forceinline__ long pidx(const Dims& d, int i, int j) { return (long)(j + d.ng2) * d.ns1 + (i + d.ng1); }

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

__device__ __forceinline__ void limit_gamma(double gammamax, double* pr, const Geom& G)
{
  if (gammamax <= 1.0) return;
  const double gamma = gamma_of(pr, G);
  if ((gamma > gammamax && (gamma != 1.0))) {
    const double pref = (gammamax * gammamax - 1.) / (gamma * gamma - 1.);
    const double f = sqrt(pref);
    pr[HARM_U1] *= f;
    pr[HARM_U2] *= f;
    pr[HARM_U3] *= f;
  }
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

__device__ __forceinline__ double symsum3(const double* a) { return (0.0 + (a[0] + a[2])) + a[1] * 1; }
__device__ __forceinline__ double symsum5(const double* a) { return ((0.0 + (a[0] + a[4])) + (a[1] + a[3])) + a[2] * 1; }

__device__ __forceinline__ double limit_weight3(double w)
{
  const double avg = 1. / 3.;
  const double w1 = 0.2 * avg, w2 = (1.0 - 0.2) * avg, w3 = (1.0 + 0.2) * avg, w4 = 1.0;
  if (w < w1) return 0.0;
  if (w < w2) return avg * (w - w1) / (w2 - w1);
  if (w < w3) return avg;
  return (w - w3) * (w4 - avg) / (w4 - w3) + avg;
}

// WENO5 centre-to-edge core (point values).
__device__ __forceinline__ void weno5(const double* u, double& lout, double& rout)
{
  const double EPS = (500.0 * DBL_EPSILON) * (500.0 * DBL_EPSILON);
  double b[3], d1[3], d2[3];
  d1[0] = -(3 * u[0] + u[2]) + 4 * u[1];
  d1[1] = u[1] - u[-1];
  d1[2] = (3 * u[0] + u[-2]) - 4 * u[-1];
  d2[0] = (u[0] + u[2]) - 2 * u[1];
  d2[1] = (u[-1] + u[1]) - 2 * u[0];
  d2[2] = (u[0] + u[-2]) - 2 * u[-1];
#pragma unroll
  for (int s = 0; s < 3; s++) b[s] = 13. / 12. * d2[s] * d2[s] + 1. / 4. * d1[s] * d1[s];
  double u2[5];
#pragma unroll
  for (int i = -2; i <= 2; i++) u2[i + 2] = u[i] * u[i];
  const double err = symsum5(u2) * EPS;
#pragma unroll
  for (int s = 0; s < 3; s++) b[s] += err;
  const double sb = symsum3(b);
#pragma unroll
  for (int s = 0; s < 3; s++) b[s] /= (sb + DBL_MIN);
  double w[3];
#pragma unroll
  for (int s = 0; s < 3; s++) {
    double t = b[s] + EPS / 1.;
    t *= t;
    w[s] = 1. / t;
  }
  double sw = symsum3(w);
#pragma unroll
  for (int s = 0; s < 3; s++) w[s] /= (sw + 0);
#pragma unroll
  for (int s = 0; s < 3; s++) w[s] = limit_weight3(w[s]);
  sw = symsum3(w);
#pragma unroll
  for (int s = 0; s < 3; s++) w[s] /= (sw + 0);
  double wr[3] = {w[0] * (5.0 / 16.0), w[1] * (5.0 / 8.0), w[2] * (1.0 / 16.0)};
  double wl[3] = {w[0] * (1.0 / 16.0), w[1] * (5.0 / 8.0), w[2] * (5.0 / 16.0)};
  sw = symsum3(wr);
#pragma unroll
  for (int s = 0; s < 3; s++) wr[s] /= (sw + DBL_MIN);
  sw = symsum3(wl);
#pragma unroll
  for (int s = 0; s < 3; s++) wl[s] /= (sw + DBL_MIN);
  const double qr0 = (3 * u[0] + 6 * u[1] - u[2]) / 8.0;
  const double qr1 = (-u[-1] + 6 * u[0] + 3 * u[1]) / 8.0;
  const double qr2 = (3 * u[-2] - 10 * u[-1] + 15 * u[0]) / 8.0;
  const double ql0 = (15 * u[0] - 10 * u[1] + 3 * u[2]) / 8.0;
  const double ql1 = (3 * u[-1] + 6 * u[0] - u[1]) / 8.0;
  const double ql2 = (-u[-2] + 6 * u[-1] + 3 * u[0]) / 8.0;
  rout = (wr[0] * qr0 + wr[2] * qr2) + wr[1] * qr1;
  lout = (wl[0] * ql0 + wl[2] * ql2) + wl[1] * ql1;
}

template <int RECON>
__device__ __forceinline__ void recon_cell(const double* y, double& l, double& r)
{
  if (RECON == HARM_RECON_PARA) para4(y, l, r);
  else if (RECON == HARM_RECON_WENO5) weno5(y, l, r);
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

}  // namespace hd
