// weno.hip -- WENO5BND centre-to-edge reconstruction of the primitives as the
// reference compiles it for the torus (lim = WENO5BND): the whole storage line
// along a direction is reconstructed (slope_lim_linetype_c2e interpline.c:306)
// with
//   * the SMONO indicator (interpline.smono.c:198, :517): where the five-point
//     interpolant's first three derivatives at the neighbouring zone centres
//     agree in sign and are well above round-off, its face values replace
//     WENO's, with a linear ramp in between;
//   * WENO5 weights (compute_cf_stencil_weights reconstructeno.c:1693):
//     Jiang-Shu indicators of the point values, desensitised by eps |u|^2,
//     normalised, 1/(beta+eps)^2, normalised, limited (limit_weight :3015)
//     and normalised, with the SMONO ramp pulling them to equal (:852);
//   * the AND stencil reduction (compute_weights_ratios :2311): a zone whose
//     outer sub-stencils weigh 1.3-1.6 times more than the same three zones
//     weigh for the neighbours two zones away blends into WENO3;
//   * the pre/post-shock reduction (apply_additional_reduction_to_weights
//     :795): a zone whose pressure-jump measure dP exceeds 1/4 takes its
//     neighbours' reduction, fully at dP >= 1/2;
//   * the optimal face weights and the blend of WENO5, WENO3 and SMONO values
//     (choose_weno_order :2442, eno_cvt :884, eno_line_reconstruct :538).
//
// Device layout.  k_weno_terms and k_weno_dp form dP for every direction at
// every zone that a face reconstruction needs (compute_df_line_new
// interpline.c:3557, from the zone's state and its two neighbours' along the
// direction).  Along x1, where a line is contiguous in memory,
// k_weno_recon_x1 gives each lane of a wave one zone of the line: the lane
// forms its zone's weights, takes its neighbours' from the lanes two away for
// the ratio test and the neighbours' ratios from the lanes one away, and
// writes the zone's faces (coalesced loads and stores, no carried state).
// Along x2 and x3, k_weno_recon sweeps segments of each storage line of one
// primitive in one thread (lanes are consecutive i: coalesced), carrying the
// last eight values and a four-zone pipeline of weights in registers.  The
// face-flux kernels read the left and right face values of zones -1 .. N
// back.  The arithmetic repeats the oracle's operation order (no contraction
// in this translation unit, IEEE divisions), so the face values agree bit for
// bit.
#include <hip/hip_runtime.h>
#include <math.h>
#include <float.h>
#include <algorithm>
#include "dgrid.h"
#include "gkern.h"
#include "hhost.h"

namespace hk {
namespace {

using hd::Geom;
using hd::State;

constexpr double kSqrtWeps = 500.0 * DBL_EPSILON;  // SQRT_WENO_EPSILON
constexpr double kWeps = kSqrtWeps * kSqrtWeps;     // WENO_EPSILON = REDUCEEPSILON
constexpr double kMinNum = DBL_MIN;                 // MINNUMREPRESENT

constexpr double wmin(double a, double b) { return a < b ? a : b; }
constexpr double wmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double wminmod(double a, double b)
{
  return (a * b <= 0) ? 0.0 : ((fabs(a) < fabs(b)) ? a : b);
}

// get_sum_of_elements with SYMMETRIZE_WENO: mirror pairs first, the middle last
constexpr double sym3(double a0, double a1, double a2)
{
  double s = 0.0;
  s += a0 + a2;
  s += a1 * 1;
  return s;
}
constexpr double sym2(double a0, double a1)
{
  double s = 0.0;
  s += a0 + a1;
  s += a1 * 0;
  return s;
}
__device__ __forceinline__ double sym5(const double* a)
{
  double s = 0.0;
  s += a[0] + a[4];
  s += a[1] + a[3];
  s += a[2] * 1;
  return s;
}

template <int ORDER>
__device__ __forceinline__ double limit_weight(double w)
{
  const double avg = 1. / ORDER, w1 = 0.2 * avg, w2 = (1.0 - 0.2) * avg, w3 = (1.0 + 0.2) * avg, w4 = 1.0;
  if (w < w1) return 0.0;
  if (w < w2) return avg * (w - w1) / (w2 - w1);
  if (w < w3) return avg;
  return (w - w3) * (w4 - avg) / (w4 - w3) + avg;
}

// The weights of the two uniform cases, formed by the same operations as the
// general code below (constant-folded, round to nearest):
//   * all five values zero: the Jiang-Shu indicators are 0, every weight is
//     1/(WENO_EPSILON^2) before normalisation, the normalised ones lie within
//     an ulp of 1/3, limit_weight maps them to its flat band's value 1./3,
//     and (1/3 + 1/3) + 1/3 rounds to 1, so the zone's weights are 1./3 each
//     and its SMONO indicator is 0 (every derivative estimate is 0);
//   * three equal values u[-1] = u[0] = u[1]: both WENO3 indicators are the
//     same, so both weights normalise to exactly 1/2, which limit_weight
//     keeps.
// The face weights of a zone whose unoptimized weights are 1./3 each, and
// the WENO3 face weights of 1/2 each, are these constants.
constexpr double kThird = 1. / 3;
constexpr double kL3n = sym3(kThird * (1. / 16.), kThird * (5. / 8.), kThird * (5. / 16.)) + kMinNum;
constexpr double kR3n = sym3(kThird * (5. / 16.), kThird * (5. / 8.), kThird * (1. / 16.)) + kMinNum;
constexpr double kL3[3] = {kThird * (1. / 16.) / kL3n, kThird * (5. / 8.) / kL3n, kThird * (5. / 16.) / kL3n};
constexpr double kR3[3] = {kThird * (5. / 16.) / kR3n, kThird * (5. / 8.) / kR3n, kThird * (1. / 16.) / kR3n};
constexpr double kL2n = sym2(0.5 * (1. / 4.), 0.5 * (3. / 4.)) + kMinNum;
constexpr double kR2n = sym2(0.5 * (3. / 4.), 0.5 * (1. / 4.)) + kMinNum;
constexpr double kL2[2] = {0.5 * (1. / 4.) / kL2n, 0.5 * (3. / 4.) / kL2n};
constexpr double kR2[2] = {0.5 * (3. / 4.) / kR2n, 0.5 * (1. / 4.) / kR2n};

// Unoptimized WENO5 weights of the zone u[0] (all three sub-stencils lie in
// the line for every zone the sweep computes).
__device__ __forceinline__ void weights3(const double* u, double* w)
{
  const double d10 = -(3 * u[0] + u[2]) + 4 * u[1], d11 = u[1] - u[-1], d12 = (3 * u[0] + u[-2]) - 4 * u[-1];
  const double d20 = (u[0] + u[2]) - 2 * u[1], d21 = (u[-1] + u[1]) - 2 * u[0], d22 = (u[0] + u[-2]) - 2 * u[-1];
  double s0 = 13. / 12. * d20 * d20 + 1. / 4. * d10 * d10;
  double s1 = 13. / 12. * d21 * d21 + 1. / 4. * d11 * d11;
  double s2 = 13. / 12. * d22 * d22 + 1. / 4. * d12 * d12;
  double q[5] = {u[-2] * u[-2], u[-1] * u[-1], u[0] * u[0], u[1] * u[1], u[2] * u[2]};
  const double err = sym5(q) * kWeps;
  s0 += err;
  s1 += err;
  s2 += err;
  double n = sym3(s0, s1, s2) + kMinNum;
  s0 /= n;
  s1 /= n;
  s2 /= n;
  double t0 = s0 + kWeps, t1 = s1 + kWeps, t2 = s2 + kWeps;
  t0 *= t0;
  t1 *= t1;
  t2 *= t2;
  double w0 = 1. / t0, w1 = 1. / t1, w2 = 1. / t2;
  n = sym3(w0, w1, w2) + 0;
  w0 /= n;
  w1 /= n;
  w2 /= n;
  w0 = limit_weight<3>(w0);
  w1 = limit_weight<3>(w1);
  w2 = limit_weight<3>(w2);
  n = sym3(w0, w1, w2) + 0;
  w[0] = w0 / n;
  w[1] = w1 / n;
  w[2] = w2 / n;
}

// WENO3 weights, unoptimized then face-optimized ({1/4, 3/4} left, mirrored right)
__device__ __forceinline__ void weights2(const double* u, double* L2, double* R2)
{
  const double d0 = u[1] - u[0], d1 = u[0] - u[-1];
  double s0 = d0 * d0, s1 = d1 * d1;
  const double err = sym3(u[-1] * u[-1], u[0] * u[0], u[1] * u[1]) * kWeps;
  s0 += err;
  s1 += err;
  double n = sym2(s0, s1) + kMinNum;
  s0 /= n;
  s1 /= n;
  double t0 = s0 + kWeps, t1 = s1 + kWeps;
  t0 *= t0;
  t1 *= t1;
  double w0 = 1. / t0, w1 = 1. / t1;
  n = sym2(w0, w1) + 0;
  w0 /= n;
  w1 /= n;
  w0 = limit_weight<2>(w0);
  w1 = limit_weight<2>(w1);
  n = sym2(w0, w1) + 0;
  w0 /= n;
  w1 /= n;
  double l0 = w0 * (1. / 4.), l1 = w1 * (3. / 4.), r0 = w0 * (3. / 4.), r1 = w1 * (1. / 4.);
  n = sym2(l0, l1) + kMinNum;
  L2[0] = l0 / n;
  L2[1] = l1 / n;
  n = sym2(r0, r1) + kMinNum;
  R2[0] = r0 / n;
  R2[1] = r1 / n;
}

// SMONO indicator (compute_mono_indicator_point_eno5).  The indicator is
// min over the three orders of |minmod(left, right)| against a round-off
// cut, clamped to [0, 1].  It is evaluated lazily, with the same values:
// minmod is 0 whenever its arguments are not strictly of one sign, and a
// quotient by 12 keeps its dividend's sign (or is 0), so an order whose
// numerators already disagree in sign gives 0 without its divisions; m = 0
// gives the indicator 0 and m >= cut gives 1 (m / cut >= 1 exactly).
__device__ __forceinline__ bool same_sign(double a, double b) { return (a > 0 && b > 0) || (a < 0 && b < 0); }
__device__ __forceinline__ double smono(const double* y)
{
  const double L3 = (((-3 * y[-2] - y[2]) + (10 * y[-1] + 6 * y[1])) - 12 * y[0]) / 2.;
  const double R3 = (((y[-2] + 3 * y[2]) - (6 * y[-1] + 10 * y[1])) + 12 * y[0]) / 2.;
  const double m3 = fabs(wminmod(L3, R3));
  if (m3 == 0.0) return 0.0;
  const double L2n = (((11 * y[-2] - y[2]) + (-20 * y[-1] + 4 * y[1])) + 6 * y[0]);
  const double R2n = (((-y[-2] + 11 * y[2]) + (4 * y[-1] - 20 * y[1])) + 6 * y[0]);
  if (!same_sign(L2n, R2n)) return 0.0;
  const double L1n = (((-3 * y[-2] + y[2]) + (-10 * y[-1] - 6 * y[1])) + 18 * y[0]);
  const double R1n = (((-y[-2] + 3 * y[2]) + (6 * y[-1] + 10 * y[1])) - 18 * y[0]);
  const double dl = y[-1] - y[-2], dr = y[2] - y[1];
  if (!same_sign(L1n, dl) || !same_sign(R1n, dr)) return 0.0;
  const double L1 = wminmod(L1n / 12., dl);
  const double R1 = wminmod(R1n / 12., dr);
  const double m2 = fabs(wminmod(L2n / 12., R2n / 12.));
  double m = kBIG;
  m = wmin(m, fabs(wminmod(L1, R1)));
  m = wmin(m, m2);
  m = wmin(m, m3);
  if (m == 0.0) return 0.0;
  const double norm = fabs(y[0]) + ((fabs(y[-1]) + fabs(y[1])) + (fabs(y[-2]) + fabs(y[2])));
  const double cut = norm * kSqrtWeps + kMinNum;
  if (m >= cut) return 1.0;
  return (m - 0.0) / (cut - 0.0);
}

// five-point interpolant at the left (RIGHT = false) or right face
template <bool RIGHT>
__device__ __forceinline__ double eno5(const double* y)
{
  double d[5];
  if (!RIGHT) {
    d[0] = -5. / 128. * y[-2];
    d[1] = 15. / 32. * y[-1];
    d[2] = 45. / 64. * y[0];
    d[3] = -5. / 32. * y[1];
    d[4] = 3. / 128. * y[2];
  } else {
    d[0] = 3. / 128. * y[-2];
    d[1] = -5. / 32. * y[-1];
    d[2] = 45. / 64. * y[0];
    d[3] = 15. / 32. * y[1];
    d[4] = -5. / 128. * y[2];
  }
  return sym5(d);
}

// eno_cvt of WENO5 (sub-stencil s covers u[-s .. 2-s]), zero-weight stencils skipped
template <bool RIGHT>
__device__ __forceinline__ double cvt3(const double* w, const double* u)
{
  double up0 = 0.0, up1 = 0.0, up2 = 0.0;
  if (!RIGHT) {
    if (w[0] != 0.0) up0 = w[0] * sym3(15. / 8. * u[0], -5. / 4. * u[1], 3. / 8. * u[2]);
    if (w[1] != 0.0) up1 = w[1] * sym3(3. / 8. * u[-1], 3. / 4. * u[0], -1. / 8. * u[1]);
    if (w[2] != 0.0) up2 = w[2] * sym3(-1. / 8. * u[-2], 3. / 4. * u[-1], 3. / 8. * u[0]);
  } else {
    if (w[0] != 0.0) up0 = w[0] * sym3(3. / 8. * u[0], 3. / 4. * u[1], -1. / 8. * u[2]);
    if (w[1] != 0.0) up1 = w[1] * sym3(-1. / 8. * u[-1], 3. / 4. * u[0], 3. / 8. * u[1]);
    if (w[2] != 0.0) up2 = w[2] * sym3(3. / 8. * u[-2], -5. / 4. * u[-1], 15. / 8. * u[0]);
  }
  return sym3(up0, up1, up2);
}
template <bool RIGHT>
__device__ __forceinline__ double cvt2(const double* w, const double* u)
{
  double up0 = 0.0, up1 = 0.0;
  if (!RIGHT) {
    if (w[0] != 0.0) up0 = w[0] * sym2(3. / 2. * u[0], -1. / 2. * u[1]);
    if (w[1] != 0.0) up1 = w[1] * sym2(1. / 2. * u[-1], 1. / 2. * u[0]);
  } else {
    if (w[0] != 0.0) up0 = w[0] * sym2(1. / 2. * u[0], 1. / 2. * u[1]);
    if (w[1] != 0.0) up1 = w[1] * sym2(-1. / 2. * u[-1], 3. / 2. * u[0]);
  }
  return sym2(up0, up1);
}

// What the sweep keeps of a zone between its weights and its face values:
// the unoptimized WENO5 weights after the SMONO ramp and the indicator.
// Everything else a zone's faces need (the SMONO face values, the WENO3
// weights, the optimized weights) is formed from the value window when its
// faces are, and the WENO3 weights only where the zone is reduced.
struct ZoneW {
  double u[3];
  double mono;
};

__device__ __forceinline__ void zone_weights(const double* y, ZoneW& z)
{
  if (y[-2] == 0.0 && y[-1] == 0.0 && y[0] == 0.0 && y[1] == 0.0 && y[2] == 0.0) {
    z.mono = 0.0;
    z.u[0] = z.u[1] = z.u[2] = kThird;
    return;
  }
  z.mono = smono(y);
  if (z.mono == 1) {
    z.u[0] = z.u[1] = z.u[2] = 1.0 / 3;
  } else {
    weights3(y, z.u);
    if (z.mono > 0.0 && z.mono < 1.0)
#pragma unroll
      for (int s = 0; s < 3; s++) z.u[s] = (1.0 - z.mono) * z.u[s] + z.mono * (1.0 / 3);
  }
}

// one side of the AND ratio test: the zone's outer sub-stencil weight against
// the neighbour's weight of the same three zones (equal weights: exactly 1)
__device__ __forceinline__ double side_ratio(double uc, double un)
{
  if (uc == un) return 1.0;
  return (uc + kWeps) / (un + kWeps);
}

// lower-order fraction from the ratio test (BIG for a side that does not exist)
__device__ __forceinline__ double ratio_lof(double wl, double wr)
{
  const double wrat = wmin(wl, wr);
  const double min_rat = 1.3, max_rat = 1.6;
  // the clamps below give exactly these at and beyond the band's ends
  if (wrat <= min_rat) return 0.0;
  if (wrat >= max_rat) return 1.0;
  double lof = (wrat - min_rat) / (max_rat - min_rat);
  lof = wmax(0., wmin(lof, 1.0));
  return wmax(lof, 0.0);  // the weights' own fraction starts at 0
}

// face values of zone y[0] from its kept weights, its final lower-order
// fraction and the SMONO ramp
__device__ __forceinline__ void zone_faces(const double* y, const ZoneW& z, double lof, double& lo, double& ro)
{
  const double m = z.mono;
  double ol = 0.0, orr = 0.0;
  if (m != 0) {
    ol = eno5<false>(y);
    orr = eno5<true>(y);
  }
  if (m >= 0.0 && m < 1.0) {
    // optimized WENO5 weights from the (ramped) unoptimized ones
    double L[3], R[3];
    if (z.u[0] == kThird && z.u[1] == kThird && z.u[2] == kThird) {
#pragma unroll
      for (int s = 0; s < 3; s++) {
        L[s] = kL3[s];
        R[s] = kR3[s];
      }
    } else {
      L[0] = z.u[0] * (1. / 16.);
      L[1] = z.u[1] * (5. / 8.);
      L[2] = z.u[2] * (5. / 16.);
      R[0] = z.u[0] * (5. / 16.);
      R[1] = z.u[1] * (5. / 8.);
      R[2] = z.u[2] * (1. / 16.);
      double n = sym3(L[0], L[1], L[2]) + kMinNum;
      L[0] /= n;
      L[1] /= n;
      L[2] /= n;
      n = sym3(R[0], R[1], R[2]) + kMinNum;
      R[0] /= n;
      R[1] /= n;
      R[2] /= n;
    }
    double vl = 0.0, vr = 0.0;
    if (0.0 < lof) {
      double L2[2], R2[2];
      if (y[-1] == y[0] && y[0] == y[1]) {
        L2[0] = kL2[0];
        L2[1] = kL2[1];
        R2[0] = kR2[0];
        R2[1] = kR2[1];
      } else {
        weights2(y, L2, R2);
      }
#pragma unroll
      for (int s = 0; s < 3; s++) {
        L[s] *= (1. - lof);
        R[s] *= (1. - lof);
      }
#pragma unroll
      for (int s = 0; s < 2; s++) {
        L2[s] *= lof;
        R2[s] *= lof;
      }
      vl += cvt3<false>(L, y);
      vr += cvt3<true>(R, y);
      vl += cvt2<false>(L2, y);
      vr += cvt2<true>(R2, y);
    } else {
      vl += cvt3<false>(L, y);
      vr += cvt3<true>(R, y);
    }
    ol = (1.0 - m) * vl + m * ol;
    orr = (1.0 - m) * vr + m * orr;
  }
  lo = ol;
  ro = orr;
}

// dP terms of one zone along DIR (compute_df_line_new, b^dir b_t = 0)
struct DpTerms {
  double rho, par, veff, upar;
};
__device__ __forceinline__ DpTerms dp_terms(const State& q, const double* pr, const Geom& G, double gam, int dir)
{
  const double rho = pr[HARM_RHO], u = pr[HARM_UU];
  const double pgas = (gam - 1.0) * u;
  const double V = G.gdet * rho * q.ucon[dir];
  const double bsq = q.bsq, btt = 0.0;
  const double ptot = pgas + 0.5 * bsq;
  DpTerms t;
  t.rho = rho;
  t.par = fabs(q.p1ud0 * q.ucon[dir]);
  t.veff = copysign(1.0, V) * t.par / (sqrt(t.par) + D_SMALL);
  t.upar = fabs((u + pgas + bsq) * (fabs(q.ucon[dir]) * q.ucov[0]) + ptot + btt);
  return t;
}

__device__ __forceinline__ void load_prim(const double* __restrict__ p, long ncell, long c, double* pr)
{
#pragma unroll
  for (int v = 0; v < HARM_NPR; v++) pr[v] = p[v * ncell + c];
}

// The zone terms of the dP measure for every active direction: tm[0] = rho,
// then (par, veff, upar) of direction d at tm[1 + 3 (d-1) ..].  One state
// per zone, over the zones -2 .. N+1 of every active direction.
__global__ __launch_bounds__(kBlock) void k_weno_terms(Dims d, Box z, const double* __restrict__ pb,
                                                         const double* __restrict__ gcent, double gam,
                                                         double* __restrict__ tm)
{
  int i, j, k;
  if (!z.at((long)blockIdx.x * kBlock + threadIdx.x, i, j, k)) return;
  Geom G;
  hd::load_geom(gcent, hd::pidx(d, i, j), d.nplane, G);
  double pr[HARM_NPR];
  const long c = hd::cidx(d, i, j, k);
  load_prim(pb, d.ncell, c, pr);
  State q;
  hd::get_state<false>(pr, G, gam, q);
  tm[c] = pr[HARM_RHO];
#pragma unroll
  for (int dir = 1; dir <= 3; dir++) {
    if ((dir == 1 ? d.n1 : dir == 2 ? d.n2 : d.n3) <= 1) continue;
    const DpTerms t = dp_terms(q, pr, G, gam, dir);
    tm[(1 + 3 * (dir - 1)) * d.ncell + c] = t.par;
    tm[(2 + 3 * (dir - 1)) * d.ncell + c] = t.veff;
    tm[(3 + 3 * (dir - 1)) * d.ncell + c] = t.upar;
  }
}

// dP of every active direction at the zones -1 .. N of every active direction
// (the zones whose faces are reconstructed), from the terms of the zone and
// its two neighbours along the direction.
__global__ __launch_bounds__(kBlock) void k_weno_dp(Dims d, Box z, const double* __restrict__ tm,
                                                      double* __restrict__ dp)
{
  int i, j, k;
  if (!z.at((long)blockIdx.x * kBlock + threadIdx.x, i, j, k)) return;
  const long c = hd::cidx(d, i, j, k);
  const double rho = tm[c];
#pragma unroll
  for (int dir = 1; dir <= 3; dir++) {
    if ((dir == 1 ? d.n1 : dir == 2 ? d.n2 : d.n3) <= 1) continue;
    const long st = dir == 1 ? 1 : dir == 2 ? (long)d.ps1 : (long)d.ps1 * d.ns2;
    const double* T = tm + (1 + 3 * (dir - 1)) * d.ncell;
    const double par = T[c], upar = T[2 * d.ncell + c];
    const double dv = T[d.ncell + c + st] - T[d.ncell + c - st];
    dp[(dir - 1) * d.ncell + c] = (fabs(rho * dv * dv) + fabs(T[2 * d.ncell + c + st] - T[2 * d.ncell + c - st])) /
                                  (fabs(rho * par) + fabs(upar) + D_SMALL);
  }
}

// Segments of the storage lines along DIR (2 or 3), one primitive each, per
// thread.  Line index c = 0 .. n-1 is the sweep coordinate m = c - 4 (m = -4
// .. N+3); weights exist for c = 2 .. n-3, ratios for c = 2 .. n-3 (the left
// comparison from c = 4, the right one up to c = n-5), faces for c = 3 ..
// n-4 (m = -1 .. N).  A thread owns the faces of zones f0 .. f1 and starts
// five zones earlier so that the ratios of f0-1 and its neighbours exist.  At
// step s the value y[s] arrives, zone s-2 gets its weights, zone s-4 its
// ratio and zone s-5 its faces.  Threads are the lines of the face box's
// transverse plane with i fastest (no lanes padded per row); blockIdx.y is
// (primitive, segment).
template <int DIR>
__global__ __launch_bounds__(kBlock) void k_weno_recon(Dims d, Box b, int seg, const double* __restrict__ pb,
                                                         const double* __restrict__ dpd, double* __restrict__ L,
                                                         double* __restrict__ R)
{
  const int n0 = b.ni;
  const int n1 = DIR == 3 ? b.nj : b.nk;
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (long)n0 * n1) return;
  const int v = blockIdx.y % HARM_NPR, sg = blockIdx.y / HARM_NPR;
  const int t0 = (int)(t % n0), t1 = (int)(t / n0);
  int i, j, k;
  if (DIR == 2) { i = b.is + t0; j = -4; k = b.ks + t1; }
  else { i = b.is + t0; j = b.js + t1; k = -4; }
  const int N = DIR == 2 ? d.n2 : d.n3;
  const int n = N + 8;
  const int f0 = 3 + sg * seg, f1 = min(f0 + seg - 1, n - 4);
  const long st = DIR == 2 ? (long)d.ps1 : (long)d.ps1 * d.ns2;
  const long c0 = hd::cidx(d, i, j, k);
  const double* __restrict__ a = pb + (long)v * d.ncell + c0;
  double* __restrict__ Lo = L + (long)v * d.ncell + c0;
  double* __restrict__ Ro = R + (long)v * d.ncell + c0;
  const double* __restrict__ dpl = dpd + c0;

  double y[8];  // y[7] = y[s], y[0] = y[s-7]
#pragma unroll
  for (int m = 0; m < 8; m++) y[m] = 0.0;
  ZoneW zw[4] = {};  // zw[0] = zone s-2 ... zw[3] = zone s-5
  double z6u0 = 0.0;  // zone s-6's first weight (its ratio partner)
  double fd[3] = {0, 0, 0};  // ratio fractions of zones s-6, s-5, s-4
  const int s0 = max(0, f0 - 5);
  // the loads run two steps (values) and one step (dP) ahead of their use,
  // so that their latency overlaps a step's arithmetic
  double ya = s0 < n ? a[(long)s0 * st] : 0.0;
  double yb = s0 + 1 < n ? a[(long)(s0 + 1) * st] : 0.0;
  double dpa = s0 - 5 >= f0 ? dpl[(long)(s0 - 5) * st] : 0.0;
#pragma unroll 1
  for (int s = s0; s <= f1 + 5; s++) {
#pragma unroll
    for (int m = 0; m < 7; m++) y[m] = y[m + 1];
    y[7] = ya;
    ya = yb;
    yb = s + 2 < n ? a[(long)(s + 2) * st] : 0.0;
    const double dPi = dpa;
    dpa = (s + 1 - 5 >= f0 && s + 1 - 5 <= f1) ? dpl[(long)(s + 1 - 5) * st] : 0.0;
    // weights of zone s-2
    z6u0 = zw[3].u[0];
#pragma unroll
    for (int m = 3; m > 0; m--) zw[m] = zw[m - 1];
    if (s >= 4 && s <= n - 1 && s - 2 >= f0 - 3) zone_weights(y + 5, zw[0]);
    // ratio of zone r = s-4 (zw[2]) against zones r-2 (z6u0) and r+2 (zw[0])
    const int r = s - 4;
    fd[0] = fd[1];
    fd[1] = fd[2];
    fd[2] = 0.0;
    if (r >= 2 && r >= f0 - 1) {
      const double wl = r >= 4 ? side_ratio(zw[2].u[2], z6u0) : kBIG;
      const double wr = r <= n - 5 ? side_ratio(zw[2].u[0], zw[0].u[2]) : kBIG;
      fd[2] = ratio_lof(wl, wr);
    }
    // faces of zone f = s-5 (zw[3], values y[0..4] centred on y[2])
    const int f = s - 5;
    if (f >= f0) {
      const double ss = wmin((fabs(dPi) - 0.5 * 0.5) / (0.5 * 0.5), 1.);
      double lof = fd[1];
      lof = wmax(lof, ss * fd[2]);
      lof = wmax(lof, ss * fd[0]);
      double lo, ro;
      zone_faces(y + 2, zw[3], lof, lo, ro);
      Lo[(long)f * st] = lo;
      Ro[(long)f * st] = ro;
    }
  }
}

// One zone per lane of a GL-lane group (g = 0 .. GL-1 its lane, c its line
// index, y[0..4] the values of zones c-2 .. c+2): the lane forms its zone's
// weights, then its ratio from the weights two lanes away, then its
// lower-order fraction from the ratios one lane away and its faces.  Lanes 3
// .. GL-4 of a group own faces (returns true and sets lo, ro); dp() reads the
// zone's dP and is called by those lanes only.  Every lane of the group must
// call it (the ratios travel by shuffles).  Same operations as the sweep.
template <int GL, class DP>
__device__ __forceinline__ bool lane_faces(const double* y, int c, int n, int g, bool act, DP dp, double& lo,
                                           double& ro)
{
  ZoneW z = {};
  if (act && c >= 2 && c <= n - 3) zone_weights(y + 2, z);
  // ratio of zone c against zones c-2 and c+2
  const double ul = __shfl_up(z.u[0], 2, GL), ur = __shfl_down(z.u[2], 2, GL);
  double fd = 0.0;
  if (act && c >= 2 && c <= n - 3 && g >= 2 && g <= GL - 3) {
    const double wl = c >= 4 ? side_ratio(z.u[2], ul) : kBIG;
    const double wr = c <= n - 5 ? side_ratio(z.u[0], ur) : kBIG;
    fd = ratio_lof(wl, wr);
  }
  const double fdm = __shfl_up(fd, 1, GL), fdp = __shfl_down(fd, 1, GL);
  if (act && c >= 3 && c <= n - 4 && g >= 3 && g <= GL - 4) {
    const double dPi = dp();
    const double ss = wmin((fabs(dPi) - 0.5 * 0.5) / (0.5 * 0.5), 1.);
    double lof = fd;
    lof = wmax(lof, ss * fdp);
    lof = wmax(lof, ss * fdm);
    zone_faces(y + 2, z, lof, lo, ro);
    return true;
  }
  return false;
}

// Direction 1 (lines contiguous in memory): a group of GL lanes holds GL
// consecutive zones c = (GL - 6) w + g of one storage line, so every load and
// store is contiguous; nw = ceil((n - 6) / (GL - 6)) groups cover a line.  A
// wave holds the same segment of two neighbouring lines j (same radii); the
// wave's segment, line pair, k and primitive are decoded once per wave in
// scalar registers.  (Along x2 and x3 the same scheme, with strided loads or
// staged through LDS, measured slower than the sweep.)
constexpr int kX1Lanes = 32;
__global__ __launch_bounds__(kBlock) void k_weno_recon_x1(Dims d, Box b, int nw, const double* __restrict__ pb,
                                                            const double* __restrict__ dpd, double* __restrict__ L,
                                                            double* __restrict__ R)
{
  constexpr int GL = kX1Lanes;
  static_assert(64 / GL == 2, "two line groups per wave");
  const int njp = (b.nj + 1) / 2;
  const int wv = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)));
  if (wv >= njp * nw * b.nk * HARM_NPR) return;  // the whole wave is past the end
  int r = wv;
  const int jp = r % njp;
  r /= njp;
  const int ws = r % nw;
  r /= nw;
  const int kk = r % b.nk, v = r / b.nk;
  const int g = threadIdx.x & (GL - 1);
  const int jj = 2 * jp + ((threadIdx.x & 63) >= GL);
  const bool act = jj < b.nj;
  const int n = d.n1 + 8;
  const int c = (GL - 6) * ws + g;
  const long c0 = hd::cidx(d, -4, b.js + (act ? jj : b.nj - 1), b.ks + kk);
  const double* __restrict__ a = pb + (long)v * d.ncell + c0;

  double y[5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const int cm = c - 2 + m;
    y[m] = (act && cm >= 0 && cm < n) ? a[cm] : 0.0;
  }
  double lo, ro;
  if (lane_faces<GL>(y, c, n, g, act, [&] { return dpd[c0 + c]; }, lo, ro)) {
    L[(long)v * d.ncell + c0 + c] = lo;
    R[(long)v * d.ncell + c0 + c] = ro;
  }
}

}  // namespace

// dP of every active direction over the zones the face reconstructions read;
// tm: 10 x ncell doubles of zone terms
void launch_weno_dp(const Dims& d, hipStream_t s, const double* pb, const double* gcent, const Params& P, double* tm,
                    double* dp)
{
  const int e0 = d.n1 > 1, e1 = d.n2 > 1, e2 = d.n3 > 1;
  Box zt{-2 * e0, -2 * e1, -2 * e2, d.n1 + 4 * e0, d.n2 + 4 * e1, d.n3 + 4 * e2};
  hipLaunchKernelGGL(k_weno_terms, dim3((unsigned)((zt.count() + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, d, zt,
                     pb, gcent, P.gam, tm);
  Box z{-e0, -e1, -e2, d.n1 + 2 * e0, d.n2 + 2 * e1, d.n3 + 2 * e2};
  hipLaunchKernelGGL(k_weno_dp, dim3((unsigned)((z.count() + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, d, z, tm,
                     dp);
}

// face values L, R (SoA like pb) of zones -1 .. N along dir over the
// transverse range of the face box b.
void launch_weno_recon(int dir, const Dims& d, const Box& b, hipStream_t s, const double* pb, const double* dp,
                       double* L, double* R)
{
  const double* dpd = dp + (long)(dir - 1) * d.ncell;
  if (dir == 1) {
    constexpr int GL = kX1Lanes;
    const int n = d.n1 + 8;
    const int nw = (n - 6 + GL - 7) / (GL - 6);
    const long nwave = (long)((b.nj + 1) / 2) * nw * b.nk * HARM_NPR;
    const long blocks = (nwave + kBlock / 64 - 1) / (kBlock / 64);
    hipLaunchKernelGGL(k_weno_recon_x1, dim3((unsigned)blocks), dim3(kBlock), 0, s, d, b, nw, pb, dpd, L, R);
    return;
  }
  // directions 2 and 3: lines cut into segments of at least kMinSeg faces,
  // about kWaveTarget waves in all (a few per SIMD)
  constexpr long kWaveTarget = 16384;  // measured: 4 segments along x3, 8 along x2 at 256x128x64
  constexpr int kMinSeg = 16;
  const int N = dir == 2 ? d.n2 : d.n3;
  const int nf = N + 2;  // zones -1 .. N
  const long nl = (long)b.ni * (dir == 3 ? b.nj : b.nk);
  const long waves1 = (nl * HARM_NPR + 63) / 64;
  int nseg = (int)std::max(1L, std::min((long)(nf / kMinSeg), (kWaveTarget + waves1 - 1) / waves1));
  const int seg = (nf + nseg - 1) / nseg;
  nseg = (nf + seg - 1) / seg;
  const dim3 grid((unsigned)((nl + kBlock - 1) / kBlock), (unsigned)(HARM_NPR * nseg));
  if (dir == 2) hipLaunchKernelGGL(k_weno_recon<2>, grid, dim3(kBlock), 0, s, d, b, seg, pb, dpd, L, R);
  if (dir == 3) hipLaunchKernelGGL(k_weno_recon<3>, grid, dim3(kBlock), 0, s, d, b, seg, pb, dpd, L, R);
}

}  // namespace hk
