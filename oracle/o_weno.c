/*
 * o_weno.c -- oracle restatement of HARM's WENO5BND centre-to-edge line
 * reconstruction of primitives, as the reference compiles it for the torus
 * (lim = WENO5BND, DOENOFLUX = NOENOFLUX).  TEST INFRASTRUCTURE ONLY (see
 * oracle.h): the checker the HIP reconstruction is compared against.
 *
 * Call path restated: slope_lim_linetype_c2e (interpline.c:306) ->
 * slope_lim_linetype_c2e_perijk_precomputeindicators (:468) ->
 * compute_df_line_new (:3557, the dP line) and compute_monotonicity_line
 * (interpline.mono.c:10 -> compute_smonotonicity_line_split
 * interpline.smono.c:75, C2E branch :198) -> pass_1d_line_multipl_weno
 * (reconstructeno.c:2877) -> eno_line_c2e (:264) -> eno_line_reconstruct
 * (:333) with compute_cf_stencil_weights (:1693),
 * compute_optimized_cf_stencil_weights (:1928), the order-2 weights of
 * compute_lower_order_stencil_weights (:1302), mono_adjust_unoptimized_weights
 * (:852), compute_weights_ratios (:2311), apply_additional_reduction_to_weights
 * (:795), choose_weno_order (:2442) and eno_cvt (:884).
 *
 * Compile-time switches in force (reconstructeno_global.h, definit.h,
 * init.fishmon.h): DO_ENO_STENCIL_REDUCTION 1 with the AND form
 * (DO_OR_STENCIL_REDUCTION 0, transition ratios 1.3..1.6),
 * DO_NORMALIZE_SMOOTHNESS_INDICATORS 1, USE_NORMU 1, DO_LIMIT_C2E_WEIGHTS 1
 * (WEIGHTSPREADAROUNDEQUAL 0.2), SYMMETRIZE_WENO 1, SMONO with DO_SMONO_C2E 1,
 * SMONO_DO_SMOOTH_TRANSITION 1, DO_MONO_1ST_DERIVATIVE 1, DO_3RD_DER 1 and no
 * cusp indicator, DO_SMOOTH_ADJUSTMENT_UNOPTIMIZED_WEIGHTS 1,
 * DO_REDUCE_POST_SHOCK/PRE_SHOCK 1 with REDUCE_DP_FRACTION 0.5,
 * STORESHOCKINDICATOR 1 (the Ficalc shock indicator is carried along the line
 * but no c2e step reads it), MERGEDC2EA2CMETHOD 0, WENO_REDUCE_A2C_LOOK_
 * OTHER_DIRECTIONS inactive in c2e (do_weno_lower_order_fraction is 0 outside
 * avg2cen_interp), preforder 5, minorder 3, no causal shift.
 *
 * Two places where the reference reads memory it never wrote, and what is
 * restated instead:
 *  - compute_df_line_new reads the line's primitives from yprim, which the
 *    precomputed-indicator path (interpline.c:468) never fills; the
 *    non-precomputed path of the same function sets yprim = yin when the
 *    interpolated quantities are the primitives (:697-701), so the dP line is
 *    formed from the line's own primitives here.
 *  - b^dir b_t in that dP line comes from the stored flux state, whose b^mu
 *    is never formed (COMPUTE4FIELDforFLUX is 0 with stored wave speeds,
 *    global.depnmemonics.h:76) and holds valueinit, 0 in a production build:
 *    the term is 0 here.
 */
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include "o_grid.h"

#define W_SMALL 1.e-300
#define W_BIG 1.e+100
/* SQRT_WENO_EPSILON, WENO_EPSILON = REDUCEEPSILON (reconstructeno_global.h) */
#define SQRT_WEPS (500.0 * DBL_EPSILON)
#define WEPS (SQRT_WEPS * SQRT_WEPS)
#define MINNUM DBL_MIN

static inline double wmin(double a, double b) { return a < b ? a : b; }
static inline double wmax(double a, double b) { return a > b ? a : b; }
static inline double wminmod(double a, double b) { return (a * b <= 0) ? 0.0 : ((fabs(a) < fabs(b)) ? a : b); }

/* get_sum_of_elements (reconstructeno.c:2261), SYMMETRIZE_WENO: mirror pairs
 * first, the middle element last */
static double symsum(int n, const double* a)
{
  double s = 0.0;
  int i;
  for (i = 0; i < n / 2; i++) s += a[i] + a[n - 1 - i];
  s += a[i] * (n % 2);
  return s;
}

/* normalize_array (:2292) */
static double normalize(int n, double eps, double* a)
{
  const double s = symsum(n, a);
  for (int i = 0; i < n; i++) a[i] /= (s + eps);
  return s;
}

/* limit_weight (:3015) */
static double limit_weight(int order, double w)
{
  const double avg = 1. / order, w1 = 0.2 * avg, w2 = (1.0 - 0.2) * avg, w3 = (1.0 + 0.2) * avg, w4 = 1.0;
  if (w < w1) return 0.0;
  if (w < w2) return avg * (w - w1) / (w2 - w1);
  if (w < w3) return avg;
  return (w - w3) * (w4 - avg) / (w4 - w3) + avg;
}

/* Optimal (linear) weights that combine the order-k sub-stencils into the
 * (2k-1)-point interpolant at the left and right faces; index = shift, the
 * sub-stencil of shift s covering cells [-s, k-1-s]. */
static const double optL3[3] = {1. / 16., 5. / 8., 5. / 16.}, optR3[3] = {5. / 16., 5. / 8., 1. / 16.};
static const double optL2[2] = {1. / 4., 3. / 4.}, optR2[2] = {3. / 4., 1. / 4.};
/* Lagrange interpolation of sub-stencil s (points u[j-s], j = 0..k-1) to the
 * left face x = -1/2 and the right face x = +1/2 */
static const double c2l3[3][3] = {{15. / 8., -5. / 4., 3. / 8.}, {3. / 8., 3. / 4., -1. / 8.}, {-1. / 8., 3. / 4., 3. / 8.}};
static const double c2r3[3][3] = {{3. / 8., 3. / 4., -1. / 8.}, {-1. / 8., 3. / 4., 3. / 8.}, {3. / 8., -5. / 4., 15. / 8.}};
static const double c2l2[2][2] = {{3. / 2., -1. / 2.}, {1. / 2., 1. / 2.}};
static const double c2r2[2][2] = {{1. / 2., 1. / 2.}, {-1. / 2., 3. / 2.}};
/* five-point interpolant at the faces, points u[-2..2] (simple_eno :1034) */
static const double c2l5[5] = {-5. / 128., 15. / 32., 45. / 64., -5. / 32., 3. / 128.};
static const double c2r5[5] = {3. / 128., -5. / 32., 45. / 64., 15. / 32., -5. / 128.};

/* The weights of one cell: order 3 (WENO5) and its order-2 (WENO3) companion. */
typedef struct {
  double u[3], L[3], R[3]; /* unoptimized (roughness) and the two face-optimized sets */
  double lof, fordpp;      /* lower-order fraction, and its copy before the dP reduction */
  double u2[2], L2[2], R2[2];
} wcell;

/* compute_cf_stencil_weights (:1693) for order 3 or 2 at the cell u[0];
 * min_shift..max_shift are the sub-stencils inside the line.  Returns 1 when
 * the SMONO indicator is 1 (equal weights, nothing else formed). */
static int cf_weights(int order, const double* u, double mono, int min_shift, int max_shift, double* w)
{
  double si[3];
  if (mono == 1) {
    for (int s = 0; s < order; s++) w[s] = 1.0 / order;
    return 1;
  }
  if (order == 3) {
    const double d1[3] = {-(3 * u[0] + u[2]) + 4 * u[1], u[1] - u[-1], (3 * u[0] + u[-2]) - 4 * u[-1]};
    const double d2[3] = {(u[0] + u[2]) - 2 * u[1], (u[-1] + u[1]) - 2 * u[0], (u[0] + u[-2]) - 2 * u[-1]};
    for (int s = 0; s < 3; s++) si[s] = 13. / 12. * d2[s] * d2[s] + 1. / 4. * d1[s] * d1[s];
  } else {
    const double d1[2] = {u[1] - u[0], u[0] - u[-1]};
    for (int s = 0; s < 2; s++) si[s] = d1[s] * d1[s];
  }
  /* desensitise_smoothness_indicators (:2001): + |u|^2 over the stencil x eps */
  double u2[5];
  for (int i = -order + 1; i < order; i++) u2[i + order - 1] = u[i] * u[i];
  const double err = symsum(2 * order - 1, u2) * WEPS;
  for (int s = 0; s < order; s++) si[s] += err;
  for (int s = 0; s < order; s++)
    if (s < min_shift || s > max_shift) si[s] = 0.0;
  normalize(order, MINNUM, si);
  /* compute_weights_from_smoothness_indicators (:2211) */
  for (int s = 0; s < order; s++) {
    double t = si[s] + WEPS;
    t *= t;
    w[s] = 1. / t;
  }
  for (int s = 0; s < order; s++)
    if (s < min_shift || s > max_shift) w[s] = 0.0;
  normalize(order, 0, w);
  for (int s = 0; s < order; s++) w[s] = limit_weight(order, w[s]);
  normalize(order, 0, w);
  return 0;
}

/* compute_optimized_cf_stencil_weights (:1928) */
static void cf_optimized(int order, const double* w, double* L, double* R)
{
  const double* oL = order == 3 ? optL3 : optL2;
  const double* oR = order == 3 ? optR3 : optR2;
  for (int s = 0; s < order; s++) {
    L[s] = w[s] * oL[s];
    R[s] = w[s] * oR[s];
  }
  normalize(order, MINNUM, L);
  normalize(order, MINNUM, R);
}

/* eno_cvt (:884): weighted sum of the sub-stencil interpolants, each formed
 * only where its weight is non-zero */
static double eno_cvt(int order, int right, const double* w, const double* u)
{
  double upd[3];
  for (int s = 0; s < order; s++) {
    upd[s] = 0.0;
    if (w[s] != 0.0) {
      double pts[3];
      for (int j = 0; j < order; j++) {
        const double m = order == 3 ? (right ? c2r3[s][j] : c2l3[s][j]) : (right ? c2r2[s][j] : c2l2[s][j]);
        pts[j] = m * u[j - s];
      }
      upd[s] = w[s] * symsum(order, pts);
    }
  }
  return symsum(order, upd);
}

/* compute_mono_indicator_point_eno5 (interpline.smono.c:517): the first,
 * second and third derivatives of the five-point interpolant at the two
 * neighbouring zone centres x = -1, +1 (first derivatives also limited by the
 * outer one-sided differences) must
 * agree in sign; the smallest of their minmods, against eps times the
 * stencil's magnitude, ramps the indicator from 0 to 1 */
double o_smono_indicator(const double* y, double eps)
{
  double L[4], R[4];
  L[1] = (((-3 * y[-2] + y[2]) + (-10 * y[-1] - 6 * y[1])) + 18 * y[0]) / 12.;
  L[2] = (((11 * y[-2] - y[2]) + (-20 * y[-1] + 4 * y[1])) + 6 * y[0]) / 12.;
  L[3] = (((-3 * y[-2] - y[2]) + (10 * y[-1] + 6 * y[1])) - 12 * y[0]) / 2.;
  R[1] = (((-y[-2] + 3 * y[2]) + (6 * y[-1] + 10 * y[1])) - 18 * y[0]) / 12.;
  R[2] = (((-y[-2] + 11 * y[2]) + (4 * y[-1] - 20 * y[1])) + 6 * y[0]) / 12.;
  R[3] = (((y[-2] + 3 * y[2]) - (6 * y[-1] + 10 * y[1])) + 12 * y[0]) / 2.;
  L[1] = wminmod(L[1], y[-1] - y[-2]);
  R[1] = wminmod(R[1], y[2] - y[1]);
  double minabs = W_BIG;
  for (int n = 1; n <= 3; n++) minabs = wmin(minabs, fabs(wminmod(L[n], R[n])));
  const double norm = fabs(y[0]) + ((fabs(y[-1]) + fabs(y[1])) + (fabs(y[-2]) + fabs(y[2])));
  const double cut = norm * eps + MINNUM;
  /* transition_function (:612) */
  return wmax(0., wmin(1., (minabs - 0.0) / (cut - 0.0)));
}

/* simple_eno (reconstructeno.c:1034) with the five-point interpolant */
static double eno5(const double* y, int right)
{
  double d[5];
  const double* c = right ? c2r5 : c2l5;
  for (int i = -2; i <= 2; i++) d[i + 2] = c[i + 2] * y[i];
  return symsum(5, d);
}

/* One line of one primitive: y[0..n-1] is the line from bs to be (the whole
 * storage row along the sweep), ng its ghost depth (>= 4), so the cells to
 * reconstruct are ps = ng-1 .. pe = n-ng (the reference's -1 .. N).  dP[c]
 * is the dP line at those cells.  yl[c], yr[c]: the value at the cell's left
 * and right face. */
static void weno_line(const double* y, int n, int ng, const double* dP, double* yl, double* yr, double* lof_out,
                      double* fordpp_out)
{
  const int bs = 0, be = n - 1, ps = ng - 1, pe = n - ng;
  double mono[n], el[n], er[n];
  wcell W[n];
  memset(mono, 0, sizeof(mono));
  memset(el, 0, sizeof(el));
  memset(er, 0, sizeof(er));
  memset(W, 0, sizeof(W));

  /* SMONO c2e (interpline.smono.c:198): cells bs+2 .. be-2 */
  for (int c = bs + 2; c <= be - 2; c++) {
    mono[c] = o_smono_indicator(y + c, SQRT_WEPS);
    if (mono[c] != 0) {
      el[c] = eno5(y + c, 0);
      er[c] = eno5(y + c, 1);
    }
  }
  /* unoptimized weights (eno_line_reconstruct :424-445) */
  int ws = bs + 2, wf = be - 2;
  if (ws > ps) ws = ps;
  if (wf < pe) wf = pe;
  for (int c = ws; c <= wf; c++) {
    wcell* w = &W[c];
    w->lof = 0.0;
    const int minidx = bs - c, maxidx = be - c;
    cf_weights(3, y + c, mono[c], 3 - 1 - maxidx, -minidx, w->u);
    /* the order-2 companion: unoptimized and optimized (skipped when SMONO
     * holds the cell: the reference leaves them unset there, and never reads
     * them) */
    if (!cf_weights(2, y + c, mono[c], 2 - 1 - maxidx, -minidx, w->u2)) cf_optimized(2, w->u2, w->L2, w->R2);
    /* mono_adjust_unoptimized_weights (:852) */
    if (mono[c] > 0.0 && mono[c] < 1.0)
      for (int s = 0; s < 3; s++) w->u[s] = (1.0 - mono[c]) * w->u[s] + mono[c] * (1.0 / 3);
  }
  /* compute_weights_ratios (:2311), AND reduction: the cell's outer
   * sub-stencil weight against the neighbour's weight of the same cells */
  for (int c = ps - 1; c <= pe + 1; c++) {
    double wl = W_BIG, wr = W_BIG;
    if (c >= bs + 4 && c <= be) wl = (W[c].u[2] + WEPS) / (W[c - 2].u[0] + WEPS);
    if (c >= bs && c <= be - 4) wr = (W[c].u[0] + WEPS) / (W[c + 2].u[2] + WEPS);
    const double wrat = wmin(wl, wr);
    const double min_rat = 1.3, max_rat = 1.6;
    double lof = (wrat - min_rat) / (max_rat - min_rat);
    lof = wmax(0., wmin(lof, 1.0));
    W[c].lof = wmax(lof, W[c].lof);
    W[c].fordpp = W[c].lof;
  }
  /* apply_additional_reduction_to_weights (:795): pre/post-shock reduction
   * from the line's pressure-jump measure */
  for (int c = ps; c <= pe; c++) {
    const double dPi = dP[c];
    const double ss = wmin((fabs(dPi) - 0.5 * 0.5) / (0.5 * 0.5), 1.);
    double lof = W[c].lof;
    lof = wmax(lof, ss * W[c + 1].fordpp);
    lof = wmax(lof, ss * W[c - 1].fordpp);
    W[c].lof = lof;
  }
  /* optimized weights from the (SMONO-adjusted) unoptimized ones */
  for (int c = ws; c <= wf; c++)
    if (!(mono[c] == 1 || mono[c] == -1)) cf_optimized(3, W[c].u, W[c].L, W[c].R);
  /* choose_weno_order (:2442) and the output loop (:538-608) */
  for (int c = ps; c <= pe; c++) {
    const wcell* w = &W[c];
    const double m = mono[c];
    double out[2] = {el[c], er[c]};
    if (m == 0.0) out[0] = out[1] = 0.0;
    if (m >= 0.0 && m < 1.0) {
      double L[3], R[3], L2[2], R2[2];
      const double lof = w->lof;
      const int two = 0.0 < lof;
      for (int s = 0; s < 3; s++) {
        L[s] = two ? w->L[s] * (1. - lof) : w->L[s];
        R[s] = two ? w->R[s] * (1. - lof) : w->R[s];
      }
      if (two)
        for (int s = 0; s < 2; s++) {
          L2[s] = w->L2[s] * lof;
          R2[s] = w->R2[s] * lof;
        }
      for (int side = 0; side < 2; side++) {
        double v = 0.0;
        v += eno_cvt(3, side, side ? R : L, y + c);
        if (two) v += eno_cvt(2, side, side ? R2 : L2, y + c);
        out[side] = (1.0 - m) * v + m * out[side];
      }
    }
    yl[c] = out[0];
    yr[c] = out[1];
  }
  if (lof_out)
    for (int c = ps; c <= pe; c++) {
      lof_out[c] = W[c].lof;
      fordpp_out[c] = W[c].fordpp;
    }
}

void o_weno5bnd_line(const double* y, int n, int ng, const double* dP, double* yl, double* yr)
{
  weno_line(y, n, ng, dP, yl, yr, NULL, NULL);
}

/* the same, also returning each reconstructed cell's final lower-order
 * fraction and the fraction before the dP reduction (for the tests) */
void o_weno5bnd_line_lof(const double* y, int n, int ng, const double* dP, double* yl, double* yr, double* lof,
                         double* fordpp)
{
  weno_line(y, n, ng, dP, yl, yr, lof, fordpp);
}

/* compute_df_line_new (interpline.c:3557) for C2E: per cell of the line, from
 * its state at the zone centre, the dP measure
 *   dP = (|rho (veff(+1) - veff(-1))^2| + |upar(+1) - upar(-1)|) / (|rho par| + |upar| + SMALL)
 * with par = |(1 + u_t) u^dir|, veff = sign(gdet rho u^dir) par / (sqrt(par) + SMALL),
 * upar = |(u + p + b^2)(|u^dir| u_t) + (p + b^2/2) + b_t b^dir| (b_t b^dir = 0,
 * see the header).  st[c*4+..] = {rho, par, veff, upar} per cell (filled by
 * o_dp_cell); dP[c] for c = 1 .. n-2. */
void o_dp_cell(const double* pr, const o_geom* G, double gam, int dir, double* out4)
{
  o_state q;
  o_get_state(pr, G, gam, &q);
  const double rho = pr[O_RHO], u = pr[O_UU];
  const double pgas = (gam - 1.0) * u;
  const double V = G->gdet * rho * q.ucon[dir];
  const double bsq = q.bsq, btt = 0.0;
  const double ptot = pgas + 0.5 * bsq;
  const double par = fabs(q.p1ud0 * q.ucon[dir]);
  const double veff = copysign(1.0, V) * par / (sqrt(par) + W_SMALL);
  const double upar = fabs((u + pgas + bsq) * (fabs(q.ucon[dir]) * q.ucov[0]) + ptot + btt);
  out4[0] = rho;
  out4[1] = par;
  out4[2] = veff;
  out4[3] = upar;
}

void o_dp_line(const double* st, int n, double* dP)
{
  for (int c = 1; c <= n - 2; c++) {
    const double* a = st + 4 * c;
    const double* l = st + 4 * (c - 1);
    const double* r = st + 4 * (c + 1);
    const double dv = r[2] - l[2];
    dP[c] = (fabs(a[0] * dv * dv) + fabs(r[3] - l[3])) / (fabs(a[0] * a[1]) + fabs(a[3]) + W_SMALL);
  }
}

/* o_dp_cell at a 27-field geometry record (for the tests) */
void o_dp_cell_rec(const double* pr, const double* rec, double gam, int dir, double* out4)
{
  o_geom G;
  o_geom_from_rec(rec, &G);
  o_dp_cell(pr, &G, gam, dir, out4);
}
