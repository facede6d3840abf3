/*
 * o_grid.c -- oracle restatement of HARM's advance() hot path on a grid:
 * boundary conditions, reconstruction (PARA, donor; WENO5BND lines in
 * o_weno.c), face Riemann
 * fluxes (LAX-F, HLL), Toth flux-CT, the conserved update with connection
 * source terms, the inversion with floors, failure averaging, the timestep
 * and the RK2/RK3 step driver.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Reference files restated (HARM tree): flux.c:852-1140 (fluxcalc_standard),
 * flux.c:1952-2110 (getplpr, check_plpr), interppoint.para.c:568-1200 (para4),
 * interpline.c:306-640 (line extraction for WENO5BND), fluxcompute.c
 * (flux_compute, HLL/LAXF), wavespeeds.c:397 (cminmax_calc), fluxct.c:186-453,
 * 1106-1161 (Toth), advance.c:140-865, 1866-2075, 2331-2620 (update, dt),
 * fixup.c:561-830, 1061-1330, 2087-2200, 2723-2850, 3255-3390, bounds.tools.c
 * :172-1215, 1417-2010, 2004-2560 (extrapolation, polar axis, pole death),
 * step_ch.c:140-235, 524-845, 1171-1310 (RK driver).
 */
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include "o_grid.h"

#define SMALL 1.e-300
#define BIG 1.e+100
#define NPR 8



static inline double minmod(double a, double b) { return ((a) * (b) <= 0) ? 0.0 : ((fabs(a) < fabs(b)) ? (a) : (b)); }

void o_dims(const o_grid* g, int* ng1, int* ng2, int* ng3, int* ns1, int* ns2, int* ns3)
{
  *ng1 = g->n1 > 1 ? g->ng : 0;
  *ng2 = g->n2 > 1 ? g->ng : 0;
  *ng3 = g->n3 > 1 ? g->ng : 0;
  *ns1 = g->n1 + 2 * (*ng1);
  *ns2 = g->n2 + 2 * (*ng2);
  *ns3 = g->n3 + 2 * (*ng3);
}

typedef struct {
  int n1, n2, n3, ng1, ng2, ng3, ns1, ns2, ns3;
} dims_t;

static dims_t D(const o_grid* g)
{
  dims_t d;
  d.n1 = g->n1; d.n2 = g->n2; d.n3 = g->n3;
  o_dims(g, &d.ng1, &d.ng2, &d.ng3, &d.ns1, &d.ns2, &d.ns3);
  return d;
}

static inline long ix(const dims_t* d, int i, int j, int k)
{
  return (((long)(k + d->ng3) * d->ns2 + (j + d->ng2)) * d->ns1 + (i + d->ng1)) * NPR;
}
static inline long ig2(const dims_t* d, int i, int j) { return (long)(j + d->ng2) * d->ns1 + (i + d->ng1); }

static void geom_at(const o_grid* g, const dims_t* d, int loc, int i, int j, o_geom* G)
{
  o_geom_from_rec(g->geom[loc] + ig2(d, i, j) * 27, G);
}

static inline double igdetnosing(double gdet) { return (gdet >= 0.0 ? 1.0 : -1.0) / (fabs(gdet) + SMALL); }

/* gamma_calc (phys.tools.c:1975) for VELREL4 */
static double gamma_of(const double* pr, const o_geom* G)
{
  const double u1 = pr[O_U1], u2 = pr[O_U2], u3 = pr[O_U3];
  double qsq = G->gcov[1][1] * u1 * u1 + G->gcov[2][2] * u2 * u2 + G->gcov[3][3] * u3 * u3
               + 2. * (G->gcov[1][2] * u1 * u2 + G->gcov[1][3] * u1 * u3);
  qsq = qsq * (qsq >= 0.0);
  return sqrt(1. + qsq);
}

/* limit_gamma (fixup.c:2723), DO_CONSERVE_D=0 */
/* returns 1 when the velocity was rescaled (limit_gamma's -1, fixup.c:2891) */
static int limit_gamma(double gammamax, double* pr, const o_geom* G)
{
  if (gammamax <= 1.0) return 0;
  const double gamma = gamma_of(pr, G);
  if ((gamma > gammamax && (gamma != 1.0))) {
    const double pref = (gammamax * gammamax - 1.) / (gamma * gamma - 1.);
    const double f = sqrt(pref);
    pr[O_U1] *= f;
    pr[O_U2] *= f;
    pr[O_U3] *= f;
    return 1;
  }
  return 0;
}

/* diag_fixup (fixup.c:220-405) on the final substep, the original HARM
 * branch: fl[pl] += dVF * (Uf[pl] - Ui[pl]) with Ui = primtoU(UDIAG, pr0),
 * Uf = primtoU(UDIAG, pr); UDIAG equals the evolved form here (WITHGDET,
 * UtoU_evolve2diag phys.tools.c:740), dVF = dx1 dx2 dx3 (coord.c:2306). */
static void diag_fixup_acc(const o_grid* g, const o_geom* G, const double* pr0, const double* pr, double* fl)
{
  o_state q;
  double Ui[NPR], Uf[NPR];
  o_get_state(pr0, G, g->gam, &q);
  o_primtoflux(pr0, &q, 0, G, Ui);
  o_get_state(pr, G, g->gam, &q);
  o_primtoflux(pr, &q, 0, G, Uf);
  const double dVF = g->dx[1] * g->dx[2] * g->dx[3];
  for (int pl = 0; pl < NPR; pl++) {
    const double deltaUavg = Uf[pl] - Ui[pl];
    fl[pl] += dVF * deltaUavg;
  }
}

/* ---------------------------------------------------------------------- */
/* Boundary conditions                                                    */
/* ---------------------------------------------------------------------- */

/* extrapfunc (bounds.tools.c:1417, HORIZONEXTRAP=OUTEREXTRAP=3) for one (j,k) line */
static void extrap_line(const o_grid* g, const dims_t* d, double* p, int up, int j, int k)
{
  const int ri = up ? d->n1 - 1 : 0;
  const int ri2 = up ? ri - 1 : ri + 1;
  const int ri3 = up ? ri - 2 : ri + 2;
  const double signdq = up ? -1.0 : 1.0;
  o_geom Gr, Gr2, Gr3;
  geom_at(g, d, 0, ri, j, &Gr);
  geom_at(g, d, 0, ri2, j, &Gr2);
  geom_at(g, d, 0, ri3, j, &Gr3);
  const double* pr = p + ix(d, ri, j, k);
  const double* pr2 = p + ix(d, ri2, j, k);
  const double* pr3 = p + ix(d, ri3, j, k);
  double Bd3 = 0.0, Bd3ri2 = 0.0, Bd3ri3 = 0.0;
  for (int jj = 1; jj < 4; jj++) Bd3 += pr[O_B1 + jj - 1] * Gr.gcov[3][jj];
  for (int jj = 1; jj < 4; jj++) Bd3ri2 += pr2[O_B1 + jj - 1] * Gr2.gcov[3][jj];
  for (int jj = 1; jj < 4; jj++) Bd3ri3 += pr3[O_B1 + jj - 1] * Gr3.gcov[3][jj];
  double Dqp = Bd3ri3 - Bd3ri2, Dqm = Bd3ri2 - Bd3, Dqc = Bd3ri3 - Bd3;
  const double dqBd3 = signdq * minmod(minmod(Dqp, Dqm), Dqc);
  o_state qr, qr2;
  o_get_state(pr, &Gr, g->gam, &qr);
  o_get_state(pr2, &Gr2, g->gam, &qr2);
  double dq[NPR], dql[NPR], dqg[NPR];
  for (int pl = 0; pl < NPR; pl++) {
    Dqp = pr3[pl] - pr2[pl];
    Dqm = pr2[pl] - pr[pl];
    Dqc = pr3[pl] - pr[pl];
    dq[pl] = signdq * minmod(minmod(Dqp, Dqm), Dqc);
  }
  for (int pl = 0; pl < NPR; pl++) {
    Dqp = log(SMALL + fabs(pr3[pl])) - log(SMALL + fabs(pr2[pl]));
    Dqm = log(SMALL + fabs(pr2[pl])) - log(SMALL + fabs(pr[pl]));
    Dqc = log(SMALL + fabs(pr3[pl])) - log(SMALL + fabs(pr[pl]));
    dql[pl] = signdq * minmod(minmod(Dqp, Dqm), Dqc);
  }
  for (int pl = 0; pl < NPR; pl++) {
    Dqp = (Gr3.gdet) * pr3[pl] - (Gr2.gdet) * pr2[pl];
    Dqm = (Gr2.gdet) * pr2[pl] - (Gr.gdet) * pr[pl];
    Dqc = (Gr3.gdet) * pr3[pl] - (Gr.gdet) * pr[pl];
    dqg[pl] = signdq * minmod(minmod(Dqp, Dqm), Dqc);
  }
  const double xdqfrac = fabs(qr.ucon[0] - qr2.ucon[0]) / qr.ucon[0];
  double ydqfrac;
  if (xdqfrac < 0.1) ydqfrac = 0.0;
  else if (xdqfrac < 0.2) ydqfrac = 0.0 + (1.0 - 0.0) / (0.2 - 0.1) * (xdqfrac - 0.1);
  else ydqfrac = 1.0;
  for (int jj = 0; jj < 4; jj++) {
    dq[O_UU + jj] = dq[O_UU + jj] * (1.0 - ydqfrac);
    dql[O_UU + jj] = dql[O_UU + jj] * (1.0 - ydqfrac);
    dqg[O_UU + jj] = dqg[O_UU + jj] * (1.0 - ydqfrac);
  }
  const int is = up ? d->n1 : -d->ng1;
  const int ie = up ? d->n1 + d->ng1 - 1 : -1;
  for (int i = is; i <= ie; i++) {
    double* pi = p + ix(d, i, j, k);
    for (int pl = 0; pl < NPR; pl++) pi[pl] = pr[pl];
    o_geom Gi;
    geom_at(g, d, 0, i, j, &Gi);
    const double di = (double)(i - ri);
    for (int pl = 0; pl <= O_UU; pl++) {
      const double ftemp = exp(-signdq * dql[pl]) - 1.1;
      double mydqlog;
      if (ftemp >= 0.0) mydqlog = log(1.1) / (-signdq);
      else mydqlog = dql[pl];
      pi[pl] = exp(log(SMALL + fabs(pr[pl])) + mydqlog * di);
    }
    for (int pl = O_U1; pl <= O_U3; pl++) pi[pl] = pr[pl] + dq[pl] * di;
    for (int pl = O_B1; pl <= O_B2; pl++) {
      const double ign = igdetnosing(Gi.gdet);
      pi[pl] = (pr[pl] * fabs((Gr.gdet) + dqg[pl] * di) * ign);
    }
    const double Bu1 = pi[O_B1], Bu2 = pi[O_B2];
    const double gcon03 = Gi.gcon[0][3], gcon13 = Gi.gcon[1][3], gcon23 = Gi.gcon[2][3], gcon33 = Gi.gcon[3][3];
    const double gcov01 = Gi.gcov[0][1], gcov02 = Gi.gcov[0][2], gcov11 = Gi.gcov[1][1];
    const double gcov12 = Gi.gcov[1][2], gcov21 = gcov12, gcov22 = Gi.gcov[2][2];
    const double gcov03 = Gi.gcov[0][3], gcov13 = Gi.gcov[1][3], gcov23 = Gi.gcov[2][3];
    const double myBd3 = Bd3 + dqBd3 * di;
    const double ftemp = (1.0 - gcon03 * gcov03 - gcon13 * gcov13 - gcon23 * gcov23);
    const double ign = igdetnosing(ftemp);
    pi[O_B3] = (myBd3 * gcon33 + Bu1 * gcon03 * gcov01 + Bu2 * gcon03 * gcov02 + Bu1 * gcon13 * gcov11
                + Bu2 * gcon13 * gcov12 + Bu1 * gcon23 * gcov21 + Bu2 * gcon23 * gcov22) * ign;
  }
}

/* inflow_check_rel4vel (fixup.c:3255) for dir 1 at the outer boundary */
static void inflow_check_x1up(double* pr, const o_geom* G, double gam, double gammamax)
{
  o_state q;
  o_get_state(pr, G, gam, &q);
  if (!(q.ucon[1] < 0.)) return;
  const double gamma = gamma_of(pr, G);
  pr[O_U1] /= gamma;
  pr[O_U2] /= gamma;
  pr[O_U3] /= gamma;
  const double alpha = 1. / sqrt(-G->gcon[0][0]);
  const double betacon = G->gcon[0][1] * alpha * alpha;
  pr[O_U1] = betacon / alpha;
  double vsq = 0.;
  for (int j = 1; j < 4; j++)
    for (int k = 1; k < 4; k++) vsq += G->gcov[j][k] * pr[O_U1 + j - 1] * pr[O_U1 + k - 1];
  if (vsq < 0.0) vsq = 0.0;
  if (vsq >= 1.0) vsq = 1.0 - 1.0 / (gammamax * gammamax);
  const double gnew = 1. / sqrt(1. - vsq);
  pr[O_U1] *= gnew;
  pr[O_U2] *= gnew;
  pr[O_U3] *= gnew;
}

/* ---- Bondi problem (TESTNUMBER 153) ------------------------------------ */
/* bondi_compute_integrals (initboundcode/bounds.sasha.c:896): the zone's
 * primitives as Kerr-Schild VEL4 (metp2met2bl: u^r = (dr/dX1) u^X1 with
 * dr/dX1 = r - R0), T = p/rho, C1 = T^n u^r r^2, C2 = (1+(1+n)T)^2
 * (1 - 2/r + (u^r)^2), K = p/rho^Gamma. */
static void bondi_integrals(const o_grid* g, const double* pr, const o_geom* G, double r, double* C1, double* C2,
                            double* K)
{
  o_state q;
  o_get_state(pr, G, g->gam, &q);
  const double uu1 = (r - g->R0) * q.ucon[1];
  const double n = 1 / (g->gam - 1);
  const double p = (g->gam - 1) * pr[O_UU];
  const double rho = pr[O_RHO];
  const double T = p / rho;
  *C1 = pow(T, n) * uu1 * r * r;
  *C2 = pow(1 + (1 + n) * T, 2.) * (1 - 2. / r + uu1 * uu1);
  *K = p / pow(rho, g->gam);
}

/* bondi_invert_integrals (bounds.sasha.c:1101-1160), its iterations as
 * written (the max_iter test after the loops never fires there either). */
static int bondi_invert(double n, double C1, double C2, double K, double r, double* primout)
{
  const double relative_accuracy = 1.0e-15;
  const int max_iter = 250;
  double T0 = 0.2, T1 = 0.1, u1;
  int iter;
  if (r > 0.35 && r < 2.5) {
    for (iter = 0; iter < max_iter && fabs((T0 - T1) / T1) > relative_accuracy; iter++) {
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
    if (T1 > T0) return 1;
  } else
    return 1;
  for (int pl = 0; pl < NPR; pl++) primout[pl] = 0.0;
  const double rho = pow(T1 / K, n);
  primout[O_RHO] = rho;
  primout[O_UU] = rho * T1 * n;
  primout[O_U1] = C1 / (pow(T1, n) * r * r);
  return 0;
}

/* bondi_generate_solution (bounds.sasha.c:946) at ghost zone (i, j): the
 * Kerr-Schild (spin 0) u^t from the normalisation (pr2ucon VEL4: g_tt =
 * -(1-2/r), g_tr = 2/r, g_rr = 1+2/r), the 4-velocity taken to the MKS
 * coordinates (mettometpface: u^X1 = u^r/(r - R0)), then the relative
 * 4-velocity u~^i = u^i - u^t g^ti/g^tt (ucon2pr). */
static void bondi_ghost(const o_grid* g, const dims_t* d, double C1, double C2, double K, int i, int j,
                        double* primout)
{
  const double r = g->R0 + exp(g->startx[1] + (i + 0.5) * g->dx[1]);
  if (bondi_invert(1 / (g->gam - 1), C1, C2, K, r, primout) != 0) {
    for (int pl = 0; pl < NPR; pl++) primout[pl] = nan("");
    return;
  }
  const double ur = primout[O_U1];
  const double gtt = -(1. - 2. / r), gtr = 2. / r, grr = 1. + 2. / r;
  const double AA = gtt, BB = 2. * gtr * ur, CC = 1. + grr * ur * ur;
  const double ut = (-BB - sqrt(BB * BB - 4. * AA * CC)) / (2. * AA);
  const double u1 = ur / (r - g->R0);
  o_geom G;
  geom_at(g, d, 0, i, j, &G);
  primout[O_U1] = u1 - ut * G.gcon[0][1] / G.gcon[0][0];
  primout[O_U2] = 0.0 - ut * G.gcon[0][2] / G.gcon[0][0];
  primout[O_U3] = 0.0 - ut * G.gcon[0][3] / G.gcon[0][0];
}

/* X1DN BONDIINTOUTFLOW and X1UP BONDIMDOTOUTFLOW (bounds.sasha.c:182-198,
 * :341-353) with BONDIEXTRAPORDER 1: the integrals of the boundary zone are
 * carried unchanged into every ghost zone (outside: C1 only, with the exact
 * C2 = 1.373125 and K = 1). */
static void bondi_x1(const o_grid* g, const dims_t* d, double* p)
{
#pragma omp parallel for collapse(2) num_threads(g->nthreads > 0 ? g->nthreads : 1)
  for (int k = 0; k < d->n3; k++)
    for (int j = 0; j < d->n2; j++)
      for (int up = 0; up < 2; up++) {
        const int ri = up ? d->n1 - 1 : 0;
        o_geom Gr;
        geom_at(g, d, 0, ri, j, &Gr);
        const double rr = g->R0 + exp(g->startx[1] + (ri + 0.5) * g->dx[1]);
        double C1, C2, K;
        bondi_integrals(g, p + ix(d, ri, j, k), &Gr, rr, &C1, &C2, &K);
        if (up) {
          C2 = 1.373125;
          K = 1.;
        }
        for (int gs = 0; gs < d->ng1; gs++) {
          const int i = up ? d->n1 + gs : -1 - gs;
          bondi_ghost(g, d, C1, C2, K, i, j, p + ix(d, i, j, k));
        }
      }
}

static void bound_x1(const o_grid* g, const dims_t* d, double* p)
{
  if (g->problem == 1) {
    bondi_x1(g, d, p);
    return;
  }
#pragma omp parallel for collapse(2) num_threads(g->nthreads > 0 ? g->nthreads : 1)
  for (int k = 0; k < d->n3; k++)
    for (int j = 0; j < d->n2; j++) {
      extrap_line(g, d, p, 0, j, k); /* X1DN: FREEOUTFLOW */
      extrap_line(g, d, p, 1, j, k); /* X1UP: OUTFLOW + inflow check */
      for (int i = d->n1; i < d->n1 + d->ng1; i++) {
        o_geom Gi;
        geom_at(g, d, 0, i, j, &Gi);
        double* pi = p + ix(d, i, j, k);
        inflow_check_x1up(pi, &Gi, g->gam, g->gammamax);
        limit_gamma(g->gammamax, pi, &Gi);
      }
    }
}

/* bound_x2dn/x2up_polaraxis (bounds.tools.c:751, :996 and the _full3d forms
 * with the pi rotation in phi when N3>1): each pole is a reflection loop over
 * all (i,k) columns followed by poledeath (:2004) over all columns, with
 * POLEDEATH=1, POLEINTERPTYPE=3, POLEGAMMADEATH=1. */
static void x2_reflect(const o_grid* g, const dims_t* d, double* p, int side)
{
  const int n2 = d->n2, n3 = d->n3;
#pragma omp parallel for collapse(2) num_threads(g->nthreads > 0 ? g->nthreads : 1)
  for (int k = 0; k < n3; k++)
    for (int i = -d->ng1; i < d->n1 + d->ng1; i++) {
      const int rk = (n3 > 1) ? (k + n3 / 2) % n3 : k;
      const int js = side == 0 ? -d->ng2 : n2, je = side == 0 ? -1 : n2 + d->ng2 - 1;
      for (int j = js; j <= je; j++) {
        const int rj = side == 0 ? -j - 1 : 2 * (n2 - 1) - j + 1;
        const double* src = p + ix(d, i, rj, rk);
        double* dst = p + ix(d, i, j, k);
        for (int pl = 0; pl < NPR; pl++) dst[pl] = src[pl];
        dst[O_U2] *= -1.;
        dst[O_B2] *= -1.;
      }
    }
}

static void x2_poledeath(const o_grid* g, const dims_t* d, double* p, int side)
{
  const int n2 = d->n2, n3 = d->n3;
  const double x2c0 = g->startx[2];
#pragma omp parallel for collapse(2) num_threads(g->nthreads > 0 ? g->nthreads : 1)
  for (int k = 0; k < n3; k++)
    for (int i = -d->ng1; i < d->n1 + d->ng1; i++) {
      const int rj = side == 0 ? 1 : n2 - 2;
      const int js = side == 0 ? -1 : n2 - 1;
      const int je = side == 0 ? 0 : n2;
      const double X0 = side == 0 ? x2c0 : x2c0 + n2 * g->dx[2];
      const double Xr = x2c0 + (rj + 0.5) * g->dx[2];
      const double* pref = p + ix(d, i, rj, k);
      for (int j = js; j <= je; j++) {
        double* pj = p + ix(d, i, j, k);
        const double X = x2c0 + (j + 0.5) * g->dx[2];
        for (int pl = O_RHO; pl <= O_U3; pl++) {
          if (pl == O_U2) continue;
          pj[pl] = pref[pl];
        }
        double ftemp = pref[O_U2];
        if ((side == 0 && ftemp > 0.0) || (side == 1 && ftemp < 0.0)) ftemp = 0.0;
        pj[O_U2] = ftemp + (X - Xr) * (ftemp - 0.0) / (Xr - X0);
        if (n3 == 1) pj[O_B3] = pref[O_B3];
      }
      for (int j = js; j <= je; j++) {
        double* pj = p + ix(d, i, j, k);
        o_geom Gj;
        geom_at(g, d, 0, i, j, &Gj);
        o_state q;
        o_get_state(pj, &Gj, g->gam, &q);
        if (q.ucon[1] > 0.0) {
          const double r = g->rcent[i + d->ng1];
          double lim;
          if (r <= 10.0) lim = 1.1;
          else lim = 1.1 * pow(r / 10.0, 1.0);
          limit_gamma(lim, pj, &Gj);
        } else {
          limit_gamma(g->gammamax, pj, &Gj);
        }
      }
    }
}

static void bound_x2(const o_grid* g, const dims_t* d, double* p)
{
  if (g->problem == 1) {  /* Bondi: POLARAXIS, a plain reflection (bounds.sasha.c:170-180) */
    x2_reflect(g, d, p, 0);
    x2_reflect(g, d, p, 1);
    return;
  }
  x2_reflect(g, d, p, 0);
  x2_poledeath(g, d, p, 0);
  x2_reflect(g, d, p, 1);
  x2_poledeath(g, d, p, 1);
}

static void bound_x3(const o_grid* g, const dims_t* d, double* p)
{
  if (d->n3 == 1) return;
#pragma omp parallel for collapse(2) num_threads(g->nthreads > 0 ? g->nthreads : 1)
  for (int j = -d->ng2; j < d->n2 + d->ng2; j++)
    for (int i = -d->ng1; i < d->n1 + d->ng1; i++) {
      for (int k = -d->ng3; k < 0; k++) memcpy(p + ix(d, i, j, k), p + ix(d, i, j, k + d->n3), sizeof(double) * NPR);
      for (int k = d->n3; k < d->n3 + d->ng3; k++)
        memcpy(p + ix(d, i, j, k), p + ix(d, i, j, k - d->n3), sizeof(double) * NPR);
    }
}

void o_bound(const o_grid* g, double* p)
{
  dims_t d = D(g);
  if (d.n1 > 1) bound_x1(g, &d, p);
  if (d.n2 > 1) bound_x2(g, &d, p);
  bound_x3(g, &d, p);
}

/* ---------------------------------------------------------------------- */
/* Reconstruction                                                          */
/* ---------------------------------------------------------------------- */

/* para4 (interppoint.para.c:568) with PARA2LIM=MC, no smoothing/steepening */
static void para4(const double* y, double* lout, double* rout)
{
  double dq1[4], dq2[3], dq[3];
  /* dq1[m+1] = y[m]-y[m-1], m=-1..2 ; dq2[m+1] = (y[m+1]-y[m-1])/2, m=-1..1 */
  for (int m = -1; m <= 1; m++) {
    dq1[m + 1] = (y[m] - y[m - 1]);
    dq2[m + 1] = 0.5 * (y[m + 1] - y[m - 1]);
  }
  dq1[3] = (y[2] - y[1]);
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
  /* checkparamonotonicity, a6COEF = 4 */
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
  *lout = l;
  *rout = r;
}

static void recon_cell(int recon, const double* y, double* l, double* r)
{
  if (recon == 2) para4(y, l, r);
  else { *l = y[0]; *r = y[0]; }
}

/* ---------------------------------------------------------------------- */
/* Fluxes                                                                  */
/* ---------------------------------------------------------------------- */

static void face_flux(const o_grid* g, int dir, const o_geom* G, double* p_l, double* p_r, double* F, double* ctop)
{
  /* getplpr: BDIRCONT, then check_plpr: limit_gamma on both states */
  const int pb = O_B1 + dir - 1;
  p_l[pb] = p_r[pb] = 0.5 * (p_l[pb] + p_r[pb]);
  limit_gamma(g->gammamax, p_l, G);
  limit_gamma(g->gammamax, p_r, G);
  o_state ql, qr;
  double F_l[NPR], F_r[NPR], U_l[NPR], U_r[NPR];
  o_get_state(p_l, G, g->gam, &ql);
  o_primtoflux(p_l, &ql, dir, G, F_l);
  o_primtoflux(p_l, &ql, 0, G, U_l);
  o_get_state(p_r, G, g->gam, &qr);
  o_primtoflux(p_r, &qr, dir, G, F_r);
  o_primtoflux(p_r, &qr, 0, G, U_r);
  double cmaxl, cminl, cmaxr, cminr;
  o_vchar(p_l, &ql, dir, G, g->gam, &cmaxl, &cminl);
  o_vchar(p_r, &qr, dir, G, g->gam, &cmaxr, &cminr);
  double lmin = fmin(cminl, cminr), lmax = fmax(cmaxl, cmaxr);
  lmin = fabs(fmax(0., -lmin));
  lmax = fabs(fmax(0., lmax));
  const double cmin = lmin, cmax = lmax, ct = fmax(lmax, lmin);
  *ctop = ct;
  if (g->flux == 1 && cmax + cmin != 0.0) {
    for (int pl = 0; pl < NPR; pl++)
      F[pl] = ((cmax * F_l[pl] + cmin * F_r[pl]) - (cmax * cmin) * (U_r[pl] - U_l[pl])) / (cmax + cmin);
  } else {
    for (int pl = 0; pl < NPR; pl++) F[pl] = (0.5 * ((F_l[pl] + F_r[pl]) - ct * (U_r[pl] - U_l[pl])));
  }
}

/* Single-cell and single-face forms of the two routines above, exported for
 * the tests that pin them on known answers. */
void o_recon_cell(int recon, const double* y5, double* l, double* r) { recon_cell(recon, y5 + 2, l, r); }

void o_face_flux(int flux, int dir, double gam, double gammamax, const double* rec, double* p_l, double* p_r,
                 double* F, double* ctop)
{
  o_grid g;
  memset(&g, 0, sizeof(g));
  g.flux = flux;
  g.gam = gam;
  g.gammamax = gammamax;
  o_geom G;
  o_geom_from_rec(rec, &G);
  face_flux(&g, dir, &G, p_l, p_r, F, ctop);
}

/* The face fluxes of one direction with WENO5BND: the reference reconstructs
 * whole storage lines along dir (slope_lim_linetype_c2e interpline.c:306,
 * line bs..be = the storage row, cells -1..N reconstructed), one transverse
 * position at a time; the face loop then takes p_l from the right edge of
 * the cell before the face and p_r from the left edge of the cell after it
 * (getplpr flux.c:1952). */
static void weno_flux_dir(const o_grid* g, const dims_t* d, int dir, const double* pb, double* F, double* ndt)
{
  const int e0 = d->n1 > 1, e1 = d->n2 > 1, e2 = d->n3 > 1;
  const int nd = dir == 1 ? d->n1 : dir == 2 ? d->n2 : d->n3;
  const int ngd = dir == 1 ? d->ng1 : dir == 2 ? d->ng2 : d->ng3;
  const int n = nd + 2 * ngd;
  /* transverse positions: the face loop's range (active, one more layer
   * each side for the corner EMFs) */
  int as = 0, ae = 0, bs2 = 0, be2 = 0;
  if (dir == 1) { as = -e1; ae = d->n2 - 1 + e1; bs2 = -e2; be2 = d->n3 - 1 + e2; }
  if (dir == 2) { as = -e0; ae = d->n1 - 1 + e0; bs2 = -e2; be2 = d->n3 - 1 + e2; }
  if (dir == 3) { as = -e0; ae = d->n1 - 1 + e0; bs2 = -e1; be2 = d->n2 - 1 + e1; }
  double mindt = BIG;
#pragma omp parallel for collapse(2) reduction(min : mindt) num_threads(g->nthreads > 0 ? g->nthreads : 1)
  for (int b = bs2; b <= be2; b++)
    for (int a = as; a <= ae; a++) {
      double* y = (double*)malloc(sizeof(double) * (size_t)n * 13 + sizeof(double) * (size_t)n * NPR * 2);
      double* st = y + n;         /* 4 per cell */
      double* dP = st + 4 * n;    /* n */
      double* yl1 = dP + n;       /* n */
      double* yr1 = yl1 + n;      /* n */
      double* yl = yr1 + n + n;   /* n * NPR */
      double* yr = yl + (size_t)n * NPR;
      int ci[3] = {0, 0, 0};
      for (int c = 0; c < n; c++) {
        const int m = c - ngd;
        if (dir == 1) { ci[0] = m; ci[1] = a; ci[2] = b; }
        if (dir == 2) { ci[0] = a; ci[1] = m; ci[2] = b; }
        if (dir == 3) { ci[0] = a; ci[1] = b; ci[2] = m; }
        o_geom G;
        geom_at(g, d, 0, ci[0], ci[1], &G);
        o_dp_cell(pb + ix(d, ci[0], ci[1], ci[2]), &G, g->gam, dir, st + 4 * c);
      }
      o_dp_line(st, n, dP);
      for (int pl = 0; pl < NPR; pl++) {
        for (int c = 0; c < n; c++) {
          const int m = c - ngd;
          if (dir == 1) { ci[0] = m; ci[1] = a; ci[2] = b; }
          if (dir == 2) { ci[0] = a; ci[1] = m; ci[2] = b; }
          if (dir == 3) { ci[0] = a; ci[1] = b; ci[2] = m; }
          y[c] = pb[ix(d, ci[0], ci[1], ci[2]) + pl];
        }
        o_weno5bnd_line(y, n, ngd, dP, yl1, yr1);
        for (int c = 0; c < n; c++) {
          yl[(size_t)c * NPR + pl] = yl1[c];
          yr[(size_t)c * NPR + pl] = yr1[c];
        }
      }
      for (int f = 0; f <= nd; f++) {
        int fi[3] = {0, 0, 0};
        if (dir == 1) { fi[0] = f; fi[1] = a; fi[2] = b; }
        if (dir == 2) { fi[0] = a; fi[1] = f; fi[2] = b; }
        if (dir == 3) { fi[0] = a; fi[1] = b; fi[2] = f; }
        o_geom G;
        geom_at(g, d, dir, fi[0], fi[1], &G);
        double p_l[NPR], p_r[NPR];
        const int cl = f - 1 + ngd, cr = f + ngd; /* line index of the cells either side */
        for (int pl = 0; pl < NPR; pl++) {
          p_l[pl] = yr[(size_t)cl * NPR + pl];
          p_r[pl] = yl[(size_t)cr * NPR + pl];
        }
        double ctop;
        face_flux(g, dir, &G, p_l, p_r, F + ix(d, fi[0], fi[1], fi[2]), &ctop);
        const double dtij = g->cour * g->dx[dir] / ctop;
        if (dtij < mindt) mindt = dtij;
      }
      free(y);
    }
  *ndt = mindt;
}

static void flux_dir(const o_grid* g, const dims_t* d, int dir, const double* pb, double* F, double* ndt)
{
  const int e1 = d->n2 > 1, e2 = d->n3 > 1, e0 = d->n1 > 1;
  int is = 0, ie = d->n1 - 1, js = 0, je = d->n2 - 1, ks = 0, ke = d->n3 - 1;
  if (dir == 1) { ie = d->n1; js -= e1; je += e1; ks -= e2; ke += e2; }
  if (dir == 2) { je = d->n2; is -= e0; ie += e0; ks -= e2; ke += e2; }
  if (dir == 3) { ke = d->n3; is -= e0; ie += e0; js -= e1; je += e1; }
  const int di = dir == 1, dj = dir == 2, dk = dir == 3;
  const long st = (dir == 1) ? NPR : (dir == 2) ? (long)d->ns1 * NPR : (long)d->ns1 * d->ns2 * NPR;
  double mindt = BIG;
  if (g->recon == 3) {
    weno_flux_dir(g, d, dir, pb, F, ndt);
    return;
  }
#pragma omp parallel for collapse(3) reduction(min : mindt) num_threads(g->nthreads > 0 ? g->nthreads : 1)
  for (int k = ks; k <= ke; k++)
    for (int j = js; j <= je; j++)
      for (int i = is; i <= ie; i++) {
        o_geom G;
        geom_at(g, d, dir, i, j, &G);
        const double* c = pb + ix(d, i, j, k);          /* cell on the right of the face */
        const double* cm = pb + ix(d, i - di, j - dj, k - dk); /* cell on the left */
        double p_l[NPR], p_r[NPR];
        for (int pl = 0; pl < NPR; pl++) {
          double y[5], l, r;
          for (int m = -2; m <= 2; m++) y[m + 2] = cm[pl + m * st];
          recon_cell(g->recon, y + 2, &l, &r);
          p_l[pl] = r; /* right edge of the left cell */
          for (int m = -2; m <= 2; m++) y[m + 2] = c[pl + m * st];
          recon_cell(g->recon, y + 2, &l, &r);
          p_r[pl] = l; /* left edge of the right cell */
        }
        double ctop;
        double* Fo = F + ix(d, i, j, k);
        face_flux(g, dir, &G, p_l, p_r, Fo, &ctop);
        const double dtij = g->cour * g->dx[dir] / ctop;
        if (dtij < mindt) mindt = dtij;
      }
  *ndt = mindt;
}

/* flux_ct (fluxct.c:186) with FLUXB=FLUXCTTOTH, CORNGDETVERSION=1 */
static void flux_ct(const o_grid* g, const dims_t* d, double* F1, double* F2, double* F3)
{
  const int N1 = d->n1, N2 = d->n2, N3 = d->n3;
  const int e0 = N1 > 1, e1 = N2 > 1, e2 = N3 > 1;
  double* Fs[4] = {NULL, F1, F2, F3};
  /* divide face B fluxes by the face gdet over all computed faces */
  for (int dir = 1; dir <= 3; dir++) {
    if ((dir == 1 && !e0) || (dir == 2 && !e1) || (dir == 3 && !e2)) continue;
    int is = 0, ie = N1 - 1, js = 0, je = N2 - 1, ks = 0, ke = N3 - 1;
    if (dir == 1) { ie = N1; js -= e1; je += e1; ks -= e2; ke += e2; }
    if (dir == 2) { je = N2; is -= e0; ie += e0; ks -= e2; ke += e2; }
    if (dir == 3) { ke = N3; is -= e0; ie += e0; js -= e1; je += e1; }
    for (int k = ks; k <= ke; k++)
      for (int j = js; j <= je; j++)
        for (int i = is; i <= ie; i++) {
          const double ign = igdetnosing(g->geom[dir][ig2(d, i, j) * 27 + 21]);
          double* F = Fs[dir] + ix(d, i, j, k);
          F[O_B1] *= ign;
          F[O_B2] *= ign;
          F[O_B3] *= ign;
        }
  }
  const double c1 = (e1 && e2) ? 0.25 : 0.5, c2 = (e0 && e2) ? 0.25 : 0.5, c3 = (e0 && e1) ? 0.25 : 0.5;
  const int ni = N1 + e0, nj = N2 + e1, nk = N3 + e2;
  double* emf = (double*)malloc(sizeof(double) * 3 * (size_t)ni * nj * nk);
#define EMF(c, i, j, k) emf[(((size_t)(c) * nk + (k)) * nj + (j)) * ni + (i)]
  for (int k = 0; k < nk; k++)
    for (int j = 0; j < nj; j++)
      for (int i = 0; i < ni; i++) {
        const int km = k - e2, jm = j - e1, im = i - e0;
        double e;
        if (e1 || e2) {
          if (e1 && e2) e = c1 * ((+F2[ix(d, i, j, k) + O_B3] + F2[ix(d, i, j, km) + O_B3]) - F3[ix(d, i, j, k) + O_B2] - F3[ix(d, i, jm, k) + O_B2]);
          else if (e1) e = c1 * (+F2[ix(d, i, j, k) + O_B3] + F2[ix(d, i, j, km) + O_B3]);
          else e = c1 * (-F3[ix(d, i, j, k) + O_B2] - F3[ix(d, i, jm, k) + O_B2]);
        } else e = 0.0;
        EMF(0, i, j, k) = e * g->geom[2][ig2(d, i, j) * 27 + 21]; /* CORN1 gdet = FACE2 gdet */
        if (e0 || e2) {
          if (e0 && e2) e = c2 * ((+F3[ix(d, i, j, k) + O_B1] + F3[ix(d, im, j, k) + O_B1]) - F1[ix(d, i, j, k) + O_B3] - F1[ix(d, i, j, km) + O_B3]);
          else if (e0) e = c2 * (-F1[ix(d, i, j, k) + O_B3] - F1[ix(d, i, j, km) + O_B3]);
          else e = c2 * (+F3[ix(d, i, j, k) + O_B1] + F3[ix(d, im, j, k) + O_B1]);
        } else e = 0.0;
        EMF(1, i, j, k) = e * g->geom[1][ig2(d, i, j) * 27 + 21]; /* CORN2 gdet = FACE1 gdet */
        if (e0 || e1) {
          if (e0 && e1) e = c3 * ((+F1[ix(d, i, j, k) + O_B2] + F1[ix(d, i, jm, k) + O_B2]) - F2[ix(d, i, j, k) + O_B1] - F2[ix(d, im, j, k) + O_B1]);
          else if (e0) e = c3 * (+F1[ix(d, i, j, k) + O_B2] + F1[ix(d, i, jm, k) + O_B2]);
          else e = c3 * (-F2[ix(d, i, j, k) + O_B1] - F2[ix(d, im, j, k) + O_B1]);
        } else e = 0.0;
        EMF(2, i, j, k) = e * g->gdet_corn3[ig2(d, i, j)];
      }
  /* rewrite face fluxes on active faces */
  for (int k = 0; k < N3; k++)
    for (int j = 0; j < N2; j++)
      for (int i = 0; i <= N1 - 1 + e0; i++) {
        if (!e0) break;
        double* F = F1 + ix(d, i, j, k);
        F[O_B1] = 0.;
        F[O_B2] = 0.5 * (EMF(2, i, j, k) + EMF(2, i, j + e1, k));
        F[O_B3] = -0.5 * (EMF(1, i, j, k) + EMF(1, i, j, k + e2));
      }
  for (int k = 0; k < N3; k++)
    for (int j = 0; j <= N2 - 1 + e1; j++)
      for (int i = 0; i < N1; i++) {
        if (!e1) break;
        double* F = F2 + ix(d, i, j, k);
        F[O_B1] = -0.5 * (EMF(2, i, j, k) + EMF(2, i + e0, j, k));
        F[O_B2] = 0.;
        F[O_B3] = 0.5 * (EMF(0, i, j, k) + EMF(0, i, j, k + e2));
      }
  for (int k = 0; k <= N3 - 1 + e2; k++)
    for (int j = 0; j < N2; j++)
      for (int i = 0; i < N1; i++) {
        if (!e2) break;
        double* F = F3 + ix(d, i, j, k);
        F[O_B1] = 0.5 * (EMF(1, i, j, k) + EMF(1, i + e0, j, k));
        F[O_B2] = -0.5 * (EMF(0, i, j, k) + EMF(0, i, j + e1, k));
        F[O_B3] = 0.;
      }
#undef EMF
  free(emf);
}

void o_fluxcalc(const o_grid* g, const double* pb, double* F1, double* F2, double* F3, double* ndt, int do_ct)
{
  dims_t d = D(g);
  ndt[1] = ndt[2] = ndt[3] = BIG;
  if (d.n1 > 1) flux_dir(g, &d, 1, pb, F1, &ndt[1]);
  if (d.n2 > 1) flux_dir(g, &d, 2, pb, F2, &ndt[2]);
  if (d.n3 > 1) flux_dir(g, &d, 3, pb, F3, &ndt[3]);
  if (do_ct) flux_ct(g, &d, F1, F2, F3);
}

/* ---------------------------------------------------------------------- */
/* Update                                                                  */
/* ---------------------------------------------------------------------- */

/* dUtodt -> compute_dt_fromsource (advance.c:2581, :2331; the dUevolve
 * arrays sized NPR, which is the intended reading of the indexing) */
static double source_dt(const o_grid* g, const dims_t* d, const o_geom* G, const double* pr, const o_state* q,
                        const double* dUtot, double dt)
{
  double U[NPR], Ugf[NPR];
  o_primtoflux(pr, q, 0, G, U);
  const double ig = igdetnosing(G->gdet);
  for (int pl = 0; pl < NPR; pl++) Ugf[pl] = U[pl] * ig;
  double dUd[4], dUu[4], mydU[4], mydx[4];
  for (int jj = 0; jj < 4; jj++) dUd[jj] = (dUtot[O_UU + jj] * dt) * ig;
  for (int a = 0; a < 4; a++) dUu[a] = G->gcon[a][0] * dUd[0] + G->gcon[a][1] * dUd[1] + G->gcon[a][2] * dUd[2] + G->gcon[a][3] * dUd[3];
  mydU[0] = dUd[0];
  mydx[0] = dt;
  for (int jj = 1; jj < 4; jj++) { mydU[jj] = dUu[jj]; mydx[jj] = g->dx[jj]; }
  /* b^mu from primitive B (bcon_calc), b^2 = b.b */
  double bcon[4], bcov[4];
  bcon[0] = pr[O_B1] * q->ucov[1] + pr[O_B2] * q->ucov[2] + pr[O_B3] * q->ucov[3];
  for (int j = 1; j < 4; j++) bcon[j] = (pr[O_B1 - 1 + j] + bcon[0] * q->ucon[j]) / q->ucon[0];
  for (int a = 0; a < 4; a++) bcov[a] = G->gcov[a][0] * bcon[0] + G->gcov[a][1] * bcon[1] + G->gcov[a][2] * bcon[2] + G->gcov[a][3] * bcon[3];
  const double bsq = bcon[0] * bcov[0] + bcon[1] * bcov[1] + bcon[2] * bcov[2] + bcon[3] * bcov[3];
  const double rho = pr[O_RHO], u = pr[O_UU];
  const double P = (g->gam - 1.0) * u;
  const double w = rho + u + P;
  const double eta = w + bsq;
  double dtsource[4];
  for (int jj = 0; jj < 4; jj++) {
    const double rhoprime = fmax(fabs(eta), fabs(Ugf[O_UU]));
    const double ag = SMALL + fabs(mydU[jj] / rhoprime);
    dtsource[jj] = g->cour * (mydx[jj] / ag);
  }
  double dtij = BIG;
  if (dtsource[0] < dtij) dtij = dtsource[0];
  if (d->n1 > 1 && dtsource[1] < dtij) dtij = dtsource[1];
  if (d->n2 > 1 && dtsource[2] < dtij) dtij = dtsource[2];
  if (d->n3 > 1 && dtsource[3] < dtij) dtij = dtsource[3];
  return dtij;
}

void o_update(const o_grid* g, const double* pi, const double* pb, const double* F1, const double* F2,
              const double* F3, double* ui, double* uf, double* ucum, const double* CUf, const double* Cunew,
              double dt, int timeorder, double* accdt)
{
  dims_t d = D(g);
  const double idx1 = 1.0 / g->dx[1], idx2 = 1.0 / g->dx[2], idx3 = 1.0 / g->dx[3];
  const long s1 = NPR, s2 = (long)d.ns1 * NPR, s3 = (long)d.ns1 * d.ns2 * NPR;
  double mindt = BIG;
#pragma omp parallel for collapse(3) reduction(min : mindt) num_threads(g->nthreads > 0 ? g->nthreads : 1)
  for (int k = 0; k < d.n3; k++)
    for (int j = 0; j < d.n2; j++)
      for (int i = 0; i < d.n1; i++) {
        const long o = ix(&d, i, j, k);
        o_geom G;
        geom_at(g, &d, 0, i, j, &G);
        o_state q;
        double Ui[NPR], dU1[NPR], dU2[NPR], dU3[NPR], dUr[NPR], dUg[NPR];
        o_get_state(pi + o, &G, g->gam, &q);
        o_primtoflux(pi + o, &q, 0, &G, Ui);
        for (int pl = 0; pl < NPR; pl++) ui[o + pl] = Ui[pl];
        for (int pl = 0; pl < NPR; pl++) {
          dU1[pl] = 0; dU2[pl] = 0; dU3[pl] = 0;
          if (d.n1 > 1) dU1[pl] = (-(F1[o + s1 + pl] - F1[o + pl]) * idx1);
          if (d.n2 > 1) dU2[pl] = (-(F2[o + s2 + pl] - F2[o + pl]) * idx2);
          if (d.n3 > 1) dU3[pl] = (-(F3[o + s3 + pl] - F3[o + pl]) * idx3);
          dUr[pl] = dU1[pl] + dU2[pl] + dU3[pl];
        }
        o_state qb;
        o_get_state(pb + o, &G, g->gam, &qb);
        o_source(pb + o, &G, &qb, g->conn + ig2(&d, i, j) * 40, dUg);
        if (timeorder == 0)
          for (int pl = 0; pl < NPR; pl++) { uf[o + pl] = 0.0; ucum[o + pl] = 0.0; }
        for (int pl = 0; pl < NPR; pl++)
          uf[o + pl] = (CUf[0] * Ui[pl] + CUf[1] * uf[o + pl] + CUf[2] * dt * (dUr[pl] + dUg[pl]));
        for (int pl = 0; pl < NPR; pl++)
          ucum[o + pl] += (Cunew[0] * Ui[pl] + Cunew[1] * dt * (dUr[pl] + dUg[pl]) + Cunew[2] * uf[o + pl]);
        double dUt[NPR];
        for (int pl = 0; pl < NPR; pl++) dUt[pl] = dUr[pl] + dUg[pl];
        const double dts = source_dt(g, &d, &G, pb + o, &qb, dUt, dt);
        if (dts < mindt) mindt = dts;
      }
  *accdt = mindt;
}

/* ---------------------------------------------------------------------- */
/* Inversion, floors, failure averaging                                    */
/* ---------------------------------------------------------------------- */

/* fixup1zone (fixup.c:561), FIXUPTYPE=1, rescaletype 0 floors; with fl (the
 * fladdterms block), the floor step (fixup.c:766, COUNTFLOORACT) and the
 * gamma limit (fixup.c:784, COUNTLIMITGAMMAACT) are each accounted from pr0,
 * the primitives on entry (fixup.c:593). */
static void fixup1zone(const o_grid* g, const o_geom* G, double r, double* pr, int32_t* pflag, int* nstroke,
                       double* fl)
{
  o_params P = {g->gam, g->gammafail, g->gammamax, g->hot2cold, g->checkinv};
  double prfloor[2], prnew[NPR], pr0[NPR];
  for (int pl = 0; pl < NPR; pl++) pr0[pl] = pr[pl];
  prfloor[0] = g->rhomin * pow(r, -1.5);
  prfloor[1] = g->uumin * pow(r, -2.5);
  for (int pl = 0; pl < 2; pl++)
    if (prfloor[pl] < SMALL) prfloor[pl] = SMALL;
  int didchange = 0;
  for (int pl = 0; pl < NPR; pl++) {
    if (pl <= O_UU && prfloor[pl] > pr[pl]) { didchange = 1; prnew[pl] = prfloor[pl]; }
    else prnew[pl] = pr[pl];
  }
  if (didchange) {
    o_state q, dq;
    o_geom G1 = *G;
    G1.gdet = 1.0; /* primtoU(UNOTHING): no gdet factor */
    double U[NPR], dpr[NPR], dU[NPR];
    o_get_state(pr, G, g->gam, &q);
    o_primtoflux(pr, &q, 0, &G1, U);
    for (int pl = 0; pl < NPR; pl++) dpr[pl] = 0.0;
    for (int pl = O_RHO; pl <= O_UU; pl++) dpr[pl] = prnew[pl] - pr[pl];
    o_get_state(dpr, G, g->gam, &dq);
    o_primtoflux(dpr, &dq, 0, &G1, dU);
    for (int pl = 0; pl < NPR; pl++) U[pl] += dU[pl];
    const int flag = o_utoprimgen_unothing(U, G, &P, pr, nstroke);
    *pflag = flag;
    if (flag > 0)
      for (int pl = 0; pl < NPR; pl++) pr[pl] = prnew[pl];
    if (fl) diag_fixup_acc(g, G, pr0, pr, fl + O_FL_FLOOR * NPR);
  }
  if (limit_gamma(g->gammamax, pr, G) && fl) diag_fixup_acc(g, G, pr0, pr, fl + O_FL_GAMMA * NPR);
}

/* fl (final substep): the fladdterms block, or NULL */
static void invert(const o_grid* g, const double* pb, const double* U, double* pf, int32_t* pflag, int finalstep,
                   int64_t* nstroke, double* fl)
{
  dims_t d = D(g);
  const long ntot = (long)d.ns1 * d.ns2 * d.ns3 * NPR;
  memcpy(pf, pb, sizeof(double) * ntot);
  int64_t ns = 0;
  o_params P = {g->gam, g->gammafail, g->gammamax, g->hot2cold, g->checkinv};
  double fls[3 * NPR] = {0};  /* categories 0..2: only FLOOR and GAMMA are reached here */
  const int acc = fl != NULL;
#pragma omp parallel for collapse(3) reduction(+ : ns, fls[:3 * NPR]) schedule(dynamic, 64) num_threads(g->nthreads > 0 ? g->nthreads : 1)
  for (int k = 0; k < d.n3; k++)
    for (int j = 0; j < d.n2; j++)
      for (int i = 0; i < d.n1; i++) {
        const long o = ix(&d, i, j, k);
        const long f = ((long)(k + d.ng3) * d.ns2 + (j + d.ng2)) * d.ns1 + (i + d.ng1);
        o_geom G;
        geom_at(g, &d, 0, i, j, &G);
        int nst = 0;
        pflag[f] = o_utoprimgen(U + o, &G, &P, pf + o, &nst);
        if (finalstep) fixup1zone(g, &G, g->rcent[i + d.ng1], pf + o, &pflag[f], &nst, acc ? fls : NULL);
        ns += nst;
      }
  if (nstroke) *nstroke = ns;
  if (acc)
    for (int c = 0; c < 3 * NPR; c++) fl[c] += fls[c];
  if (finalstep && g->failcount)
    for (int k = 0; k < d.n3; k++)
      for (int j = 0; j < d.n2; j++)
        for (int i = 0; i < d.n1; i++) {
          const int f = pflag[((long)(k + d.ng3) * d.ns2 + (j + d.ng2)) * d.ns1 + (i + d.ng1)];
          if (f > 0) g->failcount[0]++;
          if (f > 0 && (f < 9 || f > 11) && f != O_FAIL_U2AVG1 && f != O_FAIL_U2AVG2) g->failcount[1]++;
        }
}

void o_invert(const o_grid* g, const double* pb, const double* U, double* pf, int32_t* pflag, int finalstep,
              int64_t* nstroke)
{
  invert(g, pb, U, pf, pflag, finalstep, nstroke, NULL);
}

/* bound_pflag (step_ch.c:1637 -> bound_pflag_user, boundsint.c:18-330):
 * one ghost layer of flags (numbnd = 1 for BOUNDINTTYPE,
 * initbase.boundloop.c:83-88), filled in the reference's order:
 *  1. phi periodic (X3 PERIODIC, :80-92) over i and j in -1..n, which reads
 *     the i and j ghost cells as the previous call left them;
 *  2. the outer radial OUTFLOW copy of the last zone (:110-121) over active
 *     j and k (LOOPX1dir covers the active range, initbase.boundloop.c:
 *     358-363): the torus's X1UP (init.fishmon.c:247).  The inner radial
 *     FREEOUTFLOW (init.fishmon.c:248) and both BONDI*OUTFLOW boundaries
 *     (init.sasha.c:1503-1504) are not handled there, so those ghost flags
 *     keep what they held: NANPFLAG from set_arrays_multidimen.c:404;
 *  3. the polar mirror (POLARAXIS, :178-200) over i in -1..n1, active k. */
void o_bound_pflag(const o_grid* g, int32_t* pflag)
{
  dims_t d = D(g);
#define PF(i, j, k) pflag[((long)((k) + d.ng3) * d.ns2 + ((j) + d.ng2)) * d.ns1 + ((i) + d.ng1)]
  const int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  if (s3)
    for (int j = -s2; j < d.n2 + s2; j++)
      for (int i = -s1; i < d.n1 + s1; i++) {
        PF(i, j, -1) = PF(i, j, d.n3 - 1);
        PF(i, j, d.n3) = PF(i, j, 0);
      }
  if (s1 && g->problem == 0)
    for (int k = 0; k < d.n3; k++)
      for (int j = 0; j < d.n2; j++) PF(d.n1, j, k) = PF(d.n1 - 1, j, k);
  if (s2)
    for (int k = 0; k < d.n3; k++)
      for (int i = -s1; i < d.n1 + s1; i++) {
        PF(i, -1, k) = PF(i, 0, k);
        PF(i, d.n2, k) = PF(i, d.n2 - 1, k);
      }
#undef PF
}

/* IFUTOPRIMNOFAILORFIXED (global.nondepnmemonics.h:854) */
static inline int good_flag(int f) { return f >= O_FAIL_FIXEDCOLD && f <= 0; }

/* the fix-up category of a failure flag (fixuputoprim_accounting,
 * fixup.c:1605-1745); the soft failures are counted on the final substep only
 * (DOCOUNTNEG* = 1, STEPOVERNEG* = NEGDENSITY_NEVERFIXUP), the convergence
 * failures on every substep; -1 when not accounted */
static int fail_category(int f, int finalstep)
{
  if (f == 9) return finalstep ? O_FL_RHONEG : -1;
  if (f == 10 || f == O_FAIL_U2AVG1 || f == O_FAIL_U2AVG2) return finalstep ? O_FL_UNEG : -1;
  if (f == 11) return finalstep ? O_FL_RHOUNEG : -1;
  return O_FL_CONV;
}

/* fixup_utoprim (fixup.c:1061) as the torus compiles it: UTOPRIMADJUST =
 * UTOPRIMAVG, GENERALAVERAGE 1 (:1035), HANDLE*NEG 0 (:1042-1055; so
 * fixup_negdensities changes nothing, :1457), CAUSALAVG 0 (:1817),
 * HOWTOAVG_WHEN_U2AVG = MAXUPOSAVG (:1827), TODONOGOOD = NOGOODAVERAGE
 * (:2278) with USERRESETREGION 1 (init.fishmon.h:356; off for the Bondi
 * problem, whose header does not set it).  The flags are bounded first
 * (o_bound_pflag); averages read a copy of pf (ptoavg) and of the flags
 * (pflagfailorig).  Each failed zone:
 *  - u-only failures (UTOPRIMFAILU2AVG1/2, the cold retry's mark) take u
 *    from general_average's second loop: the smallest u over the symmetric
 *    neighbour pairs whose two u's are both at least the zone's own (:1990);
 *  - every other failure takes rho..U3 as the mean of the good neighbours in
 *    the 3^d box, summed pair by pair into two alternating sums (:1960-2030);
 *  - with no usable neighbour, fixup_nogood (:2291): rho..U3 (u alone for
 *    u-only failures) become the mean of the 4+4 in-plane neighbours
 *    (AVG4_1, AVG4_2 at :1812-1813), failed or not; in the reset region
 *    (j <= 1, j >= n2 - 2 or i < 10) rho and u do and the velocity is set to
 *    the normal observer's (set_atmosphere, init.tools.c:797, VELREL4: 0);
 *  - after a velocity average the Lorentz factor is capped at the zone's own
 *    pre-average value (itself capped at GAMMAMAX) (:1290-1322);
 *  - accounted zones get UTOPRIMFAILFIXEDUTOPRIM (:1761).
 * fl (final substep, the fladdterms block): each accounted zone's change from
 * its value before the averaging, in its category (diag_fixup), or NULL. */
static void post_fixup(const o_grid* g, double* pf, int32_t* pflag, int finalstep, double* fl)
{
  dims_t d = D(g);
  o_bound_pflag(g, pflag);
  const long ntot = (long)d.ns1 * d.ns2 * d.ns3;
  double* pa = (double*)malloc(sizeof(double) * ntot * NPR);
  int32_t* fo = (int32_t*)malloc(sizeof(int32_t) * ntot);
  memcpy(pa, pf, sizeof(double) * ntot * NPR);
  memcpy(fo, pflag, sizeof(int32_t) * ntot);
  const int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  const int rnx = s1 ? 3 : 1, rny = s2 ? 3 : 1;
  int npow = 1;
  for (int q = 0; q < s1 + s2 + s3; q++) npow *= 3;
  const int numupairs = (npow - 1) / 2;
#define FO(i, j, k) fo[((long)((k) + d.ng3) * d.ns2 + ((j) + d.ng2)) * d.ns1 + ((i) + d.ng1)]
#define A(i, j, k, pl) pa[ix(&d, (i), (j), (k)) + (pl)]
  for (int k = 0; k < d.n3; k++)
    for (int j = 0; j < d.n2; j++)
      for (int i = 0; i < d.n1; i++) {
        const int f = FO(i, j, k);
        int32_t* pfl = &pflag[((long)(k + d.ng3) * d.ns2 + (j + d.ng2)) * d.ns1 + (i + d.ng1)];
        if (f == O_FAIL_FIXEDCOLD || f == -2) { *pfl = -1; continue; } /* accounting of no change */
        if (f <= 0) continue;
        /* u alone for UTOPRIMFAILU2AVG1/2 and UTOPRIMFAILUPERC (:1207); the
         * min-of-pairs loop for the U2AVG flags only (:1854-1860) */
        const int urange = (f == O_FAIL_U2AVG1 || f == O_FAIL_U2AVG2 || f == 13);
        const int uonly = (f == O_FAIL_U2AVG1 || f == O_FAIL_U2AVG2);
        const int startpl = urange ? O_UU : O_RHO, endpl = urange ? O_UU : O_U3;
        double* p = pf + ix(&d, i, j, k);
        /* general_average */
        double mysum[2][NPR], lastmin[NPR];
        for (int pl = 0; pl < NPR; pl++) { mysum[0][pl] = 0.0; mysum[1][pl] = 0.0; lastmin[pl] = 1E50; }
        int numavg0 = 0, numavg1 = 0;
        const double ref = A(i, j, k, O_UU);
        for (int qq = 0; qq < numupairs; qq++) {
          const int ii = qq % rnx - s1, jj = (qq % (rnx * rny)) / rnx - s2, kk = qq / (rnx * rny) - s3;
          const int th = good_flag(FO(i + ii, j + jj, k + kk)), at = good_flag(FO(i - ii, j - jj, k - kk));
          numavg0 += th + at;
          if (!uonly) {
            for (int pl = startpl; pl <= endpl; pl++)
              mysum[qq % 2][pl] += A(i + ii, j + jj, k + kk, pl) * th + A(i - ii, j - jj, k - kk, pl) * at;
          } else {
            for (int pl = startpl; pl <= endpl; pl++) {
              const double f1 = A(i + ii, j + jj, k + kk, pl), f2 = A(i - ii, j - jj, k - kk, pl);
              if (f1 >= ref && f2 >= ref) {
                const double m1 = lastmin[pl] < f1 ? lastmin[pl] : f1;
                lastmin[pl] = m1 < f2 ? m1 : f2;
                numavg1++;
              }
            }
          }
        }
        int numavg;
        if (!uonly) {
          if (numavg0 != 0)
            for (int pl = startpl; pl <= endpl; pl++) p[pl] = (mysum[0][pl] + mysum[1][pl]) / ((double)numavg0);
          numavg = numavg0;
        } else {
          if (numavg1 != 0)
            for (int pl = startpl; pl <= endpl; pl++) p[pl] = lastmin[pl];
          numavg = numavg1;
        }
        if (f == O_FAIL_U2AVG2) numavg++;
        o_geom G;
        geom_at(g, &d, 0, i, j, &G);
        if (numavg == 0) {
          /* fixup_nogood */
          const int ip = i + s1, im = i - s1, jp = j + s2, jm = j - s2;
#define AVG8(pl) (0.5 * ((0.25 * ((A(i, jp, k, pl) + A(i, jm, k, pl)) + (A(im, j, k, pl) + A(ip, j, k, pl)))) + \
                         (0.25 * ((A(ip, jp, k, pl) + A(ip, jm, k, pl)) + (A(im, jp, k, pl) + A(im, jm, k, pl))))))
          const int reset = g->resetregion && ((j <= 1 || j >= d.n2 - 2) || (i < 10));
          if (reset) {
            for (int pl = O_RHO; pl <= O_UU; pl++) p[pl] = AVG8(pl);
            for (int pl = O_U1; pl <= O_U3; pl++) p[pl] = 0.0;
          } else {
            for (int pl = startpl; pl <= endpl; pl++) p[pl] = AVG8(pl);
          }
#undef AVG8
        }
        if (startpl <= O_U1 && endpl >= O_U3) {
          double gamma = gamma_of(pa + ix(&d, i, j, k), &G);
          if (gamma > g->gammamax) gamma = g->gammamax;
          limit_gamma(gamma, p, &G);
        }
        const int cat = fail_category(f, finalstep);
        if (cat >= 0) {
          if (fl) diag_fixup_acc(g, &G, pa + ix(&d, i, j, k), p, fl + cat * NPR);
          *pfl = -1;
        }
      }
#undef A
#undef FO
  free(fo);
  free(pa);
}

void o_post_fixup(const o_grid* g, double* pf, int32_t* pflag, int finalstep)
{
  post_fixup(g, pf, pflag, finalstep, NULL);
}

/* ---------------------------------------------------------------------- */
/* Step driver                                                             */
/* ---------------------------------------------------------------------- */

int o_rk_tables(int timeorder, double CUf[4][4], double Cunew[4][4])
{
  memset(CUf, 0, sizeof(double) * 16);
  memset(Cunew, 0, sizeof(double) * 16);
  if (timeorder == 3) {
    CUf[0][0] = 1.0; CUf[0][1] = 0.0; CUf[0][2] = 1.0;
    CUf[1][0] = 3.0 / 4.0; CUf[1][1] = 1.0 / 4.0; CUf[1][2] = 1.0 / 4.0;
    CUf[2][0] = 1.0 / 3.0; CUf[2][1] = 2.0 / 3.0; CUf[2][2] = 2.0 / 3.0;
    Cunew[1][3] = 1.0;
    Cunew[2][2] = 1.0; Cunew[2][3] = 1.0 / 4.0;
    return 3;
  }
  CUf[0][0] = 1.0; CUf[0][1] = 0.0; CUf[0][2] = 0.5;
  CUf[1][0] = 1.0; CUf[1][1] = 0.0; CUf[1][2] = 1.0;
  Cunew[1][2] = 1.0; Cunew[1][3] = 0.5;
  return 2;
}

/* effective weight of substep s's flux in the final state */
static void rk_flux_weights(int timeorder, double* w)
{
  double CUf[4][4], Cunew[4][4];
  const int n = o_rk_tables(timeorder, CUf, Cunew);
  double C[4][4] = {{0}};
  for (int s = 0; s < n; s++) {
    for (int r = 0; r < s; r++) C[s][r] = CUf[s][1] * C[s - 1][r];
    C[s][s] = CUf[s][2];
  }
  for (int r = 0; r < n; r++) w[r] = C[n - 1][r];
}

/* ---------------------------------------------------------------------- */
/* Energy-file diagnostics                                                 */
/* ---------------------------------------------------------------------- */

/* diag_flux_pureflux (diag.c:1020-1210): per boundary X1UP, X1DN, X2UP,
 * X2DN, X3UP, X3DN (global.nondepnmemonics.h:699-704) the sum over its faces of F * surface (non-finite terms
 * dropped, diag.c:1196-1201), and pcum += pdot * Dt. */
static void ener_pdot(const o_grid* g, const dims_t* d, const double* F1, const double* F2, const double* F3,
                      double Dt, double* en)
{
  const int nn[4] = {0, d->n1, d->n2, d->n3};
  for (int bd = 0; bd < 6; bd++) {
    const int dir = bd / 2 + 1, up = !(bd & 1);
    double* pdot = en + O_EN_PDOT + bd * NPR;
    for (int pl = 0; pl < NPR; pl++) pdot[pl] = 0.0;
    if (nn[dir] > 1) {
      const double* F = dir == 1 ? F1 : dir == 2 ? F2 : F3;
      const double surface = dir == 1 ? g->dx[2] * g->dx[3] : dir == 2 ? g->dx[1] * g->dx[3] : g->dx[1] * g->dx[2];
      const int is = dir == 1 ? (up ? d->n1 : 0) : 0, ie = dir == 1 ? is : d->n1 - 1;
      const int js = dir == 2 ? (up ? d->n2 : 0) : 0, je = dir == 2 ? js : d->n2 - 1;
      const int ks = dir == 3 ? (up ? d->n3 : 0) : 0, ke = dir == 3 ? ks : d->n3 - 1;
      for (int k = ks; k <= ke; k++)
        for (int j = js; j <= je; j++)
          for (int i = is; i <= ie; i++)
            for (int pl = 0; pl < NPR; pl++) {
              double Ftemp = F[ix(d, i, j, k) + pl] * surface;
              if (!isfinite(Ftemp)) Ftemp = 0.0;
              pdot[pl] += Ftemp;
            }
    }
    for (int pl = 0; pl < NPR; pl++) en[O_EN_PCUM + bd * NPR + pl] += pdot[pl] * Dt;
  }
}

/* dt4diag (advance.c:251-262, DIAGSOURCECOMPSTEP = 4) for this substep */
static double ener_dt4(double* en, double dt, int last)
{
  const long nstep = (long)en[O_EN_NSTEP];
  if (last && nstep % 4 == 0) {
    en[O_EN_DT4] = en[O_EN_WILLBE];
    en[O_EN_WILLBE] = 0.0;
  } else {
    en[O_EN_WILLBE] += dt;
    en[O_EN_DT4] = -1.0;
  }
  return en[O_EN_DT4];
}

/* diag_source_all (diag.c:1958) over the active zones: sourceadd +=
 * Dt * dVF * dUgeom(pb) when Dt > 0 */
static void ener_source(const o_grid* g, const dims_t* d, const double* pb, double Dt, double* en)
{
  if (!(Dt > 0.0)) return;
  const double dVF = g->dx[1] * g->dx[2] * g->dx[3];
  for (int k = 0; k < d->n3; k++)
    for (int j = 0; j < d->n2; j++)
      for (int i = 0; i < d->n1; i++) {
        o_geom G;
        geom_at(g, d, 0, i, j, &G);
        o_state q;
        double dUg[NPR];
        o_get_state(pb + ix(d, i, j, k), &G, g->gam, &q);
        o_source(pb + ix(d, i, j, k), &G, &q, g->conn + ig2(d, i, j) * 40, dUg);
        for (int pl = 0; pl < NPR; pl++) en[O_EN_SRCADD + pl] += Dt * dVF * dUg[pl];
      }
}

/* FLUXCTTOTH corner divergence (SETFDIVBFLUXCTTOTHPRIM,
 * global.fieldmacros.h:130-244, 383): Fg = B * gdet(CENT), averaged over
 * the transverse pairs, differenced along each direction and normalised by
 * the same averages of |Fg|.  im1mac and friends keep the index in a
 * direction with one zone. */
static double divb_at(const o_grid* g, const dims_t* d, const double* p, int i, int j, int k)
{
#define GD(ii, jj) (g->geom[0][ig2(d, ii, jj) * 27 + 21])
#define FG(ii, jj, kk, pl) (p[ix(d, ii, jj, kk) + (pl)] * GD(ii, jj))
#define AFG(ii, jj, kk, pl) fabs(p[ix(d, ii, jj, kk) + (pl)] * GD(ii, jj))
#define IM(x) (d->n1 > 1 ? (x) - 1 : (x))
#define JM(x) (d->n2 > 1 ? (x) - 1 : (x))
#define KM(x) (d->n3 > 1 ? (x) - 1 : (x))
#define AV1(F, ii, jj, kk, pl) (0.5 * (F(ii, jj, kk, pl) + F(IM(ii), jj, kk, pl)))
#define AV2(F, ii, jj, kk, pl) (0.5 * (F(ii, jj, kk, pl) + F(ii, JM(jj), kk, pl)))
#define FOR1(F, ii, jj, kk) (0.5 * (AV2(F, ii, jj, kk, O_B1) + AV2(F, ii, jj, KM(kk), O_B1)))
#define FOR2(F, ii, jj, kk) (0.5 * (AV1(F, ii, jj, kk, O_B2) + AV1(F, ii, jj, KM(kk), O_B2)))
#define FOR3(F, ii, jj, kk) (0.5 * (AV2(F, ii, jj, kk, O_B3) + AV2(F, IM(ii), jj, kk, O_B3)))
  const double diffx = FOR1(FG, i, j, k) - FOR1(FG, IM(i), j, k);
  const double diffy = FOR2(FG, i, j, k) - FOR2(FG, i, JM(j), k);
  const double diffz = FOR3(FG, i, j, k) - FOR3(FG, i, j, KM(k));
  const double normx = FOR1(AFG, i, j, k) + FOR1(AFG, IM(i), j, k);
  const double normy = FOR2(AFG, i, j, k) + FOR2(AFG, i, JM(j), k);
  const double normz = FOR3(AFG, i, j, k) + FOR3(AFG, i, j, KM(k));
  const double norm = 1.0 / (fabs(normx / g->dx[1] + normy / g->dx[2] + normz / g->dx[3]) + SMALL);
  return fabs((diffx / g->dx[1] + diffy / g->dx[2] + diffz / g->dx[3]) * norm);
#undef GD
#undef FG
#undef AFG
#undef IM
#undef JM
#undef KM
#undef AV1
#undef AV2
#undef FOR1
#undef FOR2
#undef FOR3
}

void o_ener_totals(const o_grid* g, const double* p, double* out)
{
  dims_t d = D(g);
  const double dVF = g->dx[1] * g->dx[2] * g->dx[3];
  double divbmax = 0.0, divbavg = 0.0;
  for (int pl = 0; pl < NPR; pl++) out[pl] = 0.0;
  /* constotal (dump_ener.c:882) and divbmaxavg (dump_ener.c:680) */
  for (int k = 0; k < d.n3; k++)
    for (int j = 0; j < d.n2; j++)
      for (int i = 0; i < d.n1; i++) {
        o_geom G;
        geom_at(g, &d, 0, i, j, &G);
        o_state q;
        double U[NPR];
        o_get_state(p + ix(&d, i, j, k), &G, g->gam, &q);
        o_primtoflux(p + ix(&d, i, j, k), &q, 0, &G, U);
        for (int pl = 0; pl < NPR; pl++) out[pl] += U[pl] * dVF;
        const double divb = divb_at(g, &d, p, i, j, k);
        if (divb > divbmax) divbmax = divb;
        divbavg += divb;
      }
  out[8] = divbmax;
  out[9] = divbavg;
}

/* dump_content (dump.c:439-600, CALCFARADAYANDCURRENTS=1) for active zone
 * (i,j,k): the columns after the coordinates and without the four-current:
 * p[8], pressure, cs2, Sden (idealgaseos.c:19,48,71), U[8] from ucum, divb,
 * u^mu, u_mu, b^mu, b_mu (bcon_calc), vmin/vmax per direction (vchar),
 * gdet, F^{mu nu} (current_precalc CURTYPEFARADAY, phys.c:701-708) and
 * F_{mu nu} (lowerf, phys.tools.c:1449). */
void o_dump_row(const o_grid* g, const double* p, const double* ucum, int i, int j, int k, double* out)
{
  dims_t d = D(g);
  const double* pr = p + ix(&d, i, j, k);
  o_geom G;
  geom_at(g, &d, 0, i, j, &G);
  int c = 0;
  for (int pl = 0; pl < NPR; pl++) out[c++] = pr[pl];
  const double rho0 = pr[O_RHO], u = pr[O_UU];
  const double pressure = (g->gam - 1.) * u;
  out[c++] = pressure;
  out[c++] = g->gam * pressure / (rho0 + u + pressure);
  const double indexn = 1.0 / (g->gam - 1.);
  double entropy;
  if (rho0 < SMALL && pressure < SMALL) entropy = 0.0;
  else if (rho0 < SMALL) entropy = 0.0;
  else if (pressure < SMALL) entropy = -1.e+100;
  else entropy = rho0 * log(pow(pressure, indexn) / pow(rho0, indexn + 1.0));
  out[c++] = entropy;
  for (int pl = 0; pl < NPR; pl++) out[c++] = ucum[ix(&d, i, j, k) + pl];
  out[c++] = divb_at(g, &d, p, i, j, k);
  o_state q;
  o_get_state(pr, &G, g->gam, &q);
  double bcon[4], bcov[4];
  bcon[0] = pr[O_B1] * q.ucov[1] + pr[O_B2] * q.ucov[2] + pr[O_B3] * q.ucov[3];
  for (int m = 1; m < 4; m++) bcon[m] = (pr[O_B1 - 1 + m] + bcon[0] * q.ucon[m]) / q.ucon[0];
  for (int a = 0; a < 4; a++) {
    bcov[a] = 0.0;
    for (int b = 0; b < 4; b++) bcov[a] += G.gcov[a][b] * bcon[b];
  }
  for (int m = 0; m < 4; m++) out[c++] = q.ucon[m];
  for (int m = 0; m < 4; m++) out[c++] = q.ucov[m];
  for (int m = 0; m < 4; m++) out[c++] = bcon[m];
  for (int m = 0; m < 4; m++) out[c++] = bcov[m];
  for (int dir = 1; dir <= 3; dir++) {
    double vmax, vmin;
    o_vchar(pr, &q, dir, &G, g->gam, &vmax, &vmin);
    out[c++] = vmin;
    out[c++] = vmax;
  }
  out[c++] = G.gdet;
  double f[6];
  f[0] = -1.0 / G.gdet * (bcov[3] * q.ucov[2] - bcov[2] * q.ucov[3]);
  f[1] = -1.0 / G.gdet * (-bcov[3] * q.ucov[1] + bcov[1] * q.ucov[3]);
  f[2] = -1.0 / G.gdet * (bcov[2] * q.ucov[1] - bcov[1] * q.ucov[2]);
  f[3] = -1.0 / G.gdet * (bcov[3] * q.ucov[0] - bcov[0] * q.ucov[3]);
  f[4] = -1.0 / G.gdet * (-bcov[2] * q.ucov[0] + bcov[0] * q.ucov[2]);
  f[5] = -1.0 / G.gdet * (bcov[1] * q.ucov[0] - bcov[0] * q.ucov[1]);
  double F[4][4] = {{0}};
  F[0][1] = f[0]; F[0][2] = f[1]; F[0][3] = f[2]; F[1][2] = f[3]; F[1][3] = f[4]; F[2][3] = f[5];
  F[1][0] = -f[0]; F[2][0] = -f[1]; F[3][0] = -f[2]; F[2][1] = -f[3]; F[3][1] = -f[4]; F[3][2] = -f[5];
  for (int m = 0; m < 6; m++) out[c++] = f[m];
  static const int lo[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  for (int m = 0; m < 6; m++) {
    const int a = lo[m][0], b = lo[m][1];
    double sum = 0.0;
    for (int jp = 0; jp < 4; jp++)
      for (int kp = 0; kp < 4; kp++) sum += F[jp][kp] * G.gcov[a][jp] * G.gcov[b][kp];
    out[c++] = sum;
  }
}

/* The four-current of the main dump (CALCFARADAYANDCURRENTS 1, WHICHCURRENTCALC
 * CURRENTCALC1, definit.h:595-597).  cf holds 6 slots of 3 Faraday samples
 * per storage zone, cf[zone * 18 + slot * 3 + c] (zone = i + ns1 (j + ns2 k)
 * over the storage box).
 *
 * o_current_precalc: current_doprecalc (phys.c:552) for CURTYPET, X, Y, Z at
 * zone centres over the whole storage box with current_precalc's formulas
 * (phys.c:622-686): slot 4 <- slot 5, slot 5 <- slot 0 (the two previous
 * samples of F^{it}), then slot 0 = (F^{rt}, F^{ht}, F^{pt}) now, slot 1 =
 * (F^{tr}, F^{hr}, F^{pr}), slot 2 = (F^{th}, F^{rh}, F^{ph}), slot 3 =
 * (F^{rt}, F^{rp}, F^{hp}) as the x1, x2, x3 derivatives take them. */
void o_current_precalc(const o_grid* g, const double* p, double* cf)
{
  dims_t d = D(g);
  for (int k = -d.ng3; k < d.n3 + d.ng3; k++)
    for (int j = -d.ng2; j < d.n2 + d.ng2; j++)
      for (int i = -d.ng1; i < d.n1 + d.ng1; i++) {
        const double* pr = p + ix(&d, i, j, k);
        double* f = cf + ix(&d, i, j, k) / NPR * 18;
        o_geom G;
        geom_at(g, &d, 0, i, j, &G);
        o_state q;
        o_get_state(pr, &G, g->gam, &q);
        double bcon[4], bcov[4];
        bcon[0] = pr[O_B1] * q.ucov[1] + pr[O_B2] * q.ucov[2] + pr[O_B3] * q.ucov[3];
        for (int m = 1; m < 4; m++) bcon[m] = (pr[O_B1 - 1 + m] + bcon[0] * q.ucon[m]) / q.ucon[0];
        for (int a = 0; a < 4; a++) {
          bcov[a] = 0.0;
          for (int b = 0; b < 4; b++) bcov[a] += G.gcov[a][b] * bcon[b];
        }
        const double* u = q.ucov;
        const double* b = bcov;
        const double gd = G.gdet;
        for (int c = 0; c < 3; c++) {
          f[12 + c] = f[15 + c];
          f[15 + c] = f[c];
        }
        f[0] = -1.0 / gd * (-b[3] * u[2] + b[2] * u[3]);
        f[1] = -1.0 / gd * (b[3] * u[1] - b[1] * u[3]);
        f[2] = -1.0 / gd * (-b[2] * u[1] + b[1] * u[2]);
        f[3] = -1.0 / gd * (b[3] * u[2] - b[2] * u[3]);
        f[4] = -1.0 / gd * (-b[3] * u[0] + b[0] * u[3]);
        f[5] = -1.0 / gd * (b[2] * u[0] - b[0] * u[2]);
        f[6] = -1.0 / gd * (-b[3] * u[1] + b[1] * u[3]);
        f[7] = -1.0 / gd * (b[3] * u[0] - b[0] * u[3]);
        f[8] = -1.0 / gd * (-b[1] * u[0] + b[0] * u[1]);
        f[9] = 1.0 / gd * (-b[3] * u[2] + b[2] * u[3]);
        f[10] = 1.0 / gd * (b[2] * u[0] - b[0] * u[2]);
        f[11] = 1.0 / gd * (-b[1] * u[0] + b[0] * u[1]);
      }
}

/* current_calc_1 (phys.c:846-1000) at the active zones: J^mu from centred
 * differences of gdet F over two zones (x3 differences vanish when n3 = 1:
 * kp1mac(k) = k) and, from the third call on (timeterms != 0), the time
 * derivative of F^{it} at the newest of three samples dtl and dtr apart
 * (the quadratic through them).  jcon[(i + n1 (j + n2 k)) * 4 + mu]. */
void o_current_calc(const o_grid* g, const double* cf, int timeterms, double dtl, double dtr, double* jcon)
{
  dims_t d = D(g);
  const double idx1 = 1.0 / (2.0 * g->dx[1]), idx2 = 1.0 / (2.0 * g->dx[2]), idx3 = 1.0 / (2.0 * g->dx[3]);
  const int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  for (int k = 0; k < d.n3; k++)
    for (int j = 0; j < d.n2; j++)
      for (int i = 0; i < d.n1; i++) {
        o_geom Gt, Grp, Grm, Ghp, Ghm;
        geom_at(g, &d, 0, i, j, &Gt);
        geom_at(g, &d, 0, i + s1, j, &Grp);
        geom_at(g, &d, 0, i - s1, j, &Grm);
        geom_at(g, &d, 0, i, j + s2, &Ghp);
        geom_at(g, &d, 0, i, j - s2, &Ghm);
        const double gt = Gt.gdet;
        const double* f = cf + ix(&d, i, j, k) / NPR * 18;
        const double* frp = cf + ix(&d, i + s1, j, k) / NPR * 18;
        const double* frm = cf + ix(&d, i - s1, j, k) / NPR * 18;
        const double* fhp = cf + ix(&d, i, j + s2, k) / NPR * 18;
        const double* fhm = cf + ix(&d, i, j - s2, k) / NPR * 18;
        const double* fpp = cf + ix(&d, i, j, k + s3) / NPR * 18;
        const double* fpm = cf + ix(&d, i, j, k - s3) / NPR * 18;
        double tt[4] = {0.0, 0.0, 0.0, 0.0};
        if (timeterms) {
          for (int c = 0; c < 3; c++) {
            const double fl = f[12 + c], f0 = f[15 + c], fr = f[c];
            const double AA = (dtr * (fl - f0) + dtl * (fr - f0)) / (dtl * dtr * (dtl + dtr));
            const double BB = (-dtr * dtr * (fl - f0) + dtl * dtl * (fr - f0)) / (dtl * dtr * (dtl + dtr));
            tt[1 + c] = 2.0 * AA * dtr + BB;
          }
        }
        double* J = jcon + (((long)k * d.n2 + j) * d.n1 + i) * 4;
        J[0] = +1. / gt * (Grp.gdet * frp[3] - Grm.gdet * frm[3]) * idx1
               + 1. / gt * (Ghp.gdet * fhp[6] - Ghm.gdet * fhm[6]) * idx2
               + 1. / gt * (gt * fpp[9] - gt * fpm[9]) * idx3;
        J[1] = +tt[1] + 1. / gt * (Ghp.gdet * fhp[7] - Ghm.gdet * fhm[7]) * idx2
               + 1. / gt * (gt * fpp[10] - gt * fpm[10]) * idx3;
        J[2] = +tt[2] + 1. / gt * (Grp.gdet * frp[4] - Grm.gdet * frm[4]) * idx1
               + 1. / gt * (gt * fpp[11] - gt * fpm[11]) * idx3;
        J[3] = +tt[3] + 1. / gt * (Grp.gdet * frp[5] - Grm.gdet * frm[5]) * idx1
               + 1. / gt * (Ghp.gdet * fhp[8] - Ghm.gdet * fhm[8]) * idx2;
      }
}

void o_step(const o_grid* g, double* p, double* pk0, double* pk1, double* ui, double* uf, double* ucum, double* F1,
            double* F2, double* F3, int32_t* pflag, double* dt, double* fluxacc)
{
  dims_t d = D(g);
  double CUf[4][4], Cunew[4][4], wflux[4];
  const int nto = o_rk_tables(g->timeorder, CUf, Cunew);
  rk_flux_weights(g->timeorder, wflux);
  double* pii[3] = {p, p, p};
  double* pbb[3] = {p, pk0, pk1};
  double* pff[3] = {pk0, pk1, p};
  if (nto == 2) { pbb[1] = pk0; pff[1] = p; }
  double lastndt = BIG;
  double* en = g->ener;
  for (int to = 0; to < nto; to++) {
    const int finalstep = (to == nto - 1);
    double ndt[4], accdt;
    if (en) ener_source(g, &d, pbb[to], ener_dt4(en, *dt, finalstep), en);
    o_fluxcalc(g, pbb[to], F1, F2, F3, ndt, 1);
    if (en) ener_pdot(g, &d, F1, F2, F3, (*dt) * wflux[to], en);
    if (fluxacc && d.n1 > 1) {
      /* diag_flux_pureflux: F1[RHO], F1[UU] through the radial boundaries */
      double s[4] = {0, 0, 0, 0};
      for (int k = 0; k < d.n3; k++)
        for (int j = 0; j < d.n2; j++) {
          s[0] += F1[ix(&d, 0, j, k) + O_RHO];
          s[1] += F1[ix(&d, 0, j, k) + O_UU];
          s[2] += F1[ix(&d, d.n1, j, k) + O_RHO];
          s[3] += F1[ix(&d, d.n1, j, k) + O_UU];
        }
      for (int c = 0; c < 4; c++) fluxacc[c] += s[c] * g->dx[2] * g->dx[3] * (*dt) * wflux[to];
    }
    o_update(g, pii[to], pbb[to], F1, F2, F3, ui, uf, ucum, CUf[to], Cunew[to], *dt, to, &accdt);
    int64_t ns;
    double* fl = (en && finalstep) ? en + O_EN_FLTERMS : NULL;
    invert(g, pbb[to], finalstep ? ucum : uf, pff[to], pflag, finalstep, &ns, fl);
    const double wavedt = 1. / (1. / ndt[1] + 1. / ndt[2] + 1. / ndt[3]);
    const double nd = 1.0 * fmin(wavedt, accdt);
    if (nd < lastndt) lastndt = nd;
    o_bound(g, pff[to]);
    post_fixup(g, pff[to], pflag, finalstep, fl);
    o_bound(g, pff[to]);
  }
  if (en) en[O_EN_NSTEP] += 1.0;
  double fullndt = lastndt;
  if (fullndt > g->safe * (*dt)) fullndt = g->safe * (*dt);
  *dt = fullndt;
}
