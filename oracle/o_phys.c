/*
 * o_phys.c -- oracle restatement of HARM's per-zone physics (CPU, FP64).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Configuration restated: WHICHVEL=VELREL4, REMOVERESTMASSFROMUU=2,
 * MAXWELL=PRIMMAXWELL, WHICHEOM=WITHGDET, ideal gas, UD0PLUS1FIX=1,
 * DOMIXTHETAPHI=0 (init.fishmon.h:145-223, definit.h:241, metric.h:3).
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

#define O_SMALL 1.e-300
#define O_NUMEPSILON 2.2204460492503131e-16

static const int symtab[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};

void o_geom_from_rec(const double* rec, o_geom* g)
{
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) {
      g->gcov[a][b] = rec[symtab[a][b]];
      g->gcon[a][b] = rec[10 + symtab[a][b]];
    }
  g->gpert_tt = rec[20];
  g->gdet = rec[21];
  g->alpha = rec[22];
  g->bsqoa2 = rec[23];
  g->beta[0] = 0.0;
  g->beta[1] = rec[24];
  g->beta[2] = rec[25];
  g->beta[3] = rec[26];
}

/* lower_vec, phys.tools.c:1421 */
static void lower(const double* ucon, const o_geom* g, double* ucov)
{
  for (int a = 0; a < 4; a++)
    ucov[a] = g->gcov[a][0] * ucon[0] + g->gcov[a][1] * ucon[1] + g->gcov[a][2] * ucon[2] + g->gcov[a][3] * ucon[3];
}

/* get_state: ucon_calc_rel4vel (phys.tools.c:1925), gamma_calc_fromuconrel
 * (:1975, production clamp of q^2<0), qsq_calc (:2014), lower_vec,
 * compute_1plusud0_rel4vel (:1229), bsq_calc_fromq_rel4vel (:1392),
 * ideal-gas pressure. */
void o_get_state(const double* pr, const o_geom* g, double gam, o_state* q)
{
  const double u1 = pr[O_U1], u2 = pr[O_U2], u3 = pr[O_U3];
  double qsq = g->gcov[1][1] * u1 * u1 + g->gcov[2][2] * u2 * u2 + g->gcov[3][3] * u3 * u3
               + 2. * (g->gcov[1][2] * u1 * u2 + g->gcov[1][3] * u1 * u3);
  qsq = qsq * (qsq >= 0.0);
  const double gamma = sqrt(1. + qsq);
  q->qsq = qsq;
  q->gamma = gamma;
  q->ucon[0] = gamma / g->alpha;
  q->ucon[1] = u1 - q->ucon[0] * g->beta[1];
  q->ucon[2] = u2 - q->ucon[0] * g->beta[2];
  q->ucon[3] = u3 - q->ucon[0] * g->beta[3];
  lower(q->ucon, g, q->ucov);
  /* 1 + u_t without cancellation */
  double ud0tilde = 0.0;
  ud0tilde += u1 * g->gcov[0][1];
  ud0tilde += u2 * g->gcov[0][2];
  ud0tilde += u3 * g->gcov[0][3];
  const double alpha = g->alpha, alphasq = alpha * alpha;
  q->p1ud0 = ud0tilde + (g->gpert_tt - alphasq * (g->bsqoa2 + qsq)) / (1.0 + gamma * alpha);
  /* b^2 from primitive B */
  double Bcon[4] = {0.0, pr[O_B1], pr[O_B2], pr[O_B3]}, Bcov[4];
  for (int jj = 0; jj < 4; jj++) {
    Bcov[jj] = 0.0;
    for (int kk = 1; kk < 4; kk++) Bcov[jj] += Bcon[kk] * g->gcov[jj][kk];
  }
  double Bsq = 0.0, udotB = 0.0;
  for (int kk = 1; kk < 4; kk++) Bsq += Bcon[kk] * Bcov[kk];
  for (int kk = 1; kk < 4; kk++) udotB += Bcon[kk] * q->ucov[kk];
  double bsq = (Bsq + udotB * udotB);
  bsq *= 1.0 / (q->ucon[0] * q->ucon[0]);
  q->bsq = bsq;
  q->pressure = (gam - 1.0) * pr[O_UU];
}

/* mhd_calc_norestmass_ma (phys.tools.c:1071) + mhd_calc_primfield_em (:1133) */
static void mhd_ma(const double* pr, int dir, const o_state* q, double* mhd)
{
  const double rho = pr[O_RHO], u = pr[O_UU], P = q->pressure;
  const double w = P + rho + u;
  mhd[0] = (P + u) * q->ucon[dir] * q->ucov[0] + rho * q->ucon[dir] * q->p1ud0;
  for (int j = 1; j < 4; j++) mhd[j] = w * q->ucon[dir] * q->ucov[j];
  mhd[dir] += P;
}

static void mhd_em(const double* pr, int dir, const o_geom* g, const o_state* q, double* mhd)
{
  double Bcon[4] = {0.0, pr[O_B1], pr[O_B2], pr[O_B3]}, Bcov[4];
  for (int jj = 0; jj < 4; jj++) {
    Bcov[jj] = 0.0;
    for (int kk = 1; kk < 4; kk++) Bcov[jj] += Bcon[kk] * g->gcov[jj][kk];
  }
  double Bsq = 0.0, udotB = 0.0;
  for (int kk = 1; kk < 4; kk++) Bsq += Bcon[kk] * Bcov[kk];
  for (int kk = 1; kk < 4; kk++) udotB += Bcon[kk] * q->ucov[kk];
  for (int mu = 0; mu < 4; mu++)
    mhd[mu] = (Bsq * (q->ucon[dir]) * (q->ucov[mu]) - Bcon[dir] * Bcov[mu]
               - udotB * (Bcon[dir] * (q->ucov[mu]) + Bcov[mu] * (q->ucon[dir])));
  mhd[dir] += 0.5 * (Bsq + udotB * udotB);
  const double oneovergammasq = 1.0 / ((q->ucon[0]) * (q->ucon[0]));
  for (int mu = 0; mu < 4; mu++) mhd[mu] *= oneovergammasq;
}

void o_mhd_calc(const double* pr, int dir, const o_geom* g, const o_state* q, double* mhd)
{
  double ma[4], em[4];
  mhd_ma(pr, dir, q, ma);
  mhd_em(pr, dir, g, q, em);
  for (int j = 0; j < 4; j++) mhd[j] = (ma[j] + 0.0) + em[j];
}

/* primtoflux (phys.tools.c:30): MA + EM, dualfaraday (:350, PRIMMAXWELL),
 * then UtoU_fromunothing(UEVOLVE) = x gdet (:708). */
void o_primtoflux(const double* pr, const o_state* q, int dir, const o_geom* g, double* F)
{
  double ma[4], em[4], fl[O_NPR];
  fl[O_RHO] = pr[O_RHO] * q->ucon[dir];
  mhd_ma(pr, dir, q, ma);
  mhd_em(pr, dir, g, q, em);
  for (int j = 0; j < 4; j++) fl[O_UU + j] = ma[j] + em[j];
  if (dir > 0) {
    const double Bd = pr[O_B1 + dir - 1];
    fl[O_B1] = (pr[O_B1] * q->ucon[dir] - Bd * q->ucon[1]) / q->ucon[0];
    fl[O_B2] = (pr[O_B2] * q->ucon[dir] - Bd * q->ucon[2]) / q->ucon[0];
    fl[O_B3] = (pr[O_B3] * q->ucon[dir] - Bd * q->ucon[3]) / q->ucon[0];
  } else {
    fl[O_B1] = pr[O_B1];
    fl[O_B2] = pr[O_B2];
    fl[O_B3] = pr[O_B3];
  }
  for (int pl = 0; pl < O_NPR; pl++) F[pl] = fl[pl] * g->gdet;
}

/* vchar (vchar.c:34) with simplefast (:260, USESASHAREWRITE=1).  HARM calls
 * fail() (process abort) on a negative cms2 / discriminant beyond round-off;
 * the restatement clamps instead (such states never occur in the tests). */
void o_vchar(const double* pr, const o_state* q, int dir, const o_geom* g, double gam, double* vmax, double* vmin)
{
  double bsq = q->bsq;
  double rho = pr[O_RHO], u = pr[O_UU];
  if (u < 0) u = 0;
  if (rho < 0) rho = O_SMALL;
  const double P = q->pressure;
  double WW = rho + u + P, cs2;
  if (WW > 0.0) {
    const double pressure = (gam - 1.0) * u;
    const double h = rho + u + pressure;
    cs2 = gam * pressure / h;
  } else
    WW = cs2 = O_SMALL;
  if (cs2 >= 1.0) cs2 = 1.0 - O_NUMEPSILON * 10.0;
  else if (cs2 <= 0.0) cs2 = O_SMALL;
  const double EF = O_SMALL + fabs(bsq + WW);
  const double va2 = bsq / EF;
  double cms2 = cs2 + va2 - cs2 * va2;
  if (cms2 < 0.0) cms2 = 0.0;
  if (cms2 > 1.0) cms2 = 1.0;
  /* simplefast */
  const double Asq = g->gcon[dir][dir];
  const double Bsq = g->gcon[0][0];
  const double Au = q->ucon[dir];
  const double Bu = q->ucon[0];
  const double AB = g->gcon[dir][0];
  const double Au2 = Au * Au, Bu2 = Bu * Bu, AuBu = Au * Bu;
  const double B = 2. * (AuBu * (1.0 - cms2) - AB * cms2);
  const double A = Bu2 * (1.0 - cms2) - Bsq * cms2;
  double discr = 4.0 * cms2 * ((AB * AB - Asq * Bsq) * cms2 + (2.0 * AB * Au * Bu - Asq * Bu2 - Bsq * Au2) * (cms2 - 1.0));
  if ((discr < 0.0) && (discr > -1E-10)) discr = 0.0;
  else if (discr < 0.0) discr = 0.0;
  discr = sqrt(discr);
  const double vm = -(-B + discr) / (2. * A);
  const double vp = -(-B - discr) / (2. * A);
  if (vp < vm) { *vmax = vm; *vmin = vp; }
  else { *vmax = vp; *vmin = vm; }
}

/* source (phys.tools.c:1294) -> source_conn (phys.c:132): contract the
 * stress tensor with the connection, rest-mass term for
 * REMOVERESTMASSFROMUU=2, times gdet; no other physics sources (NOCOOLING,
 * ideal gas). */
void o_source(const double* pr, const o_geom* g, const o_state* q, const double* conn, double* dU)
{
  double mhd[4][4], dUconn[O_NPR];
  for (int j = 0; j < 4; j++) o_mhd_calc(pr, j, g, q, mhd[j]);
  for (int pl = 0; pl < O_NPR; pl++) dUconn[pl] = 0;
  for (int l = 0; l < 4; l++)
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 4; k++) dUconn[O_UU + l] += mhd[j][k] * conn[k * 10 + symtab[l][j]];
  for (int l = 0; l < 4; l++)
    for (int j = 0; j < 4; j++) dUconn[O_UU + l] += -pr[O_RHO] * q->ucon[j] * conn[0 * 10 + symtab[l][j]];
  for (int pl = 0; pl < O_NPR; pl++) dU[pl] = 0.0 + g->gdet * dUconn[pl];
}
