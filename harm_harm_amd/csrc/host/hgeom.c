/*
 * hgeom.c -- MKS Kerr-Schild geometry tables, built once on the host.
 *
 * References (HARM tree): set_gcov_ksmetric metric.c:1301, bl_coord
 * LOGRSINTH coord.c:930, dxdxprim coord.c:1515, lapse/shift
 * metric.tools.c:1245-1280, connection conn_func_numerical1
 * metric.tools.c:1771 (numerical metric derivatives).
 */
#include <math.h>
#include <string.h>
#include "hgeom.h"

int hg_sym(int a, int b)
{
  static const int tab[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
  return tab[a][b];
}

void hg_ks_gcov(double a, double r, double th, double g[4][4], double* gpert_tt)
{
  const double sth = sin(th), cth = cos(th);
  const double s2 = sth * sth;
  const double rho2 = r * r + a * a * cth * cth;
  const double m = 2.0 * r / rho2; /* 2Mr/rho^2, M = 1 */
  memset(g, 0, sizeof(double) * 16);
  g[0][0] = -1.0 + m;
  g[0][1] = g[1][0] = m;
  g[0][3] = g[3][0] = -m * a * s2;
  g[1][1] = 1.0 + m;
  g[1][3] = g[3][1] = -a * s2 * (1.0 + m);
  g[2][2] = rho2;
  g[3][3] = s2 * (rho2 + a * a * s2 * (1.0 + m));
  *gpert_tt = m;
}

void hg_mks_V(const harm_config* c, const double X[4], double V[4])
{
  V[0] = X[0];
  V[1] = c->R0 + exp(X[1]);
  if (X[2] < 0.5)
    V[2] = M_PI * X[2] + 0.5 * (1.0 - c->hslope) * sin(2.0 * M_PI * X[2]);
  else
    V[2] = M_PI - M_PI * (1.0 - X[2]) - 0.5 * (1.0 - c->hslope) * sin(2.0 * M_PI * (1.0 - X[2]));
  V[3] = 2.0 * M_PI * X[3];
}

void hg_mks_dxdxp(const harm_config* c, const double X[4], double J[4][4])
{
  double V[4];
  hg_mks_V(c, X, V);
  memset(J, 0, sizeof(double) * 16);
  J[0][0] = 1.0;
  J[1][1] = V[1] - c->R0;
  if (X[2] < 0.5)
    J[2][2] = M_PI + (1.0 - c->hslope) * M_PI * cos(2.0 * M_PI * X[2]);
  else
    J[2][2] = M_PI + (1.0 - c->hslope) * M_PI * cos(2.0 * M_PI * (1.0 - X[2]));
  J[3][3] = 2.0 * M_PI;
}

void hg_mks_gcov(const harm_config* c, const double X[4], double g[4][4], double* gpert_tt)
{
  double V[4], J[4][4], gks[4][4];
  hg_mks_V(c, X, V);
  hg_mks_dxdxp(c, X, J);
  hg_ks_gcov(c->a, V[1], V[2], gks, gpert_tt);
  /* J is diagonal: g_mn = gks_mn J_mm J_nn */
  for (int m = 0; m < 4; m++)
    for (int n = 0; n < 4; n++) g[m][n] = gks[m][n] * J[m][m] * J[n][n];
}

void hg_record_from_gcov(const double g[4][4], const double gi[4][4], double gpert_tt, double gdet, double* rec)
{
  for (int a = 0; a < 4; a++)
    for (int b = a; b < 4; b++) {
      rec[HARM_G_GCOV + hg_sym(a, b)] = g[a][b];
      rec[HARM_G_GCON + hg_sym(a, b)] = gi[a][b];
    }
  rec[HARM_G_GCOVPERT_TT] = gpert_tt;
  rec[HARM_G_GDET] = gdet;
  /* alphalapse_func / betasqoalphasq_func / beta_func, metric.tools.c:1245-1280 */
  const double alpha = 1.0 / sqrt(fabs(-gi[0][0]));
  rec[HARM_G_ALPHA] = alpha;
  double bsq = 0.0;
  for (int j = 1; j < 4; j++) bsq += g[0][j] * gi[0][j];
  rec[HARM_G_BETASQOALPHASQ] = bsq;
  const double alphasq = alpha * alpha;
  rec[HARM_G_BETA1] = gi[0][1] * alphasq;
  rec[HARM_G_BETA2] = gi[0][2] * alphasq;
  rec[HARM_G_BETA3] = gi[0][3] * alphasq;
}

/* Closed-form inverse Kerr-Schild metric (M=1). */
void hg_ks_gcon(double a, double r, double th, double gi[4][4])
{
  const double sth = sin(th), cth = cos(th);
  const double rho2 = r * r + a * a * cth * cth;
  const double delta = r * r - 2.0 * r + a * a;
  memset(gi, 0, sizeof(double) * 16);
  gi[0][0] = -(1.0 + 2.0 * r / rho2);
  gi[0][1] = gi[1][0] = 2.0 * r / rho2;
  gi[1][1] = delta / rho2;
  gi[1][3] = gi[3][1] = a / rho2;
  gi[2][2] = 1.0 / rho2;
  gi[3][3] = 1.0 / (rho2 * sth * sth);
}

void hg_mks_record(const harm_config* c, const double X[4], double* rec)
{
  double g[4][4], gi[4][4], gp, V[4], J[4][4];
  hg_mks_gcov(c, X, g, &gp);
  hg_mks_V(c, X, V);
  hg_mks_dxdxp(c, X, J);
  hg_ks_gcon(c->a, V[1], V[2], gi);
  /* J diagonal: g^mn = gks^mn / (J_mm J_nn) */
  for (int m = 0; m < 4; m++)
    for (int n = 0; n < 4; n++) gi[m][n] = gi[m][n] / (J[m][m] * J[n][n]);
  const double r = V[1], th = V[2];
  const double cth = cos(th);
  const double rho2 = r * r + c->a * c->a * cth * cth;
  const double gdet = rho2 * fabs(sin(th)) * fabs(J[1][1] * J[2][2] * J[3][3]);
  hg_record_from_gcov(g, gi, gp, gdet, rec);
}

void hg_ks_record(double a, double r, double th, double* rec)
{
  double g[4][4], gi[4][4], gp;
  hg_ks_gcov(a, r, th, g, &gp);
  hg_ks_gcon(a, r, th, gi);
  const double cth = cos(th);
  const double rho2 = r * r + a * a * cth * cth;
  hg_record_from_gcov(g, gi, gp, rho2 * fabs(sin(th)), rec);
}

/* Ridders' method for a derivative (Ridders 1982, Adv. Eng. Software 4:75):
 * Neville extrapolation of central differences with step ratio 1.4. */
#define RD_CON 1.4
#define RD_NTAB 10
#define RD_SAFE 2.0
static double ridders(const harm_config* c, const double X[4], int dir, int m, int n, double h)
{
  double tab[RD_NTAB][RD_NTAB];
  double err = 1e300, ans = 0.0;
  double hh = h;
  for (int i = 0; i < RD_NTAB; i++) {
    double Xp[4], Xm[4], gp[4][4], gm[4][4], dum;
    memcpy(Xp, X, sizeof(Xp));
    memcpy(Xm, X, sizeof(Xm));
    Xp[dir] += hh;
    Xm[dir] -= hh;
    hg_mks_gcov(c, Xp, gp, &dum);
    hg_mks_gcov(c, Xm, gm, &dum);
    tab[0][i] = (gp[m][n] - gm[m][n]) / (Xp[dir] - Xm[dir]);
    if (i == 0) { ans = tab[0][0]; hh /= RD_CON; continue; }
    double fac = RD_CON * RD_CON;
    for (int j = 1; j <= i; j++) {
      tab[j][i] = (tab[j - 1][i] * fac - tab[j - 1][i - 1]) / (fac - 1.0);
      fac *= RD_CON * RD_CON;
      double e1 = fabs(tab[j][i] - tab[j - 1][i]);
      double e2 = fabs(tab[j][i] - tab[j - 1][i - 1]);
      double errt = e1 > e2 ? e1 : e2;
      if (errt <= err) { err = errt; ans = tab[j][i]; }
    }
    if (fabs(tab[i][i] - tab[i - 1][i - 1]) >= RD_SAFE * err) break;
    hh /= RD_CON;
  }
  return ans;
}

void hg_mks_conn(const harm_config* c, const double X[4], double h1, double h2, double* conn)
{
  double dg[4][4][4]; /* dg[d][m][n] = d g_mn / dX^d */
  memset(dg, 0, sizeof(dg));
  for (int m = 0; m < 4; m++)
    for (int n = m; n < 4; n++) {
      double d1 = ridders(c, X, 1, m, n, h1);
      double d2 = ridders(c, X, 2, m, n, h2);
      dg[1][m][n] = dg[1][n][m] = d1;
      dg[2][m][n] = dg[2][n][m] = d2;
    }
  double gi[4][4], V[4], J[4][4];
  hg_mks_V(c, X, V);
  hg_mks_dxdxp(c, X, J);
  hg_ks_gcon(c->a, V[1], V[2], gi);
  for (int m = 0; m < 4; m++)
    for (int n = 0; n < 4; n++) gi[m][n] = gi[m][n] / (J[m][m] * J[n][n]);
  /* Gamma^a_{bc} = 1/2 g^{ad} (d_c g_db + d_b g_dc - d_d g_bc) */
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++)
      for (int cc = b; cc < 4; cc++) {
        long double s = 0.0L;
        for (int d = 0; d < 4; d++)
          s += (long double)gi[a][d] * ((long double)dg[cc][d][b] + dg[b][d][cc] - dg[d][b][cc]);
        conn[a * 10 + hg_sym(b, cc)] = (double)(0.5L * s);
      }
}

double hg_rhor(double a)
{
  double disc = 1.0 - a * a;
  if (disc < 0.0) disc = 0.0;
  return 1.0 + sqrt(disc);
}

double hg_setRin(double a, double R0, double Rout, int n1)
{
  /* setihor: min(int(0.1 N1), MAXBND=4); setRin with ADJUSTFRACT 0.25 */
  int ihor = (int)((double)n1 * 0.1);
  if (ihor > 4) ihor = 4;
  double f = ((double)ihor + 0.25) / (double)n1;
  double Rhor = hg_rhor(a);
  return R0 + pow((Rhor - R0) / pow(Rout - R0, f), 1.0 / (1.0 - f));
}

void hg_ks_batch(int64_t n, const double* a, const double* r, const double* th, double* recsoa)
{
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < n; z++) {
    double rec[HARM_G_NFIELD];
    hg_ks_record(a[z], r[z], th[z], rec);
    for (int f = 0; f < HARM_G_NFIELD; f++) recsoa[f * n + z] = rec[f];
  }
}
