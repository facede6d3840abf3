/*
 * hinit.c -- Fishbone-Moncrief torus initial data on the host (C).
 *
 * Follows init_dsandvels_torus (initboundcode/init.fishmon.c:403-510),
 * lfish_calc (:843), the A_phi = max(rho_av/rho_max - 0.2, 0) field
 * (init_vpot_user, :680-750) taken to cell-centred B with the corner
 * differencing that keeps Toth's corner divergence zero, and the beta
 * normalisation by the minimum of p_gas/(b^2/2) (user1_normalize_field,
 * init.tools.c:745, FIELDBETANORMMETHOD 1).  The internal energy is
 * perturbed by u*(1 + randfact*(x - 0.5)) with x uniform in [0,1) from a
 * seeded splitmix64 stream drawn in (k, j, i) order.  Atmosphere and floors:
 * rho = rhomin r^-1.5, u = uumin r^-2.5 with zero relative velocity
 * (user1_set_atmosphere, init.tools.c:797, rescaletype 0).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "hgeom.h"
#include "hinit.h"

static double lfish(double a, double r)
{
  return (((pow(a, 2) - 2. * a * sqrt(r) + pow(r, 2))
           * ((-2. * a * r * (pow(a, 2) - 2. * a * sqrt(r) + pow(r, 2))) / sqrt(2. * a * sqrt(r) + (-3. + r) * r)
              + ((a + (-2. + r) * sqrt(r)) * (pow(r, 3) + pow(a, 2) * (2. + r)))
                    / sqrt(1 + (2. * a) / pow(r, 1.5) - 3. / r)))
          / (pow(r, 3) * sqrt(2. * a * sqrt(r) + (-3. + r) * r) * (pow(a, 2) + (-2. + r) * r)));
}

static uint64_t splitmix64(uint64_t* s)
{
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double uniform01(uint64_t* s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

/* FM torus state at BL (r, th); returns 0 outside the torus. */
static int torus_point(double a, double gam, double rin, double l, double kappa, double r, double th, double* rho,
                       double* u, double* up)
{
  const double sth = sin(th), cth = cos(th);
  const double DD = r * r - 2. * r + a * a;
  const double AA = (r * r + a * a) * (r * r + a * a) - DD * a * a * sth * sth;
  const double SS = r * r + a * a * cth * cth;
  const double thin = M_PI / 2., sthin = sin(thin), cthin = cos(thin);
  const double DDin = rin * rin - 2. * rin + a * a;
  const double AAin = (rin * rin + a * a) * (rin * rin + a * a) - DDin * a * a * sthin * sthin;
  const double SSin = rin * rin + a * a * cthin * cthin;
  double lnh;
  if (r >= rin) {
    lnh = 0.5 * log((1. + sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * sth * AA * sth))) / (SS * DD / AA))
          - 0.5 * sqrt(1. + 4. * (l * l * SS * SS) * DD / (AA * AA * sth * sth)) - 2. * a * r * l / AA
          - (0.5 * log((1. + sqrt(1. + 4. * (l * l * SSin * SSin) * DDin / (AAin * AAin * sthin * sthin)))
                       / (SSin * DDin / AAin))
             - 0.5 * sqrt(1. + 4. * (l * l * SSin * SSin) * DDin / (AAin * AAin * sthin * sthin))
             - 2. * a * rin * l / AAin);
  } else
    lnh = 1.;
  if (lnh < 0. || r < rin) return 0;
  const double hm1 = exp(lnh) - 1.;
  *rho = pow(hm1 * (gam - 1.) / (kappa * gam), 1. / (gam - 1.));
  *u = kappa * pow(*rho, gam) / (gam - 1.);
  const double expm2chi = SS * SS * DD / (AA * AA * sth * sth);
  const double up1 = sqrt((-1. + sqrt(1. + 4. * l * l * expm2chi)) / 2.);
  *up = 2. * a * r * sqrt(1. + up1 * up1) / sqrt(AA * SS * DD) + sqrt(SS / AA) * up1 / sth;
  return 1;
}

/* BL (u^t, 0, 0, u^phi) -> MKS relative 4-velocity */
static void bl_uphi_to_rel4(const harm_config* c, const double X[4], double r, double th, double up, double* utilde)
{
  const double a = c->a;
  const double sth = sin(th), cth = cos(th);
  const double SS = r * r + a * a * cth * cth;
  const double DD = r * r - 2. * r + a * a;
  const double AA = (r * r + a * a) * (r * r + a * a) - DD * a * a * sth * sth;
  const double gtt = -(1. - 2. * r / SS), gtp = -2. * a * r * sth * sth / SS, gpp = AA * sth * sth / SS;
  /* gtt ut^2 + 2 gtp ut up + gpp up^2 = -1 */
  const double A = gtt, B = 2. * gtp * up, C = gpp * up * up + 1.;
  const double disc = B * B - 4. * A * C;
  const double ut = (-B - sqrt(disc > 0 ? disc : 0.0)) / (2. * A);
  /* BL -> KS with u^r = 0 leaves u^t, u^phi unchanged; KS -> MKS */
  double J[4][4], rec[HARM_G_NFIELD];
  hg_mks_dxdxp(c, X, J);
  const double ucon[4] = {ut, 0.0, 0.0, up / J[3][3]};
  hg_mks_record(c, X, rec);
  const double beta[4] = {0.0, rec[HARM_G_BETA1], rec[HARM_G_BETA2], rec[HARM_G_BETA3]};
  for (int j = 1; j < 4; j++) utilde[j - 1] = ucon[j] + ucon[0] * beta[j];
}

static void zone_X(const harm_config* c, const hi_grid* G, int i, int j, int k, double X[4])
{
  X[0] = 0.0;
  X[1] = G->startx[1] + (i + 0.5) * G->dx[1];
  X[2] = G->startx[2] + (j + 0.5) * G->dx[2];
  X[3] = G->startx[3] + (k + 0.5) * G->dx[3];
}

static double bsq_of(const double* pr, const double* rec)
{
  /* b^2 = (B^2 + (B.u)^2)/(u^t)^2 */
  const double u1 = pr[HARM_U1], u2 = pr[HARM_U2], u3 = pr[HARM_U3];
  const double* g = rec;
  double qsq = g[4] * u1 * u1 + g[7] * u2 * u2 + g[9] * u3 * u3 + 2. * (g[5] * u1 * u2 + g[6] * u1 * u3);
  if (qsq < 0) qsq = 0;
  const double gamma = sqrt(1 + qsq);
  const double ut = gamma / rec[HARM_G_ALPHA];
  const double ucon[4] = {ut, u1 - ut * rec[HARM_G_BETA1], u2 - ut * rec[HARM_G_BETA2], u3 - ut * rec[HARM_G_BETA3]};
  double ucov[4];
  static const int sy[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
  for (int m = 0; m < 4; m++) {
    ucov[m] = 0;
    for (int n = 0; n < 4; n++) ucov[m] += g[sy[m][n]] * ucon[n];
  }
  const double B[4] = {0, pr[HARM_B1], pr[HARM_B2], pr[HARM_B3]};
  double Bsq = 0, udB = 0;
  for (int m = 1; m < 4; m++)
    for (int n = 1; n < 4; n++) Bsq += g[sy[m][n]] * B[m] * B[n];
  for (int m = 1; m < 4; m++) udB += B[m] * ucov[m];
  return (Bsq + udB * udB) / (ut * ut);
}

int hi_init_torus(const harm_config* c, const hi_grid* G, double rin, double rmax, double kappa, double beta,
                  double randfact, uint32_t seed, double* p)
{
  const int n1 = c->n1, n2 = c->n2, n3 = c->n3;
  const int ng1 = G->ng1, ng2 = G->ng2, ng3 = G->ng3;
  const int ns1 = n1 + 2 * ng1, ns2 = n2 + 2 * ng2;
  const double l = lfish(c->a, rmax);
  uint64_t rs = 0x5DEECE66Dull ^ (uint64_t)seed;
#define PIX(i, j, k) ((((long)((k) + ng3) * ns2 + ((j) + ng2)) * ns1 + ((i) + ng1)) * HARM_NPR)
  double rhomax = 0.0;
  /* torus or (temporary zero) atmosphere on active zones */
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        double X[4], V[4], rho, u, up;
        zone_X(c, G, i, j, k, X);
        hg_mks_V(c, X, V);
        double* pr = p + PIX(i, j, k);
        for (int pl = 0; pl < HARM_NPR; pl++) pr[pl] = 0.0;
        if (torus_point(c->a, c->gam, rin, l, kappa, V[1], V[2], &rho, &u, &up)) {
          double ut[3];
          bl_uphi_to_rel4(c, X, V[1], V[2], up, ut);
          pr[HARM_RHO] = rho;
          pr[HARM_UU] = u * (1. + randfact * (uniform01(&rs) - 0.5));
          pr[HARM_U1] = ut[0];
          pr[HARM_U2] = ut[1];
          pr[HARM_U3] = ut[2];
          if (rho > rhomax) rhomax = rho;
        }
      }
  if (rhomax <= 0.0) return -1;
  /* normalise densities to rho_max = 1 */
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        double* pr = p + PIX(i, j, k);
        pr[HARM_RHO] /= rhomax;
        pr[HARM_UU] /= rhomax;
      }
  /* A_phi at corners (i-1/2, j-1/2) from the average rho of the 4 zones around it */
  const int na1 = n1 + 1, na2 = n2 + 1;
  double* A = (double*)calloc((size_t)na1 * na2, sizeof(double));
  if (!A) return -2;
  for (int j = 0; j <= n2; j++)
    for (int i = 0; i <= n1; i++) {
      int cnt = 0;
      double s = 0.0;
      for (int dj = -1; dj <= 0; dj++)
        for (int di = -1; di <= 0; di++) {
          const int ii = i + di, jj = j + dj;
          if (ii < 0 || ii >= n1 || jj < 0 || jj >= n2) continue;
          s += p[PIX(ii, jj, 0) + HARM_RHO];
          cnt++;
        }
      const double q = (cnt ? s / cnt : 0.0) / 1.0 - 0.2;
      A[(size_t)j * na1 + i] = q > 0. ? q : 0.;
    }
  /* B from corner differences (axisymmetric: same for all k) */
  double bmin_beta = 1e300;
  for (int j = 0; j < n2; j++)
    for (int i = 0; i < n1; i++) {
      double X[4], rec[HARM_G_NFIELD];
      zone_X(c, G, i, j, 0, X);
      hg_mks_record(c, X, rec);
      const double gdet = rec[HARM_G_GDET];
      const double a00 = A[(size_t)j * na1 + i], a10 = A[(size_t)j * na1 + i + 1];
      const double a01 = A[(size_t)(j + 1) * na1 + i], a11 = A[(size_t)(j + 1) * na1 + i + 1];
      const double B1 = -(a00 - a01 + a10 - a11) / (2. * G->dx[2] * gdet);
      const double B2 = (a00 + a01 - a10 - a11) / (2. * G->dx[1] * gdet);
      for (int k = 0; k < n3; k++) {
        double* pr = p + PIX(i, j, k);
        pr[HARM_B1] = B1;
        pr[HARM_B2] = B2;
        pr[HARM_B3] = 0.0;
      }
    }
  free(A);
  /* beta normalisation: minimum p_gas/(b^2/2) over magnetised zones */
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        double X[4], rec[HARM_G_NFIELD];
        zone_X(c, G, i, j, k, X);
        hg_mks_record(c, X, rec);
        const double* pr = p + PIX(i, j, k);
        const double bsq = bsq_of(pr, rec);
        if (bsq > 0.0) {
          const double b = (c->gam - 1.0) * pr[HARM_UU] / (0.5 * bsq);
          if (b < bmin_beta) bmin_beta = b;
        }
      }
  if (bmin_beta < 1e300) {
    const double norm = sqrt(bmin_beta / beta);
    for (int k = 0; k < n3; k++)
      for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) {
          double* pr = p + PIX(i, j, k);
          pr[HARM_B1] *= norm;
          pr[HARM_B2] *= norm;
          pr[HARM_B3] *= norm;
        }
  }
  /* atmosphere and floors */
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        double X[4], V[4];
        zone_X(c, G, i, j, k, X);
        hg_mks_V(c, X, V);
        double* pr = p + PIX(i, j, k);
        const double rf = c->rhomin * pow(V[1], -1.5), uf = c->uumin * pow(V[1], -2.5);
        if (pr[HARM_RHO] < rf) pr[HARM_RHO] = rf;
        if (pr[HARM_UU] < uf) pr[HARM_UU] = uf;
      }
#undef PIX
  return 0;
}


/* ---------------------------------------------------------------------- */
/* Bondi accretion (TESTNUMBER 153, initboundcode/init.sasha.c:865, :2977) */
/* ---------------------------------------------------------------------- */
/* The reference reads the flow from initbounddata/bondigrid<N>.dat (written
 * by a Mathematica notebook for Gamma = 4/3, sonic point r_c = 8, rho_c =
 * 0.000421875).  Here it is computed: at the sonic point u_c^2 = 1/(2 r_c)
 * and the sound speed a_c^2 = u_c^2/(1 - 3 u_c^2); with a^2 = Gamma T/(1 +
 * (n+1) T), T = p/rho, n = 1/(Gamma-1), that fixes T_c; the polytropic
 * constant is K = 1 (rho_c = T_c^n), and the integrals C1 = T^n u^r r^2 and
 * C2 = (1 + (1+n) T)^2 (1 - 2/r + (u^r)^2) are those of the sonic point.  At
 * each r the temperature is the root of
 *   F(T) = 2 ln(1 + (1+n) T) + ln(1 - 2/r + (C1 / (T^n r^2))^2) - ln C2,
 * which has one minimum in T: the larger root (subsonic) outside r_c, the
 * smaller (supersonic) inside, both by bisection to the last bit. */
static const double kBondiRc = 8.0;

static void bondi_constants(double gam, double* n, double* C1, double* C2, double* Tc)
{
  *n = 1.0 / (gam - 1.0);
  const double uc2 = 1.0 / (2.0 * kBondiRc), uc = -sqrt(uc2);
  const double ac2 = uc2 / (1.0 - 3.0 * uc2);
  *Tc = ac2 / (gam - (*n + 1.0) * ac2);
  *C1 = pow(*Tc, *n) * uc * kBondiRc * kBondiRc;
  *C2 = pow(1.0 + (1.0 + *n) * (*Tc), 2.0) * (1.0 - 2.0 / kBondiRc + uc2);
}

static double bondi_F(double T, double n, double C1, double C2, double r)
{
  const double u = C1 / (pow(T, n) * r * r);
  return 2.0 * log(1.0 + (1.0 + n) * T) + log(1.0 - 2.0 / r + u * u) - log(C2);
}

static double bondi_dF(double T, double n, double C1, double r)
{
  const double u = C1 / (pow(T, n) * r * r);
  const double w = 1.0 - 2.0 / r + u * u;
  return 2.0 * (1.0 + n) / (1.0 + (1.0 + n) * T) - 2.0 * n * u * u / (T * w);
}

/* bisection of f over [lo, hi] (f(lo), f(hi) of opposite sign) until the
 * interval stops shrinking */
static double bisect(double (*f)(double, const double*), const double* par, double lo, double hi)
{
  double flo = f(lo, par);
  for (int it = 0; it < 2000; it++) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    const double fm = f(mid, par);
    if ((fm < 0) == (flo < 0)) { lo = mid; flo = fm; }
    else hi = mid;
  }
  return 0.5 * (lo + hi);
}
static double f_F(double T, const double* q) { return bondi_F(T, q[0], q[1], q[2], q[3]); }
static double f_dF(double T, const double* q) { return bondi_dF(T, q[0], q[1], q[3]); }

int hi_bondi_state(double gam, double r, double* rho, double* uu, double* ur)
{
  double n, C1, C2, Tc;
  bondi_constants(gam, &n, &C1, &C2, &Tc);
  const double q[4] = {n, C1, C2, r};
  double lo = 1e-12, hi = 1e3;
  if (!(bondi_dF(lo, n, C1, r) < 0.0 && bondi_dF(hi, n, C1, r) > 0.0)) return 1;
  const double Tmin = bisect(f_dF, q, lo, hi);
  double T;
  if (r == kBondiRc) T = Tc;
  else if (bondi_F(Tmin, n, C1, C2, r) > 0.0) return 1; /* no flow through this radius */
  else if (r > kBondiRc) T = bisect(f_F, q, Tmin, hi);
  else T = bisect(f_F, q, lo, Tmin);
  *rho = pow(T, n); /* K = 1 */
  *uu = *rho * T * n;
  *ur = C1 / (pow(T, n) * r * r);
  return 0;
}

int hi_init_bondi(const harm_config* c, const hi_grid* G, double* p)
{
  const int n1 = c->n1, n2 = c->n2, n3 = c->n3;
  const int ng1 = G->ng1, ng2 = G->ng2, ng3 = G->ng3;
  const int ns1 = n1 + 2 * ng1, ns2 = n2 + 2 * ng2;
#define PIX(i, j, k) ((((long)((k) + ng3) * ns2 + ((j) + ng2)) * ns1 + ((i) + ng1)) * HARM_NPR)
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < n1; i++) {
        double X[4], V[4], J[4][4], rec[HARM_G_NFIELD], rho, uu, ur;
        zone_X(c, G, i, j, k, X);
        hg_mks_V(c, X, V);
        if (hi_bondi_state(c->gam, V[1], &rho, &uu, &ur) != 0) return 1;
        hg_mks_dxdxp(c, X, J);
        hg_mks_record(c, X, rec);
        /* radial KS 4-velocity -> MKS relative 4-velocity (u^t > 0) */
        const double u1 = ur / J[1][1];
        const double A = rec[0], B = 2.0 * rec[1] * u1, C = 1.0 + rec[4] * u1 * u1;
        const double ut = (-B - sqrt(B * B - 4.0 * A * C)) / (2.0 * A);
        double* pr = p + PIX(i, j, k);
        for (int pl = 0; pl < HARM_NPR; pl++) pr[pl] = 0.0;
        pr[HARM_RHO] = rho;
        pr[HARM_UU] = uu;
        pr[HARM_U1] = u1 + ut * rec[HARM_G_BETA1];
        pr[HARM_U2] = ut * rec[HARM_G_BETA2];
        pr[HARM_U3] = ut * rec[HARM_G_BETA3];
      }
#undef PIX
  return 0;
}

