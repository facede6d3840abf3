/*
 * o_u2p.c -- oracle restatement of HARM's conserved->primitive inversion:
 * Utoprimgen (utoprimgen.c:32) -> Utoprim_jon_nonrelcompat_inputnorestmass
 * (utoprim_jon.c:136) -> Utoprim_new_body (:406) with WHICHHOTINVERTER=3
 * (1-D Newton on W' for E'(W'), func_Eprime_opt :2953), no line search
 * (USE_LINE_SEARCH=-2), PRECISEINVERSION tolerances (u2p_defs.h:20-60),
 * ideal gas (idealgaseos.c:295-410, utoprim_jon_eos.c:652); then, as the
 * torus configuration compiles it (HOT2COLD 1, init.fishmon.h:325), the cold
 * GRMHD retry of a failed hot solve (trycoldinversion, utoprimgen.c:597-690:
 * WHICHCOLDINVERTER=2, the P^2[W'] residual Psq_Wp :1923 / dPsqdWp_Wp :1944
 * with the cold EOS p = 0, coldeos.c), and, where the configuration turns it
 * on (CHECKONINVERSION, on for the Bondi problem's init.sasha.h:198, off for
 * the torus's init.fishmon.h:74,225), check_on_inversion with FAILIFBADCHECK
 * (utoprimgen.c:18, :697-850).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Not restated: HOT2ENTROPY (utoprimgen.c:200-226), which needs the entropy
 * variable that NPR=8 leaves out.
 */
#include <math.h>
#include <float.h>
#include <string.h>
#include "oracle.h"

#define NUMEPSILON DBL_EPSILON
#define SMALL 1.e-300
#define KINDASMALL 1.e-100
#define VERYBIG 1.e150
#define SQRTMINNUMREPRESENT 1.4916681462400413e-154 /* sqrt(DBL_MIN) */
#define GAMMASQCHECKRESID 1e2
#define ITERDAMPSTART 10
#define MAX_NEWT_ITER 200
#define NEWT_TOL (1E3 * NUMEPSILON)
#define NEWT_TOL_ULTRAREL (NUMEPSILON)
#define MIN_NEWT_TOL 1E-6
#define EXTRA_NEWT_ITER 2
#define EXTRA_NEWT_ITER_ULTRAREL 4
#define MAXNUMGUESSCHANGES 1000
#define FAILCONVRET 50

typedef struct {
  double Bsq, QdotB, QdotBsq, Qtsq, Qdotn, Qdotnp, D;
  double w[3];      /* wglobal */
  double gam;
  double gsq_too_big, utsq_too_big;
} inv_t;

static double dmax(double a, double b) { return a > b ? a : b; }

/* utoprim_jon.c:1730 */
static double utsq_calc(double W, const inv_t* s)
{
  const double S = s->QdotB, SoW = S / W;
  const double top = s->Qtsq + (s->Bsq + 2.0 * W) * SoW * SoW;
  const double bot = (s->Bsq * s->Bsq - s->Qtsq) + 2.0 * s->Bsq * W + W * W - SoW * SoW * (s->Bsq + 2.0 * W);
  return top / bot;
}

/* utoprim_jon.c:1709 */
static int vsqgtr1(double W, const inv_t* s)
{
  return (W * W * W * (W + 2. * s->Bsq) - s->QdotBsq * (2. * W + s->Bsq) <= W * W * (s->Qtsq - s->Bsq * s->Bsq));
}

/* utoprim_jon.c:1844 */
static int coldgrhd(int* pflag, double Qtsq, double D, double* Wp)
{
  const double det = D * D + Qtsq;
  if (det < 0.0) { *pflag = 3; return 1; }
  if (D > 0.0) *Wp = Qtsq / (D + sqrt(det));
  else *Wp = Qtsq / (D - sqrt(det));
  return 0;
}

/* which equations the Newton solve works on (eomtype) */
enum { EOM_HOT = 0, EOM_COLD = 1 };

static double wmrho0_utsq(double Wp, double D, double utsq, double gamma, double gammasq)
{
  return (Wp - D * utsq / (1.0 + gamma)) / gammasq;
}

/* func_Eprime_opt (utoprim_jon.c:2953) with get_kinetics_part1/2 (:2903,:2932) */
static void func_Eprime(double Wp, double* dx, double* resid, double* jac, double* f, double* df, const inv_t* s)
{
  const double D = s->D, Bsq = s->Bsq;
  const double W = Wp + D;
  const double X = Bsq + W + SMALL;
  const double X2 = X * X, X3 = X2 * X;
  const double utsq = utsq_calc(W, s);
  const double gammasq = 1.0 + utsq;
  const double gamma = sqrt(gammasq);
  const double wmrho0 = wmrho0_utsq(Wp, D, utsq, gamma, gammasq);
  double Eprime, dE;
  if (utsq < -0.9 * s->utsq_too_big || utsq != utsq) {
    Eprime = -VERYBIG;
    dE = 0.0;
    *resid = Eprime;
    *jac = dE;
    *dx = VERYBIG;
    *f = 0.5 * (Eprime * Eprime);
    *df = -2. * (*f);
    return;
  }
  const double igr = (s->gam - 1.0) / s->gam;
  const double pofchi = igr * wmrho0, dpdrho = 0.0, dpdchi = (s->gam - 1.0) / s->gam;
  Eprime = s->Qdotnp + Wp - pofchi + 0.5 * Bsq + 0.5 * (Bsq * s->Qtsq - s->QdotBsq) / X2;
  if (pofchi != pofchi || Eprime != Eprime) Eprime = -VERYBIG;
  /* dvsq_dW, utoprim_jon.c:1819 */
  const double Xd = Bsq + W, Wsq = W * W, W3 = Wsq * W, Xd3 = Xd * Xd * Xd;
  const double dvsq = -2.0 / Xd3 * (s->Qtsq + s->QdotBsq * (3.0 * W * Xd + Bsq * Bsq) / W3);
  const double dwmrho0dW = 1.0 / gammasq;
  const double drho0dW = 0.0;
  const double dwmrho0dvsq = (D * (gamma * 0.5 - 1.0) - Wp);
  const double drho0dvsq = -D * gamma * 0.5;
  const double dpdW = drho0dW * dpdrho + dwmrho0dW * dpdchi;
  const double dpdvsq = drho0dvsq * dpdrho + dwmrho0dvsq * dpdchi;
  const double dpdWp = dpdW + dpdvsq * dvsq;
  dE = (1.0 - dpdWp) - (Bsq * s->Qtsq - s->QdotBsq) / X3;
  if (dvsq != dvsq || dpdW != dpdW || dpdvsq != dpdvsq || dpdWp != dpdWp || dE != dE) dE = 0.0;
  *resid = Eprime;
  *jac = dE;
  *dx = -Eprime / dE;
  *f = 0.5 * (Eprime * Eprime);
  *df = -2. * (*f);
}

/* func_Psq (utoprim_jon.c:3103): the cold GRMHD residual W^2 P^2 of the
 * momentum equation (Psq_Wp :1923) and its W' derivative (dPsqdWp_Wp :1944);
 * neither uses the energy equation. */
static void func_Psq(double Wp, double* dx, double* resid, double* jac, double* f, double* df, const inv_t* s)
{
  const double D = s->D, Bsq = s->Bsq, Qtsq = s->Qtsq;
  const double W = Wp + D;
  const double Y = Bsq + 2.0 * W;
  const double W2 = W * W;
  const double S = s->QdotB;
  const double r = -Qtsq * W2 + Wp * (D + W) * (W2 + Bsq * Y) - Y * S * S;
  const double drdW = 2.0 * (D * ((Bsq + D) * (Bsq + D) - Qtsq) - (S * S) + ((Bsq + D) * (Bsq + 5.0 * D) - Qtsq) * Wp +
                             3.0 * (Bsq + 2.0 * D) * (Wp * Wp) + 2.0 * (Wp * Wp * Wp));
  *resid = r;
  *jac = drdW;
  *dx = -r / drdW;
  *f = 0.5 * r * r;
  *df = -2. * (*f);
}

static void residual(int eom, double Wp, double* dx, double* resid, double* jac, double* f, double* df,
                     const inv_t* s)
{
  if (eom == EOM_COLD) func_Psq(Wp, dx, resid, jac, f, df, s);
  else func_Eprime(Wp, dx, resid, jac, f, df, s);
}

/* newt_errorcheck / newt_extracheck, utoprim_jon.c:4488-4520 */
static int ultrarel(double x0, const inv_t* s) { return (fabs(s->w[1]) > s->w[2]) || (fabs(x0) > s->w[2]); }
static int errcheck(double errx, double x0, const inv_t* s)
{
  return ultrarel(x0, s) ? (fabs(errx) <= NEWT_TOL_ULTRAREL) : (fabs(errx) <= NEWT_TOL);
}
static int extracheck(double x0, const inv_t* s) { return ultrarel(x0, s) ? EXTRA_NEWT_ITER_ULTRAREL : EXTRA_NEWT_ITER; }

/* validate_Wp, utoprim_jon.c:2319 */
static void validate_Wp(double Wpold, double* Wpnew, const inv_t* s)
{
  if (*Wpnew < -s->D && s->D > 0.0) {
    int vf = 0;
    double Wp = 0.0;
    coldgrhd(&vf, s->Qtsq, s->D, &Wp);
    *Wpnew = Wp * 10.0;
  }
  *Wpnew = (fabs(*Wpnew / s->w[1]) > s->gsq_too_big) ? 0.5 * (Wpold + *Wpnew) : *Wpnew;
}

/* general_newton_raphson (utoprim_jon.c:3853), n=1, no line search,
 * JONGENNEWTSTUFF=1, NEWCONVERGE=1.  Returns retval; x updated. */
static int newton(int* pflag, double* xio, const inv_t* s, int* nstroke, int eom)
{
  double x = *xio, x_old, x_older, dx, dx_old, resid, jac, f, df;
  double errx = 1.;
  double damp = 1.0;
  int n_iter = 0, i_extra = 0, doing_extra = 0, keep = 1, diddamp = 0, lntries = 0;
  x_older = x_old = x;
  dx_old = VERYBIG;
  f = df = 1.;
  while (keep) {
    (*nstroke)++;
    lntries++;
    if (lntries > ITERDAMPSTART && diddamp == 0) damp = 0.5;
    residual(eom, x, &dx, &resid, &jac, &f, &df, s);
    errx = 0.;
    x_older = x_old;
    x_old = x;
    diddamp = 0;
    if (n_iter == 0) {
      x_older = x;
      x_old = x;
      dx_old = dx;
    } else {
      if (resid <= -VERYBIG || jac == 0.0) {
        damp = 0.5 * damp;
        x = x_older;
        dx = dx_old;
        diddamp = 1;
      }
      if (diddamp) {
        residual(eom, x, &dx, &resid, &jac, &f, &df, s);
        if (resid <= -VERYBIG || jac == 0.0) { *xio = x; return 1; }
        x_older = x;
        x_old = x;
        dx_old = dx;
      }
    }
    x += damp * dx;
    errx = (s->w[0] == 0.) ? fabs(dx_old) : fabs(dx_old / dmax(fabs(s->w[0]), fabs(x)));
    validate_Wp(x_old, &x, s);
    dx_old = dx;
    if (errcheck(errx, x, s) && (doing_extra == 0) && (extracheck(x, s) > 0)) doing_extra = 1;
    if (doing_extra == 1) i_extra++;
    if ((errcheck(errx, x, s) && (doing_extra == 0)) || (i_extra > extracheck(x, s)) || (n_iter >= (MAX_NEWT_ITER - 1)))
      keep = 0;
    n_iter++;
  }
  if (!isfinite(x) || !isfinite(f) || !isfinite(df)) {
    *pflag = 8;
    *xio = x;
    return 1;
  }
  x = x_old;
  *xio = x;
  if (errcheck(errx, x, s)) return 0;
  if (fabs(errx) > MIN_NEWT_TOL) {
    *pflag = FAILCONVRET + 4;
    return 1;
  }
  return 0;
}

/* Utoprim_new_body (utoprim_jon.c:406): U and prim are the alpha-scaled
 * U_tmp / prim_tmp of the caller; eom EOM_COLD uses the cold EOS (pressure 0,
 * pickeos_eomtype eos.c:259) in the guess and the result. */
static int new_body(int* pflag, const double* U, const o_geom* g, const o_params* P, double* prim, int* nstroke,
                    int eom)
{
  inv_t s;
  double Bcon[4], Bcov[4], Qcovp[4], Qconp[4], Qtcon[4], Qtcov[4], Qcov[4], Qcon[4], ncov[4], ncon[4];
  int retval = 0;
  *pflag = 0;
  s.gam = P->gam;
  s.gsq_too_big = P->gammafail * P->gammafail;
  s.utsq_too_big = (P->gammafail - 1.0) * (P->gammafail - 1.0);
  /* compute_setup_quantities, :716 */
  s.D = U[O_RHO];
  for (int i = O_B1; i <= O_B3; i++) prim[i] = U[i];
  Bcon[0] = 0.;
  for (int i = 1; i < 4; i++) Bcon[i] = U[O_B1 + i - 1];
  for (int a = 0; a < 4; a++) Bcov[a] = g->gcov[a][0] * Bcon[0] + g->gcov[a][1] * Bcon[1] + g->gcov[a][2] * Bcon[2] + g->gcov[a][3] * Bcon[3];
  s.Bsq = 0.;
  for (int i = 1; i < 4; i++) s.Bsq += Bcon[i] * Bcov[i];
  ncov[0] = -g->alpha; ncov[1] = ncov[2] = ncov[3] = 0.;
  for (int a = 0; a < 4; a++) ncon[a] = g->gcon[a][0] * ncov[0] + g->gcon[a][1] * ncov[1] + g->gcon[a][2] * ncov[2] + g->gcon[a][3] * ncov[3];
  for (int i = 0; i < 4; i++) Qcovp[i] = U[O_UU + i];
  for (int a = 0; a < 4; a++) Qconp[a] = g->gcon[a][0] * Qcovp[0] + g->gcon[a][1] * Qcovp[1] + g->gcon[a][2] * Qcovp[2] + g->gcon[a][3] * Qcovp[3];
  for (int j = 0; j < 4; j++) Qtcon[j] = 0.0;
  for (int j = 1; j < 4; j++)
    for (int i = 1; i < 4; i++) Qtcon[j] += U[O_UU + i] * (g->gcon[i][j] + ncon[i] * ncon[j]);
  for (int a = 0; a < 4; a++) Qtcov[a] = g->gcov[a][0] * Qtcon[0] + g->gcov[a][1] * Qtcon[1] + g->gcov[a][2] * Qtcon[2] + g->gcov[a][3] * Qtcon[3];
  s.Qtsq = 0.;
  for (int i = 0; i < 4; i++) s.Qtsq += Qtcov[i] * Qtcon[i];
  for (int i = 0; i < 4; i++) Qcov[i] = (i == 0) ? Qcovp[i] - s.D : Qcovp[i];
  for (int a = 0; a < 4; a++) Qcon[a] = g->gcon[a][0] * Qcov[0] + g->gcon[a][1] * Qcov[1] + g->gcon[a][2] * Qcov[2] + g->gcon[a][3] * Qcov[3];
  s.QdotB = 0.;
  for (int i = 1; i < 4; i++) s.QdotB += Qcov[i] * Bcon[i];
  s.QdotBsq = s.QdotB * s.QdotB;
  {
    const double alpha = g->alpha, alphasq = alpha * alpha;
    const double Dfactor = (-g->gpert_tt + alphasq * (g->bsqoa2)) / (alphasq + alpha);
    s.Qdotnp = Qconp[0] * ncov[0] + (s.D) * (Dfactor);
  }
  s.Qdotn = Qcon[0] * ncov[0];

  /* set_guess_Wp, :890 */
  double Wp_last, W_last;
  {
    double utsq = 0.;
    for (int i = 1; i < 4; i++)
      for (int j = 1; j < 4; j++) utsq += g->gcov[i][j] * prim[O_U1 + i - 1] * prim[O_U1 + j - 1];
    if ((utsq < 0.) && (fabs(utsq) < 0.9 * s.utsq_too_big)) utsq = 0.0;
    if (!(utsq >= 0. && utsq < s.utsq_too_big)) utsq = 0.0;
    const double gammasq = 1. + utsq;
    const double gamma = sqrt(gammasq);
    const double rho0 = (s.D) / gamma;
    const double u = prim[O_UU];
    const double p = eom == EOM_COLD ? 0.0 : (P->gam - 1.0) * u;
    double BBcon[4], BBcov[4], bsq = 0.0;
    BBcon[0] = 0.0;
    for (int jj = 1; jj < 4; jj++) BBcon[jj] = prim[O_B1 - 1 + jj];
    for (int a = 0; a < 4; a++) BBcov[a] = g->gcov[a][0] * BBcon[0] + g->gcov[a][1] * BBcon[1] + g->gcov[a][2] * BBcon[2] + g->gcov[a][3] * BBcon[3];
    for (int jj = 1; jj < 4; jj++) bsq += BBcon[jj] * BBcov[jj];
    bsq *= ((g->alpha * g->alpha)) / gammasq;
    Wp_last = (s.D) * utsq / (1.0 + gamma) + (u + p) * gammasq;
    W_last = Wp_last + (s.D);
    const double Wpabs = fabs((s.D) * utsq) / (1.0 + fabs(gamma)) + (fabs(u) + fabs(p)) * fabs(gammasq) + SQRTMINNUMREPRESENT;
    const double nuabs = fabs(u) + fabs(p) + fabs(bsq);
    const double etaabs = fabs(rho0) + nuabs;
    s.w[2] = GAMMASQCHECKRESID * fabs(rho0);
    if (gammasq * etaabs >= (s.w[2])) {
      double Wpcoldhd = 0.0;
      if (coldgrhd(pflag, s.Qtsq, s.D, &Wpcoldhd)) Wpcoldhd = Wp_last;
      s.w[0] = dmax(dmax(Wpcoldhd, Wp_last), Wpabs);
      Wp_last = dmax(2.0 * Wpcoldhd, Wp_last);
    } else {
      s.w[0] = Wpabs;
    }
    s.w[1] = fabs(rho0) + s.w[0] + SQRTMINNUMREPRESENT;
    /* verify_Wlast, :1120 */
    if (W_last < -(s.D) || Wp_last < 0) { W_last = -(s.D); Wp_last = 0; }
    int i_increase = 0;
    while (vsqgtr1(W_last, &s) && (i_increase < 10)) {
      Wp_last *= 10.;
      W_last = Wp_last + (s.D);
      i_increase++;
    }
    int nattempt = 0;
    while (1) {
      const double ut2 = utsq_calc(W_last, &s);
      if (ut2 >= 0.0 && ut2 == ut2) break;
      int cpf = 0;
      double coldWp = 0.0;
      coldgrhd(&cpf, s.Qtsq, s.D, &coldWp);
      if (Wp_last < coldWp) Wp_last = coldWp;
      if (nattempt > 0) Wp_last = dmax(dmax(fabs(Wp_last) * 10.0, NUMEPSILON * 100.0 * fabs(s.D)), KINDASMALL);
      W_last = Wp_last + (s.D);
      if (nattempt > MAXNUMGUESSCHANGES) { *pflag = 4; return 1; }
      nattempt++;
    }
  }

  /* find_root_1D_gen_Eprime, :3665 */
  double Wp = Wp_last;
  retval += newton(pflag, &Wp, &s, nstroke, eom);
  if (retval != 0) return retval;

  /* check_Wp, :1188: every branch returns retval (= 0 here), so it only
   * sets flags and Wp2prim still runs */
  if (Wp == 1.e100) {
    *pflag = retval + FAILCONVRET + 1;
  } else if (Wp <= -(s.D)) {
    *pflag = retval + FAILCONVRET + 1;
  } else if (Wp / s.w[1] > s.gsq_too_big) {
    *pflag = 2;
  }

  /* Wp2prim, :1346 (PRIMFROMVSQ=0, UTOPRIMRETURNADJUSTED) */
  {
    double pr0[O_NPR];
    for (int i = O_B1; i <= O_B3; i++) prim[i] = U[i];
    for (int pl = 0; pl < O_NPR; pl++) pr0[pl] = prim[pl];
    const double W = Wp + s.D;
    const double utsq = utsq_calc(W, &s);
    const double gammasq = 1.0 + utsq;
    const double gamma = sqrt(gammasq);
    const double wmrho0 = wmrho0_utsq(Wp, s.D, utsq, gamma, gammasq);
    const double rho0 = s.D / gamma;
    const double p = eom == EOM_COLD ? 0.0 : ((P->gam - 1.0) / P->gam) * wmrho0;
    const double u = wmrho0 - p;
    if ((rho0 <= 0.) || ((p < 0.) && eom != EOM_COLD)) {
      if ((rho0 <= 0.) && (p >= 0.)) *pflag = 9;
      if ((rho0 > 0.) && (p < 0.)) *pflag = 10;
      if ((rho0 <= 0.) && (p < 0.)) *pflag = 11;
    }
    prim[O_RHO] = rho0;
    prim[O_UU] = eom == EOM_COLD ? 0.0 : u;
    for (int i = 1; i < 4; i++) prim[O_U1 + i - 1] = gamma / (W + s.Bsq) * (Qtcon[i] + s.QdotB * Bcon[i] / W);
    int foundnan = 0;
    for (int pl = 0; pl < O_NPR; pl++)
      if (!isfinite(prim[pl])) foundnan = 1;
    if (foundnan) {
      *pflag = 8;
      for (int pl = 0; pl < O_NPR; pl++) prim[pl] = pr0[pl];
      return 1;
    }
  }
  return 0;
}

/* Utoprim_jon_nonrelcompat_inputnorestmass (utoprim_jon.c:136) on
 * geometry-free U. */
static int jon(const double* Ugf, const o_geom* g, const o_params* P, double* pr, int* nstroke, int eom)
{
  double Utmp[O_NPR], ptmp[O_NPR];
  int pflag = 0;
  for (int i = O_B1; i <= O_B3; i++) pr[i] = Ugf[i];
  const double alpha = g->alpha;
  for (int i = 0; i < O_NPR; i++) Utmp[i] = alpha * Ugf[i];
  for (int i = 0; i < O_NPR; i++) ptmp[i] = pr[i];
  for (int i = O_B1; i <= O_B3; i++) ptmp[i] = alpha * pr[i];
  const int ret = new_body(&pflag, Utmp, g, P, ptmp, nstroke, eom);
  if (ret == 0) {
    for (int i = 0; i <= O_U3; i++) pr[i] = ptmp[i];
  } else {
    if (pflag <= 0) pflag = ret + FAILCONVRET + 1;
  }
  return pflag;
}

#define THIRD_L 0.3333333333333333333333333333333333L /* THIRD, global.nondepnmemonics.h:746 (long double) */
#define CHECKONINVFRACFAIL 1E-1

/* check_on_inversion (utoprimgen.c:697-850) after the hot solve with
 * FAILIFBADCHECK=1 (:18): a converged zone whose U(p) differs from the input
 * by more than 10% in any of rho, u, the momenta or the fields (each against
 * its own norm, momenta made comparable through u^i/u^t) gets
 * UTOPRIMFAILCONVBADINVERTCOMPARE.  Uold: the geometry-free input. */
static int check_on_inversion(int flag, const double* Uold, const double* pr, const o_geom* g, const o_params* P)
{
  if (flag != 0) return flag;
  o_state q;
  o_geom G1 = *g;
  G1.gdet = 1.0; /* primtoU(UNOTHING) */
  double Unew[O_NPR], on[O_NPR], nn[O_NPR], fdiff[O_NPR];
  o_get_state(pr, g, P->gam, &q);
  o_primtoflux(pr, &q, 0, &G1, Unew);
  for (int pl = 0; pl < O_NPR; pl++) { on[pl] = Uold[pl]; nn[pl] = Unew[pl]; }
  for (int jj = 1; jj <= 3; jj++) {  /* REMOVERESTMASSFROMUU==2 */
    on[O_UU + jj] = Uold[O_UU + jj] * (q.ucon[jj] / q.ucon[0]);
    nn[O_UU + jj] = Unew[O_UU + jj] * (q.ucon[jj] / q.ucon[0]);
  }
  for (int pl = 0; pl < O_NPR; pl++) {
    fdiff[pl] = 0.0;
    if (pl == O_RHO || pl == O_UU) {
      fdiff[pl] = fabs(nn[pl] - on[pl]) / (fabs(nn[pl]) + fabs(on[pl]) + SMALL);
    } else if (pl >= O_U1 && pl <= O_U3) {
      /* the norm sums the U2 terms twice and leaves U1 out, as the reference does */
      double errornorm = (double)(THIRD_L * (long double)(fabs(nn[O_U2]) + fabs(on[O_U2]) + fabs(nn[O_U2]) + fabs(on[O_U2]) +
                                                         fabs(nn[O_U3]) + fabs(on[O_U3])));
      errornorm = dmax(errornorm, 0.5 * (fabs(on[O_UU]) + fabs(nn[O_UU])));
      fdiff[pl] = fabs(nn[pl] - on[pl]) / (errornorm + SMALL);
    } else {
      /* long double throughout, as THIRD makes the reference's expression */
      fdiff[pl] = (double)((long double)fabs(nn[pl] - on[pl]) /
                           (THIRD_L * (long double)(fabs(nn[O_B1]) + fabs(on[O_B1]) + fabs(nn[O_B2]) + fabs(on[O_B2]) +
                                                    fabs(nn[O_B3]) + fabs(on[O_B3])) + (long double)SMALL));
    }
  }
  int bad = 0;
  for (int pl = 0; pl < O_NPR; pl++)
    if (fdiff[pl] > CHECKONINVFRACFAIL && fabs(on[pl]) > SMALL && fabs(nn[pl]) > SMALL) bad++;
  return bad ? O_FAIL_BADINVERTCOMPARE : flag;
}

/* Utoprimgen's EOMGRMHD branch (utoprimgen.c:188-263) on geometry-free U:
 * the hot solve from the guess pr0, its check, and on any failure the cold
 * solve from the same guess (trycoldinversion :597, every failure type
 * retried: USECOLDIFHOT* = 1 :587-591).  A converged cold solve keeps the
 * cold density and velocities, puts u back to the guess's and flags the zone
 * UTOPRIMFAILU2AVG1 so that failure averaging recomputes u alone (:655-660);
 * otherwise the hot result and flag stand.  Newton strokes of both solves add
 * to *nstroke. */
static int utoprimgen_gf(const double* Ugf, const o_geom* g, const o_params* P, double* pr, int* nstroke)
{
  double pr0[O_NPR];
  for (int pl = 0; pl < O_NPR; pl++) pr0[pl] = pr[pl];
  int fl = jon(Ugf, g, P, pr, nstroke, EOM_HOT);
  if (P->checkinv) fl = check_on_inversion(fl, Ugf, pr, g, P);
  if (P->hot2cold && fl > 0) {
    double prcold[O_NPR];
    for (int pl = 0; pl < O_NPR; pl++) prcold[pl] = pr0[pl];
    const int cf = jon(Ugf, g, P, prcold, nstroke, EOM_COLD);
    if (cf >= O_FAIL_FIXEDCOLD && cf <= 0) { /* IFUTOPRIMNOFAILORFIXED */
      for (int pl = 0; pl < O_NPR; pl++) pr[pl] = prcold[pl];
      pr[O_UU] = pr0[O_UU];
      fl = O_FAIL_U2AVG1;
    }
  }
  return fl;
}

int o_utoprimgen_unothing(const double* Ugf, const o_geom* g, const o_params* P, double* pr, int* nstroke)
{
  return utoprimgen_gf(Ugf, g, P, pr, nstroke);
}

int o_utoprimgen(const double* U, const o_geom* g, const o_params* P, double* pr, int* nstroke)
{
  double Ugf[O_NPR];
  /* UtoU(UEVOLVE -> UNOTHING): x igdetnosing = sign(g)/(|g|+SMALL) */
  const double ig = (g->gdet >= 0.0 ? 1.0 : -1.0) / (fabs(g->gdet) + SMALL);
  for (int pl = 0; pl < O_NPR; pl++) Ugf[pl] = U[pl] * ig;
  return utoprimgen_gf(Ugf, g, P, pr, nstroke);
}

/* check_on_inversion alone, for the tests: geomrec is one 27-field record */
int o_check_inversion(int flag, const double* Ugf, const double* pr, const double* geomrec, double gam)
{
  o_geom g;
  o_geom_from_rec(geomrec, &g);
  o_params P = {gam, 100.0, 1.0, 0, 1};
  return check_on_inversion(flag, Ugf, pr, &g, &P);
}

/* one solve of the given equations alone (no retry, no check): the building
 * block above, exposed so the tests can pin the cold solve by itself */
int o_jon_solve(const double* Ugf, const o_geom* g, double gam, double gammafail, int cold, double* pr, int* nstroke)
{
  o_params P = {gam, gammafail, gammafail / 100.0, 0, 0};
  return jon(Ugf, g, &P, pr, nstroke, cold ? EOM_COLD : EOM_HOT);
}

void o_invert_batch(const double* U, const double* geomrec, double* p, int32_t* pflag, int32_t* niter,
                    int64_t n, double gam, double gammafail, int nthreads)
{
  o_params P = {gam, gammafail, gammafail / 100.0, 1, 0}; /* the torus configuration: HOT2COLD on */
  (void)nthreads;
#pragma omp parallel for schedule(dynamic, 1024) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t z = 0; z < n; z++) {
    double rec[27], Uz[O_NPR], pz[O_NPR];
    o_geom g;
    for (int f = 0; f < 27; f++) rec[f] = geomrec[f * n + z];
    o_geom_from_rec(rec, &g);
    for (int pl = 0; pl < O_NPR; pl++) { Uz[pl] = U[pl * n + z]; pz[pl] = p[pl * n + z]; }
    int ns = 0;
    const int fl = o_utoprimgen(Uz, &g, &P, pz, &ns);
    for (int pl = 0; pl < O_NPR; pl++) p[pl * n + z] = pz[pl];
    pflag[z] = fl;
    if (niter) niter[z] = ns;
  }
}
