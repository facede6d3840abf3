/*
 * oracle.h -- CPU restatement of HARM's advance() hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker the HIP path is compared
 * against (and, timed, the CPU baseline); nothing in the shipped library calls
 * it.  Every function cites the reference (pseudotensor/harm_harm) file:line
 * whose algorithm it restates.  Arrays are HARM layout: a[((k*ns2+j)*ns1+i)*8+pl]
 * over the ghosted storage box; geometry is the library's per-point record
 * (27 doubles) on the 2-D (i,j) storage plane.
 *
 * Pinning (tests/test_oracle_phys.py, tests/test_oracle_grid.py,
 * tests/test_oracle_weno.py, tests/test_oracle_fixup.py,
 * tests/test_oracle_current.py, tests/test_connection_closed_form.py):
 * per-zone state, fluxes, wave
 * speeds and the inversion against textbook GRMHD and round trips; the
 * WENO5BND line reconstruction against polynomial exactness, measured
 * fifth-order convergence, the SMONO derivative stencils against numpy's
 * polynomial derivatives, a hand-derived stencil-reduction case and the dP
 * ramp; the face Riemann solver, RK tables, ghost zones (the pole-death
 * velocity against its closed form), flux-CT and the update/source balance
 * against known answers and invariants; the connection table against the
 * closed-form MKS Kerr-Schild connection; the radial outflow extrapolation
 * against power laws (exact on the log-radius grid), constant gdet B^1 and
 * the minmod continuation of B_3; failure averaging against brute-force
 * means of good neighbours; the cold inversion against cold states; the
 * dump's four-current against the Levi-Civita form of F and exact
 * differences of linear and quadratic samples.
 * Parity unpinned: the reference's uninitialised reads in the WENO dP line
 * (o_weno.c states the choice made) cannot be pinned against anything.
 */
#ifndef HARM_ORACLE_H
#define HARM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define O_NPR 8
enum { O_RHO = 0, O_UU = 1, O_U1 = 2, O_U2 = 3, O_U3 = 4, O_B1 = 5, O_B2 = 6, O_B3 = 7 };

typedef struct o_geom {
  double gcov[4][4], gcon[4][4];
  double gpert_tt, gdet, alpha, bsqoa2, beta[4];
} o_geom;

typedef struct o_state {
  double ucon[4], ucov[4];
  double gamma, qsq, pressure, bsq, p1ud0;
} o_state;

/* physical parameters the per-zone kernels need */
typedef struct o_params {
  double gam;        /* adiabatic index                 */
  double gammafail;  /* GAMMAFAIL (inversion)           */
  double gammamax;   /* GAMMAMAX (limit_gamma)          */
  int hot2cold;      /* HOT2COLD: cold retry of a failed hot solve      */
  int checkinv;      /* CHECKONINVERSION with FAILIFBADCHECK           */
} o_params;

/* inversion flags beyond 0..11 (global.nondepnmemonics.h:824-844) */
enum { O_FAIL_FIXEDCOLD = -3, O_FAIL_BADINVERTCOMPARE = 6, O_FAIL_U2AVG1 = 14, O_FAIL_U2AVG2 = 15,
       O_NANPFLAG = -100 };

void o_geom_from_rec(const double* rec, o_geom* g);

/* phys.tools.c:1513 get_state with VELREL4, UD0PLUS1FIX, rel4vel bsq */
void o_get_state(const double* pr, const o_geom* g, double gam, o_state* q);
/* phys.tools.c:30 primtoflux(UEVOLVE, ...) -> gdet * flux */
void o_primtoflux(const double* pr, const o_state* q, int dir, const o_geom* g, double* F);
/* phys.tools.c:1047 mhd_calc_norestmass (MA + primitive-field EM) */
void o_mhd_calc(const double* pr, int dir, const o_geom* g, const o_state* q, double* mhd);
/* vchar.c:34 vchar + simplefast :260 */
void o_vchar(const double* pr, const o_state* q, int dir, const o_geom* g, double gam, double* vmax, double* vmin);
/* phys.tools.c:1294 source -> phys.c:132 source_conn, conn packed a*10+sym(b,c) */
void o_source(const double* pr, const o_geom* g, const o_state* q, const double* conn, double* dU);

/* utoprimgen.c:32 Utoprimgen(UEVOLVE) -> utoprim_jon.c:136; returns pflag,
 * *nstroke gets the Newton iterations (of_newtonstats.nstroke). pr is the
 * guess on entry, the solution on exit. */
int o_utoprimgen(const double* U, const o_geom* g, const o_params* P, double* pr, int* nstroke);
/* same with geometry-free U (inputtype UNOTHING), as fixup1zone uses */
int o_utoprimgen_unothing(const double* Ugf, const o_geom* g, const o_params* P, double* pr, int* nstroke);
/* one hot (cold = 0) or cold (cold = 1) solve of geometry-free U from the
 * guess pr, without the retry or the check */
int o_jon_solve(const double* Ugf, const o_geom* g, double gam, double gammafail, int cold, double* pr, int* nstroke);
/* check_on_inversion (utoprimgen.c:697) of a solution pr of geometry-free Ugf
 * at one 27-field geometry record; returns the (possibly changed) flag */
int o_check_inversion(int flag, const double* Ugf, const double* pr, const double* geomrec, double gam);

/* batched: SoA like harm_invert_batch */
void o_invert_batch(const double* U, const double* geomrec, double* p, int32_t* pflag, int32_t* niter,
                    int64_t n, double gam, double gammafail, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
