/*
 * o_grid.h -- oracle restatement of HARM's advance() substep on a grid.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Arrays are HARM layout
 * a[((k*ns2+j)*ns1+i)*8+pl] over the ghosted storage box.
 */
#ifndef HARM_O_GRID_H
#define HARM_O_GRID_H

#include <stdint.h>
#include "oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct o_grid {
  int n1, n2, n3, ng;
  int recon, flux, timeorder;
  int nthreads;
  double gam, cour, safe, gammamax, gammafail;
  double dx[4];      /* dX1..dX3 (index 1..3)                      */
  double startx[4];  /* X of the left edge of zone 0               */
  double rhomin, uumin;
  /* geometry, AoS per 2-D point: (j*ns1+i)*27 + f */
  const double* geom[4]; /* CENT, FACE1, FACE2, FACE3 */
  const double* gdet_corn3; /* (j*ns1+i) */
  const double* conn;       /* (j*ns1+i)*40 + c */
  const double* rcent;      /* r at zone centre per i (ns1) */
  int problem;              /* 0 torus, 1 Bondi: the boundary-condition set */
  double R0;                /* r = R0 + exp(X1) */
  double* ener;             /* energy-file accumulators (O_EN_*), or NULL: o_step adds to them */
  int hot2cold;             /* HOT2COLD: cold retry of failed hot inversions                 */
  int checkinv;             /* CHECKONINVERSION with FAILIFBADCHECK                          */
  int resetregion;          /* USERRESETREGION: fixup_nogood's reset region                  */
  int64_t* failcount;       /* or NULL: the final substeps' inversions add the zones they
                               leave failed ([0]) and of those the ones with neither a
                               converged hot nor a converged cold solve ([1])            */
} o_grid;

/* Energy-file accumulators (dump_ener.c:391-407): boundary flux sums of the
 * last substep and their time integral, six boundaries (X1UP, X1DN, X2UP,
 * X2DN, X3UP, X3DN) by NPR; fix-up changes by category (fladdterms,
 * NUMFAILFLOORFLAGS = 12 by NPR, global.nondepnmemonics.h:801-814); the
 * sampled source; the source sample's weight, the time gathered towards the
 * next sample, and the step count of the next step (nstep, advance.c:254). */
enum { O_EN_PDOT = 0, O_EN_PCUM = 48, O_EN_FLTERMS = 96, O_EN_SRCADD = 192, O_EN_DT4 = 200, O_EN_WILLBE = 201,
       O_EN_NSTEP = 202, O_EN_N = 256 };
/* COUNT* categories (global.nondepnmemonics.h:803-809) */
enum { O_FL_CONV = 0, O_FL_FLOOR = 1, O_FL_GAMMA = 2, O_FL_RHONEG = 4, O_FL_UNEG = 5, O_FL_RHOUNEG = 6 };

/* storage helpers */
void o_dims(const o_grid* g, int* ng1, int* ng2, int* ng3, int* ns1, int* ns2, int* ns3);

/* bound_prim_user_dir for dirs 1..3 (bounds.fishmon.c:15, bounds.tools.c) */
void o_bound(const o_grid* g, double* p);
/* fluxcalc + flux_ct (flux.c:38, fluxct.c:186); ndt[1..3] */
void o_fluxcalc(const o_grid* g, const double* pb, double* F1, double* F2, double* F3, double* ndt, int do_ct);
/* advance_standard update loop (advance.c:420-564) */
void o_update(const o_grid* g, const double* pi, const double* pb, const double* F1, const double* F2,
              const double* F3, double* ui, double* uf, double* ucum, const double* CUf, const double* Cunew,
              double dt, int timeorder, double* accdt);
/* pf = pb; invert utoinvert into pf over active zones; fixup1zone on finalstep */
void o_invert(const o_grid* g, const double* pb, const double* U, double* pf, int32_t* pflag, int finalstep,
              int64_t* nstroke);
/* bound_pflag: one ghost layer of the flag array pflag (storage box) */
void o_bound_pflag(const o_grid* g, int32_t* pflag);
/* post_fixup -> fixup_utoprim with general_average and fixup_nogood; bounds
 * the flags first and marks accounted zones UTOPRIMFAILFIXEDUTOPRIM */
void o_post_fixup(const o_grid* g, double* pf, int32_t* pflag, int finalstep);
/* one full RK step (step_ch_simplempi + post_advance + post_stepch).
 * Arrays: p (in/out), pk0, pk1, ui, uf, ucum, F1..F3 scratch, pflag.
 * *dt in/out; fluxacc[4] accumulates boundary fluxes (may be NULL). */
void o_step(const o_grid* g, double* p, double* pk0, double* pk1, double* ui, double* uf, double* ucum, double* F1,
            double* F2, double* F3, int32_t* pflag, double* dt, double* fluxacc);
/* totals over active zones (constotal dump_ener.c:882, divbmaxavg :680):
 * out[0..7] = sum U(p) dV, out[8] = max |div B|, out[9] = sum |div B| */
void o_ener_totals(const o_grid* g, const double* p, double* out);
/* The main dump's four-current (current_precalc phys.c:622, current_calc_1
 * :846; CURRENTCALC1): cf = 18 doubles per storage zone (6 slots x 3 Faraday
 * samples); jcon = 4 per active zone. */
void o_current_precalc(const o_grid* g, const double* p, double* cf);
void o_current_calc(const o_grid* g, const double* cf, int timeterms, double dtl, double dtr, double* jcon);
/* dump_content's physics columns for active zone (i,j,k): 55 values, the
 * dump's columns 10-52 and 61-72 (dump.c:439-600) */
void o_dump_row(const o_grid* g, const double* p, const double* ucum, int i, int j, int k, double* out);

/* WENO5BND centre-to-edge reconstruction of one storage line of one primitive
 * (o_weno.c): y[0..n-1] the line, ng its ghost depth (>= 4), dP the line's
 * pressure-jump measure; yl/yr the left/right face values of cells ng-1 ..
 * n-ng */
void o_weno5bnd_line(const double* y, int n, int ng, const double* dP, double* yl, double* yr);
/* SMONO monotonicity indicator of the five values y[-2..2] */
double o_smono_indicator(const double* y, double eps);
/* per-cell terms of the dP line (rho, par, veff, upar) and the line itself */
void o_dp_cell(const double* pr, const o_geom* G, double gam, int dir, double* out4);
void o_dp_line(const double* st, int n, double* dP);
/* one cell's (left, right) edge values from y5[0..4] = y[-2..2] (recon 0 donor,
 * 2 PARA, 3 WENO5), and one face's Riemann flux (flux 0 LAX-F, 1 HLL) from
 * two states at a 27-field geometry record; p_l, p_r are modified as
 * getplpr/check_plpr do (B^dir averaged, gamma limited). */
void o_recon_cell(int recon, const double* y5, double* l, double* r);
void o_face_flux(int flux, int dir, double gam, double gammamax, const double* rec, double* p_l, double* p_r,
                 double* F, double* ctop);
/* RK coefficient tables (step_ch.c:1171 setup_rktimestep) */
int o_rk_tables(int timeorder, double CUf[4][4], double Cunew[4][4]);

#ifdef __cplusplus
}
#endif
#endif
