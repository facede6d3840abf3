/*
 * harmstep.h -- C ABI of the MI355X-native HARM advance() step.
 *
 * One handle owns one grid resident in HBM (structure-of-arrays, i fastest,
 * ghost width ng on every dimension with more than one zone).  The host passes
 * HARM-layout arrays (array-of-structs, p[k][j][i][NPR] over the full ghosted
 * box) only through the set/get calls; every compute entry point works on
 * device-resident state and launches hand-written gfx950 FP64 kernels on the
 * handle's stream.  No entry point exits the process: each returns a
 * harm_status, and per-zone inversion failures land in an int32 pflag array.
 *
 * Every compute entry point names the reference function it replaces
 * (HARM, pseudotensor/harm_harm; file:line of the reference tree).
 */
#ifndef HARMSTEP_H
#define HARMSTEP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- primitive / conserved indices (HARM's NPR=8 ordering, definit.h) ---- */
enum { HARM_RHO = 0, HARM_UU = 1, HARM_U1 = 2, HARM_U2 = 3, HARM_U3 = 4,
       HARM_B1 = 5, HARM_B2 = 6, HARM_B3 = 7, HARM_NPR = 8 };

/* ---- status codes ---- */
typedef enum {
  HARM_OK = 0,
  HARM_EINVAL = 1,     /* bad argument / size / configuration            */
  HARM_EHIP = 2,       /* a HIP runtime call failed                      */
  HARM_ENOMEM = 3,     /* device allocation failed                       */
  HARM_ENOGPU = 4,     /* no usable gfx950 device                        */
  HARM_EFAIL = 5       /* numerical failure that HARM would treat as fatal */
} harm_status;

/* ---- per-zone inversion flags (global.nondepnmemonics.h:823-844) ---- */
enum {
  HARM_NANPFLAG = -100,                 /* a flag never set (ghost zones the flag bounds skip) */
  HARM_UTOPRIMFAILFIXEDCOLD = -3,
  HARM_UTOPRIMFAILFIXEDUTOPRIM = -1,    /* failure averaging fixed and accounted the zone */
  HARM_UTOPRIMNOFAIL = 0,
  HARM_UTOPRIMFAILCONV = 1,
  HARM_UTOPRIMFAILCONVW = 2,
  HARM_UTOPRIMFAILCONVUTSQ = 3,
  HARM_UTOPRIMFAILCONVGUESSUTSQ = 4,
  HARM_UTOPRIMFAILCONVBADINVERTCOMPARE = 6,
  HARM_UTOPRIMFAILNANRESULT = 8,
  HARM_UTOPRIMFAILRHONEG = 9,
  HARM_UTOPRIMFAILUNEG = 10,
  HARM_UTOPRIMFAILRHOUNEG = 11,
  HARM_UTOPRIMFAILU2AVG1 = 14,          /* the cold retry converged: u to be averaged */
  HARM_UTOPRIMFAILU2AVG2 = 15,
  HARM_UTOPRIMFAILCONVRET = 50
};

/* ---- method selectors (HARM's runtime globals lim[], fluxmethod, TIMEORDER) */
enum { HARM_RECON_DONOR = 0, HARM_RECON_MC = 1, HARM_RECON_PARA = 2, HARM_RECON_WENO5 = 3 };
enum { HARM_FLUX_LAXF = 0, HARM_FLUX_HLL = 1 };

/* ---- per-point geometry record (one of the NPG grid locations) ----------
 * gcov/gcon are packed symmetric (00,01,02,03,11,12,13,22,23,33); the other
 * fields are HARM's of_geom extras (global.structs.h:85-144).              */
enum {
  HARM_G_GCOV = 0, HARM_G_GCON = 10, HARM_G_GCOVPERT_TT = 20, HARM_G_GDET = 21,
  HARM_G_ALPHA = 22, HARM_G_BETASQOALPHASQ = 23, HARM_G_BETA1 = 24,
  HARM_G_BETA2 = 25, HARM_G_BETA3 = 26, HARM_G_NFIELD = 27
};
/* connection Gamma^a_{bc} at CENT, packed a*10 + sym(b,c): 40 values */
enum { HARM_NCONN = 40 };
/* grid locations with geometry tables */
enum { HARM_LOC_CENT = 0, HARM_LOC_FACE1 = 1, HARM_LOC_FACE2 = 2, HARM_LOC_FACE3 = 3, HARM_NLOC = 4 };

/* ---- problems: the boundary-condition sets of the two problem setups ----
 * TORUS: the FM torus (initboundcode/bounds.fishmon.c:15-216): radial outflow
 *        extrapolation with the inflow check, polar reflection with pole
 *        death, periodic phi.
 * BONDI: spherical Bondi accretion (TESTNUMBER 153,
 *        initboundcode/init.sasha.c:865, :1502; bounds.sasha.c:182, :341):
 *        inner BONDIINTOUTFLOW (the Bondi integrals C1, C2, K of the first
 *        active zone carried into the ghost zones), outer BONDIMDOTOUTFLOW
 *        (C1 of the last active zone with the exact C2 and K), plain polar
 *        reflection (POLARAXIS).  Spin 0, Gamma 4/3, 2-D (n3 = 1), and the
 *        radial ghost zones must lie in 0.35 < r < 2.5 (inner) and r > 9
 *        (outer), the ranges of the reference's integral inversion.        */
enum { HARM_PROBLEM_TORUS = 0, HARM_PROBLEM_BONDI = 1 };

/* ---- runtime configuration (HARM's init_global/init_grid globals) ---- */
typedef struct harm_config {
  int32_t n1, n2, n3;       /* active zones (N1,N2,N3)                        */
  int32_t ng;               /* ghost width (MAXBND); 4 for PARA and WENO5     */
  int32_t recon;            /* HARM_RECON_*  (lim[1..3])                      */
  int32_t flux;             /* HARM_FLUX_*   (fluxmethod)                     */
  int32_t timeorder;        /* 2 (RK2 midpoint) or 3 (RK3 TVD)               */
  int32_t problem;          /* HARM_PROBLEM_*: boundary conditions (BCtype[]) */
  double a;                 /* Kerr spin                                      */
  double gam;               /* ideal-gas adiabatic index                      */
  double R0, Rin, Rout;     /* MKS radial map r = R0 + exp(X1)                */
  double hslope;            /* MKS theta concentration                        */
  double cour;              /* Courant factor                                 */
  double safe;              /* dt growth cap (SAFE)                           */
  double rhomin, uumin;     /* floors rho=rhomin r^-1.5, u=uumin r^-2.5 (rescaletype 0) */
  double gammamax;          /* limit_gamma ceiling (GAMMAMAX)                 */
  double gammafail;         /* inversion failure Lorentz factor (GAMMAFAIL)   */
  /* failure handling switches of the problem's header (1 on, 0 off):       */
  int32_t hot2cold;         /* HOT2COLD: a failed hot inversion is retried with
                               the cold GRMHD solve (utoprimgen.c:243-262);
                               torus 1 (init.fishmon.h:325), Bondi 0
                               (init.sasha.h:379)                            */
  int32_t checkinv;         /* CHECKONINVERSION with FAILIFBADCHECK: a converged
                               zone whose U(p) is off by >10% fails
                               (utoprimgen.c:697-850); torus 0
                               (init.fishmon.h:225), Bondi 1 (init.sasha.h:198) */
  int32_t resetregion;      /* USERRESETREGION: fixup_nogood resets the velocity
                               near the poles and the inner radial edge
                               (fixup.c:2330); torus 1 (init.fishmon.h:356),
                               Bondi 0                                       */
} harm_config;

typedef struct harm_handle harm_handle;

/* Newton statistics of one inversion call (of_newtonstats nstroke, summed). */
typedef struct harm_newton_stats {
  int64_t nstroke;          /* total Newton iterations over all zones         */
  int64_t nfail;            /* zones with pflag > 0 after the call            */
} harm_newton_stats;

/* ------------------------------------------------------------------------ */
/* Library / handle lifetime                                                 */
/* ------------------------------------------------------------------------ */

/* Fills cfg with HARM's torus defaults for an n1 x n2 x n3 grid
 * (init.fishmon.c:194-290, init.tools.c:89-263, rescaletype 0 floors). */
void harm_default_config(harm_config* cfg, int32_t n1, int32_t n2, int32_t n3);
/* Fills cfg with the Bondi problem's setup on an n1 x n2 (r, theta) grid
 * (TESTNUMBER 153: init.sasha.c:865-895, :1084-1102, :1222-1228): spin 0,
 * Gamma 4/3, MKS with R0 = 0 and hslope = 1, Rin = 1.9, Rout = 20, Courant
 * 0.5, no floors (RHOMIN = UUMIN = 0), GAMMAMAX 100. */
void harm_bondi_config(harm_config* cfg, int32_t n1, int32_t n2);

/* Replaces init()'s grid/geometry part (initbase.c:7 -> set_grid.c:114):
 * builds MKS Kerr-Schild geometry and connection tables on the host and
 * uploads the 2-D (r,theta) tables; allocates all device state. */
harm_status harm_create(const harm_config* cfg, harm_handle** out);
harm_status harm_destroy(harm_handle* h);
harm_status harm_sync(harm_handle* h);
const char* harm_last_error(void);

/* Sizes of the full ghosted storage box (NSTORE1..3). */
harm_status harm_storage_dims(const harm_handle* h, int32_t dims[3]);

/* Host copies of geometry: loc in HARM_LOC_*, out[(j*ns1+i)*HARM_G_NFIELD + f]
 * over the 2-D storage plane (ns1 x ns2); conn out[(j*ns1+i)*40 + c]. */
harm_status harm_get_geometry(harm_handle* h, int32_t loc, double* out);
harm_status harm_get_connection(harm_handle* h, double* out);
/* gdet at the (i-1/2, j-1/2) corners (CORN3), out[j*ns1+i]; the Toth EMFs are
 * multiplied by the corner gdet (fluxct.c:186, CORNGDETVERSION). */
harm_status harm_get_corner_gdet(harm_handle* h, double* out);
/* MKS coordinates (X1..X3) and BL radius/theta of a zone centre. */
harm_status harm_get_coords(harm_handle* h, int32_t i, int32_t j, int32_t k, double X[4], double V[4]);

/* ------------------------------------------------------------------------ */
/* State exchange in HARM layout: a[((k*ns2 + j)*ns1 + i)*NPR + pl]           */
/* ------------------------------------------------------------------------ */
enum { HARM_ARR_P = 0,     /* pglobal (primitives)                           */
       HARM_ARR_PK0 = 1,   /* pk[0] RK temporary primitives                  */
       HARM_ARR_PK1 = 2,   /* pk[1]                                          */
       HARM_ARR_UI = 3,    /* uinitial                                       */
       HARM_ARR_UF = 4,    /* ulast                                          */
       HARM_ARR_UCUM = 5,  /* unewglobal                                     */
       HARM_ARR_F1 = 6, HARM_ARR_F2 = 7, HARM_ARR_F3 = 8, /* face fluxes     */
       HARM_NARR = 9 };
/* Both copy the whole storage box over PCIe and convert between this layout
 * and the device planes with a kernel (a staging buffer of the box's size is
 * allocated by the first call); both return when the copy is complete. */
harm_status harm_set_array(harm_handle* h, int32_t which, const double* host_aos);
harm_status harm_get_array(harm_handle* h, int32_t which, double* host_aos);
/* pflag[k][j][i] (FLAGUTOPRIMFAIL) over the storage box */
harm_status harm_get_pflag(harm_handle* h, int32_t* host_flags);
/* the clock of step() (t, dt and the step count nstep, cf. global t, dt,
 * nstep); setting it also restarts the Newton sums below */
harm_status harm_set_time(harm_handle* h, double t, double dt, int64_t nstep);
harm_status harm_get_time(harm_handle* h, double* t, double* dt, int64_t* nstep);
/* Newton statistics of the last whole step and summed over the steps since
 * harm_set_time (either pointer may be NULL); the failure counts are the
 * zones left with pflag > 0 by each substep's inversion, as counted into the
 * reference's failure diagnostics (fail.c) */
harm_status harm_get_newton_stats(harm_handle* h, harm_newton_stats* last_step, harm_newton_stats* total);

/* ------------------------------------------------------------------------ */
/* Problem setup (host init, then upload)                                    */
/* ------------------------------------------------------------------------ */
/* FM torus + atmosphere + A_phi field normalised to beta, u perturbed by
 * +-randfact/2 from HARM's ranc(seed) stream (init.fishmon.c:403-750,
 * init.tools.c:295-894).  Leaves p bounded and sets dt from set_dt(). */
harm_status harm_init_torus(harm_handle* h, double rin, double rmax, double kappa,
                            double beta, double randfact, uint32_t seed);
/* Bondi accretion flow through the sonic point r = 8 (the flow of
 * initbounddata/bondigrid<N>.dat, init.sasha.c:2977-2988): rho, u and the
 * radial Kerr-Schild 4-velocity on every active zone, fields zero; leaves p
 * bounded.  The handle must have been created with problem BONDI. */
harm_status harm_init_bondi(harm_handle* h);

/* Host-only forms (no GPU needed): the geometry tables harm_create uploads
 * (geom_aos[loc][j][i][27] for the 4 locations, conn_aos[j][i][40],
 * corner_gdet[j][i], rcent[i], grid origin and spacing in X), and the torus
 * initial data on the active zones of a HARM-layout host array (ghosts zero,
 * unbounded). */
harm_status harm_host_tables(const harm_config* cfg, double* geom_aos, double* conn_aos, double* corner_gdet,
                             double* rcent, double startx[4], double dx[4]);
harm_status harm_host_init_torus(const harm_config* cfg, double rin, double rmax, double kappa, double beta,
                                 double randfact, uint32_t seed, double* p_aos);
harm_status harm_host_init_bondi(const harm_config* cfg, double* p_aos);
/* The Bondi flow at Boyer-Lindquist/Kerr-Schild radius r (spin 0):
 * out = {rho, u, u^r}. */
harm_status harm_bondi_state(double gam, double r, double out[3]);

/* ------------------------------------------------------------------------ */
/* Hot path                                                                  */
/* ------------------------------------------------------------------------ */

/* Replaces fluxcalc() (flux.c:38 -> fluxcalc_flux :658 -> fluxcalc_standard
 * :852): reconstruct arr `which_p`, face Riemann fluxes into F1..F3 and
 * ndt[3] = min over faces of cour*dx/ctop; then flux_ct() (fluxct.c:186,
 * Toth) when do_ct != 0. */
harm_status harm_fluxcalc(harm_handle* h, int32_t which_p, int32_t do_ct, double ndt[3]);

/* Replaces the per-zone update loop of advance_standard (advance.c:420-564):
 * Ui = U(pi); dU = flux2dUavg + source(pb); Uf = UFSET; ucum += UCUMUPDATE;
 * accdt = source dt limit (dUtodt -> compute_dt_fromsource). */
harm_status harm_update(harm_handle* h, int32_t which_pi, int32_t which_pb,
                        const double CUf[4], const double Cunew[4], double dt,
                        int32_t timeorder, double* accdt);

/* Replaces advance.c:654-808: pf = pb (guess), invert (finalstep ? ucum : uf)
 * with Utoprimgen (utoprimgen.c:32 -> utoprim_jon.c:136), then fixup1zone
 * (fixup.c:561) on the final substep. */
harm_status harm_invert_grid(harm_handle* h, int32_t which_pb, int32_t which_pf,
                             int32_t finalstep, harm_newton_stats* stats);

/* Replaces bound_prim_user_dir (initboundcode/bounds.fishmon.c:15) for arr
 * `which`: x1 outflow (inner inflow check), x2 polar axis, x3 periodic. */
harm_status harm_bound(harm_handle* h, int32_t which);

/* Replaces post_fixup (fixup.c:53 -> fixup_utoprim :1061) on arr `which`:
 * bounds the flags' first ghost layer (bound_pflag, boundsint.c:18), then each
 * failed zone gets general_average (fixup.c:1830: the mean of the good
 * neighbours in the 3^d box, or for the cold retry's u-only mark the least u
 * over neighbour pairs that both exceed its own), fixup_nogood (:2291) when
 * there is none, and the Lorentz factor cap; accounted zones end flagged
 * UTOPRIMFAILFIXEDUTOPRIM (-1; the soft failures only when finalstep != 0,
 * as the reference counts them). */
harm_status harm_post_fixup(harm_handle* h, int32_t which, int32_t finalstep);

/* Replaces advance() (advance.c:85) for one RK substep. */
harm_status harm_advance(harm_handle* h, int32_t timeorder, double* ndt);

/* Replaces step_ch_full() (step_ch.c:35 -> step_ch_simplempi :524 ->
 * post_stepch): one full RK step with bounds, fixups and the dt update. */
harm_status harm_step(harm_handle* h, int32_t nsteps);

/* Accretion-rate diagnostic (diag_flux_pureflux, diag.c:1020): F1[RHO] and
 * F1[UU] summed over the inner and outer radial boundary faces times
 * dx2*dx3*dt, accumulated since the last reset. out[4] = {mdot_in, edot_in,
 * mdot_out, edot_out}. */
harm_status harm_get_boundary_fluxes(harm_handle* h, double out[4]);

/* Fix-up categories of harm_ener.fladdterms (NUMFAILFLOORFLAGS and the
 * COUNT* values, global.nondepnmemonics.h:801-814): failure averaging after
 * a convergence failure, floors, the gamma limit, the inflow check (ghost
 * zones only, so always 0 here), failure averaging after rho < 0, u < 0 and
 * both; the other slots stay 0. */
enum { HARM_FL_CONV = 0, HARM_FL_FLOOR = 1, HARM_FL_GAMMA = 2, HARM_FL_INFLOW = 3, HARM_FL_RHONEG = 4,
       HARM_FL_UNEG = 5, HARM_FL_RHOUNEG = 6, HARM_NUMFAILFLOORFLAGS = 12 };

/* One record of the energy file (dump_ener.c:391-407, the "ener.out" line).
 * Boundaries in the reference's order X1UP, X1DN, X2UP, X2DN, X3UP, X3DN
 * (global.nondepnmemonics.h:699-704); variables in
 * primitive order; U's in the diagnostic form (with gdet, T^t_t + rho u^t in
 * UU) times dV = dx1 dx2 dx3. */
typedef struct harm_ener {
  double t;
  int64_t nstep;
  double U[HARM_NPR];           /* sum of U(p) dV over active zones (constotal, dump_ener.c:882) */
  double pdot[6][HARM_NPR];     /* boundary face sums of F * surface, last substep (diag.c:1020) */
  double pcum[6][HARM_NPR];     /* pdot * the substep's dt share, summed (diag.c:1209) */
  double fladd[HARM_NPR];       /* final-substep changes from floors, gamma limits and
                                   failure averaging (diag_fixup, fixup.c:220): the sum of
                                   fladdterms over its categories */
  double fladdterms[HARM_NUMFAILFLOORFLAGS][HARM_NPR]; /* by category (HARM_FL_*) */
  double sourceadd[HARM_NPR];   /* sampled geometric source dt dV dU (diag_source_all,
                                   diag.c:1958; advance.c:251-262 sampling) */
  double divbmax, divbavg;      /* |div B| at the corners (divbmaxavg, dump_ener.c:680) */
} harm_ener;

/* Replaces the accumulation side of dump_ener()/diag(): with on != 0 every
 * later harm_step / harm_advance of this handle adds to pdot, pcum, fladd
 * and sourceadd (batch steps do not).  Switching from off to on clears the
 * accumulators; a call that finds tracking on already keeps them. */
harm_status harm_set_ener_tracking(harm_handle* h, int32_t on);
/* ---- HARM-format dumps (dump.c:256-600, 983-1060; docs/datadesc.txt) ----
 * Main dump: HARM_DUMP_NCOL columns per active zone, rows with i fastest,
 * then j, then k: ti tj tk, X1..X3 (MKS), r th ph, the 8 primitives, p,
 * cs2, Sden, U[8] (the conserved state of the last inversion), |div B|,
 * u^mu, u_mu, b^mu, b_mu, v1m v1p v2m v2p v3m v3p, gdet, j^mu, j_mu,
 * F^{mu nu} (01,02,03,12,13,23), F_{mu nu}.  The current columns j hold
 * the four-current of the last harm_currents call (0 before the first).
 * Grid dump: HARM_GDUMP_NCOL columns: ti tj tk, X, V, Gamma^a_{bc} (a
 * slowest), g^{ab}, g_{ab}, gdet, conn2 (0 with gdet in the equations),
 * dV^a/dX^b. */
enum { HARM_DUMP_NCOL = 72, HARM_GDUMP_NCOL = 126 };
/* Replaces current_doprecalc()/current_calc() (phys.c:552-1000,
 * CALCFARADAYANDCURRENTS with CURRENTCALC1): with at_start != 0 as
 * compute_currents_t0 does at the start of a run (initbase.c:1414), else as
 * post_stepch does after a step that ends at a main dump (step_ch.c:156-170):
 * the Faraday samples of every zone from HARM_ARR_P, then J^mu (the time
 * derivative over the last three samples from the third call on).  Allocates
 * its arrays on the first call. */
harm_status harm_currents(harm_handle* h, int32_t at_start);
/* Replaces dump_content() over the grid: rows[z * HARM_DUMP_NCOL + c] from
 * primitives `which` (HARM_ARR_*) on the host. */
harm_status harm_dump_columns(harm_handle* h, int32_t which, double* rows);
/* No counterpart in the reference: harm_dump_columns for callers that cannot
 * allocate (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same rows, built in a static device buffer of 65536
 * zones (37.7 MB per device, held until the process ends) instead of
 * allocated memory; grids above 65536 zones get HARM_EINVAL.  The buffer is
 * shared by every call on a device: calls that overlap on one device race. */
harm_status harm_dump_columns_noalloc(harm_handle* h, int32_t which, double* rows);
/* Replaces gdump_content() over the grid: rows[z * HARM_GDUMP_NCOL + c]. */
harm_status harm_gdump_columns(harm_handle* h, double* rows);
/* Replace dump() and gdump() (text output, dump_header_general's header line,
 * "%21.15g " per value, one row per zone). */
harm_status harm_write_dump(harm_handle* h, int32_t which, const char* path);
harm_status harm_write_gdump(harm_handle* h, const char* path);

/* ---- restart files (restart.c:163-285, readwrite_restart_header :593-850) ----
 * One text header line holding, in readwrite_restart_header's order, the
 * part of its list that this library has state for: N1 N2 N3, t, tf, nstep,
 * a, MBH, QBH, gam, gamideal, cour, dt, lim[1..3], TIMEORDER, fluxmethod,
 * defcoord, R0, Rin, Rout, hslope, binaryoutput, then the energy-file
 * accumulators (pdot[6][8], pcum[6][8], fladdterms[12][8], sourceadd[8], the
 * source-sample clock's two sums).  Doubles print as "%26.20g" (FTYPE) or
 * "%31.26g" (SFTYPE), which read back exactly.  Then per active zone, i
 * fastest: the 8 primitives of HARM_ARR_P and the 8 conserved of
 * HARM_ARR_UCUM (rdump_content).  binary = 0 (TEXTOUTPUT) writes each
 * value as "%21.15g ", one zone per line, as the reference does: 15 digits,
 * not an exact restart.  binary = 2 (MIXEDOUTPUT) writes raw native doubles:
 * stepping after harm_read_restart then gives the same bits as stepping on. */
enum { HARM_TEXTOUTPUT = 0, HARM_MIXEDOUTPUT = 2 };
harm_status harm_write_restart(harm_handle* h, const char* path, int32_t binary, double tf);
/* Replaces restart_read(): the grid sizes and the configuration in the
 * header must equal the handle's (HARM_EINVAL naming the field otherwise).
 * Loads the active zones, fills the ghost zones with the problem's boundary
 * conditions, sets t, dt and nstep, restores the energy accumulators (switch
 * tracking on before the call: switching on clears them) and starts the
 * flags as a new run's (rdump holds none).  tf may be NULL. */
harm_status harm_read_restart(harm_handle* h, const char* path, double* tf);

/* Clears pdot, pcum, fladd, sourceadd and the source-sample clock. */
harm_status harm_reset_ener(harm_handle* h);
/* Replaces the record-forming side of dump_ener(): totals and div B from
 * array `which` (HARM_ARR_*, normally HARM_ARR_P), with the accumulators and
 * the step clock. */
harm_status harm_get_ener(harm_handle* h, int32_t which, harm_ener* out);

/* ------------------------------------------------------------------------ */
/* Device memory, streams and timing (so callers need no other HIP runtime)  */
/* ------------------------------------------------------------------------ */
harm_status harm_device_count(int32_t* n);
harm_status harm_dev_malloc(void** ptr, size_t bytes);
harm_status harm_dev_free(void* ptr);
harm_status harm_memcpy_htod(void* dst, const void* src, size_t bytes);
harm_status harm_memcpy_dtoh(void* dst, const void* src, size_t bytes);
harm_status harm_memcpy_dtod(void* dst, const void* src, size_t bytes);
harm_status harm_device_synchronize(void);
/* the handle's stream (kernels of h run there) */
harm_status harm_get_stream(harm_handle* h, void** stream);
harm_status harm_event_create(void** ev);
harm_status harm_event_destroy(void* ev);
harm_status harm_event_record(void* ev, void* stream);
harm_status harm_event_elapsed_ms(void* ev0, void* ev1, float* ms);

/* Kernel timing with HIP events recorded on the handle's stream around each
 * kernel group while enabled; read back (after a sync) as total ms and launch
 * count per group.  Groups: */
enum { HARM_K_FLUX1 = 0, HARM_K_FLUX2 = 1, HARM_K_FLUX3 = 2, HARM_K_EMFCT = 3, HARM_K_UPDATE = 4,
       HARM_K_INVERT = 5, HARM_K_BOUND = 6, HARM_K_FIXUP = 7, HARM_K_OTHER = 8, HARM_K_RECON = 9, HARM_K_N = 10 };
harm_status harm_profile(harm_handle* h, int32_t enable);
harm_status harm_profile_read(harm_handle* h, double* ms_total, int64_t* count);

/* ------------------------------------------------------------------------ */
/* Batched inversion on caller-provided device buffers (config-3 microbench) */
/* ------------------------------------------------------------------------ */
/* Replaces the per-zone Utoprimgen call (utoprimgen.c:32) for n independent
 * zones.  Device pointers, SoA: U[pl*n + z] (UEVOLVE, gdet-weighted),
 * geom[f*n + z] (HARM_G_NFIELD fields), p_inout[pl*n + z] (guess in,
 * solution out), pflag[z], niter[z] (may be NULL).  stream may be NULL. */
harm_status harm_invert_batch(const double* U, const double* geom, double* p_inout,
                              int32_t* pflag, int32_t* niter, int64_t n, double gam,
                              double gammafail, void* stream);

/* Replaces primtoU(UEVOLVE, ...) (phys.tools.c:475) over n independent zones:
 * device SoA p[pl*n+z], geom[f*n+z] -> U[pl*n+z]. */
harm_status harm_prim_to_cons_batch(const double* p, const double* geom, double* U, int64_t n,
                                    double gam, void* stream);

/* Host geometry records (HARM_G_NFIELD doubles): Kerr-Schild coordinates at
 * (a, r, th) (set_gcov_ksmetric, metric.c:1301), or MKS at X for cfg. */
void harm_geom_ks_record(double a, double r, double th, double* rec);
void harm_geom_mks_record(const harm_config* cfg, const double X[4], double* rec);
/* n Kerr-Schild records into host SoA rec[f*n + z] (host threads). */
void harm_geom_ks_batch(int64_t n, const double* a, const double* r, const double* th, double* rec);

/* ------------------------------------------------------------------------ */
/* Tile tuning and batches of independent grids (no counterpart in the       */
/* reference, which steps one grid per process)                              */
/* ------------------------------------------------------------------------ */
/* Times the face-flux kernels of directions 2 and 3 with each admissible tile
 * depth (4, 8, 16 faces along the sweep) on the current primitives and keeps
 * the fastest for this handle; reps launches per candidate.  td_out[dir]
 * (may be NULL) gets the chosen depths.  Overwrites F2/F3 and the wave dt
 * slots, which every flux calculation rewrites anyway. */
harm_status harm_tune(harm_handle* h, int32_t reps, int32_t td_out[4]);

/* n independent grids that share one stream and are stepped together: every
 * stage of the step is one launch over all grids.  All grids must share the
 * reconstruction, flux, time order and dimensionality; sizes, spin, floors,
 * Courant factor and initial data may differ.  Each grid keeps its own time
 * step. */
typedef struct harm_batch harm_batch;
harm_status harm_batch_create(const harm_config* cfgs, int32_t n, harm_batch** out);
harm_status harm_batch_destroy(harm_batch* b);
/* The handle of grid i (owned by the batch): every single-grid entry point
 * (set/get arrays, init, time, the hot-path pieces) works on it. */
harm_status harm_batch_grid(harm_batch* b, int32_t i, harm_handle** h);
/* harm_tune on every grid, then rebuild the batch's launch tables. */
harm_status harm_batch_tune(harm_batch* b, int32_t reps);
/* nsteps full RK steps of every grid (harm_step's sequence, batched).  With
 * use_graph != 0 one step is captured in a hipGraph on first use and
 * replayed. */
harm_status harm_batch_step(harm_batch* b, int32_t nsteps, int32_t use_graph);
harm_status harm_batch_sync(harm_batch* b);

/* ------------------------------------------------------------------------ */
/* Learned initial guess for the inversion (no counterpart in the reference:  */
/* HARM starts every Newton solve from the previous primitives).             */
/* ------------------------------------------------------------------------ */
/* An 8 -> 64 -> 64 -> 2 perceptron with tanh hidden layers, evaluated with
 * BF16 operands on the matrix cores (FP32 accumulation), trained on the GPU
 * by this library's own forward and backward kernels with Adam.  Inputs are
 * eight scale-free features of the solve's invariants (D, B^2, Q~^2, (Q.B)^2,
 * Q.n); outputs are log(W'/D) and log(u~^2) of the solution.  The FP64 Newton
 * solve stays the judge: the network only proposes its starting W'.
 * Parameters: W1[8][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2] (FP32,
 * row-major [in][out]); standardisation: feature mean[8], feature scale[8]
 * (1/std), target mean[2], target std[2]. */
enum { HARM_MLP_NPARAM = 4866, HARM_MLP_NNORM = 20, HARM_MLP_NIN = 8, HARM_MLP_NOUT = 2,
       HARM_MLP_BATCH = 128 };
typedef struct harm_mlp harm_mlp;
/* Glorot-uniform weights from a deterministic generator, zero biases. */
harm_status harm_mlp_create(uint64_t seed, harm_mlp** out);
harm_status harm_mlp_destroy(harm_mlp* m);
/* Host copies in and out (either pointer may be NULL). */
harm_status harm_mlp_get_params(harm_mlp* m, float* params, float* norm);
harm_status harm_mlp_set_params(harm_mlp* m, const float* params, const float* norm);
/* Training pairs from solved inversions (device pointers, SoA as in
 * harm_invert_batch): X[z][8] raw features, T[z][2] targets, ok[z] = 1 when
 * the pair is usable. */
harm_status harm_mlp_pairs(const double* U, const double* geom, const double* p, int64_t n, double gam,
                           double gammafail, float* X, float* T, int32_t* ok, void* stream);
/* Mean and spread of the usable pairs' features and targets. */
harm_status harm_mlp_fit_norm(harm_mlp* m, const float* X, const float* T, const int32_t* ok, int64_t n);
/* nsteps Adam steps on minibatches of HARM_MLP_BATCH pairs taken in the order
 * order[(first + s*128 + b) % norder] (device int32 indices into the nx rows
 * of X/T, each in [0, nx)); *loss (host, may be NULL) gets the mean loss over
 * the steps (synchronises). */
harm_status harm_mlp_train(harm_mlp* m, const float* X, const float* T, int64_t nx, const int32_t* order,
                           int64_t norder, int64_t first, int32_t nsteps, float lr, float* loss, void* stream);
/* Gradients of the mean-squared loss of one minibatch (rows order[0..127] of
 * the nx rows of X/T) with respect to all HARM_MLP_NPARAM parameters into
 * device grad[], no update; *loss (host) gets the batch loss. */
harm_status harm_mlp_grad(harm_mlp* m, const float* X, const float* T, int64_t nx, const int32_t* order, float* grad,
                          float* loss);
/* Y[z][2] = network(X[z][8]) in target units (device pointers). */
harm_status harm_mlp_forward(harm_mlp* m, const float* X, float* Y, int64_t n, void* stream);
/* harm_invert_batch started from the network's W'; a zone whose seeded solve
 * fails is solved again from p_inout's guess, and *nfallback (device
 * counter, may be NULL) is incremented for it. */
harm_status harm_invert_batch_mlp(harm_mlp* m, const double* U, const double* geom, double* p_inout,
                                  int32_t* pflag, int32_t* niter, int64_t n, double gam, double gammafail,
                                  unsigned long long* nfallback, void* stream);

/* The learned guess inside the step (replaces the per-zone Utoprimgen call of
 * advance(), advance.c:748, as harm_invert_batch_mlp does for a batch): with
 * m set, every inversion of the handle's steps starts its Newton solve from
 * the network's W' and falls back to the previous primitives when that
 * fails; floors, flags and counters are those of the plain inversion.  m
 * NULL restores the plain inversion.  m must outlive its use by h; batches
 * (harm_batch_*) always invert plainly. */
harm_status harm_set_inversion_guess(harm_handle* h, const harm_mlp* m);
/* zones whose seeded solve fell back, in the last whole step and summed
 * since harm_set_time (either pointer may be NULL) */
harm_status harm_get_guess_stats(harm_handle* h, int64_t* fallbacks_last_step, int64_t* fallbacks_total);
/* Training data from a grid: n active zones drawn with a seeded hash (with
 * replacement) with their conserved state from array which_U, primitives from
 * which_p and geometry record, into device SoA buffers U[8][n], geom[27][n],
 * p[8][n] (the layout of harm_mlp_pairs and harm_invert_batch). */
harm_status harm_grid_sample_zones(harm_handle* h, int32_t which_U, int32_t which_p, int64_t n, uint64_t seed,
                                   double* U, double* geom, double* p);

#ifdef __cplusplus
}
#endif
#endif /* HARMSTEP_H */
