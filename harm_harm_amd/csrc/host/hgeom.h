/*
 * hgeom.h -- host-side geometry for the MKS Kerr-Schild grid (C).
 *
 * Init-time only: the tables built here are uploaded once and read by the
 * device kernels.  The metric is stationary and axisymmetric, so every table
 * is a 2-D (i,j) plane over the ghosted storage box.
 */
#ifndef HGEOM_H
#define HGEOM_H

#include "harmstep.h"

#ifdef __cplusplus
extern "C" {
#endif

/* packed symmetric index of (a,b), a,b in 0..3 */
int hg_sym(int a, int b);

/* Kerr-Schild (M=1) covariant metric at BL-like (r,th); gpert_tt = g_tt + 1.
 * Follows set_gcov_ksmetric (metric.c:1301-1380). */
void hg_ks_gcov(double a, double r, double th, double g[4][4], double* gpert_tt);

/* MKS map (coord.c:930-960, LOGRSINTH): V = (t, r, th, ph) from X. */
void hg_mks_V(const harm_config* c, const double X[4], double V[4]);
/* dV/dX (coord.c:1515-1525). */
void hg_mks_dxdxp(const harm_config* c, const double X[4], double J[4][4]);

/* covariant MKS metric at X. */
void hg_mks_gcov(const harm_config* c, const double X[4], double g[4][4], double* gpert_tt);

/* Full geometry record (HARM_G_NFIELD doubles) from the covariant and
 * contravariant metric and sqrt(-det). */
void hg_record_from_gcov(const double g[4][4], const double gi[4][4], double gpert_tt, double gdet, double* rec);
/* closed-form inverse Kerr-Schild metric */
void hg_ks_gcon(double a, double r, double th, double gi[4][4]);
/* n KS records into SoA recsoa[f*n+z] */
void hg_ks_batch(int64_t n, const double* a, const double* r, const double* th, double* recsoa);

/* Record at MKS point X (gdet analytic: rho^2 sin th * dr/dX1 * dth/dX2 * 2pi). */
void hg_mks_record(const harm_config* c, const double X[4], double* rec);

/* Record at a Kerr-Schild point (a, r, th) in KS coordinates themselves. */
void hg_ks_record(double a, double r, double th, double* rec);

/* Connection Gamma^a_{bc} (HARM_NCONN packed) at MKS point X; metric
 * derivatives by Ridders' polynomial extrapolation of central differences
 * with starting step h1, h2 in X1, X2. */
void hg_mks_conn(const harm_config* c, const double X[4], double h1, double h2, double* conn);

/* Event horizon radius r_+ (metric.tools.c:1140). */
double hg_rhor(double a);
/* Grid inner radius with ihor zones inside the horizon (coord.c:2334-2360). */
double hg_setRin(double a, double R0, double Rout, int n1);

#ifdef __cplusplus
}
#endif
#endif
