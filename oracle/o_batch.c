/*
 * o_batch.c -- SoA batch wrappers around the oracle's per-zone functions, so
 * tests can drive them on many independent states at once.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

static void load(const double* p, const double* geom, int64_t n, int64_t z, double* pz, o_geom* g)
{
  double rec[27];
  for (int f = 0; f < 27; f++) rec[f] = geom[f * n + z];
  o_geom_from_rec(rec, g);
  for (int pl = 0; pl < O_NPR; pl++) pz[pl] = p[pl * n + z];
}

void o_p2u_batch(const double* p, const double* geom, double* U, int64_t n, double gam)
{
  for (int64_t z = 0; z < n; z++) {
    double pz[O_NPR], Uz[O_NPR];
    o_geom g;
    o_state q;
    load(p, geom, n, z, pz, &g);
    o_get_state(pz, &g, gam, &q);
    o_primtoflux(pz, &q, 0, &g, Uz);
    for (int pl = 0; pl < O_NPR; pl++) U[pl * n + z] = Uz[pl];
  }
}

void o_flux_batch(const double* p, const double* geom, double* F, int64_t n, int dir, double gam)
{
  for (int64_t z = 0; z < n; z++) {
    double pz[O_NPR], Fz[O_NPR];
    o_geom g;
    o_state q;
    load(p, geom, n, z, pz, &g);
    o_get_state(pz, &g, gam, &q);
    o_primtoflux(pz, &q, dir, &g, Fz);
    for (int pl = 0; pl < O_NPR; pl++) F[pl * n + z] = Fz[pl];
  }
}

void o_vchar_batch(const double* p, const double* geom, double* vmax, double* vmin, int64_t n, int dir, double gam)
{
  for (int64_t z = 0; z < n; z++) {
    double pz[O_NPR];
    o_geom g;
    o_state q;
    load(p, geom, n, z, pz, &g);
    o_get_state(pz, &g, gam, &q);
    o_vchar(pz, &q, dir, &g, gam, &vmax[z], &vmin[z]);
  }
}

void o_source_batch(const double* p, const double* geom, const double* conn, double* dU, int64_t n, double gam)
{
  for (int64_t z = 0; z < n; z++) {
    double pz[O_NPR], dUz[O_NPR], cz[40];
    o_geom g;
    o_state q;
    load(p, geom, n, z, pz, &g);
    for (int c = 0; c < 40; c++) cz[c] = conn[c * n + z];
    o_get_state(pz, &g, gam, &q);
    o_source(pz, &g, &q, cz, dUz);
    for (int pl = 0; pl < O_NPR; pl++) dU[pl * n + z] = dUz[pl];
  }
}
