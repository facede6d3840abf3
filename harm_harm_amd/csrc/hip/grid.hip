// grid.hip -- the HARM advance() step on one GPU's whole domain.
//
// Device state is structure-of-arrays a[pl][k][j][i] over the ghosted storage
// box (i fastest, so every kernel's lanes walk i and HBM accesses coalesce);
// geometry is stored once per (i,j) plane and reused across phi.  Each hot loop
// of advance() is one kernel:
//   k_flux<dir>   reconstruct + Riemann flux per face, fused   (flux.c:852;
//                 in flux.hip; WENO5BND face values from weno.hip)
//   k_emf/k_ct_all Toth corner EMFs and face B fluxes          (fluxct.c:186)
//   k_update      U(pi), flux divergence, connection source,
//                 RK combine and the source dt limit           (advance.c:420)
//   k_invert      Newton inversion (hot, check, cold retry)    (advance.c:748)
//   k_floors      final substep: floors + Lorentz-factor limit (fixup.c:561)
//   k_bound_*     radial extrapolation, polar axis, periodic   (bounds.tools.c)
//   k_post_fixup  failure averaging                            (fixup.c:1061)
// The timestep lives in device memory: reductions use atomicMin on the bit
// pattern of non-negative doubles and a one-thread kernel finishes the step,
// so a whole step enqueues without a host round trip.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include <algorithm>
#include <mutex>
#include "dgrid.h"
#include "gkern.h"
#include "dinvert.h"
#include "hhost.h"
#include "../host/hgeom.h"
#include "../host/hinit.h"

using hd::Dims;
using hd::Geom;
using hd::State;
using namespace hk;

namespace {


// Toth EMFs at corners, each multiplied by the corner gdet (CORNGDETVERSION)
__device__ __forceinline__ void emf_body(Dims d, int ni, int nj, int nk, const double* __restrict__ F1,
                                                const double* __restrict__ F2, const double* __restrict__ F3,
                                                const double* __restrict__ gf1, const double* __restrict__ gf2,
                                                const double* __restrict__ gc3, double* __restrict__ emf, const int blk)
{
  const long t = (long)blk * kBlock + threadIdx.x;
  const long n = (long)ni * nj * nk;
  if (t >= n) return;
  const int i = (int)(t % ni), j = (int)((t / ni) % nj), k = (int)(t / ((long)ni * nj));
  const int e0 = d.n1 > 1, e1 = d.n2 > 1, e2 = d.n3 > 1;
  const double c1 = (e1 && e2) ? 0.25 : 0.5, c2 = (e0 && e2) ? 0.25 : 0.5, c3 = (e0 && e1) ? 0.25 : 0.5;
  const int km = k - e2, jm = j - e1, im = i - e0;
  const long nc = d.ncell;
  auto f = [&](const double* F, int pl, int ii, int jj, int kk) { return F[(long)pl * nc + hd::cidx(d, ii, jj, kk)]; };
  double e;
  if (e1 || e2) {
    if (e1 && e2) e = c1 * ((+f(F2, HARM_B3, i, j, k) + f(F2, HARM_B3, i, j, km)) - f(F3, HARM_B2, i, j, k) - f(F3, HARM_B2, i, jm, k));
    else if (e1) e = c1 * (+f(F2, HARM_B3, i, j, k) + f(F2, HARM_B3, i, j, km));
    else e = c1 * (-f(F3, HARM_B2, i, j, k) - f(F3, HARM_B2, i, jm, k));
  } else e = 0.0;
  const long pp = hd::pidx(d, i, j);
  emf[t] = e * gf2[(long)HARM_G_GDET * d.nplane + pp];
  if (e0 || e2) {
    if (e0 && e2) e = c2 * ((+f(F3, HARM_B1, i, j, k) + f(F3, HARM_B1, im, j, k)) - f(F1, HARM_B3, i, j, k) - f(F1, HARM_B3, i, j, km));
    else if (e0) e = c2 * (-f(F1, HARM_B3, i, j, k) - f(F1, HARM_B3, i, j, km));
    else e = c2 * (+f(F3, HARM_B1, i, j, k) + f(F3, HARM_B1, im, j, k));
  } else e = 0.0;
  emf[n + t] = e * gf1[(long)HARM_G_GDET * d.nplane + pp];
  if (e0 || e1) {
    if (e0 && e1) e = c3 * ((+f(F1, HARM_B2, i, j, k) + f(F1, HARM_B2, i, jm, k)) - f(F2, HARM_B1, i, j, k) - f(F2, HARM_B1, im, j, k));
    else if (e0) e = c3 * (+f(F1, HARM_B2, i, j, k) + f(F1, HARM_B2, i, jm, k));
    else e = c3 * (-f(F2, HARM_B1, i, j, k) - f(F2, HARM_B1, im, j, k));
  } else e = 0.0;
  emf[2 * n + t] = e * gc3[pp];
}

__global__ __launch_bounds__(kBlock) void k_emf(Dims d, int ni, int nj, int nk, const double* __restrict__ F1,
                                                const double* __restrict__ F2, const double* __restrict__ F3,
                                                const double* __restrict__ gf1, const double* __restrict__ gf2,
                                                const double* __restrict__ gc3, double* __restrict__ emf)
{
  emf_body(d, ni, nj, nk, F1, F2, F3, gf1, gf2, gc3, emf, xcd_tile(blockIdx.x, gridDim.x));
}

// The three directions' B fluxes from the corner EMFs (fluxct.c:1106-1161)
// in one pass: thread (i, j, k) of the union of the three face boxes writes
// face (i, j, k) of every direction whose box holds it.
// The face fluxes' own field component (F1[B1], F2[B2], F3[B3]) is left as the
// face kernel wrote it: with B^dir continuous across the face (BDIRCONT) its
// induction flux B^dir v^dir - B^dir v^dir, its jump and so its Riemann flux
// are exactly +0, the value flux_ct assigns (fluxct.c:1106-1161).
__device__ __forceinline__ void ct_all_body(Dims d, int ni, int nj, int nk, const double* __restrict__ emf,
                                            double* __restrict__ F1, double* __restrict__ F2,
                                            double* __restrict__ F3, const int blk)
{
  const int e0 = d.n1 > 1, e1 = d.n2 > 1, e2 = d.n3 > 1;
  const int fi = d.n1 + e0, fj = d.n2 + e1, fk = d.n3 + e2;
  const int t = blk * kBlock + threadIdx.x;
  if (t >= fi * fj * fk) return;
  const int r = t / fi, i = t - r * fi, k = r / fj, j = r - k * fj;
  const int n = ni * nj * nk;
  auto E = [&](int c, int ii, int jj, int kk) { return emf[c * n + (kk * nj + jj) * ni + ii]; };
  const long o = hd::cidx(d, i, j, k);
  const long nc = d.ncell;
  const bool in_i = i < d.n1, in_j = j < d.n2, in_k = k < d.n3;
  if (e0 && in_j && in_k) {
    F1[HARM_B2 * nc + o] = 0.5 * (E(2, i, j, k) + E(2, i, j + e1, k));
    F1[HARM_B3 * nc + o] = -0.5 * (E(1, i, j, k) + E(1, i, j, k + e2));
  }
  if (e1 && in_i && in_k) {
    F2[HARM_B1 * nc + o] = -0.5 * (E(2, i, j, k) + E(2, i + e0, j, k));
    F2[HARM_B3 * nc + o] = 0.5 * (E(0, i, j, k) + E(0, i, j, k + e2));
  }
  if (e2 && in_i && in_j) {
    F3[HARM_B1 * nc + o] = 0.5 * (E(1, i, j, k) + E(1, i + e0, j, k));
    F3[HARM_B2 * nc + o] = -0.5 * (E(0, i, j, k) + E(0, i, j + e1, k));
  }
}

__global__ __launch_bounds__(kBlock) void k_ct_all(Dims d, int ni, int nj, int nk, const double* __restrict__ emf,
                                                   double* __restrict__ F1, double* __restrict__ F2,
                                                   double* __restrict__ F3)
{
  ct_all_body(d, ni, nj, nk, emf, F1, F2, F3, xcd_tile(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------------------
// update: Ui = U(pi); dU = -div F + source(pb); Uf, ucum; source dt limit
// ---------------------------------------------------------------------------
// DM: bit m-1 set when dimension m has more than one zone (a compile-time
// mask keeps the flux-divergence loads free of branches so they issue
// together); DM < 0 tests the dimensions at run time.
// MODE selects which of the RK arrays are read and written:
//   kUpdAll    the reference's loop as is: Ui = U(pi) stored in ui; uf and
//              ucum read (unless timeorder == 0) and written;
//   kUpdFirst  first substep of a step: Ui = U(pi) stored in ui, uf written,
//              ucum left alone (its coefficients are zero, so the reference
//              keeps it at 0 there);
//   kUpdMiddle later substep before the last: Ui read back from ui (pi is the
//              same state as on the first substep, so the value is identical),
//              uf read and written, ucum left alone;
//   kUpdLast   last substep: Ui from ui, uf read, ucum = 0 + (combination)
//              written without reading the zero it holds in the reference; uf
//              is not written (nothing reads it before the next step rebuilds
//              it).
enum { kUpdAll = 0, kUpdFirst = 1, kUpdMiddle = 2, kUpdLast = 3 };

template <int DM, int MODE>
__device__ __forceinline__ void update_body(Dims d, Params P, const double* __restrict__ pi,
                                                   const double* __restrict__ pb, const double* __restrict__ F1,
                                                   const double* __restrict__ F2, const double* __restrict__ F3,
                                                   double* __restrict__ ui, double* __restrict__ uf,
                                                   double* __restrict__ ucum, const double* __restrict__ gcent,
                                                   const double* __restrict__ conn, double cuf0, double cuf1,
                                                   double cuf2, double cun0, double cun1, double cun2, int timeorder,
                                                   double* __restrict__ ds, const int blk)
{
  // Work is ordered so that few values are live at once: flux divergence,
  // then everything that needs the state of pb (source, source dt), then the
  // state of pi (Ui) and the RK combination.
  const int t = blk * kBlock + threadIdx.x;
  const int nact = d.n1 * d.n2 * d.n3;
  double dts = kBIG;
  if (t < nact) {
    int i, j, k;
    zone_ijk(d, t, i, j, k);
    const long o = hd::cidx(d, i, j, k);
    const long pp = hd::pidx(d, i, j);
    const long nc = d.ncell;
    const double dt = ds[DS_DT];
    Geom G;
    hd::load_geom(gcent, pp, d.nplane, G);
    // flux divergence (flux2dUavg)
    const double idx1 = 1.0 / P.dx[1], idx2 = 1.0 / P.dx[2], idx3 = 1.0 / P.dx[3];
    const long s2 = d.ps1, s3 = (long)d.ps1 * d.ns2;
    double dUt[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) {
      double dU1 = 0, dU2 = 0, dU3 = 0;
      const bool e1 = DM < 0 ? d.n1 > 1 : (DM & 1), e2 = DM < 0 ? d.n2 > 1 : (DM & 2) != 0,
                 e3 = DM < 0 ? d.n3 > 1 : (DM & 4) != 0;
      if (e1) dU1 = (-(F1[pl * nc + o + 1] - F1[pl * nc + o]) * idx1);
      if (e2) dU2 = (-(F2[pl * nc + o + s2] - F2[pl * nc + o]) * idx2);
      if (e3) dU3 = (-(F3[pl * nc + o + s3] - F3[pl * nc + o]) * idx3);
      dUt[pl] = dU1 + dU2 + dU3;
    }
    // keep the phases apart: the scheduler would otherwise hoist every load of
    // the kernel to the top and run out of registers
    __builtin_amdgcn_sched_barrier(0);
    {
      double pbz[HARM_NPR];
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) pbz[pl] = pb[pl * nc + o];
      State qb;
      hd::get_state(pbz, G, P.gam, qb);
      double T00;
      double dUg[4];
      const double* cz = conn + pp;
      const long np = d.nplane;
      hd::source_rows(pbz, G, qb, [&](int c) { return cz[(long)c * np]; }, dUg, T00);
#pragma unroll
      for (int l = 0; l < 4; l++) dUt[HARM_UU + l] = dUt[HARM_UU + l] + dUg[l];
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++)
        if (pl < HARM_UU || pl > HARM_U3) dUt[pl] = dUt[pl] + (0.0 + G.gdet * 0.0);
      // dUtodt -> compute_dt_fromsource
      const double ig = hd::igdetnosing(G.gdet);
      const double UggUU = (T00 * G.gdet) * ig;
      double dUd[4], dUu[4], mydU[4], mydx[4];
#pragma unroll
      for (int jj = 0; jj < 4; jj++) dUd[jj] = (dUt[HARM_UU + jj] * dt) * ig;
      hd::raise(dUd, G, dUu);
      mydU[0] = dUd[0];
      mydx[0] = dt;
#pragma unroll
      for (int jj = 1; jj < 4; jj++) { mydU[jj] = dUu[jj]; mydx[jj] = P.dx[jj]; }
      double bcon[4], bcov[4];
      bcon[0] = pbz[HARM_B1] * qb.ucov[1] + pbz[HARM_B2] * qb.ucov[2] + pbz[HARM_B3] * qb.ucov[3];
#pragma unroll
      for (int jj = 1; jj < 4; jj++) bcon[jj] = (pbz[HARM_B1 - 1 + jj] + bcon[0] * qb.ucon[jj]) / qb.ucon[0];
      hd::lower(bcon, G, bcov);
      const double bsq = bcon[0] * bcov[0] + bcon[1] * bcov[1] + bcon[2] * bcov[2] + bcon[3] * bcov[3];
      const double rho = pbz[HARM_RHO], u = pbz[HARM_UU];
      const double Pg = (P.gam - 1.0) * u;
      const double w = rho + u + Pg;
      const double eta = w + bsq;
      const double rhoprime = fmax(fabs(eta), fabs(UggUU));
      double dtsource[4];
#pragma unroll
      for (int jj = 0; jj < 4; jj++) {
        const double ag = D_SMALL + fabs(mydU[jj] / rhoprime);
        dtsource[jj] = P.cour * (mydx[jj] / ag);
      }
      if (dtsource[0] < dts) dts = dtsource[0];
      const bool e1 = DM < 0 ? d.n1 > 1 : (DM & 1), e2 = DM < 0 ? d.n2 > 1 : (DM & 2) != 0,
                 e3 = DM < 0 ? d.n3 > 1 : (DM & 4) != 0;
      if (e1 && dtsource[1] < dts) dts = dtsource[1];
      if (e2 && dtsource[2] < dts) dts = dtsource[2];
      if (e3 && dtsource[3] < dts) dts = dtsource[3];
    }
    __builtin_amdgcn_sched_barrier(0);
    double Ui[HARM_NPR];
    if (MODE == kUpdAll || MODE == kUpdFirst) {
      double pz[HARM_NPR];
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) pz[pl] = pi[pl * nc + o];
      State q;
      hd::get_state(pz, G, P.gam, q);
      hd::primtoflux(pz, q, 0, G, Ui);
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) ui[pl * nc + o] = Ui[pl];
    } else {
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) Ui[pl] = ui[pl * nc + o];
    }
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) {
      const bool rd = MODE == kUpdAll ? timeorder != 0 : MODE != kUpdFirst;
      const double ufo = rd ? uf[pl * nc + o] : 0.0;
      const double ufn = (cuf0 * Ui[pl] + cuf1 * ufo + cuf2 * dt * dUt[pl]);
      if (MODE != kUpdLast) uf[pl * nc + o] = ufn;
      if (MODE == kUpdAll) {
        const double uco = (timeorder == 0) ? 0.0 : ucum[pl * nc + o];
        ucum[pl * nc + o] = uco + (cun0 * Ui[pl] + cun1 * dt * dUt[pl] + cun2 * ufn);
      } else if (MODE == kUpdLast) {
        ucum[pl * nc + o] = 0.0 + (cun0 * Ui[pl] + cun1 * dt * dUt[pl] + cun2 * ufn);
      }
    }
  }
  block_min_to(dts, ds + DS_ACCDT);
}

template <int DM, int MODE>
__global__ __launch_bounds__(kBlock) void k_update(Dims d, Params P, const double* __restrict__ pi,
                                                   const double* __restrict__ pb, const double* __restrict__ F1,
                                                   const double* __restrict__ F2, const double* __restrict__ F3,
                                                   double* __restrict__ ui, double* __restrict__ uf,
                                                   double* __restrict__ ucum, const double* __restrict__ gcent,
                                                   const double* __restrict__ conn, double cuf0, double cuf1,
                                                   double cuf2, double cun0, double cun1, double cun2, int timeorder,
                                                   double* __restrict__ ds)
{
  update_body<DM, MODE>(d, P, pi, pb, F1, F2, F3, ui, uf, ucum, gcent, conn, cuf0, cuf1, cuf2, cun0, cun1, cun2, timeorder, ds, xcd_tile(blockIdx.x, gridDim.x));
}

// The grid inversion is two passes.  The solve pass (every substep) is
// invert_zone without the floors: the hot Newton solve, the inversion check
// and the cold retry.  On the final substep the floors pass then applies
// fixup1zone's floors and the Lorentz-factor limit to every zone
// (floors_zone), reading the solve's primitives and flags back; kept apart,
// the solve does not carry the floors' own inversion in its registers.  The
// failure count of the final substep is taken after the floors.
__device__ __forceinline__ void invert_body(Dims d, Params P, const double* __restrict__ U,
                                            const double* __restrict__ pg, double* __restrict__ pf,
                                            int32_t* __restrict__ pflag, const double* __restrict__ gcent,
                                            int finalstep, unsigned long long* __restrict__ di, const int blk)
{
  const int t = blk * kBlock + threadIdx.x;
  const int nact = d.n1 * d.n2 * d.n3;
  int ns = 0, nf = 0;
  if (t < nact) {
    int i, j, k;
    zone_ijk(d, t, i, j, k);
    const long o = hd::cidx(d, i, j, k);
    const long pp = hd::pidx(d, i, j);
    const long nc = d.ncell;
    double Uz[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) Uz[pl] = U[pl * nc + o];
    const int fl = invert_zone<false, false, false>(d, P, Uz, U, pg, pf, pflag, gcent, nullptr, finalstep, o, pp, ns);
    nf = !finalstep && fl > 0;
  }
  newton_counts_to(ns, nf, di);
}

// DIAG: also keeps the states the fixes started from (floors_zone; fx = {p0,
// after-floor, flags}).
template <bool DIAG = false>
__device__ __forceinline__ void floors_body(Dims d, Params P, double* __restrict__ pf, int32_t* __restrict__ pflag,
                                            const double* __restrict__ gcent, const double* __restrict__ floors,
                                            unsigned long long* __restrict__ di, const int blk,
                                            double* __restrict__ fx_p0 = nullptr,
                                            double* __restrict__ fx_pfl = nullptr,
                                            uint8_t* __restrict__ fx_flag = nullptr)
{
  const int t = blk * kBlock + threadIdx.x;
  const int nact = d.n1 * d.n2 * d.n3;
  int ns = 0, nf = 0;
  if (t < nact) {
    int i, j, k;
    zone_ijk(d, t, i, j, k);
    const long o = hd::cidx(d, i, j, k);
    const long pp = hd::pidx(d, i, j);
    const long nc = d.ncell;
    Geom G;
    hd::load_geom(gcent, pp, d.nplane, G);
    double pz[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) pz[pl] = pf[pl * nc + o];
    const int fl0 = pflag[o];
    bool changed;
    const int fl = floors_zone<DIAG>(d, P, G, floors, pz, fl0, o, pp, ns, &changed, fx_p0, fx_pfl, fx_flag);
    if (changed) {
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) pf[pl * nc + o] = pz[pl];
    }
    if (fl != fl0) pflag[o] = fl;
    nf = fl > 0;
  }
  newton_counts_to(ns, nf, di);
}

// Four waves per SIMD (128 VGPRs): measured faster than three on the torus
// step.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_invert(
    Dims d, Params P, const double* __restrict__ U, const double* __restrict__ pg, double* __restrict__ pf,
    int32_t* __restrict__ pflag, const double* __restrict__ gcent, int finalstep, unsigned long long* __restrict__ di)
{
  invert_body(d, P, U, pg, pf, pflag, gcent, finalstep, di, xcd_tile(blockIdx.x, gridDim.x));
}

template <bool DIAG>
__global__ __launch_bounds__(kBlock) void k_floors(Dims d, Params P, double* __restrict__ pf,
                                                   int32_t* __restrict__ pflag, const double* __restrict__ gcent,
                                                   const double* __restrict__ floors,
                                                   unsigned long long* __restrict__ di, double* __restrict__ fx_p0,
                                                   double* __restrict__ fx_pfl, uint8_t* __restrict__ fx_flag)
{
  floors_body<DIAG>(d, P, pf, pflag, gcent, floors, di, xcd_tile(blockIdx.x, gridDim.x), fx_p0, fx_pfl, fx_flag);
}

// ---------------------------------------------------------------------------
// failure averaging (post_fixup, fixup.c:53 -> fixup_utoprim :1061)
// ---------------------------------------------------------------------------
// bound_pflag (step_ch.c:1637 -> bound_pflag_user, boundsint.c:18-330): one
// ghost layer of flags (numbnd = 1 for BOUNDINTTYPE,
// initbase.boundloop.c:83-88), filled in the reference's order by two launches:
//  phase 0: phi periodic over the (i, j) plane of -1..n (boundsint.c:80-92),
//           which reads the i and j ghost flags as the previous call left them;
//  phase 1: at active k, the outer radial OUTFLOW copy (torus X1UP,
//           init.fishmon.c:247; over active j, initbase.boundloop.c:358-363)
//           and the polar mirror (POLARAXIS, boundsint.c:178-200) over i in
//           -1..n1, which sees the radial copy: the inner radial FREEOUTFLOW
//           and both Bondi radial boundaries are not handled there, so those
//           ghost flags keep NANPFLAG (set_arrays_multidimen.c:404).
// Item t of a phase: phase 0 one (i, j) column; phase 1 first the radial
// copies (n2 n3 of them, torus only), then one (i, k) pair of polar mirrors.
__host__ __device__ __forceinline__ long bound_pflag_items(const Dims& d, int phase)
{
  const int s1 = d.n1 > 1, s2 = d.n2 > 1;
  if (phase == 0) return d.n3 > 1 ? (long)(d.n1 + 2 * s1) * (d.n2 + 2 * s2) : 0;
  return (long)d.n2 * d.n3 + (s2 ? (long)(d.n1 + 2 * s1) * d.n3 : 0);
}

__device__ __forceinline__ void bound_pflag_body(const Dims& d, int bc, int32_t* __restrict__ pflag, int phase, long t)
{
  const int s1 = d.n1 > 1, s2 = d.n2 > 1;
  if (phase == 0) {
    const int w = d.n1 + 2 * s1;
    const int i = (int)(t % w) - s1, j = (int)(t / w) - s2;
    pflag[hd::cidx(d, i, j, -1)] = pflag[hd::cidx(d, i, j, d.n3 - 1)];
    pflag[hd::cidx(d, i, j, d.n3)] = pflag[hd::cidx(d, i, j, 0)];
    return;
  }
  const long nr = (long)d.n2 * d.n3;
  if (t < nr) {
    if (bc == HARM_PROBLEM_TORUS && s1) {
      const int j = (int)(t % d.n2), k = (int)(t / d.n2);
      pflag[hd::cidx(d, d.n1, j, k)] = pflag[hd::cidx(d, d.n1 - 1, j, k)];
    }
    return;
  }
  t -= nr;
  const int w = d.n1 + 2 * s1;
  const int i = (int)(t % w) - s1, k = (int)(t / w);
  // the value at (i, 0 or n2-1, k) once the radial copy is done; the copy at
  // i = n1 is taken from its source so that this launch reads no flag it writes
  const int ii = (i == d.n1 && bc == HARM_PROBLEM_TORUS) ? d.n1 - 1 : i;
  pflag[hd::cidx(d, i, -1, k)] = pflag[hd::cidx(d, ii, 0, k)];
  pflag[hd::cidx(d, i, d.n2, k)] = pflag[hd::cidx(d, ii, d.n2 - 1, k)];
}

__global__ __launch_bounds__(kBlock) void k_bound_pflag(Dims d, int bc, int32_t* __restrict__ pflag, int phase)
{
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  if (t < bound_pflag_items(d, phase)) bound_pflag_body(d, bc, pflag, phase, t);
}

// IFUTOPRIMNOFAILORFIXED (global.nondepnmemonics.h:854)
__device__ __forceinline__ int good_flag(int f) { return f >= HARM_UTOPRIMFAILFIXEDCOLD && f <= 0; }

// Whether failure averaging accounts a zone with flag f (fixuputoprim_accounting,
// fixup.c:1605-1745): the convergence failures on every substep, the soft
// ones (rho < 0, u < 0, both, the u-only marks) on the final substep only
// (DOCOUNTNEG* = 1, STEPOVERNEG* = NEGDENSITY_NEVERFIXUP).  Accounted zones
// end with UTOPRIMFAILFIXEDUTOPRIM (:1761).
__device__ __forceinline__ bool accounted(int f, int finalstep)
{
  if (f == HARM_UTOPRIMFAILRHONEG || f == HARM_UTOPRIMFAILUNEG || f == HARM_UTOPRIMFAILRHOUNEG ||
      f == HARM_UTOPRIMFAILU2AVG1 || f == HARM_UTOPRIMFAILU2AVG2)
    return finalstep != 0;
  return f > 0 || f == HARM_UTOPRIMFAILFIXEDCOLD || f == -2;
}

// fixup_utoprim's averaging of one failed zone (fixup.c:1061-1330), the flags
// bounded: UTOPRIMADJUST = UTOPRIMAVG, GENERALAVERAGE 1 (:1035), HANDLE*NEG 0
// (so fixup_negdensities changes nothing, :1457), CAUSALAVG 0 (:1817),
// HOWTOAVG_WHEN_U2AVG = MAXUPOSAVG (:1827), TODONOGOOD = NOGOODAVERAGE
// (:2278); P.resetregion is USERRESETREGION (init.fishmon.h:356).  Reads pf
// and the flags as they stand (ptoavg, pflagfailorig); writes the zone's new
// primitives to pout at its own place, for the write-back launch.
//  - u-only failures (the U2AVG marks of the cold retry) take u from
//    general_average's second loop: the smallest u over the symmetric
//    neighbour pairs whose two u's are both at least the zone's own (:1990);
//  - every other failure takes rho..U3 as the mean of the good neighbours in
//    the 3^d box, summed pair by pair into two alternating sums (:1960-2030);
//  - with no usable neighbour, fixup_nogood (:2291): the mean of the 4+4
//    in-plane neighbours (AVG4_1, AVG4_2, :1812-1813), failed or not; in the
//    reset region (j <= 1, j >= n2 - 2 or i < 10) rho and u so and the
//    velocity the normal observer's (set_atmosphere VELREL4, init.tools.c:797);
//  - after a velocity average the Lorentz factor is capped at the zone's own
//    pre-average value, itself capped at GAMMAMAX (:1290-1322).
// DIAG (final substep): an accounted zone's change dV (U(after) - U(before))
// goes to dfl and `true` is returned, *cat being its category slot.
template <bool DIAG = false>
__device__ __forceinline__ bool post_fixup_body(Dims d, Params P, const double* __restrict__ pf,
                                                const int32_t* __restrict__ pflag, double* __restrict__ pout,
                                                const double* __restrict__ gcent, int finalstep, const int blk,
                                                double* dfl = nullptr, int* cat = nullptr)
{
  const long t = (long)blk * kBlock + threadIdx.x;
  const long nact = (long)d.n1 * d.n2 * d.n3;
  if (t >= nact) return false;
  const int i = (int)(t % d.n1), j = (int)((t / d.n1) % d.n2), k = (int)(t / ((long)d.n1 * d.n2));
  const long o = hd::cidx(d, i, j, k);
  const int f = pflag[o];
  if (f <= 0) return false;
  const long nc = d.ncell;
  const int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  const int rnx = s1 ? 3 : 1, rny = s2 ? 3 : 1;
  const int npow = (s1 + s2 + s3) == 3 ? 27 : (s1 + s2 + s3) == 2 ? 9 : 3;
  const int numupairs = (npow - 1) / 2;
  const bool urange = f == HARM_UTOPRIMFAILU2AVG1 || f == HARM_UTOPRIMFAILU2AVG2 || f == 13;  // UPERC
  const bool uonly = f == HARM_UTOPRIMFAILU2AVG1 || f == HARM_UTOPRIMFAILU2AVG2;
  const int startpl = urange ? HARM_UU : HARM_RHO, endpl = urange ? HARM_UU : HARM_U3;
  auto A = [&](int pl, int ii, int jj, int kk) { return pf[pl * nc + hd::cidx(d, ii, jj, kk)]; };
  double pz[HARM_NPR];
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) pz[pl] = pf[pl * nc + o];
  // general_average (fixup.c:1830)
  double sum0[HARM_NPR], sum1[HARM_NPR], lastmin[HARM_NPR];
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) { sum0[pl] = 0.0; sum1[pl] = 0.0; lastmin[pl] = 1E50; }
  int numavg0 = 0, numavg1 = 0;
  const double ref = pz[HARM_UU];
  for (int qq = 0; qq < numupairs; qq++) {
    const int ii = qq % rnx - s1, jj = (qq % (rnx * rny)) / rnx - s2, kk = qq / (rnx * rny) - s3;
    const int th = good_flag(pflag[hd::cidx(d, i + ii, j + jj, k + kk)]);
    const int at = good_flag(pflag[hd::cidx(d, i - ii, j - jj, k - kk)]);
    numavg0 += th + at;
    for (int pl = startpl; pl <= endpl; pl++) {
      const double f1 = A(pl, i + ii, j + jj, k + kk), f2 = A(pl, i - ii, j - jj, k - kk);
      if (!uonly) {
        const double v = f1 * th + f2 * at;
        if (qq & 1) sum1[pl] += v;
        else sum0[pl] += v;
      } else if (f1 >= ref && f2 >= ref) {
        const double m1 = lastmin[pl] < f1 ? lastmin[pl] : f1;
        lastmin[pl] = m1 < f2 ? m1 : f2;
        numavg1++;
      }
    }
  }
  int numavg;
  if (!uonly) {
    if (numavg0 != 0)
      for (int pl = startpl; pl <= endpl; pl++) pz[pl] = (sum0[pl] + sum1[pl]) / ((double)numavg0);
    numavg = numavg0;
  } else {
    if (numavg1 != 0)
      for (int pl = startpl; pl <= endpl; pl++) pz[pl] = lastmin[pl];
    numavg = numavg1;
  }
  if (f == HARM_UTOPRIMFAILU2AVG2) numavg++;
  if (numavg == 0) {  // fixup_nogood
    const int ip = i + s1, im = i - s1, jp = j + s2, jm = j - s2;
    auto avg8 = [&](int pl) {
      return 0.5 * ((0.25 * ((A(pl, i, jp, k) + A(pl, i, jm, k)) + (A(pl, im, j, k) + A(pl, ip, j, k)))) +
                    (0.25 * ((A(pl, ip, jp, k) + A(pl, ip, jm, k)) + (A(pl, im, jp, k) + A(pl, im, jm, k)))));
    };
    const bool reset = P.resetregion && ((j <= 1 || j >= d.n2 - 2) || (i < 10));
    if (reset) {
      for (int pl = HARM_RHO; pl <= HARM_UU; pl++) pz[pl] = avg8(pl);
      for (int pl = HARM_U1; pl <= HARM_U3; pl++) pz[pl] = 0.0;
    } else {
      for (int pl = startpl; pl <= endpl; pl++) pz[pl] = avg8(pl);
    }
  }
  Geom G;
  hd::load_geom(gcent, hd::pidx(d, i, j), d.nplane, G);
  if (startpl <= HARM_U1 && endpl >= HARM_U3) {
    double p0[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) p0[pl] = pf[pl * nc + o];
    double gamma = hd::gamma_of(p0, G);
    if (gamma > P.gammamax) gamma = P.gammamax;
    hd::limit_gamma(gamma, pz, G);
  }
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) pout[pl * nc + o] = pz[pl];
  if (DIAG && accounted(f, finalstep)) {
    double p0[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) p0[pl] = pf[pl * nc + o];
    fix_delta_to(P, G, p0, pz, dfl);
    *cat = fail_slot(f);
    return true;
  }
  return false;
}

template <bool DIAG>
__global__ __launch_bounds__(kBlock) void k_post_fixup(Dims d, Params P, const double* __restrict__ pf,
                                                       const int32_t* __restrict__ pflag, double* __restrict__ pout,
                                                       const double* __restrict__ gcent, int finalstep,
                                                       double* __restrict__ fl_part)
{
  double dfl[HARM_NPR] = {0, 0, 0, 0, 0, 0, 0, 0};
  int cat = 0;
  const bool ch = post_fixup_body<DIAG>(d, P, pf, pflag, pout, gcent, finalstep, blockIdx.x, dfl, &cat);
  if (DIAG && !__syncthreads_or(ch)) {  // no accounted zone in the block
    if (threadIdx.x < 32) fl_part[(long)(threadIdx.x >> 3) * gridDim.x * 8 + (long)blockIdx.x * 8 + (threadIdx.x & 7)] = 0.0;
    return;
  }
  if (DIAG)  // one partial per failure category: fl_part + c * nblk * 8
    for (int c = 0; c < 4; c++) {
      const bool mine = ch && cat == c;
      double v[HARM_NPR];
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) v[pl] = mine ? dfl[pl] : 0.0;
      block_sum8_part(v, mine, fl_part + (long)c * gridDim.x * 8, blockIdx.x);
    }
}

// The write-back after k_post_fixup: every failed zone takes its averaged
// primitives, and accounted zones (also the fixed-flag ones, whose
// accounting changes nothing) get UTOPRIMFAILFIXEDUTOPRIM.
__device__ __forceinline__ void fixup_apply_body(const Dims& d, double* __restrict__ pf, int32_t* __restrict__ pflag,
                                                 const double* __restrict__ pin, int finalstep, const int blk)
{
  const long t = (long)blk * kBlock + threadIdx.x;
  const long nact = (long)d.n1 * d.n2 * d.n3;
  if (t >= nact) return;
  const int i = (int)(t % d.n1), j = (int)((t / d.n1) % d.n2), k = (int)(t / ((long)d.n1 * d.n2));
  const long o = hd::cidx(d, i, j, k);
  const int f = pflag[o];
  if (f == 0) return;
  if (f > 0) {
    const long nc = d.ncell;
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) pf[pl * nc + o] = pin[pl * nc + o];
  }
  if (accounted(f, finalstep)) pflag[o] = HARM_UTOPRIMFAILFIXEDUTOPRIM;
}

__global__ __launch_bounds__(kBlock) void k_fixup_apply(Dims d, double* __restrict__ pf, int32_t* __restrict__ pflag,
                                                        const double* __restrict__ pin, int finalstep)
{
  fixup_apply_body(d, pf, pflag, pin, finalstep, blockIdx.x);
}

// ---------------------------------------------------------------------------
// boundary conditions
// ---------------------------------------------------------------------------
// Radial extrapolation into ghost zone number `gsel` (0 .. ng1-1, counted
// outward from the boundary) of the (j, k) line; the slopes come from the
// three boundary zones, which no ghost write touches, so the ghost zones of a
// line can be filled by separate threads.
__device__ void extrap_line(const Dims& d, const Params& P, double* p, const double* gcent, int up, int j, int k,
                            int gsel)
{
  const long nc = d.ncell;
  const int ri = up ? d.n1 - 1 : 0;
  const int ri2 = up ? ri - 1 : ri + 1;
  const int ri3 = up ? ri - 2 : ri + 2;
  const double signdq = up ? -1.0 : 1.0;
  Geom Gr, Gr2, Gr3;
  hd::load_geom(gcent, hd::pidx(d, ri, j), d.nplane, Gr);
  hd::load_geom(gcent, hd::pidx(d, ri2, j), d.nplane, Gr2);
  hd::load_geom(gcent, hd::pidx(d, ri3, j), d.nplane, Gr3);
  const long o1 = hd::cidx(d, ri, j, k), o2 = hd::cidx(d, ri2, j, k), o3 = hd::cidx(d, ri3, j, k);
  double pr[HARM_NPR], pr2[HARM_NPR], pr3[HARM_NPR];
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) {
    pr[pl] = p[pl * nc + o1];
    pr2[pl] = p[pl * nc + o2];
    pr3[pl] = p[pl * nc + o3];
  }
  double Bd3 = 0.0, Bd3ri2 = 0.0, Bd3ri3 = 0.0;
#pragma unroll
  for (int jj = 1; jj < 4; jj++) Bd3 += pr[HARM_B1 + jj - 1] * hd::gco(Gr, 3, jj);
#pragma unroll
  for (int jj = 1; jj < 4; jj++) Bd3ri2 += pr2[HARM_B1 + jj - 1] * hd::gco(Gr2, 3, jj);
#pragma unroll
  for (int jj = 1; jj < 4; jj++) Bd3ri3 += pr3[HARM_B1 + jj - 1] * hd::gco(Gr3, 3, jj);
  double Dqp = Bd3ri3 - Bd3ri2, Dqm = Bd3ri2 - Bd3, Dqc = Bd3ri3 - Bd3;
  const double dqBd3 = signdq * hd::minmod(hd::minmod(Dqp, Dqm), Dqc);
  State qr, qr2;
  hd::get_state(pr, Gr, P.gam, qr);
  hd::get_state(pr2, Gr2, P.gam, qr2);
  double dq[HARM_NPR], dql[HARM_NPR], dqg[HARM_NPR];
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) {
    Dqp = pr3[pl] - pr2[pl];
    Dqm = pr2[pl] - pr[pl];
    Dqc = pr3[pl] - pr[pl];
    dq[pl] = signdq * hd::minmod(hd::minmod(Dqp, Dqm), Dqc);
    Dqp = log(D_SMALL + fabs(pr3[pl])) - log(D_SMALL + fabs(pr2[pl]));
    Dqm = log(D_SMALL + fabs(pr2[pl])) - log(D_SMALL + fabs(pr[pl]));
    Dqc = log(D_SMALL + fabs(pr3[pl])) - log(D_SMALL + fabs(pr[pl]));
    dql[pl] = signdq * hd::minmod(hd::minmod(Dqp, Dqm), Dqc);
    Dqp = (Gr3.gdet) * pr3[pl] - (Gr2.gdet) * pr2[pl];
    Dqm = (Gr2.gdet) * pr2[pl] - (Gr.gdet) * pr[pl];
    Dqc = (Gr3.gdet) * pr3[pl] - (Gr.gdet) * pr[pl];
    dqg[pl] = signdq * hd::minmod(hd::minmod(Dqp, Dqm), Dqc);
  }
  const double xdqfrac = fabs(qr.ucon[0] - qr2.ucon[0]) / qr.ucon[0];
  double ydqfrac;
  if (xdqfrac < 0.1) ydqfrac = 0.0;
  else if (xdqfrac < 0.2) ydqfrac = 0.0 + (1.0 - 0.0) / (0.2 - 0.1) * (xdqfrac - 0.1);
  else ydqfrac = 1.0;
#pragma unroll
  for (int jj = 0; jj < 4; jj++) {
    dq[HARM_UU + jj] = dq[HARM_UU + jj] * (1.0 - ydqfrac);
    dql[HARM_UU + jj] = dql[HARM_UU + jj] * (1.0 - ydqfrac);
    dqg[HARM_UU + jj] = dqg[HARM_UU + jj] * (1.0 - ydqfrac);
  }
  {
    const int i = up ? d.n1 + gsel : -1 - gsel;
    double pi[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) pi[pl] = pr[pl];
    Geom Gi;
    hd::load_geom(gcent, hd::pidx(d, i, j), d.nplane, Gi);
    const double di = (double)(i - ri);
    for (int pl = 0; pl <= HARM_UU; pl++) {
      const double ftemp = exp(-signdq * dql[pl]) - 1.1;
      double mydqlog;
      if (ftemp >= 0.0) mydqlog = log(1.1) / (-signdq);
      else mydqlog = dql[pl];
      pi[pl] = exp(log(D_SMALL + fabs(pr[pl])) + mydqlog * di);
    }
    for (int pl = HARM_U1; pl <= HARM_U3; pl++) pi[pl] = pr[pl] + dq[pl] * di;
    for (int pl = HARM_B1; pl <= HARM_B2; pl++) {
      const double ign = hd::igdetnosing(Gi.gdet);
      pi[pl] = (pr[pl] * fabs((Gr.gdet) + dqg[pl] * di) * ign);
    }
    const double Bu1 = pi[HARM_B1], Bu2 = pi[HARM_B2];
    const double gcon03 = hd::gcn(Gi, 0, 3), gcon13 = hd::gcn(Gi, 1, 3), gcon23 = hd::gcn(Gi, 2, 3),
                 gcon33 = hd::gcn(Gi, 3, 3);
    const double gcov01 = hd::gco(Gi, 0, 1), gcov02 = hd::gco(Gi, 0, 2), gcov11 = hd::gco(Gi, 1, 1);
    const double gcov12 = hd::gco(Gi, 1, 2), gcov21 = gcov12, gcov22 = hd::gco(Gi, 2, 2);
    const double gcov03 = hd::gco(Gi, 0, 3), gcov13 = hd::gco(Gi, 1, 3), gcov23 = hd::gco(Gi, 2, 3);
    const double myBd3 = Bd3 + dqBd3 * di;
    const double ftemp = (1.0 - gcon03 * gcov03 - gcon13 * gcov13 - gcon23 * gcov23);
    const double ign = hd::igdetnosing(ftemp);
    pi[HARM_B3] = (myBd3 * gcon33 + Bu1 * gcon03 * gcov01 + Bu2 * gcon03 * gcov02 + Bu1 * gcon13 * gcov11
                   + Bu2 * gcon13 * gcov12 + Bu1 * gcon23 * gcov21 + Bu2 * gcon23 * gcov22) * ign;
    if (up) {
      hd::inflow_check_x1up(pi, Gi, P.gam, P.gammamax);
      hd::limit_gamma(P.gammamax, pi, Gi);
    }
    const long oi = hd::cidx(d, i, j, k);
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) p[pl * nc + oi] = pi[pl];
  }
}

// Bondi radial boundaries (bounds.sasha.c:182-198, :341-353; BONDIEXTRAPORDER
// 1): the integrals of the boundary zone (inner: C1, C2 and K; outer: C1 with
// the exact C2 and K) give the state of ghost zone `gsel` of the (j, k) line;
// the fields are zero there (the reference zeroes every primitive first).
__device__ void bondi_line(const Dims& d, const Params& P, double* p, const double* gcent, int up, int j, int k,
                           int gsel)
{
  const long nc = d.ncell;
  const int ri = up ? d.n1 - 1 : 0;
  const int i = up ? d.n1 + gsel : -1 - gsel;
  double pr[HARM_NPR];
  const long o1 = hd::cidx(d, ri, j, k);
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) pr[pl] = p[pl * nc + o1];
  Geom Gr;
  hd::load_geom(gcent, hd::pidx(d, ri, j), d.nplane, Gr);
  const double rr = P.R0 + exp(P.startx[1] + (ri + 0.5) * P.dx[1]);
  double C1, C2, K;
  hd::bondi_integrals(pr, Gr, rr, P.R0, P.gam, C1, C2, K);
  if (up) {
    C2 = hd::kBondiC2x;
    K = hd::kBondiKx;
  }
  const double r = P.R0 + exp(P.startx[1] + (i + 0.5) * P.dx[1]);
  double pi[HARM_NPR];
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) pi[pl] = 0.0;
  double ur = 0.0;
  hd::bondi_invert(1 / (P.gam - 1), C1, C2, K, r, pi[HARM_RHO], pi[HARM_UU], ur);
  Geom Gi;
  hd::load_geom(gcent, hd::pidx(d, i, j), d.nplane, Gi);
  hd::radial_ks_to_rel4(ur, r, P.R0, Gi, pi);
  const long oi = hd::cidx(d, i, j, k);
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) p[pl * nc + oi] = pi[pl];
}

__device__ __forceinline__ void bound_x1_body(Dims d, Params P, double* __restrict__ p,
                                                     const double* __restrict__ gcent, const int blk)
{
  const long t = (long)blk * kBlock + threadIdx.x;
  if (t >= 2L * d.ng1 * d.n2 * d.n3) return;
  const int up = (int)(t & 1);
  const long l = t >> 1;
  const int gsel = (int)(l % d.ng1);
  const long line = l / d.ng1;
  const int j = (int)(line % d.n2), k = (int)(line / d.n2);
  if (P.bc == HARM_PROBLEM_BONDI) bondi_line(d, P, p, gcent, up, j, k, gsel);
  else extrap_line(d, P, p, gcent, up, j, k, gsel);
}

__global__ __launch_bounds__(kBlock) void k_bound_x1(Dims d, Params P, double* __restrict__ p,
                                                     const double* __restrict__ gcent)
{
  bound_x1_body(d, P, p, gcent, blockIdx.x);
}

// one ghost zone of the polar reflection: ghost row j = -ng2 + g (side 0) or
// n2 + g (side 1) of column col = (i + ng1) + ns1 * k
__device__ __forceinline__ void x2_reflect_zone(const Dims& d, double* __restrict__ p, int side, int g, long col)
{
  const int i = (int)(col % d.ns1) - d.ng1, k = (int)(col / d.ns1);
  const int n2 = d.n2, n3 = d.n3;
  const int rk = (n3 > 1) ? (k + n3 / 2) % n3 : k;
  const int j = side == 0 ? -d.ng2 + g : n2 + g;
  const int rj = side == 0 ? -j - 1 : 2 * (n2 - 1) - j + 1;
  const long nc = d.ncell;
  const long so = hd::cidx(d, i, rj, rk), doff = hd::cidx(d, i, j, k);
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) {
    double v = p[pl * nc + so];
    if (pl == HARM_U2 || pl == HARM_B2) v *= -1.;
    p[pl * nc + doff] = v;
  }
}

__device__ __forceinline__ void x2_reflect_body(Dims d, double* __restrict__ p, int side, const int blk)
{
  const long t = (long)blk * kBlock + threadIdx.x;
  const long ncol = (long)d.ns1 * d.n3;
  if (t >= ncol * d.ng2) return;
  x2_reflect_zone(d, p, side, (int)(t / ncol), t % ncol);
}

__global__ __launch_bounds__(kBlock) void k_x2_reflect(Dims d, double* __restrict__ p, int side)
{
  x2_reflect_body(d, p, side, blockIdx.x);
}

// pole death of one column col = (i + ng1) + ns1 * k at one pole: the first
// active row and the first ghost row are rebuilt from the second active row.
// The field components of the ghost row are kept as the reflection left
// them; with mirror_ghost they are taken from the mirrored first active row
// directly (the reflection of that ghost row has not run).
__device__ __forceinline__ void x2_poledeath_col(const Dims& d, const Params& P, double* __restrict__ p,
                                                 const double* __restrict__ gcent, const double* __restrict__ rcent,
                                                 int side, long col, bool mirror_ghost = false)
{
  const int i = (int)(col % d.ns1) - d.ng1, k = (int)(col / d.ns1);
  const int n2 = d.n2, n3 = d.n3;
  const long nc = d.ncell;
  const double x2c0 = P.startx[2];
  const int rj = side == 0 ? 1 : n2 - 2;
  const int js = side == 0 ? -1 : n2 - 1;
  const int je = side == 0 ? 0 : n2;
  const double X0 = side == 0 ? x2c0 : x2c0 + n2 * P.dx[2];
  const double Xr = x2c0 + (rj + 0.5) * P.dx[2];
  const long ro = hd::cidx(d, i, rj, k);
  double pref[HARM_NPR];
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) pref[pl] = p[pl * nc + ro];
  for (int j = js; j <= je; j++) {
    const long o = hd::cidx(d, i, j, k);
    double pj[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) pj[pl] = p[pl * nc + o];
    if (mirror_ghost && j == (side == 0 ? js : je)) {
      // x2_reflect_zone of this ghost row, field components only (pole death
      // replaces the rest, and never writes the fields of the mirrored row)
      const int rk = (n3 > 1) ? (k + n3 / 2) % n3 : k;
      const long so = hd::cidx(d, i, side == 0 ? 0 : n2 - 1, rk);
      pj[HARM_B1] = p[HARM_B1 * nc + so];
      pj[HARM_B2] = -p[HARM_B2 * nc + so];
      pj[HARM_B3] = p[HARM_B3 * nc + so];
    }
    const double X = x2c0 + (j + 0.5) * P.dx[2];
    for (int pl = HARM_RHO; pl <= HARM_U3; pl++) {
      if (pl == HARM_U2) continue;
      pj[pl] = pref[pl];
    }
    double ftemp = pref[HARM_U2];
    if ((side == 0 && ftemp > 0.0) || (side == 1 && ftemp < 0.0)) ftemp = 0.0;
    pj[HARM_U2] = ftemp + (X - Xr) * (ftemp - 0.0) / (Xr - X0);
    if (n3 == 1) pj[HARM_B3] = pref[HARM_B3];
    Geom Gj;
    hd::load_geom(gcent, hd::pidx(d, i, j), d.nplane, Gj);
    State q;
    hd::get_state(pj, Gj, P.gam, q);
    if (q.ucon[1] > 0.0) {
      const double r = rcent[i + d.ng1];
      double lim;
      if (r <= 10.0) lim = 1.1;
      else lim = 1.1 * pow(r / 10.0, 1.0);
      hd::limit_gamma(lim, pj, Gj);
    } else {
      hd::limit_gamma(P.gammamax, pj, Gj);
    }
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) p[pl * nc + o] = pj[pl];
  }
}

__device__ __forceinline__ void x2_poledeath_body(Dims d, Params P, double* __restrict__ p,
                                                         const double* __restrict__ gcent,
                                                         const double* __restrict__ rcent, int side, const int blk)
{
  const long t = (long)blk * kBlock + threadIdx.x;
  const long ncol = (long)d.ns1 * d.n3;
  if (t >= ncol || P.bc == HARM_PROBLEM_BONDI) return;  // Bondi: plain reflection only
  x2_poledeath_col(d, P, p, gcent, rcent, side, t);
}

// Both poles in one pass, for n2 > ng2: the pole-death columns (rows -1, 0
// and n2-1, n2), then the reflection of the ghost rows beyond the first
// (-ng2..-2 and n2+1..n2+ng2-1).  The reflection reads rows 1..ng2-1 and
// n2-ng2..n2-2, pole death reads rows 1 and n2-2 and the field components of
// rows 0 and n2-1 (which it never writes), none of which this pass writes
// when n2 > ng2, so the result equals reflect-then-pole-death at one pole and
// then the other: there pole death overwrites all but the fields of the
// reflected rows -1 and n2, and those fields are mirrored here directly.
__device__ __forceinline__ void x2_poles_body(Dims d, Params P, double* __restrict__ p,
                                              const double* __restrict__ gcent, const double* __restrict__ rcent,
                                              const int blk)
{
  const long t = (long)blk * kBlock + threadIdx.x;
  const long ncol = (long)d.ns1 * d.n3;
  if (t < 2 * ncol) {
    const int side = t >= ncol;
    if (P.bc == HARM_PROBLEM_BONDI)  // plain polar reflection (POLARAXIS): rows -1 and n2 too
      x2_reflect_zone(d, p, side, side == 0 ? d.ng2 - 1 : 0, t - side * ncol);
    else
      x2_poledeath_col(d, P, p, gcent, rcent, side, t - side * ncol, true);
    return;
  }
  const long u = t - 2 * ncol, nside = ncol * (d.ng2 - 1);
  if (u >= 2 * nside) return;
  const int side = u >= nside;
  const long v = u - side * nside;
  x2_reflect_zone(d, p, side, (int)(v / ncol) + side, v % ncol);
}

__global__ __launch_bounds__(kBlock) void k_x2_poles(Dims d, Params P, double* __restrict__ p,
                                                     const double* __restrict__ gcent,
                                                     const double* __restrict__ rcent)
{
  x2_poles_body(d, P, p, gcent, rcent, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_x2_poledeath(Dims d, Params P, double* __restrict__ p,
                                                         const double* __restrict__ gcent,
                                                         const double* __restrict__ rcent, int side)
{
  x2_poledeath_body(d, P, p, gcent, rcent, side, blockIdx.x);
}

__device__ __forceinline__ void bound_x3_body(Dims d, double* __restrict__ p, const int blk)
{
  const long t = (long)blk * kBlock + threadIdx.x;
  const long nplane = d.nplane;
  if (t >= nplane * 2 * d.ng3) return;
  const long pp = t % nplane;
  const int g = (int)(t / nplane);
  const int k = g < d.ng3 ? -d.ng3 + g : d.n3 + (g - d.ng3);
  const int ks = k < 0 ? k + d.n3 : k - d.n3;
  const long nc = d.ncell;
  const long o = (long)(k + d.ng3) * nplane + pp, so = (long)(ks + d.ng3) * nplane + pp;
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) p[pl * nc + o] = p[pl * nc + so];
}

__global__ __launch_bounds__(kBlock) void k_bound_x3(Dims d, double* __restrict__ p)
{
  bound_x3_body(d, p, blockIdx.x);
}

// ---------------------------------------------------------------------------
// scalars: reset, per-substep dt, end of step
// ---------------------------------------------------------------------------
__global__ void k_scal_begin_flux(double* ds)
{
  ds[DS_NDT1] = kBIG;
  ds[DS_NDT2] = kBIG;
  ds[DS_NDT3] = kBIG;
  ds[DS_ACCDT] = kBIG;
}
__global__ void k_scal_begin_step(double* ds, unsigned long long* di)
{
  ds[DS_LASTNDT] = kBIG;
  di[DI_NSTROKE] = 0;
  di[DI_NFAIL] = 0;
  di[DI_NFALLBACK] = 0;
}
__global__ void k_scal_substep_dt(double* ds)
{
  const double wavedt = 1. / (1. / ds[DS_NDT1] + 1. / ds[DS_NDT2] + 1. / ds[DS_NDT3]);
  const double nd = 1.0 * fmin(wavedt, ds[DS_ACCDT]);
  if (nd < ds[DS_LASTNDT]) ds[DS_LASTNDT] = nd;
}
__global__ void k_scal_end_step(double* ds, unsigned long long* di, double safe)
{
  double fullndt = ds[DS_LASTNDT];
  const double dt = ds[DS_DT];
  if (fullndt > safe * dt) fullndt = safe * dt;
  ds[DS_DTSTEP] = dt;
  ds[DS_T] = ds[DS_T] + dt;
  ds[DS_DT] = fullndt;
  ds[DS_NSTEP] = ds[DS_NSTEP] + 1.0;
  di[DI_NSTROKE_SUM] += di[DI_NSTROKE];
  di[DI_NFAIL_SUM] += di[DI_NFAIL];
  di[DI_NFALLBACK_SUM] += di[DI_NFALLBACK];
}

// Boundary flux sums through the inner and outer radial faces
// (diag_flux_pureflux), deterministic: kFaccBlocks blocks each reduce a fixed
// slice of the (j, k) faces in a fixed tree order into `part`, then one block
// adds the partial sums in block order.
constexpr int kFaccBlocks = 64;

__device__ __forceinline__ void fluxacc_part_body(Dims d, const double* __restrict__ F1,
                                                         double* __restrict__ part, const int blk)
{
  __shared__ double red[4][kBlock];
  double s[4] = {0, 0, 0, 0};
  const int nf = d.n2 * d.n3, chunk = (nf + kFaccBlocks - 1) / kFaccBlocks;
  const int t0 = blk * chunk, t1 = min(nf, t0 + chunk);
  const long nc = d.ncell;
  for (int t = t0 + threadIdx.x; t < t1; t += kBlock) {
    const int k = t / d.n2, j = t - k * d.n2;
    const long oi = hd::cidx(d, 0, j, k), oo = hd::cidx(d, d.n1, j, k);
    s[0] += F1[HARM_RHO * nc + oi];
    s[1] += F1[HARM_UU * nc + oi];
    s[2] += F1[HARM_RHO * nc + oo];
    s[3] += F1[HARM_UU * nc + oo];
  }
  for (int c = 0; c < 4; c++) red[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int c = 0; c < 4; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[blk * 4 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(kBlock) void k_fluxacc_part(Dims d, const double* __restrict__ F1,
                                                         double* __restrict__ part)
{
  fluxacc_part_body(d, F1, part, blockIdx.x);
}

__global__ void k_fluxacc_final(Params P, const double* __restrict__ part, double* __restrict__ ds, double weight)
{
  const int c = threadIdx.x;
  if (c >= 4) return;
  double v = 0.0;
  for (int b = 0; b < kFaccBlocks; b++) v += part[b * 4 + c];
  ds[DS_FLUXACC + c] += v * P.dx[2] * P.dx[3] * ds[DS_DT] * weight;
}

inline unsigned nblocks(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

// ===========================================================================
// handle
// ===========================================================================
struct harm_handle {
  harm_config cfg;
  Dims d;
  Params P;
  hipStream_t stream = nullptr;
  int dev = 0;  // the device current at harm_create (the stream's)
  double* arr[HARM_NARR] = {nullptr};
  int32_t* pflag = nullptr;  // over the storage box; the one-layer ghost band holds bound_pflag's values
  double* pfix = nullptr;    // failure averaging's new primitives of the failed zones (SoA, own place)
  double* geom[HARM_NLOC] = {nullptr};
  double* gcorn3 = nullptr;
  double* conn = nullptr;
  double* floors = nullptr;
  double* rcent = nullptr;
  double* emf = nullptr;
  double* wl = nullptr;    // WENO5BND: face values of the lines, one SoA block (like p) per direction
  double* wr = nullptr;
  hipStream_t side[2] = {nullptr, nullptr};  // WENO5BND: directions 2 and 3 of fluxcalc, forked from stream
  hipEvent_t evfork = nullptr, evjoin[2] = {nullptr, nullptr};
  double* cf = nullptr;    // the dump's four-current: Faraday samples (18 x ncell) and J^mu (4 x ncell),
  double* jcon = nullptr;  // allocated by the first harm_currents
  double* xfer = nullptr;  // HARM's interleaved layout of one array (NPR x storage cells), staged for set/get
  long cur_calls = 0;      // current_calc_1's call count and the dt of its last call
  double cur_lastdt = 0.0;
  double* wdp = nullptr;   // WENO5BND: dP measure per direction (3 x ncell)
  double* wtm = nullptr;   // WENO5BND: per-zone terms of dP (10 x ncell)
  double* facc = nullptr;  // boundary-flux partial sums
  double* en = nullptr;      // energy-file accumulators (EN_*)
  double* enpart = nullptr;  // their per-block partial sums
  double* entot = nullptr;   // on-demand totals: U[8], divb max, divb sum
  bool ener_on = false;      // steps add to the energy-file accumulators
  double* fx_p0 = nullptr;   // with ener_on: states the final substep's fixes started from
  double* fx_pfl = nullptr;
  uint8_t* fx_flag = nullptr;
  const harm_mlp* guess = nullptr;  // learned inversion guess (harm_set_inversion_guess), or none
  int tile_td[4] = {0, 0, 0, 0};  // sweep depth of the face tiles of dirs 2,3 (0: cost model)
  hipGraph_t sgraph = nullptr;      // one full step, captured on first use when not profiling
  hipGraphExec_t sexec = nullptr;
  FaceTile ft[4];                 // planned face tiles per direction (index 1..3)
  double* ds = nullptr;
  unsigned long long* di = nullptr;
  int emf_ni = 0, emf_nj = 0, emf_nk = 0;
  std::vector<double> hgeom[HARM_NLOC];  // AoS host copies
  std::vector<double> hconn;
  std::vector<double> hrcent;
  std::vector<double> hgc3;
  // kernel timing
  bool prof = false;
  std::vector<hipEvent_t> evpool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> evrec;
  double prof_ms[HARM_K_N] = {0};
  long long prof_n[HARM_K_N] = {0};
  hipEvent_t cur0 = nullptr;
};

namespace {

harm_status upload(harm_handle* h, double* dst, const double* src, size_t n)
{
  return hh::check_hip(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice), "upload");
}

harm_status alloc(void** p, size_t bytes) { return hh::check_hip(hipMalloc(p, bytes ? bytes : 8), "hipMalloc"); }

// every flag NANPFLAG, as a run starts (set_arrays_multidimen.c:404): ghost
// flags that bound_pflag never sets keep that value, and the flags' phi ghost
// corners are read before they are first set
harm_status reset_pflag(harm_handle* h)
{
  return hh::check_hip(hipMemsetD32((hipDeviceptr_t)h->pflag, HARM_NANPFLAG, (size_t)h->d.ncell), "hipMemsetD32");
}

void make_dims(const harm_config* cfg, Dims& d, Params& P)
{
  d.n1 = cfg->n1; d.n2 = cfg->n2; d.n3 = cfg->n3;
  d.ng1 = d.n1 > 1 ? cfg->ng : 0;
  d.ng2 = d.n2 > 1 ? cfg->ng : 0;
  d.ng3 = d.n3 > 1 ? cfg->ng : 0;
  d.ns1 = d.n1 + 2 * d.ng1; d.ns2 = d.n2 + 2 * d.ng2; d.ns3 = d.n3 + 2 * d.ng3;
  d.ps1 = (d.ns1 + 15) & ~15;
  d.nplane = (long)d.ps1 * d.ns2;
  d.ncell = d.nplane * d.ns3;
  P.gam = cfg->gam; P.cour = cfg->cour; P.safe = cfg->safe; P.gammamax = cfg->gammamax; P.gammafail = cfg->gammafail;
  P.R0 = cfg->R0;
  P.bc = cfg->problem;
  P.hot2cold = cfg->hot2cold != 0;
  P.checkinv = cfg->checkinv != 0;
  P.resetregion = cfg->resetregion != 0;
  P.startx[0] = 0.0; P.dx[0] = 0.0;
  P.startx[1] = log(cfg->Rin - cfg->R0);
  P.dx[1] = (log(cfg->Rout - cfg->R0) - log(cfg->Rin - cfg->R0)) / cfg->n1;
  P.startx[2] = 0.0; P.dx[2] = 1.0 / cfg->n2;
  P.startx[3] = 0.0; P.dx[3] = 1.0 / cfg->n3;
}

// Host geometry tables over the 2-D storage plane (AoS per point).
void host_tables(const harm_config& c, const Dims& d, const Params& P, std::vector<double>* geom_aos,
                 std::vector<double>& conn_aos, std::vector<double>& gc3, std::vector<double>& rcent)
{
  const long np = (long)d.ns1 * d.ns2;  // host tables cover the logical storage plane
  for (int loc = 0; loc < HARM_NLOC; loc++) {
    geom_aos[loc].assign((size_t)np * HARM_G_NFIELD, 0.0);
#pragma omp parallel for collapse(2)
    for (int jj = 0; jj < d.ns2; jj++)
      for (int ii = 0; ii < d.ns1; ii++) {
        const int i = ii - d.ng1, j = jj - d.ng2;
        double X[4] = {0.0, P.startx[1] + (i + (loc == HARM_LOC_FACE1 ? 0.0 : 0.5)) * P.dx[1],
                       P.startx[2] + (j + (loc == HARM_LOC_FACE2 ? 0.0 : 0.5)) * P.dx[2], P.startx[3] + 0.5 * P.dx[3]};
        hg_mks_record(&c, X, &geom_aos[loc][((long)jj * d.ns1 + ii) * HARM_G_NFIELD]);
      }
  }
  gc3.assign(np, 0.0);
  for (int jj = 0; jj < d.ns2; jj++)
    for (int ii = 0; ii < d.ns1; ii++) {
      const int i = ii - d.ng1, j = jj - d.ng2;
      double X[4] = {0.0, P.startx[1] + i * P.dx[1], P.startx[2] + j * P.dx[2], 0.0};
      double rec[HARM_G_NFIELD];
      hg_mks_record(&c, X, rec);
      gc3[(size_t)jj * d.ns1 + ii] = rec[HARM_G_GDET];
    }
  conn_aos.assign((size_t)np * HARM_NCONN, 0.0);
#pragma omp parallel for collapse(2) schedule(dynamic, 16)
  for (int jj = 0; jj < d.ns2; jj++)
    for (int ii = 0; ii < d.ns1; ii++) {
      const int i = ii - d.ng1, j = jj - d.ng2;
      double X[4] = {0.0, P.startx[1] + (i + 0.5) * P.dx[1], P.startx[2] + (j + 0.5) * P.dx[2],
                     P.startx[3] + 0.5 * P.dx[3]};
      hg_mks_conn(&c, X, 0.5 * P.dx[1], 0.5 * P.dx[2], &conn_aos[((long)jj * d.ns1 + ii) * HARM_NCONN]);
    }
  rcent.assign(d.ns1, 0.0);
  for (int ii = 0; ii < d.ns1; ii++) {
    double X[4] = {0.0, P.startx[1] + (ii - d.ng1 + 0.5) * P.dx[1], 0.5, 0.0}, V[4];
    hg_mks_V(&c, X, V);
    rcent[ii] = V[1];
  }
}

harm_status build_geometry(harm_handle* h)
{
  const Dims& d = h->d;
  const harm_config& c = h->cfg;
  const long np = d.nplane;  // pitched device plane
  host_tables(c, d, h->P, h->hgeom, h->hconn, h->hgc3, h->hrcent);
  // device position of logical storage point (ii, jj), and its host index
  auto dev = [&](int ii, int jj) { return (size_t)jj * d.ps1 + ii; };
  auto hst = [&](int ii, int jj) { return (size_t)jj * d.ns1 + ii; };
  std::vector<double> soa((size_t)std::max((int)HARM_G_NFIELD, (int)HARM_NCONN) * np, 0.0);
  for (int loc = 0; loc < HARM_NLOC; loc++) {
    for (int jj = 0; jj < d.ns2; jj++)
      for (int ii = 0; ii < d.ns1; ii++)
        for (int f = 0; f < HARM_G_NFIELD; f++)
          soa[(size_t)f * np + dev(ii, jj)] = h->hgeom[loc][hst(ii, jj) * HARM_G_NFIELD + f];
    HH_TRY(alloc((void**)&h->geom[loc], sizeof(double) * HARM_G_NFIELD * np));
    HH_TRY(upload(h, h->geom[loc], soa.data(), (size_t)HARM_G_NFIELD * np));
  }
  for (int jj = 0; jj < d.ns2; jj++)
    for (int ii = 0; ii < d.ns1; ii++)
      for (int f = 0; f < HARM_NCONN; f++) soa[(size_t)f * np + dev(ii, jj)] = h->hconn[hst(ii, jj) * HARM_NCONN + f];
  HH_TRY(alloc((void**)&h->conn, sizeof(double) * HARM_NCONN * np));
  HH_TRY(upload(h, h->conn, soa.data(), (size_t)HARM_NCONN * np));
  std::vector<double> gc(np, 0.0), fl(2 * np, 0.0);
  for (int jj = 0; jj < d.ns2; jj++)
    for (int ii = 0; ii < d.ns1; ii++) {
      gc[dev(ii, jj)] = h->hgc3[hst(ii, jj)];
      const double r = h->hrcent[ii];
      fl[dev(ii, jj)] = c.rhomin * pow(r, -1.5);
      fl[np + dev(ii, jj)] = c.uumin * pow(r, -2.5);
    }
  HH_TRY(alloc((void**)&h->gcorn3, sizeof(double) * np));
  HH_TRY(upload(h, h->gcorn3, gc.data(), np));
  HH_TRY(alloc((void**)&h->floors, sizeof(double) * 2 * np));
  HH_TRY(upload(h, h->floors, fl.data(), fl.size()));
  HH_TRY(alloc((void**)&h->rcent, sizeof(double) * d.ns1));
  HH_TRY(upload(h, h->rcent, h->hrcent.data(), h->hrcent.size()));
  return HARM_OK;
}

hipEvent_t ev_get(harm_handle* h)
{
  if (h->evpool.empty()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  hipEvent_t e = h->evpool.back();
  h->evpool.pop_back();
  return e;
}

void prof_begin(harm_handle* h)
{
  if (!h->prof) return;
  h->cur0 = ev_get(h);
  if (h->cur0) (void)hipEventRecord(h->cur0, h->stream);
}

void prof_end(harm_handle* h, int kind)
{
  if (!h->prof || !h->cur0) return;
  hipEvent_t e1 = ev_get(h);
  if (!e1) return;
  (void)hipEventRecord(e1, h->stream);
  h->evrec.push_back({kind, {h->cur0, e1}});
  h->cur0 = nullptr;
}

harm_status check_handle(const harm_handle* h) { return h ? HARM_OK : hh::set_error(HARM_EINVAL, "null handle"); }

void drop_step_graph(harm_handle* h)
{
  if (h->sexec) (void)hipGraphExecDestroy(h->sexec);
  if (h->sgraph) (void)hipGraphDestroy(h->sgraph);
  h->sexec = nullptr;
  h->sgraph = nullptr;
}
harm_status check_which(int which)
{
  return (which >= 0 && which < HARM_NARR) ? HARM_OK : hh::set_error(HARM_EINVAL, "bad array index %d", which);
}

// ---- enqueue helpers (no host synchronisation) ----
// Tile shape of the face kernels.  Direction 1: the largest face count per
// block whose cells fit the LDS tile.  Directions 2 and 3: td in {4, 8, 16}
// (tc = 256/td) minimising blocks x (recon passes x 1 + face pass x 12), the
// rough cost ratio of one (cell, variable) reconstruction to one face solve;
// the i extent is covered by ceil(ni/tc) tiles.
FaceTile plan_face_tile(const Dims& d, int dir, int td_force = 0)
{
  FaceTile T{};
  T.b = face_box(d, dir);
  const Box& b = T.b;
  if (dir == 1) {
    T.nd = b.ni;
    T.ncol = b.nj * b.nk;
    int fpb = kBlock;
    while (fpb > 1 && fpb + (fpb + T.nd - 1) / T.nd + 1 > kCellMax) fpb--;
    T.fpb = fpb;
    T.nblk = ((long)T.nd * T.ncol + fpb - 1) / fpb;
    return T;
  }
  T.nd = dir == 2 ? b.nj : b.nk;
  T.nplane = dir == 2 ? b.nk : b.nj;
  T.i0 = -d.ng1;
  const int iext = b.is + b.ni - T.i0;  // i span from the aligned origin to the last face
  double best = 1e300;
  for (int td : {4, 8, 16}) {
    const int tc = kBlock / td;
    if ((td + 1) * tc > kCellMax) continue;
    if (td_force && td != td_force) continue;
    const long nti = (iext + tc - 1) / tc, ntd = (T.nd + td - 1) / td;
    const int passes = ((td + 1) * tc * HARM_NPR + kBlock - 1) / kBlock;
    const double cost = (double)(nti * ntd * T.nplane) * (passes + 12.0);
    if (cost < best) {
      best = cost;
      T.td = td;
      T.tc = tc;
      T.nti = (int)nti;
      T.ntd = (int)ntd;
    }
  }
  T.nblk = (long)T.nti * T.ntd * T.nplane;
  return T;
}

// WENO5BND face values of direction dir (its block of wl, wr)
double* wl_dir(harm_handle* h, int dir) { return h->wl + (size_t)(dir - 1) * HARM_NPR * h->d.ncell; }
double* wr_dir(harm_handle* h, int dir) { return h->wr + (size_t)(dir - 1) * HARM_NPR * h->d.ncell; }

harm_status enq_fluxcalc(harm_handle* h, int which_p, int do_ct)
{
  const Dims& d = h->d;
  hipStream_t s = h->stream;
  const double* pb = h->arr[which_p];
  hipLaunchKernelGGL(k_scal_begin_flux, dim3(1), dim3(1), 0, s, h->ds);
  const int recon = h->cfg.recon, ft = h->cfg.flux;
  const bool weno = recon == HARM_RECON_WENO5;
  if (weno) {
    prof_begin(h);
    launch_weno_dp(d, s, pb, h->geom[HARM_LOC_CENT], h->P, h->wtm, h->wdp);
    prof_end(h, HARM_K_RECON);
  }
  // WENO5BND: each direction's lines are reconstructed into its own block of
  // wl, wr just before its face kernel reads them.  Outside the timed
  // profiling run the three directions go out on three streams (forked after
  // the dP pass, joined before the CT), so the line passes along x2 and x3,
  // which wait on long division chains, share the GPU with the others; in a
  // captured step they are parallel branches of the graph.
  const bool fork = weno && !h->prof && h->side[0];
  if (fork) HH_HIP(hipEventRecord(h->evfork, s));
  auto dir_stream = [&](int dir) -> hipStream_t {
    if (!fork || dir == 1) return s;
    hipStream_t q = h->side[dir - 2];
    (void)hipStreamWaitEvent(q, h->evfork, 0);
    return q;
  };
#define LAUNCH_FLUX(DIR, F, G)                                                                              \
  do {                                                                                                      \
    hipStream_t sd = dir_stream(DIR);                                                                       \
    double* wl = weno ? wl_dir(h, DIR) : nullptr;                                                           \
    double* wr = weno ? wr_dir(h, DIR) : nullptr;                                                           \
    if (weno) {                                                                                             \
      prof_begin(h);                                                                                        \
      launch_weno_recon(DIR, d, h->ft[DIR].b, sd, pb, h->wdp, wl, wr);                                      \
      prof_end(h, HARM_K_RECON);                                                                            \
    }                                                                                                       \
    prof_begin(h);                                                                                          \
    launch_flux(DIR, recon, h->ft[DIR], sd, d, pb, F, G, h->P, ft, do_ct, h->ds, wl, wr);                   \
    prof_end(h, HARM_K_FLUX1 + DIR - 1);                                                                    \
    if (fork && DIR > 1) HH_HIP(hipEventRecord(h->evjoin[DIR - 2], sd));                                    \
  } while (0)
  if (d.n1 > 1) LAUNCH_FLUX(1, h->arr[HARM_ARR_F1], h->geom[HARM_LOC_FACE1]);
  if (d.n2 > 1) LAUNCH_FLUX(2, h->arr[HARM_ARR_F2], h->geom[HARM_LOC_FACE2]);
  if (d.n3 > 1) LAUNCH_FLUX(3, h->arr[HARM_ARR_F3], h->geom[HARM_LOC_FACE3]);
#undef LAUNCH_FLUX
  if (fork) {
    if (d.n2 > 1) HH_HIP(hipStreamWaitEvent(s, h->evjoin[0], 0));
    if (d.n3 > 1) HH_HIP(hipStreamWaitEvent(s, h->evjoin[1], 0));
  }
  HH_TRY(hh::check_launch("k_flux"));
  if (do_ct) {
    prof_begin(h);
    const int ni = h->emf_ni, nj = h->emf_nj, nk = h->emf_nk;
    hipLaunchKernelGGL(k_emf, dim3(nblocks((long)ni * nj * nk)), dim3(kBlock), 0, s, d, ni, nj, nk,
                       h->arr[HARM_ARR_F1], h->arr[HARM_ARR_F2], h->arr[HARM_ARR_F3], h->geom[HARM_LOC_FACE1],
                       h->geom[HARM_LOC_FACE2], h->gcorn3, h->emf);
    const long nct = (long)(d.n1 + (d.n1 > 1)) * (d.n2 + (d.n2 > 1)) * (d.n3 + (d.n3 > 1));
    hipLaunchKernelGGL(k_ct_all, dim3(nblocks(nct)), dim3(kBlock), 0, s, d, ni, nj, nk, h->emf,
                       h->arr[HARM_ARR_F1], h->arr[HARM_ARR_F2], h->arr[HARM_ARR_F3]);
    prof_end(h, HARM_K_EMFCT);
    HH_TRY(hh::check_launch("k_emf/k_ct_all"));
  }
  return HARM_OK;
}

harm_status enq_update(harm_handle* h, int which_pi, int which_pb, const double* CUf, const double* Cunew,
                       int timeorder, int mode)
{
  const Dims& d = h->d;
  prof_begin(h);
  const int dm = (d.n1 > 1) | ((d.n2 > 1) << 1) | ((d.n3 > 1) << 2);
#define LAUNCH_UPDATE(DM, MODE)                                                                                     \
  hipLaunchKernelGGL((k_update<DM, MODE>), dim3(nblocks((long)d.n1 * d.n2 * d.n3)), dim3(kBlock), 0, h->stream, d, \
                     h->P, h->arr[which_pi], h->arr[which_pb], h->arr[HARM_ARR_F1], h->arr[HARM_ARR_F2],           \
                     h->arr[HARM_ARR_F3], h->arr[HARM_ARR_UI], h->arr[HARM_ARR_UF], h->arr[HARM_ARR_UCUM],          \
                     h->geom[HARM_LOC_CENT], h->conn, CUf[0], CUf[1], CUf[2], Cunew[0], Cunew[1], Cunew[2],         \
                     timeorder, h->ds)
#define LAUNCH_UPDATE_MODES(DM)                         \
  switch (mode) {                                       \
    case kUpdFirst: LAUNCH_UPDATE(DM, kUpdFirst); break;   \
    case kUpdMiddle: LAUNCH_UPDATE(DM, kUpdMiddle); break; \
    case kUpdLast: LAUNCH_UPDATE(DM, kUpdLast); break;     \
    default: LAUNCH_UPDATE(DM, kUpdAll); break;            \
  }
  if (dm == 7) { LAUNCH_UPDATE_MODES(7); }
  else if (dm == 3) { LAUNCH_UPDATE_MODES(3); }
  else { LAUNCH_UPDATE_MODES(-1); }
#undef LAUNCH_UPDATE_MODES
#undef LAUNCH_UPDATE
  prof_end(h, HARM_K_UPDATE);
  return hh::check_launch("k_update");
}

// copy_all: pf = pb over the whole storage box first, as advance() does
// (advance.c:654).  Inside a step the copy is skipped: the inversion reads its
// guess from pb and writes every active zone of pf, and the boundary pass that
// follows rewrites every ghost zone of pf before anything reads it.
// ener: the final substep's floor and gamma-limit changes go to their fladdterms
harm_status enq_invert(harm_handle* h, int which_pb, int which_pf, int finalstep, bool copy_all, bool ener = false)
{
  const Dims& d = h->d;
  const size_t bytes = sizeof(double) * HARM_NPR * d.ncell;
  if (copy_all && which_pf != which_pb)
    HH_HIP(hipMemcpyAsync(h->arr[which_pf], h->arr[which_pb], bytes, hipMemcpyDeviceToDevice, h->stream));
  const double* U = h->arr[finalstep ? HARM_ARR_UCUM : HARM_ARR_UF];
  const bool diag = ener && finalstep;
  const unsigned nb = nblocks((long)d.n1 * d.n2 * d.n3);
  prof_begin(h);
  if (h->guess) {
    HH_TRY(launch_invert_grid_mlp(h->guess, d, h->P, U, h->arr[which_pb], h->arr[which_pf], h->pflag,
                                  h->geom[HARM_LOC_CENT], h->floors, finalstep, h->di, h->stream,
                                  diag ? h->fx_p0 : nullptr, h->fx_pfl, h->fx_flag));
  } else {
    hipLaunchKernelGGL(k_invert, dim3(nb), dim3(kBlock), 0, h->stream, d, h->P, U, h->arr[which_pb],
                       h->arr[which_pf], h->pflag, h->geom[HARM_LOC_CENT], finalstep, h->di);
    if (finalstep && diag)
      hipLaunchKernelGGL(k_floors<true>, dim3(nb), dim3(kBlock), 0, h->stream, d, h->P, h->arr[which_pf], h->pflag,
                         h->geom[HARM_LOC_CENT], h->floors, h->di, h->fx_p0, h->fx_pfl, h->fx_flag);
    else if (finalstep)
      hipLaunchKernelGGL(k_floors<false>, dim3(nb), dim3(kBlock), 0, h->stream, d, h->P, h->arr[which_pf], h->pflag,
                         h->geom[HARM_LOC_CENT], h->floors, h->di, nullptr, nullptr, nullptr);
  }
  prof_end(h, HARM_K_INVERT);
  if (diag)
    launch_ener_fix(d, h->P, h->arr[which_pf], h->geom[HARM_LOC_CENT], h->fx_p0, h->fx_pfl, h->fx_flag, h->enpart,
                    h->en, h->stream);
  return hh::check_launch("k_invert");
}

harm_status enq_bound(harm_handle* h, int which)
{
  const Dims& d = h->d;
  hipStream_t s = h->stream;
  double* p = h->arr[which];
  prof_begin(h);
  if (d.n1 > 1)
    hipLaunchKernelGGL(k_bound_x1, dim3(nblocks(2L * d.ng1 * d.n2 * d.n3)), dim3(kBlock), 0, s, d, h->P, p,
                       h->geom[HARM_LOC_CENT]);
  if (d.n2 > d.ng2) {
    hipLaunchKernelGGL(k_x2_poles, dim3(nblocks(2L * d.ns1 * d.n3 * d.ng2)), dim3(kBlock), 0, s, d, h->P, p,
                       h->geom[HARM_LOC_CENT], h->rcent);
  } else if (d.n2 > 1) {  // so few rows that the two poles' stencils overlap: one pole after the other
    for (int side = 0; side < 2; side++) {
      hipLaunchKernelGGL(k_x2_reflect, dim3(nblocks((long)d.ns1 * d.n3 * d.ng2)), dim3(kBlock), 0, s, d, p, side);
      hipLaunchKernelGGL(k_x2_poledeath, dim3(nblocks((long)d.ns1 * d.n3)), dim3(kBlock), 0, s, d, h->P, p,
                         h->geom[HARM_LOC_CENT], h->rcent, side);
    }
  }
  if (d.n3 > 1) hipLaunchKernelGGL(k_bound_x3, dim3(nblocks(d.nplane * 2L * d.ng3)), dim3(kBlock), 0, s, d, p);
  prof_end(h, HARM_K_BOUND);
  return hh::check_launch("bounds");
}

// post_fixup (fixup.c:53): bound the flags, average the failed zones into the
// scratch pfix (reading pf and the flags unchanged: failed neighbours are read
// too), then write them back and mark the accounted zones.
// ener: the averaged zones' changes go to their fladdterms (final substep only)
harm_status enq_post_fixup(harm_handle* h, int which, int finalstep, bool ener = false)
{
  const Dims& d = h->d;
  const unsigned nb = nblocks((long)d.n1 * d.n2 * d.n3);
  prof_begin(h);
  for (int ph = 0; ph < 2; ph++) {
    const long n = bound_pflag_items(d, ph);
    if (n > 0)
      hipLaunchKernelGGL(k_bound_pflag, dim3(nblocks(n)), dim3(kBlock), 0, h->stream, d, h->cfg.problem, h->pflag,
                         ph);
  }
  if (ener)
    hipLaunchKernelGGL(k_post_fixup<true>, dim3(nb), dim3(kBlock), 0, h->stream, d, h->P, h->arr[which], h->pflag,
                       h->pfix, h->geom[HARM_LOC_CENT], finalstep, h->enpart);
  else
    hipLaunchKernelGGL(k_post_fixup<false>, dim3(nb), dim3(kBlock), 0, h->stream, d, h->P, h->arr[which], h->pflag,
                       h->pfix, h->geom[HARM_LOC_CENT], finalstep, nullptr);
  hipLaunchKernelGGL(k_fixup_apply, dim3(nb), dim3(kBlock), 0, h->stream, d, h->arr[which], h->pflag, h->pfix,
                     finalstep);
  prof_end(h, HARM_K_FIXUP);
  if (ener) {
    int dst[4];
    for (int c = 0; c < 4; c++) dst[c] = EN_FLTERMS + kFailCats[c] * HARM_NPR;
    launch_ener_reduce(h->enpart, nb, h->en, dst, 4, h->stream);
  }
  return hh::check_launch("k_post_fixup");
}

int rk_tables(int timeorder, double CUf[4][4], double Cunew[4][4])
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

// one RK substep of advance() + post_advance, all enqueued
// in_step: the substep is part of a whole step enqueued in order (harm_step),
// so the reduced-traffic update modes apply; a lone substep (harm_advance)
// runs the reference loop as is.
harm_status enq_substep(harm_handle* h, int to, int nto, const double CUf[4][4], const double Cunew[4][4],
                        const double* wflux, bool diag, bool in_step)
{
  const int pii[3] = {HARM_ARR_P, HARM_ARR_P, HARM_ARR_P};
  int pbb[3] = {HARM_ARR_P, HARM_ARR_PK0, HARM_ARR_PK1};
  int pff[3] = {HARM_ARR_PK0, HARM_ARR_PK1, HARM_ARR_P};
  if (nto == 2) { pbb[1] = HARM_ARR_PK0; pff[1] = HARM_ARR_P; }
  const int finalstep = (to == nto - 1);
  const bool ener = h->ener_on;
  if (ener) {
    // the source sample of this substep, if any, from its starting state pb
    launch_ener_dt4(h->en, h->ds, finalstep, h->stream);
    if (finalstep)
      launch_ener_source(h->d, h->P, h->arr[pbb[to]], h->geom[HARM_LOC_CENT], h->conn, h->enpart, h->en,
                         h->stream);
  }
  HH_TRY(enq_fluxcalc(h, pbb[to], 1));
  if (ener)
    launch_ener_pdot(h->d, h->P, h->arr[HARM_ARR_F1], h->arr[HARM_ARR_F2], h->arr[HARM_ARR_F3], h->enpart, h->en,
                     h->ds, wflux[to], h->stream);
  if (diag && h->d.n1 > 1) {
    hipLaunchKernelGGL(k_fluxacc_part, dim3(kFaccBlocks), dim3(kBlock), 0, h->stream, h->d, h->arr[HARM_ARR_F1],
                       h->facc);
    hipLaunchKernelGGL(k_fluxacc_final, dim3(1), dim3(64), 0, h->stream, h->P, h->facc, h->ds, wflux[to]);
  }
  // The reduced-traffic modes rely on the tables keeping ucum at zero until
  // the last substep (true of the RK2 and RK3 tables); otherwise run the
  // reference loop as is.
  bool zero_ucum = in_step;
  for (int s = 0; s + 1 < nto; s++)
    zero_ucum = zero_ucum && Cunew[s][0] == 0.0 && Cunew[s][1] == 0.0 && Cunew[s][2] == 0.0;
  const int mode = !zero_ucum ? kUpdAll : to == 0 ? kUpdFirst : finalstep ? kUpdLast : kUpdMiddle;
  HH_TRY(enq_update(h, pii[to], pbb[to], CUf[to], Cunew[to], to, mode));
  HH_TRY(enq_invert(h, pbb[to], pff[to], finalstep, false, ener));
  hipLaunchKernelGGL(k_scal_substep_dt, dim3(1), dim3(1), 0, h->stream, h->ds);
  HH_TRY(enq_bound(h, pff[to]));
  HH_TRY(enq_post_fixup(h, pff[to], finalstep, ener && finalstep));
  HH_TRY(enq_bound(h, pff[to]));
  return hh::check_launch("substep");
}

void rk_flux_weights(int timeorder, double* w)
{
  double CUf[4][4], Cunew[4][4];
  const int n = rk_tables(timeorder, CUf, Cunew);
  double C[4][4] = {{0}};
  for (int s = 0; s < n; s++) {
    for (int r = 0; r < s; r++) C[s][r] = CUf[s][1] * C[s - 1][r];
    C[s][s] = CUf[s][2];
  }
  for (int r = 0; r < n; r++) w[r] = C[n - 1][r];
}

// ---------------------------------------------------------------------------
// Batches of independent grids stepped together.
//
// Every stage of the step (face fluxes per direction, EMFs, CT, update,
// inversion, each boundary pass, failure averaging, the dt scalars) is one
// launch over all grids of the batch: a block finds its grid by binary search
// in the stage's prefix sums of per-grid block counts and runs the same
// kernel body as the single-grid launch, on that grid's arrays.  Grids of
// different sizes, spins, field strengths and seeds share the launches, so a
// batch of small 2-D problems fills the GPU that one of them alone could not.
// A whole RK step is captured once in a hipGraph (kernel nodes only) and
// replayed.
// ---------------------------------------------------------------------------
enum { BS_FLUX1 = 0, BS_FLUX2, BS_FLUX3, BS_EMF, BS_CT, BS_FACC, BS_ZONE, BS_BX1, BS_X2R, BS_X2P,
       BS_X2M, BS_BX3, BS_ONE, BS_PFB, BS_N };

__device__ __forceinline__ void batch_decode(const int* __restrict__ pre, int ng, int b, int& gi, int& lb)
{
  int lo = 0, hi = ng - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  // the same for every thread of the block: keep it in scalar registers, so
  // the grid's descriptor is read with scalar loads
  gi = __builtin_amdgcn_readfirstlane(lo);
  lb = __builtin_amdgcn_readfirstlane(b - pre[lo]);
}

#define BATCH_DECODE(REMAP)                                                        \
  int gi, lb;                                                                      \
  batch_decode(pre, ng, (REMAP) ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x, gi, lb); \
  const GridRef& g = G[gi];

__global__ __launch_bounds__(kBlock) void kb_emf(const GridRef* __restrict__ G, const int* __restrict__ pre, int ng)
{
  BATCH_DECODE(true)
  emf_body(g.d, g.emf_ni, g.emf_nj, g.emf_nk, gmem(g.arr[HARM_ARR_F1]), gmem(g.arr[HARM_ARR_F2]), gmem(g.arr[HARM_ARR_F3]),
           gmem(g.geom[HARM_LOC_FACE1]), gmem(g.geom[HARM_LOC_FACE2]), gmem(g.gcorn3), gmem(g.emf), lb);
}

__global__ __launch_bounds__(kBlock) void kb_ct_all(const GridRef* __restrict__ G, const int* __restrict__ pre,
                                                    int ng)
{
  BATCH_DECODE(true)
  ct_all_body(g.d, g.emf_ni, g.emf_nj, g.emf_nk, gmem(g.emf), gmem(g.arr[HARM_ARR_F1]), gmem(g.arr[HARM_ARR_F2]),
              gmem(g.arr[HARM_ARR_F3]), lb);
}

__global__ __launch_bounds__(kBlock) void kb_fluxacc_part(const GridRef* __restrict__ G, const int* __restrict__ pre,
                                                          int ng)
{
  BATCH_DECODE(false)
  fluxacc_part_body(g.d, gmem(g.arr[HARM_ARR_F1]), gmem(g.facc), lb);
}

__global__ void kb_fluxacc_final(const GridRef* __restrict__ G, double weight)
{
  const GridRef& g = G[blockIdx.x];
  const int c = threadIdx.x;
  if (c >= 4 || g.d.n1 < 2) return;
  double v = 0.0;
  for (int b = 0; b < kFaccBlocks; b++) v += gmem(g.facc)[b * 4 + c];
  gmem(g.ds)[DS_FLUXACC + c] += v * g.P.dx[2] * g.P.dx[3] * gmem(g.ds)[DS_DT] * weight;
}


template <int DM, int MODE>
__global__ __launch_bounds__(kBlock) void kb_update(const GridRef* __restrict__ G, const int* __restrict__ pre, int ng,
                                                    int which_pi, int which_pb, double cuf0, double cuf1, double cuf2,
                                                    double cun0, double cun1, double cun2, int timeorder)
{
  BATCH_DECODE(true)
  update_body<DM, MODE>(g.d, g.P, gmem(g.arr[which_pi]), gmem(g.arr[which_pb]), gmem(g.arr[HARM_ARR_F1]),
                        gmem(g.arr[HARM_ARR_F2]), gmem(g.arr[HARM_ARR_F3]), gmem(g.arr[HARM_ARR_UI]),
                        gmem(g.arr[HARM_ARR_UF]), gmem(g.arr[HARM_ARR_UCUM]), gmem(g.geom[HARM_LOC_CENT]),
                        gmem(g.conn), cuf0, cuf1, cuf2, cun0, cun1, cun2, timeorder, gmem(g.ds), lb);
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void kb_invert(
    const GridRef* __restrict__ G, const int* __restrict__ pre, int ng, int which_pb, int which_pf, int finalstep)
{
  BATCH_DECODE(true)
  invert_body(g.d, g.P, gmem(g.arr[finalstep ? HARM_ARR_UCUM : HARM_ARR_UF]), gmem(g.arr[which_pb]),
              gmem(g.arr[which_pf]), gmem(g.pflag), gmem(g.geom[HARM_LOC_CENT]), finalstep, gmem(g.di), lb);
}

__global__ __launch_bounds__(kBlock) void kb_floors(const GridRef* __restrict__ G, const int* __restrict__ pre, int ng,
                                                    int which_pf)
{
  BATCH_DECODE(true)
  floors_body(g.d, g.P, gmem(g.arr[which_pf]), gmem(g.pflag), gmem(g.geom[HARM_LOC_CENT]), gmem(g.floors),
              gmem(g.di), lb);
}

__global__ __launch_bounds__(kBlock) void kb_bound_pflag(const GridRef* __restrict__ G, const int* __restrict__ pre,
                                                         int ng, int phase)
{
  BATCH_DECODE(false)
  const long t = (long)lb * kBlock + threadIdx.x;
  if (t < bound_pflag_items(g.d, phase)) bound_pflag_body(g.d, g.P.bc, gmem(g.pflag), phase, t);
}

__global__ __launch_bounds__(kBlock) void kb_post_fixup(const GridRef* __restrict__ G, const int* __restrict__ pre,
                                                        int ng, int which, int finalstep)
{
  BATCH_DECODE(false)
  post_fixup_body(g.d, g.P, gmem(g.arr[which]), gmem(g.pflag), gmem(g.pfix), gmem(g.geom[HARM_LOC_CENT]), finalstep,
                  lb);
}

__global__ __launch_bounds__(kBlock) void kb_fixup_apply(const GridRef* __restrict__ G, const int* __restrict__ pre,
                                                         int ng, int which, int finalstep)
{
  BATCH_DECODE(false)
  fixup_apply_body(g.d, gmem(g.arr[which]), gmem(g.pflag), gmem(g.pfix), finalstep, lb);
}

__global__ __launch_bounds__(kBlock) void kb_bound_x1(const GridRef* __restrict__ G, const int* __restrict__ pre,
                                                      int ng, int which)
{
  BATCH_DECODE(false)
  bound_x1_body(g.d, g.P, gmem(g.arr[which]), gmem(g.geom[HARM_LOC_CENT]), lb);
}

__global__ __launch_bounds__(kBlock) void kb_x2_reflect(const GridRef* __restrict__ G, const int* __restrict__ pre,
                                                        int ng, int which, int side)
{
  BATCH_DECODE(false)
  x2_reflect_body(g.d, gmem(g.arr[which]), side, lb);
}

__global__ __launch_bounds__(kBlock) void kb_x2_poledeath(const GridRef* __restrict__ G, const int* __restrict__ pre,
                                                          int ng, int which, int side)
{
  BATCH_DECODE(false)
  x2_poledeath_body(g.d, g.P, gmem(g.arr[which]), gmem(g.geom[HARM_LOC_CENT]), gmem(g.rcent), side, lb);
}

__global__ __launch_bounds__(kBlock) void kb_x2_poles(const GridRef* __restrict__ G, const int* __restrict__ pre,
                                                      int ng, int which)
{
  BATCH_DECODE(false)
  x2_poles_body(g.d, g.P, gmem(g.arr[which]), gmem(g.geom[HARM_LOC_CENT]), gmem(g.rcent), lb);
}

__global__ __launch_bounds__(kBlock) void kb_bound_x3(const GridRef* __restrict__ G, const int* __restrict__ pre,
                                                      int ng, int which)
{
  BATCH_DECODE(false)
  bound_x3_body(g.d, gmem(g.arr[which]), lb);
}

// one thread per grid
__global__ void kb_scal(const GridRef* __restrict__ G, int ng, int what)
{
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= ng) return;
  double* ds = G[gi].ds;
  if (what == 0) {  // begin flux
    ds[DS_NDT1] = kBIG; ds[DS_NDT2] = kBIG; ds[DS_NDT3] = kBIG; ds[DS_ACCDT] = kBIG;
  } else if (what == 1) {  // begin step
    ds[DS_LASTNDT] = kBIG;
    G[gi].di[DI_NSTROKE] = 0;
    G[gi].di[DI_NFAIL] = 0;
    G[gi].di[DI_NFALLBACK] = 0;
  } else if (what == 2) {  // substep dt
    const double wavedt = 1. / (1. / ds[DS_NDT1] + 1. / ds[DS_NDT2] + 1. / ds[DS_NDT3]);
    const double nd = 1.0 * fmin(wavedt, ds[DS_ACCDT]);
    if (nd < ds[DS_LASTNDT]) ds[DS_LASTNDT] = nd;
  } else {  // end step
    const double safe = G[gi].P.safe;
    double fullndt = ds[DS_LASTNDT];
    const double dt = ds[DS_DT];
    if (fullndt > safe * dt) fullndt = safe * dt;
    ds[DS_DTSTEP] = dt;
    ds[DS_T] = ds[DS_T] + dt;
    ds[DS_DT] = fullndt;
    ds[DS_NSTEP] = ds[DS_NSTEP] + 1.0;
    unsigned long long* di = G[gi].di;
    di[DI_NSTROKE_SUM] += di[DI_NSTROKE];
    di[DI_NFAIL_SUM] += di[DI_NFAIL];
    di[DI_NFALLBACK_SUM] += di[DI_NFALLBACK];
  }
}

GridRef make_ref(const harm_handle* h)
{
  GridRef r{};
  r.d = h->d;
  r.P = h->P;
  r.fluxtype = h->cfg.flux;
  for (int a = 0; a < HARM_NARR; a++) r.arr[a] = h->arr[a];
  r.pflag = h->pflag;
  r.pfix = h->pfix;
  for (int l = 0; l < HARM_NLOC; l++) r.geom[l] = h->geom[l];
  r.gcorn3 = h->gcorn3;
  r.conn = h->conn;
  r.floors = h->floors;
  r.rcent = h->rcent;
  r.emf = h->emf;
  r.emf_ni = h->emf_ni;
  r.emf_nj = h->emf_nj;
  r.emf_nk = h->emf_nk;
  r.ds = h->ds;
  r.di = h->di;
  r.facc = h->facc;
  r.wl = h->wl;
  r.wr = h->wr;
  for (int dir = 0; dir < 4; dir++) r.ft[dir] = h->ft[dir];
  return r;
}

}  // namespace

struct harm_batch {
  std::vector<harm_handle*> grids;
  hipStream_t stream = nullptr;
  GridRef* dref = nullptr;
  int* dpre = nullptr;  // BS_N prefix arrays of ng+1 ints
  int ntot[BS_N] = {0};
  int recon = 0, timeorder = 0, dm = 0;
  bool x2_merged = true;  // every grid has n2 > ng2 (or no polar boundary): both poles in one launch
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
};

namespace {

harm_status batch_build(harm_batch* b)
{
  const int ng = (int)b->grids.size();
  std::vector<GridRef> refs(ng);
  std::vector<int> pre((size_t)BS_N * (ng + 1), 0);
  for (int gi = 0; gi < ng; gi++) {
    const harm_handle* h = b->grids[gi];
    const Dims& d = h->d;
    refs[gi] = make_ref(h);
    long cnt[BS_N] = {0};
    const int e0 = d.n1 > 1, e1 = d.n2 > 1, e2 = d.n3 > 1;
    cnt[BS_FLUX1] = e0 ? h->ft[1].nblk : 0;
    cnt[BS_FLUX2] = e1 ? h->ft[2].nblk : 0;
    cnt[BS_FLUX3] = e2 ? h->ft[3].nblk : 0;
    cnt[BS_EMF] = nblocks((long)h->emf_ni * h->emf_nj * h->emf_nk);
    cnt[BS_CT] = nblocks((long)(d.n1 + e0) * (d.n2 + e1) * (d.n3 + e2));
    cnt[BS_FACC] = e0 ? kFaccBlocks : 0;
    cnt[BS_ZONE] = nblocks((long)d.n1 * d.n2 * d.n3);
    cnt[BS_BX1] = e0 ? nblocks(2L * d.ng1 * d.n2 * d.n3) : 0;
    cnt[BS_X2R] = e1 ? nblocks((long)d.ns1 * d.n3 * d.ng2) : 0;
    cnt[BS_X2P] = e1 ? nblocks((long)d.ns1 * d.n3) : 0;
    cnt[BS_X2M] = e1 ? nblocks(2L * d.ns1 * d.n3 * d.ng2) : 0;
    if (e1 && d.n2 <= d.ng2) b->x2_merged = false;
    cnt[BS_BX3] = e2 ? nblocks(d.nplane * 2L * d.ng3) : 0;
    cnt[BS_ONE] = 1;
    cnt[BS_PFB] = nblocks(std::max(bound_pflag_items(d, 0), bound_pflag_items(d, 1)));
    for (int st = 0; st < BS_N; st++) {
      const long nx = (long)pre[(size_t)st * (ng + 1) + gi] + cnt[st];
      if (nx > 0x7fffffffL) return hh::set_error(HARM_EINVAL, "harm_batch: too many blocks in one stage");
      pre[(size_t)st * (ng + 1) + gi + 1] = (int)nx;
    }
  }
  for (int st = 0; st < BS_N; st++) b->ntot[st] = pre[(size_t)st * (ng + 1) + ng];
  if (!b->dref) HH_HIP(hipMalloc((void**)&b->dref, sizeof(GridRef) * ng));
  if (!b->dpre) HH_HIP(hipMalloc((void**)&b->dpre, sizeof(int) * pre.size()));
  HH_HIP(hipMemcpy(b->dref, refs.data(), sizeof(GridRef) * ng, hipMemcpyHostToDevice));
  HH_HIP(hipMemcpy(b->dpre, pre.data(), sizeof(int) * pre.size(), hipMemcpyHostToDevice));
  return HARM_OK;
}

harm_status enq_batch_substep(harm_batch* b, int to, int nto, const double CUf[4][4], const double Cunew[4][4],
                              const double* wflux)
{
  const int ng = (int)b->grids.size();
  hipStream_t s = b->stream;
  const GridRef* G = b->dref;
  auto P = [&](int st) { return b->dpre + (size_t)st * (ng + 1); };
  const dim3 blk(kBlock);
  const int pii[3] = {HARM_ARR_P, HARM_ARR_P, HARM_ARR_P};
  int pbb[3] = {HARM_ARR_P, HARM_ARR_PK0, HARM_ARR_PK1};
  int pff[3] = {HARM_ARR_PK0, HARM_ARR_PK1, HARM_ARR_P};
  if (nto == 2) { pbb[1] = HARM_ARR_PK0; pff[1] = HARM_ARR_P; }
  const int finalstep = (to == nto - 1);
  const unsigned g1 = (unsigned)((ng + 63) / 64);
  hipLaunchKernelGGL(kb_scal, dim3(g1), dim3(64), 0, s, G, ng, 0);
  // fluxes
  // WENO5BND: every grid's lines of a direction are reconstructed (one
  // launch per grid) before the batched face kernel of that direction
  const bool weno = b->recon == HARM_RECON_WENO5;
  if (weno)
    for (harm_handle* h : b->grids) launch_weno_dp(h->d, s, h->arr[pbb[to]], h->geom[HARM_LOC_CENT], h->P, h->wtm, h->wdp);
#define BFLUX(DIR)                                                                                         \
  if (b->ntot[BS_FLUX1 + DIR - 1] > 0) {                                                                    \
    if (weno)                                                                                               \
      for (harm_handle* h : b->grids)                                                                       \
        if (h->ft[DIR].nblk > 0 && (DIR == 1 ? h->d.n1 : DIR == 2 ? h->d.n2 : h->d.n3) > 1)                 \
          launch_weno_recon(DIR, h->d, h->ft[DIR].b, s, h->arr[pbb[to]], h->wdp, wl_dir(h, DIR), wr_dir(h, DIR)); \
    launch_flux_batch(DIR, b->recon, (unsigned)b->ntot[BS_FLUX1 + DIR - 1], s, G, P(BS_FLUX1 + DIR - 1), ng, \
                      pbb[to], 1);                                                                          \
  }
  BFLUX(1)
  BFLUX(2)
  BFLUX(3)
#undef BFLUX
  hipLaunchKernelGGL(kb_emf, dim3((unsigned)b->ntot[BS_EMF]), blk, 0, s, G, P(BS_EMF), ng);
  if (b->ntot[BS_CT]) hipLaunchKernelGGL(kb_ct_all, dim3((unsigned)b->ntot[BS_CT]), blk, 0, s, G, P(BS_CT), ng);
  if (b->ntot[BS_FACC]) {
    hipLaunchKernelGGL(kb_fluxacc_part, dim3((unsigned)b->ntot[BS_FACC]), blk, 0, s, G, P(BS_FACC), ng);
    hipLaunchKernelGGL(kb_fluxacc_final, dim3((unsigned)ng), dim3(64), 0, s, G, wflux[to]);
  }
  // update (the reduced-traffic modes: the RK2/RK3 tables keep ucum at zero
  // until the last substep)
  bool zero_ucum = true;
  for (int st = 0; st + 1 < nto; st++)
    zero_ucum = zero_ucum && Cunew[st][0] == 0.0 && Cunew[st][1] == 0.0 && Cunew[st][2] == 0.0;
  const int mode = !zero_ucum ? kUpdAll : to == 0 ? kUpdFirst : finalstep ? kUpdLast : kUpdMiddle;
  const dim3 gz((unsigned)b->ntot[BS_ZONE]);
#define BUPD(DM, MODE)                                                                                            \
  hipLaunchKernelGGL((kb_update<DM, MODE>), gz, blk, 0, s, G, P(BS_ZONE), ng, pii[to], pbb[to], CUf[to][0],      \
                     CUf[to][1], CUf[to][2], Cunew[to][0], Cunew[to][1], Cunew[to][2], to)
#define BUPD_MODES(DM)                                   \
  switch (mode) {                                        \
    case kUpdFirst: BUPD(DM, kUpdFirst); break;          \
    case kUpdMiddle: BUPD(DM, kUpdMiddle); break;        \
    case kUpdLast: BUPD(DM, kUpdLast); break;            \
    default: BUPD(DM, kUpdAll); break;                   \
  }
  if (b->dm == 7) { BUPD_MODES(7); }
  else if (b->dm == 3) { BUPD_MODES(3); }
  else { BUPD_MODES(-1); }
#undef BUPD_MODES
#undef BUPD
  hipLaunchKernelGGL(kb_invert, gz, blk, 0, s, G, P(BS_ZONE), ng, pbb[to], pff[to], finalstep);
  if (finalstep) hipLaunchKernelGGL(kb_floors, gz, blk, 0, s, G, P(BS_ZONE), ng, pff[to]);
  hipLaunchKernelGGL(kb_scal, dim3(g1), dim3(64), 0, s, G, ng, 2);
  // bounds, failure averaging, bounds
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      for (int ph = 0; ph < 2; ph++)
        hipLaunchKernelGGL(kb_bound_pflag, dim3((unsigned)b->ntot[BS_PFB]), blk, 0, s, G, P(BS_PFB), ng, ph);
      hipLaunchKernelGGL(kb_post_fixup, gz, blk, 0, s, G, P(BS_ZONE), ng, pff[to], finalstep);
      hipLaunchKernelGGL(kb_fixup_apply, gz, blk, 0, s, G, P(BS_ZONE), ng, pff[to], finalstep);
    }
    if (b->ntot[BS_BX1]) hipLaunchKernelGGL(kb_bound_x1, dim3((unsigned)b->ntot[BS_BX1]), blk, 0, s, G, P(BS_BX1), ng,
                                            pff[to]);
    if (b->ntot[BS_X2M] && b->x2_merged)
      hipLaunchKernelGGL(kb_x2_poles, dim3((unsigned)b->ntot[BS_X2M]), blk, 0, s, G, P(BS_X2M), ng, pff[to]);
    else if (b->ntot[BS_X2R])
      for (int side = 0; side < 2; side++) {
        hipLaunchKernelGGL(kb_x2_reflect, dim3((unsigned)b->ntot[BS_X2R]), blk, 0, s, G, P(BS_X2R), ng, pff[to], side);
        hipLaunchKernelGGL(kb_x2_poledeath, dim3((unsigned)b->ntot[BS_X2P]), blk, 0, s, G, P(BS_X2P), ng, pff[to],
                           side);
      }
    if (b->ntot[BS_BX3]) hipLaunchKernelGGL(kb_bound_x3, dim3((unsigned)b->ntot[BS_BX3]), blk, 0, s, G, P(BS_BX3), ng,
                                            pff[to]);
  }
  return hh::check_launch("batch substep");
}

harm_status enq_batch_step(harm_batch* b)
{
  const int ng = (int)b->grids.size();
  double CUf[4][4], Cunew[4][4], w[4];
  const int nto = rk_tables(b->timeorder, CUf, Cunew);
  rk_flux_weights(b->timeorder, w);
  hipLaunchKernelGGL(kb_scal, dim3((unsigned)((ng + 63) / 64)), dim3(64), 0, b->stream, b->dref, ng, 1);
  for (int to = 0; to < nto; to++) HH_TRY(enq_batch_substep(b, to, nto, CUf, Cunew, w));
  hipLaunchKernelGGL(kb_scal, dim3((unsigned)((ng + 63) / 64)), dim3(64), 0, b->stream, b->dref, ng, 3);
  return hh::check_launch("batch step");
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" harm_status harm_create(const harm_config* cfg, harm_handle** out)
{
  if (!cfg || !out) return hh::set_error(HARM_EINVAL, "harm_create: null argument");
  if (cfg->n1 < 1 || cfg->n2 < 1 || cfg->n3 < 1 || cfg->ng < 3)
    return hh::set_error(HARM_EINVAL, "harm_create: bad sizes n=(%d,%d,%d) ng=%d", cfg->n1, cfg->n2, cfg->n3, cfg->ng);
  if (cfg->n1 > 1 && cfg->n1 < 3) return hh::set_error(HARM_EINVAL, "harm_create: n1 must be >= 3");
  if (cfg->n2 > 1 && cfg->n2 < 4) return hh::set_error(HARM_EINVAL, "harm_create: n2 must be >= 4 (pole death)");
  if (cfg->n3 > 1 && (cfg->n3 % 2 != 0 || cfg->n3 / 2 < cfg->ng))
    return hh::set_error(HARM_EINVAL, "harm_create: n3 must be even and n3/2 >= ng for the polar rotation");
  if (cfg->timeorder != 2 && cfg->timeorder != 3) return hh::set_error(HARM_EINVAL, "harm_create: timeorder 2 or 3");
  if (cfg->problem != HARM_PROBLEM_TORUS && cfg->problem != HARM_PROBLEM_BONDI)
    return hh::set_error(HARM_EINVAL, "harm_create: unknown problem %d", cfg->problem);
  if (cfg->problem == HARM_PROBLEM_BONDI) {
    // the boundary integrals and their inversion hold for this flow only
    if (cfg->a != 0.0 || fabs(cfg->gam - 4.0 / 3.0) > 1e-15 || cfg->n3 != 1 || cfg->n1 < 2)
      return hh::set_error(HARM_EINVAL, "harm_create: the Bondi problem needs a = 0, gam = 4/3, n3 = 1");
    const double dx1 = (log(cfg->Rout - cfg->R0) - log(cfg->Rin - cfg->R0)) / cfg->n1;
    const double r_in_lo = cfg->R0 + (cfg->Rin - cfg->R0) * exp(-(cfg->ng - 0.5) * dx1);
    const double r_in_hi = cfg->R0 + (cfg->Rin - cfg->R0) * exp(-0.5 * dx1);
    const double r_out_lo = cfg->R0 + (cfg->Rout - cfg->R0) * exp(0.5 * dx1);
    if (!(r_in_lo > 0.35 && r_in_hi < 2.5 && r_out_lo > 9.0))
      return hh::set_error(HARM_EINVAL,
                           "harm_create: Bondi ghost zones at r in [%g, %g] and beyond %g; the integral inversion "
                           "needs 0.35 < r < 2.5 inside and r > 9 outside", r_in_lo, r_in_hi, r_out_lo);
  }
  if (cfg->recon < HARM_RECON_DONOR || cfg->recon > HARM_RECON_WENO5 || cfg->recon == HARM_RECON_MC)
    return hh::set_error(HARM_EINVAL, "harm_create: recon %d not supported", cfg->recon);
  if (cfg->recon == HARM_RECON_WENO5 && cfg->ng != 4)
    return hh::set_error(HARM_EINVAL, "harm_create: WENO5BND lines are built for 4 ghost zones (MAXBND 4), ng = %d",
                         cfg->ng);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return hh::set_error(HARM_ENOGPU, "harm_create: no GPU");
  harm_handle* h = new harm_handle();
  h->cfg = *cfg;
  if (hipGetDevice(&h->dev) != hipSuccess) {
    delete h;
    return hh::set_error(HARM_ENOGPU, "harm_create: hipGetDevice failed");
  }
  Dims& d = h->d;
  make_dims(cfg, h->d, h->P);
  harm_status st = HARM_OK;
#define TRYC(x)                    \
  do {                             \
    st = (x);                      \
    if (st != HARM_OK) goto fail;  \
  } while (0)
  TRYC(hh::check_hip(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate"));
  for (int a = 0; a < HARM_NARR; a++) {
    TRYC(alloc((void**)&h->arr[a], sizeof(double) * HARM_NPR * d.ncell));
    TRYC(hh::check_hip(hipMemset(h->arr[a], 0, sizeof(double) * HARM_NPR * d.ncell), "hipMemset"));
  }
  TRYC(alloc((void**)&h->pflag, sizeof(int32_t) * d.ncell));
  TRYC(reset_pflag(h));
  TRYC(alloc((void**)&h->pfix, sizeof(double) * HARM_NPR * d.ncell));
  h->emf_ni = d.n1 + (d.n1 > 1);
  h->emf_nj = d.n2 + (d.n2 > 1);
  h->emf_nk = d.n3 + (d.n3 > 1);
  TRYC(alloc((void**)&h->emf, sizeof(double) * 3 * (size_t)h->emf_ni * h->emf_nj * h->emf_nk));
  if (cfg->recon == HARM_RECON_WENO5) {
    TRYC(alloc((void**)&h->wl, sizeof(double) * 3 * HARM_NPR * d.ncell));
    TRYC(alloc((void**)&h->wr, sizeof(double) * 3 * HARM_NPR * d.ncell));
    for (int q = 0; q < 2; q++) {
      TRYC(hh::check_hip(hipStreamCreateWithFlags(&h->side[q], hipStreamNonBlocking), "hipStreamCreate"));
      TRYC(hh::check_hip(hipEventCreateWithFlags(&h->evjoin[q], hipEventDisableTiming), "hipEventCreate"));
    }
    TRYC(hh::check_hip(hipEventCreateWithFlags(&h->evfork, hipEventDisableTiming), "hipEventCreate"));
    TRYC(alloc((void**)&h->wdp, sizeof(double) * 3 * d.ncell));
    TRYC(alloc((void**)&h->wtm, sizeof(double) * 10 * d.ncell));
    TRYC(hh::check_hip(hipMemset(h->wl, 0, sizeof(double) * 3 * HARM_NPR * d.ncell), "hipMemset"));
    TRYC(hh::check_hip(hipMemset(h->wr, 0, sizeof(double) * 3 * HARM_NPR * d.ncell), "hipMemset"));
    TRYC(hh::check_hip(hipMemset(h->wdp, 0, sizeof(double) * 3 * d.ncell), "hipMemset"));
  }
  TRYC(alloc((void**)&h->ds, sizeof(double) * DS_N));
  TRYC(alloc((void**)&h->facc, sizeof(double) * 4 * kFaccBlocks));
  TRYC(alloc((void**)&h->en, sizeof(double) * EN_N));
  TRYC(hh::check_hip(hipMemset(h->en, 0, sizeof(double) * EN_N), "hipMemset"));
  TRYC(alloc((void**)&h->enpart, sizeof(double) * 4 * 8 * ener_part_blocks(d)));  // 4 channels
  TRYC(alloc((void**)&h->entot, sizeof(double) * 10));
  TRYC(hh::check_hip(hipMemset(h->ds, 0, sizeof(double) * DS_N), "hipMemset"));
  TRYC(alloc((void**)&h->di, sizeof(unsigned long long) * DI_N));
  TRYC(hh::check_hip(hipMemset(h->di, 0, sizeof(unsigned long long) * DI_N), "hipMemset"));
  TRYC(build_geometry(h));
  for (int dir = 1; dir <= 3; dir++) h->ft[dir] = plan_face_tile(d, dir);
#undef TRYC
  *out = h;
  return HARM_OK;
fail:
  harm_destroy(h);
  return st;
}

extern "C" harm_status harm_destroy(harm_handle* h)
{
  if (!h) return HARM_OK;
  // teardown: errors here are not actionable, the handle is gone either way
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_step_graph(h);
  for (int a = 0; a < HARM_NARR; a++) (void)hipFree(h->arr[a]);
  for (int l = 0; l < HARM_NLOC; l++) (void)hipFree(h->geom[l]);
  for (auto& r : h->evrec) {
    (void)hipEventDestroy(r.second.first);
    (void)hipEventDestroy(r.second.second);
  }
  for (hipEvent_t e : h->evpool) (void)hipEventDestroy(e);
  void* bufs[] = {h->pflag, h->pfix, h->gcorn3, h->conn, h->floors, h->rcent, h->emf, h->wl, h->wr, h->wdp, h->wtm, h->facc,
                  h->cf, h->jcon, h->xfer,
                  h->en, h->enpart, h->entot, h->fx_p0, h->fx_pfl, h->fx_flag, h->ds, h->di};
  for (void* b : bufs) (void)hipFree(b);
  for (int q = 0; q < 2; q++) {
    if (h->side[q]) (void)hipStreamDestroy(h->side[q]);
    if (h->evjoin[q]) (void)hipEventDestroy(h->evjoin[q]);
  }
  if (h->evfork) (void)hipEventDestroy(h->evfork);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return HARM_OK;
}

extern "C" harm_status harm_sync(harm_handle* h)
{
  HH_TRY(check_handle(h));
  return hh::check_hip(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
}

extern "C" harm_status harm_get_stream(harm_handle* h, void** stream)
{
  HH_TRY(check_handle(h));
  *stream = (void*)h->stream;
  return HARM_OK;
}

extern "C" harm_status harm_storage_dims(const harm_handle* h, int32_t dims[3])
{
  HH_TRY(check_handle(h));
  dims[0] = h->d.ns1;
  dims[1] = h->d.ns2;
  dims[2] = h->d.ns3;
  return HARM_OK;
}

extern "C" harm_status harm_get_geometry(harm_handle* h, int32_t loc, double* out)
{
  HH_TRY(check_handle(h));
  if (loc < 0 || loc >= HARM_NLOC) return hh::set_error(HARM_EINVAL, "bad location %d", loc);
  memcpy(out, h->hgeom[loc].data(), sizeof(double) * h->hgeom[loc].size());
  return HARM_OK;
}

extern "C" harm_status harm_get_connection(harm_handle* h, double* out)
{
  HH_TRY(check_handle(h));
  memcpy(out, h->hconn.data(), sizeof(double) * h->hconn.size());
  return HARM_OK;
}

extern "C" harm_status harm_get_corner_gdet(harm_handle* h, double* out)
{
  HH_TRY(check_handle(h));
  memcpy(out, h->hgc3.data(), sizeof(double) * h->hgc3.size());
  return HARM_OK;
}

extern "C" harm_status harm_get_coords(harm_handle* h, int32_t i, int32_t j, int32_t k, double X[4], double V[4])
{
  HH_TRY(check_handle(h));
  X[0] = 0.0;
  X[1] = h->P.startx[1] + (i + 0.5) * h->P.dx[1];
  X[2] = h->P.startx[2] + (j + 0.5) * h->P.dx[2];
  X[3] = h->P.startx[3] + (k + 0.5) * h->P.dx[3];
  hg_mks_V(&h->cfg, X, V);
  return HARM_OK;
}

// device position of logical storage cell c = (kk*ns2 + jj)*ns1 + ii
static inline size_t dev_cell(const Dims& d, long c)
{
  const long row = c / d.ns1;
  return (size_t)row * d.ps1 + (size_t)(c - row * d.ns1);
}

namespace {

// HARM's interleaved prim/cons layout (a[c][pl], c = (kk*ns2 + jj)*ns1 + ii
// over the storage box) <-> the device planes (a[pl][row*ps1 + ii]).  One
// thread per device cell: the planes are read or written coalesced, the
// interleaved side as 64 contiguous bytes per thread.  Row padding
// (ii >= ns1) is written as zero on the way in.
template <bool TO_PLANES>
__global__ __launch_bounds__(kBlock) void k_xfer(Dims d, const double* __restrict__ src, double* __restrict__ dst)
{
  const long o = (long)blockIdx.x * kBlock + threadIdx.x;
  if (o >= d.ncell) return;
  const long row = o / d.ps1, ii = o - row * d.ps1;
  const bool in = row < (long)d.ns2 * d.ns3 && ii < d.ns1;
  const long c = row * d.ns1 + ii;
  if constexpr (TO_PLANES) {
    for (int pl = 0; pl < HARM_NPR; pl++) dst[(long)pl * d.ncell + o] = in ? src[c * HARM_NPR + pl] : 0.0;
  } else {
    if (!in) return;
    for (int pl = 0; pl < HARM_NPR; pl++) dst[c * HARM_NPR + pl] = src[(long)pl * d.ncell + o];
  }
}

harm_status xfer_buffer(harm_handle* h)
{
  if (!h->xfer) HH_TRY(alloc((void**)&h->xfer, sizeof(double) * HARM_NPR * (size_t)h->d.ns1 * h->d.ns2 * h->d.ns3));
  return HARM_OK;
}

}  // namespace

extern "C" harm_status harm_set_array(harm_handle* h, int32_t which, const double* host_aos)
{
  HH_TRY(check_handle(h));
  HH_TRY(check_which(which));
  const Dims& d = h->d;
  const size_t n = (size_t)HARM_NPR * d.ns1 * d.ns2 * d.ns3;
  HH_TRY(xfer_buffer(h));
  HH_HIP(hipMemcpyAsync(h->xfer, host_aos, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_xfer<true>, dim3((unsigned)((d.ncell + kBlock - 1) / kBlock)), dim3(kBlock), 0, h->stream, d,
                     (const double*)h->xfer, h->arr[which]);
  HH_HIP(hipGetLastError());
  HH_HIP(hipStreamSynchronize(h->stream));
  return HARM_OK;
}

extern "C" harm_status harm_get_array(harm_handle* h, int32_t which, double* host_aos)
{
  HH_TRY(check_handle(h));
  HH_TRY(check_which(which));
  const Dims& d = h->d;
  const size_t n = (size_t)HARM_NPR * d.ns1 * d.ns2 * d.ns3;
  HH_TRY(xfer_buffer(h));
  hipLaunchKernelGGL(k_xfer<false>, dim3((unsigned)((d.ncell + kBlock - 1) / kBlock)), dim3(kBlock), 0, h->stream, d,
                     (const double*)h->arr[which], h->xfer);
  HH_HIP(hipGetLastError());
  HH_HIP(hipMemcpyAsync(host_aos, h->xfer, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  HH_HIP(hipStreamSynchronize(h->stream));
  return HARM_OK;
}

extern "C" harm_status harm_get_pflag(harm_handle* h, int32_t* host_flags)
{
  HH_TRY(check_handle(h));
  const Dims& d = h->d;
  std::vector<int32_t> f(d.ncell);
  HH_HIP(hipStreamSynchronize(h->stream));
  HH_HIP(hipMemcpy(f.data(), h->pflag, sizeof(int32_t) * d.ncell, hipMemcpyDeviceToHost));
  const long nl = (long)d.ns1 * d.ns2 * d.ns3;
  for (long c = 0; c < nl; c++) host_flags[c] = f[dev_cell(d, c)];
  return HARM_OK;
}

extern "C" harm_status harm_set_time(harm_handle* h, double t, double dt, int64_t nstep)
{
  HH_TRY(check_handle(h));
  double v[DS_N];
  HH_HIP(hipStreamSynchronize(h->stream));
  HH_HIP(hipMemcpy(v, h->ds, sizeof(v), hipMemcpyDeviceToHost));
  v[DS_DT] = dt;
  v[DS_T] = t;
  v[DS_NSTEP] = (double)nstep;
  HH_HIP(hipMemcpy(h->ds, v, sizeof(v), hipMemcpyHostToDevice));
  unsigned long long z[DI_N];  // the Newton sums restart with the clock
  HH_HIP(hipMemcpy(z, h->di, sizeof(z), hipMemcpyDeviceToHost));
  z[DI_NSTROKE_SUM] = z[DI_NFAIL_SUM] = z[DI_NFALLBACK_SUM] = 0;
  return hh::check_hip(hipMemcpy(h->di, z, sizeof(z), hipMemcpyHostToDevice), "set_time");
}

extern "C" harm_status harm_get_time(harm_handle* h, double* t, double* dt, int64_t* nstep)
{
  HH_TRY(check_handle(h));
  double v[DS_N];
  HH_HIP(hipStreamSynchronize(h->stream));
  HH_HIP(hipMemcpy(v, h->ds, sizeof(v), hipMemcpyDeviceToHost));
  if (t) *t = v[DS_T];
  if (dt) *dt = v[DS_DT];
  if (nstep) *nstep = (int64_t)v[DS_NSTEP];
  return HARM_OK;
}

extern "C" harm_status harm_get_newton_stats(harm_handle* h, harm_newton_stats* last_step, harm_newton_stats* total)
{
  HH_TRY(check_handle(h));
  unsigned long long z[DI_N];
  HH_HIP(hipStreamSynchronize(h->stream));
  HH_HIP(hipMemcpy(z, h->di, sizeof(z), hipMemcpyDeviceToHost));
  if (last_step) {
    last_step->nstroke = (int64_t)z[DI_NSTROKE];
    last_step->nfail = (int64_t)z[DI_NFAIL];
  }
  if (total) {
    total->nstroke = (int64_t)z[DI_NSTROKE_SUM];
    total->nfail = (int64_t)z[DI_NFAIL_SUM];
  }
  return HARM_OK;
}

// n zones drawn (with replacement, a seeded hash of the draw index) from the
// active zones: U and p at the zone, and its 27-field geometry record, into
// SoA buffers of n entries (the layout of harm_invert_batch and
// harm_mlp_pairs).
__global__ __launch_bounds__(kBlock) void k_sample_zones(Dims d, const double* __restrict__ U,
                                                         const double* __restrict__ p,
                                                         const double* __restrict__ gcent, long n, uint64_t seed,
                                                         double* __restrict__ Uo, double* __restrict__ go,
                                                         double* __restrict__ po)
{
  const long s = (long)blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  uint64_t x = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(s + 1);  // splitmix64 of the draw index
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  const int nact = d.n1 * d.n2 * d.n3;
  int i, j, k;
  zone_ijk(d, (int)(x % (uint64_t)nact), i, j, k);
  const long o = hd::cidx(d, i, j, k), pp = hd::pidx(d, i, j);
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) {
    Uo[pl * n + s] = U[pl * d.ncell + o];
    po[pl * n + s] = p[pl * d.ncell + o];
  }
  for (int f = 0; f < HARM_G_NFIELD; f++) go[f * n + s] = gcent[f * d.nplane + pp];
}

extern "C" harm_status harm_grid_sample_zones(harm_handle* h, int32_t which_U, int32_t which_p, int64_t n,
                                              uint64_t seed, double* U, double* geom, double* p)
{
  HH_TRY(check_handle(h));
  HH_TRY(check_which(which_U));
  HH_TRY(check_which(which_p));
  if (n <= 0 || !U || !geom || !p) return hh::set_error(HARM_EINVAL, "harm_grid_sample_zones: bad arguments");
  hipLaunchKernelGGL(k_sample_zones, dim3(nblocks(n)), dim3(kBlock), 0, h->stream, h->d, h->arr[which_U],
                     h->arr[which_p], h->geom[HARM_LOC_CENT], (long)n, seed, U, geom, p);
  HH_TRY(hh::check_launch("k_sample_zones"));
  return harm_sync(h);
}

extern "C" harm_status harm_set_inversion_guess(harm_handle* h, const harm_mlp* m)
{
  HH_TRY(check_handle(h));
  HH_HIP(hipStreamSynchronize(h->stream));
  h->guess = m;
  drop_step_graph(h);  // the captured step holds the other inversion kernel
  return HARM_OK;
}

extern "C" harm_status harm_get_guess_stats(harm_handle* h, int64_t* fallbacks_last_step, int64_t* fallbacks_total)
{
  HH_TRY(check_handle(h));
  unsigned long long z[DI_N];
  HH_HIP(hipStreamSynchronize(h->stream));
  HH_HIP(hipMemcpy(z, h->di, sizeof(z), hipMemcpyDeviceToHost));
  if (fallbacks_last_step) *fallbacks_last_step = (int64_t)z[DI_NFALLBACK];
  if (fallbacks_total) *fallbacks_total = (int64_t)z[DI_NFALLBACK_SUM];
  return HARM_OK;
}

extern "C" harm_status harm_init_torus(harm_handle* h, double rin, double rmax, double kappa, double beta,
                                       double randfact, uint32_t seed)
{
  HH_TRY(check_handle(h));
  const Dims& d = h->d;
  std::vector<double> p((size_t)HARM_NPR * d.ns1 * d.ns2 * d.ns3, 0.0);
  hi_grid G;
  G.ng1 = d.ng1; G.ng2 = d.ng2; G.ng3 = d.ng3;
  for (int a = 0; a < 4; a++) { G.startx[a] = h->P.startx[a]; G.dx[a] = h->P.dx[a]; }
  if (hi_init_torus(&h->cfg, &G, rin, rmax, kappa, beta, randfact, seed, p.data()) != 0)
    return hh::set_error(HARM_EINVAL, "harm_init_torus: no torus zones on this grid");
  HH_TRY(reset_pflag(h));  // a new run's flags
  HH_TRY(harm_set_array(h, HARM_ARR_P, p.data()));
  HH_TRY(enq_bound(h, HARM_ARR_P));
  return harm_sync(h);
}

extern "C" harm_status harm_fluxcalc(harm_handle* h, int32_t which_p, int32_t do_ct, double ndt[3])
{
  HH_TRY(check_handle(h));
  HH_TRY(check_which(which_p));
  HH_TRY(enq_fluxcalc(h, which_p, do_ct));
  if (ndt) {
    double v[DS_N];
    HH_HIP(hipStreamSynchronize(h->stream));
    HH_HIP(hipMemcpy(v, h->ds, sizeof(v), hipMemcpyDeviceToHost));
    ndt[0] = v[DS_NDT1];
    ndt[1] = v[DS_NDT2];
    ndt[2] = v[DS_NDT3];
  }
  return HARM_OK;
}

extern "C" harm_status harm_update(harm_handle* h, int32_t which_pi, int32_t which_pb, const double CUf[4],
                                   const double Cunew[4], double dt, int32_t timeorder, double* accdt)
{
  HH_TRY(check_handle(h));
  HH_TRY(check_which(which_pi));
  HH_TRY(check_which(which_pb));
  double v[DS_N];
  HH_HIP(hipStreamSynchronize(h->stream));
  HH_HIP(hipMemcpy(v, h->ds, sizeof(v), hipMemcpyDeviceToHost));
  v[DS_DT] = dt;
  v[DS_ACCDT] = kBIG;
  HH_HIP(hipMemcpy(h->ds, v, sizeof(v), hipMemcpyHostToDevice));
  HH_TRY(enq_update(h, which_pi, which_pb, CUf, Cunew, timeorder, kUpdAll));
  if (accdt) {
    HH_HIP(hipStreamSynchronize(h->stream));
    HH_HIP(hipMemcpy(v, h->ds, sizeof(v), hipMemcpyDeviceToHost));
    *accdt = v[DS_ACCDT];
  }
  return HARM_OK;
}

extern "C" harm_status harm_invert_grid(harm_handle* h, int32_t which_pb, int32_t which_pf, int32_t finalstep,
                                        harm_newton_stats* stats)
{
  HH_TRY(check_handle(h));
  HH_TRY(check_which(which_pb));
  HH_TRY(check_which(which_pf));
  unsigned long long z[DI_N];
  HH_HIP(hipStreamSynchronize(h->stream));
  HH_HIP(hipMemcpy(z, h->di, sizeof(z), hipMemcpyDeviceToHost));
  z[DI_NSTROKE] = z[DI_NFAIL] = z[DI_NFALLBACK] = 0;  // this call's counts (the step sums stay)
  HH_HIP(hipMemcpy(h->di, z, sizeof(z), hipMemcpyHostToDevice));
  HH_TRY(enq_invert(h, which_pb, which_pf, finalstep, true));
  if (stats) {
    HH_HIP(hipStreamSynchronize(h->stream));
    HH_HIP(hipMemcpy(z, h->di, sizeof(z), hipMemcpyDeviceToHost));
    stats->nstroke = (int64_t)z[DI_NSTROKE];
    stats->nfail = (int64_t)z[DI_NFAIL];
  }
  return HARM_OK;
}

extern "C" harm_status harm_bound(harm_handle* h, int32_t which)
{
  HH_TRY(check_handle(h));
  HH_TRY(check_which(which));
  return enq_bound(h, which);
}

extern "C" harm_status harm_post_fixup(harm_handle* h, int32_t which, int32_t finalstep)
{
  HH_TRY(check_handle(h));
  HH_TRY(check_which(which));
  return enq_post_fixup(h, which, finalstep != 0);
}

extern "C" harm_status harm_advance(harm_handle* h, int32_t timeorder, double* ndt)
{
  HH_TRY(check_handle(h));
  double CUf[4][4], Cunew[4][4], w[4];
  const int nto = rk_tables(h->cfg.timeorder, CUf, Cunew);
  if (timeorder < 0 || timeorder >= nto) return hh::set_error(HARM_EINVAL, "harm_advance: bad timeorder");
  rk_flux_weights(h->cfg.timeorder, w);
  if (timeorder == 0) hipLaunchKernelGGL(k_scal_begin_step, dim3(1), dim3(1), 0, h->stream, h->ds, h->di);
  HH_TRY(enq_substep(h, timeorder, nto, CUf, Cunew, w, false, false));
  if (ndt) {
    double v[DS_N];
    HH_HIP(hipStreamSynchronize(h->stream));
    HH_HIP(hipMemcpy(v, h->ds, sizeof(v), hipMemcpyDeviceToHost));
    *ndt = v[DS_LASTNDT];
  }
  return HARM_OK;
}

namespace {
harm_status enq_step(harm_handle* h)
{
  double CUf[4][4], Cunew[4][4], w[4];
  const int nto = rk_tables(h->cfg.timeorder, CUf, Cunew);
  rk_flux_weights(h->cfg.timeorder, w);
  hipLaunchKernelGGL(k_scal_begin_step, dim3(1), dim3(1), 0, h->stream, h->ds, h->di);
  for (int to = 0; to < nto; to++) HH_TRY(enq_substep(h, to, nto, CUf, Cunew, w, true, true));
  hipLaunchKernelGGL(k_scal_end_step, dim3(1), dim3(1), 0, h->stream, h->ds, h->di, h->P.safe);
  return hh::check_launch("harm_step");
}
}  // namespace

// A step is a fixed sequence of ~75 kernel launches with arguments that do not
// change between steps, so it is captured once in a hipGraph (kernel nodes
// only) and replayed; with kernel timing enabled the launches go out one by one
// so that the timing events can bracket them.
extern "C" harm_status harm_step(harm_handle* h, int32_t nsteps)
{
  HH_TRY(check_handle(h));
  if (nsteps < 0) return hh::set_error(HARM_EINVAL, "harm_step: nsteps < 0");
  if (h->prof) {
    for (int n = 0; n < nsteps; n++) HH_TRY(enq_step(h));
    return HARM_OK;
  }
  if (nsteps == 0) return HARM_OK;
  if (!h->sexec) {
    HH_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const harm_status st = enq_step(h);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (st != HARM_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    HH_HIP(e);
    h->sgraph = g;
    HH_HIP(hipGraphInstantiate(&h->sexec, h->sgraph, nullptr, nullptr, 0));
  }
  for (int n = 0; n < nsteps; n++) HH_HIP(hipGraphLaunch(h->sexec, h->stream));
  return HARM_OK;
}

extern "C" harm_status harm_get_boundary_fluxes(harm_handle* h, double out[4])
{
  HH_TRY(check_handle(h));
  double v[DS_N];
  HH_HIP(hipStreamSynchronize(h->stream));
  HH_HIP(hipMemcpy(v, h->ds, sizeof(v), hipMemcpyDeviceToHost));
  for (int c = 0; c < 4; c++) out[c] = v[DS_FLUXACC + c];
  return HARM_OK;
}

extern "C" harm_status harm_set_ener_tracking(harm_handle* h, int32_t on)
{
  HH_TRY(check_handle(h));
  HH_HIP(hipStreamSynchronize(h->stream));
  if (on) {  // the fix-up states, allocated on first use (each one that is missing)
    const size_t n = (size_t)HARM_NPR * h->d.ncell;
    if (!h->fx_p0) HH_TRY(alloc((void**)&h->fx_p0, sizeof(double) * n));
    if (!h->fx_pfl) HH_TRY(alloc((void**)&h->fx_pfl, sizeof(double) * n));
    if (!h->fx_flag) HH_TRY(alloc((void**)&h->fx_flag, (size_t)h->d.ncell));
  }
  const bool was_on = h->ener_on;
  if (was_on != (on != 0)) drop_step_graph(h);  // the captured step holds the other kernel set
  h->ener_on = on != 0;
  // switching tracking on starts the sums from zero; a call that leaves it on
  // keeps them (the sums a restart file restored, for one)
  return (!was_on && h->ener_on) ? harm_reset_ener(h) : HARM_OK;
}

extern "C" harm_status harm_reset_ener(harm_handle* h)
{
  HH_TRY(check_handle(h));
  HH_HIP(hipStreamSynchronize(h->stream));
  return hh::check_hip(hipMemset(h->en, 0, sizeof(double) * EN_N), "harm_reset_ener");
}

extern "C" harm_status harm_get_ener(harm_handle* h, int32_t which, harm_ener* out)
{
  HH_TRY(check_handle(h));
  if (!out) return hh::set_error(HARM_EINVAL, "harm_get_ener: null argument");
  if (which < 0 || which >= HARM_NARR) return hh::set_error(HARM_EINVAL, "harm_get_ener: bad array %d", which);
  launch_ener_totals(h->d, h->P, h->arr[which], h->geom[HARM_LOC_CENT], h->enpart, h->entot, h->stream);
  HH_TRY(hh::check_launch("harm_get_ener"));
  double en[EN_N], tot[10], v[DS_N];
  HH_HIP(hipStreamSynchronize(h->stream));
  HH_HIP(hipMemcpy(en, h->en, sizeof(en), hipMemcpyDeviceToHost));
  HH_HIP(hipMemcpy(tot, h->entot, sizeof(tot), hipMemcpyDeviceToHost));
  HH_HIP(hipMemcpy(v, h->ds, sizeof(v), hipMemcpyDeviceToHost));
  out->t = v[DS_T];
  out->nstep = (int64_t)v[DS_NSTEP];
  for (int pl = 0; pl < HARM_NPR; pl++) {
    out->U[pl] = tot[pl];
    out->fladd[pl] = 0.0;
    for (int c = 0; c < FL_N; c++) {
      out->fladdterms[c][pl] = en[EN_FLTERMS + c * HARM_NPR + pl];
      out->fladd[pl] += out->fladdterms[c][pl];
    }
    out->sourceadd[pl] = en[EN_SRCADD + pl];
    for (int b = 0; b < 6; b++) {
      out->pdot[b][pl] = en[EN_PDOT + b * HARM_NPR + pl];
      out->pcum[b][pl] = en[EN_PCUM + b * HARM_NPR + pl];
    }
  }
  out->divbmax = tot[8];
  out->divbavg = tot[9] / ((double)h->d.n1 * h->d.n2 * h->d.n3);
  return HARM_OK;
}

// ---------------------------------------------------------------------------
// HARM-format dumps (dump.c)
// ---------------------------------------------------------------------------
namespace {
// coordinates of a zone centre, columns ti tj tk X1..X3 r th ph (bl_coord_ijk_2)
void dump_coords(harm_handle* h, int i, int j, int k, double* row)
{
  double X[4], V[4];
  harm_get_coords(h, i, j, k, X, V);
  row[0] = i; row[1] = j; row[2] = k;
  for (int m = 1; m <= 3; m++) { row[2 + m] = X[m]; row[5 + m] = V[m]; }
}

// the 2-D host geometry record of active zone (i, j)
const double* host_rec(const harm_handle* h, int loc, int i, int j)
{
  const Dims& d = h->d;
  return &h->hgeom[loc][((size_t)(j + d.ng2) * d.ns1 + (i + d.ng1)) * HARM_G_NFIELD];
}

// dump_header_general (dump.c:321-431), text form
harm_status dump_header(harm_handle* h, FILE* f, int whichdump, int ncol)
{
  double v[DS_N];
  // the clock after the work already on the handle's (non-blocking) stream
  HH_HIP(hipStreamSynchronize(h->stream));
  HH_HIP(hipMemcpy(v, h->ds, sizeof(v), hipMemcpyDeviceToHost));
  const harm_config& c = h->cfg;
  fprintf(f,
          "%21.15g %d %d %d %21.15g %21.15g %21.15g %21.15g %21.15g %21.15g %ld %21.15g %21.15g %21.15g %21.15g "
          "%21.15g %21.15g %21.15g %d %21.15g %21.15g %d %d %d %d %d %d %d %d %d\n",
          v[DS_T], c.n1, c.n2, c.n3, h->P.startx[1], h->P.startx[2], h->P.startx[3], h->P.dx[1], h->P.dx[2],
          h->P.dx[3], (long)v[DS_NSTEP], c.gam, c.a, c.R0, c.Rin, c.Rout, c.hslope, v[DS_DT], 0 /* LOGRSINTH */,
          1.0 /* MBH */, 0.0 /* QBH */, 0, c.n1 - 1, 0, c.n2 - 1, 0, c.n3 - 1, whichdump, 0, ncol);
  return HARM_OK;
}

harm_status write_rows(FILE* f, const double* rows, long nz, int ncol)
{
  for (long z = 0; z < nz; z++) {
    const double* r = rows + z * ncol;
    for (int c = 0; c < ncol; c++) fprintf(f, "%21.15g ", r[c]);
    fprintf(f, "\n");
  }
  return ferror(f) ? hh::set_error(HARM_EINVAL, "dump: write failed") : HARM_OK;
}
}  // namespace

extern "C" harm_status harm_dump_columns(harm_handle* h, int32_t which, double* rows)
{
  HH_TRY(check_handle(h));
  if (!rows) return hh::set_error(HARM_EINVAL, "harm_dump_columns: null argument");
  if (which < 0 || which >= HARM_NARR) return hh::set_error(HARM_EINVAL, "harm_dump_columns: bad array %d", which);
  const Dims& d = h->d;
  const long nz = (long)d.n1 * d.n2 * d.n3;
  double* dcol = nullptr;
  HH_TRY(alloc((void**)&dcol, sizeof(double) * DC_N * nz));
  launch_dump_cols(d, h->P, h->arr[which], h->arr[HARM_ARR_UCUM], h->geom[HARM_LOC_CENT], h->jcon, dcol, h->stream);
  std::vector<double> col((size_t)DC_N * nz);
  harm_status st = hh::check_launch("k_dump_cols");
  if (st == HARM_OK) st = hh::check_hip(hipStreamSynchronize(h->stream), "harm_dump_columns");
  if (st == HARM_OK)
    st = hh::check_hip(hipMemcpy(col.data(), dcol, sizeof(double) * col.size(), hipMemcpyDeviceToHost), "copy");
  (void)hipFree(dcol);
  if (st != HARM_OK) return st;
  // dump_content's column order: coordinates, then the device columns
  for (long t = 0; t < nz; t++) {
    const int i = (int)(t % d.n1), j = (int)((t / d.n1) % d.n2), k = (int)(t / ((long)d.n1 * d.n2));
    double* r = rows + t * HARM_DUMP_NCOL;
    dump_coords(h, i, j, k, r);
    for (int q = 0; q < DC_N; q++) r[kDumpCoordCols + q] = col[(size_t)q * nz + t];
  }
  return HARM_OK;
}

// ---------------------------------------------------------------------------
// harm_dump_columns with no allocation: the rows are built on the device in a
// static buffer of kNoallocZones rows (one per device, 37.7 MB, held until the
// process ends) and copied into the caller's rows.  Calls that overlap on one
// device share the buffer and race.
// ---------------------------------------------------------------------------
namespace {
constexpr long kNoallocZones = 65536;
static __device__ __align__(256) double scratch[kNoallocZones * HARM_DUMP_NCOL];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {nullptr};

// the device address of `scratch` on device dev, looked up once per device
// with that device current (each device holds its own copy of the array)
harm_status scratch_on(int dev, double** out)
{
  if (dev < 0 || dev >= kMaxDevices) return hh::set_error(HARM_EINVAL, "noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    HH_HIP(hipGetDevice(&cur));
    if (cur != dev) HH_HIP(hipSetDevice(dev));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) HH_HIP(hipSetDevice(cur));
    HH_TRY(hh::check_hip(e, "hipGetSymbolAddress(scratch)"));
    g_scratch_ptr[dev] = (double*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return HARM_OK;
}
}  // namespace

extern "C" harm_status harm_dump_columns_noalloc(harm_handle* h, int32_t which, double* rows)
{
  HH_TRY(check_handle(h));
  if (!rows) return hh::set_error(HARM_EINVAL, "harm_dump_columns_noalloc: null argument");
  if (which < 0 || which >= HARM_NARR)
    return hh::set_error(HARM_EINVAL, "harm_dump_columns_noalloc: bad array %d", which);
  const Dims& d = h->d;
  const long nz = (long)d.n1 * d.n2 * d.n3;
  if (nz > kNoallocZones)
    return hh::set_error(HARM_EINVAL, "harm_dump_columns_noalloc: %ld zones, above the static buffer's %ld", nz,
                         kNoallocZones);
  double* dev = nullptr;
  HH_TRY(scratch_on(h->dev, &dev));
  // every column of every row the copy takes is written: the kernel fills
  // the physics columns (the four-current included), the host the
  // coordinates after the copy
  launch_dump_rows(d, h->P, h->arr[which], h->arr[HARM_ARR_UCUM], h->geom[HARM_LOC_CENT], h->jcon, dev, h->stream);
  HH_TRY(hh::check_launch("k_dump_cols"));
  HH_HIP(hipMemcpyAsync(rows, dev, sizeof(double) * HARM_DUMP_NCOL * nz, hipMemcpyDeviceToHost, h->stream));
  HH_HIP(hipStreamSynchronize(h->stream));
  for (long t = 0; t < nz; t++) {
    const int i = (int)(t % d.n1), j = (int)((t / d.n1) % d.n2), k = (int)(t / ((long)d.n1 * d.n2));
    dump_coords(h, i, j, k, rows + t * HARM_DUMP_NCOL);
  }
  return HARM_OK;
}

// The four-current the main dump carries (CALCFARADAYANDCURRENTS 1,
// CURRENTCALC1): at the start of a run (compute_currents_t0,
// initbase.c:1414, CURRENTST0) and after each step that ends at a main dump
// (post_stepch's block, step_ch.c:156-170): the Faraday samples of every
// zone, then J^mu from centred differences and, from the third call on, the
// time derivative over the last three samples (dtl: the step dt at the
// previous call, dtr: now).
extern "C" harm_status harm_currents(harm_handle* h, int32_t at_start)
{
  HH_TRY(check_handle(h));
  const Dims& d = h->d;
  if (!h->cf) {
    HH_TRY(alloc((void**)&h->cf, sizeof(double) * 18 * d.ncell));
    HH_TRY(alloc((void**)&h->jcon, sizeof(double) * 4 * d.ncell));
    HH_HIP(hipMemsetAsync(h->cf, 0, sizeof(double) * 18 * d.ncell, h->stream));
    HH_HIP(hipMemsetAsync(h->jcon, 0, sizeof(double) * 4 * d.ncell, h->stream));
  }
  double v[DS_N];
  HH_HIP(hipStreamSynchronize(h->stream));
  HH_HIP(hipMemcpy(v, h->ds, sizeof(v), hipMemcpyDeviceToHost));
  const double dt = at_start ? v[DS_DT] : v[DS_DTSTEP];
  launch_current_precalc(d, h->P, h->arr[HARM_ARR_P], h->geom[HARM_LOC_CENT], h->cf, h->stream);
  launch_current_calc(d, h->P, h->cf, h->geom[HARM_LOC_CENT], h->cur_calls >= 2, h->cur_lastdt, dt, h->jcon,
                      h->stream);
  HH_TRY(hh::check_launch("k_current"));
  h->cur_calls++;
  h->cur_lastdt = dt;
  return HARM_OK;
}

extern "C" harm_status harm_gdump_columns(harm_handle* h, double* rows)
{
  HH_TRY(check_handle(h));
  if (!rows) return hh::set_error(HARM_EINVAL, "harm_gdump_columns: null argument");
  const Dims& d = h->d;
  const long nz = (long)d.n1 * d.n2 * d.n3;
  for (long t = 0; t < nz; t++) {
    const int i = (int)(t % d.n1), j = (int)((t / d.n1) % d.n2), k = (int)(t / ((long)d.n1 * d.n2));
    double* r = rows + t * HARM_GDUMP_NCOL;
    dump_coords(h, i, j, k, r);
    int c = 9;
    const double* cn = &h->hconn[((size_t)(j + d.ng2) * d.ns1 + (i + d.ng1)) * HARM_NCONN];
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++)
        for (int e = 0; e < 4; e++) r[c++] = cn[a * 10 + hg_sym(b, e)];
    const double* g = host_rec(h, HARM_LOC_CENT, i, j);
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) r[c++] = g[HARM_G_GCON + hg_sym(a, b)];
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) r[c++] = g[HARM_G_GCOV + hg_sym(a, b)];
    r[c++] = g[HARM_G_GDET];
    for (int a = 0; a < 4; a++) r[c++] = 0.0;  // conn2: no second connection with gdet in the equations
    double X[4], V[4], J[4][4];
    harm_get_coords(h, i, j, k, X, V);
    hg_mks_dxdxp(&h->cfg, X, J);
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) r[c++] = J[a][b];
  }
  return HARM_OK;
}

extern "C" harm_status harm_write_dump(harm_handle* h, int32_t which, const char* path)
{
  HH_TRY(check_handle(h));
  if (!path) return hh::set_error(HARM_EINVAL, "harm_write_dump: null path");
  const long nz = (long)h->d.n1 * h->d.n2 * h->d.n3;
  std::vector<double> rows((size_t)nz * HARM_DUMP_NCOL);
  HH_TRY(harm_dump_columns(h, which, rows.data()));
  FILE* f = fopen(path, "w");
  if (!f) return hh::set_error(HARM_EINVAL, "harm_write_dump: cannot open %s", path);
  harm_status st = dump_header(h, f, 3 /* MAINDUMPTYPE */, HARM_DUMP_NCOL);
  if (st == HARM_OK) st = write_rows(f, rows.data(), nz, HARM_DUMP_NCOL);
  fclose(f);
  return st;
}

extern "C" harm_status harm_write_gdump(harm_handle* h, const char* path)
{
  HH_TRY(check_handle(h));
  if (!path) return hh::set_error(HARM_EINVAL, "harm_write_gdump: null path");
  const long nz = (long)h->d.n1 * h->d.n2 * h->d.n3;
  std::vector<double> rows((size_t)nz * HARM_GDUMP_NCOL);
  HH_TRY(harm_gdump_columns(h, rows.data()));
  FILE* f = fopen(path, "w");
  if (!f) return hh::set_error(HARM_EINVAL, "harm_write_gdump: cannot open %s", path);
  harm_status st = dump_header(h, f, 4 /* GRIDDUMPTYPE */, HARM_GDUMP_NCOL);
  if (st == HARM_OK) st = write_rows(f, rows.data(), nz, HARM_GDUMP_NCOL);
  fclose(f);
  return st;
}

// ---------------------------------------------------------------------------
// Restart files (restart.c)
// ---------------------------------------------------------------------------
namespace {
// lim[] codes of global.nondepnmemonics.h:527-541 for HARM_RECON_*
const int kLimCode[4] = {0 /* DONOR */, 3 /* MC */, 4 /* PARA */, 10 /* WENO5 */};
// pdot, pcum, fladdterms, sourceadd, the source-sample clock (en[0..EN_WILLBE])
constexpr int kRestartEn = EN_WILLBE + 1;
static_assert(EN_PCUM == 48 && EN_FLTERMS == 96 && EN_SRCADD == 192 && EN_DT4 == 200, "restart block order");

struct RestartHead {
  int n[3], lim, timeorder, fluxmethod, defcoord, binary;
  long nstep;
  double t, tf, a, MBH, QBH, gam, gamideal, cour, dt, R0, Rin, Rout, hslope;
  double en[kRestartEn];
};

// the active zones of `which` in rdump order (i fastest), `per` values apart
void active_rows(const harm_handle* h, const std::vector<double>& aos, double* out, int per, int off)
{
  const Dims& d = h->d;
  const long nz = (long)d.n1 * d.n2 * d.n3;
  for (long t = 0; t < nz; t++) {
    const long i = t % d.n1, j = (t / d.n1) % d.n2, k = t / ((long)d.n1 * d.n2);
    const long c = ((k + d.ng3) * d.ns2 + (j + d.ng2)) * d.ns1 + (i + d.ng1);
    for (int pl = 0; pl < HARM_NPR; pl++) out[t * per + off + pl] = aos[c * HARM_NPR + pl];
  }
}
}  // namespace

extern "C" harm_status harm_write_restart(harm_handle* h, const char* path, int32_t binary, double tf)
{
  HH_TRY(check_handle(h));
  if (!path) return hh::set_error(HARM_EINVAL, "harm_write_restart: null path");
  if (binary != HARM_TEXTOUTPUT && binary != HARM_MIXEDOUTPUT)
    return hh::set_error(HARM_EINVAL, "harm_write_restart: binary must be 0 (text) or 2 (mixed), got %d", binary);
  const Dims& d = h->d;
  const long nl = (long)d.ns1 * d.ns2 * d.ns3, nz = (long)d.n1 * d.n2 * d.n3;
  std::vector<double> aos((size_t)nl * HARM_NPR), rows((size_t)nz * 2 * HARM_NPR);
  HH_TRY(harm_get_array(h, HARM_ARR_P, aos.data()));
  active_rows(h, aos, rows.data(), 2 * HARM_NPR, 0);
  HH_TRY(harm_get_array(h, HARM_ARR_UCUM, aos.data()));
  active_rows(h, aos, rows.data(), 2 * HARM_NPR, HARM_NPR);
  double v[DS_N], en[EN_N];
  HH_HIP(hipMemcpy(v, h->ds, sizeof(v), hipMemcpyDeviceToHost));
  HH_HIP(hipMemcpy(en, h->en, sizeof(en), hipMemcpyDeviceToHost));
  const harm_config& c = h->cfg;
  FILE* f = fopen(path, binary ? "wb" : "w");
  if (!f) return hh::set_error(HARM_EINVAL, "harm_write_restart: cannot open %s", path);
  const int lim = kLimCode[c.recon];
  fprintf(f, "%d %d %d %31.26g %31.26g %ld %31.26g %31.26g %31.26g %26.20g %26.20g %26.20g %31.26g %d %d %d %d %d ",
          c.n1, c.n2, c.n3, v[DS_T], tf, (long)v[DS_NSTEP], c.a, 1.0 /* MBH */, 0.0 /* QBH */, c.gam, c.gam, c.cour,
          v[DS_DT], lim, lim, lim, c.timeorder, c.flux);
  fprintf(f, "%d %26.20g %26.20g %26.20g %26.20g %d ", 0 /* defcoord LOGRSINTH */, c.R0, c.Rin, c.Rout, c.hslope,
          binary);
  for (int q = 0; q < kRestartEn; q++) fprintf(f, "%31.26g ", en[q]);
  fprintf(f, "\n");
  if (binary) {
    fwrite(rows.data(), sizeof(double), rows.size(), f);
  } else {
    for (long z = 0; z < nz; z++) {
      for (int q = 0; q < 2 * HARM_NPR; q++) fprintf(f, "%21.15g ", rows[z * 2 * HARM_NPR + q]);
      fprintf(f, "\n");
    }
  }
  const bool bad = ferror(f) != 0;
  fclose(f);
  return bad ? hh::set_error(HARM_EINVAL, "harm_write_restart: write failed on %s", path) : HARM_OK;
}

extern "C" harm_status harm_read_restart(harm_handle* h, const char* path, double* tf)
{
  HH_TRY(check_handle(h));
  if (!path) return hh::set_error(HARM_EINVAL, "harm_read_restart: null path");
  FILE* f = fopen(path, "rb");
  if (!f) return hh::set_error(HARM_EINVAL, "harm_read_restart: cannot open %s", path);
  RestartHead r;
  int lim2, lim3;
  int got = fscanf(f, "%d %d %d %lf %lf %ld %lf %lf %lf %lf %lf %lf %lf %d %d %d %d %d %d %lf %lf %lf %lf %d", &r.n[0],
                   &r.n[1], &r.n[2], &r.t, &r.tf, &r.nstep, &r.a, &r.MBH, &r.QBH, &r.gam, &r.gamideal, &r.cour, &r.dt,
                   &r.lim, &lim2, &lim3, &r.timeorder, &r.fluxmethod, &r.defcoord, &r.R0, &r.Rin, &r.Rout, &r.hslope,
                   &r.binary);
  int q = 0;
  while (got == 24 && q < kRestartEn && fscanf(f, "%lf", &r.en[q]) == 1) q++;
  got += q;
  int ch;
  while (got == 24 + kRestartEn && (ch = fgetc(f)) != EOF && ch != '\n') {
  }
  if (got != 24 + kRestartEn) {
    fclose(f);
    return hh::set_error(HARM_EINVAL, "harm_read_restart: %s: header has %d of %d fields", path, got,
                         24 + kRestartEn);
  }
  const harm_config& c = h->cfg;
  const char* bad = nullptr;
  if (r.n[0] != c.n1) bad = "N1";
  else if (r.n[1] != c.n2) bad = "N2";
  else if (r.n[2] != c.n3) bad = "N3";
  else if (r.a != c.a) bad = "a";
  else if (r.gam != c.gam) bad = "gam";
  else if (r.lim != kLimCode[c.recon] || lim2 != r.lim || lim3 != r.lim) bad = "lim";
  else if (r.timeorder != c.timeorder) bad = "TIMEORDER";
  else if (r.fluxmethod != c.flux) bad = "fluxmethod";
  else if (r.defcoord != 0) bad = "defcoord";
  else if (r.R0 != c.R0 || r.Rin != c.Rin || r.Rout != c.Rout || r.hslope != c.hslope) bad = "grid map";
  else if (r.binary != HARM_TEXTOUTPUT && r.binary != HARM_MIXEDOUTPUT) bad = "binaryoutput";
  if (bad) {
    fclose(f);
    return hh::set_error(HARM_EINVAL, "harm_read_restart: %s: %s differs from the handle's", path, bad);
  }
  const Dims& d = h->d;
  const long nl = (long)d.ns1 * d.ns2 * d.ns3, nz = (long)d.n1 * d.n2 * d.n3;
  std::vector<double> rows((size_t)nz * 2 * HARM_NPR);
  size_t nread = 0;
  if (r.binary) {
    nread = fread(rows.data(), sizeof(double), rows.size(), f);
  } else {
    while (nread < rows.size() && fscanf(f, "%lf", &rows[nread]) == 1) nread++;
  }
  fclose(f);
  if (nread != rows.size())
    return hh::set_error(HARM_EINVAL, "harm_read_restart: %s: %zu of %zu values", path, nread, rows.size());
  std::vector<double> aos((size_t)nl * HARM_NPR, 0.0);
  for (int half = 0; half < 2; half++) {
    for (long t = 0; t < nz; t++) {
      const long i = t % d.n1, j = (t / d.n1) % d.n2, k = t / ((long)d.n1 * d.n2);
      const long cc = ((k + d.ng3) * d.ns2 + (j + d.ng2)) * d.ns1 + (i + d.ng1);
      for (int pl = 0; pl < HARM_NPR; pl++) aos[cc * HARM_NPR + pl] = rows[t * 2 * HARM_NPR + half * HARM_NPR + pl];
    }
    HH_TRY(harm_set_array(h, half ? HARM_ARR_UCUM : HARM_ARR_P, aos.data()));
  }
  HH_TRY(reset_pflag(h));  // a restarted run's flags start as a new run's (rdump holds none)
  HH_TRY(enq_bound(h, HARM_ARR_P));
  HH_TRY(hh::check_launch("harm_read_restart: bound"));
  HH_TRY(harm_set_time(h, r.t, r.dt, r.nstep));
  double en[EN_N] = {0.0};
  for (int q = 0; q < kRestartEn; q++) en[q] = r.en[q];
  HH_HIP(hipMemcpy(h->en, en, sizeof(en), hipMemcpyHostToDevice));
  if (tf) *tf = r.tf;
  return HARM_OK;
}

extern "C" harm_status harm_host_tables(const harm_config* cfg, double* geom_aos, double* conn_aos,
                                        double* corner_gdet, double* rcent, double startx[4], double dx[4])
{
  if (!cfg) return hh::set_error(HARM_EINVAL, "harm_host_tables: null config");
  Dims d;
  Params P;
  make_dims(cfg, d, P);
  std::vector<double> g[HARM_NLOC], c, gc3, rc;
  host_tables(*cfg, d, P, g, c, gc3, rc);
  if (geom_aos)
    for (int loc = 0; loc < HARM_NLOC; loc++)
      memcpy(geom_aos + (size_t)loc * d.ns1 * d.ns2 * HARM_G_NFIELD, g[loc].data(), sizeof(double) * g[loc].size());
  if (conn_aos) memcpy(conn_aos, c.data(), sizeof(double) * c.size());
  if (corner_gdet) memcpy(corner_gdet, gc3.data(), sizeof(double) * gc3.size());
  if (rcent) memcpy(rcent, rc.data(), sizeof(double) * rc.size());
  for (int a = 0; a < 4; a++) {
    if (startx) startx[a] = P.startx[a];
    if (dx) dx[a] = P.dx[a];
  }
  return HARM_OK;
}

extern "C" harm_status harm_host_init_torus(const harm_config* cfg, double rin, double rmax, double kappa,
                                            double beta, double randfact, uint32_t seed, double* p_aos)
{
  if (!cfg || !p_aos) return hh::set_error(HARM_EINVAL, "harm_host_init_torus: null argument");
  Dims d;
  Params P;
  make_dims(cfg, d, P);
  hi_grid G;
  G.ng1 = d.ng1; G.ng2 = d.ng2; G.ng3 = d.ng3;
  for (int a = 0; a < 4; a++) { G.startx[a] = P.startx[a]; G.dx[a] = P.dx[a]; }
  memset(p_aos, 0, sizeof(double) * HARM_NPR * d.ns1 * d.ns2 * d.ns3);
  if (hi_init_torus(cfg, &G, rin, rmax, kappa, beta, randfact, seed, p_aos) != 0)
    return hh::set_error(HARM_EINVAL, "harm_host_init_torus: no torus zones on this grid");
  return HARM_OK;
}

extern "C" harm_status harm_host_init_bondi(const harm_config* cfg, double* p_aos)
{
  if (!cfg || !p_aos) return hh::set_error(HARM_EINVAL, "harm_host_init_bondi: null argument");
  Dims d;
  Params P;
  make_dims(cfg, d, P);
  hi_grid G;
  G.ng1 = d.ng1; G.ng2 = d.ng2; G.ng3 = d.ng3;
  for (int a = 0; a < 4; a++) { G.startx[a] = P.startx[a]; G.dx[a] = P.dx[a]; }
  memset(p_aos, 0, sizeof(double) * HARM_NPR * d.ns1 * d.ns2 * d.ns3);
  if (hi_init_bondi(cfg, &G, p_aos) != 0)
    return hh::set_error(HARM_EINVAL, "harm_host_init_bondi: no Bondi flow through some zone of this grid");
  return HARM_OK;
}

extern "C" harm_status harm_init_bondi(harm_handle* h)
{
  HH_TRY(check_handle(h));
  if (h->cfg.problem != HARM_PROBLEM_BONDI)
    return hh::set_error(HARM_EINVAL, "harm_init_bondi: the handle was not created for the Bondi problem");
  const Dims& d = h->d;
  std::vector<double> p((size_t)HARM_NPR * d.ns1 * d.ns2 * d.ns3, 0.0);
  HH_TRY(harm_host_init_bondi(&h->cfg, p.data()));
  HH_TRY(reset_pflag(h));  // a new run's flags
  HH_TRY(harm_set_array(h, HARM_ARR_P, p.data()));
  HH_TRY(enq_bound(h, HARM_ARR_P));
  return harm_sync(h);
}

extern "C" harm_status harm_bondi_state(double gam, double r, double out[3])
{
  if (!out) return hh::set_error(HARM_EINVAL, "harm_bondi_state: null argument");
  if (hi_bondi_state(gam, r, &out[0], &out[1], &out[2]) != 0)
    return hh::set_error(HARM_EINVAL, "harm_bondi_state: no flow through r = %g", r);
  return HARM_OK;
}

extern "C" harm_status harm_profile(harm_handle* h, int32_t enable)
{
  HH_TRY(check_handle(h));
  h->prof = enable != 0;
  return HARM_OK;
}

extern "C" harm_status harm_profile_read(harm_handle* h, double* ms_total, int64_t* count)
{
  HH_TRY(check_handle(h));
  HH_HIP(hipStreamSynchronize(h->stream));
  for (auto& r : h->evrec) {
    float ms = 0.f;
    HH_HIP(hipEventElapsedTime(&ms, r.second.first, r.second.second));
    h->prof_ms[r.first] += ms;
    h->prof_n[r.first] += 1;
    h->evpool.push_back(r.second.first);
    h->evpool.push_back(r.second.second);
  }
  h->evrec.clear();
  for (int k = 0; k < HARM_K_N; k++) {
    if (ms_total) ms_total[k] = h->prof_ms[k];
    if (count) count[k] = h->prof_n[k];
    h->prof_ms[k] = 0.0;
    h->prof_n[k] = 0;
  }
  return HARM_OK;
}

// ---------------------------------------------------------------------------
// Tile autotuning and batches
// ---------------------------------------------------------------------------
extern "C" harm_status harm_tune(harm_handle* h, int32_t reps, int32_t td_out[4])
{
  HH_TRY(check_handle(h));
  if (reps < 1) reps = 1;
  const Dims& d = h->d;
  hipEvent_t e0, e1;
  HH_HIP(hipEventCreate(&e0));
  HH_HIP(hipEventCreate(&e1));
  const int recon = h->cfg.recon, ft = h->cfg.flux;
  hipStream_t s = h->stream;
  const double* pb = h->arr[HARM_ARR_P];
  harm_status st = HARM_OK;
  for (int dir = 2; dir <= 3 && st == HARM_OK; dir++) {
    if ((dir == 2 && d.n2 < 2) || (dir == 3 && d.n3 < 2)) continue;
    double best = 1e300;
    int best_td = 0;
    for (int td : {4, 8, 16}) {
      if ((td + 1) * (kBlock / td) > kCellMax) continue;
      const FaceTile T = plan_face_tile(d, dir, td);
      double* F = h->arr[HARM_ARR_F1 + dir - 1];
      const double* G = h->geom[dir];
      if (recon == HARM_RECON_WENO5) {  // the face kernel reads this direction's line reconstruction
        launch_weno_dp(d, s, pb, h->geom[HARM_LOC_CENT], h->P, h->wtm, h->wdp);
        launch_weno_recon(dir, d, T.b, s, pb, h->wdp, wl_dir(h, dir), wr_dir(h, dir));
      }
      auto launch = [&]() {
        launch_flux(dir, recon, T, s, d, pb, F, G, h->P, ft, 1, h->ds, h->wl ? wl_dir(h, dir) : nullptr,
                    h->wr ? wr_dir(h, dir) : nullptr);
      };
      launch();  // warm
      (void)hipEventRecord(e0, s);
      for (int r = 0; r < reps; r++) launch();
      (void)hipEventRecord(e1, s);
      if ((st = hh::check_launch("harm_tune")) != HARM_OK) break;
      if ((st = hh::check_hip(hipEventSynchronize(e1), "harm_tune")) != HARM_OK) break;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) {
        best = ms;
        best_td = td;
      }
    }
    if (st == HARM_OK && best_td) {
      h->tile_td[dir] = best_td;
      h->ft[dir] = plan_face_tile(d, dir, best_td);
      drop_step_graph(h);  // the captured step holds the old tile shapes
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (td_out)
    for (int dir = 0; dir < 4; dir++) td_out[dir] = dir >= 2 ? h->ft[dir].td : 0;
  return st;
}

extern "C" harm_status harm_batch_destroy(harm_batch* b)
{
  if (!b) return HARM_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (harm_handle* h : b->grids) {
    if (!h) continue;
    h->stream = nullptr;  // the batch's stream, destroyed below
    harm_destroy(h);
  }
  if (b->dref) (void)hipFree(b->dref);
  if (b->dpre) (void)hipFree(b->dpre);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return HARM_OK;
}

extern "C" harm_status harm_batch_create(const harm_config* cfgs, int32_t n, harm_batch** out)
{
  if (!cfgs || !out || n < 1) return hh::set_error(HARM_EINVAL, "harm_batch_create: bad arguments");
  const harm_config& c0 = cfgs[0];
  const int dm0 = (c0.n1 > 1) | ((c0.n2 > 1) << 1) | ((c0.n3 > 1) << 2);
  for (int i = 0; i < n; i++) {
    const harm_config& c = cfgs[i];
    const int dm = (c.n1 > 1) | ((c.n2 > 1) << 1) | ((c.n3 > 1) << 2);
    if (c.recon != c0.recon || c.flux != c0.flux || c.timeorder != c0.timeorder || dm != dm0)
      return hh::set_error(HARM_EINVAL,
                           "harm_batch_create: grid %d differs in reconstruction, flux, time order or dimensionality", i);
  }
  harm_batch* b = new harm_batch();
  b->recon = c0.recon;
  b->timeorder = c0.timeorder;
  b->dm = dm0;
  harm_status st = hh::check_hip(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
  for (int i = 0; i < n && st == HARM_OK; i++) {
    harm_handle* h = nullptr;
    st = harm_create(&cfgs[i], &h);
    if (st != HARM_OK) break;
    (void)hipStreamDestroy(h->stream);
    h->stream = b->stream;
    b->grids.push_back(h);
  }
  if (st == HARM_OK) st = batch_build(b);
  if (st != HARM_OK) {
    harm_batch_destroy(b);
    return st;
  }
  *out = b;
  return HARM_OK;
}

extern "C" harm_status harm_batch_grid(harm_batch* b, int32_t i, harm_handle** h)
{
  if (!b || !h || i < 0 || i >= (int32_t)b->grids.size()) return hh::set_error(HARM_EINVAL, "harm_batch_grid: bad index");
  *h = b->grids[i];
  return HARM_OK;
}

extern "C" harm_status harm_batch_tune(harm_batch* b, int32_t reps)
{
  if (!b) return hh::set_error(HARM_EINVAL, "harm_batch_tune: null batch");
  for (harm_handle* h : b->grids) HH_TRY(harm_tune(h, reps, nullptr));
  HH_HIP(hipStreamSynchronize(b->stream));
  if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
  return batch_build(b);
}

extern "C" harm_status harm_batch_step(harm_batch* b, int32_t nsteps, int32_t use_graph)
{
  if (!b || nsteps < 0) return hh::set_error(HARM_EINVAL, "harm_batch_step: bad arguments");
  if (!use_graph) {
    for (int n = 0; n < nsteps; n++) HH_TRY(enq_batch_step(b));
    return hh::check_launch("harm_batch_step");
  }
  if (!b->exec) {
    HH_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const harm_status st = enq_batch_step(b);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st != HARM_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    HH_HIP(e);
    b->graph = g;
    HH_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  for (int n = 0; n < nsteps; n++) HH_HIP(hipGraphLaunch(b->exec, b->stream));
  return HARM_OK;
}

extern "C" harm_status harm_batch_sync(harm_batch* b)
{
  if (!b) return hh::set_error(HARM_EINVAL, "harm_batch_sync: null batch");
  return hh::check_hip(hipStreamSynchronize(b->stream), "harm_batch_sync");
}
