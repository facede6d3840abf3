// dinvert.h -- one zone of the step's inversion: Utoprimgen (utoprimgen.c:32:
// Jon's hot Newton solve, the inversion check and the cold retry as the
// configuration turns them on) and, on the final substep, fixup1zone's floors
// (fixup.c:561), as the inversion kernels of the step run it (the plain one in
// grid.hip, the learned-guess one in mlp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "dgrid.h"
#include "gkern.h"

namespace hk {

using hd::Dims;
using hd::Geom;
using hd::State;

// dfl[pl] += dV (U(pa)[pl] - U(pb)[pl]), U with gdet: the conserved change of
// one fix as the energy file accounts it (diag_fixup, fixup.c:303-405)
__device__ __forceinline__ void fix_delta_to(const Params& P, const Geom& G, const double* pb, const double* pa,
                                             double* dfl)
{
  const double dV = P.dx[1] * P.dx[2] * P.dx[3];
  State q;
  double Ub[HARM_NPR], Ua[HARM_NPR];
  hd::get_state(pb, G, P.gam, q);
  hd::primtoflux(pb, q, 0, G, Ub);
  hd::get_state(pa, G, P.gam, q);
  hd::primtoflux(pa, q, 0, G, Ua);
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) dfl[pl] += dV * (Ua[pl] - Ub[pl]);
}

// ---------------------------------------------------------------------------
// inversion + fixup1zone (floors via ZAMO-frame mass/energy addition)
// ---------------------------------------------------------------------------
// fixup1zone on the final substep (fixup.c:561-830): the density and
// internal-energy floors (mass and energy added in the ZAMO frame, then the
// floors' own Utoprimgen from the primitives so far) and the Lorentz-factor
// limit, on pz with flag fl (replaced by the floors' inversion flag when the
// floors changed the zone).  G: the zone's centre geometry.  Returns the
// flag; the floors' Newton strokes are added to ns.  *changed reports
// whether pz changed.  DIAG: as invert_zone.
template <bool DIAG = false>
__device__ __forceinline__ int floors_zone(const Dims& d, const Params& P, const Geom& G,
                                           const double* __restrict__ floors, double* pz, int fl, long o, long pp,
                                           int& ns, bool* changed, double* __restrict__ fx_p0,
                                           double* __restrict__ fx_pfl, uint8_t* __restrict__ fx_flag)
{
  const long nc = d.ncell;
  double prfloor[2] = {floors[pp], floors[d.nplane + pp]};
  for (int pl = 0; pl < 2; pl++)
    if (prfloor[pl] < D_SMALL) prfloor[pl] = D_SMALL;
  double prnew[HARM_NPR];
  bool didchange = false;
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) {
    if (pl <= HARM_UU && prfloor[pl] > pz[pl]) { didchange = true; prnew[pl] = prfloor[pl]; }
    else prnew[pl] = pz[pl];
  }
  if (didchange) {
    if (DIAG) {
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) fx_p0[pl * nc + o] = pz[pl];
    }
    Geom G1 = G;
    G1.gdet = 1.0;
    State q, dq;
    double Ua[HARM_NPR], dpr[HARM_NPR], dUa[HARM_NPR];
    hd::get_state(pz, G, P.gam, q);
    hd::primtoflux(pz, q, 0, G1, Ua);
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) dpr[pl] = 0.0;
    for (int pl = HARM_RHO; pl <= HARM_UU; pl++) dpr[pl] = prnew[pl] - pz[pl];
    hd::get_state(dpr, G, P.gam, dq);
    hd::primtoflux(dpr, dq, 0, G1, dUa);
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) Ua[pl] += dUa[pl];
    // the floors' own Utoprimgen (UNOTHING), from the primitives so far;
    // any flag > 0, the cold retry's mark included, restores prnew
    double p0[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) p0[pl] = pz[pl];
    fl = hd::jon(Ua, G, P.gam, P.gammafail, pz, ns, __builtin_nan(""));
    if (hd::needs_after_hot(fl, P.hot2cold, P.checkinv))
      fl = hd::after_hot(fl, Ua, G, P.gam, P.gammafail, p0, pz, ns, P.hot2cold, P.checkinv);
    if (fl > 0) {
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) pz[pl] = prnew[pl];
    }
  }
  double pb4[HARM_NPR];  // DIAG: the primitives ahead of the gamma limit
  if (DIAG) {
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) pb4[pl] = pz[pl];
  }
  const bool lg = hd::limit_gamma(P.gammamax, pz, G);
  if (DIAG) {
    if (lg) {
      double* dst = didchange ? fx_pfl : fx_p0;  // with no floor step these are also the primitives ahead of both
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) dst[pl * nc + o] = pb4[pl];
    }
    fx_flag[o] = (uint8_t)((didchange ? 1 : 0) | (lg ? 2 : 0));
  }
  *changed = didchange || lg;
  return fl;
}

// One zone: conserved Uz (with gdet) in registers, the guess read from pg at
// offset o, the result written to pf and pflag.  Returns the flag; the Newton
// stroke count is added to ns.
//
// SEEDED: the Newton solve starts from W' = wp_seed (the learned guess); when
// that solve fails, the zone is solved again from the primitives' guess as
// the plain inversion does (Newton from the previous primitives is the
// fallback and decides), and *fellback is set.
// DIAG (final substep): the states each fix started from are kept for the
// energy-file accounting (diag_fixup, fixup.c:220): fx_flag[o] gets bit 0
// when the floors changed the zone (fixup.c:766) and bit 1 when the gamma
// limit did (fixup.c:784); fx_p0 (SoA, stride ncell) then holds the
// primitives ahead of both fixes and, when the floors changed them, fx_pfl
// the primitives after the floors.  Only fixed zones store their states.
//
// FLOORS = false: the floors and the Lorentz-factor limit are left to a
// separate pass over the zones (floors_zone), as the plain grid inversion
// does on the final substep so that this solve keeps its registers.
//
// Ug: the conserved state in global memory (Uz's source, SoA at o); the rare
// checks and retries after the hot solve read it again rather than keeping
// Uz in registers through the Newton iterations.
template <bool SEEDED = false, bool DIAG = false, bool FLOORS = true>
__device__ __forceinline__ int invert_zone(const Dims& d, const Params& P, double* Uz, const double* __restrict__ Ug,
                                           const double* __restrict__ pg, double* __restrict__ pf,
                                           int32_t* __restrict__ pflag, const double* __restrict__ gcent,
                                           const double* __restrict__ floors, int finalstep, long o, long pp, int& ns,
                                           double wp_seed = 0.0, int* fellback = nullptr,
                                           double* __restrict__ fx_p0 = nullptr, double* __restrict__ fx_pfl = nullptr,
                                           uint8_t* __restrict__ fx_flag = nullptr)
{
  const long nc = d.ncell;
  Geom G;
  hd::load_geom(gcent, pp, d.nplane, G);
  double pz[HARM_NPR];
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) pz[pl] = pg[pl * nc + o];  // guess: the substep's starting primitives
  int fl = hd::utoprimgen(Uz, G, P.gam, P.gammafail, pz, ns, SEEDED ? wp_seed : __builtin_nan(""));
  if (SEEDED && fl != 0) {
    *fellback = 1;
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) pz[pl] = pg[pl * nc + o];
    fl = hd::utoprimgen(Uz, G, P.gam, P.gammafail, pz, ns, __builtin_nan(""));
  }
  // The geometry is needed again only for the floors below.  Reloading it
  // through a pointer the compiler cannot see through frees its 54
  // registers during the Newton iterations (they come back from L1/L2).
  hd::load_geom(hd::fresh(gcent), pp, d.nplane, G);
  if (hd::needs_after_hot(fl, P.hot2cold, P.checkinv)) {
    double Ur[HARM_NPR], Ugf[HARM_NPR], p0[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) {
      Ur[pl] = Ug[pl * nc + o];
      p0[pl] = pg[pl * nc + o];
    }
    hd::geometry_free(Ur, G, Ugf);
    fl = hd::after_hot(fl, Ugf, G, P.gam, P.gammafail, p0, pz, ns, P.hot2cold, P.checkinv);
  }
  if (FLOORS && finalstep) {
    bool changed;
    fl = floors_zone<DIAG>(d, P, G, floors, pz, fl, o, pp, ns, &changed, fx_p0, fx_pfl, fx_flag);
  } else if (FLOORS && DIAG) {
    fx_flag[o] = 0;
  }
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) pf[pl * nc + o] = pz[pl];
  pflag[o] = fl;
  return fl;
}

// Block sums of the Newton strokes and failures into di (integer sums are
// order independent).  Every thread of the block calls it.
__device__ __forceinline__ void newton_counts_to(int ns, int nf, unsigned long long* __restrict__ di)
{
  __shared__ unsigned long long red[2];
  if (threadIdx.x == 0) { red[0] = 0; red[1] = 0; }
  __syncthreads();
  if (ns) atomicAdd(&red[0], (unsigned long long)ns);
  if (nf) atomicAdd(&red[1], 1ull);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (red[0]) atomicAdd(&di[DI_NSTROKE], red[0]);
    if (red[1]) atomicAdd(&di[DI_NFAIL], red[1]);
  }
}

}  // namespace hk
