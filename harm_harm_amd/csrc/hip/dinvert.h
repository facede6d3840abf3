// dinvert.h -- one zone of the step's inversion: Jon's Newton solve
// (Utoprimgen, utoprimgen.c:32) and, on the final substep, fixup1zone's
// floors (fixup.c:561), as the inversion kernels of the step run it (the
// plain one in grid.hip, the learned-guess one in mlp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "dgrid.h"
#include "gkern.h"

namespace hk {

using hd::Dims;
using hd::Geom;
using hd::State;

// ---------------------------------------------------------------------------
// inversion + fixup1zone (floors via ZAMO-frame mass/energy addition)
// ---------------------------------------------------------------------------
// One zone: conserved Uz (with gdet) in registers, the guess read from pg at
// offset o, the result written to pf and pflag.  Returns the flag; the Newton
// stroke count is added to ns.
//
// SEEDED: the Newton solve starts from W' = wp_seed (the learned guess); when
// that solve fails, the zone is solved again from the primitives' guess as
// the plain inversion does (Newton from the previous primitives is the
// fallback and decides), and *fellback is set.
// DIAG: on the final substep, the conserved change of each fix is added as
// diag_fixup does (fixup.c:220, finalstep > 0): dV (U(after) - U(before)),
// with U = primtoU with gdet and `before` the primitives ahead of the floors,
// for the floor step (fixup.c:766) to dfl[0..7] and for the gamma limit
// (fixup.c:784) to dfl[8..15]; changed[0], changed[1] say whether each was.
// dfl[pl] += dV (U(pa)[pl] - U(pb)[pl]), U with gdet (diag_fixup, fixup.c:303-405)
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

template <bool SEEDED = false, bool DIAG = false>
__device__ __forceinline__ int invert_zone(const Dims& d, const Params& P, double* Uz,
                                           const double* __restrict__ pg, double* __restrict__ pf,
                                           int32_t* __restrict__ pflag, const double* __restrict__ gcent,
                                           const double* __restrict__ floors, int finalstep, long o, long pp, int& ns,
                                           double wp_seed = 0.0, int* fellback = nullptr, double* dfl = nullptr,
                                           bool* changed = nullptr)
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
  if (finalstep) {
    double p0[HARM_NPR];
    if (DIAG) {
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) p0[pl] = pz[pl];
    }
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
      fl = hd::jon(Ua, G, P.gam, P.gammafail, pz, ns, __builtin_nan(""));
      if (fl > 0) {
#pragma unroll
        for (int pl = 0; pl < HARM_NPR; pl++) pz[pl] = prnew[pl];
      }
      if (DIAG) { fix_delta_to(P, G, p0, pz, dfl); changed[0] = true; }
    }
    const bool lg = hd::limit_gamma(P.gammamax, pz, G);
    if (DIAG && lg) { fix_delta_to(P, G, p0, pz, dfl + HARM_NPR); changed[1] = true; }
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
