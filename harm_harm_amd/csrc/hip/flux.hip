// flux.hip -- the face-flux kernels of the step (reconstruction and Riemann
// solve, flux.c:852 fluxcalc_standard), one direction per launch, for one
// grid or for a batch.  Like the rest of the step they keep the reference's
// operation order without contraction and divide where it divides, so face
// fluxes equal the CPU restatement's bit for bit and the per-zone parity of
// a step does not depend on where cancellation amplifies rounding.  (Fused
// multiply-adds and reciprocals in place of divisions by positive
// quantities, kFast below, were measured 20 % faster per flux sweep but
// differ from the IEEE sequence by up to a few 1e-12 in zones whose flux
// cancels.)
#include <hip/hip_runtime.h>
#include <math.h>
#include "dgrid.h"
#include "gkern.h"
#include "hhost.h"

namespace hk {
namespace {

using hd::Geom;
using hd::State;

constexpr bool kFast = false;

// Riemann flux of one face from its two reconstructed states
// (getplpr/check_plpr flux.c:1952-2110, flux_compute fluxcompute.c:19-402,
// cminmax_calc wavespeeds.c:397).  Returns ctop.
template <int DIR>
__device__ __forceinline__ double face_riemann(double* p_l, double* p_r, const Geom& G, const Params& P, int fluxtype,
                                               int do_ct, double* Fo)
{
  const int pbd = HARM_B1 + DIR - 1;
  p_l[pbd] = p_r[pbd] = 0.5 * (p_l[pbd] + p_r[pbd]);
  hd::limit_gamma(P.gammamax, p_l, G);
  hd::limit_gamma(P.gammamax, p_r, G);
  double F_l[HARM_NPR], U_l[HARM_NPR], F_r[HARM_NPR], U_r[HARM_NPR];
  double cmaxl, cminl, cmaxr, cminr;
  {
    State q;
    hd::get_state<kFast>(p_l, G, P.gam, q);
    hd::primtoflux<kFast>(p_l, q, DIR, G, F_l);
    hd::primtoflux<kFast>(p_l, q, 0, G, U_l);
    hd::vchar<kFast>(p_l, q, DIR, G, P.gam, cmaxl, cminl);
  }
  {
    State q;
    hd::get_state<kFast>(p_r, G, P.gam, q);
    hd::primtoflux<kFast>(p_r, q, DIR, G, F_r);
    hd::primtoflux<kFast>(p_r, q, 0, G, U_r);
    hd::vchar<kFast>(p_r, q, DIR, G, P.gam, cmaxr, cminr);
  }
  double lmin = fmin(cminl, cminr), lmax = fmax(cmaxl, cmaxr);
  lmin = fabs(fmax(0., -lmin));
  lmax = fabs(fmax(0., lmax));
  const double cmin = lmin, cmax = lmax, ct = fmax(lmax, lmin);
  if (fluxtype == HARM_FLUX_HLL && cmax + cmin != 0.0) {
    if (kFast) {
      const double ic = 1.0 / (cmax + cmin);
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++)
        Fo[pl] = ((cmax * F_l[pl] + cmin * F_r[pl]) - (cmax * cmin) * (U_r[pl] - U_l[pl])) * ic;
    } else {
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++)
        Fo[pl] = ((cmax * F_l[pl] + cmin * F_r[pl]) - (cmax * cmin) * (U_r[pl] - U_l[pl])) / (cmax + cmin);
    }
  } else {
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) Fo[pl] = (0.5 * ((F_l[pl] + F_r[pl]) - ct * (U_r[pl] - U_l[pl])));
  }
  if (do_ct) {
    // igdetnosing with the reciprocal: |gdet| + SMALL is positive and normal
    const double ign = kFast ? (G.gdet >= 0.0 ? 1.0 : -1.0) * hd::recip_nr(fabs(G.gdet) + D_SMALL)
                             : hd::igdetnosing(G.gdet);
    Fo[HARM_B1] *= ign;
    Fo[HARM_B2] *= ign;
    Fo[HARM_B3] *= ign;
  }
  return ct;
}

// Three waves per SIMD: the LDS tile allows three blocks per CU.
template <int DIR, int RECON>
__device__ __forceinline__ void flux_body(Dims d, FaceTile T, const double* __restrict__ pb,
                                                 double* __restrict__ F, const double* __restrict__ gface, Params P,
                                                 int fluxtype, int do_ct, double* __restrict__ ds, const int blk,
                                                 const double* __restrict__ wl, const double* __restrict__ wr)
{
  __shared__ double s_l[RECON == HARM_RECON_WENO5 ? 1 : HARM_NPR * kCellMax],
      s_r[RECON == HARM_RECON_WENO5 ? 1 : HARM_NPR * kCellMax];
  const int tile = blk;
  const int st = (DIR == 1) ? 1 : (DIR == 2) ? d.ps1 : d.ps1 * d.ns2;
  const Box& b = T.b;

  // ---- phase 1: reconstruct the tile's cells ----
  // (WENO5BND: the line reconstruction already wrote every cell's face
  // values, and phase 2 reads them straight from wl, wr)
  // (32-bit index arithmetic: a face box stays far below 2^31 faces)
  int f0 = 0, g_lo = 0, ncell, bc = 0, bd = 0, pl = 0;
  const int rowc = T.nd + 1;  // dir 1: cells per row
  if (DIR == 1) {
    const int nf = T.nd * T.ncol;
    f0 = tile * T.fpb;
    const int fl = min(f0 + T.fpb, nf) - 1;
    g_lo = (f0 / T.nd) * rowc + f0 % T.nd;
    ncell = (fl / T.nd) * rowc + fl % T.nd + 1 - g_lo + 1;
  } else {
    // bc: i tile, bd: sweep tile, pl: transverse plane
    if (DIR == 2) {
      pl = tile % T.nplane;
      const int r = tile / T.nplane;
      bd = r % T.ntd;
      bc = r / T.ntd;
    } else {
      bd = tile % T.ntd;
      const int r = tile / T.ntd;
      pl = r % T.nplane;
      bc = r / T.nplane;
    }
    ncell = (T.td + 1) * T.tc;
  }
  // storage offset of every cell of the tile (-1: outside the face box),
  // decoded once per cell so the loop over (variable, cell) items below needs
  // no integer division
  __shared__ int s_off[RECON == HARM_RECON_WENO5 ? 1 : kCellMax];
  if constexpr (RECON != HARM_RECON_WENO5) {
  for (int slot = threadIdx.x; slot < ncell; slot += kBlock) {
    int i, j, k, off = -1;
    if (DIR == 1) {
      const int g = g_lo + slot;
      const int row = g / rowc;
      i = b.is - 1 + (g - row * rowc);
      const int rj = row / b.nk;
      k = b.ks + (row - rj * b.nk);
      j = b.js + rj;
      off = (int)hd::cidx(d, i, j, k);
    } else {
      const int dd = slot / T.tc, cc = slot - dd * T.tc;
      i = T.i0 + bc * T.tc + cc;
      const int pd = bd * T.td - 1 + dd;  // sweep index of the cell relative to the first face
      if (i >= b.is && i < b.is + b.ni && pd <= T.nd - 1) {
        if (DIR == 2) { k = b.ks + pl; j = b.js + pd; }
        else { j = b.js + pl; k = b.ks + pd; }
        off = (int)hd::cidx(d, i, j, k);
      }
    }
    s_off[slot] = off;
  }
  __syncthreads();
  // item it = v * ncell + slot, advanced by kBlock per pass
  int v = 0, slot = threadIdx.x;
  while (slot >= ncell) { slot -= ncell; v++; }
  // One item per pass: every cell is reconstructed by the same instruction
  // sequence whatever the tile shape, so tile shapes never change results.
  // (Two items per pass in one basic block measured 1.5 % faster, but the
  // compiler may contract the two copies differently; likewise no unrolling.)
#pragma clang loop unroll(disable)
  while (v < HARM_NPR) {
    const int off = s_off[slot];
    if (off >= 0) {
      double l, r;
      const double* a = pb + (long)v * d.ncell + off;
      double y[5];
#pragma unroll
      for (int m = 0; m < 5; m++) y[m] = a[(m - 2) * st];
      hd::recon_cell<RECON>(y + 2, l, r);
      s_l[v * kCellMax + slot] = l;
      s_r[v * kCellMax + slot] = r;
    }
    slot += kBlock;
    if (slot >= ncell) {  // ncell >= kBlock except in small tiles: one subtraction
      slot -= ncell;
      v++;
      while (slot >= ncell) { slot -= ncell; v++; }
    }
  }
  __syncthreads();
  }

  // ---- phase 2: one face per lane ----
  double dtij = kBIG;
  int i, j, k, sl = 0, sr = 0;
  bool ok;
  if (DIR == 1) {
    const int f = f0 + (int)threadIdx.x;
    ok = (int)threadIdx.x < T.fpb && f < T.nd * T.ncol;
    if (ok) {
      const int row = f / T.nd;
      const int fi = f - row * T.nd;
      sr = row * rowc + fi + 1 - g_lo;
      sl = sr - 1;
      const int rj = row / b.nk;
      i = b.is + fi;
      k = b.ks + (row - rj * b.nk);
      j = b.js + rj;
    }
  } else {
    const int dd = (int)threadIdx.x / T.tc, cc = (int)threadIdx.x - dd * T.tc;
    const int fd = bd * T.td + dd;
    i = T.i0 + bc * T.tc + cc;
    ok = dd < T.td && i >= b.is && i < b.is + b.ni && fd < T.nd;
    if (ok) {
      sl = dd * T.tc + cc;
      sr = sl + T.tc;
      if (DIR == 2) { k = b.ks + pl; j = b.js + fd; }
      else { j = b.js + pl; k = b.ks + fd; }
    }
  }
  if (ok) {
    double p_l[HARM_NPR], p_r[HARM_NPR];
    const long c = hd::cidx(d, i, j, k);
    if constexpr (RECON == HARM_RECON_WENO5) {
      // the right face value of the cell behind the face, the left one of the
      // cell ahead (weno.hip)
#pragma unroll
      for (int v = 0; v < HARM_NPR; v++) {
        p_l[v] = wr[(long)v * d.ncell + c - st];
        p_r[v] = wl[(long)v * d.ncell + c];
      }
    } else {
#pragma unroll
      for (int v = 0; v < HARM_NPR; v++) {
        p_l[v] = s_r[v * kCellMax + sl];
        p_r[v] = s_l[v * kCellMax + sr];
      }
    }
    Geom G;
    hd::load_geom(gface, hd::pidx(d, i, j), d.nplane, G);
    double Fo[HARM_NPR];
    const double ct = face_riemann<DIR>(p_l, p_r, G, P, fluxtype, do_ct, Fo);
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) F[(long)pl * d.ncell + c] = Fo[pl];
    dtij = P.cour * P.dx[DIR] / ct;
  }
  block_min_to(dtij, ds + DS_NDT1 + DIR - 1);
}

template <int DIR, int RECON>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_flux(Dims d, FaceTile T, const double* __restrict__ pb,
                                                 double* __restrict__ F, const double* __restrict__ gface, Params P,
                                                 int fluxtype, int do_ct, double* __restrict__ ds,
                                                 const double* __restrict__ wl, const double* __restrict__ wr)
{
  flux_body<DIR, RECON>(d, T, pb, F, gface, P, fluxtype, do_ct, ds, xcd_tile(blockIdx.x, gridDim.x), wl, wr);
}

template <int DIR, int RECON>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) void kb_flux(
    const GridRef* __restrict__ G, const int* __restrict__ pre, int ng, int which, int do_ct)
{
  int gi, lb;
  {
    const int bb = xcd_tile(blockIdx.x, gridDim.x);
    int lo = 0, hi = ng - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pre[mid] <= bb) lo = mid;
      else hi = mid - 1;
    }
    gi = __builtin_amdgcn_readfirstlane(lo);  // block-uniform: scalar registers
    lb = __builtin_amdgcn_readfirstlane(bb - pre[lo]);
  }
  const GridRef& g = G[gi];
  flux_body<DIR, RECON>(g.d, g.ft[DIR], gmem(g.arr[which]), gmem(g.arr[HARM_ARR_F1 + DIR - 1]), gmem(g.geom[DIR]), g.P, g.fluxtype,
                        do_ct, gmem(g.ds), lb,
                        RECON == HARM_RECON_WENO5 ? gmem(g.wl) + (long)(DIR - 1) * HARM_NPR * g.d.ncell : nullptr,
                        RECON == HARM_RECON_WENO5 ? gmem(g.wr) + (long)(DIR - 1) * HARM_NPR * g.d.ncell : nullptr);
}

}  // namespace

void launch_flux(int dir, int recon, const FaceTile& T, hipStream_t s, const Dims& d, const double* pb, double* F,
                 const double* gface, const Params& P, int fluxtype, int do_ct, double* ds, const double* wl,
                 const double* wr)
{
  const dim3 grid((unsigned)T.nblk), blk(kBlock);
#define LF(DIR)                                                                                                   \
  if (recon == HARM_RECON_WENO5)                                                                                  \
    hipLaunchKernelGGL((k_flux<DIR, HARM_RECON_WENO5>), grid, blk, 0, s, d, T, pb, F, gface, P, fluxtype, do_ct, ds, \
                       wl, wr);                                                                                   \
  else if (recon == HARM_RECON_PARA)                                                                              \
    hipLaunchKernelGGL((k_flux<DIR, HARM_RECON_PARA>), grid, blk, 0, s, d, T, pb, F, gface, P, fluxtype, do_ct, ds, \
                       nullptr, nullptr);                                                                         \
  else                                                                                                            \
    hipLaunchKernelGGL((k_flux<DIR, HARM_RECON_DONOR>), grid, blk, 0, s, d, T, pb, F, gface, P, fluxtype, do_ct, ds, \
                       nullptr, nullptr);
  if (dir == 1) { LF(1) } else if (dir == 2) { LF(2) } else { LF(3) }
#undef LF
}

void launch_flux_batch(int dir, int recon, unsigned nblk, hipStream_t s, const GridRef* G, const int* pre, int ng,
                       int which, int do_ct)
{
  const dim3 grid(nblk), blk(kBlock);
#define LB(DIR)                                                                                                   \
  if (recon == HARM_RECON_WENO5)                                                                                  \
    hipLaunchKernelGGL((kb_flux<DIR, HARM_RECON_WENO5>), grid, blk, 0, s, G, pre, ng, which, do_ct);               \
  else if (recon == HARM_RECON_PARA)                                                                              \
    hipLaunchKernelGGL((kb_flux<DIR, HARM_RECON_PARA>), grid, blk, 0, s, G, pre, ng, which, do_ct);                \
  else                                                                                                            \
    hipLaunchKernelGGL((kb_flux<DIR, HARM_RECON_DONOR>), grid, blk, 0, s, G, pre, ng, which, do_ct);
  if (dir == 1) { LB(1) } else if (dir == 2) { LB(2) } else { LB(3) }
#undef LB
}

}  // namespace hk
