// ener.hip -- the energy-file diagnostics of one grid on the device
// (dump_ener.c:5-420, diag.c:1020-1250, 1958-2000, fixup.c:220-420):
//
//   k_ener_pdot      boundary flux sums of the substep (diag_flux_pureflux)
//                    over the six boundaries and eight variables, and their
//                    time integral pcum
//   k_ener_dt4       the weight of the sampled source (dt4diag,
//                    advance.c:251-262)
//   k_ener_source    the sampled geometric source (diag_source_all)
//   k_ener_totals    conserved totals sum U dV (constotal, dump_ener.c:882)
//   k_ener_divb      max and sum of |div B| at the corners (divbmaxavg,
//                    dump_ener.c:680, SETFDIVBFLUXCTTOTHPRIM)
//   k_ener_reduce    fixed-order sums of per-block partials, one block per
//                    channel
//   k_dump_cols      the physics columns of HARM's main dump (dump.c:439)
//   k_current_*      the dump's four-current (phys.c:552-1000)
//
//   k_ener_fix       the final substep's floor and gamma-limit changes from
//                    the states the inversion kept (diag_fixup, fixup.c:220)
//
// The failure-averaging changes are summed by the failure-averaging kernel
// itself (its DIAG variant) into per-block partials.  Every sum is reduced
// in a fixed order, so the same state gives the same bits.
#include <hip/hip_runtime.h>
#include "gkern.h"
#include "dinvert.h"

namespace hk {

using hd::Geom;
using hd::State;

namespace {

inline unsigned nblk(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// One block per (boundary, chunk): boundary bd = X1UP, X1DN, X2UP, X2DN,
// X3UP, X3DN (the reference's order, global.nondepnmemonics.h:699); each face adds F[pl] * surface (non-finite terms count as 0, as
// with PRODUCTION > 0 at diag.c:1196-1201).  A direction with one zone has no
// faces to sum (its fluxes are not computed).
__global__ __launch_bounds__(kBlock) void k_ener_pdot(Dims d, Params P, const double* __restrict__ F1,
                                                      const double* __restrict__ F2,
                                                      const double* __restrict__ F3, double* __restrict__ part)
{
  const int bd = blockIdx.x / kEnFaceBlocks, chunk = blockIdx.x % kEnFaceBlocks;
  const int dir = bd / 2 + 1, up = !(bd & 1);
  double v[HARM_NPR] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int nd = dir == 1 ? d.n1 : dir == 2 ? d.n2 : d.n3;
  if (nd > 1) {
    const double* F = dir == 1 ? F1 : dir == 2 ? F2 : F3;
    const double surface = dir == 1 ? P.dx[2] * P.dx[3] : dir == 2 ? P.dx[1] * P.dx[3] : P.dx[1] * P.dx[2];
    // the face plane: (a, b) = (j, k), (i, k) or (i, j)
    const int na = dir == 1 ? d.n2 : d.n1, nb = dir == 3 ? d.n2 : d.n3;
    const long nf = (long)na * nb, per = (nf + kEnFaceBlocks - 1) / kEnFaceBlocks;
    const long f0 = chunk * per, f1 = min(nf, f0 + per);
    const long nc = d.ncell;
    for (long f = f0 + threadIdx.x; f < f1; f += kBlock) {
      const int a = (int)(f % na), b = (int)(f / na);
      int i, j, k;
      if (dir == 1) { i = up ? d.n1 : 0; j = a; k = b; }
      else if (dir == 2) { i = a; j = up ? d.n2 : 0; k = b; }
      else { i = a; j = b; k = up ? d.n3 : 0; }
      const long o = hd::cidx(d, i, j, k);
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) {
        const double x = F[pl * nc + o] * surface;
        v[pl] += isfinite(x) ? x : 0.0;
      }
    }
  }
  block_sum8_part(v, true, part, blockIdx.x);
}

// pdot of each boundary from its kEnFaceBlocks partials; pcum += pdot * Dt
// with Dt = dt * weight, the substep's share of the step (diag.c:1209).
__global__ void k_ener_pdot_final(const double* __restrict__ part, double* __restrict__ en,
                                  const double* __restrict__ ds, double weight)
{
  const int t = threadIdx.x;
  if (t >= 6 * HARM_NPR) return;
  const int bd = t / HARM_NPR, c = t % HARM_NPR;
  double s = 0.0;
  for (int q = 0; q < kEnFaceBlocks; q++) s += part[(long)(bd * kEnFaceBlocks + q) * 8 + c];
  en[EN_PDOT + t] = s;
  en[EN_PCUM + t] += s * (ds[DS_DT] * weight);
}

// dt4diag (advance.c:251-262): on the last substep of every
// kDiagSourceStep-th step the source is sampled with the time gathered
// since the previous sample; every other substep adds dt to that time.
__global__ void k_ener_dt4(double* __restrict__ en, const double* __restrict__ ds, int last)
{
  if (threadIdx.x != 0) return;
  const long nstep = (long)ds[DS_NSTEP];
  if (last && nstep % kDiagSourceStep == 0) {
    en[EN_DT4] = en[EN_WILLBE];
    en[EN_WILLBE] = 0.0;
  } else {
    en[EN_WILLBE] += ds[DS_DT];
    en[EN_DT4] = -1.0;
  }
}

// diag_source_all: sourceadd += Dt dV dUgeom with Dt = dt4diag > 0; dUgeom
// is the connection source of pb (source_conn, as the update adds it), zero
// for the density and the field.
__global__ __launch_bounds__(kBlock) void k_ener_source(Dims d, Params P, const double* __restrict__ pb,
                                                        const double* __restrict__ gcent,
                                                        const double* __restrict__ conn,
                                                        const double* __restrict__ en, double* __restrict__ part)
{
  const double Dt = en[EN_DT4];
  const int t = blockIdx.x * kBlock + threadIdx.x;
  double v[HARM_NPR] = {0, 0, 0, 0, 0, 0, 0, 0};
  const bool on = Dt > 0.0 && t < d.n1 * d.n2 * d.n3;
  if (on) {
    int i, j, k;
    zone_ijk(d, t, i, j, k);
    const long o = hd::cidx(d, i, j, k), pp = hd::pidx(d, i, j), nc = d.ncell, np = d.nplane;
    Geom G;
    hd::load_geom(gcent, pp, np, G);
    double pz[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) pz[pl] = pb[pl * nc + o];
    State q;
    hd::get_state(pz, G, P.gam, q);
    double dUg[4], T00;
    const double* cz = conn + pp;
    hd::source_rows(pz, G, q, [&](int c) { return cz[(long)c * np]; }, dUg, T00);
    const double w = Dt * (P.dx[1] * P.dx[2] * P.dx[3]);
#pragma unroll
    for (int l = 0; l < 4; l++) v[HARM_UU + l] = w * dUg[l];
  }
  block_sum8_part(v, on, part, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_ener_totals(Dims d, Params P, const double* __restrict__ p,
                                                        const double* __restrict__ gcent,
                                                        double* __restrict__ part)
{
  const int t = blockIdx.x * kBlock + threadIdx.x;
  double v[HARM_NPR] = {0, 0, 0, 0, 0, 0, 0, 0};
  const bool on = t < d.n1 * d.n2 * d.n3;
  if (on) {
    int i, j, k;
    zone_ijk(d, t, i, j, k);
    const long o = hd::cidx(d, i, j, k), nc = d.ncell;
    Geom G;
    hd::load_geom(gcent, hd::pidx(d, i, j), d.nplane, G);
    double pz[HARM_NPR], U[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) pz[pl] = p[pl * nc + o];
    State q;
    hd::get_state(pz, G, P.gam, q);
    hd::primtoflux(pz, q, 0, G, U);
    const double dV = P.dx[1] * P.dx[2] * P.dx[3];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) v[pl] = U[pl] * dV;
  }
  block_sum8_part(v, on, part, blockIdx.x);
}

// dV (U(after) - U(before)) of each fix: floors (flag bit 0) from fx_p0 to
// the after-floor state (fx_pfl when the gamma limit followed, else pf); the
// gamma limit (bit 1) from fx_p0 to pf.  Per block into part (floors) and
// part + nblk * 8 (gamma limits).
__global__ __launch_bounds__(kBlock) void k_ener_fix(Dims d, Params P, const double* __restrict__ pf,
                                                     const double* __restrict__ gcent,
                                                     const double* __restrict__ fx_p0,
                                                     const double* __restrict__ fx_pfl,
                                                     const uint8_t* __restrict__ fx_flag, double* __restrict__ part)
{
  const int t = blockIdx.x * kBlock + threadIdx.x;
  double vf[HARM_NPR] = {0, 0, 0, 0, 0, 0, 0, 0}, vg[HARM_NPR] = {0, 0, 0, 0, 0, 0, 0, 0};
  int fl = 0;
  if (t < d.n1 * d.n2 * d.n3) {
    int i, j, k;
    zone_ijk(d, t, i, j, k);
    const long o = hd::cidx(d, i, j, k), nc = d.ncell;
    fl = fx_flag[o];
    if (fl) {
      Geom G;
      hd::load_geom(gcent, hd::pidx(d, i, j), d.nplane, G);
      double p0[HARM_NPR], pa[HARM_NPR];
#pragma unroll
      for (int pl = 0; pl < HARM_NPR; pl++) p0[pl] = fx_p0[pl * nc + o];
      if (fl & 1) {
        const double* src = (fl & 2) ? fx_pfl : pf;
#pragma unroll
        for (int pl = 0; pl < HARM_NPR; pl++) pa[pl] = src[pl * nc + o];
        fix_delta_to(P, G, p0, pa, vf);
      }
      if (fl & 2) {
#pragma unroll
        for (int pl = 0; pl < HARM_NPR; pl++) pa[pl] = pf[pl * nc + o];
        fix_delta_to(P, G, p0, pa, vg);
      }
    }
  }
  block_sum8_part(vf, (fl & 1) != 0, part, blockIdx.x);
  block_sum8_part(vg, (fl & 2) != 0, part + (long)gridDim.x * 8, blockIdx.x);
}

// |div B| at the corner (i-1/2, j-1/2, k-1/2) of zone (i, j, k)
// (DIVBFLUXCTTOTHPRIM and its normalisation, global.fieldmacros.h:223-244):
// gdet B averaged onto the corner-centred faces, differenced, divided by the
// same average of |gdet B|.  A direction with one zone uses the zone itself
// as its neighbour (im1mac and friends), which zeroes its difference.
__device__ __forceinline__ double corner_divb(const Dims& d, const Params& P, const double* __restrict__ p,
                                              const double* __restrict__ gcent, int i, int j, int k)
{
  const long nc = d.ncell, gd = (long)HARM_G_GDET * d.nplane;
  const int im = d.n1 > 1 ? i - 1 : i, jm = d.n2 > 1 ? j - 1 : j, km = d.n3 > 1 ? k - 1 : k;
  // Fg over the 8 zones (ii, jj, kk) in {im, i} x {jm, j} x {km, k}
  double Fb[3][2][2][2], Fa[3][2][2][2];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int e = 0; e < 2; e++) {
          const int ii = a ? i : im, jj = b ? j : jm, kk = e ? k : km;
          const double g = gcent[gd + hd::pidx(d, ii, jj)];
          const double x = p[(HARM_B1 + c) * nc + hd::cidx(d, ii, jj, kk)] * g;
          Fb[c][a][b][e] = x;
          Fa[c][a][b][e] = fabs(p[(HARM_B1 + c) * nc + hd::cidx(d, ii, jj, kk)] * g);
        }
  // AVG_2 / AVG_1 then the second average, as the macros nest them
  auto for1 = [&](const double (&F)[3][2][2][2], int a) {  // AVG_for1 at i (a=1) or i-1 (a=0)
    const double s_k = 0.5 * (F[0][a][1][1] + F[0][a][0][1]);
    const double s_km = 0.5 * (F[0][a][1][0] + F[0][a][0][0]);
    return 0.5 * (s_k + s_km);
  };
  auto for2 = [&](const double (&F)[3][2][2][2], int b) {  // AVG_for2 at j (b=1) or j-1 (b=0)
    const double s_k = 0.5 * (F[1][1][b][1] + F[1][0][b][1]);
    const double s_km = 0.5 * (F[1][1][b][0] + F[1][0][b][0]);
    return 0.5 * (s_k + s_km);
  };
  auto for3 = [&](const double (&F)[3][2][2][2], int e) {  // AVG_for3 at k (e=1) or k-1 (e=0)
    const double s_i = 0.5 * (F[2][1][1][e] + F[2][1][0][e]);
    const double s_im = 0.5 * (F[2][0][1][e] + F[2][0][0][e]);
    return 0.5 * (s_i + s_im);
  };
  const double dfx = for1(Fb, 1) - for1(Fb, 0), dfy = for2(Fb, 1) - for2(Fb, 0), dfz = for3(Fb, 1) - for3(Fb, 0);
  const double nx = for1(Fa, 1) + for1(Fa, 0), ny = for2(Fa, 1) + for2(Fa, 0), nz = for3(Fa, 1) + for3(Fa, 0);
  const double norm = 1.0 / (fabs(nx / P.dx[1] + ny / P.dx[2] + nz / P.dx[3]) + D_SMALL);
  return fabs((dfx / P.dx[1] + dfy / P.dx[2] + dfz / P.dx[3]) * norm);
}

// per block: part[blk][0] = max, part[blk][1] = sum of |div B| over its zones
__global__ __launch_bounds__(kBlock) void k_ener_divb(Dims d, Params P, const double* __restrict__ p,
                                                      const double* __restrict__ gcent, double* __restrict__ part)
{
  __shared__ double rm[kBlock / 64], rs[kBlock / 64];
  const int t = blockIdx.x * kBlock + threadIdx.x;
  double m = 0.0, s = 0.0;
  if (t < d.n1 * d.n2 * d.n3) {
    int i, j, k;
    zone_ijk(d, t, i, j, k);
    m = s = corner_divb(d, P, p, gcent, i, j, k);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    m = fmax(m, __shfl_down(m, off, 64));
    s += __shfl_down(s, off, 64);
  }
  if (lane == 0) { rm[wid] = m; rs[wid] = s; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; w++) { m = fmax(m, rm[w]); s += rs[w]; }
    part[(long)blockIdx.x * 2] = m;
    part[(long)blockIdx.x * 2 + 1] = s;
  }
}

// One block per channel k < nchan: en[dst.v[k] + c] += sum_b part_k[b][c],
// part_k = part + k * n * 8.  128 groups of 8 lanes each sum every 128th
// block in increasing order (loads unrolled so they overlap), then lane c
// adds the 128 group sums in order: a fixed order, so fixed bits.
struct Dst4 {
  int v[4];
};
constexpr int kRedThreads = 1024;
__global__ __launch_bounds__(kRedThreads) void k_ener_reduce(const double* __restrict__ part, long n,
                                                             double* __restrict__ en, Dst4 dst, int assign)
{
  __shared__ double red[kRedThreads / 8][8];
  const double* pk = part + (long)blockIdx.x * n * 8;
  const int c = threadIdx.x & 7, g = threadIdx.x >> 3;
  double s = 0.0;
#pragma unroll 8
  for (long b = g; b < n; b += kRedThreads / 8) s += pk[b * 8 + c];
  red[g][c] = s;
  __syncthreads();
  if (threadIdx.x < 8) {
    double x = red[0][c];
    for (int q = 1; q < kRedThreads / 8; q++) x += red[q][c];
    const int d = dst.v[blockIdx.x];
    if (assign) en[d + c] = x;
    else en[d + c] += x;
  }
}

// divb max and sum from the per-block pairs: 64 lanes each take every 64th
// block in increasing order, then lane sums are combined in lane order
__global__ void k_ener_divb_final(const double* __restrict__ part, long n, double* __restrict__ out)
{
  __shared__ double rm[64], rs[64];
  const int t = threadIdx.x;
  double m = 0.0, s = 0.0;
  for (long b = t; b < n; b += 64) {
    m = fmax(m, part[b * 2]);
    s += part[b * 2 + 1];
  }
  rm[t] = m;
  rs[t] = s;
  __syncthreads();
  if (t == 0) {
    for (int q = 1; q < 64; q++) { m = fmax(m, rm[q]); s += rs[q]; }
    out[8] = m;
    out[9] = s;
  }
}

// The physics columns of one zone of HARM's main dump (dump_content,
// dump.c:439-600, CALCFARADAYANDCURRENTS): column c (DC_* order) of zone t =
// i + n1 (j + n2 k), the dump's row order, at out[c * cs + t * rs].  U is the
// conserved state the last inversion consumed (udump = unewglobal,
// initbase.c:885); J^mu from jcon (0 without it) and J_mu = lower_vec.
__global__ __launch_bounds__(kBlock) void k_dump_cols(Dims d, Params P, const double* __restrict__ p,
                                                      const double* __restrict__ ucum,
                                                      const double* __restrict__ gcent,
                                                      const double* __restrict__ jcon, double* __restrict__ out,
                                                      long cs, long rs)
{
  const long nz = (long)d.n1 * d.n2 * d.n3;
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nz) return;
  const int i = (int)(t % d.n1), j = (int)((t / d.n1) % d.n2), k = (int)(t / ((long)d.n1 * d.n2));
  const long o = hd::cidx(d, i, j, k), nc = d.ncell;
  Geom G;
  hd::load_geom(gcent, hd::pidx(d, i, j), d.nplane, G);
  double pz[HARM_NPR];
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) pz[pl] = p[pl * nc + o];
  auto put = [&](int c, double v) { out[(long)c * cs + t * rs] = v; };
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) put(DC_P + pl, pz[pl]);
  // ideal-gas EOS columns (idealgaseos.c:19, 48, 71)
  const double rho0 = pz[HARM_RHO], u = pz[HARM_UU];
  const double pressure = (P.gam - 1.) * u;
  put(DC_PRESS, pressure);
  put(DC_CS2, P.gam * pressure / (rho0 + u + pressure));
  const double indexn = 1.0 / (P.gam - 1.);
  double Sden;
  if (rho0 < D_SMALL && pressure < D_SMALL) Sden = 0.0;
  else if (rho0 < D_SMALL) Sden = 0.0;
  else if (pressure < D_SMALL) Sden = -kBIG;
  else Sden = rho0 * log(pow(pressure, indexn) / pow(rho0, indexn + 1.0));
  put(DC_SDEN, Sden);
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) put(DC_U + pl, ucum[pl * nc + o]);
  put(DC_DIVB, corner_divb(d, P, p, gcent, i, j, k));
  State q;
  hd::get_state(pz, G, P.gam, q);
  double bcon[4], bcov[4];
  bcon[0] = pz[HARM_B1] * q.ucov[1] + pz[HARM_B2] * q.ucov[2] + pz[HARM_B3] * q.ucov[3];
#pragma unroll
  for (int m = 1; m < 4; m++) bcon[m] = (pz[HARM_B1 - 1 + m] + bcon[0] * q.ucon[m]) / q.ucon[0];
  hd::lower(bcon, G, bcov);
#pragma unroll
  for (int m = 0; m < 4; m++) {
    put(DC_UCON + m, q.ucon[m]);
    put(DC_UCOV + m, q.ucov[m]);
    put(DC_BCON + m, bcon[m]);
    put(DC_BCOV + m, bcov[m]);
  }
#pragma unroll
  for (int dir = 1; dir <= 3; dir++) {
    double vmax, vmin;
    hd::vchar(pz, q, dir, G, P.gam, vmax, vmin);
    put(DC_VCHAR + 2 * (dir - 1), vmin);
    put(DC_VCHAR + 2 * (dir - 1) + 1, vmax);
  }
  put(DC_GDET, G.gdet);
  {
    double jc[4] = {0.0, 0.0, 0.0, 0.0}, jl[4];
    if (jcon) {
#pragma unroll
      for (int m = 0; m < 4; m++) jc[m] = jcon[(long)m * nc + o];
    }
    hd::lower(jc, G, jl);
#pragma unroll
    for (int m = 0; m < 4; m++) {
      put(DC_JCON + m, jc[m]);
      put(DC_JCOV + m, jl[m]);
    }
  }
  // Faraday F^{mu nu} at the centre (current_precalc CURTYPEFARADAY,
  // phys.c:701-708) and F_{mu nu} (lowerf, phys.tools.c:1449)
  const double ig = -1.0 / G.gdet;
  double f[6];
  f[0] = ig * (bcov[3] * q.ucov[2] - bcov[2] * q.ucov[3]);
  f[1] = ig * (-bcov[3] * q.ucov[1] + bcov[1] * q.ucov[3]);
  f[2] = ig * (bcov[2] * q.ucov[1] - bcov[1] * q.ucov[2]);
  f[3] = ig * (bcov[3] * q.ucov[0] - bcov[0] * q.ucov[3]);
  f[4] = ig * (-bcov[2] * q.ucov[0] + bcov[0] * q.ucov[2]);
  f[5] = ig * (bcov[1] * q.ucov[0] - bcov[0] * q.ucov[1]);
  double F[4][4];
  F[0][0] = F[1][1] = F[2][2] = F[3][3] = 0.0;
  F[0][1] = f[0]; F[0][2] = f[1]; F[0][3] = f[2]; F[1][2] = f[3]; F[1][3] = f[4]; F[2][3] = f[5];
  F[1][0] = -f[0]; F[2][0] = -f[1]; F[3][0] = -f[2]; F[2][1] = -f[3]; F[3][1] = -f[4]; F[3][2] = -f[5];
  const int lo[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
#pragma unroll
  for (int c = 0; c < 6; c++) {
    const int a = lo[c][0], b = lo[c][1];
    double s = 0.0;
#pragma unroll
    for (int jp = 0; jp < 4; jp++)
#pragma unroll
      for (int kp = 0; kp < 4; kp++) s += F[jp][kp] * hd::gco(G, a, jp) * hd::gco(G, b, kp);
    put(DC_FCON + c, f[c]);
    put(DC_FCOV + c, s);
  }
}

}  // namespace

void launch_dump_cols(const hd::Dims& d, const Params& P, const double* p, const double* ucum, const double* gcent,
                      const double* jcon, double* out, hipStream_t s)
{
  const long nz = (long)d.n1 * d.n2 * d.n3;
  hipLaunchKernelGGL(k_dump_cols, dim3(nblk(nz)), dim3(kBlock), 0, s, d, P, p, ucum, gcent, jcon, out, nz, 1L);
}

void launch_dump_rows(const hd::Dims& d, const Params& P, const double* p, const double* ucum, const double* gcent,
                      const double* jcon, double* rows, hipStream_t s)
{
  const long nz = (long)d.n1 * d.n2 * d.n3;
  hipLaunchKernelGGL(k_dump_cols, dim3(nblk(nz)), dim3(kBlock), 0, s, d, P, p, ucum, gcent, jcon,
                     rows + kDumpCoordCols, 1L, (long)HARM_DUMP_NCOL);
}

namespace {
// current_precalc (phys.c:622-686) for CURTYPET, X, Y, Z at the centre of
// every storage zone: the two previous F^{it} samples move to slots 4 and 5,
// slot 0 takes F^{it} now and slots 1-3 the components the x1, x2, x3
// derivatives of current_calc_1 read (slot 3's first sample is the
// reference's F^{tr}, phys.c:679).
__global__ __launch_bounds__(kBlock) void k_current_precalc(Dims d, Params P, const double* __restrict__ p,
                                                            const double* __restrict__ gcent,
                                                            double* __restrict__ cf)
{
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  const long nz = (long)d.ns1 * d.ns2 * d.ns3;
  if (t >= nz) return;
  const int i = (int)(t % d.ns1) - d.ng1, j = (int)((t / d.ns1) % d.ns2) - d.ng2,
            k = (int)(t / ((long)d.ns1 * d.ns2)) - d.ng3;
  const long o = hd::cidx(d, i, j, k), nc = d.ncell;
  Geom G;
  hd::load_geom(gcent, hd::pidx(d, i, j), d.nplane, G);
  double pz[HARM_NPR];
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) pz[pl] = p[pl * nc + o];
  State q;
  hd::get_state(pz, G, P.gam, q);
  double bcon[4], b[4];
  bcon[0] = pz[HARM_B1] * q.ucov[1] + pz[HARM_B2] * q.ucov[2] + pz[HARM_B3] * q.ucov[3];
#pragma unroll
  for (int m = 1; m < 4; m++) bcon[m] = (pz[HARM_B1 - 1 + m] + bcon[0] * q.ucon[m]) / q.ucon[0];
  hd::lower(bcon, G, b);
  const double* u = q.ucov;
  const double gd = G.gdet;
  auto F = [&](int slot, int c) -> double& { return cf[(long)(slot * 3 + c) * nc + o]; };
#pragma unroll
  for (int c = 0; c < 3; c++) {
    F(4, c) = F(5, c);
    F(5, c) = F(0, c);
  }
  F(0, 0) = -1.0 / gd * (-b[3] * u[2] + b[2] * u[3]);
  F(0, 1) = -1.0 / gd * (b[3] * u[1] - b[1] * u[3]);
  F(0, 2) = -1.0 / gd * (-b[2] * u[1] + b[1] * u[2]);
  F(1, 0) = -1.0 / gd * (b[3] * u[2] - b[2] * u[3]);
  F(1, 1) = -1.0 / gd * (-b[3] * u[0] + b[0] * u[3]);
  F(1, 2) = -1.0 / gd * (b[2] * u[0] - b[0] * u[2]);
  F(2, 0) = -1.0 / gd * (-b[3] * u[1] + b[1] * u[3]);
  F(2, 1) = -1.0 / gd * (b[3] * u[0] - b[0] * u[3]);
  F(2, 2) = -1.0 / gd * (-b[1] * u[0] + b[0] * u[1]);
  F(3, 0) = 1.0 / gd * (-b[3] * u[2] + b[2] * u[3]);
  F(3, 1) = 1.0 / gd * (b[2] * u[0] - b[0] * u[2]);
  F(3, 2) = 1.0 / gd * (-b[1] * u[0] + b[0] * u[1]);
}

// current_calc_1 (phys.c:846-1000) at the active zones: centred differences
// of gdet F over two zones (none along a dimension of one zone), and from
// the third call on the time derivative at the newest of three F^{it}
// samples dtl and dtr apart.
__global__ __launch_bounds__(kBlock) void k_current_calc(Dims d, Params P, const double* __restrict__ cf,
                                                         const double* __restrict__ gcent, int timeterms, double dtl,
                                                         double dtr, double* __restrict__ jcon)
{
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (long)d.n1 * d.n2 * d.n3) return;
  const int i = (int)(t % d.n1), j = (int)((t / d.n1) % d.n2), k = (int)(t / ((long)d.n1 * d.n2));
  const int s1 = d.n1 > 1, s2 = d.n2 > 1, s3 = d.n3 > 1;
  const long nc = d.ncell, gd = (long)HARM_G_GDET * d.nplane;
  const long o = hd::cidx(d, i, j, k);
  const long orp = hd::cidx(d, i + s1, j, k), orm = hd::cidx(d, i - s1, j, k);
  const long ohp = hd::cidx(d, i, j + s2, k), ohm = hd::cidx(d, i, j - s2, k);
  const long opp = hd::cidx(d, i, j, k + s3), opm = hd::cidx(d, i, j, k - s3);
  const double gt = gcent[gd + hd::pidx(d, i, j)];
  const double grp = gcent[gd + hd::pidx(d, i + s1, j)], grm = gcent[gd + hd::pidx(d, i - s1, j)];
  const double ghp = gcent[gd + hd::pidx(d, i, j + s2)], ghm = gcent[gd + hd::pidx(d, i, j - s2)];
  auto F = [&](int slot, int c, long at) { return cf[(long)(slot * 3 + c) * nc + at]; };
  const double idx1 = 1.0 / (2.0 * P.dx[1]), idx2 = 1.0 / (2.0 * P.dx[2]), idx3 = 1.0 / (2.0 * P.dx[3]);
  double tt[4] = {0.0, 0.0, 0.0, 0.0};
  if (timeterms) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double fl = F(4, c, o), f0 = F(5, c, o), fr = F(0, c, o);
      const double AA = (dtr * (fl - f0) + dtl * (fr - f0)) / (dtl * dtr * (dtl + dtr));
      const double BB = (-dtr * dtr * (fl - f0) + dtl * dtl * (fr - f0)) / (dtl * dtr * (dtl + dtr));
      tt[1 + c] = 2.0 * AA * dtr + BB;
    }
  }
  jcon[0 * nc + o] = +1. / gt * (grp * F(1, 0, orp) - grm * F(1, 0, orm)) * idx1
                     + 1. / gt * (ghp * F(2, 0, ohp) - ghm * F(2, 0, ohm)) * idx2
                     + 1. / gt * (gt * F(3, 0, opp) - gt * F(3, 0, opm)) * idx3;
  jcon[1 * nc + o] = +tt[1] + 1. / gt * (ghp * F(2, 1, ohp) - ghm * F(2, 1, ohm)) * idx2
                     + 1. / gt * (gt * F(3, 1, opp) - gt * F(3, 1, opm)) * idx3;
  jcon[2 * nc + o] = +tt[2] + 1. / gt * (grp * F(1, 1, orp) - grm * F(1, 1, orm)) * idx1
                     + 1. / gt * (gt * F(3, 2, opp) - gt * F(3, 2, opm)) * idx3;
  jcon[3 * nc + o] = +tt[3] + 1. / gt * (grp * F(1, 2, orp) - grm * F(1, 2, orm)) * idx1
                     + 1. / gt * (ghp * F(2, 2, ohp) - ghm * F(2, 2, ohm)) * idx2;
}
}  // namespace

void launch_current_precalc(const hd::Dims& d, const Params& P, const double* p, const double* gcent, double* cf,
                            hipStream_t s)
{
  hipLaunchKernelGGL(k_current_precalc, dim3(nblk((long)d.ns1 * d.ns2 * d.ns3)), dim3(kBlock), 0, s, d, P, p, gcent,
                     cf);
}

void launch_current_calc(const hd::Dims& d, const Params& P, const double* cf, const double* gcent, int timeterms,
                         double dtl, double dtr, double* jcon, hipStream_t s)
{
  hipLaunchKernelGGL(k_current_calc, dim3(nblk((long)d.n1 * d.n2 * d.n3)), dim3(kBlock), 0, s, d, P, cf, gcent,
                     timeterms, dtl, dtr, jcon);
}

long ener_part_blocks(const hd::Dims& d)
{
  const long nz = nblk((long)d.n1 * d.n2 * d.n3);
  return nz > 6 * kEnFaceBlocks ? nz : 6 * kEnFaceBlocks;
}

void launch_ener_pdot(const hd::Dims& d, const Params& P, const double* F1, const double* F2, const double* F3,
                      double* part, double* en, const double* ds, double weight, hipStream_t s)
{
  hipLaunchKernelGGL(k_ener_pdot, dim3(6 * kEnFaceBlocks), dim3(kBlock), 0, s, d, P, F1, F2, F3, part);
  hipLaunchKernelGGL(k_ener_pdot_final, dim3(1), dim3(64), 0, s, part, en, ds, weight);
}

void launch_ener_reduce(const double* part, long nb, double* en, const int* dst, int nchan, hipStream_t s)
{
  Dst4 d{{0, 0, 0, 0}};
  for (int k = 0; k < nchan && k < 4; k++) d.v[k] = dst[k];
  hipLaunchKernelGGL(k_ener_reduce, dim3(nchan), dim3(kRedThreads), 0, s, part, nb, en, d, 0);
}

void launch_ener_fix(const hd::Dims& d, const Params& P, const double* pf, const double* gcent, const double* fx_p0,
                     const double* fx_pfl, const uint8_t* fx_flag, double* part, double* en, hipStream_t s)
{
  const unsigned nb = nblk((long)d.n1 * d.n2 * d.n3);
  hipLaunchKernelGGL(k_ener_fix, dim3(nb), dim3(kBlock), 0, s, d, P, pf, gcent, fx_p0, fx_pfl, fx_flag, part);
  const int dst[2] = {EN_FLTERMS + FL_FLOOR * HARM_NPR, EN_FLTERMS + FL_GAMMA * HARM_NPR};
  launch_ener_reduce(part, nb, en, dst, 2, s);
}

void launch_ener_dt4(double* en, const double* ds, int last, hipStream_t s)
{
  hipLaunchKernelGGL(k_ener_dt4, dim3(1), dim3(1), 0, s, en, ds, last);
}

void launch_ener_source(const hd::Dims& d, const Params& P, const double* pb, const double* gcent,
                        const double* conn, double* part, double* en, hipStream_t s)
{
  const unsigned nb = nblk((long)d.n1 * d.n2 * d.n3);
  hipLaunchKernelGGL(k_ener_source, dim3(nb), dim3(kBlock), 0, s, d, P, pb, gcent, conn, en, part);
  hipLaunchKernelGGL(k_ener_reduce, dim3(1), dim3(kRedThreads), 0, s, part, (long)nb, en, Dst4{{EN_SRCADD, 0, 0, 0}},
                     0);
}

void launch_ener_totals(const hd::Dims& d, const Params& P, const double* p, const double* gcent, double* part,
                        double* out, hipStream_t s)
{
  const unsigned nb = nblk((long)d.n1 * d.n2 * d.n3);
  hipLaunchKernelGGL(k_ener_totals, dim3(nb), dim3(kBlock), 0, s, d, P, p, gcent, part);
  hipLaunchKernelGGL(k_ener_reduce, dim3(1), dim3(kRedThreads), 0, s, part, (long)nb, out, Dst4{{0, 0, 0, 0}}, 1);
  hipLaunchKernelGGL(k_ener_divb, dim3(nb), dim3(kBlock), 0, s, d, P, p, gcent, part);
  hipLaunchKernelGGL(k_ener_divb_final, dim3(1), dim3(64), 0, s, part, (long)nb, out);
}

}  // namespace hk
