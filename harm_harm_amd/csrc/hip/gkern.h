// gkern.h -- definitions shared by the grid step's translation units: the
// device scalar slots, run parameters, face-tile plans, the XCD-aware block
// order, block reductions and the per-grid descriptor of batched launches.
// The face-flux kernels live in their own translation unit (flux.hip), the
// WENO5BND line reconstruction in weno.hip; everything else of the step is in
// grid.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "dgrid.h"

namespace hk {

using hd::Dims;

constexpr int kBlock = 256;  // threads per block of every grid kernel
constexpr double kBIG = 1.e+100;

// device scalar slots
enum { DS_DT = 0, DS_T = 1, DS_NDT1 = 2, DS_NDT2 = 3, DS_NDT3 = 4, DS_ACCDT = 5, DS_LASTNDT = 6, DS_NSTEP = 7,
       DS_FLUXACC = 8, DS_DTSTEP = 12, DS_N = 16 };  // DS_DTSTEP: the dt of the last whole step
// Newton counters: the current step's, and their sums over the steps since
// the time was last set
// and the zones whose learned-guess solve fell back to the plain guess
enum { DI_NSTROKE = 0, DI_NFAIL = 1, DI_NSTROKE_SUM = 2, DI_NFAIL_SUM = 3, DI_NFALLBACK = 4, DI_NFALLBACK_SUM = 5,
       DI_N = 8 };
// Energy-file accumulators (dump_ener.c:391-407), per grid: boundary flux
// sums of the last substep (pdot) and their time integral (pcum), six
// boundaries X1UP, X1DN, X2UP, X2DN, X3UP, X3DN by eight variables; the
// conserved change from the fix-ups by category (fladdterms, twelve
// categories by eight variables, their sum being fladd); the sampled
// geometric source (sourceadd); the source sample's weight (dt4diag) and the
// time gathered towards the next sample.
enum { EN_PDOT = 0, EN_PCUM = 48, EN_FLTERMS = 96, EN_SRCADD = 192, EN_DT4 = 200, EN_WILLBE = 201, EN_N = 256 };
// fix-up categories (NUMFAILFLOORFLAGS, global.nondepnmemonics.h:801-814)
enum { FL_CONV = 0, FL_FLOOR = 1, FL_GAMMA = 2, FL_INFLOW = 3, FL_RHONEG = 4, FL_UNEG = 5, FL_RHOUNEG = 6, FL_N = 12 };
// the failure categories failure averaging accounts to (fixuputoprim_accounting, fixup.c:1603-1720)
constexpr int kFailCats[4] = {FL_CONV, FL_RHONEG, FL_UNEG, FL_RHOUNEG};
__host__ __device__ inline int fail_slot(int pflag)  // index into kFailCats
{
  return pflag == HARM_UTOPRIMFAILRHONEG ? 1
         : (pflag == HARM_UTOPRIMFAILUNEG || pflag == HARM_UTOPRIMFAILU2AVG1 || pflag == HARM_UTOPRIMFAILU2AVG2) ? 2
         : pflag == HARM_UTOPRIMFAILRHOUNEG ? 3 : 0;
}
constexpr int kEnFaceBlocks = 32;  // blocks per boundary of the boundary-flux sums
constexpr int kDiagSourceStep = 4;  // DIAGSOURCECOMPSTEP (definit.h:19): full steps between source samples

struct Params {
  double gam, cour, safe, gammamax, gammafail;
  double dx[4], startx[4];
  double R0;  // r = R0 + exp(X1)
  int bc;     // HARM_PROBLEM_*: which radial and polar boundary conditions
  int hot2cold, checkinv, resetregion;  // the configuration's failure-handling switches (harm_config)
};

__device__ __forceinline__ void atomic_min_pos(double* addr, double v)
{
  // valid for v >= 0 (IEEE order of non-negative doubles equals integer order)
  atomicMin((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

// block-wide minimum followed by one atomic per block
__device__ __forceinline__ void block_min_to(double v, double* addr)
{
  __shared__ double red[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < kBlock / 64; w++) m = fmin(m, red[w]);
    atomic_min_pos(addr, m);
  }
}

// Per-block sums of eight values into part[blk][8], in a fixed tree order (the
// same inputs give the same bits).  `any`: whether this thread holds a nonzero
// value; a block where no thread does writes zeros without the tree.  Every
// thread of the block calls it.
__device__ __forceinline__ void block_sum8_part(const double* v, bool any, double* __restrict__ part, long blk)
{
  __shared__ double red[8][kBlock / 64];
  if (!__syncthreads_or(any)) {
    if (threadIdx.x < 8) part[blk * 8 + threadIdx.x] = 0.0;
    return;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    double x = v[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[c][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    double s = red[threadIdx.x][0];
    for (int w = 1; w < kBlock / 64; w++) s += red[threadIdx.x][w];
    part[blk * 8 + threadIdx.x] = s;
  }
}

// ---------------------------------------------------------------------------
// fluxes
// ---------------------------------------------------------------------------

struct Box {
  int is, js, ks, ni, nj, nk;
  __device__ __forceinline__ bool at(long t, int& i, int& j, int& k) const
  {
    const long n = (long)ni * nj * nk;
    if (t >= n) return false;
    i = is + (int)(t % ni);
    const long r = t / ni;
    j = js + (int)(r % nj);
    k = ks + (int)(r / nj);
    return true;
  }
  long count() const { return (long)ni * nj * nk; }
};

inline Box face_box(const Dims& d, int dir)
{
  const int e0 = d.n1 > 1, e1 = d.n2 > 1, e2 = d.n3 > 1;
  int is = 0, ie = d.n1 - 1, js = 0, je = d.n2 - 1, ks = 0, ke = d.n3 - 1;
  if (dir == 1) { ie = d.n1; js -= e1; je += e1; ks -= e2; ke += e2; }
  if (dir == 2) { je = d.n2; is -= e0; ie += e0; ks -= e2; ke += e2; }
  if (dir == 3) { ke = d.n3; is -= e0; ie += e0; js -= e1; je += e1; }
  Box b{is, js, ks, ie - is + 1, je - js + 1, ke - ks + 1};
  return b;
}

// Face tiles.  A block first reconstructs every cell its faces touch, one
// (cell, variable) pair per lane, into LDS (each cell once, both edges), then
// solves one face per lane.  Direction 1 runs along i, the contiguous
// dimension: a block takes `fpb` consecutive faces of the face box flattened
// as (i, k, j) (rows of ni faces own ni+1 cells, so a block needs
// fpb + rows + 1 cells).  Directions 2 and 3 take tc consecutive i by td
// faces along the sweep in one transverse plane (k for direction 2, j for
// direction 3), i.e. (td+1)*tc cells.  Tiles are ordered so that neighbours
// in the launch reuse the same (i, j) rows of the 2-D face geometry: phi
// fastest for directions 1 and 2, the sweep fastest for direction 3.
constexpr int kCellMax = 320;

struct FaceTile {
  Box b;
  int ncol;       // dir 1: rows (nj*nk)
  int nd;         // faces along the sweep
  int fpb;        // dir 1: faces per block
  int tc, td;     // dirs 2,3: tile shape
  int nti, ntd;   // dirs 2,3: tiles along i / along the sweep
  int nplane;     // dirs 2,3: transverse planes (nk for dir 2, nj for dir 3)
  int i0;         // dirs 2,3: i of the first tile's first column (-ng1: tiles start on 128-byte lines)
  long nblk;
};

// Blocks are dealt to the 8 XCDs round-robin; remap so that each XCD works on
// a contiguous range of tiles (neighbouring tiles share stencil halos in L2).
__device__ __forceinline__ int xcd_tile(int b, int n)
{
  const int x = b & 7, loc = b >> 3, q = n >> 3, r = n & 7;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + loc;
}

struct GridRef {
  Dims d;
  Params P;
  int fluxtype;
  double* arr[HARM_NARR];
  int32_t* pflag;
  double* pfix;
  const double* geom[HARM_NLOC];
  const double* gcorn3;
  const double* conn;
  const double* floors;
  const double* rcent;
  double* emf;
  int emf_ni, emf_nj, emf_nk;
  double* ds;
  unsigned long long* di;
  double* facc;
  const double* wl;  // WENO5BND face values, one block of HARM_NPR x ncell per direction (or null)
  const double* wr;
  FaceTile ft[4];
};

// A pointer read from memory (a GridRef) is generic to the compiler, so loads
// and stores through it would be flat instructions.  This says that it points
// to global memory (global_load/global_store): the pointer, the same in every
// lane (a grid's base address), is taken into scalar registers and passed
// through an empty asm typed as a global pointer, which the compiler cannot
// fold back into the generic one.
template <class T>
__device__ __forceinline__ T* gmem(T* p)
{
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  auto g = (__attribute__((address_space(1))) T*)(((unsigned long long)hi << 32) | lo);
  asm("" : "+s"(g));
  return (T*)g;
}

// Active-zone order for the per-zone kernels: i fastest (coalesced), then k,
// then j.  Consecutive rows are consecutive phi planes of the same (i, j)
// row, so the 2-D geometry and connection rows stay in L2 across phi; blocks
// are remapped (xcd_tile) so each XCD works on a contiguous range of rows.
__device__ __forceinline__ void zone_ijk(const Dims& d, int t, int& i, int& j, int& k)
{
  const int r = t / d.n1;
  i = t - r * d.n1;
  j = r / d.n3;
  k = r - j * d.n3;
}

// The step's inversion started from a learned W' (mlp.hip): harm_step's
// k_invert with the network in front of the Newton solve.
// With fx_p0, the states the final substep's fixes started from are kept
// as k_invert<true> keeps them (invert_zone).
harm_status launch_invert_grid_mlp(const harm_mlp* m, const hd::Dims& d, const Params& P, const double* U,
                                   const double* pg, double* pf, int32_t* pflag, const double* gcent,
                                   const double* floors, int finalstep, unsigned long long* di, hipStream_t s,
                                   double* fx_p0 = nullptr, double* fx_pfl = nullptr, uint8_t* fx_flag = nullptr);

// Energy-file diagnostics (ener.hip).  `part` holds per-block partial sums:
// at least 4 * ener_part_blocks(d) * 8 doubles.
long ener_part_blocks(const hd::Dims& d);
// boundary flux sums of F1..F3 into en[EN_PDOT..], and en[EN_PCUM..] += pdot * dt * weight
void launch_ener_pdot(const hd::Dims& d, const Params& P, const double* F1, const double* F2, const double* F3,
                      double* part, double* en, const double* ds, double weight, hipStream_t s);
// for each channel k < nchan (at most 4): en[dst[k] + c] += sum over nblk
// blocks of part[k][b][c] (fixed order), c < 8
void launch_ener_reduce(const double* part, long nblk, double* en, const int* dst, int nchan, hipStream_t s);
// the final substep's floor and gamma-limit changes into their fladdterms
// from the states the inversion kept (fx_*) and the primitives pf after it
void launch_ener_fix(const hd::Dims& d, const Params& P, const double* pf, const double* gcent, const double* fx_p0,
                     const double* fx_pfl, const uint8_t* fx_flag, double* part, double* en, hipStream_t s);
// The main dump's physics columns (dump_content, dump.c:439-600), device
// SoA out[c][nz] in the dump's row order (i fastest, then j, then k): the
// primitives, pressure, cs2, entropy density, U (from ucum), |div B|, u^mu,
// u_mu, b^mu, b_mu, (vmin, vmax) per direction, gdet, F^{mu nu}, F_{mu nu}.
// The coordinates and the current columns are the host's.
enum { DC_P = 0, DC_PRESS = 8, DC_CS2 = 9, DC_SDEN = 10, DC_U = 11, DC_DIVB = 19, DC_UCON = 20, DC_UCOV = 24,
       DC_BCON = 28, DC_BCOV = 32, DC_VCHAR = 36, DC_GDET = 42, DC_JCON = 43, DC_JCOV = 47, DC_FCON = 51,
       DC_FCOV = 57, DC_N = 63 };
void launch_dump_cols(const hd::Dims& d, const Params& P, const double* p, const double* ucum, const double* gcent,
                      const double* jcon, double* out, hipStream_t s);
// The same columns straight into the dump's rows: rows[z * HARM_DUMP_NCOL +
// c], columns kDumpCoordCols.. (the coordinate columns before them are left
// as they are).  jcon: J^mu (4 x ncell, harm_currents), or null for zeros.
enum { kDumpCoordCols = 9 };
static_assert(kDumpCoordCols + DC_N == HARM_DUMP_NCOL, "dump row layout");
void launch_dump_rows(const hd::Dims& d, const Params& P, const double* p, const double* ucum, const double* gcent,
                      const double* jcon, double* rows, hipStream_t s);
// The main dump's four-current (CURRENTCALC1): the Faraday samples of every
// storage zone (current_precalc phys.c:622; cf: 18 x ncell, slot * 3 + c) and
// J^mu of the active zones from them (current_calc_1 phys.c:846; jcon: 4 x
// ncell), the time term from the third call on.
void launch_current_precalc(const hd::Dims& d, const Params& P, const double* p, const double* gcent, double* cf,
                            hipStream_t s);
void launch_current_calc(const hd::Dims& d, const Params& P, const double* cf, const double* gcent, int timeterms,
                         double dtl, double dtr, double* jcon, hipStream_t s);
// source sample weight at the end of substep `to` of nto (advance.c:251-262)
void launch_ener_dt4(double* en, const double* ds, int last, hipStream_t s);
// sourceadd += dt4diag * dV * dUgeom(pb) over active zones when dt4diag > 0
void launch_ener_source(const hd::Dims& d, const Params& P, const double* pb, const double* gcent,
                        const double* conn, double* part, double* en, hipStream_t s);
// totals over active zones: U[8] = sum U(p) dV into out[0..7]; divb max and
// sum into out[8], out[9] (device array of 10 doubles)
void launch_ener_totals(const hd::Dims& d, const Params& P, const double* p, const double* gcent, double* part,
                        double* out, hipStream_t s);

// Launchers of the face-flux kernels (flux.hip): one direction of one grid,
// or one direction of every grid of a batch.  With WENO5BND the face values
// come from wl, wr (launch_weno_recon) instead of an in-tile reconstruction.
void launch_flux(int dir, int recon, const FaceTile& T, hipStream_t s, const Dims& d, const double* pb, double* F,
                 const double* gface, const Params& P, int fluxtype, int do_ct, double* ds,
                 const double* wl = nullptr, const double* wr = nullptr);
void launch_flux_batch(int dir, int recon, unsigned nblk, hipStream_t s, const GridRef* G, const int* pre, int ng,
                       int which, int do_ct);
// WENO5BND (weno.hip): the dP measure of every active direction, then the
// face values of one direction's lines
// (tm: 10 x ncell doubles of per-zone terms, dp: 3 x ncell)
void launch_weno_dp(const Dims& d, hipStream_t s, const double* pb, const double* gcent, const Params& P, double* tm,
                    double* dp);
void launch_weno_recon(int dir, const Dims& d, const Box& b, hipStream_t s, const double* pb, const double* dp,
                       double* L, double* R);

}  // namespace hk
