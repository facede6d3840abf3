// batch.hip -- per-zone kernels over n independent zones (SoA device buffers).
//
// harm_invert_batch: Utoprimgen (utoprimgen.c:32) for the inversion
// microbench, as the torus configuration compiles it (the hot solve, then the
// cold retry of a failed one; no inversion check); harm_prim_to_cons_batch: primtoU(UEVOLVE) (phys.tools.c:475).
#include <hip/hip_runtime.h>
#include "dphys.h"
#include "hhost.h"

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_invert_batch(const double* __restrict__ U, const double* __restrict__ geom,
                                                         double* __restrict__ p, int32_t* __restrict__ pflag,
                                                         int32_t* __restrict__ niter, long n, double gam,
                                                         double gammafail, const double* __restrict__ wpseed)
{
  const long z = (long)blockIdx.x * kBlock + threadIdx.x;
  if (z >= n) return;
  hd::Geom g;
  hd::load_geom(geom, z, n, g);
  double Uz[HARM_NPR], pz[HARM_NPR];
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) {
    Uz[pl] = U[pl * n + z];
    pz[pl] = p[pl * n + z];
  }
  int ns = 0;
  const double seed = wpseed ? wpseed[z] : __builtin_nan("");
  int fl = hd::utoprimgen(Uz, g, gam, gammafail, pz, ns, seed);
  if (hd::needs_after_hot(fl, true, false)) {  // HOT2COLD (the torus configuration)
    double Ugf[HARM_NPR], p0[HARM_NPR];
#pragma unroll
    for (int pl = 0; pl < HARM_NPR; pl++) {
      Ugf[pl] = U[pl * n + z];
      p0[pl] = p[pl * n + z];
    }
    hd::geometry_free(Ugf, g, Ugf);
    fl = hd::after_hot(fl, Ugf, g, gam, gammafail, p0, pz, ns, true, false);
  }
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) p[pl * n + z] = pz[pl];
  pflag[z] = fl;
  if (niter) niter[z] = ns;
}

__global__ __launch_bounds__(kBlock) void k_p2u_batch(const double* __restrict__ p, const double* __restrict__ geom,
                                                      double* __restrict__ U, long n, double gam)
{
  const long z = (long)blockIdx.x * kBlock + threadIdx.x;
  if (z >= n) return;
  hd::Geom g;
  hd::load_geom(geom, z, n, g);
  double pz[HARM_NPR], Uz[HARM_NPR];
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) pz[pl] = p[pl * n + z];
  hd::State q;
  hd::get_state(pz, g, gam, q);
  hd::primtoflux(pz, q, 0, g, Uz);
#pragma unroll
  for (int pl = 0; pl < HARM_NPR; pl++) U[pl * n + z] = Uz[pl];
}

}  // namespace

namespace hh {

harm_status invert_batch_seeded(const double* U, const double* geom, double* p, int32_t* pflag, int32_t* niter,
                                int64_t n, double gam, double gammafail, const double* wpseed, hipStream_t s)
{
  if (n <= 0 || !U || !geom || !p || !pflag) return set_error(HARM_EINVAL, "invert_batch: bad arguments");
  const long nb = (long)((n + kBlock - 1) / kBlock);
  if (nb > 0x7fffffffL) return set_error(HARM_EINVAL, "invert_batch: n too large");
  hipLaunchKernelGGL(k_invert_batch, dim3((unsigned)nb), dim3(kBlock), 0, s, U, geom, p, pflag, niter, (long)n, gam,
                     gammafail, wpseed);
  return check_launch("k_invert_batch");
}

}  // namespace hh

extern "C" harm_status harm_invert_batch(const double* U, const double* geom, double* p_inout, int32_t* pflag,
                                         int32_t* niter, int64_t n, double gam, double gammafail, void* stream)
{
  return hh::invert_batch_seeded(U, geom, p_inout, pflag, niter, n, gam, gammafail, nullptr, (hipStream_t)stream);
}

extern "C" harm_status harm_prim_to_cons_batch(const double* p, const double* geom, double* U, int64_t n, double gam,
                                               void* stream)
{
  if (n <= 0 || !p || !geom || !U) return hh::set_error(HARM_EINVAL, "prim_to_cons_batch: bad arguments");
  const long nb = (long)((n + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_p2u_batch, dim3((unsigned)nb), dim3(kBlock), 0, (hipStream_t)stream, p, geom, U, (long)n, gam);
  return hh::check_launch("k_p2u_batch");
}
