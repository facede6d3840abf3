// common.hip -- error reporting and host helpers of the C ABI.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include "hhost.h"
#include "../host/hgeom.h"

namespace {
thread_local char g_err[512] = "";
}

namespace hh {

harm_status set_error(harm_status st, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return st;
}

harm_status check_hip(hipError_t e, const char* what)
{
  if (e == hipSuccess) return HARM_OK;
  return set_error(e == hipErrorOutOfMemory ? HARM_ENOMEM : HARM_EHIP, "%s: %s", what, hipGetErrorString(e));
}

harm_status check_launch(const char* kernel)
{
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) return HARM_OK;
  return set_error(HARM_EHIP, "launch of %s failed: %s", kernel, hipGetErrorString(e));
}

}  // namespace hh

extern "C" const char* harm_last_error(void) { return g_err; }

extern "C" void harm_geom_ks_record(double a, double r, double th, double* rec) { hg_ks_record(a, r, th, rec); }

extern "C" void harm_geom_mks_record(const harm_config* cfg, const double X[4], double* rec)
{
  hg_mks_record(cfg, X, rec);
}

extern "C" void harm_geom_ks_batch(int64_t n, const double* a, const double* r, const double* th, double* rec)
{
  hg_ks_batch(n, a, r, th, rec);
}

extern "C" void harm_default_config(harm_config* c, int32_t n1, int32_t n2, int32_t n3)
{
  memset(c, 0, sizeof(*c));
  c->n1 = n1;
  c->n2 = n2;
  c->n3 = n3;
  c->ng = 4;
  c->recon = HARM_RECON_PARA;
  c->flux = HARM_FLUX_LAXF;
  c->timeorder = 2;
  c->a = 0.9375;
  c->gam = 4.0 / 3.0;
  c->R0 = 0.0;
  c->Rout = 40.0;
  c->Rin = hg_setRin(c->a, c->R0, c->Rout, n1);
  c->hslope = 1.04 * pow(0.3, 2.0 / 3.0);
  c->cour = 0.8;
  c->safe = 1.3;
  c->rhomin = 1e-4;
  c->uumin = 1e-6;
  c->gammamax = 25.0;
  c->gammafail = 2500.0;
  c->hot2cold = 1;     // init.fishmon.h:325
  c->checkinv = 0;     // init.fishmon.h:74, 225
  c->resetregion = 1;  // init.fishmon.h:356
}

extern "C" void harm_bondi_config(harm_config* c, int32_t n1, int32_t n2)
{
  harm_default_config(c, n1, n2, 1);
  c->problem = HARM_PROBLEM_BONDI;
  c->a = 0.0;
  c->R0 = 0.0;
  c->Rin = 1.9;
  c->Rout = 20.0;
  c->hslope = 1.0;
  c->cour = 0.5;
  c->rhomin = 0.0;
  c->uumin = 0.0;
  c->gammamax = 100.0;
  c->gammafail = 100.0 * c->gammamax;
  c->hot2cold = 0;     // init.sasha.h:186, 379
  c->checkinv = 1;     // init.sasha.h:73, 198
  c->resetregion = 0;  // init.sasha.h leaves USERRESETREGION unset
}

extern "C" harm_status harm_device_count(int32_t* n)
{
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *n = c;
  return HARM_OK;
}

extern "C" harm_status harm_dev_malloc(void** ptr, size_t bytes)
{
  if (!ptr) return hh::set_error(HARM_EINVAL, "dev_malloc: null");
  return hh::check_hip(hipMalloc(ptr, bytes ? bytes : 1), "hipMalloc");
}

extern "C" harm_status harm_dev_free(void* ptr) { return hh::check_hip(hipFree(ptr), "hipFree"); }

extern "C" harm_status harm_memcpy_htod(void* dst, const void* src, size_t bytes)
{
  return hh::check_hip(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy(HtoD)");
}

extern "C" harm_status harm_memcpy_dtoh(void* dst, const void* src, size_t bytes)
{
  return hh::check_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy(DtoH)");
}

extern "C" harm_status harm_memcpy_dtod(void* dst, const void* src, size_t bytes)
{
  return hh::check_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy(DtoD)");
}

extern "C" harm_status harm_device_synchronize(void) { return hh::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

extern "C" harm_status harm_event_create(void** ev)
{
  return hh::check_hip(hipEventCreate((hipEvent_t*)ev), "hipEventCreate");
}

extern "C" harm_status harm_event_destroy(void* ev) { return hh::check_hip(hipEventDestroy((hipEvent_t)ev), "hipEventDestroy"); }

extern "C" harm_status harm_event_record(void* ev, void* stream)
{
  return hh::check_hip(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream), "hipEventRecord");
}

extern "C" harm_status harm_event_elapsed_ms(void* ev0, void* ev1, float* ms)
{
  HH_HIP(hipEventSynchronize((hipEvent_t)ev1));
  return hh::check_hip(hipEventElapsedTime(ms, (hipEvent_t)ev0, (hipEvent_t)ev1), "hipEventElapsedTime");
}
