// hhost.h -- host-side helpers shared by the library's .hip translation units.
#pragma once
#include <hip/hip_runtime.h>
#include "harmstep.h"

namespace hh {

harm_status set_error(harm_status st, const char* fmt, ...);
harm_status check_hip(hipError_t e, const char* what);
harm_status check_launch(const char* kernel);

harm_status invert_batch_seeded(const double* U, const double* geom, double* p, int32_t* pflag, int32_t* niter,
                                int64_t n, double gam, double gammafail, const double* wpseed, hipStream_t s);

}  // namespace hh

#define HH_TRY(expr)                          \
  do {                                        \
    harm_status _st = (expr);                 \
    if (_st != HARM_OK) return _st;           \
  } while (0)

#define HH_HIP(expr) HH_TRY(hh::check_hip((expr), #expr))
