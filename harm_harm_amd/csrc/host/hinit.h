/* hinit.h -- host-side initial data for the torus problem. */
#ifndef HINIT_H
#define HINIT_H

#include <stdint.h>
#include "harmstep.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hi_grid {
  int ng1, ng2, ng3;
  double startx[4], dx[4];
} hi_grid;

/* Fills the active zones of the HARM-layout array p (ghosts untouched). */
int hi_init_torus(const harm_config* c, const hi_grid* G, double rin, double rmax, double kappa, double beta,
                  double randfact, uint32_t seed, double* p);

#ifdef __cplusplus
}
#endif
#endif
