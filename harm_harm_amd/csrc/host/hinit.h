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

/* Bondi flow (Gamma, sonic point r = 8, K = 1) at radius r: rho, u and the
 * Kerr-Schild radial 4-velocity; nonzero when no flow passes r. */
int hi_bondi_state(double gam, double r, double* rho, double* uu, double* ur);
/* Fills the active zones with the Bondi flow (fields zero). */
int hi_init_bondi(const harm_config* c, const hi_grid* G, double* p);

#ifdef __cplusplus
}
#endif
#endif
