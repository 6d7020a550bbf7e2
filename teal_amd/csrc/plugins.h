// plugins.h -- device ids of the host plugin callbacks (internal to the library).
#pragma once

#include "equations.h"

#ifdef __cplusplus
extern "C" {
#endif

int teal_flux_id(Convective *fn);       // TealFlux id, -1 if the callback is not ours
int teal_viscous_id(Viscous *fn);       // 1 for the central viscous flux, 0 none, -1 foreign
int teal_bc_id(Boundary *fn);           // TealBoundary id, -1 if foreign
int teal_timestep_id(Timestep *fn);     // 0 Euler, 1 Navier-Stokes, -1 foreign
int teal_convert_is_euler(Convert *prim2cons, Convert *cons2prim);
// the polar file of euler_polar (src/euler/polar.c:60-79): alpha, the face centres, the
// pressure coefficients and the lift and drag coefficients (io.c)
void teal_io_write_polar(const char *fname, double alpha_deg, const Vector *center,
                         const double *pressure_c, int num, double lift, double drag);

#ifdef __cplusplus
}
#endif
