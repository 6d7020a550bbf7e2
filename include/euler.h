// euler.h -- compressible Euler plugin. Replaces the reference interface src/euler.h:1-36.
#pragma once

#include "equations.h"
#include "simulation.h"
#include "teal.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double density;
    Vector velocity;
    double pressure;
} EulerPrimitive;

typedef struct {
    double density;
    Vector momentum;
    double energy;
} EulerConserved;

enum EulerProperty {
    EULER_HEAT_CAPACITY_RATIO,
};

Timestep euler_timestep;                 // src/euler.h:23
ConvectiveSelect euler_convective;       // godunov roe hll hllc hlle ausmd ausmdv lxf central
BoundarySelect euler_boundary;           // symmetry/slipwall, supersonic/subsonic in/outflow,
                                         // pressure outflow, farfield
Convert euler_primitive, euler_conserved; // src/euler.h:26

Equations *euler_create(const Mesh *mesh); // src/euler.h:29

EulerPrimitive euler_riemann(EulerPrimitive left, EulerPrimitive right, double gamma,
                             double loc);  // src/euler.h:32

// Airfoil polar post-processing (src/euler.h:35, src/euler/polar.c:22-83): force sums on the
// device, lift and drag printed, <prefix>_polar_<alpha>.h5 written when the prefix is set.
void euler_polar(const Simulation *sim, const char *entity, EulerPrimitive reference,
                 double length, double time);

#ifdef __cplusplus
}
#endif
