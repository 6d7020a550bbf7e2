// navierstokes.h -- Navier-Stokes plugin. Replaces the reference interface
// src/navierstokes.h:1-30.
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
} NavierStokesPrimitive;

typedef struct {
    double density;
    Vector momentum;
    double energy;
} NavierStokesConserved;

enum NavierStokesProperty {
    NAVIERSTOKES_HEAT_CAPACITY_RATIO,
    NAVIERSTOKES_DYNAMIC_VISCOSITY,
    NAVIERSTOKES_PRANDTL,
};

Timestep navierstokes_timestep;       // src/navierstokes.h:25
ViscousSelect navierstokes_viscous;   // "central"
BoundarySelect navierstokes_boundary; // "adiabatic wall"/"wall" + the Euler conditions

Equations *navierstokes_create(const Mesh *mesh); // src/navierstokes.h:30

#ifdef __cplusplus
}
#endif
