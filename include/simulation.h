// simulation.h -- time integration driver. Replaces the reference interface
// src/simulation.h:1-83 (same names, struct layout and signatures).
#pragma once

#include "equations.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef double Advance(const Equations *eqns, double *time, void *residual, double max_step,
                       double courant, const void *context);

typedef struct {
    double max;
    double out;
} SimulationTime;

typedef struct {
    int max;
    int out;
} SimulationIter;

typedef struct {
    String condition;
    int variable;
    double threshold;
} SimulationTermination;

typedef struct {
    int time_order;
    int num_stages;
} RungeKutta;

typedef struct {
    double newton_tolerance;
    double krylov_tolerance;
    int krylov_dimension;
} NewtonKrylov;

typedef struct {
    String name;
    double courant;
    void *context;
    Advance *method;
} SimulationAdvance;

typedef struct {
    const Equations *eqns;
    String prefix;
    SimulationTime time;
    SimulationIter iter;
    SimulationTermination termination;
    SimulationAdvance advance;
} Simulation;

// Advance methods selectable by name (src/simulation/advance.h:5): "euler", "lserk",
// "implicit euler".  They advance the GPU-resident state of eqns.
Advance euler, lserk, implicit_euler;

Simulation *simulation_create(const Equations *eqns, const char *prefix);       // :53
void simulation_set_max_time(Simulation *sim, double time);                     // :56
void simulation_set_out_time(Simulation *sim, double time);                     // :59
void simulation_set_max_iter(Simulation *sim, int iter);                        // :62
void simulation_set_out_iter(Simulation *sim, int iter);                        // :65
void simulation_set_termination(Simulation *sim, const char *condition, double threshold); // :68
void simulation_set_advance(Simulation *sim, const char *name, double courant,
                            const void *context);                              // :71
void simulation_summary(const Simulation *sim);                                 // :74
double simulation_run(Simulation *sim);                                         // :77
void simulation_error(const Simulation *sim, double time, void *norm);          // :80
void simulation_destroy(Simulation *sim);                                       // :83

#ifdef __cplusplus
}
#endif
