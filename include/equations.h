// equations.h -- the generic finite-volume operators and the plugin contract.
// Replaces the reference interface src/equations.h:1-194 (same typedefs, struct layout,
// names and signatures).  The operators run on the GPU: equations_derivative and friends
// accept host pointers (copy in / copy out); the Advance methods keep the state resident.
#pragma once

#include "matrix.h"
#include "mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct Equations Equations;

// src/equations.h:9-41 -- plugin callback types
typedef double Timestep(const void *primitive, const double *property, double volume,
                        Vector projection);
typedef void Convective(void *flux, const void *left, const void *right, const double *property,
                        Matrix basis);
typedef Convective *ConvectiveSelect(const char *name);
typedef void Viscous(void *flux, const void *primitive, const void *gradient,
                     const double *property, Matrix basis);
typedef Viscous *ViscousSelect(const char *name);
typedef void Compute(void *, const double *property, Vector center, double time,
                     const void *context);
typedef void Boundary(void *outer, const void *inner, const double *property, Matrix basis,
                      const void *context);
typedef Boundary *BoundarySelect(const char *name);
typedef void Convert(void *, const void *, const double *property);
typedef void Prepare(const Equations *, const void *primitive);

typedef struct {
    Timestep *compute;
} EquationsTimestep;

typedef struct {
    String name;
    Convective *flux;
    ConvectiveSelect *select;
} EquationsConvective;

typedef struct {
    String name;
    Viscous *flux;
    ViscousSelect *select;
} EquationsViscous;

typedef struct {
    int kind;
    String name;
    double *parameter;
} EquationsLimiter;

typedef struct {
    int num;
    String *name;
    Compute **compute;
    Boundary **condition;
    const void **context;
    BoundarySelect *select;
} EquationsBoundary;

typedef struct {
    int num;
    int stride;
    String *name;
    int *dimension;
    void *data;
    union {
        Convert *convert;
        Compute *compute;
    } func;
} EquationsVariables;

typedef struct {
    int num;
    String *name;
    double *data;
} EquationsProperties;

typedef struct {
    Compute *compute;
    Prepare *prepare;
} EquationsSource;

// src/equations.h:97-111, plus `device`: the GPU mirror owned by this library (opaque).
struct Equations {
    const Mesh *mesh;
    String name;
    int space_order;
    EquationsTimestep timestep;
    EquationsConvective convective;
    EquationsViscous viscous;
    EquationsLimiter limiter;
    EquationsBoundary boundary;
    EquationsVariables primitive;
    EquationsVariables conserved;
    EquationsVariables reference;
    EquationsProperties properties;
    EquationsSource source;
    void *device;
};

// limiter kinds (src/equations/limiter.h:5-9)
enum LimiterKind {
    NONE,
    BARTH_JESPERSEN,
    VENKATAKRISHNAN,
};

Equations *equations_create(const Mesh *mesh, const char *name, Timestep *timestep,
                            ConvectiveSelect *convective, ViscousSelect *viscous,
                            BoundarySelect *boundary);                          // :114
void equations_create_primitive(Equations *eqns, const char **name, const int *dimension,
                                Convert *convert, int num);                     // :119
void equations_create_conserved(Equations *eqns, const char **name, const int *dimension,
                                Convert *convert, int num);                     // :123
void equations_create_reference(Equations *eqns, Compute *compute);             // :127
void equations_create_properties(Equations *eqns, const char **name, const double *property,
                                 int num);                                      // :130
void equations_set_space_order(Equations *eqns, int space_order);               // :134
void equations_set_convective_flux(Equations *eqns, const char *name);          // :137
void equations_set_viscous_flux(Equations *eqns, const char *name);             // :140
void equations_set_limiter(Equations *eqns, const char *name, double coefficient); // :143
void equations_set_boundary_condition(Equations *eqns, const char *entity, const char *name,
                                      Compute *compute, const void *context);   // :146
void equations_set_initial_condition(Equations *eqns, const char *entity, Compute *compute,
                                     const void *initial);                      // :150
void equations_set_property(Equations *eqns, const char *name, double property); // :154
void equations_set_source(Equations *eqns, Compute *compute, Prepare *prepare); // :157
void equations_summary(const Equations *eqns);                                  // :160

// The residual path (src/equations.h:163-178), computed on the GPU.
double equations_timestep(const Equations *eqns, const void *primitive);
void equations_boundary(const Equations *eqns, void *primitive, double time);
void equations_gradient(const Equations *eqns, void *primitive, void *gradient);
void equations_limiter(const Equations *eqns, const void *primitive, void *gradient);
void equations_derivative(const Equations *eqns, void *primitive, void *derivative, double time);
void equations_residual(const Equations *eqns, const void *derivative, void *residual);

void equations_average(const Equations *eqns, const char *entity, const void *primitive,
                       void *average);                                          // :181
void equations_norm(const Equations *eqns, void *norm, double time);            // :184
// VTKHDF output and nearest-centre restart (src/equations.h:187-190), serial HDF5.
void equations_write(const Equations *eqns, const char *name, double time, int index);
void equations_read(const Equations *eqns, const char *fname, double *time, int *index);
void equations_destroy(Equations *eqns);                                        // :194

#ifdef __cplusplus
}
#endif
