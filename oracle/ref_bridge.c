/* CPU ORACLE (test infrastructure only) -- bridge into the reference build.
 *
 * Compiled by oracle/Makefile against the reference's own headers and linked with the
 * reference sources of the residual path into oracle/_ref/libteal_ref.so.  It exposes
 * flat entry points so tests can run the reference's own equations_derivative, lserk,
 * implicit_euler, gmres, mesh_create and mesh_validate.
 *
 * src/mesh/generate.c needs metis.h (METIS is not in this image) and is not built: meshes
 * handed to the reference come from the oracle's restatement (omesh_generate), whose
 * struct layout is identical to src/mesh.h:13-72 (checked below). */
#include <stddef.h>
#include <string.h>

#include "equations.h"
#include "euler.h"
#include "mesh.h"
#include "navierstokes.h"
#include "simulation.h"
#include "simulation/advance.h"
#include "simulation/gmres.h"
#include "teal.h"

#include "oracle.h"

_Static_assert(sizeof(Mesh) == sizeof(OMesh), "oracle mesh layout");
_Static_assert(offsetof(Mesh, faces) == offsetof(OMesh, faces), "oracle mesh layout");
_Static_assert(offsetof(Mesh, entities) == offsetof(OMesh, entities), "oracle mesh layout");
_Static_assert(offsetof(Mesh, neighbors) == offsetof(OMesh, neighbors), "oracle mesh layout");
_Static_assert(offsetof(MeshFaces, correction) == offsetof(OFaces, correction), "faces layout");
_Static_assert(offsetof(MeshCells, offset) == offsetof(OCells, offset), "cells layout");
_Static_assert(sizeof(MeshEntities) == sizeof(OEntities), "entities layout");

static int initialized = 0;

void ref_init(void)
{
    if (initialized) {
        return;
    }
    static char arg0[] = "teal_ref";
    static char arg1[] = "-q";
    static char *argv_[] = {arg0, arg1, 0};
    int argc = 2;
    char **argv = argv_;
    teal_init(&argc, &argv);
    initialized = 1;
}

Mesh *ref_mesh_create(const double *min, const double *max, const int *n, const int *periodic)
{
    ref_init();
    Vector lo = {min[0], min[1], min[2]};
    Vector hi = {max[0], max[1], max[2]};
    return mesh_create(lo, hi, (Triple){n[0], n[1], n[2]}, (Triple){periodic[0], periodic[1], periodic[2]});
}

void ref_mesh_validate(const void *mesh)
{
    ref_init();
    mesh_validate((const Mesh *)mesh);
}

/* Equations on an oracle-generated mesh (layout-identical, see header) */
Equations *ref_eqns_create(const void *omesh, int navierstokes)
{
    ref_init();
    const Mesh *mesh = (const Mesh *)omesh;
    return navierstokes ? navierstokes_create(mesh) : euler_create(mesh);
}

void ref_eqns_config(Equations *eqns, int space_order, const char *flux, int viscous,
                     const char *limiter, double coefficient, double gamma, double mu,
                     double prandtl)
{
    equations_set_space_order(eqns, space_order);
    equations_set_convective_flux(eqns, flux);
    if (eqns->viscous.select) {
        equations_set_viscous_flux(eqns, viscous ? "central" : 0);
    }
    equations_set_limiter(eqns, limiter, coefficient);
    equations_set_property(eqns, "heat capacity ratio", gamma);
    if (eqns->viscous.select) {
        equations_set_property(eqns, "dynamic viscosity", mu);
        equations_set_property(eqns, "prandtl number", prandtl);
    }
}

static EulerPrimitive contexts[64];
static int num_contexts = 0;

void ref_eqns_set_bc(Equations *eqns, const char *entity, const char *name, const double *ctx)
{
    const void *context = 0;
    if (ctx) {
        EulerPrimitive *c = &contexts[num_contexts++ % 64];
        c->density = ctx[0];
        c->velocity = (Vector){ctx[1], ctx[2], ctx[3]};
        c->pressure = ctx[4];
        context = c;
    }
    equations_set_boundary_condition(eqns, entity, name, 0, context);
}

/* The periodic-hill driver's source term (run/periodic_hill/run.c:8-21) with its bulk
 * velocity and relaxation time as parameters: the reference calls the Prepare hook once per
 * equations_derivative and the source per inner cell (src/equations/derivative.c:229-267). */
static double forcing_velocity = 1, forcing_relaxation = 0.001;
static NavierStokesPrimitive forcing_average;

static void forcing_prepare(const Equations *eqns, const void *primitive)
{
    equations_average(eqns, "domain", primitive, &forcing_average);
}

static void forcing_source(void *source_, const double *property, Vector center, double time,
                           const void *context)
{
    (void)property, (void)center, (void)time;
    NavierStokesConserved *source = source_;
    const NavierStokesPrimitive *primitive = context;
    double s1 = primitive->density * (forcing_velocity - forcing_average.velocity.x) /
                forcing_relaxation;
    source->momentum.x = s1;
    source->energy = primitive->velocity.x * s1;
}

void ref_eqns_set_forcing(Equations *eqns, double velocity, double relaxation)
{
    forcing_velocity = velocity;
    forcing_relaxation = relaxation;
    equations_set_source(eqns, forcing_source, forcing_prepare);
}

void ref_average(const Equations *eqns, const char *entity, const double *p, double *avg)
{
    equations_average(eqns, entity, p, avg);
}

double *ref_eqns_primitive(Equations *eqns) { return eqns->primitive.data; }
double *ref_eqns_conserved(Equations *eqns) { return eqns->conserved.data; }
double *ref_eqns_eps2(Equations *eqns) { return eqns->limiter.parameter; }

void ref_boundary(const Equations *eqns, double *p) { equations_boundary(eqns, p, 0); }
void ref_gradient(const Equations *eqns, double *p, double *g) { equations_gradient(eqns, p, g); }
void ref_limiter(const Equations *eqns, const double *p, double *g) { equations_limiter(eqns, p, g); }
void ref_derivative(const Equations *eqns, double *p, double *d) { equations_derivative(eqns, p, d, 0); }
double ref_timestep(const Equations *eqns, const double *p) { return equations_timestep(eqns, p); }
void ref_residual(const Equations *eqns, const double *d, double *res) { equations_residual(eqns, d, res); }

double ref_euler(const Equations *eqns, double *time, double *res, double max_step, double courant)
{
    return euler(eqns, time, res, max_step, courant, 0);
}

double ref_lserk(const Equations *eqns, double *time, double *res, double max_step, double courant,
                 int time_order, int num_stages)
{
    RungeKutta ctx = {time_order, num_stages};
    return lserk(eqns, time, res, max_step, courant, &ctx);
}

double ref_implicit_euler(const Equations *eqns, double *time, double *res, double max_step,
                          double courant, double newton_tol, double krylov_tol, int dim)
{
    NewtonKrylov ctx = {newton_tol, krylov_tol, dim};
    return implicit_euler(eqns, time, res, max_step, courant, &ctx);
}

void ref_gmres(const Equations *eqns, const double *u, const double *d, const double *r, double *inc,
               double step, double norm, double fd_scale, double krylov_tol, int dim)
{
    NewtonKrylov ctx = {0.1, krylov_tol, dim};
    gmres(eqns, u, d, r, inc, 0, step, norm, fd_scale, &ctx);
}

static Matrix mat_of(const double *b)
{
    return (Matrix){{b[0], b[1], b[2]}, {b[3], b[4], b[5]}, {b[6], b[7], b[8]}};
}

void ref_flux(const char *name, double *flux, const double *l, const double *r,
              const double *property, const double *basis)
{
    ref_init();
    euler_convective(name)(flux, l, r, property, mat_of(basis));
}

void ref_viscous(double *flux, const double *p, const double *g, const double *property,
                 const double *basis)
{
    ref_init();
    navierstokes_viscous("central")(flux, p, g, property, mat_of(basis));
}

void ref_bc(const char *name, double *outer, const double *inner, const double *property,
            const double *basis, const double *ctx)
{
    ref_init();
    EulerPrimitive c = {0};
    if (ctx) {
        c.density = ctx[0];
        c.velocity = (Vector){ctx[1], ctx[2], ctx[3]};
        c.pressure = ctx[4];
    }
    navierstokes_boundary(name)(outer, inner, property, mat_of(basis), &c);
}

void ref_riemann(const double *l, const double *r, double gamma, double loc, double *out)
{
    ref_init();
    EulerPrimitive lp = {l[0], {l[1], l[2], l[3]}, l[4]};
    EulerPrimitive rp = {r[0], {r[1], r[2], r[3]}, r[4]};
    EulerPrimitive o = euler_riemann(lp, rp, gamma, loc);
    out[0] = o.density;
    out[1] = o.velocity.x;
    out[2] = o.velocity.y;
    out[3] = o.velocity.z;
    out[4] = o.pressure;
}

void ref_prim2cons(double *u, const double *p, const double *property) { euler_conserved(u, p, property); }
void ref_cons2prim(double *p, const double *u, const double *property) { euler_primitive(p, u, property); }
