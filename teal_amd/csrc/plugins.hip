// plugins.hip -- host side of the Euler / Navier-Stokes plugins (src/euler/*.c,
// src/navierstokes/*.c): the select-by-name functions return host callbacks with the
// reference signatures, built from the same templates (physics.h) the device kernels use.
// Each callback also has a device id (teal_*_id) so the GPU path can dispatch to the
// matching compile-time kernel.  Compiled with -ffp-contract=off (bit-compatible FP64).
#include <cmath>
#include <cstring>

#include "euler.h"
#include "navierstokes.h"
#include "physics.h"
#include "plugins.h"
#include "simulation.h"
#include "teal.h"
#include "teal_gpu.h"

using namespace fv;

static M3<double> mat(Matrix b) { return m3<double>(&b.x.x); }

template <int FLUX>
static void conv(void *flux, const void *left, const void *right, const double *property,
                 Matrix basis)
{
    int err = 0;
    convective<FLUX, double>((double *)flux, (const double *)left, (const double *)right,
                             property[EULER_HEAT_CAPACITY_RATIO], mat(basis), &err);
    if (err == 1) {
        teal_error("vacuum state detected in riemann solver");
    }
    if (err == 2) {
        teal_error("newton solver did not converge (%d)", 20);
    }
}

static const struct {
    const char *name;
    Convective *fn;
    int id;
} flux_table[] = {
    {"godunov", conv<FLUX_GODUNOV>, FLUX_GODUNOV}, {"roe", conv<FLUX_ROE>, FLUX_ROE},
    {"hll", conv<FLUX_HLL>, FLUX_HLL},             {"hllc", conv<FLUX_HLLC>, FLUX_HLLC},
    {"hlle", conv<FLUX_HLLE>, FLUX_HLLE},          {"ausmd", conv<FLUX_AUSMD>, FLUX_AUSMD},
    {"ausmdv", conv<FLUX_AUSMDV>, FLUX_AUSMDV},    {"lxf", conv<FLUX_LXF>, FLUX_LXF},
    {"central", conv<FLUX_CENTRAL>, FLUX_CENTRAL},
};

// src/euler/convective.c:593-623
extern "C" Convective *euler_convective(const char *name)
{
    for (const auto &e : flux_table) {
        if (!strcmp(name, e.name)) {
            return e.fn;
        }
    }
    teal_error("invalid convective flux (%s)", name);
}

extern "C" int teal_flux_id(Convective *fn)
{
    if (!fn) {
        return FLUX_NONE;
    }
    for (const auto &e : flux_table) {
        if (e.fn == fn) {
            return e.id;
        }
    }
    return -1;
}

// src/navierstokes/viscous.c:11-47
static void viscous(void *flux, const void *primitive, const void *gradient,
                    const double *property, Matrix basis)
{
    viscous_central<double>((double *)flux, (const double *)primitive,
                            (const V3<double> *)gradient, property[NAVIERSTOKES_HEAT_CAPACITY_RATIO],
                            property[NAVIERSTOKES_DYNAMIC_VISCOSITY], property[NAVIERSTOKES_PRANDTL],
                            v3(basis.x.x, basis.x.y, basis.x.z));
}

extern "C" Viscous *navierstokes_viscous(const char *name)
{
    if (!strcmp(name, "central")) {
        return viscous;
    }
    teal_error("invalid viscous flux (%s)", name);
}

extern "C" int teal_viscous_id(Viscous *fn) { return !fn ? 0 : (fn == viscous ? 1 : -1); }

template <int BC>
static void bc(void *outer, const void *inner, const double *property, Matrix basis,
               const void *context)
{
    static const double zero[5] = {0, 0, 0, 0, 0};
    const double *ctx = context ? (const double *)context : zero;
    boundary<double>(BC, (double *)outer, (const double *)inner, property[0], mat(basis), ctx);
}

static const struct {
    const char *name;
    Boundary *fn;
    int id;
} bc_table[] = {
    {"symmetry", bc<BC_SYMMETRY>, BC_SYMMETRY},
    {"slipwall", bc<BC_SYMMETRY>, BC_SYMMETRY},
    {"supersonic inflow", bc<BC_SUPERSONIC_INFLOW>, BC_SUPERSONIC_INFLOW},
    {"supersonic outflow", bc<BC_SUPERSONIC_OUTFLOW>, BC_SUPERSONIC_OUTFLOW},
    {"subsonic inflow", bc<BC_SUBSONIC_INFLOW>, BC_SUBSONIC_INFLOW},
    {"subsonic outflow", bc<BC_SUBSONIC_OUTFLOW>, BC_SUBSONIC_OUTFLOW},
    {"pressure outflow", bc<BC_PRESSURE_OUTFLOW>, BC_PRESSURE_OUTFLOW},
    {"farfield", bc<BC_FARFIELD>, BC_FARFIELD},
    {"adiabatic wall", bc<BC_ADIABATIC_WALL>, BC_ADIABATIC_WALL},
    {"wall", bc<BC_ADIABATIC_WALL>, BC_ADIABATIC_WALL},
};

// src/euler/boundary.c:273-297 (the wall conditions belong to navierstokes_boundary)
extern "C" Boundary *euler_boundary(const char *name)
{
    for (const auto &e : bc_table) {
        if (e.id != BC_ADIABATIC_WALL && !strcmp(name, e.name)) {
            return e.fn;
        }
    }
    teal_error("invalid boundary condition (%s)", name);
}

// src/navierstokes/boundary.c:20-26
extern "C" Boundary *navierstokes_boundary(const char *name)
{
    if (!strcmp(name, "adiabatic wall") || !strcmp(name, "wall")) {
        return bc<BC_ADIABATIC_WALL>;
    }
    return euler_boundary(name);
}

extern "C" int teal_bc_id(Boundary *fn)
{
    for (const auto &e : bc_table) {
        if (e.fn == fn) {
            return e.id;
        }
    }
    return -1;
}

// src/euler/create.c:6-44
extern "C" double euler_timestep(const void *primitive, const double *property, double volume,
                                 Vector projection)
{
    return euler_dt<double>((const double *)primitive, property[0], volume,
                            v3(projection.x, projection.y, projection.z));
}

extern "C" void euler_primitive(void *primitive, const void *conserved, const double *property)
{
    cons2prim<double>((double *)primitive, (const double *)conserved, property[0]);
}

extern "C" void euler_conserved(void *conserved, const void *primitive, const double *property)
{
    prim2cons<double>((double *)conserved, (const double *)primitive, property[0]);
}

// src/navierstokes/create.c:7-26
extern "C" double navierstokes_timestep(const void *primitive, const double *property,
                                        double volume, Vector projection)
{
    return ns_dt<double>((const double *)primitive, property[0], property[1], property[2], volume,
                         v3(projection.x, projection.y, projection.z));
}

extern "C" int teal_timestep_id(Timestep *fn)
{
    if (fn == euler_timestep) return 0;
    if (fn == navierstokes_timestep) return 1;
    return -1;
}

extern "C" int teal_convert_is_euler(Convert *prim2cons, Convert *cons2prim)
{
    return prim2cons == euler_conserved && cons2prim == euler_primitive;
}

// src/euler/riemann.c:162-197
extern "C" EulerPrimitive euler_riemann(EulerPrimitive left, EulerPrimitive right, double gamma,
                                        double loc)
{
    double l[5] = {left.density, left.velocity.x, left.velocity.y, left.velocity.z, left.pressure};
    double r[5] = {right.density, right.velocity.x, right.velocity.y, right.velocity.z,
                   right.pressure};
    double o[5];
    int err = 0;
    riemann<double>(l, r, gamma, loc, o, &err);
    if (err == 1) {
        teal_error("vacuum state detected in riemann solver");
    }
    if (err == 2) {
        teal_error("newton solver did not converge (%d)", 20);
    }
    EulerPrimitive out;
    out.density = o[0];
    out.velocity.x = o[1];
    out.velocity.y = o[2];
    out.velocity.z = o[3];
    out.pressure = o[4];
    return out;
}

// src/euler/create.c:46-67
extern "C" Equations *euler_create(const Mesh *mesh)
{
    Equations *eqns =
        equations_create(mesh, "euler", euler_timestep, euler_convective, 0, euler_boundary);
    const char *primitive[] = {"density", "velocity", "pressure"};
    int dims[] = {1, 3, 1};
    equations_create_primitive(eqns, primitive, dims, euler_primitive, 3);
    const char *conserved[] = {"density", "momentum", "energy"};
    equations_create_conserved(eqns, conserved, dims, euler_conserved, 3);
    const char *property[] = {"heat capacity ratio"};
    double values[] = {1.4};
    equations_create_properties(eqns, property, values, 1);
    equations_set_space_order(eqns, 2);
    equations_set_convective_flux(eqns, "hllc");
    equations_set_limiter(eqns, "minmod", 0);
    return eqns;
}

// src/navierstokes/create.c:28-51
extern "C" Equations *navierstokes_create(const Mesh *mesh)
{
    Equations *eqns = equations_create(mesh, "navier-stokes", navierstokes_timestep,
                                       euler_convective, navierstokes_viscous,
                                       navierstokes_boundary);
    const char *primitive[] = {"density", "velocity", "pressure"};
    int dims[] = {1, 3, 1};
    equations_create_primitive(eqns, primitive, dims, euler_primitive, 3);
    const char *conserved[] = {"_density", "momentum", "energy"};
    equations_create_conserved(eqns, conserved, dims, euler_conserved, 3);
    const char *property[] = {"heat capacity ratio", "dynamic viscosity", "prandtl number"};
    double values[] = {1.4, 0, 0.71};
    equations_create_properties(eqns, property, values, 3);
    equations_set_space_order(eqns, 2);
    equations_set_convective_flux(eqns, "hllc");
    equations_set_viscous_flux(eqns, "central");
    equations_set_limiter(eqns, "minmod", 0);
    return eqns;
}

// src/euler/polar.c:22-83: the boundary pass and the face sums on the device (teal_gpu_polar),
// lift and drag from the force coefficients, and with an output prefix the per-face pressure
// coefficients in <prefix>_polar_<alpha in degrees>.h5
extern "C" void euler_polar(const Simulation *sim, const char *entity, EulerPrimitive reference,
                            double length, double time)
{
    (void)time;  // the device boundary conditions do not depend on time
    const Equations *eqns = sim->eqns;
    const Mesh *mesh = eqns->mesh;
    int idx = -1;
    for (int i = 0; i < mesh->entities.off_boundary && idx < 0; i++) {
        if (!strcmp(mesh->entities.name[i], entity)) {
            idx = i;
        }
    }
    if (idx < 0) {
        teal_error("invalid entity (%s)", entity);
    }
    const int f0 = mesh->entities.face_off[idx];
    const int num = mesh->entities.face_off[idx + 1] - f0;
    const double ref[5] = {reference.density, reference.velocity.x, reference.velocity.y,
                           reference.velocity.z, reference.pressure};
    double coef[2];
    double *pressure_c = (double *)teal_calloc(num + 1, sizeof(double));
    if (teal_gpu_polar(teal_equations_device(eqns), idx, (double *)eqns->primitive.data, ref,
                       length, coef, pressure_c)) {
        teal_error("%s", teal_gpu_last_error());
    }
    const double alpha = atan2(reference.velocity.y, reference.velocity.x);
    const double lift = (cos(alpha) * coef[1]) - (sin(alpha) * coef[0]);
    const double drag = (cos(alpha) * coef[0]) + (sin(alpha) * coef[1]);
    teal_print("Airfoil polar");
    teal_print("\t lift coefficient : %g", lift);
    teal_print("\t drag coefficient : %g", drag);
    if (sim->prefix[0]) {
        const double alpha_deg = alpha * 180 / acos(-1.0);
        char fname[512];
        snprintf(fname, sizeof(fname), "%s_polar_%g.h5", sim->prefix, alpha_deg);
        teal_io_write_polar(fname, alpha_deg, &mesh->faces.center[f0], pressure_c, num, lift, drag);
    }
    teal_free(pressure_c);
}
