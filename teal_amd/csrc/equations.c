// equations.c -- the equations half of the teal API (include/equations.h).
//
// Creation and setters follow src/equations/create.c:9-99 and src/equations/set.c:8-148.
// The residual operators (src/equations.h:163-178) run on the GPU: each call maps the
// plugin callbacks to device ids, keeps the device mirror (eqns->device) configured, and
// moves the caller's host arrays in and out.  There is no CPU fallback: a callback with no
// device implementation is an error reported through teal_error.
#include <assert.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "equations.h"
#include "plugins.h"
#include "teal.h"
#include "teal_gpu.h"

// ------------------------------------------------------------------ creation
Equations *equations_create(const Mesh *mesh, const char *name, Timestep *timestep,
                            ConvectiveSelect *convective, ViscousSelect *viscous,
                            BoundarySelect *boundary)
{
    assert(mesh && name && timestep);
    Equations *eqns = teal_calloc(1, sizeof(*eqns));
    eqns->mesh = mesh;
    strcpy(eqns->name, name);
    eqns->timestep.compute = timestep;
    eqns->convective.select = convective;
    eqns->viscous.select = viscous;
    eqns->boundary.select = boundary;
    int num = mesh->entities.off_boundary - mesh->entities.num_inner;
    eqns->boundary.num = num;
    eqns->boundary.name = teal_calloc(num, sizeof(String));
    eqns->boundary.context = teal_calloc(num, sizeof(*eqns->boundary.context));
    eqns->boundary.compute = teal_calloc(num, sizeof(*eqns->boundary.compute));
    eqns->boundary.condition = teal_calloc(num, sizeof(*eqns->boundary.condition));
    return eqns;
}

static void create_variables(EquationsVariables *var, const Equations *eqns, const int *dim,
                             int num)
{
    int stride = 0;
    for (int i = 0; i < num; i++) {
        assert(dim[i] > 0);
        stride += dim[i];
    }
    var->num = num;
    var->stride = stride;
    var->dimension = teal_calloc(num, sizeof(int));
    memcpy(var->dimension, dim, num * sizeof(int));
    var->data = teal_calloc(eqns->mesh->cells.num, sizeof(double) * stride);
}

static void copy_names(String **dst, const char **name, int num)
{
    *dst = teal_calloc(num, sizeof(String));
    for (int i = 0; i < num; i++) {
        strcpy((*dst)[i], name[i]);
    }
}

void equations_create_primitive(Equations *eqns, const char **name, const int *dimension,
                                Convert *convert, int num)
{
    assert(eqns && name && dimension && num > 0);
    copy_names(&eqns->primitive.name, name, num);
    create_variables(&eqns->primitive, eqns, dimension, num);
    eqns->primitive.func.convert = convert;
}

void equations_create_conserved(Equations *eqns, const char **name, const int *dimension,
                                Convert *convert, int num)
{
    assert(eqns && name && dimension && num > 0);
    copy_names(&eqns->conserved.name, name, num);
    create_variables(&eqns->conserved, eqns, dimension, num);
    eqns->conserved.func.convert = convert;
}

void equations_create_reference(Equations *eqns, Compute *compute)
{
    assert(eqns && compute);
    int num = eqns->primitive.num;
    eqns->reference.name = teal_calloc(num, sizeof(String));
    for (int i = 0; i < num; i++) {
        snprintf(eqns->reference.name[i], sizeof(String), "%s-ref", eqns->primitive.name[i]);
    }
    create_variables(&eqns->reference, eqns, eqns->primitive.dimension, num);
    eqns->reference.func.compute = compute;
}

void equations_create_properties(Equations *eqns, const char **name, const double *property,
                                 int num)
{
    assert(eqns && name && property && num > 0);
    eqns->properties.num = num;
    copy_names(&eqns->properties.name, name, num);
    eqns->properties.data = teal_calloc(num, sizeof(double));
    memcpy(eqns->properties.data, property, num * sizeof(double));
}

// ------------------------------------------------------------------ setters
void equations_set_space_order(Equations *eqns, int space_order)
{
    assert(eqns && 1 <= space_order && space_order <= 2);
    eqns->space_order = space_order;
}

void equations_set_convective_flux(Equations *eqns, const char *name)
{
    assert(eqns);
    if (!name) {
        eqns->convective.name[0] = 0;
        eqns->convective.flux = 0;
        return;
    }
    strcpy(eqns->convective.name, name);
    eqns->convective.flux = eqns->convective.select(name);
}

void equations_set_viscous_flux(Equations *eqns, const char *name)
{
    assert(eqns);
    if (!name) {
        eqns->viscous.name[0] = 0;
        eqns->viscous.flux = 0;
        return;
    }
    strcpy(eqns->viscous.name, name);
    eqns->viscous.flux = eqns->viscous.select(name);
}

// Venkatakrishnan parameter eps^2 = (K cbrt(V))^3 per inner cell (src/equations/limiter.c:8-21)
static double *venkatakrishnan_parameter(const Equations *eqns, double coefficient)
{
    int ni = eqns->mesh->cells.num_inner;
    double *eps2 = teal_calloc(ni, sizeof(double));
    for (int i = 0; i < ni; i++) {
        eps2[i] = cb(coefficient * cbrt(eqns->mesh->cells.volume[i]));
    }
    return eps2;
}

void equations_set_limiter(Equations *eqns, const char *name, double coefficient)
{
    assert(eqns);
    teal_free(eqns->limiter.parameter);
    eqns->limiter.parameter = 0;
    if (!name) {
        eqns->limiter.name[0] = 0;
        eqns->limiter.kind = NONE;
        return;
    }
    strcpy(eqns->limiter.name, name);
    if (!strcmp(name, "barth jespersen") || !strcmp(name, "minmod")) {
        eqns->limiter.kind = BARTH_JESPERSEN;
    }
    else if (!strcmp(name, "venkatakrishnan")) {
        eqns->limiter.kind = VENKATAKRISHNAN;
        eqns->limiter.parameter = venkatakrishnan_parameter(eqns, coefficient);
    }
    else {
        teal_error("invalid limiter (%s)", name);
    }
}

void equations_set_boundary_condition(Equations *eqns, const char *entity, const char *name,
                                      Compute *compute, const void *context)
{
    assert(eqns && entity);
    int num_inner = eqns->mesh->entities.num_inner;
    for (int i = 0; i < eqns->boundary.num; i++) {
        if (!strcmp(eqns->mesh->entities.name[num_inner + i], entity)) {
            if (!name || !strcmp(name, "custom")) {
                assert(compute);
                strcpy(eqns->boundary.name[i], "custom");
                eqns->boundary.compute[i] = compute;
                eqns->boundary.condition[i] = 0;
            }
            else {
                strcpy(eqns->boundary.name[i], name);
                eqns->boundary.compute[i] = 0;
                eqns->boundary.condition[i] = eqns->boundary.select(name);
            }
            eqns->boundary.context[i] = context;
            return;
        }
    }
    teal_error("invalid entity (%s)", entity);
}

void equations_set_initial_condition(Equations *eqns, const char *entity, Compute *compute,
                                     const void *initial)
{
    assert(eqns && entity);
    int num_inner = eqns->mesh->entities.num_inner;
    int *cell_off = eqns->mesh->entities.cell_off;
    int stride = eqns->primitive.stride;
    double *primitive = eqns->primitive.data;
    for (int i = 0; i < num_inner; i++) {
        if (!strcmp(eqns->mesh->entities.name[i], entity)) {
            for (int j = cell_off[i]; j < cell_off[i + 1]; j++) {
                double *p = &primitive[(size_t)j * stride];
                if (compute) {
                    compute(p, eqns->properties.data, eqns->mesh->cells.center[j], 0, 0);
                }
                else {
                    memcpy(p, initial, stride * sizeof(double));
                }
            }
            return;
        }
    }
    teal_error("invalid entity (%s)", entity);
}

void equations_set_property(Equations *eqns, const char *name, double property)
{
    assert(eqns && name);
    for (int i = 0; i < eqns->properties.num; i++) {
        if (!strcmp(eqns->properties.name[i], name)) {
            eqns->properties.data[i] = property;
            return;
        }
    }
    teal_error("invalid property (%s)", name);
}

void equations_set_source(Equations *eqns, Compute *compute, Prepare *prepare)
{
    assert(eqns && compute);
    eqns->source.compute = compute;
    eqns->source.prepare = prepare;
}

void equations_summary(const Equations *eqns)
{
    assert(eqns);
    int num_inner = eqns->mesh->entities.num_inner;
    teal_print("Equations summary");
    teal_print("\t name            : %s", eqns->name);
    teal_print("\t space order     : %d", eqns->space_order);
    teal_print("\t convective flux : %s", eqns->convective.name);
    teal_print("\t viscous flux    : %s", eqns->viscous.name);
    teal_print("\t limiter         : %s", eqns->limiter.name);
    teal_print("\t Boundary conditions");
    for (int i = 0; i < eqns->boundary.num; i++) {
        teal_print("\t\t %-20s : %s", eqns->mesh->entities.name[num_inner + i],
                   eqns->boundary.name[i]);
    }
    teal_print("\t Properties");
    for (int i = 0; i < eqns->properties.num; i++) {
        teal_print("\t\t %-20s : %g", eqns->properties.name[i], eqns->properties.data[i]);
    }
}

// ------------------------------------------------------------------ device mirror
typedef struct {
    TealGpu *gpu;
    int precision;
    TealGpuConfig cfg;  // last configuration handed to the device
    const double *eps2; // parameter pointer it was built from
    int *bc;
    double *ctx;
} Device;

static Device *device_state(const Equations *eqns)
{
    return (Device *)eqns->device;
}

void teal_equations_set_precision(Equations *eqns, int precision)
{
    Device *dev = device_state(eqns);
    if (dev && dev->gpu && dev->precision != precision) {
        teal_gpu_destroy(dev->gpu);
        dev->gpu = 0;
    }
    if (!dev) {
        dev = teal_calloc(1, sizeof(Device));
        eqns->device = dev;
    }
    dev->precision = precision;
}

// build the device configuration from the plugin callbacks; teal_error when a callback has
// no device counterpart
static void device_config(const Equations *eqns, TealGpuConfig *cfg, int *bc, double *ctx)
{
    memset(cfg, 0, sizeof(*cfg));
    if (eqns->primitive.stride != 5 || eqns->conserved.stride != 5 ||
        !teal_convert_is_euler(eqns->conserved.func.convert, eqns->primitive.func.convert)) {
        teal_error("equations \"%s\": the device path implements the Euler/Navier-Stokes "
                   "state (5 variables)", eqns->name);
    }
    int flux = teal_flux_id(eqns->convective.flux);
    if (flux < 0) {
        teal_error("convective flux \"%s\" has no device implementation", eqns->convective.name);
    }
    int visc = teal_viscous_id(eqns->viscous.flux);
    if (visc < 0) {
        teal_error("viscous flux \"%s\" has no device implementation", eqns->viscous.name);
    }
    int ts = teal_timestep_id(eqns->timestep.compute);
    if (ts < 0) {
        teal_error("time step callback of \"%s\" has no device implementation", eqns->name);
    }
    if (eqns->source.compute || eqns->source.prepare) {
        teal_error("source terms are host callbacks with no device implementation");
    }
    cfg->space_order = eqns->space_order;
    cfg->flux = flux;
    cfg->viscous = visc;
    cfg->limiter = eqns->limiter.kind;
    cfg->navierstokes = ts;
    cfg->gamma = eqns->properties.data[0];
    cfg->mu = (eqns->properties.num > 1) ? eqns->properties.data[1] : 0;
    cfg->prandtl = (eqns->properties.num > 2) ? eqns->properties.data[2] : 0;
    cfg->num_boundary = eqns->boundary.num;
    for (int i = 0; i < eqns->boundary.num; i++) {
        int entity = eqns->mesh->entities.num_inner + i;
        if (eqns->boundary.compute[i]) {
            teal_error("custom boundary condition on \"%s\" has no device implementation",
                       eqns->mesh->entities.name[entity]);
        }
        if (!eqns->boundary.condition[i]) {
            teal_error("no boundary condition set on \"%s\"", eqns->mesh->entities.name[entity]);
        }
        bc[i] = teal_bc_id(eqns->boundary.condition[i]);
        if (bc[i] < 0) {
            teal_error("boundary condition \"%s\" has no device implementation",
                       eqns->boundary.name[i]);
        }
        const double *c = eqns->boundary.context[i];
        for (int v = 0; v < 5; v++) {
            ctx[(i * 5) + v] = c ? c[v] : 0;
        }
    }
    cfg->bc = bc;
    cfg->bc_context = ctx;
    cfg->eps2 = eqns->limiter.parameter;
}

// the configured device mirror of eqns (created on first use)
TealGpu *teal_equations_device(const Equations *eqns)
{
    Equations *mut = (Equations *)eqns;
    Device *dev = device_state(eqns);
    if (!dev) {
        dev = teal_calloc(1, sizeof(Device));
        dev->precision = TEAL_FP64;
        mut->device = dev;
    }
    if (!dev->gpu) {
        dev->gpu = teal_gpu_create(eqns->mesh, dev->precision);
        if (!dev->gpu) {
            teal_error("%s", teal_gpu_last_error());
        }
        memset(&dev->cfg, 0xff, sizeof(dev->cfg));
    }
    int nb = eqns->boundary.num;
    if (!dev->bc) {
        dev->bc = teal_calloc(nb + 1, sizeof(int));
        dev->ctx = teal_calloc(5 * (nb + 1), sizeof(double));
    }
    TealGpuConfig cfg;
    int *bc = teal_calloc(nb + 1, sizeof(int));
    double *ctx = teal_calloc(5 * (nb + 1), sizeof(double));
    device_config(eqns, &cfg, bc, ctx);
    int same = dev->cfg.space_order == cfg.space_order && dev->cfg.flux == cfg.flux &&
               dev->cfg.viscous == cfg.viscous && dev->cfg.limiter == cfg.limiter &&
               dev->cfg.navierstokes == cfg.navierstokes && dev->cfg.gamma == cfg.gamma &&
               dev->cfg.mu == cfg.mu && dev->cfg.prandtl == cfg.prandtl &&
               dev->eps2 == cfg.eps2 && !memcmp(dev->bc, bc, nb * sizeof(int)) &&
               !memcmp(dev->ctx, ctx, 5 * nb * sizeof(double));
    if (!same) {
        if (teal_gpu_configure(dev->gpu, &cfg)) {
            teal_error("%s", teal_gpu_last_error());
        }
        dev->cfg = cfg;
        dev->eps2 = cfg.eps2;
        memcpy(dev->bc, bc, nb * sizeof(int));
        memcpy(dev->ctx, ctx, 5 * nb * sizeof(double));
    }
    teal_free(bc);
    teal_free(ctx);
    return dev->gpu;
}

static void check(int rc)
{
    if (rc) {
        teal_error("%s", teal_gpu_last_error());
    }
}

// ------------------------------------------------------------------ residual path
double equations_timestep(const Equations *eqns, const void *primitive)
{
    assert(eqns && primitive);
    double dt = teal_gpu_timestep(teal_equations_device(eqns), primitive);
    if (isnan(dt)) {
        teal_error("%s", teal_gpu_last_error());
    }
    return dt;
}

void equations_boundary(const Equations *eqns, void *primitive, double time)
{
    assert(eqns && primitive);
    (void)time;
    check(teal_gpu_boundary(teal_equations_device(eqns), primitive));
}

void equations_gradient(const Equations *eqns, void *primitive, void *gradient)
{
    assert(eqns && primitive && gradient);
    check(teal_gpu_gradient(teal_equations_device(eqns), primitive, gradient));
}

void equations_limiter(const Equations *eqns, const void *primitive, void *gradient)
{
    assert(eqns && primitive && gradient);
    if (eqns->limiter.kind == NONE) {
        teal_error("invalid limiter kind (%d)", eqns->limiter.kind);
    }
    check(teal_gpu_limiter(teal_equations_device(eqns), primitive, gradient));
}

void equations_derivative(const Equations *eqns, void *primitive, void *derivative, double time)
{
    assert(eqns && primitive && derivative);
    (void)time;
    check(teal_gpu_derivative(teal_equations_device(eqns), primitive, derivative));
}

void equations_residual(const Equations *eqns, const void *derivative, void *residual)
{
    assert(eqns && derivative && residual);
    check(teal_gpu_residual(teal_equations_device(eqns), derivative, residual));
}

// ------------------------------------------------------------------ diagnostics (host)
// volume-weighted entity average of a host array (src/equations/average.c:8-41)
void equations_average(const Equations *eqns, const char *entity, const void *primitive_,
                       void *average_)
{
    assert(eqns && entity);
    const int stride = eqns->primitive.stride;
    const double *primitive = primitive_;
    double *average = average_;
    memset(average, 0, stride * sizeof(double));
    for (int i = 0; i < eqns->mesh->entities.num_inner; i++) {
        if (!strcmp(eqns->mesh->entities.name[i], entity)) {
            for (int j = eqns->mesh->entities.cell_off[i]; j < eqns->mesh->entities.cell_off[i + 1];
                 j++) {
                for (int k = 0; k < stride; k++) {
                    average[k] += eqns->mesh->cells.volume[j] * primitive[((size_t)j * stride) + k];
                }
            }
            for (int k = 0; k < stride; k++) {
                average[k] /= eqns->mesh->cells.sum_volume;
            }
            return;
        }
    }
    teal_error("invalid entity (%s)", entity);
}

// L2 error against the reference solution callback (src/equations/norm.c:8-38)
void equations_norm(const Equations *eqns, void *norm_, double time)
{
    assert(eqns && norm_);
    const int stride = eqns->primitive.stride;
    const double *primitive = eqns->primitive.data;
    double *reference = eqns->reference.data;
    double *norm = norm_;
    memset(norm, 0, stride * sizeof(double));
    if (!eqns->reference.func.compute) {
        teal_error("no reference solution registered");
    }
    for (int i = 0; i < eqns->mesh->cells.num_inner; i++) {
        double *ref = &reference[(size_t)i * stride];
        const double *p = &primitive[(size_t)i * stride];
        eqns->reference.func.compute(ref, eqns->properties.data, eqns->mesh->cells.center[i], time,
                                     p);
        for (int j = 0; j < stride; j++) {
            norm[j] += eqns->mesh->cells.volume[i] * sq(ref[j] - p[j]);
        }
    }
    for (int j = 0; j < stride; j++) {
        norm[j] = sqrt(norm[j] / eqns->mesh->cells.sum_volume);
    }
}

void equations_write(const Equations *eqns, const char *name, double time, int index)
{
    (void)eqns, (void)time;
    static int warned = 0;
    if (!warned) {
        teal_print("\t (VTKHDF output is not provided by this library: %s_%05d skipped)", name, index);
        warned = 1;
    }
}

void equations_read(const Equations *eqns, const char *fname, double *time, int *index)
{
    (void)eqns, (void)time, (void)index;
    teal_error("restart input is not provided by this library (%s)", fname);
}

void equations_destroy(Equations *eqns)
{
    assert(eqns);
    Device *dev = device_state(eqns);
    if (dev) {
        if (dev->gpu) {
            teal_gpu_destroy(dev->gpu);
        }
        teal_free(dev->bc);
        teal_free(dev->ctx);
        teal_free(dev);
    }
    teal_free(eqns->boundary.name);
    teal_free(eqns->boundary.context);
    teal_free(eqns->boundary.compute);
    teal_free(eqns->boundary.condition);
    EquationsVariables *vars[] = {&eqns->primitive, &eqns->conserved, &eqns->reference};
    for (int i = 0; i < 3; i++) {
        teal_free(vars[i]->name);
        teal_free(vars[i]->dimension);
        teal_free(vars[i]->data);
    }
    teal_free(eqns->properties.name);
    teal_free(eqns->properties.data);
    teal_free(eqns->limiter.parameter);
    teal_free(eqns);
}
