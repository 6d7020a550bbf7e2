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

// ------------------------------------------------------------------ helpers
// index of `name` among names[first, last), or -1
static int find_name(const String *names, int first, int last, const char *name)
{
    for (int i = first; i < last; i++) {
        if (strcmp(names[i], name) == 0) {
            return i;
        }
    }
    return -1;
}

// a String array of `num` names, each formatted from name[i] with `fmt` ("%s", "%s-ref", ...)
static String *name_list(const char *const *name, int num, const char *fmt)
{
    String *out = teal_calloc(num, sizeof(String));
    for (int i = 0; i < num; i++) {
        snprintf(out[i], sizeof(String), fmt, name[i]);
    }
    return out;
}

// a variable set: names, per-variable dimensions, their sum (the stride) and one zeroed row of
// `stride` doubles per mesh cell (src/equations/create.c:33-50: ghost rows included)
static void variable_set(EquationsVariables *var, String *names, const int *dimension, int num,
                         const Mesh *mesh)
{
    var->num = num;
    var->name = names;
    var->dimension = teal_calloc(num, sizeof(int));
    var->stride = 0;
    for (int i = 0; i < num; i++) {
        assert(dimension[i] > 0);
        var->dimension[i] = dimension[i];
        var->stride += dimension[i];
    }
    var->data = teal_calloc((size_t)mesh->cells.num * var->stride, sizeof(double));
}

static const char *entity_name(const Equations *eqns, int entity)
{
    return eqns->mesh->entities.name[entity];
}

// ------------------------------------------------------------------ creation
Equations *equations_create(const Mesh *mesh, const char *name, Timestep *timestep,
                            ConvectiveSelect *convective, ViscousSelect *viscous,
                            BoundarySelect *boundary)
{
    assert(mesh && name && timestep);
    Equations *eqns = teal_calloc(1, sizeof(Equations));
    *eqns = (Equations){.mesh = mesh};
    snprintf(eqns->name, sizeof(String), "%s", name);
    eqns->timestep.compute = timestep;
    eqns->convective.select = convective;
    eqns->viscous.select = viscous;
    // one slot per boundary entity (the entities between num_inner and off_boundary)
    EquationsBoundary *bc = &eqns->boundary;
    bc->select = boundary;
    bc->num = mesh->entities.off_boundary - mesh->entities.num_inner;
    bc->name = teal_calloc(bc->num, sizeof(String));
    bc->compute = teal_calloc(bc->num, sizeof(Compute *));
    bc->condition = teal_calloc(bc->num, sizeof(Boundary *));
    bc->context = teal_calloc(bc->num, sizeof(void *));
    return eqns;
}

void equations_create_primitive(Equations *eqns, const char **name, const int *dimension,
                                Convert *convert, int num)
{
    assert(eqns && name && dimension && num > 0);
    variable_set(&eqns->primitive, name_list(name, num, "%s"), dimension, num, eqns->mesh);
    eqns->primitive.func.convert = convert;
}

void equations_create_conserved(Equations *eqns, const char **name, const int *dimension,
                                Convert *convert, int num)
{
    assert(eqns && name && dimension && num > 0);
    variable_set(&eqns->conserved, name_list(name, num, "%s"), dimension, num, eqns->mesh);
    eqns->conserved.func.convert = convert;
}

// the reference solution mirrors the primitive variables, named "<primitive>-ref"
void equations_create_reference(Equations *eqns, Compute *compute)
{
    assert(eqns && compute);
    const EquationsVariables *prim = &eqns->primitive;
    const char *names[64];
    assert(prim->num <= 64);
    for (int i = 0; i < prim->num; i++) {
        names[i] = prim->name[i];
    }
    variable_set(&eqns->reference, name_list(names, prim->num, "%s-ref"), prim->dimension,
                 prim->num, eqns->mesh);
    eqns->reference.func.compute = compute;
}

void equations_create_properties(Equations *eqns, const char **name, const double *property,
                                 int num)
{
    assert(eqns && name && property && num > 0);
    EquationsProperties *prop = &eqns->properties;
    prop->num = num;
    prop->name = name_list(name, num, "%s");
    prop->data = teal_calloc(num, sizeof(double));
    memcpy(prop->data, property, sizeof(double) * num);
}

// ------------------------------------------------------------------ setters
void equations_set_space_order(Equations *eqns, int space_order)
{
    assert(eqns && (space_order == 1 || space_order == 2));
    eqns->space_order = space_order;
}

// a plugin chosen by name: the name is kept for the summary; null clears both
void equations_set_convective_flux(Equations *eqns, const char *name)
{
    assert(eqns);
    snprintf(eqns->convective.name, sizeof(String), "%s", name ? name : "");
    eqns->convective.flux = name ? eqns->convective.select(name) : 0;
}

void equations_set_viscous_flux(Equations *eqns, const char *name)
{
    assert(eqns);
    snprintf(eqns->viscous.name, sizeof(String), "%s", name ? name : "");
    eqns->viscous.flux = name ? eqns->viscous.select(name) : 0;
}

// limiter names (src/equations/set.c:52-68): minmod is Barth-Jespersen
static const struct {
    const char *name;
    int kind;
} limiter_names[] = {
    {"barth jespersen", BARTH_JESPERSEN},
    {"minmod", BARTH_JESPERSEN},
    {"venkatakrishnan", VENKATAKRISHNAN},
};

void equations_set_limiter(Equations *eqns, const char *name, double coefficient)
{
    assert(eqns);
    EquationsLimiter *lim = &eqns->limiter;
    teal_free(lim->parameter);
    lim->parameter = 0;
    lim->kind = NONE;
    snprintf(lim->name, sizeof(String), "%s", name ? name : "");
    if (!name) {
        return;
    }
    for (size_t i = 0; i < sizeof(limiter_names) / sizeof(limiter_names[0]); i++) {
        if (strcmp(limiter_names[i].name, name) == 0) {
            lim->kind = limiter_names[i].kind;
        }
    }
    if (lim->kind == NONE) {
        teal_error("invalid limiter (%s)", name);
    }
    if (lim->kind == VENKATAKRISHNAN) {
        // eps^2 = (K cbrt(V))^3 per inner cell (src/equations/limiter.c:8-21)
        const int ni = eqns->mesh->cells.num_inner;
        lim->parameter = teal_calloc(ni, sizeof(double));
        for (int i = 0; i < ni; i++) {
            lim->parameter[i] = cb(coefficient * cbrt(eqns->mesh->cells.volume[i]));
        }
    }
}

// A named condition from the equations' table, or a user callback ("custom" or no name); the
// context is borrowed (it must outlive the equations).
void equations_set_boundary_condition(Equations *eqns, const char *entity, const char *name,
                                      Compute *compute, const void *context)
{
    assert(eqns && entity);
    const Mesh *mesh = eqns->mesh;
    int e = find_name(mesh->entities.name, mesh->entities.num_inner, mesh->entities.off_boundary,
                      entity);
    if (e < 0) {
        teal_error("invalid entity (%s)", entity);
    }
    const int slot = e - mesh->entities.num_inner;
    const int custom = !name || strcmp(name, "custom") == 0;
    assert(!custom || compute);
    EquationsBoundary *bc = &eqns->boundary;
    snprintf(bc->name[slot], sizeof(String), "%s", custom ? "custom" : name);
    bc->compute[slot] = custom ? compute : 0;
    bc->condition[slot] = custom ? 0 : bc->select(name);
    bc->context[slot] = context;
}

// every cell of an inner entity from a callback of its centre, or a copy of one state
void equations_set_initial_condition(Equations *eqns, const char *entity, Compute *compute,
                                     const void *initial)
{
    assert(eqns && entity);
    const Mesh *mesh = eqns->mesh;
    int e = find_name(mesh->entities.name, 0, mesh->entities.num_inner, entity);
    if (e < 0) {
        teal_error("invalid entity (%s)", entity);
    }
    const int stride = eqns->primitive.stride;
    double *rows = eqns->primitive.data;
    for (int c = mesh->entities.cell_off[e]; c < mesh->entities.cell_off[e + 1]; c++) {
        double *row = &rows[(size_t)c * stride];
        if (compute) {
            compute(row, eqns->properties.data, mesh->cells.center[c], 0, 0);
        }
        else {
            memcpy(row, initial, sizeof(double) * stride);
        }
    }
}

void equations_set_property(Equations *eqns, const char *name, double property)
{
    assert(eqns && name);
    int i = find_name(eqns->properties.name, 0, eqns->properties.num, name);
    if (i < 0) {
        teal_error("invalid property (%s)", name);
    }
    eqns->properties.data[i] = property;
}

void equations_set_source(Equations *eqns, Compute *compute, Prepare *prepare)
{
    assert(eqns && compute);
    eqns->source = (EquationsSource){.compute = compute, .prepare = prepare};
}

void equations_summary(const Equations *eqns)
{
    assert(eqns);
    const char *label[] = {"name", "space order", "convective flux", "viscous flux", "limiter"};
    char order[16];
    snprintf(order, sizeof(order), "%d", eqns->space_order);
    const char *value[] = {eqns->name, order, eqns->convective.name, eqns->viscous.name,
                           eqns->limiter.name};
    teal_print("Equations summary");
    for (int i = 0; i < 5; i++) {
        teal_print("\t %-15s : %s", label[i], value[i]);
    }
    teal_print("\t Boundary conditions");
    for (int i = 0; i < eqns->boundary.num; i++) {
        teal_print("\t\t %-20s : %s", entity_name(eqns, eqns->mesh->entities.num_inner + i),
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
    // bulk-velocity forcing (teal_equations_set_forcing): entity, {U, tau}, and what the
    // device handle currently has (-1: nothing applied yet)
    int forcing_entity;
    double forcing[2];
    int applied_source;
    double applied_forcing[2];
} Device;

static Device *device_state(const Equations *eqns)
{
    Equations *mut = (Equations *)eqns;
    if (!mut->device) {
        Device *dev = teal_calloc(1, sizeof(Device));
        dev->precision = TEAL_FP64;
        dev->applied_source = -1;
        mut->device = dev;
    }
    return (Device *)mut->device;
}

static void device_drop(Device *dev)
{
    if (dev->gpu) {
        teal_gpu_destroy(dev->gpu);
        dev->gpu = 0;
    }
    dev->applied_source = -1;
}

void teal_equations_set_precision(Equations *eqns, int precision)
{
    Device *dev = device_state(eqns);
    if (dev->precision != precision) {
        device_drop(dev);
    }
    dev->precision = precision;
}

// ------------------------------------------------------------------ forcing source
// The bulk-velocity forcing of run/periodic_hill/run.c:8-21 as library callbacks: the device
// path recognises this pair and runs it as TEAL_SOURCE_FORCING.  The host bodies follow the
// reference driver (its Prepare keeps the average in a file-scope record).
static struct {
    double velocity, relaxation, average;
} forcing_host = {0, 1, 0};

void teal_forcing_prepare(const Equations *eqns, const void *primitive)
{
    Device *dev = device_state(eqns);
    double avg[5];
    equations_average(eqns, entity_name(eqns, dev->forcing_entity), primitive, avg);
    forcing_host.velocity = dev->forcing[0];
    forcing_host.relaxation = dev->forcing[1];
    forcing_host.average = avg[1];
}

void teal_forcing_source(void *source, const double *property, Vector center, double time,
                         const void *context)
{
    (void)property, (void)center, (void)time;
    const double *p = context;  // the cell's primitive state {rho, u, v, w, p}
    double *s = source;
    const double s1 = (p[0] * (forcing_host.velocity - forcing_host.average)) /
                      forcing_host.relaxation;
    s[0] = s[2] = s[3] = 0;
    s[1] = s1;
    s[4] = p[1] * s1;
}

void teal_equations_set_forcing(Equations *eqns, const char *entity, double velocity,
                                double relaxation)
{
    assert(eqns && entity && relaxation != 0);
    int e = find_name(eqns->mesh->entities.name, 0, eqns->mesh->entities.num_inner, entity);
    if (e < 0) {
        teal_error("invalid entity (%s)", entity);
    }
    Device *dev = device_state(eqns);
    dev->forcing_entity = e;
    dev->forcing[0] = velocity;
    dev->forcing[1] = relaxation;
    equations_set_source(eqns, teal_forcing_source, teal_forcing_prepare);
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
    if ((eqns->source.compute || eqns->source.prepare) &&
        (eqns->source.compute != teal_forcing_source || eqns->source.prepare != teal_forcing_prepare)) {
        teal_error("source term callbacks have no device implementation (the device source is "
                   "teal_equations_set_forcing)");
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
    Device *dev = device_state(eqns);
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
    const int source = eqns->source.compute ? TEAL_SOURCE_FORCING : TEAL_SOURCE_NONE;
    if (source != dev->applied_source || memcmp(dev->forcing, dev->applied_forcing, sizeof(dev->forcing))) {
        if (teal_gpu_set_source(dev->gpu, source, dev->forcing_entity, dev->forcing)) {
            teal_error("%s", teal_gpu_last_error());
        }
        dev->applied_source = source;
        memcpy(dev->applied_forcing, dev->forcing, sizeof(dev->forcing));
    }
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
    if (eqns->boundary.num == 0) {  // no boundary entities: no ghost states to compute
        return;
    }
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

// ------------------------------------------------------------------ diagnostics
// volume-weighted entity average of a host array (src/equations/average.c:8-41), on the device
void equations_average(const Equations *eqns, const char *entity, const void *primitive,
                       void *average)
{
    assert(eqns && entity && average);
    int e = find_name(eqns->mesh->entities.name, 0, eqns->mesh->entities.num_inner, entity);
    if (e < 0) {
        teal_error("invalid entity (%s)", entity);
    }
    check(teal_gpu_average(teal_equations_device(eqns), e, primitive, average));
}

// L2 error against the reference solution (src/equations/norm.c:8-38): the reference callback
// fills eqns->reference on the host (it is user code), the sum runs on the device
void equations_norm(const Equations *eqns, void *norm, double time)
{
    assert(eqns && norm);
    Compute *exact = eqns->reference.func.compute;
    if (!exact) {
        teal_error("no reference solution registered");
    }
    const int stride = eqns->primitive.stride;
    const double *prim = eqns->primitive.data;
    double *ref = eqns->reference.data;
    for (int i = 0; i < eqns->mesh->cells.num_inner; i++) {
        exact(&ref[(size_t)i * stride], eqns->properties.data, eqns->mesh->cells.center[i], time,
              &prim[(size_t)i * stride]);
    }
    check(teal_gpu_norm(teal_equations_device(eqns), prim, ref, norm));
}

void equations_destroy(Equations *eqns)
{
    assert(eqns);
    if (eqns->device) {
        Device *dev = device_state(eqns);
        device_drop(dev);
        teal_free(dev->bc);
        teal_free(dev->ctx);
        teal_free(dev);
    }
    EquationsVariables *sets[] = {&eqns->primitive, &eqns->conserved, &eqns->reference};
    void *owned[] = {eqns->boundary.name, eqns->boundary.context, eqns->boundary.compute,
                     eqns->boundary.condition, eqns->properties.name, eqns->properties.data,
                     eqns->limiter.parameter};
    for (size_t i = 0; i < sizeof(owned) / sizeof(owned[0]); i++) {
        teal_free(owned[i]);
    }
    for (int i = 0; i < 3; i++) {
        teal_free(sets[i]->name);
        teal_free(sets[i]->dimension);
        teal_free(sets[i]->data);
    }
    teal_free(eqns);
}
