// simulation.c -- the simulation half of the teal API (include/simulation.h).
//
// simulation_create / set_* / summary / error follow src/simulation/create.c:9-26,
// set.c:8-131, summary.c:9-48 and error.c:8-35.  The time loop (src/simulation/run.c:22-113)
// keeps the state on the GPU: the primitive variables are uploaded once, every step runs
// as device kernels, and only the step size and the 5 residual norms come back per step;
// the host copy is refreshed at output points and at the end.
//
// The public Advance methods euler / lserk / implicit_euler (src/simulation/advance.c) can
// also be called directly on eqns->primitive.data: they upload it, advance one step on the
// device and download the result.
#define _POSIX_C_SOURCE 200809L
#include <assert.h>
#include <float.h>
#include <limits.h>
#include <math.h>
#include <signal.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include "simulation.h"
#include "teal.h"
#include "teal_gpu.h"

TealGpu *teal_equations_device(const Equations *eqns);

static void check(int rc)
{
    if (rc) {
        teal_error("%s", teal_gpu_last_error());
    }
}

static double checked(double step0)
{
    if (isnan(step0)) {
        teal_error("%s", teal_gpu_last_error());
    }
    return step0;
}

// one step on the resident device state
static double advance_device(const Equations *eqns, Advance *method, double *time,
                             double *residual, double max_step, double courant,
                             const void *context)
{
    TealGpu *gpu = teal_equations_device(eqns);
    if (method == lserk) {
        const RungeKutta *rk = context;
        return checked(teal_gpu_lserk(gpu, time, residual, max_step, courant, rk->time_order,
                                      rk->num_stages));
    }
    if (method == euler) {
        return checked(teal_gpu_euler(gpu, time, residual, max_step, courant));
    }
    if (method == implicit_euler) {
        const NewtonKrylov *nk = context;
        return checked(teal_gpu_implicit_euler(gpu, time, residual, max_step, courant,
                                               nk->newton_tolerance, nk->krylov_tolerance,
                                               nk->krylov_dimension));
    }
    teal_error("invalid advance method");
}

static double advance_host(const Equations *eqns, Advance *method, double *time, void *residual,
                           double max_step, double courant, const void *context)
{
    TealGpu *gpu = teal_equations_device(eqns);
    check(teal_gpu_upload(gpu, eqns->primitive.data));
    double step0 = advance_device(eqns, method, time, residual, max_step, courant, context);
    check(teal_gpu_download(gpu, eqns->primitive.data));
    return step0;
}

double euler(const Equations *eqns, double *time, void *residual, double max_step, double courant,
             const void *context)
{
    return advance_host(eqns, euler, time, residual, max_step, courant, context);
}

double lserk(const Equations *eqns, double *time, void *residual, double max_step, double courant,
             const void *context)
{
    return advance_host(eqns, lserk, time, residual, max_step, courant, context);
}

double implicit_euler(const Equations *eqns, double *time, void *residual, double max_step,
                      double courant, const void *context)
{
    return advance_host(eqns, implicit_euler, time, residual, max_step, courant, context);
}

// ------------------------------------------------------------------ setup
Simulation *simulation_create(const Equations *eqns, const char *prefix)
{
    assert(eqns && prefix);
    Simulation *sim = teal_calloc(1, sizeof(Simulation));
    sim->eqns = eqns;
    snprintf(sim->prefix, sizeof(String), "%s", prefix);
    sim->time.max = sim->time.out = DBL_MAX;  // no limit until a setter gives one
    sim->iter.max = sim->iter.out = INT_MAX;
    // first order in space steps with forward Euler, second with the 2nd-order 3-stage LSERK
    simulation_set_advance(sim, (eqns->space_order == 1) ? "euler" : "lserk", 0.99, 0);
    return sim;
}

void simulation_set_max_time(Simulation *sim, double time)
{
    assert(sim && time > 0);
    sim->time.max = time;
}

void simulation_set_out_time(Simulation *sim, double time)
{
    assert(sim && time > 0);
    sim->time.out = time;
}

void simulation_set_max_iter(Simulation *sim, int iter)
{
    assert(sim && iter > 0);
    sim->iter.max = iter;
}

void simulation_set_out_iter(Simulation *sim, int iter)
{
    assert(sim && iter > 0);
    sim->iter.out = iter;
}

// "maximum", or a conserved variable named by a prefix of its name, with "-x" / "-y" / "-z"
// selecting a vector component (src/simulation/set.c:42-72)
void simulation_set_termination(Simulation *sim, const char *condition, double threshold)
{
    assert(sim && condition && threshold > 0);
    snprintf(sim->termination.condition, sizeof(String), "%s", condition);
    sim->termination.threshold = threshold;
    if (strcmp(condition, "maximum") == 0) {
        sim->termination.variable = -1;
        return;
    }
    size_t len = strlen(condition);
    int component = 0;
    if (len > 2 && condition[len - 2] == '-') {
        const char *axes = "xyz";
        const char *at = strchr(axes, condition[len - 1]);
        if (!at || !*at) {
            teal_error("invalid component index (%c)", condition[len - 1]);
        }
        component = (int)(at - axes);
        len -= 2;
    }
    const EquationsVariables *cons = &sim->eqns->conserved;
    int first = 0;  // first column of variable i
    for (int i = 0; i < cons->num; first += cons->dimension[i], i++) {
        if (strncmp(cons->name[i], condition, len) == 0) {
            sim->termination.variable = first + component;
            return;
        }
    }
    teal_error("invalid termination condition (%s)", condition);
}

// Advance methods by name (src/simulation/set.c:74-131): each with the context it copies, the
// defaults when none is given, and the checks of the context's values.
static void lserk_defaults(void *ctx, const Simulation *sim)
{
    RungeKutta *rk = ctx;
    rk->time_order = sim->eqns->space_order;
    rk->num_stages = sim->eqns->space_order + 1;
}

static void lserk_check(const void *ctx)
{
    const RungeKutta *rk = ctx;
    if (rk->time_order < 1 || rk->time_order > 3) {
        teal_error("invalid time order (%d)", rk->time_order);
    }
    if (rk->num_stages < 1 || rk->num_stages > 6) {
        teal_error("invalid number of stages (%d)", rk->num_stages);
    }
}

static void newton_defaults(void *ctx, const Simulation *sim)
{
    (void)sim;
    *(NewtonKrylov *)ctx = (NewtonKrylov){.newton_tolerance = 0.1, .krylov_tolerance = 0.3,
                                          .krylov_dimension = 32};
}

static void newton_check(const void *ctx)
{
    const NewtonKrylov *nk = ctx;
    if (!(nk->newton_tolerance > 0)) {
        teal_error("invalid newton tolerance (%g)", nk->newton_tolerance);
    }
    if (!(nk->krylov_tolerance > 0)) {
        teal_error("invalid krylov tolerance (%g)", nk->krylov_tolerance);
    }
    if (nk->krylov_dimension <= 0) {
        teal_error("invalid krylov dimension (%d)", nk->krylov_dimension);
    }
}

static const struct {
    const char *name;
    Advance *method;
    size_t context_size;
    void (*defaults)(void *ctx, const Simulation *sim);
    void (*check)(const void *ctx);
} advance_methods[] = {
    {"euler", euler, 0, 0, 0},
    {"lserk", lserk, sizeof(RungeKutta), lserk_defaults, lserk_check},
    {"implicit euler", implicit_euler, sizeof(NewtonKrylov), newton_defaults, newton_check},
};

void simulation_set_advance(Simulation *sim, const char *name, double courant, const void *context)
{
    assert(sim && name && courant > 0);
    snprintf(sim->advance.name, sizeof(String), "%s", name);
    sim->advance.courant = courant;
    teal_free(sim->advance.context);
    sim->advance.context = 0;
    for (size_t i = 0; i < sizeof(advance_methods) / sizeof(advance_methods[0]); i++) {
        if (strcmp(advance_methods[i].name, name) != 0) {
            continue;
        }
        sim->advance.method = advance_methods[i].method;
        if (advance_methods[i].context_size) {
            void *ctx = teal_calloc(1, advance_methods[i].context_size);
            if (context) {
                memcpy(ctx, context, advance_methods[i].context_size);
            }
            else {
                advance_methods[i].defaults(ctx, sim);
            }
            advance_methods[i].check(ctx);
            sim->advance.context = ctx;
        }
        return;
    }
    teal_error("invalid advance method (%s)", name);
}

void simulation_summary(const Simulation *sim)
{
    assert(sim);
    teal_print("Simulation summary");
    teal_print("\t %-16s : %s", "prefix", sim->prefix);
    if (sim->time.max < DBL_MAX) teal_print("\t %-16s : %g", "max time", sim->time.max);
    if (sim->time.out < DBL_MAX) teal_print("\t %-16s : %g", "out time", sim->time.out);
    if (sim->iter.max < INT_MAX) teal_print("\t %-16s : %d", "max iter", sim->iter.max);
    if (sim->iter.out < INT_MAX) teal_print("\t %-16s : %d", "out iter", sim->iter.out);
    if (sim->termination.condition[0]) {
        teal_print("\t %-16s : %s < %g", "termination", sim->termination.condition,
                   sim->termination.threshold);
    }
    teal_print("\t %-16s : %s", "advance method", sim->advance.name);
    teal_print("\t %-16s : %g", "courant number", sim->advance.courant);
    if (sim->advance.method == lserk) {
        const RungeKutta *rk = sim->advance.context;
        teal_print("\t %-16s : %d", "time order", rk->time_order);
        teal_print("\t %-16s : %d", "number of stages", rk->num_stages);
    }
    else if (sim->advance.method == implicit_euler) {
        const NewtonKrylov *nk = sim->advance.context;
        teal_print("\t %-16s : %g", "newton tolerance", nk->newton_tolerance);
        teal_print("\t %-16s : %g", "krylov tolerance", nk->krylov_tolerance);
        teal_print("\t %-16s : %d", "krylov dimension", nk->krylov_dimension);
    }
}

// ------------------------------------------------------------------ run loop
static volatile sig_atomic_t stop_requested = 0;

static void request_stop(int sig)
{
    (void)sig;
    stop_requested = 1;
}

static double wtime(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + (1e-9 * (double)ts.tv_nsec);
}

// Where the loop stands: simulated time and steps, the next output point (time and step) and
// the index of the next output file.
typedef struct {
    double time;
    int iter;
    int index;
    double out_time;
    int out_iter;
} RunState;

static double largest(const double *x, int num)
{
    double m = 0;  // residual norms are non-negative
    for (int i = 0; i < num; i++) {
        m = (x[i] > m) ? x[i] : m;
    }
    return m;
}

// an output file of the state at this time: the resident device state comes back first
static void write_output(const Simulation *sim, TealGpu *gpu, RunState *run)
{
    if (!sim->prefix[0]) {
        return;
    }
    check(teal_gpu_download(gpu, sim->eqns->primitive.data));
    equations_write(sim->eqns, sim->prefix, run->time, run->index++);
}

// The time loop of src/simulation/run.c:22-113 with the state resident on the device: steps
// until the iteration or time limit, the termination criterion or a signal; output (and the
// progress line) at every output time or step and at the end.
double simulation_run(Simulation *sim)
{
    assert(sim);
    const Equations *eqns = sim->eqns;
    const int stride = eqns->conserved.stride;
    RunState run = {.time = 0, .iter = 0, .index = 0, .out_time = sim->time.out,
                    .out_iter = sim->iter.out};
    if (teal.restart) {
        equations_read(eqns, teal.restart, &run.time, &run.index);
    }
    if (sim->prefix[0]) {
        mesh_write(eqns->mesh, sim->prefix);
        equations_write(eqns, sim->prefix, run.time, run.index++);
    }
    run.out_time += run.time;
    double residual[16] = {0};
    if (stride > 16) {
        teal_error("too many conserved variables (%d)", stride);
    }
    TealGpu *gpu = teal_equations_device(eqns);
    check(teal_gpu_upload(gpu, eqns->primitive.data));

    teal_print("Running simulation");
    teal_print("\t %13s %13s %13s %13s %13s", "iter", "time", "step", "residual", "wtime");
    stop_requested = 0;
    signal(SIGINT, request_stop);
    signal(SIGTERM, request_stop);
    const double wtime_beg = wtime();
    double wtime_last = wtime_beg;
    int converged = 0;
    for (;;) {
        if (run.iter >= sim->iter.max || !is_less(run.time, sim->time.max) || converged ||
            stop_requested) {
            break;
        }
        const double horizon = fmin(sim->time.max, run.out_time);
        const double step0 = advance_device(eqns, sim->advance.method, &run.time, residual,
                                            horizon - run.time, sim->advance.courant,
                                            sim->advance.context);
        run.iter += 1;
        const double max_residual = largest(residual, stride);
        if (sim->termination.condition[0]) {
            const int var = sim->termination.variable;
            converged = ((var >= 0) ? residual[var] : max_residual) < sim->termination.threshold;
        }
        const int at_time = is_greater_equal(run.time, fmin(sim->time.max, run.out_time));
        const int at_iter = run.iter >= ((run.out_iter < sim->iter.max) ? run.out_iter : sim->iter.max);
        if (!(at_time || at_iter || converged || stop_requested)) {
            continue;
        }
        const double now = wtime();
        write_output(sim, gpu, &run);
        if (sim->time.out < DBL_MAX) run.out_time += sim->time.out;
        if (sim->iter.out < INT_MAX) run.out_iter += sim->iter.out;
        teal_print("\t %13d %13g %13g %13g %13g", run.iter, run.time, step0, max_residual,
                   now - wtime_last);
        wtime_last = now;
    }
    const double wtime_end = wtime();
    check(teal_gpu_download(gpu, eqns->primitive.data));
    signal(SIGINT, SIG_DFL);
    signal(SIGTERM, SIG_DFL);
    teal_print("\t computation time : %g", wtime_end - wtime_beg);
    return run.time;
}

// src/simulation/error.c:8-35: the error norms, one line per primitive variable
void simulation_error(const Simulation *sim, double time, void *norm)
{
    assert(sim);
    const EquationsVariables *prim = &sim->eqns->primitive;
    double *out = norm ? norm : teal_calloc(prim->stride, sizeof(double));
    equations_norm(sim->eqns, out, time);
    teal_print("Simulation error");
    const double *value = out;
    for (int i = 0; i < prim->num; value += prim->dimension[i], i++) {
        char line[256] = "";
        size_t used = 0;
        for (int k = 0; k < prim->dimension[i] && used < sizeof(line); k++) {
            used += (size_t)snprintf(&line[used], sizeof(line) - used, " %g", value[k]);
        }
        teal_print("\t %-20s :%s", prim->name[i], line);
    }
    if (!norm) {
        teal_free(out);
    }
}

void simulation_destroy(Simulation *sim)
{
    assert(sim);
    teal_free(sim->advance.context);
    teal_free(sim);
}
