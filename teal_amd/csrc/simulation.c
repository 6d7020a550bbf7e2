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
    Simulation *sim = teal_calloc(1, sizeof(*sim));
    sim->eqns = eqns;
    strcpy(sim->prefix, prefix);
    sim->time.max = DBL_MAX;
    sim->time.out = DBL_MAX;
    sim->iter.max = INT_MAX;
    sim->iter.out = INT_MAX;
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

static int component_index(char chr)
{
    switch (chr) {
        case 'x': return 0;
        case 'y': return 1;
        case 'z': return 2;
        default: teal_error("invalid component index (%c)", chr);
    }
}

void simulation_set_termination(Simulation *sim, const char *condition, double threshold)
{
    assert(sim && condition && threshold > 0);
    strcpy(sim->termination.condition, condition);
    sim->termination.threshold = threshold;
    if (!strcmp(condition, "maximum")) {
        sim->termination.variable = -1;
        return;
    }
    int idx = 0;
    size_t len = strlen(condition);
    if (len > 2 && condition[len - 2] == '-') {
        idx = component_index(condition[len - 1]);
        len -= 2;
    }
    const EquationsVariables *cons = &sim->eqns->conserved;
    for (int i = 0, j = 0; i < cons->num; j += cons->dimension[i++]) {
        if (!strncmp(cons->name[i], condition, len)) {
            sim->termination.variable = j + idx;
            return;
        }
    }
    teal_error("invalid termination condition (%s)", condition);
}

void simulation_set_advance(Simulation *sim, const char *name, double courant, const void *context)
{
    assert(sim && name && courant > 0);
    strcpy(sim->advance.name, name);
    sim->advance.courant = courant;
    teal_free(sim->advance.context);
    sim->advance.context = 0;
    if (!strcmp(name, "euler")) {
        sim->advance.method = euler;
        return;
    }
    if (!strcmp(name, "lserk")) {
        RungeKutta *rk = teal_calloc(1, sizeof(*rk));
        if (context) {
            memcpy(rk, context, sizeof(*rk));
        }
        else {
            rk->time_order = sim->eqns->space_order;
            rk->num_stages = sim->eqns->space_order + 1;
        }
        if (!(1 <= rk->time_order && rk->time_order <= 3)) {
            teal_error("invalid time order (%d)", rk->time_order);
        }
        if (!(1 <= rk->num_stages && rk->num_stages <= 6)) {
            teal_error("invalid number of stages (%d)", rk->num_stages);
        }
        sim->advance.context = rk;
        sim->advance.method = lserk;
        return;
    }
    if (!strcmp(name, "implicit euler")) {
        NewtonKrylov *nk = teal_calloc(1, sizeof(*nk));
        if (context) {
            memcpy(nk, context, sizeof(*nk));
        }
        else {
            nk->newton_tolerance = 0.1;
            nk->krylov_tolerance = 0.3;
            nk->krylov_dimension = 32;
        }
        if (nk->newton_tolerance <= 0) {
            teal_error("invalid newton tolerance (%g)", nk->newton_tolerance);
        }
        if (nk->krylov_tolerance <= 0) {
            teal_error("invalid krylov tolerance (%g)", nk->krylov_tolerance);
        }
        if (nk->krylov_dimension <= 0) {
            teal_error("invalid krylov dimension (%d)", nk->krylov_dimension);
        }
        sim->advance.context = nk;
        sim->advance.method = implicit_euler;
        return;
    }
    teal_error("invalid advance method (%s)", name);
}

void simulation_summary(const Simulation *sim)
{
    assert(sim);
    teal_print("Simulation summary");
    teal_print("\t prefix           : %s", sim->prefix);
    if (sim->time.max < DBL_MAX) {
        teal_print("\t max time         : %g", sim->time.max);
    }
    if (sim->time.out < DBL_MAX) {
        teal_print("\t out time         : %g", sim->time.out);
    }
    if (sim->iter.max < INT_MAX) {
        teal_print("\t max iter         : %d", sim->iter.max);
    }
    if (sim->iter.out < INT_MAX) {
        teal_print("\t out iter         : %d", sim->iter.out);
    }
    if (sim->termination.condition[0]) {
        teal_print("\t termination      : %s < %g", sim->termination.condition,
                   sim->termination.threshold);
    }
    teal_print("\t advance method   : %s", sim->advance.name);
    teal_print("\t courant number   : %g", sim->advance.courant);
    if (sim->advance.method == lserk) {
        RungeKutta *rk = sim->advance.context;
        teal_print("\t time order       : %d", rk->time_order);
        teal_print("\t number of stages : %d", rk->num_stages);
    }
    if (sim->advance.method == implicit_euler) {
        NewtonKrylov *nk = sim->advance.context;
        teal_print("\t newton tolerance : %g", nk->newton_tolerance);
        teal_print("\t krylov tolerance : %g", nk->krylov_tolerance);
        teal_print("\t krylov dimension : %d", nk->krylov_dimension);
    }
}

// ------------------------------------------------------------------ run loop
static volatile sig_atomic_t sig_terminate = 0;

static void handler(int sig)
{
    (void)sig;
    sig_terminate = 1;
}

static double wtime(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + (1e-9 * (double)ts.tv_nsec);
}

double simulation_run(Simulation *sim)
{
    assert(sim);
    const Equations *eqns = sim->eqns;
    int stride = eqns->conserved.stride;
    const char *prefix = sim->prefix;
    double max_time = sim->time.max;
    double out_time = sim->time.out;
    int max_iter = sim->iter.max;
    int out_iter = sim->iter.out;
    double time = 0;
    int index = 0;
    if (teal.restart) {
        equations_read(eqns, teal.restart, &time, &index);
    }
    if (prefix[0]) {
        equations_write(eqns, prefix, time, index++);
    }
    out_time += time;
    double residual[8] = {0};
    if (stride > 8) {
        teal_error("too many conserved variables (%d)", stride);
    }

    TealGpu *gpu = teal_equations_device(eqns);
    check(teal_gpu_upload(gpu, eqns->primitive.data));

    teal_print("Running simulation");
    teal_print("\t %13s %13s %13s %13s %13s", "iter", "time", "step", "residual", "wtime");
    sig_terminate = 0;
    signal(SIGINT, handler);
    signal(SIGTERM, handler);
    double wtime_beg = wtime();
    double wtime_last = wtime_beg;
    int iter = 0;
    int has_converged = 0;
    while (iter < max_iter && is_less(time, max_time) && !has_converged && !sig_terminate) {
        double max_step = fmin(max_time, out_time) - time;
        double step0 = advance_device(eqns, sim->advance.method, &time, residual, max_step,
                                      sim->advance.courant, sim->advance.context);
        iter += 1;
        double max_residual = 0;
        for (int i = 0; i < stride; i++) {
            if (residual[i] > max_residual) {
                max_residual = residual[i];
            }
        }
        if (sim->termination.condition[0]) {
            int var = sim->termination.variable;
            has_converged = ((var >= 0) ? residual[var] : max_residual) < sim->termination.threshold;
        }
        int has_reached_time = is_greater_equal(time, fmin(max_time, out_time));
        int has_reached_iter = iter >= ((out_iter < max_iter) ? out_iter : max_iter);
        if (has_reached_time || has_reached_iter || has_converged || sig_terminate) {
            double wtime_now = wtime();
            if (prefix[0]) {
                check(teal_gpu_download(gpu, eqns->primitive.data));
                equations_write(eqns, prefix, time, index++);
            }
            if (sim->time.out < DBL_MAX) {
                out_time += sim->time.out;
            }
            if (sim->iter.out < INT_MAX) {
                out_iter += sim->iter.out;
            }
            teal_print("\t %13d %13g %13g %13g %13g", iter, time, step0, max_residual,
                       wtime_now - wtime_last);
            wtime_last = wtime_now;
        }
    }
    double wtime_end = wtime();
    check(teal_gpu_download(gpu, eqns->primitive.data));
    signal(SIGINT, SIG_DFL);
    signal(SIGTERM, SIG_DFL);
    teal_print("\t computation time : %g", wtime_end - wtime_beg);
    return time;
}

void simulation_error(const Simulation *sim, double time, void *norm_)
{
    assert(sim);
    const EquationsVariables *prim = &sim->eqns->primitive;
    double *norm = norm_ ? norm_ : teal_calloc(prim->stride, sizeof(double));
    equations_norm(sim->eqns, norm, time);
    teal_print("Simulation error");
    for (int i = 0, j = 0; i < prim->num; j += prim->dimension[i++]) {
        int pos = 0;
        char buf[128];
        for (int k = 0; k < prim->dimension[i]; k++) {
            pos += snprintf(&buf[pos], sizeof(buf) - pos, " %g", norm[j + k]);
        }
        teal_print("\t %-20s :%s", prim->name[i], buf);
    }
    if (!norm_) {
        teal_free(norm);
    }
}

void simulation_destroy(Simulation *sim)
{
    assert(sim);
    teal_free(sim->advance.context);
    teal_free(sim);
}
