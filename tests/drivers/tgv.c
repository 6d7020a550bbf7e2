// A teal driver written against the reference API only (include/*.h), linked with
// libteal_amd.so: the Navier-Stokes Taylor-Green vortex on an n^3 periodic box, run through
// simulation_run, with the state written before and after the run for comparison.
//
//     tgv <n> <iterations> <output prefix> [fp64|mixed]
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "navierstokes.h"
#include "simulation.h"
#include "sync.h"
#include "teal.h"
#include "teal_gpu.h"

static const double mach = 0.1;

static void initial(void *primitive_, const double *property, Vector x, double time,
                    const void *context)
{
    double *p = primitive_;  // density, velocity (3), pressure
    double gamma = property[0];
    double p0 = 1 / (gamma * mach * mach);
    p[0] = 1;
    p[1] = sin(x.x) * cos(x.y) * cos(x.z);
    p[2] = -cos(x.x) * sin(x.y) * cos(x.z);
    p[3] = 0;
    p[4] = p0 + ((cos(2 * x.x) + cos(2 * x.y)) * (cos(2 * x.z) + 2) / 16);
}

static void dump(const Equations *eqns, const char *prefix, const char *what)
{
    char name[512];
    snprintf(name, sizeof(name), "%s_%s.bin", prefix, what);
    FILE *f = fopen(name, "wb");
    if (!f) {
        teal_error("cannot open %s", name);
    }
    size_t num = (size_t)eqns->mesh->cells.num * eqns->primitive.stride;
    if (fwrite(eqns->primitive.data, sizeof(double), num, f) != num) {
        teal_error("cannot write %s", name);
    }
    fclose(f);
}

int main(int argc, char **argv)
{
    teal_init(&argc, &argv);
    if (argc < 4) {
        teal_error("usage: tgv <n> <iterations> <output prefix> [fp64|mixed]");
    }
    int n = (int)strtol(argv[1], 0, 10);
    int iterations = (int)strtol(argv[2], 0, 10);
    const char *prefix = argv[3];
    int mixed = (argc > 4) && !strcmp(argv[4], "mixed");

    Vector lo = {0, 0, 0};
    Vector hi = {2 * M_PI, 2 * M_PI, 2 * M_PI};
    Mesh *mesh = mesh_create(lo, hi, (Triple){n, n, n}, (Triple){1, 1, 1});
    mesh_generate(mesh);

    Equations *eqns = navierstokes_create(mesh);
    equations_set_initial_condition(eqns, "domain", initial, 0);
    equations_set_property(eqns, "dynamic viscosity", 1.0 / 1600);
    if (mixed) {
        teal_equations_set_precision(eqns, TEAL_MIXED);
    }
    dump(eqns, prefix, "initial");

    Simulation *sim = simulation_create(eqns, "");
    simulation_set_max_iter(sim, iterations);
    simulation_set_advance(sim, "lserk", 0.5, 0);
    double time = simulation_run(sim);
    dump(eqns, prefix, "final");
    teal_print("ranks %d, final time %.17g", sync.size, time);

    simulation_destroy(sim);
    equations_destroy(eqns);
    mesh_destroy(mesh);
    teal_deinit();
}
