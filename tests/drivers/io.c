// VTKHDF output and restart through the teal API (host only on a fully periodic mesh):
//   io write <prefix> <n>      mesh_write + equations_write (index 3, time 1.25) of a seeded
//                              state on an n^3 periodic box; the state to <prefix>_state.bin
//   io read <file> <n> <out>   equations_read into a fresh n^3 box; the state to <out>.bin,
//                              "time <t> index <i> mu <mu>" on stdout
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "navierstokes.h"
#include "teal.h"

static Mesh *box(int n)
{
    Vector lo = {0, 0, 0}, hi = {1, 2, 0.5};
    Triple num = {n, n, n}, periodic = {1, 1, 1};
    Mesh *mesh = mesh_create(lo, hi, num, periodic);
    mesh_generate(mesh);
    return mesh;
}

static void dump(const char *path, const double *x, size_t n)
{
    FILE *f = fopen(path, "wb");
    if (!f || fwrite(x, sizeof(double), n, f) != n) {
        fprintf(stderr, "cannot write %s\n", path);
        exit(1);
    }
    fclose(f);
}

int main(int argc, char **argv)
{
    teal_init(&argc, &argv);
    if (argc < 4) {
        fprintf(stderr, "usage: io write <prefix> <n> | io read <file> <n> <out>\n");
        return 2;
    }
    Mesh *mesh = box(atoi(argv[3]));
    Equations *eqns = navierstokes_create(mesh);
    int ni = mesh->cells.num_inner, nc = mesh->cells.num;
    double *P = eqns->primitive.data;
    char path[512];
    if (!strcmp(argv[1], "write")) {
        for (int c = 0; c < ni; c++) {
            Vector x = mesh->cells.center[c];
            P[5 * c + 0] = 1 + 0.1 * sin(6.0 * x.x + x.y);
            P[5 * c + 1] = cos(3.0 * x.y);
            P[5 * c + 2] = sin(2.0 * x.z) * x.x;
            P[5 * c + 3] = 0.25 * x.y;
            P[5 * c + 4] = 2 + x.x * x.z;
        }
        equations_set_property(eqns, "dynamic viscosity", 0.0125);
        mesh_write(mesh, argv[2]);
        equations_write(eqns, argv[2], 1.25, 3);
        snprintf(path, sizeof(path), "%s_state.bin", argv[2]);
        dump(path, P, (size_t)nc * 5);
    }
    else {
        double time = 0;
        int index = 0;
        equations_read(eqns, argv[2], &time, &index);
        printf("time %.17g index %d mu %.17g\n", time, index, eqns->properties.data[1]);
        snprintf(path, sizeof(path), "%s.bin", argv[4]);
        dump(path, P, (size_t)nc * 5);
    }
    equations_destroy(eqns);
    mesh_destroy(mesh);
    teal_deinit();
    return 0;
}
