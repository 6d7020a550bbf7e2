// krylov.hip -- implicit Euler with Jacobian-free Newton-GMRES on the device
// (src/simulation/advance.c:119-195, src/simulation/gmres.c:12-149).
#include <cmath>

#include "device.h"

struct Krylov {
    int dim = 0;
};

void teal_krylov_release(TealGpu *g)
{
    if (g && g->krylov) {
        delete g->krylov;
        g->krylov = nullptr;
    }
}

extern "C" double teal_gpu_implicit_euler(TealGpu *g, double *time, double *residual,
                                          double max_step, double courant, double newton_tol,
                                          double krylov_tol, int krylov_dim)
{
    (void)g, (void)time, (void)residual, (void)max_step, (void)courant, (void)newton_tol,
        (void)krylov_tol, (void)krylov_dim;
    teal_gpu_set_error("implicit euler: not available yet");
    return NAN;
}
