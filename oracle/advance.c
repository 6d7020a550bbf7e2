/* CPU ORACLE (test infrastructure only) -- time integrators.
 * Restates src/simulation/advance.c:11-195 and src/simulation/gmres.c:12-149 on one rank
 * (sync_norm / sync_dot are serial sums, src/core/sync.c:66-79). */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static double dot(const double *a, const double *b, long n) /* src/core/sync.c:66-74 */
{
    double s = 0;
    for (long i = 0; i < n; i++) {
        s += a[i] * b[i];
    }
    return s;
}

static double norm(const double *a, long n) { return sqrt(dot(a, a, n)); }

/* src/simulation/advance.c:11-45 */
double o_euler_step(OEqns *e, double *time, double *residual, double max_step, double courant)
{
    int ni = e->mesh->cells.num_inner;
    double step0 = courant * o_timestep(e, (const double (*)[5])e->primitive);
    double step = fmin(step0, max_step);
    double (*d)[5] = o_calloc(ni, sizeof(*d));
    o_derivative(e, e->primitive, d);
    for (int i = 0; i < ni; i++) {
        o_prim2cons(e->conserved[i], e->primitive[i], e->property);
        for (int j = 0; j < 5; j++) {
            e->conserved[i][j] += d[i][j] * step;
        }
        o_cons2prim(e->primitive[i], e->conserved[i], e->property);
    }
    *time += step;
    o_residual(e, (const double (*)[5])d, residual);
    free(d);
    return step0;
}

/* src/simulation/advance.c:50-75 (4-digit literals, kept as written) */
static const double lserk_alpha[3][6][7] = {
    {
        {0.0000, 1.0000, 0.0000, 0.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.3333, 1.0000, 0.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.1481, 0.4000, 1.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.0833, 0.2069, 0.4265, 1.0000, 0.0000, 0.0000},
        {0.0000, 0.0533, 0.1263, 0.2375, 0.4414, 1.0000, 0.0000},
        {0.0000, 0.0370, 0.0851, 0.1521, 0.2562, 0.4512, 1.0000},
    },
    {
        {0.0000, 1.0000, 0.0000, 0.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.4242, 1.0000, 0.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.1918, 0.4929, 1.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.1084, 0.2602, 0.5052, 1.0000, 0.0000, 0.0000},
        {0.0000, 0.0695, 0.1602, 0.2898, 0.5060, 1.0000, 0.0000},
        {0.0000, 0.0482, 0.1085, 0.1885, 0.3050, 0.5063, 1.0000},
    },
    {
        {0.0000, 1.0000, 0.0000, 0.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.6612, 1.0000, 0.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.2884, 0.5010, 1.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.1666, 0.3027, 0.5275, 1.0000, 0.0000, 0.0000},
        {0.0000, 0.1067, 0.1979, 0.3232, 0.5201, 1.0000, 0.0000},
        {0.0000, 0.0742, 0.1393, 0.2198, 0.3302, 0.5181, 1.0000},
    },
};

/* src/simulation/advance.c:47-117 */
double o_lserk(OEqns *e, double *time, double *residual, double max_step, double courant,
               int time_order, int num_stages)
{
    const double *alpha = lserk_alpha[time_order - 1][num_stages - 1];
    int ni = e->mesh->cells.num_inner;
    double step0 = courant * o_timestep(e, (const double (*)[5])e->primitive);
    double step = fmin(step0, max_step);
    double (*u0)[5] = o_calloc(ni, sizeof(*u0));
    for (int i = 0; i < ni; i++) {
        o_prim2cons(u0[i], e->primitive[i], e->property);
    }
    double (*d)[5] = o_calloc(ni, sizeof(*d));
    for (int s = 0; s < num_stages; s++) {
        o_derivative(e, e->primitive, d);
        for (int j = 0; j < ni; j++) {
            for (int k = 0; k < 5; k++) {
                e->conserved[j][k] = u0[j][k] + (d[j][k] * alpha[s + 1] * step);
            }
            o_cons2prim(e->primitive[j], e->conserved[j], e->property);
        }
    }
    *time += step;
    o_residual(e, (const double (*)[5])d, residual);
    free(d);
    free(u0);
    return step0;
}

/* matvec, src/simulation/gmres.c:12-50 */
static void matvec(const OEqns *e, const double (*u)[5], const double (*d)[5],
                   const double (*v)[5], double (*out)[5], double step, double fd_scale)
{
    const OMesh *m = e->mesh;
    int ni = m->cells.num_inner;
    double eps = fd_scale / norm(&v[0][0], (long)ni * 5);
    double (*p1)[5] = o_calloc(m->cells.num, sizeof(*p1));
    double pert[5];
    for (int i = 0; i < ni; i++) {
        for (int j = 0; j < 5; j++) {
            pert[j] = u[i][j] + (eps * v[i][j]);
        }
        o_cons2prim(p1[i], pert, e->property);
    }
    double (*d1)[5] = o_calloc(ni, sizeof(*d1));
    o_derivative(e, p1, d1);
    for (int i = 0; i < ni; i++) {
        for (int j = 0; j < 5; j++) {
            out[i][j] = v[i][j] - (step * (d1[i][j] - d[i][j]) / eps);
        }
    }
    free(d1);
    free(p1);
}

/* src/simulation/gmres.c:52-149 */
void o_gmres(const OEqns *e, const double (*u)[5], const double (*d)[5], const double (*r)[5],
             double (*inc)[5], double step, double nrm, double fd_scale, double tol, int dim)
{
    int ni = e->mesh->cells.num_inner;
    long n = (long)ni * 5;
    memset(inc, 0, (size_t)ni * sizeof(*inc));
    if (nrm <= 0) {
        return;
    }
    double *rhs = o_calloc(dim + 1, sizeof(double));
    rhs[0] = nrm;
    double (*basis)[5] = o_calloc((size_t)(dim + 1) * ni, sizeof(*basis));
    for (int i = 0; i < ni; i++) {
        for (int j = 0; j < 5; j++) {
            basis[i][j] = -r[i][j] / nrm;
        }
    }
    double tol_norm = tol * nrm;
    double (*res)[5] = o_calloc(ni, sizeof(*res));
    double *hess = o_calloc((size_t)(dim + 1) * dim, sizeof(double));
    double *cs = o_calloc(dim, sizeof(double));
    double *sn = o_calloc(dim, sizeof(double));
#define H(i, j) hess[((i) * dim) + (j)]
    int iter = 0;
    for (; iter < dim; iter++) {
        double (*bi)[5] = &basis[(size_t)iter * ni];
        matvec(e, u, d, (const double (*)[5])bi, res, step, fd_scale);
        for (int i = 0; i < iter + 1; i++) {
            double *b = &basis[(size_t)i * ni][0];
            H(i, iter) = dot(b, &res[0][0], n);
            for (long j = 0; j < n; j++) {
                (&res[0][0])[j] -= H(i, iter) * b[j];
            }
        }
        H(iter + 1, iter) = norm(&res[0][0], n);
        for (int i = 0; i < iter; i++) {
            double tmp = (cs[i] * H(i, iter)) + (sn[i] * H(i + 1, iter));
            H(i + 1, iter) = (-sn[i] * H(i, iter)) + (cs[i] * H(i + 1, iter));
            H(i, iter) = tmp;
        }
        double rot = sqrt((H(iter, iter) * H(iter, iter)) + (H(iter + 1, iter) * H(iter + 1, iter)));
        cs[iter] = H(iter, iter) / rot;
        sn[iter] = H(iter + 1, iter) / rot;
        H(iter, iter) = rot;
        rhs[iter + 1] = -sn[iter] * rhs[iter];
        rhs[iter] = cs[iter] * rhs[iter];
        if (fabs(rhs[iter + 1]) < tol_norm) {
            iter += 1;
            break;
        }
        double *next = &basis[(size_t)(iter + 1) * ni][0];
        for (long j = 0; j < n; j++) {
            next[j] = (&res[0][0])[j] / H(iter + 1, iter);
        }
    }
    double *coef = o_calloc(dim, sizeof(double));
    for (int i = iter - 1; i >= 0; i--) {
        coef[i] = rhs[i];
        for (int j = i + 1; j < iter; j++) {
            coef[i] -= H(i, j) * coef[j];
        }
        coef[i] /= H(i, i);
        double *b = &basis[(size_t)i * ni][0];
        for (long j = 0; j < n; j++) {
            (&inc[0][0])[j] += coef[i] * b[j];
        }
    }
#undef H
    free(coef);
    free(sn);
    free(cs);
    free(hess);
    free(res);
    free(basis);
    free(rhs);
}

/* src/simulation/advance.c:119-195 */
double o_implicit_euler(OEqns *e, double *time, double *residual, double max_step,
                        double courant, double newton_tol, double krylov_tol, int krylov_dim)
{
    int ni = e->mesh->cells.num_inner;
    long n = (long)ni * 5;
    double step0 = courant * o_timestep(e, (const double (*)[5])e->primitive);
    double step = fmin(step0, max_step);
    double (*u0)[5] = o_calloc(ni, sizeof(*u0));
    for (int i = 0; i < ni; i++) {
        o_prim2cons(e->conserved[i], e->primitive[i], e->property);
    }
    memcpy(u0, e->conserved, (size_t)ni * sizeof(*u0));
    *time += step;
    double (*d)[5] = o_calloc(ni, sizeof(*d));
    o_derivative(e, e->primitive, d);
    double (*nr)[5] = o_calloc(ni, sizeof(*nr));
    for (int i = 0; i < ni; i++) {
        for (int j = 0; j < 5; j++) {
            nr[i][j] = -d[i][j] * step;
        }
    }
    double nrm = norm(&nr[0][0], n);
    double cons_norm = norm(&e->conserved[0][0], n);
    double fd_scale = 1e-6 * fmax(1, cons_norm) / fmax(1, nrm);
    double tol_norm = newton_tol * nrm;
    int max_iter = 128 * (ni + 1);
    double (*inc)[5] = o_calloc(ni, sizeof(*inc));
    for (int it = 0; it < max_iter && nrm > tol_norm; it++) {
        o_gmres(e, (const double (*)[5])e->conserved, (const double (*)[5])d,
                (const double (*)[5])nr, inc, step, nrm, fd_scale, krylov_tol, krylov_dim);
        for (int i = 0; i < ni; i++) {
            for (int j = 0; j < 5; j++) {
                e->conserved[i][j] += inc[i][j];
            }
            o_cons2prim(e->primitive[i], e->conserved[i], e->property);
        }
        o_derivative(e, e->primitive, d);
        for (int i = 0; i < ni; i++) {
            for (int j = 0; j < 5; j++) {
                nr[i][j] = e->conserved[i][j] - u0[i][j] - (d[i][j] * step);
            }
        }
        nrm = norm(&nr[0][0], n);
        cons_norm = norm(&e->conserved[0][0], n);
        fd_scale = 1e-6 * fmax(1, cons_norm) / fmax(1, nrm);
    }
    o_residual(e, (const double (*)[5])d, residual);
    free(inc);
    free(nr);
    free(d);
    free(u0);
    return step0;
}
