/* CPU ORACLE (test infrastructure only) -- the residual path on one rank.
 *
 * Restates src/equations/derivative.c, gradient.c, limiter.c, boundary.c,
 * exchange.c (periodic self-exchange), timestep.c, residual.c and the defaults
 * of src/euler/create.c:46-64 / src/navierstokes/create.c:28-51.
 * Accumulation order is the reference's: faces in index order, viscous pass
 * before convective pass, `R -= F*A` / `R += F*A` per side. */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

OEqns *oeqns_create(const OMesh *mesh, int navierstokes)
{
    OEqns *e = o_calloc(1, sizeof(*e));
    e->mesh = mesh;
    e->navierstokes = navierstokes;
    e->space_order = 2;
    e->flux = O_FLUX_HLLC;
    e->viscous = navierstokes;
    e->limiter = O_LIM_BJ; /* "minmod" */
    e->property[0] = 1.4;
    e->property[1] = 0;
    e->property[2] = 0.71;
    e->num_bc = mesh->entities.off_boundary - mesh->entities.num_inner;
    e->bc = o_calloc(e->num_bc + 1, sizeof(int));
    e->bc_ctx = o_calloc(e->num_bc + 1, sizeof(double[5]));
    e->primitive = o_calloc(mesh->cells.num, sizeof(double[5]));
    e->conserved = o_calloc(mesh->cells.num, sizeof(double[5]));
    return e;
}

void oeqns_destroy(OEqns *e)
{
    if (!e) {
        return;
    }
    free(e->eps2);
    free(e->bc);
    free(e->bc_ctx);
    free(e->primitive);
    free(e->conserved);
    free(e);
}

int o_limiter_by_name(const char *name) /* src/equations/set.c:45-70 */
{
    if (!name) {
        return O_LIM_NONE;
    }
    if (!strcmp(name, "barth jespersen") || !strcmp(name, "minmod")) {
        return O_LIM_BJ;
    }
    if (!strcmp(name, "venkatakrishnan")) {
        return O_LIM_VK;
    }
    o_fatal("invalid limiter (%s)", name);
}

/* src/equations/set.c:45-70 and venkatakrishnan_parameter, limiter.c:8-21 */
void oeqns_set_limiter(OEqns *e, int kind, double coefficient)
{
    free(e->eps2);
    e->eps2 = 0;
    e->limiter = kind;
    if (kind == O_LIM_VK) {
        int ni = e->mesh->cells.num_inner;
        e->eps2 = o_calloc(ni, sizeof(double));
        for (int i = 0; i < ni; i++) {
            double v = coefficient * cbrt(e->mesh->cells.volume[i]);
            e->eps2[i] = v * v * v;
        }
    }
}

void oeqns_set_bc(OEqns *e, int entity, int kind, const double *ctx)
{
    e->bc[entity] = kind;
    if (ctx) {
        memcpy(e->bc_ctx[entity], ctx, sizeof(double[5]));
    }
}

/* src/equations/exchange.c:8-51 on one rank: ghost slots of each periodic block receive the
 * values of the opposite side's send list */
static void exchange(const OMesh *mesh, double *buf, int len)
{
    const ONeighbors *nb = &mesh->neighbors;
    for (int e = 0; e < nb->num; e++) {
        int s = -1;
        for (int k = 0; k < nb->num; k++) {
            if (nb->tag[k][1] == nb->tag[e][0]) {
                s = k;
            }
        }
        for (int j = nb->recv_off[e]; j < nb->recv_off[e + 1]; j++) {
            int src = nb->send.idx[nb->send.off[s] + (j - nb->recv_off[e])];
            memcpy(&buf[(size_t)j * len], &buf[(size_t)src * len], len * sizeof(double));
        }
    }
}

/* src/equations/timestep.c:7-30 */
double o_timestep(const OEqns *e, const double (*p)[5])
{
    const OMesh *m = e->mesh;
    double min_step = DBL_MAX;
    for (int i = 0; i < m->cells.num_inner; i++) {
        double step = e->navierstokes
                          ? o_ns_timestep(p[i], e->property, m->cells.volume[i],
                                          m->cells.projection[i])
                          : o_euler_timestep(p[i], e->property, m->cells.volume[i],
                                             m->cells.projection[i]);
        if (step < min_step) {
            min_step = step;
        }
    }
    return min_step;
}

/* src/equations/boundary.c:5-43 */
void o_boundary(const OEqns *e, double (*p)[5])
{
    const OMesh *m = e->mesh;
    const int *face_off = &m->entities.face_off[m->entities.num_inner];
    for (int i = 0; i < e->num_bc; i++) {
        for (int j = face_off[i]; j < face_off[i + 1]; j++) {
            int left = m->faces.cell_idx[j].left;
            int right = m->faces.cell_idx[j].right;
            o_bc(e->bc[i], p[right], p[left], e->property, m->faces.basis[j], e->bc_ctx[i]);
        }
    }
}

/* src/equations/gradient.c:7-58 */
void o_gradient(const OEqns *e, double (*p)[5], OVec (*g)[5])
{
    const OMesh *m = e->mesh;
    exchange(m, &p[0][0], 5);
    memset(g, 0, (size_t)m->cells.num * sizeof(*g));
    for (int i = 0; i < m->faces.num; i++) {
        int left = m->faces.cell_idx[i].left;
        int right = m->faces.cell_idx[i].right;
        for (int j = 0; j < 5; j++) {
            double diff = p[right][j] - p[left][j];
            OVec wl = ov_mul(diff, m->faces.weight[i].left);
            g[left][j].x += wl.x;
            g[left][j].y += wl.y;
            g[left][j].z += wl.z;
            if (i < m->faces.num_inner) {
                OVec wr = ov_mul(-diff, m->faces.weight[i].right);
                g[right][j].x += wr.x;
                g[right][j].y += wr.y;
                g[right][j].z += wr.z;
            }
        }
    }
}

/* src/equations/limiter.c:23-123 */
void o_limiter(const OEqns *e, const double (*p)[5], OVec (*g)[5])
{
    const OMesh *m = e->mesh;
    const int *off = m->cells.cell.off;
    const int *idx = m->cells.cell.idx;
    for (int i = 0; i < m->cells.num_inner; i++) {
        double mn[5];
        double mx[5];
        double psi[5];
        for (int k = 0; k < 5; k++) {
            mn[k] = mx[k] = p[i][k];
        }
        for (int j = off[i]; j < off[i + 1]; j++) {
            for (int k = 0; k < 5; k++) {
                mn[k] = fmin(mn[k], p[idx[j]][k]);
                mx[k] = fmax(mx[k], p[idx[j]][k]);
            }
        }
        for (int k = 0; k < 5; k++) {
            psi[k] = 1;
        }
        for (int j = off[i]; j < off[i + 1]; j++) {
            for (int k = 0; k < 5; k++) {
                double d2 = ov_dot(g[i][k], m->cells.offset[j]);
                if (o_is_close(d2, 0)) {
                    continue;
                }
                double d1 = ((d2 > 0) ? mx[k] : mn[k]) - p[i][k];
                if (e->limiter == O_LIM_BJ) {
                    psi[k] = fmin(psi[k], fmin(1, d1 / d2));
                }
                else {
                    double eps2 = e->eps2[i];
                    double top = (((d1 * d1) + eps2) * d2) + (2 * (d2 * d2) * d1);
                    double bot = (d1 * d1) + (2 * (d2 * d2)) + (d1 * d2) + eps2;
                    psi[k] = fmin(psi[k], top / bot / d2);
                }
            }
        }
        for (int k = 0; k < 5; k++) {
            g[i][k].x *= psi[k];
            g[i][k].y *= psi[k];
            g[i][k].z *= psi[k];
        }
    }
}

/* src/equations/derivative.c:18-69 */
static void convective_o1(const OEqns *e, double (*p)[5], double (*r)[5])
{
    const OMesh *m = e->mesh;
    exchange(m, &p[0][0], 5);
    double flux[5];
    for (int i = 0; i < m->faces.num; i++) {
        int left = m->faces.cell_idx[i].left;
        int right = m->faces.cell_idx[i].right;
        o_convective(e->flux, flux, p[left], p[right], e->property, m->faces.basis[i]);
        for (int j = 0; j < 5; j++) {
            r[left][j] += flux[j] * m->faces.area[i];
            if (i < m->faces.num_inner) {
                r[right][j] -= flux[j] * m->faces.area[i];
            }
        }
    }
}

/* src/equations/derivative.c:71-156 */
static void viscous_o2(const OEqns *e, const double (*p)[5], double (*r)[5], OVec (*g)[5])
{
    const OMesh *m = e->mesh;
    exchange(m, &g[0][0].x, 15);
    double pm[5];
    OVec gm[5];
    double flux[5];
    for (int i = 0; i < m->faces.num; i++) {
        int left = m->faces.cell_idx[i].left;
        int right = m->faces.cell_idx[i].right;
        int boundary = (m->faces.num_inner <= i && i < m->faces.off_boundary);
        for (int k = 0; k < 5; k++) {
            pm[k] = (p[left][k] + p[right][k]) / 2;
        }
        for (int k = 0; k < 5; k++) {
            if (boundary) {
                gm[k] = g[left][k];
            }
            else {
                gm[k].x = (g[left][k].x + g[right][k].x) / 2;
                gm[k].y = (g[left][k].y + g[right][k].y) / 2;
                gm[k].z = (g[left][k].z + g[right][k].z) / 2;
            }
        }
        OVecNorm c = m->faces.correction[i];
        for (int k = 0; k < 5; k++) { /* correct, derivative.c:81-89 */
            double proj = ov_dot(gm[k], c.unit);
            double adjust = proj - ((p[right][k] - p[left][k]) / c.norm);
            OVec a = ov_mul(adjust, c.unit);
            gm[k].x -= a.x;
            gm[k].y -= a.y;
            gm[k].z -= a.z;
        }
        o_viscous(flux, pm, gm, e->property, m->faces.basis[i]);
        for (int j = 0; j < 5; j++) {
            r[left][j] -= flux[j] * m->faces.area[i];
            if (i < m->faces.num_inner) {
                r[right][j] += flux[j] * m->faces.area[i];
            }
        }
    }
}

/* src/equations/derivative.c:158-227 */
static void convective_o2(const OEqns *e, const double (*p)[5], double (*r)[5], OVec (*g)[5])
{
    const OMesh *m = e->mesh;
    exchange(m, &g[0][0].x, 15);
    double pl[5];
    double pr[5];
    double flux[5];
    for (int i = 0; i < m->faces.num; i++) {
        int left = m->faces.cell_idx[i].left;
        int right = m->faces.cell_idx[i].right;
        int boundary = (m->faces.num_inner <= i && i < m->faces.off_boundary);
        for (int k = 0; k < 5; k++) {
            pl[k] = p[left][k] + ov_dot(g[left][k], m->faces.offset[i].left);
            pr[k] = boundary ? p[right][k]
                             : (p[right][k] + ov_dot(g[right][k], m->faces.offset[i].right));
        }
        o_convective(e->flux, flux, pl, pr, e->property, m->faces.basis[i]);
        for (int j = 0; j < 5; j++) {
            r[left][j] += flux[j] * m->faces.area[i];
            if (i < m->faces.num_inner) {
                r[right][j] -= flux[j] * m->faces.area[i];
            }
        }
    }
}

/* src/equations/average.c:8-41: volume-weighted sum over the cells of an inner entity,
 * divided by the volume of the whole mesh (sum_volume), in cell order */
void o_average(const OEqns *e, const char *entity, const double (*p)[5], double *avg)
{
    const OMesh *m = e->mesh;
    memset(avg, 0, 5 * sizeof(double));
    for (int i = 0; i < m->entities.num_inner; i++) {
        if (!strcmp(m->entities.name[i], entity)) {
            for (int j = m->entities.cell_off[i]; j < m->entities.cell_off[i + 1]; j++) {
                for (int k = 0; k < 5; k++) {
                    avg[k] += m->cells.volume[j] * p[j][k];
                }
            }
            for (int k = 0; k < 5; k++) {
                avg[k] /= m->cells.sum_volume;
            }
            return;
        }
    }
    o_fatal("invalid entity (%s)", entity);
}

/* src/equations/norm.c:8-38 with the reference solution given per inner cell */
void o_norm(const OEqns *e, const double (*ref)[5], const double (*p)[5], double *norm)
{
    const OMesh *m = e->mesh;
    memset(norm, 0, 5 * sizeof(double));
    for (int i = 0; i < m->cells.num_inner; i++) {
        for (int j = 0; j < 5; j++) {
            double d = ref[i][j] - p[i][j];
            norm[j] += m->cells.volume[i] * (d * d);
        }
    }
    for (int j = 0; j < 5; j++) {
        norm[j] = sqrt(norm[j] / m->cells.sum_volume);
    }
}

/* src/euler/polar.c:22-83 without the printing and the file: after the boundary pass, the
 * force coefficients (x, y) of the entity's faces summed in face order and each face's
 * pressure coefficient; reference = free-stream primitive state */
void o_polar(const OEqns *e, int entity, double (*p)[5], const double *ref, double length,
             double *coef, double *pc)
{
    const OMesh *m = e->mesh;
    o_boundary(e, p);
    double dyn = ref[0] * (((ref[1] * ref[1]) + (ref[2] * ref[2])) + (ref[3] * ref[3])) / 2;
    coef[0] = coef[1] = 0;
    for (int i = m->entities.face_off[entity], j = 0; i < m->entities.face_off[entity + 1];
         i++, j++) {
        int left = m->faces.cell_idx[i].left;
        int right = m->faces.cell_idx[i].right;
        double area = m->faces.area[i];
        OVec normal = m->faces.basis[i].x;
        double pressure = (p[left][4] + p[right][4]) / 2;
        pc[j] = (pressure - ref[4]) / dyn;
        coef[0] += (pressure * area * normal.x) / (dyn * length);
        coef[1] += (pressure * area * normal.y) / (dyn * length);
    }
}

/* finalize_derivative, src/equations/derivative.c:229-267: R = -R/V, plus the source term of
 * the cell's primitive state when one is set.  The forcing source is the periodic-hill
 * driver's (run/periodic_hill/run.c:8-21): its Prepare hook averages the state over the
 * entity "domain" once per evaluation; per cell s1 = rho (U - <u>) / tau goes to the x
 * momentum and u s1 to the energy, the other components stay 0. */
static void finalize(const OEqns *e, const double (*p)[5], double (*r)[5])
{
    const OMesh *m = e->mesh;
    int ni = m->cells.num_inner;
    if (e->source == O_SOURCE_NONE) {
        for (int i = 0; i < ni; i++) {
            for (int j = 0; j < 5; j++) {
                r[i][j] = -r[i][j] / m->cells.volume[i];
            }
        }
        return;
    }
    double avg[5];
    o_average(e, "domain", p, avg);
    double src[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < ni; i++) {
        double s1 = p[i][0] * (e->forcing[0] - avg[1]) / e->forcing[1];
        src[1] = s1;
        src[4] = p[i][1] * s1;
        for (int j = 0; j < 5; j++) {
            r[i][j] = (-r[i][j] / m->cells.volume[i]) + src[j];
        }
    }
}

/* src/equations/derivative.c:269-305 */
void o_derivative(const OEqns *e, double (*p)[5], double (*r)[5])
{
    const OMesh *m = e->mesh;
    int ni = m->cells.num_inner;
    memset(r, 0, (size_t)ni * sizeof(*r));
    o_boundary(e, p);
    if (e->space_order == 1) {
        convective_o1(e, p, r);
    }
    else {
        OVec (*g)[5] = o_calloc(m->cells.num, sizeof(*g));
        o_gradient(e, p, g);
        if (e->viscous) {
            viscous_o2(e, (const double (*)[5])p, r, g);
        }
        if (e->limiter != O_LIM_NONE) {
            o_limiter(e, (const double (*)[5])p, g);
        }
        if (e->flux != O_FLUX_NONE) {
            convective_o2(e, (const double (*)[5])p, r, g);
        }
        free(g);
    }
    finalize(e, (const double (*)[5])p, r);
}

/* src/equations/residual.c:8-33 */
void o_residual(const OEqns *e, const double (*d)[5], double *res)
{
    const OMesh *m = e->mesh;
    memset(res, 0, 5 * sizeof(double));
    for (int i = 0; i < m->cells.num_inner; i++) {
        for (int j = 0; j < 5; j++) {
            res[j] += m->cells.volume[i] * (d[i][j] * d[i][j]);
        }
    }
    for (int j = 0; j < 5; j++) {
        res[j] = sqrt(res[j] / m->cells.sum_volume);
    }
}
