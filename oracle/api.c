/* CPU ORACLE (test infrastructure only) -- flat entry points for the Python tests
 * (ctypes) and for the CPU-baseline leg of bench.py.  No algorithm lives here. */
#define _POSIX_C_SOURCE 199309L
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

static OMat mat_of(const double *b)
{
    return (OMat){{b[0], b[1], b[2]}, {b[3], b[4], b[5]}, {b[6], b[7], b[8]}};
}

OMesh *oracle_mesh_create(const double *min, const double *max, const int *n, const int *periodic)
{
    return omesh_create(min, max, n, periodic);
}

static void *dup(const void *src, size_t bytes)
{
    void *dst = o_calloc(1, bytes ? bytes : 1);
    if (src && bytes) {
        memcpy(dst, src, bytes);
    }
    return dst;
}

/* Copy of a created, not yet generated mesh of the same layout (nodes, cell->node graph,
 * entities, neighbour blocks), e.g. a tetrahedral mesh built by another library, so the
 * oracle can generate and own it. */
OMesh *oracle_mesh_copy_created(const OMesh *src)
{
    OMesh *m = o_calloc(1, sizeof(OMesh));
    int nn = src->nodes.num, nc = src->cells.num, ne = src->entities.num;
    int nb = src->neighbors.num;
    m->nodes.num = nn;
    m->nodes.num_inner = src->nodes.num_inner;
    m->nodes.global = dup(src->nodes.global, nn * sizeof(long));
    m->nodes.coord = dup(src->nodes.coord, nn * sizeof(OVec));
    m->cells.num = nc;
    m->cells.num_inner = src->cells.num_inner;
    m->cells.off_boundary = src->cells.off_boundary;
    m->cells.off_periodic = src->cells.off_periodic;
    m->cells.node.off = dup(src->cells.node.off, (nc + 1) * sizeof(int));
    m->cells.node.idx = dup(src->cells.node.idx, src->cells.node.off[nc] * sizeof(int));
    m->entities.num = ne;
    m->entities.num_inner = src->entities.num_inner;
    m->entities.off_boundary = src->entities.off_boundary;
    m->entities.name = dup(src->entities.name, ne * sizeof(OString));
    m->entities.cell_off = dup(src->entities.cell_off, (ne + 1) * sizeof(int));
    m->entities.rotation = dup(src->entities.rotation, ne * sizeof(OMat));
    m->entities.translation = dup(src->entities.translation, ne * sizeof(OVec));
    m->neighbors.num = nb;
    m->neighbors.tag = dup(src->neighbors.tag, (nb + 1) * sizeof(int[2]));
    m->neighbors.rank = dup(src->neighbors.rank, (nb + 1) * sizeof(int));
    m->neighbors.recv_off = dup(src->neighbors.recv_off, (nb + 1) * sizeof(int));
    m->neighbors.send.off = dup(src->neighbors.send.off, (nb + 1) * sizeof(int));
    int nsend = src->neighbors.send.off ? src->neighbors.send.off[nb] : 0;
    m->neighbors.send.idx = dup(src->neighbors.send.idx, (nsend + 1) * sizeof(int));
    return m;
}

/* A created (not yet generated) mesh from plain arrays: nodes, the cell->node graph of the
 * inner cells followed by one ghost cell per boundary face, and the entities (the inner
 * entity first, then one boundary entity per ghost block).  No periodic sides, no
 * neighbour blocks.  For test meshes that mesh_create cannot make (mixed elements). */
OMesh *oracle_mesh_from_cells(int num_nodes, const double *coord, int num_cells, int num_inner,
                              const int *node_off, const int *node_idx, int num_entities,
                              const char *names, const int *entity_cell_off)
{
    OMesh *m = o_calloc(1, sizeof(OMesh));
    m->nodes.num = num_nodes;
    m->nodes.num_inner = num_nodes;
    m->nodes.coord = dup(coord, num_nodes * sizeof(OVec));
    m->nodes.global = o_calloc(num_nodes, sizeof(long));
    for (int i = 0; i < num_nodes; i++) {
        m->nodes.global[i] = i;
    }
    m->cells.num = num_cells;
    m->cells.num_inner = num_inner;
    m->cells.off_boundary = num_cells;
    m->cells.off_periodic = num_cells;
    m->cells.node.off = dup(node_off, (num_cells + 1) * sizeof(int));
    m->cells.node.idx = dup(node_idx, node_off[num_cells] * sizeof(int));
    m->entities.num = num_entities;
    m->entities.num_inner = 1;
    m->entities.off_boundary = num_entities;
    m->entities.name = o_calloc(num_entities, sizeof(OString));
    for (int i = 0; i < num_entities; i++) {
        strncpy(m->entities.name[i], &names[i * sizeof(OString)], sizeof(OString) - 1);
    }
    m->entities.cell_off = dup(entity_cell_off, (num_entities + 1) * sizeof(int));
    m->entities.rotation = o_calloc(num_entities, sizeof(OMat));
    m->entities.translation = o_calloc(num_entities, sizeof(OVec));
    m->neighbors.tag = o_calloc(1, sizeof(int[2]));
    m->neighbors.rank = o_calloc(1, sizeof(int));
    m->neighbors.recv_off = o_calloc(1, sizeof(int));
    m->neighbors.recv_off[0] = num_cells;
    m->neighbors.send.off = o_calloc(1, sizeof(int));
    m->neighbors.send.idx = o_calloc(1, sizeof(int));
    return m;
}

void oracle_mesh_generate(OMesh *mesh) { omesh_generate(mesh); }
void oracle_mesh_destroy(OMesh *mesh) { omesh_destroy(mesh); }
int oracle_periodic_source(const OMesh *mesh, int ghost) { return omesh_periodic_source(mesh, ghost); }

void oracle_mesh_to_dual(int ne, int nn, const long *eptr, const long *eind, int ncommon,
                         long *xadj_out, long *adjncy_out, long cap)
{
    long *xadj;
    long *adjncy;
    o_mesh_to_dual(ne, nn, eptr, eind, ncommon, &xadj, &adjncy);
    memcpy(xadj_out, xadj, (ne + 1) * sizeof(long));
    long cnt = (xadj[ne] < cap) ? xadj[ne] : cap;
    memcpy(adjncy_out, adjncy, cnt * sizeof(long));
    free(xadj);
    free(adjncy);
}

OEqns *oracle_eqns_create(const OMesh *mesh, int navierstokes)
{
    return oeqns_create(mesh, navierstokes);
}

void oracle_eqns_destroy(OEqns *e) { oeqns_destroy(e); }

/* flux/limiter names as accepted by the reference (NULL disables) */
void oracle_eqns_config(OEqns *e, int space_order, const char *flux, int viscous,
                        const char *limiter, double coefficient, double gamma, double mu,
                        double prandtl)
{
    e->space_order = space_order;
    e->flux = flux ? o_flux_by_name(flux) : O_FLUX_NONE;
    e->viscous = viscous;
    e->property[0] = gamma;
    e->property[1] = mu;
    e->property[2] = prandtl;
    oeqns_set_limiter(e, o_limiter_by_name(limiter), coefficient);
}

void oracle_eqns_set_bc(OEqns *e, const char *entity, const char *name, const double *ctx)
{
    const OMesh *m = e->mesh;
    for (int i = 0; i < e->num_bc; i++) {
        if (!strcmp(m->entities.name[m->entities.num_inner + i], entity)) {
            oeqns_set_bc(e, i, o_bc_by_name(name), ctx);
            return;
        }
    }
    o_fatal("invalid entity (%s)", entity);
}

void oracle_eqns_set_forcing(OEqns *e, double velocity, double relaxation)
{
    e->source = O_SOURCE_FORCING;
    e->forcing[0] = velocity;
    e->forcing[1] = relaxation;
}

void oracle_average(const OEqns *e, const char *entity, const double *p, double *avg)
{
    o_average(e, entity, (const double (*)[5])p, avg);
}

void oracle_norm(const OEqns *e, const double *ref, const double *p, double *norm)
{
    o_norm(e, (const double (*)[5])ref, (const double (*)[5])p, norm);
}

void oracle_polar(const OEqns *e, int entity, double *p, const double *ref, double length,
                  double *coef, double *pc)
{
    o_polar(e, entity, (double (*)[5])p, ref, length, coef, pc);
}

double *oracle_eqns_primitive(OEqns *e) { return &e->primitive[0][0]; }
double *oracle_eqns_conserved(OEqns *e) { return &e->conserved[0][0]; }
double *oracle_eqns_eps2(OEqns *e) { return e->eps2; }

void oracle_boundary(const OEqns *e, double *p) { o_boundary(e, (double (*)[5])p); }

void oracle_gradient(const OEqns *e, double *p, double *g)
{
    o_gradient(e, (double (*)[5])p, (OVec (*)[5])g);
}

void oracle_limiter(const OEqns *e, const double *p, double *g)
{
    o_limiter(e, (const double (*)[5])p, (OVec (*)[5])g);
}

void oracle_derivative(const OEqns *e, double *p, double *d)
{
    o_derivative(e, (double (*)[5])p, (double (*)[5])d);
}

double oracle_timestep(const OEqns *e, const double *p)
{
    return o_timestep(e, (const double (*)[5])p);
}

void oracle_residual(const OEqns *e, const double *d, double *res)
{
    o_residual(e, (const double (*)[5])d, res);
}

double oracle_euler(OEqns *e, double *time, double *res, double max_step, double courant)
{
    return o_euler_step(e, time, res, max_step, courant);
}

double oracle_lserk(OEqns *e, double *time, double *res, double max_step, double courant,
                    int time_order, int num_stages)
{
    return o_lserk(e, time, res, max_step, courant, time_order, num_stages);
}

double oracle_implicit_euler(OEqns *e, double *time, double *res, double max_step,
                             double courant, double newton_tol, double krylov_tol, int dim)
{
    return o_implicit_euler(e, time, res, max_step, courant, newton_tol, krylov_tol, dim);
}

void oracle_gmres(const OEqns *e, const double *u, const double *d, const double *r, double *inc,
                  double step, double norm, double fd_scale, double krylov_tol, int dim)
{
    o_gmres(e, (const double (*)[5])u, (const double (*)[5])d, (const double (*)[5])r,
            (double (*)[5])inc, step, norm, fd_scale, krylov_tol, dim);
}

void oracle_flux(const char *name, double *flux, const double *l, const double *r,
                 const double *property, const double *basis)
{
    o_convective(o_flux_by_name(name), flux, l, r, property, mat_of(basis));
}

void oracle_viscous(double *flux, const double *p, const double *g, const double *property,
                    const double *basis)
{
    o_viscous(flux, p, (const OVec *)g, property, mat_of(basis));
}

void oracle_bc(const char *name, double *outer, const double *inner, const double *property,
               const double *basis, const double *ctx)
{
    o_bc(o_bc_by_name(name), outer, inner, property, mat_of(basis), ctx);
}

void oracle_riemann(const double *l, const double *r, double gamma, double loc, double *out)
{
    o_riemann(l, r, gamma, loc, out);
}

void oracle_prim2cons(double *u, const double *p, const double *property)
{
    o_prim2cons(u, p, property);
}

void oracle_cons2prim(double *p, const double *u, const double *property)
{
    o_cons2prim(p, u, property);
}

/* wall-clock helper for the CPU baseline (monotonic seconds) */
double oracle_wtime(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + (1e-9 * (double)ts.tv_nsec);
}
