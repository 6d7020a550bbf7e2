/*
 * oracle.h -- CPU ORACLE (test infrastructure only).
 *
 * This directory is a plain-C, single-rank, FP64 restatement of the reference
 * solver's residual path (hheinzer/teal).  It exists only to check the HIP
 * product path: tests, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it; the product library never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Operation order
 * and association follow the reference exactly (compiled with
 * -ffp-contract=off), so results are bit-identical to the reference wherever
 * the reference itself is deterministic.
 *
 * Pinning (what checks each part; tests/ holds the checks):
 *   - mesh create, equations path (boundary, gradient, limiter, fluxes, viscous
 *     flux, derivative, time step, residual), integrators (euler, lserk, every
 *     order/stage pair, implicit_euler) and GMRES: bit-for-bit against the
 *     reference's own sources compiled into oracle/_ref (see oracle/Makefile)
 *     and against the golden fixtures written from that build
 *     (the .npz files under tests/golden); plus closed forms (flux consistency, supersonic
 *     upwinding, rotation and mirror invariance, Toro's exact-Riemann star
 *     states, BC invariants, viscous stress and Fourier heat flux, LSQ
 *     exactness on linear fields and the reference's convergence-order test,
 *     limiter bounds, free-stream preservation, discrete conservation, the
 *     uniform-hex time step).
 *   - dual graph (METIS_MeshToDual; METIS is third-party and absent here, its
 *     version unpinned by the reference Makefile): restated from the published
 *     METIS 5.1.0 CreateGraphDual / FindCommonElements rule; checked by brute
 *     force on small meshes and by the reference's own mesh_validate.
 *   - mesh generation (src/mesh/generate.c needs metis.h, so the reference
 *     build cannot compile it): the integer renumbering (connect_cells, the
 *     breadth-first cell order with its stable sort by descending degree,
 *     the node order, face creation and the stable face sort) against an
 *     independent step-by-step restatement on hex boxes and on mixed
 *     hex/prism meshes, whose cells differ in degree; geometry bit for bit
 *     against closed forms on boxes where every quantity the reference
 *     computes is exactly representable, and to 1e-14 on other boxes; closed
 *     cells and volume sums on jittered tets; the reference's mesh_validate.
 *   - parity unpinned: only the neighbour order real METIS would return
 *     (checked against the published rule, not against METIS itself).
 */
#pragma once

#include <stddef.h>

typedef struct {
    double x, y, z;
} OVec;

typedef struct {
    OVec left, right;
} OVecPair;

typedef struct {
    OVec unit;
    double norm;
} OVecNorm;

typedef struct {
    OVec x, y, z; /* row-major */
} OMat;

typedef struct {
    int left, right;
} OPair;

typedef struct {
    int *off, *idx;
} OGraph;

typedef char OString[128];

/* Field order and types follow src/mesh.h:13-72 (the arrays the algorithm is
 * defined over), so an OMesh can be handed to the reference build unchanged. */
typedef struct {
    int num;
    int num_inner;
    long *global;
    OVec *coord;
} ONodes;

typedef struct {
    int num;
    int num_inner;
    int off_boundary;
    int off_periodic;
    OGraph node;
    OGraph cell;
    double *volume, sum_volume;
    OVec *center;
    OVec *projection;
    OVec *offset;
} OCells;

typedef struct {
    int num;
    int num_inner;
    int off_boundary;
    OGraph node;
    OPair *cell_idx;
    double *area;
    OVec *center;
    OMat *basis;
    OVecPair *weight;
    OVecPair *offset;
    OVecNorm *correction;
} OFaces;

typedef struct {
    int num;
    int num_inner;
    int off_boundary;
    OString *name;
    int *cell_off;
    int *face_off;
    OMat *rotation;
    OVec *translation;
} OEntities;

typedef struct {
    int num;
    int (*tag)[2];
    int *rank;
    int *recv_off;
    OGraph send;
} ONeighbors;

typedef struct {
    ONodes nodes;
    OCells cells;
    OFaces faces;
    OEntities entities;
    ONeighbors neighbors;
} OMesh;

enum { O_MAX_CELL_NODES = 8, O_MAX_CELL_FACES = 6, O_MAX_FACE_NODES = 4 };

/* ---- math (src/core/vector.c, matrix.c, utils.c) ---- */
OVec ov_add(OVec a, OVec b);
OVec ov_sub(OVec a, OVec b);
OVec ov_mul(double s, OVec a);
OVec ov_div(OVec a, double s);
double ov_dot(OVec a, OVec b);
double ov_norm(OVec a);
double ov_norm2(OVec a);
OVec ov_cross(OVec a, OVec b);
OVec ov_abs(OVec a);
OVec om_vec(OMat m, OVec v);
OMat om_transpose(OMat m);
double om_det(OMat m);
int o_is_close(double a, double b);
void *o_calloc(size_t num, size_t size);
void o_fatal(const char *fmt, ...) __attribute__((noreturn, format(printf, 1, 2)));

/* ---- mesh (src/mesh/create.c, generate.c, reorder.c) ---- */
OMesh *omesh_create(const double *min_coord, const double *max_coord, const int *num_cells,
                    const int *periodic);
void omesh_generate(OMesh *mesh);
void omesh_destroy(OMesh *mesh);
int omesh_periodic_source(const OMesh *mesh, int ghost);

/* METIS 5.1.0 MeshToDual rule, restated (third-party; see header comment). */
void o_mesh_to_dual(int num_elems, int num_nodes, const long *eptr, const long *eind,
                    int ncommon, long **xadj, long **adjncy);

/* ---- equations (src/equations, src/euler, src/navierstokes) ---- */
enum { O_LIM_NONE = 0, O_LIM_BJ = 1, O_LIM_VK = 2 };
enum {
    O_BC_NONE = 0,
    O_BC_SYMMETRY,
    O_BC_SUPERSONIC_INFLOW,
    O_BC_SUPERSONIC_OUTFLOW,
    O_BC_SUBSONIC_INFLOW,
    O_BC_SUBSONIC_OUTFLOW,
    O_BC_PRESSURE_OUTFLOW,
    O_BC_FARFIELD,
    O_BC_ADIABATIC_WALL,
};
enum {
    O_FLUX_NONE = 0,
    O_FLUX_GODUNOV,
    O_FLUX_ROE,
    O_FLUX_HLL,
    O_FLUX_HLLC,
    O_FLUX_HLLE,
    O_FLUX_AUSMD,
    O_FLUX_AUSMDV,
    O_FLUX_LXF,
    O_FLUX_CENTRAL,
};

typedef struct {
    const OMesh *mesh;
    int navierstokes;    /* 1: viscous central flux + NS time step */
    int space_order;     /* 1 or 2 */
    int flux;            /* O_FLUX_* */
    int viscous;         /* 0/1 */
    int limiter;         /* O_LIM_* */
    double *eps2;        /* VK parameter per inner cell */
    double property[3];  /* gamma, mu, Pr */
    int num_bc;
    int *bc;             /* O_BC_* per boundary entity */
    double (*bc_ctx)[5]; /* context state per boundary entity */
    double (*primitive)[5];
    double (*conserved)[5];
    /* source term: O_SOURCE_FORCING is the bulk-velocity forcing of
     * run/periodic_hill/run.c:8-21 (velocity, relaxation time) */
    int source;
    double forcing[2];
} OEqns;

enum { O_SOURCE_NONE = 0, O_SOURCE_FORCING = 1 };

OEqns *oeqns_create(const OMesh *mesh, int navierstokes);
void oeqns_destroy(OEqns *eqns);
int o_flux_by_name(const char *name);
int o_bc_by_name(const char *name);
int o_limiter_by_name(const char *name);
void oeqns_set_limiter(OEqns *eqns, int kind, double coefficient);
void oeqns_set_bc(OEqns *eqns, int entity, int kind, const double *ctx);

double o_timestep(const OEqns *eqns, const double (*primitive)[5]);
void o_boundary(const OEqns *eqns, double (*primitive)[5]);
void o_gradient(const OEqns *eqns, double (*primitive)[5], OVec (*gradient)[5]);
void o_limiter(const OEqns *eqns, const double (*primitive)[5], OVec (*gradient)[5]);
void o_derivative(const OEqns *eqns, double (*primitive)[5], double (*derivative)[5]);
void o_residual(const OEqns *eqns, const double (*derivative)[5], double *residual);
void o_average(const OEqns *eqns, const char *entity, const double (*primitive)[5],
               double *average);
void o_norm(const OEqns *eqns, const double (*reference)[5], const double (*primitive)[5],
            double *norm);
void o_polar(const OEqns *eqns, int entity, double (*primitive)[5], const double *reference,
             double length, double *coef, double *pressure_c);

/* physics plugins (src/euler/convective.c, boundary.c, create.c, riemann.c; src/navierstokes) */
void o_convective(int kind, double *flux, const double *left, const double *right,
                  const double *property, OMat basis);
void o_viscous(double *flux, const double *primitive, const OVec *gradient,
               const double *property, OMat basis);
void o_bc(int kind, double *outer, const double *inner, const double *property, OMat basis,
          const double *ctx);
void o_prim2cons(double *conserved, const double *primitive, const double *property);
void o_cons2prim(double *primitive, const double *conserved, const double *property);
double o_euler_timestep(const double *primitive, const double *property, double volume,
                        OVec projection);
double o_ns_timestep(const double *primitive, const double *property, double volume,
                     OVec projection);
void o_riemann(const double *left, const double *right, double gamma, double loc, double *out);

/* ---- integrators (src/simulation/advance.c, gmres.c) ---- */
double o_euler_step(OEqns *eqns, double *time, double *residual, double max_step,
                    double courant);
double o_lserk(OEqns *eqns, double *time, double *residual, double max_step, double courant,
               int time_order, int num_stages);
double o_implicit_euler(OEqns *eqns, double *time, double *residual, double max_step,
                        double courant, double newton_tol, double krylov_tol, int krylov_dim);
void o_gmres(const OEqns *eqns, const double (*conserved)[5], const double (*derivative)[5],
             const double (*residual)[5], double (*increment)[5], double step, double norm,
             double fd_scale, double krylov_tol, int krylov_dim);
