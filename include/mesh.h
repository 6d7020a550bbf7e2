// mesh.h -- unstructured mesh data model and constructors of the drop-in library.
// Replaces the reference interface src/mesh.h:1-99.  The struct layout is the reference's
// (drivers read and modify these arrays directly, e.g. run/periodic_hill/run.c:30-34).
// One process owns the whole mesh: periodic sides are self-neighbour ghost blocks
// (src/mesh/create.c:388-440) and there are no rank-halo cells.
#pragma once

#include "matrix.h"
#include "utils.h"
#include "vector.h"

#ifdef __cplusplus
extern "C" {
#endif

enum MeshLimits {
    MAX_CELL_NODES = 8,
    MAX_CELL_FACES = 6,
    MAX_FACE_NODES = 4,
};

typedef struct {
    int num;
    int num_inner;
    long *global;
    Vector *coord;
} MeshNodes;

// cells: [inner | boundary ghosts | periodic ghosts] (src/mesh.h:20-31)
typedef struct {
    int num;
    int num_inner;
    int off_boundary;
    int off_periodic;
    Graph node;
    Graph cell;
    double *volume, sum_volume;
    Vector *center;
    Vector *projection;
    Vector *offset;
} MeshCells;

// faces: [inner | boundary | periodic], left < right, basis.x = unit normal left->right
typedef struct {
    int num;
    int num_inner;
    int off_boundary;
    Graph node;
    Pair *cell_idx;
    double *area;
    Vector *center;
    Matrix *basis;
    VectorPair *weight;
    VectorPair *offset;
    VectorNorm *correction;
} MeshFaces;

typedef struct {
    int num;
    int num_inner;
    int off_boundary;
    String *name;
    int *cell_off;
    int *face_off;
    Matrix *rotation;
    Vector *translation;
} MeshEntities;

typedef struct {
    int num;
    int (*tag)[2];
    int *rank;
    int *recv_off;
    Graph send;
} MeshNeighbors;

typedef struct {
    MeshNodes nodes;
    MeshCells cells;
    MeshFaces faces;
    MeshEntities entities;
    MeshNeighbors neighbors;
} Mesh;

// Cartesian hex mesh with optional per-axis periodicity (src/mesh.h:75, create.c:509-544).
Mesh *mesh_create(Vector min_coord, Vector max_coord, Triple num_cells, Triple periodic);

// Gmsh 4.1 input, ASCII or binary, with periodic links (src/mesh.h:78, src/mesh/read.c:901-918,
// src/mesh/gmsh.c:135-904), read on one process.  Errors end the program through teal_error.
Mesh *mesh_read(const char *fname);

// Modify all node coordinates (src/mesh.h:81, src/mesh/modify.c:5-11).
void mesh_modify(Mesh *mesh, void (*modify)(Vector *coord));

// Split an entity by a plane (src/mesh.h:84, src/mesh/split.c:88-94).
void mesh_split(Mesh *mesh, const char *entity, Vector root, Vector normal);

// Connectivity (dual graph, BFS renumbering, faces) and geometry (src/mesh.h:87,
// src/mesh/generate.c:779-815).
void mesh_generate(Mesh *mesh);

// Invariant checks; teal_error on failure (src/mesh.h:90, src/mesh/validate.c:257-265).
void mesh_validate(const Mesh *mesh);

// Print a mesh summary (src/mesh.h:93, src/mesh/summary.c:7-29).
void mesh_summary(const Mesh *mesh);

// VTKHDF output (src/mesh.h:96, src/mesh/write.c:185-211): <name>_mesh.hdf.
void mesh_write(const Mesh *mesh, const char *name);

// Release all mesh resources (src/mesh.h:99).
void mesh_destroy(Mesh *mesh);

// Tetrahedral test mesh (no counterpart in the reference, which reads tets from Gmsh
// files): an n^3 hex grid on [0,1]^3, each hex split into 6 tets, interior nodes jittered
// by +-jitter*h with a seeded generator; boundary sides named as in mesh_create.
Mesh *mesh_create_tets(int n, double jitter, unsigned long long seed);

// Source cell of each periodic ghost of a one-process mesh (the exchange pairing of
// src/equations/exchange.c:19-33): src[ghost] = source cell for every ghost that receives one;
// other entries are left unchanged.  No counterpart in the reference (its exchange is MPI).
void teal_mesh_periodic_sources(const Mesh *mesh, int *src);

#ifdef __cplusplus
}
#endif
