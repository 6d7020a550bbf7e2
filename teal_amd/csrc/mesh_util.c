// mesh_util.c -- mesh checks, summary, modification, splitting, destruction and the
// tetrahedral test-mesh generator.
//   mesh_validate  src/mesh/validate.c:257-265 (same invariants; point uniqueness by a
//                  sorted sweep instead of the reference's kd-tree)
//   mesh_summary   src/mesh/summary.c:7-29
//   mesh_modify    src/mesh/modify.c:5-11
//   mesh_split     src/mesh/split.c:88-94
//   mesh_destroy   src/mesh/destroy.c:6-45
#include <assert.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mesh.h"
#include "teal.h"

void teal_mesh_renumber_cells(Mesh *mesh, const int *map, int beg, int end);

#define ensure(expr) \
    ((expr) ? (void)0 : teal_error("`%s` failed (%s:%d: %s)", #expr, __FILE__, __LINE__, __func__))

static const Vector *sort_points;

static int compare_point_x(const void *lhs, const void *rhs)
{
    double a = sort_points[*(const int *)lhs].x;
    double b = sort_points[*(const int *)rhs].x;
    return (a > b) - (a < b);
}

// no two points may coincide within is_close in every coordinate (validate.c:8-25)
static void validate_unique(const Vector *point, int num)
{
    int *order = teal_calloc(num, sizeof(int));
    for (int i = 0; i < num; i++) {
        order[i] = i;
    }
    sort_points = point;
    qsort(order, num, sizeof(int), compare_point_x);
    for (int a = 0; a < num; a++) {
        Vector p = point[order[a]];
        for (int b = a + 1; b < num; b++) {
            Vector q = point[order[b]];
            if (!is_close(p.x, q.x)) {
                break;
            }
            ensure(!(is_close(p.y, q.y) && is_close(p.z, q.z)));
        }
    }
    teal_free(order);
}

static void validate_graph(Graph g, int num, int min_deg, int max_deg, int min_idx, int max_idx)
{
    ensure(g.off[0] == 0);
    for (int i = 0; i < num; i++) {
        ensure(g.off[i] <= g.off[i + 1]);
        if (min_deg >= 0) {
            ensure(g.off[i + 1] - g.off[i] >= min_deg);
        }
        if (max_deg >= 0) {
            ensure(g.off[i + 1] - g.off[i] <= max_deg);
        }
        for (int j = g.off[i]; j < g.off[i + 1]; j++) {
            ensure(min_idx <= g.idx[j] && g.idx[j] <= max_idx);
            for (int k = j + 1; k < g.off[i + 1]; k++) {
                ensure(g.idx[j] != g.idx[k]);
            }
        }
    }
}

static int adjacent(Graph g, int lhs, int rhs)
{
    for (int i = g.off[lhs]; i < g.off[lhs + 1]; i++) {
        if (g.idx[i] == rhs) {
            return 1;
        }
    }
    return 0;
}

void mesh_validate(const Mesh *mesh)
{
    assert(mesh);
    // nodes (validate.c:27-60)
    ensure(mesh->nodes.num > 0);
    ensure(mesh->nodes.num_inner >= 0 && mesh->nodes.num_inner <= mesh->nodes.num);
    for (int i = 0; i < mesh->nodes.num; i++) {
        ensure(mesh->nodes.global[i] >= 0 && mesh->nodes.global[i] < mesh->nodes.num);
    }
    validate_unique(mesh->nodes.coord, mesh->nodes.num);

    // cells (validate.c:93-122)
    const MeshCells *c = &mesh->cells;
    ensure(c->num > 0 && c->num_inner > 0);
    ensure(c->num_inner <= c->off_boundary && c->off_boundary <= c->off_periodic &&
           c->off_periodic <= c->num);
    validate_graph(c->node, c->num, 3, MAX_CELL_NODES, 0, mesh->nodes.num - 1);
    validate_graph(c->cell, c->num, 1, MAX_CELL_FACES, 0, c->num - 1);
    for (int i = 0; i < c->num; i++) {
        for (int j = c->cell.off[i]; j < c->cell.off[i + 1]; j++) {
            ensure(c->cell.idx[j] != i);
            ensure(adjacent(c->cell, c->cell.idx[j], i));
        }
        if (i < c->num_inner) {
            ensure(!is_close(c->volume[i], 0) && c->volume[i] > 0);
        }
        else {
            ensure(is_close(c->volume[i], 0));
        }
    }
    validate_unique(c->center, c->num);

    // faces (validate.c:124-170)
    const MeshFaces *f = &mesh->faces;
    ensure(f->num > 0 && f->num_inner >= 0);
    ensure(f->num_inner <= f->off_boundary && f->off_boundary <= f->num);
    validate_graph(f->node, f->num, 3, MAX_FACE_NODES, 0, mesh->nodes.num - 1);
    for (int i = 0; i < f->num; i++) {
        int l = f->cell_idx[i].left;
        int r = f->cell_idx[i].right;
        ensure(l < r && 0 <= l && l < c->num_inner);
        if (i < f->num_inner) {
            ensure(0 <= r && r < c->num_inner);
        }
        else if (i < f->off_boundary) {
            ensure(c->num_inner <= r && r < c->off_boundary);
        }
        else {
            ensure(c->off_boundary <= r && r < c->num);
        }
        ensure(adjacent(c->cell, l, r) && adjacent(c->cell, r, l));
        ensure(!is_close(f->area[i], 0) && f->area[i] > 0);
        Matrix b = f->basis[i];
        ensure(is_close(vector_norm(b.x), 1) && is_close(vector_norm(b.y), 1) &&
               is_close(vector_norm(b.z), 1));
        ensure(is_close(vector_dot(b.x, b.y), 0) && is_close(vector_dot(b.y, b.z), 0) &&
               is_close(vector_dot(b.z, b.x), 0));
        double proj = vector_dot(b.x, vector_sub(c->center[r], c->center[l]));
        ensure(!is_close(proj, 0) && proj > 0);
    }
    validate_unique(f->center, f->num);

    // entities (validate.c:172-216)
    const MeshEntities *e = &mesh->entities;
    ensure(e->num > 0 && e->num_inner > 0);
    ensure(e->num_inner <= e->off_boundary && e->off_boundary <= e->num);
    ensure(e->cell_off[0] == 0);
    for (int i = 0; i < e->num; i++) {
        ensure(strlen(e->name[i]) > 0);
        ensure(e->cell_off[i] <= e->cell_off[i + 1]);
        if (i < e->num_inner) {
            ensure(e->face_off[i] == f->num_inner);
        }
        ensure(e->face_off[i] <= e->face_off[i + 1]);
    }
    ensure(e->cell_off[e->num] == c->off_periodic);
    ensure(e->cell_off[e->num_inner] == c->num_inner);
    ensure(e->cell_off[e->off_boundary] == c->off_boundary);
    ensure(e->face_off[e->num] <= f->num);
    ensure(e->face_off[e->num_inner] == f->num_inner);
    ensure(e->face_off[e->off_boundary] == f->off_boundary);
    for (int i = e->num_inner; i < e->num; i++) {
        ensure(e->cell_off[i + 1] - e->cell_off[i] == e->face_off[i + 1] - e->face_off[i]);
    }
    for (int i = 0; i < e->num; i++) {
        if (i < e->off_boundary) {
            ensure(is_close(matrix_determinant(e->rotation[i]), 0));
            ensure(is_close(vector_norm(e->translation[i]), 0));
        }
        else {
            ensure(is_close(matrix_determinant(e->rotation[i]), 1));
        }
    }

    // neighbours (validate.c:218-255): one process, every exchange is with itself
    const MeshNeighbors *n = &mesh->neighbors;
    ensure(n->num >= 0);
    ensure(n->recv_off[0] == c->off_boundary);
    for (int i = 0; i < n->num; i++) {
        ensure(n->rank[i] == 0);
        ensure(n->recv_off[i] <= n->recv_off[i + 1]);
    }
    ensure(n->recv_off[n->num] == c->num);
    validate_graph(n->send, n->num, -1, -1, 0, c->num_inner - 1);
}

void mesh_summary(const Mesh *mesh)
{
    assert(mesh);
    teal_print("Mesh summary");
    teal_print("\t number of inner nodes : %ld", (long)mesh->nodes.num_inner);
    teal_print("\t number of inner cells : %ld", (long)mesh->cells.num_inner);
    teal_print("\t number of sync cells  : %ld", 0L);
}

void mesh_modify(Mesh *mesh, void (*modify)(Vector *coord))
{
    assert(mesh && modify);
    for (int i = 0; i < mesh->nodes.num; i++) {
        modify(&mesh->nodes.coord[i]);
    }
}

// split entity `entity` into "-a" (centroid on the negative side of the plane, or on it)
// and "-b" (src/mesh/split.c:20-94)
void mesh_split(Mesh *mesh, const char *entity, Vector root, Vector normal)
{
    assert(mesh && entity);
    int index = -1;
    for (int i = 0; i < mesh->entities.off_boundary; i++) {
        if (!strcmp(mesh->entities.name[i], entity)) {
            index = i;
            break;
        }
    }
    if (index < 0) {
        teal_error("invalid entity (%s)", entity);
    }
    int beg = mesh->entities.cell_off[index];
    int end = mesh->entities.cell_off[index + 1];
    int num = end - beg;
    int *map = teal_calloc(num + 1, sizeof(int));
    int lhs_count = 0;
    for (int i = beg; i < end; i++) {
        Vector mean = {0, 0, 0};
        for (int j = mesh->cells.node.off[i]; j < mesh->cells.node.off[i + 1]; j++) {
            vector_iadd(&mean, mesh->nodes.coord[mesh->cells.node.idx[j]]);
        }
        vector_idiv(&mean, mesh->cells.node.off[i + 1] - mesh->cells.node.off[i]);
        map[i - beg] = (vector_dot(vector_sub(mean, root), normal) <= 0);
        lhs_count += map[i - beg];
    }
    int lhs = 0;
    int rhs = 0;
    for (int i = 0; i < num; i++) {
        map[i] = map[i] ? lhs++ : (lhs_count + rhs++);
    }
    teal_mesh_renumber_cells(mesh, map, beg, end);
    teal_free(map);

    int ne = mesh->entities.num + 1;
    int tail = mesh->entities.num - index;
    mesh->entities.name = teal_realloc(mesh->entities.name, ne, sizeof(String));
    memmove(&mesh->entities.name[index + 1], &mesh->entities.name[index], tail * sizeof(String));
    strcat(mesh->entities.name[index], "-a");
    strcat(mesh->entities.name[index + 1], "-b");
    mesh->entities.cell_off = teal_realloc(mesh->entities.cell_off, ne + 1, sizeof(int));
    memmove(&mesh->entities.cell_off[index + 2], &mesh->entities.cell_off[index + 1],
            tail * sizeof(int));
    mesh->entities.cell_off[index + 1] = mesh->entities.cell_off[index] + lhs_count;
    mesh->entities.rotation = teal_realloc(mesh->entities.rotation, ne, sizeof(Matrix));
    memmove(&mesh->entities.rotation[index + 1], &mesh->entities.rotation[index],
            tail * sizeof(Matrix));
    mesh->entities.translation = teal_realloc(mesh->entities.translation, ne, sizeof(Vector));
    memmove(&mesh->entities.translation[index + 1], &mesh->entities.translation[index],
            tail * sizeof(Vector));
    mesh->entities.num = ne;
    if (index < mesh->entities.num_inner) {
        mesh->entities.num_inner += 1;
    }
    mesh->entities.off_boundary += 1;
}

void mesh_destroy(Mesh *mesh)
{
    assert(mesh);
    void *ptrs[] = {mesh->nodes.global,      mesh->nodes.coord,        mesh->cells.node.off,
                    mesh->cells.node.idx,    mesh->cells.cell.off,     mesh->cells.cell.idx,
                    mesh->cells.volume,      mesh->cells.center,       mesh->cells.projection,
                    mesh->cells.offset,      mesh->faces.node.off,     mesh->faces.node.idx,
                    mesh->faces.cell_idx,    mesh->faces.area,         mesh->faces.center,
                    mesh->faces.basis,       mesh->faces.weight,       mesh->faces.offset,
                    mesh->faces.correction,  mesh->entities.name,      mesh->entities.cell_off,
                    mesh->entities.face_off, mesh->entities.rotation,  mesh->entities.translation,
                    mesh->neighbors.tag,     mesh->neighbors.rank,     mesh->neighbors.recv_off,
                    mesh->neighbors.send.off, mesh->neighbors.send.idx};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) {
        teal_free(ptrs[i]);
    }
    teal_free(mesh);
}

// The ghost block of neighbour entry e receives the send list of the entry whose send tag is
// e's receive tag (src/equations/exchange.c:19-33, src/mesh/create.c:409-419): src[ghost] =
// its source cell; the other entries of src are left as they are.
void teal_mesh_periodic_sources(const Mesh *mesh, int *src)
{
    const MeshNeighbors *nb = &mesh->neighbors;
    for (int e = 0; e < nb->num; e++) {
        int s = -1;
        for (int k = 0; k < nb->num; k++) {
            if (nb->tag[k][1] == nb->tag[e][0]) {
                s = k;
            }
        }
        if (s < 0) {
            continue;
        }
        for (int j = nb->recv_off[e]; j < nb->recv_off[e + 1]; j++) {
            src[j] = nb->send.idx[nb->send.off[s] + (j - nb->recv_off[e])];
        }
    }
}

// ------------------------------------------------------------------ tetrahedral mesh
static uint64_t splitmix64(uint64_t *state)
{
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static double uniform(uint64_t *state)  // [0, 1)
{
    return (double)(splitmix64(state) >> 11) * 0x1.0p-53;
}

static double signed_volume(Vector a, Vector b, Vector c, Vector d)
{
    return vector_dot(vector_sub(b, a), vector_cross(vector_sub(c, a), vector_sub(d, a))) / 6;
}

// The n^3 unit-cube hexes split into 6 tetrahedra each around the (0,0,0)-(1,1,1) diagonal
// of every hex (Kuhn/Freudenthal split, conforming across hexes).  Interior nodes move by
// U(-jitter h, +jitter h) per coordinate; a node whose move would flatten one of its tets
// below a quarter of its undisturbed volume is redrawn at half the amplitude (seeded, so
// the mesh is reproducible).  Boundary faces become triangle ghost cells grouped into the
// entities left/right/bottom/top/back/front as in mesh_create.
Mesh *mesh_create_tets(int n, double jitter, unsigned long long seed)
{
    if (n < 1) {
        teal_error("invalid tet mesh size (%d)", n);
    }
    const int np = n + 1;
    const int num_nodes = np * np * np;
    const double h = 1.0 / n;
#define NID(i, j, k) ((i) + (np * ((j) + (np * (k)))))
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    Mesh *mesh = teal_calloc(1, sizeof(*mesh));
    mesh->nodes.num = num_nodes;
    mesh->nodes.num_inner = num_nodes;
    mesh->nodes.global = teal_calloc(num_nodes, sizeof(long));
    mesh->nodes.coord = teal_calloc(num_nodes, sizeof(Vector));
    for (int k = 0; k < np; k++) {
        for (int j = 0; j < np; j++) {
            for (int i = 0; i < np; i++) {
                int id = NID(i, j, k);
                mesh->nodes.global[id] = id;
                mesh->nodes.coord[id] = (Vector){i * h, j * h, k * h};
            }
        }
    }
    const int num_inner = 6 * n * n * n;
    const int num_ghost = 2 * 6 * n * n;
    const int nc = num_inner + num_ghost;
    int *off = teal_calloc(nc + 1, sizeof(int));
    int *idx = teal_calloc((4 * num_inner) + (3 * num_ghost), sizeof(int));
    int num = 0;
    for (int k = 0; k < n; k++) {
        for (int j = 0; j < n; j++) {
            for (int i = 0; i < n; i++) {
                for (int t = 0; t < 6; t++) {
                    int p[3] = {i, j, k};
                    int *q = &idx[off[num]];
                    q[0] = NID(p[0], p[1], p[2]);
                    for (int s = 0; s < 3; s++) {
                        p[perm[t][s]] += 1;
                        q[s + 1] = NID(p[0], p[1], p[2]);
                    }
                    off[num + 1] = off[num] + 4;
                    num += 1;
                }
            }
        }
    }
    // jitter the interior nodes
    uint64_t state = seed;
    if (jitter > 0) {
        for (int k = 1; k < n; k++) {
            for (int j = 1; j < n; j++) {
                for (int i = 1; i < n; i++) {
                    Vector *c = &mesh->nodes.coord[NID(i, j, k)];
                    c->x += (2 * uniform(&state) - 1) * jitter * h;
                    c->y += (2 * uniform(&state) - 1) * jitter * h;
                    c->z += (2 * uniform(&state) - 1) * jitter * h;
                }
            }
        }
        const double vmin = h * h * h / 24;  // a quarter of the undisturbed tet volume h^3/6
        static const double parity[6] = {1, -1, -1, 1, 1, -1};  // sign of each axis permutation
        for (int round = 0; round < 64; round++) {
            int bad = 0;
            for (int c = 0; c < num_inner; c++) {
                const int *q = &idx[off[c]];
                Vector *x = mesh->nodes.coord;
                if (signed_volume(x[q[0]], x[q[1]], x[q[2]], x[q[3]]) * parity[c % 6] >= vmin) {
                    continue;
                }
                bad += 1;
                for (int s = 0; s < 4; s++) {
                    int id = q[s];
                    int i = id % np;
                    int j = (id / np) % np;
                    int k = id / (np * np);
                    if (i == 0 || j == 0 || k == 0 || i == n || j == n || k == n) {
                        continue;
                    }
                    double amp = jitter * h * ldexp(1.0, -(round + 1));
                    x[id].x = (i * h) + ((2 * uniform(&state) - 1) * amp);
                    x[id].y = (j * h) + ((2 * uniform(&state) - 1) * amp);
                    x[id].z = (k * h) + ((2 * uniform(&state) - 1) * amp);
                }
            }
            if (!bad) {
                break;
            }
            if (round == 63) {
                teal_error("tet mesh: could not untangle the jittered nodes");
            }
        }
    }
    // boundary triangles: each boundary quad (a, b, c, d) with diagonal a -> d gives
    // (a, b, d) and (a, c, d), the faces of the adjacent Kuhn tets
    static const char *names[] = {"left", "right", "bottom", "top", "back", "front"};
    int ne = 7;
    mesh->entities.num = ne;
    mesh->entities.num_inner = 1;
    mesh->entities.off_boundary = ne;
    mesh->entities.name = teal_calloc(ne, sizeof(String));
    mesh->entities.cell_off = teal_calloc(ne + 1, sizeof(int));
    mesh->entities.rotation = teal_calloc(ne, sizeof(Matrix));
    mesh->entities.translation = teal_calloc(ne, sizeof(Vector));
    strcpy(mesh->entities.name[0], "domain");
    mesh->entities.cell_off[1] = num_inner;
    for (int side = 0; side < 6; side++) {
        strcpy(mesh->entities.name[side + 1], names[side]);
        int axis = side / 2;
        int fix = (side % 2) ? n : 0;
        int u = (axis + 1) % 3;
        int v = (axis + 2) % 3;
        for (int a = 0; a < n; a++) {
            for (int b = 0; b < n; b++) {
                int p[3];
                p[axis] = fix;
                int corner[4];
                for (int s = 0; s < 4; s++) {
                    p[u] = a + (s & 1);
                    p[v] = b + (s >> 1);
                    corner[s] = NID(p[0], p[1], p[2]);
                }
                int tri[2][3] = {{corner[0], corner[1], corner[3]}, {corner[0], corner[2], corner[3]}};
                for (int t = 0; t < 2; t++) {
                    memcpy(&idx[off[num]], tri[t], 3 * sizeof(int));
                    off[num + 1] = off[num] + 3;
                    num += 1;
                }
            }
        }
        mesh->entities.cell_off[side + 2] = num;
    }
#undef NID
    assert(num == nc);
    mesh->cells.num = nc;
    mesh->cells.num_inner = num_inner;
    mesh->cells.off_boundary = nc;
    mesh->cells.off_periodic = nc;
    mesh->cells.node.off = off;
    mesh->cells.node.idx = idx;
    mesh->neighbors.num = 0;
    mesh->neighbors.tag = teal_calloc(1, sizeof(int[2]));
    mesh->neighbors.rank = teal_calloc(1, sizeof(int));
    mesh->neighbors.recv_off = teal_calloc(1, sizeof(int));
    mesh->neighbors.recv_off[0] = nc;
    mesh->neighbors.send.off = teal_calloc(1, sizeof(int));
    mesh->neighbors.send.idx = teal_calloc(1, sizeof(int));
    return mesh;
}
