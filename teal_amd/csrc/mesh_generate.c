// mesh_generate.c -- connectivity and geometry of a mesh (src/mesh/generate.c:779-815).
//
// Connectivity (integer work, bit-exact with the reference):
//   1. dual graph: cells sharing >= 3 nodes (ghost cells get one private dummy node so two
//      ghosts never connect; ghosts keep only inner neighbours).  The neighbour order is the
//      one METIS_MeshToDual produces (src/mesh/generate.c:39-40; METIS 5.1.0
//      CreateGraphDual/FindCommonElements): first encounter walking the cell's nodes in
//      order and, per node, its cells by increasing index.
//   2. inner cells renumbered breadth-first from cell 0, each cell's undiscovered
//      neighbours taken in descending degree, ties in list order (generate.c:68-133)
//   3. nodes renumbered by first use (generate.c:171-204)
//   4. faces for every pair left < right, inner faces ordered by left cell, the rest by
//      right (ghost) cell, ties in creation order (generate.c:206-321)
// Geometry (FP64, same formulas and association as generate.c:336-777).
#include <assert.h>
#include <math.h>
#include <string.h>

#include "mesh.h"
#include "teal.h"

// ------------------------------------------------------------------ dual graph
static void dual_graph(Mesh *mesh)
{
    const int nc = mesh->cells.num;
    const int ni = mesh->cells.num_inner;
    const int *noff = mesh->cells.node.off;
    const int *nidx = mesh->cells.node.idx;
    const int num_nodes = mesh->nodes.num;

    // node -> cells, each list in increasing cell order (ghost dummy nodes are private to
    // one cell and so never create an adjacency: they are left out of these lists)
    int *cnt = teal_calloc(num_nodes + 1, sizeof(int));
    for (int j = 0; j < noff[nc]; j++) {
        cnt[nidx[j] + 1] += 1;
    }
    for (int i = 0; i < num_nodes; i++) {
        cnt[i + 1] += cnt[i];
    }
    int *cells_of = teal_calloc(noff[nc] + 1, sizeof(int));
    int *fill = teal_calloc(num_nodes, sizeof(int));
    for (int c = 0; c < nc; c++) {
        for (int j = noff[c]; j < noff[c + 1]; j++) {
            int node = nidx[j];
            cells_of[cnt[node] + fill[node]++] = c;
        }
    }
    teal_free(fill);

    int *shared = teal_calloc(nc, sizeof(int));
    int *seen = teal_calloc(nc, sizeof(int));
    int *off = teal_calloc(nc + 1, sizeof(int));
    int *idx = teal_calloc(nc * MAX_CELL_FACES, sizeof(int));
    for (int c = 0; c < nc; c++) {
        // element length as METIS sees it: ghosts carry one extra (dummy) node
        int len_c = (noff[c + 1] - noff[c]) + (c >= ni);
        int k = 0;
        for (int j = noff[c]; j < noff[c + 1]; j++) {
            int node = nidx[j];
            for (int q = cnt[node]; q < cnt[node + 1]; q++) {
                int other = cells_of[q];
                if (shared[other]++ == 0) {
                    seen[k++] = other;
                }
            }
        }
        off[c + 1] = off[c];
        for (int q = 0; q < k; q++) {
            int other = seen[q];
            int common = shared[other];
            shared[other] = 0;
            if (other == c) {
                continue;
            }
            int len_o = (noff[other + 1] - noff[other]) + (other >= ni);
            int adjacent = common >= 3 || common >= len_c - 1 || common >= len_o - 1;
            if (adjacent && (c < ni || other < ni)) {
                if (off[c + 1] - off[c] >= MAX_CELL_FACES) {
                    teal_error("cell %d has more than %d neighbours", c, MAX_CELL_FACES);
                }
                idx[off[c + 1]++] = other;
            }
        }
    }
    mesh->cells.cell.off = off;
    mesh->cells.cell.idx = teal_realloc(idx, off[nc] + 1, sizeof(int));
    teal_free(shared);
    teal_free(seen);
    teal_free(cnt);
    teal_free(cells_of);
}

// move cell i (beg <= i < end) to beg + map[i - beg]; adjacency lists keep their order and
// cell references are renamed (src/mesh/reorder.c:76-135)
static void renumber_cells(Mesh *mesh, const int *map, int beg, int end)
{
    int num = end - beg;
    int *noff = mesh->cells.node.off;
    int *coff = mesh->cells.cell.off;
    int *inv = teal_calloc(num, sizeof(int));
    for (int i = 0; i < num; i++) {
        inv[map[i]] = i;
    }
    int *nodes = teal_calloc(noff[end] - noff[beg] + 1, sizeof(int));
    int *cells = coff ? teal_calloc(coff[end] - coff[beg] + 1, sizeof(int)) : 0;
    int *new_noff = teal_calloc(num + 1, sizeof(int));
    int *new_coff = teal_calloc(num + 1, sizeof(int));
    for (int i = 0; i < num; i++) {
        int old = beg + inv[i];
        int nn = noff[old + 1] - noff[old];
        memcpy(&nodes[new_noff[i]], &mesh->cells.node.idx[noff[old]], nn * sizeof(int));
        new_noff[i + 1] = new_noff[i] + nn;
        if (coff) {
            int ncc = coff[old + 1] - coff[old];
            memcpy(&cells[new_coff[i]], &mesh->cells.cell.idx[coff[old]], ncc * sizeof(int));
            new_coff[i + 1] = new_coff[i] + ncc;
        }
    }
    int nbase = noff[beg];
    memcpy(&mesh->cells.node.idx[nbase], nodes, new_noff[num] * sizeof(int));
    for (int i = 0; i < num; i++) {
        noff[beg + i + 1] = nbase + new_noff[i + 1];
    }
    if (coff) {
        int cbase = coff[beg];
        memcpy(&mesh->cells.cell.idx[cbase], cells, new_coff[num] * sizeof(int));
        for (int i = 0; i < num; i++) {
            coff[beg + i + 1] = cbase + new_coff[i + 1];
        }
        for (int j = 0; j < coff[mesh->cells.num]; j++) {
            int c = mesh->cells.cell.idx[j];
            if (beg <= c && c < end) {
                mesh->cells.cell.idx[j] = beg + map[c - beg];
            }
        }
    }
    for (int j = 0; j < mesh->neighbors.send.off[mesh->neighbors.num]; j++) {
        int c = mesh->neighbors.send.idx[j];
        if (beg <= c && c < end) {
            mesh->neighbors.send.idx[j] = beg + map[c - beg];
        }
    }
    teal_free(inv);
    teal_free(nodes);
    teal_free(cells);
    teal_free(new_noff);
    teal_free(new_coff);
}

// exported for mesh_split (src/mesh/reorder.c:76)
void teal_mesh_renumber_cells(Mesh *mesh, const int *map, int beg, int end)
{
    renumber_cells(mesh, map, beg, end);
}

// breadth-first renumbering of the inner cells (src/mesh/generate.c:68-133)
static void bfs_order(Mesh *mesh)
{
    const int ni = mesh->cells.num_inner;
    const int *off = mesh->cells.cell.off;
    const int *idx = mesh->cells.cell.idx;
    int *map = teal_calloc(ni, sizeof(int));
    int *queue = teal_calloc(ni, sizeof(int));
    for (int i = 0; i < ni; i++) {
        map[i] = -1;
    }
    int seed = 0;
    int num = 0;
    int components = 0;
    while (num < ni) {
        while (map[seed] != -1) {
            seed = (seed + 1) % ni;
        }
        components += 1;
        int head = 0;
        int tail = 0;
        map[seed] = num++;
        queue[tail++] = seed;
        while (head < tail) {
            int cur = queue[head++];
            int child[MAX_CELL_FACES];
            int nchild = 0;
            for (int j = off[cur]; j < off[cur + 1]; j++) {
                int c = idx[j];
                if (c < ni && map[c] == -1) {
                    map[c] = -2;
                    // insert keeping descending degree, equal degrees in list order
                    int deg = off[c + 1] - off[c];
                    int pos = nchild;
                    while (pos > 0 && (off[child[pos - 1] + 1] - off[child[pos - 1]]) < deg) {
                        child[pos] = child[pos - 1];
                        pos -= 1;
                    }
                    child[pos] = c;
                    nchild += 1;
                }
            }
            for (int q = 0; q < nchild; q++) {
                map[child[q]] = num++;
                queue[tail++] = child[q];
            }
        }
    }
    if (components > 1) {
        teal_verbose("disconnected domain (%d)", components);
    }
    renumber_cells(mesh, map, 0, ni);
    teal_free(map);
    teal_free(queue);
}

// nodes numbered by first use walking the cells in order (src/mesh/generate.c:135-204)
static void first_use_nodes(Mesh *mesh)
{
    int nn = mesh->nodes.num;
    int *map = teal_calloc(nn, sizeof(int));
    for (int i = 0; i < nn; i++) {
        map[i] = -1;
    }
    int num = 0;
    for (int pass = 0; pass < 2; pass++) {
        for (int j = 0; j < mesh->cells.node.off[mesh->cells.num]; j++) {
            int node = mesh->cells.node.idx[j];
            if ((node < mesh->nodes.num_inner) == (pass == 0) && map[node] == -1) {
                map[node] = num++;
            }
        }
    }
    if (num != nn) {
        teal_error("mesh has %d unused nodes", nn - num);
    }
    Vector *coord = teal_calloc(nn, sizeof(Vector));
    for (int i = 0; i < nn; i++) {
        coord[map[i]] = mesh->nodes.coord[i];
    }
    memcpy(mesh->nodes.coord, coord, nn * sizeof(Vector));
    teal_free(coord);
    for (int i = 0; i < nn; i++) {
        mesh->nodes.global[i] = i;
    }
    for (int j = 0; j < mesh->cells.node.off[mesh->cells.num]; j++) {
        mesh->cells.node.idx[j] = map[mesh->cells.node.idx[j]];
    }
    teal_free(map);
}

// ------------------------------------------------------------------ faces
// make a quad's node cycle convex (src/mesh/generate.c:245-273)
static void orient_quad(int *node, int num, const Vector *coord)
{
    if (num == 3) {
        return;
    }
    if (num != 4) {
        teal_error("invalid number of nodes (%d)", num);
    }
    for (int i = 0; i < 3; i++) {
        Vector a = coord[node[0]];
        Vector ab = vector_sub(coord[node[1]], a);
        Vector ac = vector_sub(coord[node[2]], a);
        Vector ad = vector_sub(coord[node[3]], a);
        if (vector_dot(vector_cross(ab, ac), vector_cross(ac, ad)) > 0) {
            return;
        }
        if (i + 2 < num) {
            int t = node[i + 1];
            node[i + 1] = node[i + 2];
            node[i + 2] = t;
        }
    }
    teal_error("quad vertex order could not be made valid");
}

static void build_faces(Mesh *mesh)
{
    const int ni = mesh->cells.num_inner;
    const int nc = mesh->cells.num;
    const int *coff = mesh->cells.cell.off;
    const int *cidx = mesh->cells.cell.idx;
    const int *noff = mesh->cells.node.off;
    const int *nidx = mesh->cells.node.idx;

    // creation order: for each inner cell, its higher-numbered neighbours in list order
    int nf = 0;
    for (int c = 0; c < ni; c++) {
        for (int j = coff[c]; j < coff[c + 1]; j++) {
            nf += (c < cidx[j]);
        }
    }
    // stable bucket order: inner faces by left cell, then the others by right cell
    int *bucket = teal_calloc(nc + nc + 1, sizeof(int));
    for (int c = 0; c < ni; c++) {
        for (int j = coff[c]; j < coff[c + 1]; j++) {
            int r = cidx[j];
            if (c < r) {
                bucket[((r < ni) ? c : (nc + r)) + 1] += 1;
            }
        }
    }
    for (int i = 0; i < nc + nc; i++) {
        bucket[i + 1] += bucket[i];
    }
    Pair *cell_idx = teal_calloc(nf, sizeof(Pair));
    int *fnn = teal_calloc(nf, sizeof(int));
    int *fnodes = teal_calloc(nf * MAX_FACE_NODES, sizeof(int));
    for (int c = 0; c < ni; c++) {
        for (int j = coff[c]; j < coff[c + 1]; j++) {
            int r = cidx[j];
            if (c >= r) {
                continue;
            }
            int f = bucket[(r < ni) ? c : (nc + r)]++;
            cell_idx[f] = (Pair){c, r};
            int *fn = &fnodes[f * MAX_FACE_NODES];
            int k = 0;
            for (int a = noff[c]; a < noff[c + 1]; a++) {  // shared nodes in the left cell's order
                for (int b = noff[r]; b < noff[r + 1]; b++) {
                    if (nidx[a] == nidx[b]) {
                        if (k >= MAX_FACE_NODES) {
                            teal_error("face with more than %d nodes", MAX_FACE_NODES);
                        }
                        fn[k++] = nidx[a];
                    }
                }
            }
            orient_quad(fn, k, mesh->nodes.coord);
            fnn[f] = k;
        }
    }
    teal_free(bucket);
    int *foff = teal_calloc(nf + 1, sizeof(int));
    int num_inner = 0;
    int num_boundary = 0;
    for (int f = 0; f < nf; f++) {
        foff[f + 1] = foff[f] + fnn[f];
        num_inner += cell_idx[f].right < ni;
        num_boundary += (cell_idx[f].right >= ni && cell_idx[f].right < mesh->cells.off_boundary);
    }
    int *fidx = teal_calloc(foff[nf] + 1, sizeof(int));
    for (int f = 0; f < nf; f++) {
        memcpy(&fidx[foff[f]], &fnodes[f * MAX_FACE_NODES], fnn[f] * sizeof(int));
    }
    teal_free(fnn);
    teal_free(fnodes);
    mesh->faces.num = nf;
    mesh->faces.num_inner = num_inner;
    mesh->faces.off_boundary = num_inner + num_boundary;
    mesh->faces.node.off = foff;
    mesh->faces.node.idx = fidx;
    mesh->faces.cell_idx = cell_idx;

    // face ranges of the entities (src/mesh/generate.c:323-334)
    int ne = mesh->entities.num;
    int *face_off = teal_calloc(ne + 1, sizeof(int));
    face_off[0] = num_inner;
    for (int e = 0; e < ne; e++) {
        face_off[e + 1] = face_off[e];
        if (e >= mesh->entities.num_inner) {
            face_off[e + 1] += mesh->entities.cell_off[e + 1] - mesh->entities.cell_off[e];
        }
    }
    mesh->entities.face_off = face_off;
}

// ------------------------------------------------------------------ geometry
static double tri_area(Vector a, Vector b, Vector c)
{
    return vector_norm(vector_cross(vector_sub(b, a), vector_sub(c, a))) / 2;
}

static Vector tri_center(Vector a, Vector b, Vector c)
{
    return vector_div(vector_add(a, vector_add(b, c)), 3);
}

static double poly_area(const Vector *p, int n)  // generate.c:336-351
{
    if (n == 3) {
        return tri_area(p[0], p[1], p[2]);
    }
    return tri_area(p[0], p[1], p[2]) + tri_area(p[0], p[2], p[3]);
}

static Vector poly_center(const Vector *p, int n)  // generate.c:367-385
{
    if (n == 3) {
        return tri_center(p[0], p[1], p[2]);
    }
    double a0 = tri_area(p[0], p[1], p[2]);
    double a1 = tri_area(p[0], p[2], p[3]);
    Vector m0 = vector_mul(a0, tri_center(p[0], p[1], p[2]));
    Vector m1 = vector_mul(a1, tri_center(p[0], p[2], p[3]));
    return vector_div(vector_add(m0, m1), a0 + a1);
}

static Vector poly_normal(const Vector *p, int n)  // generate.c:400-418 (Newell for quads)
{
    Vector normal;
    if (n == 3) {
        normal = vector_cross(vector_sub(p[1], p[0]), vector_sub(p[2], p[0]));
    }
    else {
        normal = (Vector){0, 0, 0};
        for (int i = 0; i < 4; i++) {
            vector_iadd(&normal, vector_cross(p[i], p[(i + 1) % 4]));
        }
    }
    return vector_div(normal, vector_norm(normal));
}

static Matrix normal_basis(Vector n)  // generate.c:420-435
{
    Matrix b = {.x = n};
    double nqz = hypot(n.x, n.y);
    double nqy = hypot(n.x, n.z);
    if (nqz > nqy) {
        b.y = (Vector){-n.y / nqz, n.x / nqz, 0};
        b.z = (Vector){-n.x * n.z / nqz, -n.y * n.z / nqz, nqz};
    }
    else {
        b.y = (Vector){-n.x * n.y / nqy, nqy, -n.y * n.z / nqy};
        b.z = (Vector){-n.z / nqy, 0, n.x / nqy};
    }
    return b;
}

static double tet_volume(Vector a, Vector b, Vector c, Vector d)
{
    return fabs(vector_dot(vector_sub(b, a), vector_cross(vector_sub(c, a), vector_sub(d, a)))) / 6;
}

static Vector tet_center(Vector a, Vector b, Vector c, Vector d)
{
    return vector_div(vector_add(vector_add(a, b), vector_add(c, d)), 4);
}

// volume and centroid by the reference's decompositions (generate.c:468-550):
// pyramid = 2 tets, prism = pyramid + tet, hex = 2 prisms
static void solid(const Vector *p, int n, double *vol, Vector *cen)
{
    Vector q[6];
    double v0, v1;
    Vector c0, c1;
    switch (n) {
        case 4:
            *vol = tet_volume(p[0], p[1], p[2], p[3]);
            *cen = tet_center(p[0], p[1], p[2], p[3]);
            return;
        case 5:
            v0 = tet_volume(p[0], p[1], p[2], p[4]);
            v1 = tet_volume(p[0], p[2], p[3], p[4]);
            c0 = tet_center(p[0], p[1], p[2], p[4]);
            c1 = tet_center(p[0], p[2], p[3], p[4]);
            break;
        case 6:
            q[0] = p[0], q[1] = p[1], q[2] = p[4], q[3] = p[3], q[4] = p[5];
            solid(q, 5, &v0, &c0);
            v1 = tet_volume(p[0], p[1], p[2], p[5]);
            c1 = tet_center(p[0], p[1], p[2], p[5]);
            break;
        case 8:
            q[0] = p[0], q[1] = p[1], q[2] = p[2], q[3] = p[4], q[4] = p[5], q[5] = p[6];
            solid(q, 6, &v0, &c0);
            q[0] = p[0], q[1] = p[2], q[2] = p[3], q[3] = p[4], q[4] = p[6], q[5] = p[7];
            solid(q, 6, &v1, &c1);
            break;
        default: teal_error("invalid number of nodes (%d)", n);
    }
    *vol = v0 + v1;
    *cen = vector_div(vector_add(vector_mul(v0, c0), vector_mul(v1, c1)), v0 + v1);
}

static int node_coords(const Mesh *m, const Graph *g, int i, Vector *p)
{
    int n = 0;
    for (int j = g->off[i]; j < g->off[i + 1]; j++) {
        p[n++] = m->nodes.coord[g->idx[j]];
    }
    return n;
}

static void face_geometry(Mesh *mesh)
{
    int nf = mesh->faces.num;
    mesh->faces.area = teal_calloc(nf, sizeof(double));
    mesh->faces.center = teal_calloc(nf, sizeof(Vector));
    mesh->faces.basis = teal_calloc(nf, sizeof(Matrix));
    for (int f = 0; f < nf; f++) {
        Vector p[MAX_FACE_NODES];
        int n = node_coords(mesh, &mesh->faces.node, f, p);
        mesh->faces.area[f] = poly_area(p, n);
        mesh->faces.center[f] = poly_center(p, n);
        Matrix b = normal_basis(poly_normal(p, n));
        // point the normal away from the left cell (generate.c:437-451)
        int left = mesh->faces.cell_idx[f].left;
        Vector q[MAX_CELL_NODES];
        int nq = node_coords(mesh, &mesh->cells.node, left, q);
        Vector mean = {0, 0, 0};
        for (int k = 0; k < nq; k++) {
            vector_iadd(&mean, q[k]);
        }
        vector_idiv(&mean, nq);
        if (vector_dot(vector_sub(mean, mesh->faces.center[f]), b.x) > 0) {
            vector_imul(&b.x, -1);
            vector_imul(&b.y, -1);
        }
        mesh->faces.basis[f] = b;
    }
}

// periodic ghost -> inner source (exchange pairing of src/equations/exchange.c:19-33)
static int periodic_source(const Mesh *m, int ghost)
{
    const MeshNeighbors *nb = &m->neighbors;
    for (int e = 0; e < nb->num; e++) {
        if (nb->recv_off[e] <= ghost && ghost < nb->recv_off[e + 1]) {
            for (int s = 0; s < nb->num; s++) {
                if (nb->tag[s][1] == nb->tag[e][0]) {
                    return nb->send.idx[nb->send.off[s] + (ghost - nb->recv_off[e])];
                }
            }
        }
    }
    return -1;
}

static void cell_geometry(Mesh *mesh)
{
    const int nc = mesh->cells.num;
    const int ni = mesh->cells.num_inner;
    mesh->cells.volume = teal_calloc(nc, sizeof(double));
    mesh->cells.center = teal_calloc(nc, sizeof(Vector));
    mesh->cells.projection = teal_calloc(nc, sizeof(Vector));
    double sum = 0;
    for (int c = 0; c < ni; c++) {
        Vector p[MAX_CELL_NODES];
        int n = node_coords(mesh, &mesh->cells.node, c, p);
        solid(p, n, &mesh->cells.volume[c], &mesh->cells.center[c]);
        sum += mesh->cells.volume[c];
    }
    mesh->cells.sum_volume = sum;
    Vector *center = mesh->cells.center;
    // boundary ghosts: mirror image across the face (generate.c:592-599)
    for (int f = mesh->faces.num_inner; f < mesh->faces.off_boundary; f++) {
        int l = mesh->faces.cell_idx[f].left;
        int r = mesh->faces.cell_idx[f].right;
        Vector n = mesh->faces.basis[f].x;
        double factor = 2 * vector_dot(n, vector_sub(mesh->faces.center[f], center[l]));
        center[r] = vector_add(center[l], vector_mul(factor, n));
    }
    // periodic ghosts: the source centre, rotated and translated (generate.c:552-611)
    for (int g = mesh->cells.off_boundary; g < nc; g++) {
        int s = periodic_source(mesh, g);
        if (s >= 0) {
            center[g] = center[s];
        }
    }
    for (int e = mesh->entities.off_boundary; e < mesh->entities.num; e++) {
        for (int c = mesh->entities.cell_off[e]; c < mesh->entities.cell_off[e + 1]; c++) {
            center[c] = matrix_vector(mesh->entities.rotation[e], center[c]);
            vector_iadd(&center[c], mesh->entities.translation[e]);
        }
    }
    // projections sum A/2 |n| (generate.c:613-627)
    for (int f = 0; f < mesh->faces.num; f++) {
        int l = mesh->faces.cell_idx[f].left;
        int r = mesh->faces.cell_idx[f].right;
        Vector inc = vector_mul(mesh->faces.area[f] / 2, vector_abs(mesh->faces.basis[f].x));
        vector_iadd(&mesh->cells.projection[l], inc);
        if (r < ni) {
            vector_iadd(&mesh->cells.projection[r], inc);
        }
    }
}

// per inner cell, the faces it touches (<= MAX_CELL_FACES), in face order
static int *cell_faces(const Mesh *mesh, int *count)
{
    const int ni = mesh->cells.num_inner;
    int *faces = teal_calloc(ni * MAX_CELL_FACES, sizeof(int));
    for (int f = 0; f < mesh->faces.num; f++) {
        int l = mesh->faces.cell_idx[f].left;
        int r = mesh->faces.cell_idx[f].right;
        faces[(l * MAX_CELL_FACES) + count[l]++] = f;
        if (r < ni) {
            faces[(r * MAX_CELL_FACES) + count[r]++] = f;
        }
    }
    return faces;
}

// offsets from each inner cell to the shared face of every neighbour (generate.c:629-667)
static void neighbour_offsets(Mesh *mesh, const int *faces, const int *count)
{
    const int ni = mesh->cells.num_inner;
    const int *coff = mesh->cells.cell.off;
    mesh->cells.offset = teal_calloc(coff[ni] + 1, sizeof(Vector));
    for (int c = 0; c < ni; c++) {
        for (int j = coff[c]; j < coff[c + 1]; j++) {
            int other = mesh->cells.cell.idx[j];
            int f = -1;
            for (int k = 0; k < count[c]; k++) {
                Pair p = mesh->faces.cell_idx[faces[(c * MAX_CELL_FACES) + k]];
                if ((p.left == c && p.right == other) || (p.right == c && p.left == other)) {
                    f = faces[(c * MAX_CELL_FACES) + k];
                }
            }
            if (f < 0) {
                teal_error("no face between cells %d and %d", c, other);
            }
            mesh->cells.offset[j] = vector_sub(mesh->faces.center[f], mesh->cells.center[c]);
        }
    }
}

// least-squares weights from a per-cell QR factorization (generate.c:669-752)
typedef struct {
    double r11, r12, r22, r13, r23, r33;
} Qr;

static Vector lsq_weight(Vector d, const Qr *q)
{
    Vector w = {0, 0, 0};
    double beta = ((q->r12 * q->r23) - (q->r13 * q->r22)) / (q->r11 * q->r22);
    Vector a;
    a.x = d.x / sq(q->r11);
    a.y = (d.y - (q->r12 / q->r11 * d.x)) / sq(q->r22);
    a.z = (d.z - (q->r23 / q->r22 * d.y) + (beta * d.x)) / sq(q->r33);
    if (!is_close(d.x, 0)) {
        w.x += a.x;
    }
    if (!is_close(d.y, 0)) {
        w.x += -q->r12 / q->r11 * a.y;
        w.y += a.y;
    }
    if (!is_close(d.z, 0)) {
        w.x += beta * a.z;
        w.y += -q->r23 / q->r22 * a.z;
        w.z += a.z;
    }
    vector_imul(&w, 1 / vector_norm2(d));
    return w;
}

static void face_weights(Mesh *mesh)
{
    const int ni = mesh->cells.num_inner;
    const int nf = mesh->faces.num;
    const Vector *center = mesh->cells.center;
    Qr *q = teal_calloc(ni, sizeof(Qr));
    for (int f = 0; f < nf; f++) {
        int l = mesh->faces.cell_idx[f].left;
        int r = mesh->faces.cell_idx[f].right;
        Vector d = vector_sub(center[r], center[l]);
        double t2 = 1 / vector_norm2(d);
        for (int side = 0; side < 2; side++) {
            int c = side ? r : l;
            if (c >= ni) {
                break;
            }
            q[c].r11 += t2 * d.x * d.x;
            q[c].r12 += t2 * d.x * d.y;
            q[c].r22 += t2 * d.y * d.y;
            q[c].r13 += t2 * d.x * d.z;
            q[c].r23 += t2 * d.y * d.z;
            q[c].r33 += t2 * d.z * d.z;
        }
    }
    for (int c = 0; c < ni; c++) {
        q[c].r11 = sqrt(q[c].r11);
        q[c].r12 = q[c].r12 / q[c].r11;
        q[c].r22 = sqrt(q[c].r22 - sq(q[c].r12));
        q[c].r13 = q[c].r13 / q[c].r11;
        q[c].r23 = (q[c].r23 - (q[c].r12 * q[c].r13)) / q[c].r22;
        q[c].r33 = sqrt(q[c].r33 - (sq(q[c].r13) + sq(q[c].r23)));
    }
    mesh->faces.weight = teal_calloc(nf, sizeof(VectorPair));
    mesh->faces.offset = teal_calloc(nf, sizeof(VectorPair));
    mesh->faces.correction = teal_calloc(nf, sizeof(VectorNorm));
    for (int f = 0; f < nf; f++) {
        int l = mesh->faces.cell_idx[f].left;
        int r = mesh->faces.cell_idx[f].right;
        Vector d = vector_sub(center[r], center[l]);
        mesh->faces.weight[f].left = lsq_weight(d, &q[l]);
        if (r < ni) {
            mesh->faces.weight[f].right = lsq_weight(vector_mul(-1, d), &q[r]);
        }
        mesh->faces.offset[f].left = vector_sub(mesh->faces.center[f], center[l]);
        mesh->faces.offset[f].right = vector_sub(mesh->faces.center[f], center[r]);
        mesh->faces.correction[f].norm = vector_norm(d);
        mesh->faces.correction[f].unit = vector_div(d, mesh->faces.correction[f].norm);
    }
    teal_free(q);
}

void mesh_generate(Mesh *mesh)
{
    assert(mesh);
    dual_graph(mesh);
    bfs_order(mesh);
    first_use_nodes(mesh);
    build_faces(mesh);
    face_geometry(mesh);
    cell_geometry(mesh);
    int *count = teal_calloc(mesh->cells.num_inner, sizeof(int));
    int *faces = cell_faces(mesh, count);
    neighbour_offsets(mesh, faces, count);
    teal_free(faces);
    teal_free(count);
    face_weights(mesh);
}
