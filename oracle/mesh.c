/* CPU ORACLE (test infrastructure only) -- single-rank mesh construction.
 *
 * Restates src/mesh/create.c (mesh_create on one MPI rank: periodic sides become
 * self-neighbour ghost blocks), src/mesh/reorder.c and src/mesh/generate.c.
 * The dual graph that the reference obtains from METIS_MeshToDual
 * (src/mesh/generate.c:39-40) is restated in o_mesh_to_dual from the published
 * METIS 5.1.0 algorithm (libmetis/mesh.c: CreateGraphDual, FindCommonElements).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define PROC_NULL (-1)

/* ---------------------------------------------------------------- create.c */

static int side_count(const int *n, int idx) /* src/mesh/create.c:177-185 */
{
    switch (idx / 2) {
        case 0: return n[1] * n[2];
        case 1: return n[2] * n[0];
        default: return n[0] * n[1];
    }
}

/* src/mesh/create.c:75-94 with dims = 1, coords = 0 */
static void split_bounds(double *min, double *max, int *num, int dim, int ndims)
{
    if (dim < ndims) {
        double width = (*max - *min) / *num;
        int offset = 0;
        *min = *min + (offset * width);
        *max = *min + (*num * width);
    }
    else {
        *num = 1;
        if (o_is_close(*min, *max)) {
            *max = *min + 1;
        }
    }
}

static void reorder_cells_range(OMesh *mesh, const int *map, int beg, int end);

/* src/mesh/create.c:509-544 (one rank) */
OMesh *omesh_create(const double *min_coord, const double *max_coord, const int *num_cells_,
                    const int *periodic)
{
    int n[3] = {num_cells_[0], num_cells_[1], num_cells_[2]};
    int ndims = (n[0] > 1) + (n[1] > 1) + (n[2] > 1);

    int neighbor[6] = {PROC_NULL, PROC_NULL, PROC_NULL, PROC_NULL, PROC_NULL, PROC_NULL};
    for (int i = 0; i < ndims; i++) {
        if (periodic[i]) {
            neighbor[2 * i] = neighbor[(2 * i) + 1] = 0; /* MPI_Cart_shift onto itself */
        }
    }

    double lo[3] = {min_coord[0], min_coord[1], min_coord[2]};
    double hi[3] = {max_coord[0], max_coord[1], max_coord[2]};
    double del[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]}; /* create.c:529 */
    for (int d = 0; d < 3; d++) {
        split_bounds(&lo[d], &hi[d], &n[d], d, ndims);
    }

    OMesh *mesh = o_calloc(1, sizeof(*mesh));

    /* create_nodes, create.c:143-175 */
    int num_nodes = (n[0] + 1) * (n[1] + 1) * (n[2] + 1);
    OVec *coord = o_calloc(num_nodes, sizeof(*coord));
    OVec dc = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    dc.x /= n[0];
    dc.y /= n[1];
    dc.z /= n[2];
    int num = 0;
    for (int k = 0; k < n[2] + 1; k++) {
        for (int j = 0; j < n[1] + 1; j++) {
            for (int i = 0; i < n[0] + 1; i++) {
                coord[num].x = lo[0] + (i * dc.x);
                coord[num].y = lo[1] + (j * dc.y);
                coord[num].z = lo[2] + (k * dc.z);
                num += 1;
            }
        }
    }
    mesh->nodes.num = num_nodes;
    mesh->nodes.num_inner = num_nodes;
    mesh->nodes.coord = coord;
    mesh->nodes.global = o_calloc(num_nodes, sizeof(long));
    for (int i = 0; i < num_nodes; i++) {
        mesh->nodes.global[i] = i;
    }

    /* create_cells, create.c:195-287 */
    int num_inner = n[0] * n[1] * n[2];
    int num_boundary = 0;
    int num_periodic = 0;
    for (int i = 0; i < 2 * ndims; i++) {
        if (neighbor[i] == PROC_NULL) {
            num_boundary += side_count(n, i);
        }
        else {
            num_periodic += side_count(n, i);
        }
    }
    int num_cells = num_inner + num_boundary + num_periodic;
    int *node_off = o_calloc(num_cells + 1, sizeof(int));
    int *node_idx = o_calloc((size_t)num_cells * O_MAX_CELL_NODES, sizeof(int));
    int nnx = n[0] + 1;
    int nny = n[1] + 1;
#define NODE(a, b, c) ((a) + (nnx * ((b) + (nny * (c)))))
    num = 0;
    for (int k = 0; k < n[2]; k++) {
        for (int j = 0; j < n[1]; j++) {
            for (int i = 0; i < n[0]; i++) {
                int *p = &node_idx[node_off[num]];
                p[0] = NODE(i, j, k);
                p[1] = NODE(i + 1, j, k);
                p[2] = NODE(i + 1, j + 1, k);
                p[3] = NODE(i, j + 1, k);
                p[4] = NODE(i, j, k + 1);
                p[5] = NODE(i + 1, j, k + 1);
                p[6] = NODE(i + 1, j + 1, k + 1);
                p[7] = NODE(i, j + 1, k + 1);
                node_off[num + 1] = node_off[num] + 8;
                num += 1;
            }
        }
    }
    if (ndims >= 1) {
        for (int i = 0; i <= n[0]; i += n[0]) {
            for (int j = 0; j < n[1]; j++) {
                for (int k = 0; k < n[2]; k++) {
                    int *p = &node_idx[node_off[num]];
                    p[0] = NODE(i, j, k);
                    p[1] = NODE(i, j + 1, k);
                    p[2] = NODE(i, j + 1, k + 1);
                    p[3] = NODE(i, j, k + 1);
                    node_off[num + 1] = node_off[num] + 4;
                    num += 1;
                }
            }
        }
    }
    if (ndims >= 2) {
        for (int j = 0; j <= n[1]; j += n[1]) {
            for (int k = 0; k < n[2]; k++) {
                for (int i = 0; i < n[0]; i++) {
                    int *p = &node_idx[node_off[num]];
                    p[0] = NODE(i, j, k);
                    p[1] = NODE(i + 1, j, k);
                    p[2] = NODE(i + 1, j, k + 1);
                    p[3] = NODE(i, j, k + 1);
                    node_off[num + 1] = node_off[num] + 4;
                    num += 1;
                }
            }
        }
    }
    if (ndims >= 3) {
        for (int k = 0; k <= n[2]; k += n[2]) {
            for (int i = 0; i < n[0]; i++) {
                for (int j = 0; j < n[1]; j++) {
                    int *p = &node_idx[node_off[num]];
                    p[0] = NODE(i, j, k);
                    p[1] = NODE(i + 1, j, k);
                    p[2] = NODE(i + 1, j + 1, k);
                    p[3] = NODE(i, j + 1, k);
                    node_off[num + 1] = node_off[num] + 4;
                    num += 1;
                }
            }
        }
    }
#undef NODE
    mesh->cells.num = num_cells;
    mesh->cells.num_inner = num_inner;
    mesh->cells.off_boundary = num_inner + num_boundary;
    mesh->cells.off_periodic = num_inner + num_boundary + num_periodic;
    mesh->cells.node.off = node_off;
    mesh->cells.node.idx = node_idx;

    /* create_entities, create.c:300-349 */
    static const char *entity[] = {"domain", "left", "right", "bottom", "top", "back", "front"};
    int num_entities = 1 + (2 * ndims);
    mesh->entities.num = num_entities;
    mesh->entities.num_inner = 1;
    mesh->entities.name = o_calloc(num_entities, sizeof(OString));
    mesh->entities.cell_off = o_calloc(num_entities + 1, sizeof(int));
    mesh->entities.rotation = o_calloc(num_entities, sizeof(OMat));
    mesh->entities.translation = o_calloc(num_entities, sizeof(OVec));
    strcpy(mesh->entities.name[0], entity[0]);
    mesh->entities.cell_off[1] = num_inner;
    num = 1;
    int num_bnd_entities = 0;
    for (int i = 0; i < 2 * ndims; i++) {
        if (!periodic[i / 2]) {
            strcpy(mesh->entities.name[num], entity[i + 1]);
            mesh->entities.cell_off[num + 1] = mesh->entities.cell_off[num] + side_count(n, i);
            num_bnd_entities += 1;
            num += 1;
        }
    }
    for (int i = 0; i < 2 * ndims; i++) {
        if (periodic[i / 2]) {
            strcpy(mesh->entities.name[num], entity[i + 1]);
            mesh->entities.cell_off[num + 1] = mesh->entities.cell_off[num] + side_count(n, i);
            OMat *rot = &mesh->entities.rotation[num];
            rot->x.x = rot->y.y = rot->z.z = 1;
            double sign = (i % 2 == 0) ? -1 : +1; /* compute_translation, create.c:289-298 */
            OVec t = {0, 0, 0};
            if (i / 2 == 0) {
                t.x = sign * del[0];
            }
            else if (i / 2 == 1) {
                t.y = sign * del[1];
            }
            else {
                t.z = sign * del[2];
            }
            mesh->entities.translation[num] = t;
            num += 1;
        }
    }
    mesh->entities.off_boundary = 1 + num_bnd_entities;

    /* create_neighbors, create.c:351-440 */
    int num_neighbors = 0;
    int num_indices = 0;
    for (int i = 0; i < 2 * ndims; i++) {
        if (neighbor[i] != PROC_NULL) {
            num_neighbors += 1;
            num_indices += side_count(n, i);
        }
    }
    mesh->neighbors.num = num_neighbors;
    mesh->neighbors.tag = o_calloc(num_neighbors + 1, sizeof(int[2]));
    mesh->neighbors.rank = o_calloc(num_neighbors + 1, sizeof(int));
    mesh->neighbors.recv_off = o_calloc(num_neighbors + 1, sizeof(int));
    mesh->neighbors.send.off = o_calloc(num_neighbors + 1, sizeof(int));
    mesh->neighbors.send.idx = o_calloc(num_indices + 1, sizeof(int));
    mesh->neighbors.recv_off[0] = mesh->cells.off_boundary;
    num = 0;
    for (int i = 0; i < 2 * ndims; i++) {
        if (neighbor[i] == PROC_NULL) {
            continue;
        }
        mesh->neighbors.tag[num][0] = i;
        mesh->neighbors.tag[num][1] = (i % 2 == 0) ? (i + 1) : (i - 1);
        mesh->neighbors.rank[num] = 0;
        mesh->neighbors.recv_off[num + 1] = mesh->neighbors.recv_off[num] + side_count(n, i);
        int *off = &mesh->neighbors.send.off[num + 1];
        int *idx = mesh->neighbors.send.idx;
        *off = mesh->neighbors.send.off[num];
        if (i / 2 == 0) { /* append_side_send_idx, create.c:351-386 */
            int ii = (i % 2 == 0) ? 0 : (n[0] - 1);
            for (int j = 0; j < n[1]; j++) {
                for (int k = 0; k < n[2]; k++) {
                    idx[(*off)++] = ii + (n[0] * (j + (n[1] * k)));
                }
            }
        }
        else if (i / 2 == 1) {
            int jj = (i % 2 == 0) ? 0 : (n[1] - 1);
            for (int k = 0; k < n[2]; k++) {
                for (int ii = 0; ii < n[0]; ii++) {
                    idx[(*off)++] = ii + (n[0] * (jj + (n[1] * k)));
                }
            }
        }
        else {
            int k = (i % 2 == 0) ? 0 : (n[2] - 1);
            for (int ii = 0; ii < n[0]; ii++) {
                for (int j = 0; j < n[1]; j++) {
                    idx[(*off)++] = ii + (n[0] * (j + (n[1] * k)));
                }
            }
        }
        num += 1;
    }

    /* reorder_cells, create.c:471-507: ghosts by side, boundary block first */
    int *map = o_calloc(num_cells, sizeof(int));
    num = 0;
    for (int i = 0; i < num_inner; i++) {
        map[num++] = i;
    }
    int nb = 0;
    int np = 0;
    for (int i = 0; i < 2 * ndims; i++) {
        for (int j = 0; j < side_count(n, i); j++) {
            if (neighbor[i] == PROC_NULL) {
                map[num++] = num_inner + nb++;
            }
            else {
                map[num++] = mesh->cells.off_boundary + np++;
            }
        }
    }
    reorder_cells_range(mesh, map, 0, num_cells);
    free(map);
    return mesh;
}

/* --------------------------------------------------------------- reorder.c */

/* src/mesh/reorder.c:76-135: cell (beg+i) moves to (beg+map[i]); lists keep their order */
static void reorder_cells_range(OMesh *mesh, const int *map, int beg, int end)
{
    int num = end - beg;
    int has_cell = mesh->cells.cell.off && mesh->cells.cell.idx;
    int *nn = o_calloc(num, sizeof(int));
    int *nodes = o_calloc((size_t)num * O_MAX_CELL_NODES, sizeof(int));
    int *nc = o_calloc(num, sizeof(int));
    int *cells = o_calloc((size_t)num * O_MAX_CELL_FACES, sizeof(int));
    for (int i = beg; i < end; i++) {
        int dst = map[i - beg];
        for (int j = mesh->cells.node.off[i]; j < mesh->cells.node.off[i + 1]; j++) {
            nodes[(dst * O_MAX_CELL_NODES) + nn[dst]++] = mesh->cells.node.idx[j];
        }
        if (has_cell) {
            for (int j = mesh->cells.cell.off[i]; j < mesh->cells.cell.off[i + 1]; j++) {
                cells[(dst * O_MAX_CELL_FACES) + nc[dst]++] = mesh->cells.cell.idx[j];
            }
        }
    }
    for (int i = beg; i < end; i++) {
        int src = i - beg;
        mesh->cells.node.off[i + 1] = mesh->cells.node.off[i];
        for (int j = 0; j < nn[src]; j++) {
            mesh->cells.node.idx[mesh->cells.node.off[i + 1]++] = nodes[(src * O_MAX_CELL_NODES) + j];
        }
        if (has_cell) {
            mesh->cells.cell.off[i + 1] = mesh->cells.cell.off[i];
            for (int j = 0; j < nc[src]; j++) {
                mesh->cells.cell.idx[mesh->cells.cell.off[i + 1]++] =
                    cells[(src * O_MAX_CELL_FACES) + j];
            }
        }
    }
    if (has_cell) {
        for (int i = 0; i < mesh->cells.cell.off[mesh->cells.num]; i++) {
            int idx = mesh->cells.cell.idx[i];
            if (beg <= idx && idx < end) {
                mesh->cells.cell.idx[i] = beg + map[idx - beg];
            }
        }
    }
    for (int i = 0; i < mesh->neighbors.send.off[mesh->neighbors.num]; i++) {
        int idx = mesh->neighbors.send.idx[i];
        if (beg <= idx && idx < end) {
            mesh->neighbors.send.idx[i] = beg + map[idx - beg];
        }
    }
    free(nn);
    free(nodes);
    free(nc);
    free(cells);
}

/* src/mesh/reorder.c:34-67 */
static void reorder_nodes_range(OMesh *mesh, const int *map, int beg, int end)
{
    int num = end - beg;
    OVec *coord = o_calloc(num, sizeof(OVec));
    long *global = o_calloc(num, sizeof(long));
    for (int i = beg; i < end; i++) {
        coord[map[i - beg]] = mesh->nodes.coord[i];
        global[map[i - beg]] = mesh->nodes.global[i];
    }
    for (int i = beg; i < end; i++) {
        mesh->nodes.coord[i] = coord[i - beg];
        mesh->nodes.global[i] = global[i - beg];
    }
    for (int i = 0; i < mesh->cells.node.off[mesh->cells.num]; i++) {
        int idx = mesh->cells.node.idx[i];
        if (beg <= idx && idx < end) {
            mesh->cells.node.idx[i] = beg + map[idx - beg];
        }
    }
    free(coord);
    free(global);
}

/* ------------------------------------------------------------ METIS dual */

/* METIS 5.1.0 CreateGraphDual + FindCommonElements (libmetis/mesh.c), restated.
 * Neighbours of element e are listed in first-encounter order: walk e's nodes in
 * eind order and, for each node, its incident elements in increasing id.  An
 * element l is kept if it shares >= ncommon nodes with e, or >= len(e)-1, or
 * >= len(l)-1 nodes; e itself is dropped. */
void o_mesh_to_dual(int num_elems, int num_nodes, const long *eptr, const long *eind,
                    int ncommon, long **xadj_, long **adjncy_)
{
    long *nptr = o_calloc(num_nodes + 1, sizeof(long));
    long *nind = o_calloc(eptr[num_elems] + 1, sizeof(long));
    for (int i = 0; i < num_elems; i++) {
        for (long j = eptr[i]; j < eptr[i + 1]; j++) {
            nptr[eind[j] + 1] += 1;
        }
    }
    for (int i = 0; i < num_nodes; i++) {
        nptr[i + 1] += nptr[i];
    }
    long *fill = o_calloc(num_nodes + 1, sizeof(long));
    for (int i = 0; i < num_elems; i++) {
        for (long j = eptr[i]; j < eptr[i + 1]; j++) {
            long node = eind[j];
            nind[nptr[node] + fill[node]++] = i;
        }
    }
    free(fill);

    int *marker = o_calloc(num_elems, sizeof(int));
    long *nbrs = o_calloc(num_elems, sizeof(long));
    long *xadj = o_calloc(num_elems + 1, sizeof(long));
    long cap = 0;
    long *adjncy = 0;

    for (int pass = 0; pass < 2; pass++) {
        for (int e = 0; e < num_elems; e++) {
            long elen = eptr[e + 1] - eptr[e];
            long k = 0;
            for (long i = eptr[e]; i < eptr[e + 1]; i++) {
                long node = eind[i];
                for (long ii = nptr[node]; ii < nptr[node + 1]; ii++) {
                    long jj = nind[ii];
                    if (marker[jj] == 0) {
                        nbrs[k++] = jj;
                    }
                    marker[jj] += 1;
                }
            }
            if (marker[e] == 0) {
                nbrs[k++] = e;
            }
            marker[e] = 0;
            long cnt = 0;
            for (long i = 0; i < k; i++) {
                long l = nbrs[i];
                long overlap = marker[l];
                if (overlap >= ncommon || overlap >= elen - 1 ||
                    overlap >= (eptr[l + 1] - eptr[l]) - 1) {
                    if (pass == 1) {
                        adjncy[xadj[e] + cnt] = l;
                    }
                    cnt += 1;
                }
                marker[l] = 0;
            }
            if (pass == 0) {
                xadj[e + 1] = xadj[e] + cnt;
            }
        }
        if (pass == 0) {
            cap = xadj[num_elems];
            adjncy = o_calloc(cap + 1, sizeof(long));
        }
    }
    free(marker);
    free(nbrs);
    free(nptr);
    free(nind);
    *xadj_ = xadj;
    *adjncy_ = adjncy;
}

/* ------------------------------------------------------------ generate.c */

/* src/mesh/generate.c:15-66 */
static void connect_cells(OMesh *mesh)
{
    int num_cells = mesh->cells.num;
    int num_nodes = mesh->nodes.num;
    long *eptr = o_calloc(num_cells + 1, sizeof(long));
    long *eind = o_calloc((size_t)num_cells * (O_MAX_CELL_NODES + 1), sizeof(long));
    for (int i = 0; i < num_cells; i++) {
        eptr[i + 1] = eptr[i];
        for (int j = mesh->cells.node.off[i]; j < mesh->cells.node.off[i + 1]; j++) {
            eind[eptr[i + 1]++] = mesh->cells.node.idx[j];
        }
        if (i >= mesh->cells.num_inner) {
            eind[eptr[i + 1]++] = num_nodes++; /* dummy node to enforce ncommon=3 */
        }
    }
    long *xadj;
    long *adjncy;
    o_mesh_to_dual(num_cells, num_nodes, eptr, eind, 3, &xadj, &adjncy);

    int *off = o_calloc(num_cells + 1, sizeof(int));
    int *idx = o_calloc((size_t)num_cells * O_MAX_CELL_FACES, sizeof(int));
    for (int i = 0; i < num_cells; i++) {
        off[i + 1] = off[i];
        for (long j = xadj[i]; j < xadj[i + 1]; j++) {
            if (i < mesh->cells.num_inner || adjncy[j] < mesh->cells.num_inner) {
                if (off[i + 1] - off[i] >= O_MAX_CELL_FACES) {
                    o_fatal("oracle: too many cell neighbours");
                }
                idx[off[i + 1]++] = (int)adjncy[j];
            }
        }
    }
    mesh->cells.cell.off = off;
    mesh->cells.cell.idx = idx;
    free(eptr);
    free(eind);
    free(xadj);
    free(adjncy);
}

/* src/mesh/generate.c:68-133 (BFS from cell 0, children by descending degree) */
static void bfs_cells(OMesh *mesh)
{
    int ni = mesh->cells.num_inner;
    int *map = o_calloc(ni, sizeof(int));
    int *degree = o_calloc(ni, sizeof(int));
    int *queue = o_calloc(ni, sizeof(int));
    for (int i = 0; i < ni; i++) {
        map[i] = -1;
        degree[i] = mesh->cells.cell.off[i + 1] - mesh->cells.cell.off[i];
    }
    int seed = 0;
    int num = 0;
    while (num < ni) {
        while (map[seed] != -1) {
            seed = (seed + 1) % ni;
        }
        int beg = 0;
        int end = 0;
        map[seed] = num++;
        queue[end++] = seed;
        while (beg < end) {
            int cur = queue[beg++];
            int cell[O_MAX_CELL_FACES];
            int cnt = 0;
            for (int i = mesh->cells.cell.off[cur]; i < mesh->cells.cell.off[cur + 1]; i++) {
                int idx = mesh->cells.cell.idx[i];
                if (idx < ni && map[idx] == -1) {
                    cell[cnt++] = idx;
                    map[idx] = -2;
                }
            }
            for (int i = 1; i < cnt; i++) { /* stable insertion sort, degree descending */
                int idx = cell[i];
                int deg = degree[idx];
                int j = i - 1;
                while (j >= 0 && degree[cell[j]] < deg) {
                    cell[j + 1] = cell[j];
                    j -= 1;
                }
                cell[j + 1] = idx;
            }
            for (int i = 0; i < cnt; i++) {
                map[cell[i]] = num++;
                queue[end++] = cell[i];
            }
        }
    }
    reorder_cells_range(mesh, map, 0, ni);
    free(map);
    free(degree);
    free(queue);
}

/* src/mesh/generate.c:135-204 (single rank: every node is inner, globals = index) */
static void renumber_nodes(OMesh *mesh)
{
    int *map = o_calloc(mesh->nodes.num, sizeof(int));
    for (int i = 0; i < mesh->nodes.num; i++) {
        map[i] = -1;
    }
    int num = 0;
    for (int pass = 0; pass < 2; pass++) {
        for (int i = 0; i < mesh->cells.num; i++) {
            for (int j = mesh->cells.node.off[i]; j < mesh->cells.node.off[i + 1]; j++) {
                int idx = mesh->cells.node.idx[j];
                int inner = idx < mesh->nodes.num_inner;
                if (inner == (pass == 0) && map[idx] == -1) {
                    map[idx] = num++;
                }
            }
        }
    }
    if (num != mesh->nodes.num) {
        o_fatal("oracle: unused mesh nodes");
    }
    reorder_nodes_range(mesh, map, 0, mesh->nodes.num);
    for (int i = 0; i < mesh->nodes.num_inner; i++) {
        mesh->nodes.global[i] = i;
    }
    free(map);
}

/* src/mesh/generate.c:245-273 */
static void correct_face_node_order(int *node, int num_nodes, const OVec *coord)
{
    if (num_nodes == 3) {
        return;
    }
    if (num_nodes != 4) {
        o_fatal("invalid number of nodes (%d)", num_nodes);
    }
    for (int i = 0; i < 3; i++) {
        OVec c[4];
        for (int j = 0; j < 4; j++) {
            c[j] = coord[node[j]];
        }
        OVec a2b = ov_sub(c[1], c[0]);
        OVec a2c = ov_sub(c[2], c[0]);
        OVec a2d = ov_sub(c[3], c[0]);
        if (ov_dot(ov_cross(a2b, a2c), ov_cross(a2c, a2d)) > 0) {
            return;
        }
        if (i + 2 < num_nodes) {
            int swap = node[i + 1];
            node[i + 1] = node[i + 2];
            node[i + 2] = swap;
        }
    }
    o_fatal("quad vertex order could not be made valid");
}

/* src/mesh/generate.c:206-321; the qsort (glibc merge sort, stable) becomes an
 * explicit stable bucket sort on the same key */
static void create_faces(OMesh *mesh)
{
    int ni = mesh->cells.num_inner;
    int max_faces = ni * O_MAX_CELL_FACES;
    OPair *pair = o_calloc(max_faces, sizeof(OPair));
    int *nn = o_calloc(max_faces, sizeof(int));
    int *nodes = o_calloc((size_t)max_faces * O_MAX_FACE_NODES, sizeof(int));
    int num_faces = 0;
    for (int i = 0; i < ni; i++) {
        for (int j = mesh->cells.cell.off[i]; j < mesh->cells.cell.off[i + 1]; j++) {
            int left = i;
            int right = mesh->cells.cell.idx[j];
            if (left < right) {
                pair[num_faces] = (OPair){left, right};
                int *fn = &nodes[num_faces * O_MAX_FACE_NODES];
                int cnt = 0;
                for (int a = mesh->cells.node.off[left]; a < mesh->cells.node.off[left + 1]; a++) {
                    for (int b = mesh->cells.node.off[right]; b < mesh->cells.node.off[right + 1];
                         b++) {
                        if (mesh->cells.node.idx[a] == mesh->cells.node.idx[b]) {
                            if (cnt >= O_MAX_FACE_NODES) {
                                o_fatal("oracle: too many face nodes");
                            }
                            fn[cnt++] = mesh->cells.node.idx[a];
                        }
                    }
                }
                correct_face_node_order(fn, cnt, mesh->nodes.coord);
                nn[num_faces] = cnt;
                num_faces += 1;
            }
        }
    }

    /* compare_face (generate.c:214-225): inner faces first ordered by left, the rest by right */
    int nc = mesh->cells.num;
    int *bucket = o_calloc((size_t)2 * nc + 1, sizeof(int));
    for (int f = 0; f < num_faces; f++) {
        int key = (pair[f].right < ni) ? pair[f].left : (nc + pair[f].right);
        bucket[key + 1] += 1;
    }
    for (int i = 0; i < 2 * nc; i++) {
        bucket[i + 1] += bucket[i];
    }
    int *order = o_calloc(num_faces, sizeof(int));
    for (int f = 0; f < num_faces; f++) {
        int key = (pair[f].right < ni) ? pair[f].left : (nc + pair[f].right);
        order[bucket[key]++] = f;
    }
    free(bucket);

    int *node_off = o_calloc(num_faces + 1, sizeof(int));
    int *node_idx = o_calloc((size_t)num_faces * O_MAX_FACE_NODES, sizeof(int));
    OPair *cell_idx = o_calloc(num_faces, sizeof(OPair));
    int num_inner = 0;
    int num_boundary = 0;
    for (int i = 0; i < num_faces; i++) {
        int f = order[i];
        if (pair[f].right < ni) {
            num_inner += 1;
        }
        else if (pair[f].right < mesh->cells.off_boundary) {
            num_boundary += 1;
        }
        node_off[i + 1] = node_off[i];
        for (int j = 0; j < nn[f]; j++) {
            node_idx[node_off[i + 1]++] = nodes[(f * O_MAX_FACE_NODES) + j];
        }
        cell_idx[i] = pair[f];
    }
    mesh->faces.num = num_faces;
    mesh->faces.num_inner = num_inner;
    mesh->faces.off_boundary = num_inner + num_boundary;
    mesh->faces.node.off = node_off;
    mesh->faces.node.idx = node_idx;
    mesh->faces.cell_idx = cell_idx;
    free(order);
    free(pair);
    free(nn);
    free(nodes);
}

/* src/mesh/generate.c:323-334 */
static void create_face_entities(OMesh *mesh)
{
    int *face_off = o_calloc(mesh->entities.num + 1, sizeof(int));
    face_off[0] = mesh->faces.num_inner;
    for (int i = 0; i < mesh->entities.num; i++) {
        face_off[i + 1] = face_off[i];
        if (i >= mesh->entities.num_inner) {
            face_off[i + 1] += mesh->entities.cell_off[i + 1] - mesh->entities.cell_off[i];
        }
    }
    mesh->entities.face_off = face_off;
}

/* src/mesh/generate.c:336-351 */
static double face_area(const OVec *c, int n)
{
    if (n == 3) {
        return ov_norm(ov_cross(ov_sub(c[1], c[0]), ov_sub(c[2], c[0]))) / 2;
    }
    OVec lhs[3] = {c[0], c[1], c[2]};
    OVec rhs[3] = {c[0], c[2], c[3]};
    return face_area(lhs, 3) + face_area(rhs, 3);
}

/* src/mesh/generate.c:367-385 */
static OVec face_center(const OVec *c, int n)
{
    if (n == 3) {
        return ov_div(ov_add(c[0], ov_add(c[1], c[2])), 3);
    }
    OVec lhs[3] = {c[0], c[1], c[2]};
    OVec rhs[3] = {c[0], c[2], c[3]};
    double area[2] = {face_area(lhs, 3), face_area(rhs, 3)};
    OVec cen[2] = {face_center(lhs, 3), face_center(rhs, 3)};
    OVec mul[2] = {ov_mul(area[0], cen[0]), ov_mul(area[1], cen[1])};
    return ov_div(ov_add(mul[0], mul[1]), area[0] + area[1]);
}

/* src/mesh/generate.c:400-418 */
static OVec face_normal(const OVec *c, int n)
{
    if (n == 3) {
        OVec normal = ov_cross(ov_sub(c[1], c[0]), ov_sub(c[2], c[0]));
        return ov_div(normal, ov_norm(normal));
    }
    OVec normal = {0, 0, 0};
    for (int i = 0; i < 4; i++) {
        OVec cr = ov_cross(c[i], c[(i + 1) % 4]);
        normal.x += cr.x;
        normal.y += cr.y;
        normal.z += cr.z;
    }
    return ov_div(normal, ov_norm(normal));
}

/* src/mesh/generate.c:420-435 */
static OMat face_basis(const OVec *c, int n)
{
    OVec normal = face_normal(c, n);
    OMat basis = {.x = normal};
    double nqz = hypot(normal.x, normal.y);
    double nqy = hypot(normal.x, normal.z);
    if (nqz > nqy) {
        basis.y = (OVec){-normal.y / nqz, normal.x / nqz, 0};
        basis.z = (OVec){-normal.x * normal.z / nqz, -normal.y * normal.z / nqz, nqz};
    }
    else {
        basis.y = (OVec){-normal.x * normal.y / nqy, nqy, -normal.y * normal.z / nqy};
        basis.z = (OVec){-normal.z / nqy, 0, normal.x / nqy};
    }
    return basis;
}

/* src/mesh/generate.c:468-490 */
static double cell_volume(const OVec *c, int n)
{
    switch (n) {
        case 4: {
            OVec a2b = ov_sub(c[1], c[0]);
            OVec a2c = ov_sub(c[2], c[0]);
            OVec a2d = ov_sub(c[3], c[0]);
            return fabs(ov_dot(a2b, ov_cross(a2c, a2d))) / 6;
        }
        case 5: {
            OVec lhs[4] = {c[0], c[1], c[2], c[4]};
            OVec rhs[4] = {c[0], c[2], c[3], c[4]};
            return cell_volume(lhs, 4) + cell_volume(rhs, 4);
        }
        case 6: {
            OVec lhs[5] = {c[0], c[1], c[4], c[3], c[5]};
            OVec rhs[4] = {c[0], c[1], c[2], c[5]};
            return cell_volume(lhs, 5) + cell_volume(rhs, 4);
        }
        case 8: {
            OVec lhs[6] = {c[0], c[1], c[2], c[4], c[5], c[6]};
            OVec rhs[6] = {c[0], c[2], c[3], c[4], c[6], c[7]};
            return cell_volume(lhs, 6) + cell_volume(rhs, 6);
        }
        default: o_fatal("invalid number of nodes (%d)", n);
    }
}

/* src/mesh/generate.c:517-550 */
static OVec cell_center(const OVec *c, int n)
{
    switch (n) {
        case 4: {
            OVec sum[2] = {ov_add(c[0], c[1]), ov_add(c[2], c[3])};
            return ov_div(ov_add(sum[0], sum[1]), 4);
        }
        case 5: {
            OVec lhs[4] = {c[0], c[1], c[2], c[4]};
            OVec rhs[4] = {c[0], c[2], c[3], c[4]};
            double vol[2] = {cell_volume(lhs, 4), cell_volume(rhs, 4)};
            OVec cen[2] = {cell_center(lhs, 4), cell_center(rhs, 4)};
            OVec mul[2] = {ov_mul(vol[0], cen[0]), ov_mul(vol[1], cen[1])};
            return ov_div(ov_add(mul[0], mul[1]), vol[0] + vol[1]);
        }
        case 6: {
            OVec lhs[5] = {c[0], c[1], c[4], c[3], c[5]};
            OVec rhs[4] = {c[0], c[1], c[2], c[5]};
            double vol[2] = {cell_volume(lhs, 5), cell_volume(rhs, 4)};
            OVec cen[2] = {cell_center(lhs, 5), cell_center(rhs, 4)};
            OVec mul[2] = {ov_mul(vol[0], cen[0]), ov_mul(vol[1], cen[1])};
            return ov_div(ov_add(mul[0], mul[1]), vol[0] + vol[1]);
        }
        case 8: {
            OVec lhs[6] = {c[0], c[1], c[2], c[4], c[5], c[6]};
            OVec rhs[6] = {c[0], c[2], c[3], c[4], c[6], c[7]};
            double vol[2] = {cell_volume(lhs, 6), cell_volume(rhs, 6)};
            OVec cen[2] = {cell_center(lhs, 6), cell_center(rhs, 6)};
            OVec mul[2] = {ov_mul(vol[0], cen[0]), ov_mul(vol[1], cen[1])};
            return ov_div(ov_add(mul[0], mul[1]), vol[0] + vol[1]);
        }
        default: o_fatal("invalid number of nodes (%d)", n);
    }
}

static int gather_coords(const OMesh *mesh, const OGraph *g, int i, OVec *c)
{
    int n = 0;
    for (int j = g->off[i]; j < g->off[i + 1]; j++) {
        c[n++] = mesh->nodes.coord[g->idx[j]];
    }
    return n;
}

/* face area/center/basis: generate.c:353-466 */
static void face_geometry(OMesh *mesh)
{
    int nf = mesh->faces.num;
    mesh->faces.area = o_calloc(nf, sizeof(double));
    mesh->faces.center = o_calloc(nf, sizeof(OVec));
    mesh->faces.basis = o_calloc(nf, sizeof(OMat));
    for (int i = 0; i < nf; i++) {
        OVec c[O_MAX_FACE_NODES];
        int n = gather_coords(mesh, &mesh->faces.node, i, c);
        mesh->faces.area[i] = face_area(c, n);
    }
    for (int i = 0; i < nf; i++) {
        OVec c[O_MAX_FACE_NODES];
        int n = gather_coords(mesh, &mesh->faces.node, i, c);
        mesh->faces.center[i] = face_center(c, n);
    }
    for (int i = 0; i < nf; i++) {
        OVec c[O_MAX_FACE_NODES];
        int n = gather_coords(mesh, &mesh->faces.node, i, c);
        OMat basis = face_basis(c, n);
        /* correct_face_basis, generate.c:437-451 */
        int left = mesh->faces.cell_idx[i].left;
        OVec mean = {0, 0, 0};
        for (int j = mesh->cells.node.off[left]; j < mesh->cells.node.off[left + 1]; j++) {
            OVec p = mesh->nodes.coord[mesh->cells.node.idx[j]];
            mean.x += p.x;
            mean.y += p.y;
            mean.z += p.z;
        }
        double cnt = mesh->cells.node.off[left + 1] - mesh->cells.node.off[left];
        mean.x /= cnt;
        mean.y /= cnt;
        mean.z /= cnt;
        if (ov_dot(ov_sub(mean, mesh->faces.center[i]), basis.x) > 0) {
            basis.x.x *= -1;
            basis.x.y *= -1;
            basis.x.z *= -1;
            basis.y.x *= -1;
            basis.y.y *= -1;
            basis.y.z *= -1;
        }
        mesh->faces.basis[i] = basis;
    }
}

int omesh_periodic_source(const OMesh *mesh, int ghost)
{
    /* the ghost block of entry e is filled by the send list of the entry whose
     * send tag equals e's receive tag (src/equations/exchange.c:28-33) */
    const ONeighbors *nb = &mesh->neighbors;
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

/* generate.c:492-627 */
static void cell_geometry(OMesh *mesh)
{
    int nc = mesh->cells.num;
    int ni = mesh->cells.num_inner;
    mesh->cells.volume = o_calloc(nc, sizeof(double));
    double sum_volume = 0;
    for (int i = 0; i < ni; i++) {
        OVec c[O_MAX_CELL_NODES];
        int n = gather_coords(mesh, &mesh->cells.node, i, c);
        mesh->cells.volume[i] = cell_volume(c, n);
        sum_volume += mesh->cells.volume[i];
    }
    mesh->cells.sum_volume = sum_volume;

    OVec *center = o_calloc(nc, sizeof(OVec));
    for (int i = 0; i < ni; i++) {
        OVec c[O_MAX_CELL_NODES];
        int n = gather_coords(mesh, &mesh->cells.node, i, c);
        center[i] = cell_center(c, n);
    }
    for (int i = mesh->faces.num_inner; i < mesh->faces.off_boundary; i++) {
        int left = mesh->faces.cell_idx[i].left;
        int right = mesh->faces.cell_idx[i].right;
        OVec normal = mesh->faces.basis[i].x;
        OVec offset = ov_sub(mesh->faces.center[i], center[left]);
        double factor = 2 * ov_dot(normal, offset);
        center[right] = ov_add(center[left], ov_mul(factor, normal));
    }
    /* collect_centers (generate.c:552-577): periodic ghost <- opposite-side inner centre */
    for (int g = mesh->cells.off_boundary; g < nc; g++) {
        int src = omesh_periodic_source(mesh, g);
        if (src >= 0) {
            center[g] = center[src];
        }
    }
    for (int i = mesh->entities.off_boundary; i < mesh->entities.num; i++) {
        for (int j = mesh->entities.cell_off[i]; j < mesh->entities.cell_off[i + 1]; j++) {
            center[j] = om_vec(mesh->entities.rotation[i], center[j]);
            OVec t = mesh->entities.translation[i];
            center[j].x += t.x;
            center[j].y += t.y;
            center[j].z += t.z;
        }
    }
    mesh->cells.center = center;

    /* compute_cell_projections, generate.c:613-627 */
    OVec *projection = o_calloc(nc, sizeof(OVec));
    for (int i = 0; i < mesh->faces.num; i++) {
        int left = mesh->faces.cell_idx[i].left;
        int right = mesh->faces.cell_idx[i].right;
        OVec inc = ov_mul(mesh->faces.area[i] / 2, ov_abs(mesh->faces.basis[i].x));
        projection[left].x += inc.x;
        projection[left].y += inc.y;
        projection[left].z += inc.z;
        if (right < ni) {
            projection[right].x += inc.x;
            projection[right].y += inc.y;
            projection[right].z += inc.z;
        }
    }
    mesh->cells.projection = projection;
}

/* compute_cell_offsets, generate.c:629-667: offset to the centre of the face shared with
 * each neighbour */
static void cell_offsets(OMesh *mesh)
{
    int ni = mesh->cells.num_inner;
    int *cnt = o_calloc(ni, sizeof(int));
    int *faces = o_calloc((size_t)ni * O_MAX_CELL_FACES, sizeof(int));
    for (int f = 0; f < mesh->faces.num; f++) {
        int l = mesh->faces.cell_idx[f].left;
        int r = mesh->faces.cell_idx[f].right;
        faces[(l * O_MAX_CELL_FACES) + cnt[l]++] = f;
        if (r < ni) {
            faces[(r * O_MAX_CELL_FACES) + cnt[r]++] = f;
        }
    }
    OVec *offset = o_calloc(mesh->cells.cell.off[ni] + 1, sizeof(OVec));
    for (int i = 0; i < ni; i++) {
        for (int j = mesh->cells.cell.off[i]; j < mesh->cells.cell.off[i + 1]; j++) {
            int other = mesh->cells.cell.idx[j];
            int lo = (i < other) ? i : other;
            int hi = (i > other) ? i : other;
            int found = -1;
            for (int k = 0; k < cnt[i]; k++) {
                OPair p = mesh->faces.cell_idx[faces[(i * O_MAX_CELL_FACES) + k]];
                if (p.left == lo && p.right == hi) {
                    found = faces[(i * O_MAX_CELL_FACES) + k];
                }
            }
            if (found < 0) {
                o_fatal("oracle: face of neighbour pair not found");
            }
            offset[j] = ov_sub(mesh->faces.center[found], mesh->cells.center[i]);
        }
    }
    mesh->cells.offset = offset;
    free(cnt);
    free(faces);
}

/* compute_face_weight, generate.c:669-693 */
static OVec face_weight(OVec delta, const double *r, int idx)
{
    double r11 = r[(6 * idx) + 0];
    double r12 = r[(6 * idx) + 1];
    double r22 = r[(6 * idx) + 2];
    double r13 = r[(6 * idx) + 3];
    double r23 = r[(6 * idx) + 4];
    double r33 = r[(6 * idx) + 5];
    OVec weight = {0, 0, 0};
    double beta = ((r12 * r23) - (r13 * r22)) / (r11 * r22);
    OVec alpha;
    alpha.x = delta.x / (r11 * r11);
    alpha.y = (delta.y - (r12 / r11 * delta.x)) / (r22 * r22);
    alpha.z = (delta.z - (r23 / r22 * delta.y) + (beta * delta.x)) / (r33 * r33);
    if (!o_is_close(delta.x, 0)) {
        weight.x += alpha.x;
    }
    if (!o_is_close(delta.y, 0)) {
        weight.x += -r12 / r11 * alpha.y;
        weight.y += alpha.y;
    }
    if (!o_is_close(delta.z, 0)) {
        weight.x += beta * alpha.z;
        weight.y += -r23 / r22 * alpha.z;
        weight.z += alpha.z;
    }
    double theta2 = 1 / ov_norm2(delta);
    weight.x *= theta2;
    weight.y *= theta2;
    weight.z *= theta2;
    return weight;
}

/* compute_face_weights / offsets / corrections, generate.c:695-777 */
static void face_weights(OMesh *mesh)
{
    int ni = mesh->cells.num_inner;
    double *r = o_calloc((size_t)6 * ni, sizeof(double));
    for (int i = 0; i < mesh->faces.num; i++) {
        int left = mesh->faces.cell_idx[i].left;
        int right = mesh->faces.cell_idx[i].right;
        OVec d = ov_sub(mesh->cells.center[right], mesh->cells.center[left]);
        double theta2 = 1 / ov_norm2(d);
        double inc[6] = {theta2 * d.x * d.x, theta2 * d.x * d.y, theta2 * d.y * d.y,
                         theta2 * d.x * d.z, theta2 * d.y * d.z, theta2 * d.z * d.z};
        for (int k = 0; k < 6; k++) {
            r[(6 * left) + k] += inc[k];
        }
        if (right < ni) {
            for (int k = 0; k < 6; k++) {
                r[(6 * right) + k] += inc[k];
            }
        }
    }
    for (int i = 0; i < ni; i++) {
        double *q = &r[6 * i]; /* r11 r12 r22 r13 r23 r33 */
        q[0] = sqrt(q[0]);
        q[1] = q[1] / q[0];
        q[2] = sqrt(q[2] - (q[1] * q[1]));
        q[3] = q[3] / q[0];
        q[4] = (q[4] - (q[1] * q[3])) / q[2];
        q[5] = sqrt(q[5] - ((q[3] * q[3]) + (q[4] * q[4])));
    }
    mesh->faces.weight = o_calloc(mesh->faces.num, sizeof(OVecPair));
    for (int i = 0; i < mesh->faces.num; i++) {
        int left = mesh->faces.cell_idx[i].left;
        int right = mesh->faces.cell_idx[i].right;
        OVec d = ov_sub(mesh->cells.center[right], mesh->cells.center[left]);
        mesh->faces.weight[i].left = face_weight(d, r, left);
        if (right < ni) {
            d.x *= -1;
            d.y *= -1;
            d.z *= -1;
            mesh->faces.weight[i].right = face_weight(d, r, right);
        }
    }
    free(r);

    mesh->faces.offset = o_calloc(mesh->faces.num, sizeof(OVecPair));
    mesh->faces.correction = o_calloc(mesh->faces.num, sizeof(OVecNorm));
    for (int i = 0; i < mesh->faces.num; i++) {
        int left = mesh->faces.cell_idx[i].left;
        int right = mesh->faces.cell_idx[i].right;
        mesh->faces.offset[i].left = ov_sub(mesh->faces.center[i], mesh->cells.center[left]);
        mesh->faces.offset[i].right = ov_sub(mesh->faces.center[i], mesh->cells.center[right]);
    }
    for (int i = 0; i < mesh->faces.num; i++) {
        int left = mesh->faces.cell_idx[i].left;
        int right = mesh->faces.cell_idx[i].right;
        OVec d = ov_sub(mesh->cells.center[right], mesh->cells.center[left]);
        mesh->faces.correction[i].norm = ov_norm(d);
        mesh->faces.correction[i].unit = ov_div(d, mesh->faces.correction[i].norm);
    }
}

/* src/mesh/generate.c:779-815 */
void omesh_generate(OMesh *mesh)
{
    connect_cells(mesh);
    bfs_cells(mesh);
    renumber_nodes(mesh);
    create_faces(mesh);
    create_face_entities(mesh);
    face_geometry(mesh);
    cell_geometry(mesh);
    cell_offsets(mesh);
    face_weights(mesh);
}

void omesh_destroy(OMesh *mesh)
{
    if (!mesh) {
        return;
    }
    free(mesh->nodes.global);
    free(mesh->nodes.coord);
    free(mesh->cells.node.off);
    free(mesh->cells.node.idx);
    free(mesh->cells.cell.off);
    free(mesh->cells.cell.idx);
    free(mesh->cells.volume);
    free(mesh->cells.center);
    free(mesh->cells.projection);
    free(mesh->cells.offset);
    free(mesh->faces.node.off);
    free(mesh->faces.node.idx);
    free(mesh->faces.cell_idx);
    free(mesh->faces.area);
    free(mesh->faces.center);
    free(mesh->faces.basis);
    free(mesh->faces.weight);
    free(mesh->faces.offset);
    free(mesh->faces.correction);
    free(mesh->entities.name);
    free(mesh->entities.cell_off);
    free(mesh->entities.face_off);
    free(mesh->entities.rotation);
    free(mesh->entities.translation);
    free(mesh->neighbors.tag);
    free(mesh->neighbors.rank);
    free(mesh->neighbors.recv_off);
    free(mesh->neighbors.send.off);
    free(mesh->neighbors.send.idx);
    free(mesh);
}
