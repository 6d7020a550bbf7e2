// mesh_create.c -- Cartesian hex meshes (src/mesh/create.c:509-544) for one process.
//
// One process owns the whole box, so the reference's rank decomposition collapses to
// dims = {1,1,1}: a periodic axis makes each of its two sides a self-neighbour ghost block
// (create_neighbors, create.c:388-440) and a non-periodic side a boundary ghost block.
// Cells are [inner | boundary ghosts | periodic ghosts]; ghosts are quads ordered side by
// side (x-, x+, y-, y+, z-, z+), boundary sides before periodic ones (create.c:471-507).
#include <assert.h>
#include <string.h>

#include "mesh.h"
#include "teal.h"

enum { NO_NEIGHBOR = -1 };

static int side_cells(const int *n, int side)
{
    switch (side / 2) {
        case 0: return n[1] * n[2];
        case 1: return n[2] * n[0];
        default: return n[0] * n[1];
    }
}

// node index of grid point (i, j, k)
static int node_at(const int *n, int i, int j, int k)
{
    return i + ((n[0] + 1) * (j + ((n[1] + 1) * k)));
}

// the ghost quad of side `side` at in-plane position (a, b), nodes in the reference order
static void ghost_quad(const int *n, int side, int a, int b, int *q)
{
    switch (side / 2) {
        case 0: {  // x sides: (j = a, k = b)
            int i = (side % 2) ? n[0] : 0;
            q[0] = node_at(n, i, a, b);
            q[1] = node_at(n, i, a + 1, b);
            q[2] = node_at(n, i, a + 1, b + 1);
            q[3] = node_at(n, i, a, b + 1);
            return;
        }
        case 1: {  // y sides: (k = a, i = b)
            int j = (side % 2) ? n[1] : 0;
            q[0] = node_at(n, b, j, a);
            q[1] = node_at(n, b + 1, j, a);
            q[2] = node_at(n, b + 1, j, a + 1);
            q[3] = node_at(n, b, j, a + 1);
            return;
        }
        default: {  // z sides: (i = a, j = b)
            int k = (side % 2) ? n[2] : 0;
            q[0] = node_at(n, a, b, k);
            q[1] = node_at(n, a + 1, b, k);
            q[2] = node_at(n, a + 1, b + 1, k);
            q[3] = node_at(n, a, b + 1, k);
            return;
        }
    }
}

// in-plane extents of a side: ghosts are enumerated a-outer, b-inner (create.c:238-279)
static void side_extent(const int *n, int side, int *na, int *nb)
{
    switch (side / 2) {
        case 0: *na = n[1], *nb = n[2]; return;
        case 1: *na = n[2], *nb = n[0]; return;
        default: *na = n[0], *nb = n[1]; return;
    }
}

// the inner cell next to side `side` at in-plane position (a, b): the send lists of
// append_side_send_idx (create.c:351-386)
static int side_inner_cell(const int *n, int side, int a, int b)
{
    int hi = side % 2;
    switch (side / 2) {
        case 0: return (hi ? n[0] - 1 : 0) + (n[0] * (a + (n[1] * b)));
        case 1: return b + (n[0] * ((hi ? n[1] - 1 : 0) + (n[1] * a)));
        default: return a + (n[0] * (b + (n[1] * (hi ? n[2] - 1 : 0))));
    }
}

Mesh *mesh_create(Vector min_coord, Vector max_coord, Triple num_cells, Triple periodic)
{
    int n[3] = {num_cells.x, num_cells.y, num_cells.z};
    int per[3] = {periodic.x, periodic.y, periodic.z};
    int ndims = (n[0] > 1) + (n[1] > 1) + (n[2] > 1);
    double lo[3] = {min_coord.x, min_coord.y, min_coord.z};
    double hi[3] = {max_coord.x, max_coord.y, max_coord.z};
    double span[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};  // create.c:529

    // split_bounds with one part per axis (create.c:75-94)
    for (int d = 0; d < 3; d++) {
        if (d < ndims) {
            double width = (hi[d] - lo[d]) / n[d];
            lo[d] = lo[d] + (0 * width);
            hi[d] = lo[d] + (n[d] * width);
        }
        else {
            n[d] = 1;
            if (is_close(lo[d], hi[d])) {
                hi[d] = lo[d] + 1;
            }
        }
    }
    int neighbor[6];
    for (int s = 0; s < 6; s++) {
        neighbor[s] = (s < 2 * ndims && per[s / 2]) ? 0 : NO_NEIGHBOR;
    }

    Mesh *mesh = teal_calloc(1, sizeof(*mesh));

    // nodes (create.c:143-175): one process owns them all
    int num_nodes = (n[0] + 1) * (n[1] + 1) * (n[2] + 1);
    Vector del = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    del.x /= n[0];
    del.y /= n[1];
    del.z /= n[2];
    mesh->nodes.num = num_nodes;
    mesh->nodes.num_inner = num_nodes;
    mesh->nodes.global = teal_calloc(num_nodes, sizeof(long));
    mesh->nodes.coord = teal_calloc(num_nodes, sizeof(Vector));
    for (int k = 0; k <= n[2]; k++) {
        for (int j = 0; j <= n[1]; j++) {
            for (int i = 0; i <= n[0]; i++) {
                int id = node_at(n, i, j, k);
                mesh->nodes.global[id] = id;
                mesh->nodes.coord[id].x = lo[0] + (i * del.x);
                mesh->nodes.coord[id].y = lo[1] + (j * del.y);
                mesh->nodes.coord[id].z = lo[2] + (k * del.z);
            }
        }
    }

    // cells: inner hexes, then the ghost blocks already in final order
    int num_inner = n[0] * n[1] * n[2];
    int num_boundary = 0;
    int num_periodic = 0;
    for (int s = 0; s < 2 * ndims; s++) {
        if (neighbor[s] == NO_NEIGHBOR) {
            num_boundary += side_cells(n, s);
        }
        else {
            num_periodic += side_cells(n, s);
        }
    }
    int nc = num_inner + num_boundary + num_periodic;
    mesh->cells.num = nc;
    mesh->cells.num_inner = num_inner;
    mesh->cells.off_boundary = num_inner + num_boundary;
    mesh->cells.off_periodic = nc;
    int *off = teal_calloc(nc + 1, sizeof(int));
    int *idx = teal_calloc((num_inner * 8) + ((num_boundary + num_periodic) * 4), sizeof(int));
    int num = 0;
    for (int k = 0; k < n[2]; k++) {
        for (int j = 0; j < n[1]; j++) {
            for (int i = 0; i < n[0]; i++) {
                int *h = &idx[off[num]];
                h[0] = node_at(n, i, j, k);
                h[1] = node_at(n, i + 1, j, k);
                h[2] = node_at(n, i + 1, j + 1, k);
                h[3] = node_at(n, i, j + 1, k);
                h[4] = node_at(n, i, j, k + 1);
                h[5] = node_at(n, i + 1, j, k + 1);
                h[6] = node_at(n, i + 1, j + 1, k + 1);
                h[7] = node_at(n, i, j + 1, k + 1);
                off[num + 1] = off[num] + 8;
                num += 1;
            }
        }
    }
    for (int pass = 0; pass < 2; pass++) {  // boundary sides first, then periodic sides
        for (int s = 0; s < 2 * ndims; s++) {
            if ((neighbor[s] == NO_NEIGHBOR) != (pass == 0)) {
                continue;
            }
            int na, nb;
            side_extent(n, s, &na, &nb);
            for (int a = 0; a < na; a++) {
                for (int b = 0; b < nb; b++) {
                    ghost_quad(n, s, a, b, &idx[off[num]]);
                    off[num + 1] = off[num] + 4;
                    num += 1;
                }
            }
        }
    }
    assert(num == nc);
    mesh->cells.node.off = off;
    mesh->cells.node.idx = idx;

    // entities (create.c:300-349): "domain", non-periodic sides, periodic sides
    static const char *names[] = {"domain", "left", "right", "bottom", "top", "back", "front"};
    int ne = 1 + (2 * ndims);
    mesh->entities.num = ne;
    mesh->entities.num_inner = 1;
    mesh->entities.name = teal_calloc(ne, sizeof(String));
    mesh->entities.cell_off = teal_calloc(ne + 1, sizeof(int));
    mesh->entities.rotation = teal_calloc(ne, sizeof(Matrix));
    mesh->entities.translation = teal_calloc(ne, sizeof(Vector));
    strcpy(mesh->entities.name[0], names[0]);
    mesh->entities.cell_off[1] = num_inner;
    int e = 1;
    for (int pass = 0; pass < 2; pass++) {
        for (int s = 0; s < 2 * ndims; s++) {
            if ((per[s / 2] != 0) != (pass == 1)) {
                continue;
            }
            strcpy(mesh->entities.name[e], names[s + 1]);
            mesh->entities.cell_off[e + 1] = mesh->entities.cell_off[e] + side_cells(n, s);
            if (pass == 1) {
                Matrix *rot = &mesh->entities.rotation[e];
                rot->x.x = rot->y.y = rot->z.z = 1;
                double sign = (s % 2 == 0) ? -1 : +1;  // compute_translation, create.c:289-298
                Vector t = {0, 0, 0};
                if (s / 2 == 0) {
                    t.x = sign * span[0];
                }
                else if (s / 2 == 1) {
                    t.y = sign * span[1];
                }
                else {
                    t.z = sign * span[2];
                }
                mesh->entities.translation[e] = t;
            }
            e += 1;
        }
        if (pass == 0) {
            mesh->entities.off_boundary = e;
        }
    }

    // self-neighbour exchange lists of the periodic sides (create.c:388-440)
    int nn = 0;
    int nidx = 0;
    for (int s = 0; s < 2 * ndims; s++) {
        if (neighbor[s] != NO_NEIGHBOR) {
            nn += 1;
            nidx += side_cells(n, s);
        }
    }
    mesh->neighbors.num = nn;
    mesh->neighbors.tag = teal_calloc(nn + 1, sizeof(int[2]));
    mesh->neighbors.rank = teal_calloc(nn + 1, sizeof(int));
    mesh->neighbors.recv_off = teal_calloc(nn + 1, sizeof(int));
    mesh->neighbors.send.off = teal_calloc(nn + 1, sizeof(int));
    mesh->neighbors.send.idx = teal_calloc(nidx + 1, sizeof(int));
    mesh->neighbors.recv_off[0] = mesh->cells.off_boundary;
    int m = 0;
    for (int s = 0; s < 2 * ndims; s++) {
        if (neighbor[s] == NO_NEIGHBOR) {
            continue;
        }
        mesh->neighbors.tag[m][0] = s;
        mesh->neighbors.tag[m][1] = (s % 2 == 0) ? s + 1 : s - 1;
        mesh->neighbors.recv_off[m + 1] = mesh->neighbors.recv_off[m] + side_cells(n, s);
        int *cnt = &mesh->neighbors.send.off[m + 1];
        *cnt = mesh->neighbors.send.off[m];
        int na, nb;
        side_extent(n, s, &na, &nb);
        for (int a = 0; a < na; a++) {
            for (int b = 0; b < nb; b++) {
                mesh->neighbors.send.idx[(*cnt)++] = side_inner_cell(n, s, a, b);
            }
        }
        m += 1;
    }
    return mesh;
}
