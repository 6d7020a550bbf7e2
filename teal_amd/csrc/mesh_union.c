// mesh_union.c -- the disjoint union of generated meshes, for advancing many independent
// instances in one set of launches (no counterpart in the reference, which advances one
// mesh per program).
//
// The union keeps the layout conventions of src/mesh.h:13-72 across all instances: cells
// [inner (instance 0, 1, ...) | boundary ghosts (instance 0, 1, ...) | periodic ghosts (...)],
// faces [inner | boundary | periodic] in the same instance order, entities [inner | boundary |
// periodic] likewise, periodic neighbour blocks with instance-distinct tags.  Every array is
// copied (geometry) or renumbered (indices); no geometry is recomputed, so each instance's
// part of the union is bit-identical to the instance.
#include <string.h>

#include "mesh.h"
#include "teal.h"

enum { TAG_STRIDE = 1 << 16 };  // neighbour tags of instance i: tag + i * TAG_STRIDE

typedef struct {
    int inner, boundary, periodic;  // union offsets of the three cell (or face) groups
} Base;

static int map_index(const Base *b, int idx, int num_inner, int off_boundary)
{
    if (idx < num_inner) {
        return b->inner + idx;
    }
    if (idx < off_boundary) {
        return b->boundary + (idx - num_inner);
    }
    return b->periodic + (idx - off_boundary);
}

static int map_cell(const Base *b, const Mesh *m, int c)
{
    return map_index(b, c, m->cells.num_inner, m->cells.off_boundary);
}

static int map_face(const Base *b, const Mesh *m, int f)
{
    return map_index(b, f, m->faces.num_inner, m->faces.off_boundary);
}

Mesh *teal_mesh_union(const Mesh *const *meshes, int num)
{
    if (num < 1) {
        teal_error("mesh union: no meshes (%d)", num);
    }
    int nn = 0, nc = 0, ni = 0, nb = 0, nf = 0, nfi = 0, nfb = 0, ne = 0, nei = 0, neb = 0;
    int nnb = 0, nsend = 0, nnode_c = 0, nnode_f = 0, ncell_c = 0, noff = 0;
    for (int i = 0; i < num; i++) {
        const Mesh *m = meshes[i];
        if (!m->faces.area) {
            teal_error("mesh union: mesh %d is not generated", i);
        }
        nn += m->nodes.num;
        nc += m->cells.num;
        ni += m->cells.num_inner;
        nb += m->cells.off_boundary - m->cells.num_inner;
        nf += m->faces.num;
        nfi += m->faces.num_inner;
        nfb += m->faces.off_boundary - m->faces.num_inner;
        ne += m->entities.num;
        nei += m->entities.num_inner;
        neb += m->entities.off_boundary - m->entities.num_inner;
        nnb += m->neighbors.num;
        nsend += m->neighbors.send.off[m->neighbors.num];
        nnode_c += m->cells.node.off[m->cells.num];
        nnode_f += m->faces.node.off[m->faces.num];
        ncell_c += m->cells.cell.off[m->cells.num];
        noff += m->cells.cell.off[m->cells.num_inner];
    }
    Mesh *u = teal_calloc(1, sizeof(*u));
    u->nodes.num = nn;
    u->nodes.num_inner = nn;
    u->nodes.global = teal_calloc(nn, sizeof(long));
    u->nodes.coord = teal_calloc(nn, sizeof(Vector));
    u->cells.num = nc;
    u->cells.num_inner = ni;
    u->cells.off_boundary = ni + nb;
    u->cells.off_periodic = nc;
    u->cells.node.off = teal_calloc(nc + 1, sizeof(int));
    u->cells.node.idx = teal_calloc(nnode_c, sizeof(int));
    u->cells.cell.off = teal_calloc(nc + 1, sizeof(int));
    u->cells.cell.idx = teal_calloc(ncell_c, sizeof(int));
    u->cells.volume = teal_calloc(nc, sizeof(double));
    u->cells.center = teal_calloc(nc, sizeof(Vector));
    u->cells.projection = teal_calloc(nc, sizeof(Vector));
    u->cells.offset = teal_calloc(noff, sizeof(Vector));
    u->faces.num = nf;
    u->faces.num_inner = nfi;
    u->faces.off_boundary = nfi + nfb;
    u->faces.node.off = teal_calloc(nf + 1, sizeof(int));
    u->faces.node.idx = teal_calloc(nnode_f, sizeof(int));
    u->faces.cell_idx = teal_calloc(nf, sizeof(Pair));
    u->faces.area = teal_calloc(nf, sizeof(double));
    u->faces.center = teal_calloc(nf, sizeof(Vector));
    u->faces.basis = teal_calloc(nf, sizeof(Matrix));
    u->faces.weight = teal_calloc(nf, sizeof(VectorPair));
    u->faces.offset = teal_calloc(nf, sizeof(VectorPair));
    u->faces.correction = teal_calloc(nf, sizeof(VectorNorm));
    u->entities.num = ne;
    u->entities.num_inner = nei;
    u->entities.off_boundary = nei + neb;
    u->entities.name = teal_calloc(ne, sizeof(String));
    u->entities.cell_off = teal_calloc(ne + 1, sizeof(int));
    u->entities.face_off = teal_calloc(ne + 1, sizeof(int));
    u->entities.rotation = teal_calloc(ne, sizeof(Matrix));
    u->entities.translation = teal_calloc(ne, sizeof(Vector));
    u->neighbors.num = nnb;
    u->neighbors.tag = teal_calloc(nnb + 1, sizeof(int[2]));
    u->neighbors.rank = teal_calloc(nnb + 1, sizeof(int));
    u->neighbors.recv_off = teal_calloc(nnb + 1, sizeof(int));
    u->neighbors.send.off = teal_calloc(nnb + 1, sizeof(int));
    u->neighbors.send.idx = teal_calloc(nsend + 1, sizeof(int));

    // group bases of every instance
    Base *cb = teal_calloc(num, sizeof(Base));
    Base *fb = teal_calloc(num, sizeof(Base));
    Base *eb = teal_calloc(num, sizeof(Base));
    int *node0 = teal_calloc(num, sizeof(int));
    Base c = {0, ni, ni + nb}, f = {0, nfi, nfi + nfb}, e = {0, nei, nei + neb};
    int node = 0;
    for (int i = 0; i < num; i++) {
        const Mesh *m = meshes[i];
        cb[i] = c;
        fb[i] = f;
        eb[i] = e;
        node0[i] = node;
        c.inner += m->cells.num_inner;
        c.boundary += m->cells.off_boundary - m->cells.num_inner;
        c.periodic += m->cells.num - m->cells.off_boundary;
        f.inner += m->faces.num_inner;
        f.boundary += m->faces.off_boundary - m->faces.num_inner;
        f.periodic += m->faces.num - m->faces.off_boundary;
        e.inner += m->entities.num_inner;
        e.boundary += m->entities.off_boundary - m->entities.num_inner;
        e.periodic += m->entities.num - m->entities.off_boundary;
        node += m->nodes.num;
    }

    // nodes
    for (int i = 0; i < num; i++) {
        const Mesh *m = meshes[i];
        for (int k = 0; k < m->nodes.num; k++) {
            u->nodes.coord[node0[i] + k] = m->nodes.coord[k];
            u->nodes.global[node0[i] + k] = node0[i] + k;
        }
    }

    // cells: sizes first (CSR offsets in union order), then the rows
    for (int i = 0; i < num; i++) {
        const Mesh *m = meshes[i];
        for (int k = 0; k < m->cells.num; k++) {
            int r = map_cell(&cb[i], m, k);
            u->cells.node.off[r + 1] = m->cells.node.off[k + 1] - m->cells.node.off[k];
            u->cells.cell.off[r + 1] = m->cells.cell.off[k + 1] - m->cells.cell.off[k];
        }
    }
    for (int r = 0; r < nc; r++) {
        u->cells.node.off[r + 1] += u->cells.node.off[r];
        u->cells.cell.off[r + 1] += u->cells.cell.off[r];
    }
    for (int i = 0; i < num; i++) {
        const Mesh *m = meshes[i];
        for (int k = 0; k < m->cells.num; k++) {
            int r = map_cell(&cb[i], m, k);
            int dn = u->cells.node.off[r];
            for (int j = m->cells.node.off[k]; j < m->cells.node.off[k + 1]; j++) {
                u->cells.node.idx[dn++] = node0[i] + m->cells.node.idx[j];
            }
            int dc = u->cells.cell.off[r];
            for (int j = m->cells.cell.off[k]; j < m->cells.cell.off[k + 1]; j++) {
                if (k < m->cells.num_inner) {  // neighbour offsets of inner cells (CSR order)
                    u->cells.offset[dc] = m->cells.offset[j];
                }
                u->cells.cell.idx[dc++] = map_cell(&cb[i], m, m->cells.cell.idx[j]);
            }
            u->cells.volume[r] = m->cells.volume[k];
            u->cells.center[r] = m->cells.center[k];
            u->cells.projection[r] = m->cells.projection[k];
        }
        u->cells.sum_volume += m->cells.sum_volume;
    }

    // faces
    for (int i = 0; i < num; i++) {
        const Mesh *m = meshes[i];
        for (int k = 0; k < m->faces.num; k++) {
            int r = map_face(&fb[i], m, k);
            u->faces.node.off[r + 1] = m->faces.node.off[k + 1] - m->faces.node.off[k];
        }
    }
    for (int r = 0; r < nf; r++) {
        u->faces.node.off[r + 1] += u->faces.node.off[r];
    }
    for (int i = 0; i < num; i++) {
        const Mesh *m = meshes[i];
        for (int k = 0; k < m->faces.num; k++) {
            int r = map_face(&fb[i], m, k);
            int dn = u->faces.node.off[r];
            for (int j = m->faces.node.off[k]; j < m->faces.node.off[k + 1]; j++) {
                u->faces.node.idx[dn++] = node0[i] + m->faces.node.idx[j];
            }
            u->faces.cell_idx[r].left = map_cell(&cb[i], m, m->faces.cell_idx[k].left);
            u->faces.cell_idx[r].right = map_cell(&cb[i], m, m->faces.cell_idx[k].right);
            u->faces.area[r] = m->faces.area[k];
            u->faces.center[r] = m->faces.center[k];
            u->faces.basis[r] = m->faces.basis[k];
            u->faces.weight[r] = m->faces.weight[k];
            u->faces.offset[r] = m->faces.offset[k];
            u->faces.correction[r] = m->faces.correction[k];
        }
    }

    // entities: [inner | boundary | periodic] of all instances; cell ranges follow the cell
    // groups, face ranges the face groups (inner entities own no faces, generate.c:323-334)
    for (int i = 0; i < num; i++) {
        const Mesh *m = meshes[i];
        for (int k = 0; k < m->entities.num; k++) {
            int r = map_index(&eb[i], k, m->entities.num_inner, m->entities.off_boundary);
            strcpy(u->entities.name[r], m->entities.name[k]);
            u->entities.rotation[r] = m->entities.rotation[k];
            u->entities.translation[r] = m->entities.translation[k];
            int cells = m->entities.cell_off[k + 1] - m->entities.cell_off[k];
            u->entities.cell_off[r + 1] = cells;
        }
    }
    for (int r = 0; r < ne; r++) {
        u->entities.cell_off[r + 1] += u->entities.cell_off[r];
    }
    u->entities.face_off[0] = nfi;
    for (int r = 0; r < ne; r++) {
        u->entities.face_off[r + 1] = u->entities.face_off[r];
        if (r >= nei) {
            u->entities.face_off[r + 1] += u->entities.cell_off[r + 1] - u->entities.cell_off[r];
        }
    }

    // periodic neighbour blocks, instance by instance (receive ranges follow the periodic
    // ghost group)
    int nbi = 0, sent = 0;
    u->neighbors.recv_off[0] = ni + nb;
    for (int i = 0; i < num; i++) {
        const Mesh *m = meshes[i];
        for (int k = 0; k < m->neighbors.num; k++) {
            u->neighbors.tag[nbi][0] = m->neighbors.tag[k][0] + (i * TAG_STRIDE);
            u->neighbors.tag[nbi][1] = m->neighbors.tag[k][1] + (i * TAG_STRIDE);
            u->neighbors.rank[nbi] = 0;
            u->neighbors.recv_off[nbi + 1] = u->neighbors.recv_off[nbi] +
                                             (m->neighbors.recv_off[k + 1] - m->neighbors.recv_off[k]);
            for (int j = m->neighbors.send.off[k]; j < m->neighbors.send.off[k + 1]; j++) {
                u->neighbors.send.idx[sent++] = map_cell(&cb[i], m, m->neighbors.send.idx[j]);
            }
            u->neighbors.send.off[nbi + 1] = sent;
            nbi += 1;
        }
    }
    teal_free(cb);
    teal_free(fb);
    teal_free(eb);
    teal_free(node0);
    return u;
}
