// mesh_read.c -- mesh_read (src/mesh/read.c:901-918): a Gmsh 4.1 mesh file, ASCII or binary
// (src/mesh/gmsh.c:135-904), read on one process into the layout mesh_generate takes.
//
// What the reference does on one MPI rank, restated without MPI:
//   - sections $MeshFormat, $PhysicalNames, $Entities, $Nodes, $Elements, $Periodic, in any
//     order; $End* markers are skipped; any other section is an error (gmsh.c:470-506);
//   - every element block is an entity with exactly one physical group; the blocks are
//     ordered unlinked before periodic-linked, then by dimension (volumes first), physical
//     tag and entity tag (gmsh.c:619-652); cells follow that block order, each block in file
//     order: [inner (volume elements) | boundary faces | periodic faces];
//   - consecutive blocks of one physical group form one inner or boundary entity; each
//     periodic block is an entity of its own, with the affine transform of its link (inverted
//     on the master side) and its side's link nodes (gmsh.c:735-866);
//   - a periodic face's exchange source is the inner cell behind the face its nodes map to
//     through the link (dual.c:154-351, read.c:294-416): the ghost block of entity E receives
//     the inner cells behind the partner's matching faces, in E's face order;
//   - the partition and redistribution steps keep the file order on one rank; the nodes are
//     renumbered to those the cells use, in file order (read.c:770-846).
// Errors end the program through teal_error with the reference's messages.
#include <assert.h>
#include <ctype.h>
#include <errno.h>
#include <inttypes.h>
#include <limits.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "matrix.h"
#include "mesh.h"
#include "teal.h"
#include "utils.h"

// ------------------------------------------------------------------ tokens and values
typedef struct {
    FILE *file;
    const char *name;
    int binary;
    int swap;
} Source;

enum { T_I32, T_U64, T_F64 };

// The next whitespace-delimited (or quoted) token into buf; one delimiter after it is
// consumed, so binary data may start right after a section name.  Returns its length, or 0
// at the end of the file.
static int next_word(Source *s, char *buf, int size)
{
    int ch;
    do {
        ch = fgetc(s->file);
    } while (ch != EOF && isspace(ch));
    if (ch == EOF) {
        memset(buf, 0, size);
        return 0;
    }
    int len = 0;
    if (ch == '"' || ch == '\'' || ch == '`') {
        const int quote = ch;
        while ((ch = fgetc(s->file)) != EOF && ch != quote) {
            if (len + 1 >= size) {
                teal_error("missing delimiter (buffer too small)");
            }
            buf[len++] = (char)ch;
        }
        if (ch == EOF) {
            teal_error("unterminated quoted string");
        }
    }
    else {
        while (ch != EOF && !isspace(ch)) {
            if (len + 1 >= size) {
                teal_error("missing delimiter (buffer too small)");
            }
            buf[len++] = (char)ch;
            ch = fgetc(s->file);
        }
    }
    memset(buf + len, 0, size - len);
    return len;
}

static void ascii_value(Source *s, void *dst, int idx, int type)
{
    char tok[128];
    if (!next_word(s, tok, sizeof(tok))) {
        teal_error("short read (%d / %d)", idx, idx + 1);
    }
    char *end = 0;
    errno = 0;
    if (type == T_I32) {
        long long v = strtoll(tok, &end, 10);
        if (v < INT32_MIN || INT32_MAX < v) {
            teal_error("value %s out of range [%lld, %lld]", tok, (long long)INT32_MIN,
                       (long long)INT32_MAX);
        }
        ((int32_t *)dst)[idx] = (int32_t)v;
    }
    else if (type == T_U64) {
        if (tok[0] == '-') {
            teal_error("value %s out of range [0, %llu]", tok, (unsigned long long)UINT64_MAX);
        }
        ((uint64_t *)dst)[idx] = (uint64_t)strtoull(tok, &end, 10);
    }
    else {
        ((double *)dst)[idx] = strtod(tok, &end);
    }
    if (errno != 0 || !end || end == tok || *end != 0) {
        teal_error("could not parse token (%s) as %s", tok,
                   type == T_I32 ? "int32" : type == T_U64 ? "uint64" : "double");
    }
}

static size_t type_size(int type) { return (type == T_I32) ? 4 : 8; }

static void swap_bytes(void *buf, size_t num, size_t size)
{
    unsigned char *p = buf;
    for (size_t i = 0; i < num; i++, p += size) {
        for (size_t a = 0, b = size - 1; a < b; a++, b--) {
            unsigned char t = p[a];
            p[a] = p[b];
            p[b] = t;
        }
    }
}

// `num` values of `type`, as text or as the file's binary encoding (src/core/parse.c)
static void read_values(Source *s, void *dst, size_t num, int type, int binary)
{
    if (num == 0) {
        return;
    }
    if (binary) {
        size_t got = fread(dst, type_size(type), num, s->file);
        if (got != num) {
            teal_error("short read (%zu / %zu)", got, num);
        }
        if (s->swap) {
            swap_bytes(dst, num, type_size(type));
        }
        return;
    }
    for (size_t i = 0; i < num; i++) {
        ascii_value(s, dst, (int)i, type);
    }
}

static int32_t get_i32(Source *s)
{
    int32_t v;
    read_values(s, &v, 1, T_I32, s->binary);
    return v;
}

static uint64_t get_u64(Source *s)
{
    uint64_t v;
    read_values(s, &v, 1, T_U64, s->binary);
    return v;
}

static int as_int(uint64_t v)
{
    if (v > INT_MAX) {
        teal_error("value %" PRIu64 " out of range [0, %d]", v, INT_MAX);
    }
    return (int)v;
}

static void *alloc_of(uint64_t num, size_t size)
{
    return teal_calloc(as_int(num > 0 ? num : 1), size);
}

// ------------------------------------------------------------------ file contents
typedef struct {
    int32_t dim, tag;
    String name;
} Physical;

typedef struct {
    int32_t tag;
    int num_physical;
    int32_t *physical;
} EntityTags;

typedef struct {
    int32_t dim, tag, parametric;
    uint64_t num;
    uint64_t *tag_of;
    double *coord;  // [num][3 + parametric dims]
} NodeBlock;

typedef struct {
    int32_t dim, tag, type;
    uint64_t num;
    int len;
    uint64_t *node;  // [num][len] node tags
} ElementBlock;

typedef struct {
    int32_t dim, tag, tag_master;
    uint64_t num_affine;
    double *affine;
    uint64_t num;
    uint64_t *node, *node_master;
} Link;

typedef struct {
    int num_physicals;
    Physical *physical;
    int num_surfaces, num_volumes;
    EntityTags *surface, *volume;
    int num_node_blocks;
    NodeBlock *node_block;
    uint64_t num_nodes;
    int num_element_blocks;
    ElementBlock *element_block;
    uint64_t num_elements;
    int num_links;
    Link *link;
} Gmsh;

static void read_format(Source *s)
{
    double version;
    int32_t file_type, data_size;
    read_values(s, &version, 1, T_F64, 0);
    if (!is_close(version, 4.1)) {
        teal_error("unsupported version (%g)", version);
    }
    read_values(s, &file_type, 1, T_I32, 0);
    read_values(s, &data_size, 1, T_I32, 0);
    if (data_size != 8) {
        teal_error("unsupported data-size (%d)", data_size);
    }
    s->binary = (file_type == 1);
    s->swap = 0;
    if (s->binary) {
        int32_t one;
        read_values(s, &one, 1, T_I32, 1);
        s->swap = (one != 1);
    }
}

static void read_physicals(Source *s, Gmsh *g)
{
    int32_t num;
    read_values(s, &num, 1, T_I32, 0);  // text in both encodings
    g->num_physicals = num;
    g->physical = teal_calloc(num > 0 ? num : 1, sizeof(Physical));
    for (int i = 0; i < num; i++) {
        read_values(s, &g->physical[i].dim, 1, T_I32, 0);
        read_values(s, &g->physical[i].tag, 1, T_I32, 0);
        next_word(s, g->physical[i].name, sizeof(String));
    }
}

// one point / curve / surface / volume record of $Entities (gmsh.c:171-303); the bounding
// entities are read and dropped, the physical tags of surfaces and volumes kept
static void read_entity(Source *s, int dim, EntityTags *keep)
{
    int32_t tag = get_i32(s);
    double box[6];
    read_values(s, box, (dim == 0) ? 3 : 6, T_F64, s->binary);
    uint64_t np = get_u64(s);
    int32_t *phys = alloc_of(np, sizeof(int32_t));
    read_values(s, phys, np, T_I32, s->binary);
    if (dim > 0) {
        uint64_t nb = get_u64(s);
        int32_t *bound = alloc_of(nb, sizeof(int32_t));
        read_values(s, bound, nb, T_I32, s->binary);
        teal_free(bound);
    }
    if (keep) {
        keep->tag = tag;
        keep->num_physical = as_int(np);
        keep->physical = phys;
    }
    else {
        teal_free(phys);
    }
}

static void read_entities(Source *s, Gmsh *g)
{
    uint64_t num[4];
    for (int d = 0; d < 4; d++) num[d] = get_u64(s);
    g->num_surfaces = as_int(num[2]);
    g->num_volumes = as_int(num[3]);
    g->surface = alloc_of(num[2], sizeof(EntityTags));
    g->volume = alloc_of(num[3], sizeof(EntityTags));
    for (int d = 0; d < 4; d++) {
        for (uint64_t i = 0; i < num[d]; i++) {
            EntityTags *keep = (d == 2) ? &g->surface[i] : (d == 3) ? &g->volume[i] : 0;
            read_entity(s, d, keep);
        }
    }
}

static void read_nodes(Source *s, Gmsh *g)
{
    uint64_t head[4];  // blocks, nodes, min tag, max tag
    for (int k = 0; k < 4; k++) head[k] = get_u64(s);
    g->num_node_blocks = as_int(head[0]);
    g->num_nodes = head[1];
    g->node_block = alloc_of(head[0], sizeof(NodeBlock));
    uint64_t total = 0;
    for (int i = 0; i < g->num_node_blocks; i++) {
        NodeBlock *b = &g->node_block[i];
        b->dim = get_i32(s);
        b->tag = get_i32(s);
        b->parametric = get_i32(s);
        b->num = get_u64(s);
        int len = 3 + (b->parametric ? b->dim : 0);
        b->tag_of = alloc_of(b->num, sizeof(uint64_t));
        read_values(s, b->tag_of, b->num, T_U64, s->binary);
        b->coord = alloc_of(b->num * len, sizeof(double));
        read_values(s, b->coord, b->num * len, T_F64, s->binary);
        total += b->num;
    }
    if (total != g->num_nodes) {
        teal_error("node count mismatch (%" PRIu64 " / %" PRIu64 ")", total, g->num_nodes);
    }
}

static int element_nodes(int32_t type)  // gmsh.c:369-380
{
    switch (type) {
        case 2: return 3;  // triangle
        case 3:            // quadrangle
        case 4: return 4;  // tetrahedron
        case 5: return 8;  // hexahedron
        case 6: return 6;  // prism
        case 7: return 5;  // pyramid
        default: teal_error("unsupported element type (%d)", type);
    }
}

static void read_elements(Source *s, Gmsh *g)
{
    uint64_t head[4];
    for (int k = 0; k < 4; k++) head[k] = get_u64(s);
    g->num_element_blocks = as_int(head[0]);
    g->num_elements = head[1];
    g->element_block = alloc_of(head[0], sizeof(ElementBlock));
    uint64_t total = 0;
    for (int i = 0; i < g->num_element_blocks; i++) {
        ElementBlock *b = &g->element_block[i];
        b->dim = get_i32(s);
        b->tag = get_i32(s);
        b->type = get_i32(s);
        b->num = get_u64(s);
        b->len = element_nodes(b->type);
        uint64_t *rec = alloc_of(b->num * (1 + b->len), sizeof(uint64_t));
        read_values(s, rec, b->num * (1 + b->len), T_U64, s->binary);
        b->node = alloc_of(b->num * b->len, sizeof(uint64_t));
        for (uint64_t j = 0; j < b->num; j++) {
            memcpy(&b->node[j * b->len], &rec[(j * (1 + b->len)) + 1], sizeof(uint64_t) * b->len);
        }
        teal_free(rec);
        total += b->num;
    }
    if (total != g->num_elements) {
        teal_error("element count mismatch (%" PRIu64 " / %" PRIu64 ")", total, g->num_elements);
    }
}

static void read_periodic(Source *s, Gmsh *g)
{
    uint64_t num = get_u64(s);
    g->num_links = as_int(num);
    g->link = alloc_of(num, sizeof(Link));
    for (int i = 0; i < g->num_links; i++) {
        Link *l = &g->link[i];
        l->dim = get_i32(s);
        l->tag = get_i32(s);
        l->tag_master = get_i32(s);
        l->num_affine = get_u64(s);
        l->affine = alloc_of(l->num_affine, sizeof(double));
        read_values(s, l->affine, l->num_affine, T_F64, s->binary);
        l->num = get_u64(s);
        uint64_t *pairs = alloc_of(2 * l->num, sizeof(uint64_t));
        read_values(s, pairs, 2 * l->num, T_U64, s->binary);
        l->node = alloc_of(l->num, sizeof(uint64_t));
        l->node_master = alloc_of(l->num, sizeof(uint64_t));
        for (uint64_t j = 0; j < l->num; j++) {
            l->node[j] = pairs[2 * j];
            l->node_master[j] = pairs[(2 * j) + 1];
        }
        teal_free(pairs);
    }
}

static void gmsh_free(Gmsh *g)
{
    teal_free(g->physical);
    for (int i = 0; i < g->num_surfaces; i++) teal_free(g->surface[i].physical);
    for (int i = 0; i < g->num_volumes; i++) teal_free(g->volume[i].physical);
    teal_free(g->surface);
    teal_free(g->volume);
    for (int i = 0; i < g->num_node_blocks; i++) {
        teal_free(g->node_block[i].tag_of);
        teal_free(g->node_block[i].coord);
    }
    teal_free(g->node_block);
    for (int i = 0; i < g->num_element_blocks; i++) teal_free(g->element_block[i].node);
    teal_free(g->element_block);
    for (int i = 0; i < g->num_links; i++) {
        teal_free(g->link[i].affine);
        teal_free(g->link[i].node);
        teal_free(g->link[i].node_master);
    }
    teal_free(g->link);
}

static void read_file(const char *fname, Gmsh *g)
{
    Source s = {fopen(fname, "rb"), fname, 0, 0};
    if (!s.file) {
        teal_error("could not open file (%s)", fname);
    }
    char section[128];
    while (next_word(&s, section, sizeof(section))) {
        if (!strcmp(section, "$MeshFormat")) {
            read_format(&s);
        }
        else if (!strcmp(section, "$PhysicalNames")) {
            read_physicals(&s, g);
        }
        else if (!strcmp(section, "$Entities")) {
            read_entities(&s, g);
        }
        else if (!strcmp(section, "$Nodes")) {
            read_nodes(&s, g);
        }
        else if (!strcmp(section, "$Elements")) {
            read_elements(&s, g);
        }
        else if (!strcmp(section, "$Periodic")) {
            read_periodic(&s, g);
        }
        else if (!strncmp(section, "$End", 4)) {
            continue;
        }
        else {
            teal_error("unsupported mesh section (%s)", section);
        }
    }
    fclose(s.file);
}

// ------------------------------------------------------------------ entities
typedef struct {
    const ElementBlock *block;
    const Physical *physical;
    const Link *link;
} Part;

static const Physical *physical_of(const ElementBlock *b, const Gmsh *g)  // gmsh.c:557-595
{
    const EntityTags *list = 0;
    int num = 0;
    if (b->dim == 2) {
        list = g->surface;
        num = g->num_surfaces;
    }
    else if (b->dim == 3) {
        list = g->volume;
        num = g->num_volumes;
    }
    else {
        teal_error("unsupported element block dimension (%d)", b->dim);
    }
    const EntityTags *e = 0;
    for (int i = 0; i < num && !e; i++) {
        if (list[i].tag == b->tag) {
            e = &list[i];
        }
    }
    if (!e) {
        teal_error("could not find %s entity (%d)", (b->dim == 2) ? "surface" : "volume", b->tag);
    }
    if (e->num_physical != 1) {
        teal_error("unsupported number of physical tags (%d)", e->num_physical);
    }
    for (int i = 0; i < g->num_physicals; i++) {
        if (g->physical[i].dim == b->dim && g->physical[i].tag == e->physical[0]) {
            return &g->physical[i];
        }
    }
    teal_error("could not find physical name (%d, %d)", b->dim, e->physical[0]);
}

static const Link *link_of(const ElementBlock *b, const Gmsh *g)  // gmsh.c:597-613
{
    for (int i = 0; i < g->num_links; i++) {
        const Link *l = &g->link[i];
        if (l->dim == b->dim && (l->tag == b->tag || l->tag_master == b->tag)) {
            return l;
        }
    }
    return 0;
}

// unlinked before linked, volumes before surfaces, then physical tag, then entity tag
static int part_before(const Part *a, const Part *b)
{
    if (!a->link != !b->link) {
        return !a->link;
    }
    if (a->physical->dim != b->physical->dim) {
        return a->physical->dim > b->physical->dim;
    }
    if (a->physical->tag != b->physical->tag) {
        return a->physical->tag < b->physical->tag;
    }
    return a->block->tag < b->block->tag;
}

// ------------------------------------------------------------------ assembly
typedef struct {
    uint64_t tag;
    int idx;
} TagIdx;

static int compare_tag(const void *lhs_, const void *rhs_)
{
    const TagIdx *lhs = lhs_, *rhs = rhs_;
    return (lhs->tag > rhs->tag) - (lhs->tag < rhs->tag);
}

static int node_index(const TagIdx *map, int num, uint64_t tag)
{
    TagIdx key = {tag, 0};
    const TagIdx *hit = bsearch(&key, map, num, sizeof(*map), compare_tag);
    if (!hit) {
        teal_error("node tag (%" PRIu64 ") was not converted to index", tag);
    }
    return hit->idx;
}

// sorted node list of cell c (at most MAX_FACE_NODES of a face element)
typedef struct {
    int num;
    int node[MAX_CELL_NODES];
    int cell;
} Key;

static int compare_key(const void *lhs_, const void *rhs_)
{
    const Key *lhs = lhs_, *rhs = rhs_;
    if (lhs->num != rhs->num) {
        return (lhs->num < rhs->num) ? -1 : +1;
    }
    for (int i = 0; i < lhs->num; i++) {
        if (lhs->node[i] != rhs->node[i]) {
            return (lhs->node[i] < rhs->node[i]) ? -1 : +1;
        }
    }
    return 0;
}

static void sort_ints(int *v, int n)
{
    for (int i = 1; i < n; i++) {
        for (int j = i; j > 0 && v[j - 1] > v[j]; j--) {
            int t = v[j];
            v[j] = v[j - 1];
            v[j - 1] = t;
        }
    }
}

// The inner cell that shares at least three nodes with face element f (the one neighbour the
// reference's dual graph gives a face element, ncommon = 3: dual.c:15-42).
static int inner_cell_of(int f, const int *off, const int *idx, const int *node_cells_off,
                         const int *node_cells)
{
    int best = -1;
    for (int a = off[f]; a < off[f + 1]; a++) {
        const int n = idx[a];
        for (int q = node_cells_off[n]; q < node_cells_off[n + 1]; q++) {
            const int c = node_cells[q];
            if (c == best) {
                continue;
            }
            int shared = 0;
            for (int b = off[f]; b < off[f + 1]; b++) {
                for (int k = off[c]; k < off[c + 1]; k++) {
                    shared += (idx[k] == idx[b]);
                }
            }
            if (shared >= 3) {
                if (best >= 0) {
                    teal_error("periodic face (%d) touches more than one inner cell", f);
                }
                best = c;
            }
        }
    }
    if (best < 0) {
        teal_error("periodic face (%d) touches no inner cell", f);
    }
    return best;
}

Mesh *mesh_read(const char *fname)
{
    assert(fname);
    const char *ext = strrchr(fname, '.');
    if (!ext) {
        teal_error("missing file extension (%s)", fname);
    }
    if (strcmp(ext, ".msh") != 0) {  // grid.c:11-22
        teal_error("unsupported file extension (%s)", ext);
    }
    Gmsh g = {0};
    read_file(fname, &g);
    if (g.num_node_blocks == 0 || g.num_nodes == 0) {
        teal_error("mesh does not contain any nodes");
    }
    if (g.num_element_blocks == 0 || g.num_elements == 0) {
        teal_error("mesh does not contain any elements");
    }

    // element blocks in entity order (stable insertion sort; few blocks)
    const int nblk = g.num_element_blocks;
    Part *part = teal_calloc(nblk, sizeof(Part));
    for (int i = 0; i < nblk; i++) {
        part[i].block = &g.element_block[i];
        part[i].physical = physical_of(part[i].block, &g);
        part[i].link = link_of(part[i].block, &g);
    }
    for (int i = 1; i < nblk; i++) {
        for (int j = i; j > 0 && part_before(&part[j], &part[j - 1]); j--) {
            Part t = part[j];
            part[j] = part[j - 1];
            part[j - 1] = t;
        }
    }
    if (part[0].physical->dim != 3) {
        teal_error("mesh does not contain any volumes");
    }

    // nodes in file order and the tag -> index map
    const int nn_file = as_int(g.num_nodes);
    Vector *coord_file = teal_calloc(nn_file, sizeof(Vector));
    TagIdx *tmap = teal_calloc(nn_file, sizeof(TagIdx));
    int n = 0;
    for (int i = 0; i < g.num_node_blocks; i++) {
        const NodeBlock *b = &g.node_block[i];
        const int len = 3 + (b->parametric ? b->dim : 0);
        for (uint64_t j = 0; j < b->num; j++, n++) {
            coord_file[n] = (Vector){b->coord[j * len], b->coord[(j * len) + 1],
                                     b->coord[(j * len) + 2]};
            tmap[n] = (TagIdx){b->tag_of[j], n};
        }
    }
    qsort(tmap, nn_file, sizeof(TagIdx), compare_tag);

    // cells in block order, node lists as file-order node indices
    const int nc = as_int(g.num_elements);
    int *off = teal_calloc(nc + 1, sizeof(int));
    int *idx = teal_calloc(nc * MAX_CELL_NODES, sizeof(int));
    int c = 0;
    for (int i = 0; i < nblk; i++) {
        const ElementBlock *b = part[i].block;
        for (uint64_t j = 0; j < b->num; j++, c++) {
            off[c + 1] = off[c] + b->len;
            for (int k = 0; k < b->len; k++) {
                idx[off[c] + k] = node_index(tmap, nn_file, b->node[(j * b->len) + k]);
            }
        }
    }

    // entities (gmsh.c:735-866): consecutive blocks of one physical group merge; every
    // linked block is an entity of its own
    int num_inner = 1, num_boundary = 0, num_periodic = 0;
    for (int i = 1; i < nblk; i++) {
        if (!part[i].link) {
            if (part[i].physical != part[i - 1].physical) {
                if (part[i].physical->dim == 3) {
                    num_inner += 1;
                }
                else if (part[i].physical->dim == 2) {
                    num_boundary += 1;
                }
                else {
                    teal_error("unsupported physical dimension (%d)", part[i].physical->dim);
                }
            }
        }
        else {
            if (part[i].link->dim != 2) {
                teal_error("unsupported periodic dimension (%d)", part[i].link->dim);
            }
            num_periodic += 1;
        }
    }
    const int ne = num_inner + num_boundary + num_periodic;
    const int off_bnd_ent = num_inner + num_boundary;
    String *name = teal_calloc(ne, sizeof(String));
    int *cell_off = teal_calloc(ne + 1, sizeof(int));
    int *partner = teal_calloc(ne, sizeof(int));
    Matrix *rotation = teal_calloc(ne, sizeof(Matrix));
    Vector *translation = teal_calloc(ne, sizeof(Vector));
    const Part **link_part = teal_calloc(ne, sizeof(Part *));
    int p = 0;
    for (int e = 0; e < ne; e++) {
        cell_off[e + 1] = cell_off[e];
        strcpy(name[e], part[p].physical->name);
        if (!part[p].link) {
            do {
                cell_off[e + 1] += as_int(part[p].block->num);
                p += 1;
            } while (p < nblk && !part[p].link && part[p].physical == part[p - 1].physical);
            continue;
        }
        const Part *q = &part[p];
        cell_off[e + 1] += as_int(q->block->num);
        link_part[e] = q;
        // the entity of the other side of the link (gmsh.c:758-778)
        const int32_t other = (q->link->tag == q->block->tag) ? q->link->tag_master : q->link->tag;
        partner[e] = -1;
        for (int k = off_bnd_ent, pk = p - (e - off_bnd_ent); k < ne; k++, pk++) {
            if (part[pk].block->tag == other) {
                partner[e] = k;
            }
        }
        if (partner[e] < 0) {
            teal_error("could not find periodic link (%d)", other);
        }
        // affine transform, inverted on the master side (gmsh.c:780-813)
        if (q->link->num_affine != 16) {
            teal_error("unsupported affine transformation size (%" PRIu64 ")", q->link->num_affine);
        }
        const double *a = q->link->affine;
        Matrix r = {{a[0], a[1], a[2]}, {a[4], a[5], a[6]}, {a[8], a[9], a[10]}};
        Vector t = {a[3], a[7], a[11]};
        if (!is_close(matrix_determinant(r), 1)) {
            teal_error("invalid affine transformation determinant (%g)", matrix_determinant(r));
        }
        if (q->block->tag == q->link->tag_master) {
            r = matrix_inverse(r);
            t = vector_mul(-1, matrix_vector(r, t));
        }
        rotation[e] = r;
        translation[e] = t;
        p += 1;
    }
    assert(p == nblk && cell_off[ne] == nc);
    const int cells_inner = cell_off[num_inner];
    const int cells_periodic = cell_off[off_bnd_ent];  // first periodic face

    // periodic exchange lists: ghost k of entity E takes the inner cell behind the partner's
    // face that E's face k maps to through the link nodes
    int *source = teal_calloc((nc - cells_periodic) + 1, sizeof(int));
    if (num_periodic > 0) {
        int *ncount = teal_calloc(nn_file + 1, sizeof(int));
        for (int q = 0; q < cells_inner; q++) {
            for (int k = off[q]; k < off[q + 1]; k++) ncount[idx[k] + 1] += 1;
        }
        for (int i = 0; i < nn_file; i++) ncount[i + 1] += ncount[i];
        int *ncells = teal_calloc(ncount[nn_file] + 1, sizeof(int));
        int *fill = teal_calloc(nn_file + 1, sizeof(int));
        memcpy(fill, ncount, sizeof(int) * (nn_file + 1));
        for (int q = 0; q < cells_inner; q++) {
            for (int k = off[q]; k < off[q + 1]; k++) ncells[fill[idx[k]]++] = q;
        }
        for (int e = off_bnd_ent; e < ne; e++) {
            const Link *l = link_part[e]->link;
            const int slave = (l->tag == link_part[e]->block->tag);
            const int pe = partner[e];
            // node of side e -> node of the partner side
            TagIdx *nmap = teal_calloc(as_int(l->num) + 1, sizeof(TagIdx));
            for (uint64_t j = 0; j < l->num; j++) {
                const uint64_t mine = slave ? l->node[j] : l->node_master[j];
                const uint64_t peer = slave ? l->node_master[j] : l->node[j];
                nmap[j] = (TagIdx){(uint64_t)node_index(tmap, nn_file, mine),
                                   node_index(tmap, nn_file, peer)};
            }
            qsort(nmap, as_int(l->num), sizeof(TagIdx), compare_tag);
            // the partner's faces by sorted node list
            const int np_ = cell_off[pe + 1] - cell_off[pe];
            Key *pk = teal_calloc(np_ + 1, sizeof(Key));
            for (int k = 0; k < np_; k++) {
                const int f = cell_off[pe] + k;
                pk[k].num = off[f + 1] - off[f];
                memcpy(pk[k].node, &idx[off[f]], sizeof(int) * pk[k].num);
                sort_ints(pk[k].node, pk[k].num);
                pk[k].cell = f;
            }
            qsort(pk, np_, sizeof(Key), compare_key);
            for (int f = cell_off[e]; f < cell_off[e + 1]; f++) {
                Key key = {off[f + 1] - off[f], {0}, f};
                for (int k = 0; k < key.num; k++) {
                    TagIdx probe = {(uint64_t)idx[off[f] + k], 0};
                    const TagIdx *hit = bsearch(&probe, nmap, as_int(l->num), sizeof(TagIdx),
                                                compare_tag);
                    if (!hit) {
                        teal_error("missing periodic peer node (%d, %d)", e, idx[off[f] + k]);
                    }
                    key.node[k] = hit->idx;
                }
                sort_ints(key.node, key.num);
                const Key *m = bsearch(&key, pk, np_, sizeof(Key), compare_key);
                if (!m) {
                    teal_error("missing periodic peer cell (%d, %d)", pe, f);
                }
                source[f - cells_periodic] = inner_cell_of(m->cell, off, idx, ncount, ncells);
            }
            teal_free(pk);
            teal_free(nmap);
        }
        teal_free(ncount);
        teal_free(ncells);
        teal_free(fill);
    }

    // nodes the cells use, in file order (read.c:770-846)
    int *new_of = teal_calloc(nn_file, sizeof(int));
    for (int i = 0; i < nn_file; i++) new_of[i] = -1;
    for (int k = 0; k < off[nc]; k++) new_of[idx[k]] = 0;
    int nn = 0;
    for (int i = 0; i < nn_file; i++) {
        if (new_of[i] == 0) {
            new_of[i] = nn++;
        }
    }

    Mesh *mesh = teal_calloc(1, sizeof(Mesh));
    mesh->nodes.num = nn;
    mesh->nodes.num_inner = nn;
    mesh->nodes.global = teal_calloc(nn, sizeof(long));
    mesh->nodes.coord = teal_calloc(nn, sizeof(Vector));
    for (int i = 0; i < nn_file; i++) {
        if (new_of[i] >= 0) {
            mesh->nodes.coord[new_of[i]] = coord_file[i];
            mesh->nodes.global[new_of[i]] = new_of[i];
        }
    }
    for (int k = 0; k < off[nc]; k++) idx[k] = new_of[idx[k]];
    mesh->cells.num = nc;
    mesh->cells.num_inner = cells_inner;
    mesh->cells.off_boundary = cells_periodic;
    mesh->cells.off_periodic = nc;
    mesh->cells.node.off = off;
    mesh->cells.node.idx = teal_realloc(idx, off[nc] > 0 ? off[nc] : 1, sizeof(int));

    mesh->entities.num = ne;
    mesh->entities.num_inner = num_inner;
    mesh->entities.off_boundary = off_bnd_ent;
    mesh->entities.name = name;
    mesh->entities.cell_off = cell_off;
    mesh->entities.rotation = rotation;
    mesh->entities.translation = translation;

    // one exchange entry per periodic entity: tag {E, partner}; E's ghosts receive the send
    // list of the entry whose send tag is E (the partner), which is E's own sources
    const int nper = nc - cells_periodic;
    mesh->neighbors.num = num_periodic;
    mesh->neighbors.tag = teal_calloc(num_periodic + 1, sizeof(int[2]));
    mesh->neighbors.rank = teal_calloc(num_periodic + 1, sizeof(int));
    mesh->neighbors.recv_off = teal_calloc(num_periodic + 1, sizeof(int));
    mesh->neighbors.send.off = teal_calloc(num_periodic + 1, sizeof(int));
    mesh->neighbors.send.idx = teal_calloc(nper + 1, sizeof(int));
    mesh->neighbors.recv_off[0] = cells_periodic;
    for (int m = 0; m < num_periodic; m++) {
        const int e = off_bnd_ent + m;
        const int pe = partner[e];
        mesh->neighbors.tag[m][0] = e;
        mesh->neighbors.tag[m][1] = pe;
        mesh->neighbors.recv_off[m + 1] = cell_off[e + 1];
        const int cnt = cell_off[e + 1] - cell_off[e];
        if (cell_off[pe + 1] - cell_off[pe] != cnt) {
            teal_error("periodic entities %d and %d differ in faces", e, pe);
        }
        mesh->neighbors.send.off[m + 1] = mesh->neighbors.send.off[m] + cnt;
        // the send list of E is what E's partner's ghosts receive: the partner's sources
        for (int k = 0; k < cnt; k++) {
            mesh->neighbors.send.idx[mesh->neighbors.send.off[m] + k] =
                source[(cell_off[pe] - cells_periodic) + k];
        }
    }

    teal_free(new_of);
    teal_free(source);
    teal_free(link_part);
    teal_free(partner);
    teal_free(tmap);
    teal_free(coord_file);
    teal_free(part);
    gmsh_free(&g);
    return mesh;
}
