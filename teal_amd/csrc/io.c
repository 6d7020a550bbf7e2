// io.c -- VTKHDF output and restart (src/mesh/write.c:185-211, src/equations/write.c:123-144,
// src/equations/read.c:180-197) on one process with serial HDF5.
//
// Files and layout are the reference's, so ParaView and the reference's own restart read them:
//   <prefix>_mesh.hdf   /VTKHDF (attributes Version = {1, 0}, Type = "UnstructuredGrid"):
//                       NumberOfPoints, NumberOfCells, NumberOfConnectivityIds, Points [n][3],
//                       Offsets, Connectivity, Types (VTK cell types), CellData/{vtkGhostType,
//                       entity, rank, volume, center [n][3]} over the cells [inner | boundary
//                       ghosts] (periodic ghosts are not written)
//   <prefix>_%05d.hdf   /VTKHDF with the same attributes, external links to the mesh file's
//                       datasets, FieldData/{time, one dataset per property} and CellData with
//                       the mesh links plus one dataset per primitive variable ([n] or [n][dim])
//                       and, with a reference solution, one per reference variable
// A restart reads the time and the properties, then gives each inner cell the state of the
// file's cell whose centre is nearest (any mesh of the same domain can restart from a file).
//
// HDF5 is loaded on first use from the image's serial build (TEAL_HDF5_LIB overrides the
// path), not linked: that keeps its directory off the library's search path, whose older
// C++ runtime would otherwise shadow the one the HIP runtime needs.
#define _GNU_SOURCE
#include <assert.h>
#include <dlfcn.h>
#include <float.h>
#include <hdf5.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "equations.h"
#include "mesh.h"
#include "plugins.h"
#include "teal.h"

// ------------------------------------------------------------------ HDF5 helpers
#define H5_FUNCTIONS(X)                                                                          \
    X(H5open) X(H5Fcreate) X(H5Fopen) X(H5Fclose) X(H5Gcreate2) X(H5Gopen2) X(H5Gclose)         \
    X(H5Screate_simple) X(H5Screate) X(H5Sclose) X(H5Sget_simple_extent_dims) X(H5Dcreate2)      \
    X(H5Dopen2) X(H5Dwrite) X(H5Dread) X(H5Dclose) X(H5Dget_space) X(H5Acreate2) X(H5Awrite)     \
    X(H5Aclose) X(H5Tcopy) X(H5Tset_size) X(H5Tset_strpad) X(H5Tclose) X(H5Lcreate_external)

// the values of H5F_ACC_TRUNC and H5F_ACC_RDONLY (whose macros call into the library)
enum { FILE_READ_ONLY = 0x0000u, FILE_TRUNCATE = 0x0002u };

static struct {
    void *lib;
#define FIELD(name) __typeof__(name) *name;
    H5_FUNCTIONS(FIELD)
#undef FIELD
    hid_t t_int, t_long, t_double, t_uchar, t_c_s1;
} h5;

static void *h5_symbol(const char *name)
{
    void *sym = dlsym(h5.lib, name);
    if (!sym) {
        teal_error("HDF5 symbol %s not found", name);
    }
    return sym;
}

static void h5_load(void)
{
    if (h5.lib) {
        return;
    }
    const char *path = getenv("TEAL_HDF5_LIB");
    path = path ? path : "/opt/conda/lib/libhdf5.so.103";
    h5.lib = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h5.lib) {
        teal_error("VTKHDF input/output needs HDF5 (%s)", dlerror());
    }
#define LOAD(name) *(void **)&h5.name = h5_symbol(#name);
    H5_FUNCTIONS(LOAD)
#undef LOAD
    if (h5.H5open() < 0) {
        teal_error("HDF5 initialization failed");
    }
    h5.t_int = *(hid_t *)h5_symbol("H5T_NATIVE_INT_g");
    h5.t_long = *(hid_t *)h5_symbol("H5T_NATIVE_LONG_g");
    h5.t_double = *(hid_t *)h5_symbol("H5T_NATIVE_DOUBLE_g");
    h5.t_uchar = *(hid_t *)h5_symbol("H5T_NATIVE_UCHAR_g");
    h5.t_c_s1 = *(hid_t *)h5_symbol("H5T_C_S1_g");
}

// Every HDF5 call is checked; a failure ends the program through teal_error, as the
// reference's error handler does (src/core/h5io.c:8-16).
static hid_t ok(hid_t id, const char *what, const char *name)
{
    if (id < 0) {
        teal_error("HDF5: %s failed (%s)", what, name);
    }
    return id;
}

static hid_t string_type(size_t len)
{
    hid_t t = ok(h5.H5Tcopy(h5.t_c_s1), "H5Tcopy", "string");
    h5.H5Tset_size(t, len);
    h5.H5Tset_strpad(t, H5T_STR_NULLPAD);
    return t;
}

// a dataset of `rows` x `cols` values (1-D when cols == 1)
static void put(hid_t loc, const char *name, const void *data, hsize_t rows, hsize_t cols, hid_t type)
{
    hsize_t dims[2] = {rows, cols};
    hid_t space = ok(h5.H5Screate_simple((cols == 1) ? 1 : 2, dims, 0), "H5Screate", name);
    hid_t set = ok(h5.H5Dcreate2(loc, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                   "H5Dcreate", name);
    if (rows * cols > 0) {
        ok(h5.H5Dwrite(set, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "H5Dwrite", name);
    }
    h5.H5Dclose(set);
    h5.H5Sclose(space);
}

// the first dimension of a dataset (and its second, or 1)
static hsize_t extent(hid_t loc, const char *name, hsize_t *cols)
{
    hid_t set = ok(h5.H5Dopen2(loc, name, H5P_DEFAULT), "H5Dopen", name);
    hid_t space = h5.H5Dget_space(set);
    hsize_t dims[2] = {0, 1};
    int rank = h5.H5Sget_simple_extent_dims(space, dims, 0);
    h5.H5Sclose(space);
    h5.H5Dclose(set);
    if (rank < 0 || rank > 2) {
        teal_error("HDF5: unexpected rank of %s", name);
    }
    if (cols) {
        *cols = (rank == 2) ? dims[1] : 1;
    }
    return (rank == 0) ? 1 : dims[0];
}

static void get(hid_t loc, const char *name, void *data, hid_t type)
{
    hid_t set = ok(h5.H5Dopen2(loc, name, H5P_DEFAULT), "H5Dopen", name);
    ok(h5.H5Dread(set, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "H5Dread", name);
    h5.H5Dclose(set);
}

static void vtkhdf_attributes(hid_t group)
{
    const int version[2] = {1, 0};
    hsize_t two = 2;
    hid_t space = h5.H5Screate_simple(1, &two, 0);
    hid_t attr = ok(h5.H5Acreate2(group, "Version", h5.t_int, space, H5P_DEFAULT, H5P_DEFAULT),
                    "H5Acreate", "Version");
    h5.H5Awrite(attr, h5.t_int, version);
    h5.H5Aclose(attr);
    h5.H5Sclose(space);
    const char type_name[16] = "UnstructuredGrid";  // 16 characters, no terminator
    hid_t str = string_type(sizeof(type_name));
    space = h5.H5Screate(H5S_SCALAR);
    attr = ok(h5.H5Acreate2(group, "Type", str, space, H5P_DEFAULT, H5P_DEFAULT), "H5Acreate", "Type");
    h5.H5Awrite(attr, str, type_name);
    h5.H5Aclose(attr);
    h5.H5Sclose(space);
    h5.H5Tclose(str);
}

static hid_t new_file(const char *fname, hid_t *vtkhdf)
{
    h5_load();
    hid_t file = h5.H5Fcreate(fname, FILE_TRUNCATE, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) {
        teal_error("could not create file (%s)", fname);
    }
    *vtkhdf = ok(h5.H5Gcreate2(file, "VTKHDF", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Gcreate",
                 "VTKHDF");
    vtkhdf_attributes(*vtkhdf);
    return file;
}

// ------------------------------------------------------------------ mesh file
// VTK cell types (vtkCellType.h) by node count: volume cells for the inner cells, faces for
// the boundary ghosts
static unsigned char vtk_type(int num_nodes, int volume)
{
    static const unsigned char by_nodes[2][9] = {
        {0, 0, 0, 5, 9, 0, 0, 0, 0},     // triangle, quad
        {0, 0, 0, 0, 10, 14, 13, 0, 12}, // tetra, pyramid, wedge, hexahedron
    };
    unsigned char t = (num_nodes >= 0 && num_nodes <= 8) ? by_nodes[volume][num_nodes] : 0;
    if (!t) {
        teal_error("invalid number of nodes (%d)", num_nodes);
    }
    return t;
}

void mesh_write(const Mesh *mesh, const char *name)
{
    assert(mesh && name);
    char fname[512];
    snprintf(fname, sizeof(fname), "%s_mesh.hdf", name);
    hid_t vtkhdf;
    hid_t file = new_file(fname, &vtkhdf);

    const int nc = mesh->cells.off_periodic;  // inner cells and boundary ghosts
    const int nn = mesh->nodes.num_inner;
    const int nidx = mesh->cells.node.off[nc];
    const long counts[3] = {nn, nc, nidx};
    put(vtkhdf, "NumberOfPoints", &counts[0], 1, 1, h5.t_long);
    put(vtkhdf, "NumberOfCells", &counts[1], 1, 1, h5.t_long);
    put(vtkhdf, "NumberOfConnectivityIds", &counts[2], 1, 1, h5.t_long);
    put(vtkhdf, "Points", mesh->nodes.coord, nn, 3, h5.t_double);
    put(vtkhdf, "Offsets", mesh->cells.node.off, nc + 1, 1, h5.t_int);
    int *conn = teal_calloc(nidx + 1, sizeof(int));
    for (int i = 0; i < nidx; i++) {  // one process: node numbers are the global numbers
        conn[i] = (int)mesh->nodes.global[mesh->cells.node.idx[i]];
    }
    put(vtkhdf, "Connectivity", conn, nidx, 1, h5.t_int);
    teal_free(conn);

    unsigned char *type = teal_calloc(nc + 1, 1);
    unsigned char *ghost = teal_calloc(nc + 1, 1);
    int *entity = teal_calloc(nc + 1, sizeof(int));
    int *rank = teal_calloc(nc + 1, sizeof(int));  // one process: rank 0 everywhere
    for (int c = 0; c < nc; c++) {
        const int inner = c < mesh->cells.num_inner;
        type[c] = vtk_type(mesh->cells.node.off[c + 1] - mesh->cells.node.off[c], inner);
        ghost[c] = inner ? 0 : (16 | 32);  // boundary ghosts: exterior and hidden
    }
    for (int e = 0; e < mesh->entities.num; e++) {
        for (int c = mesh->entities.cell_off[e]; c < mesh->entities.cell_off[e + 1] && c < nc; c++) {
            entity[c] = e;
        }
    }
    put(vtkhdf, "Types", type, nc, 1, h5.t_uchar);
    hid_t cells = ok(h5.H5Gcreate2(vtkhdf, "CellData", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                     "H5Gcreate", "CellData");
    put(cells, "vtkGhostType", ghost, nc, 1, h5.t_uchar);
    put(cells, "entity", entity, nc, 1, h5.t_int);
    put(cells, "rank", rank, nc, 1, h5.t_int);
    put(cells, "volume", mesh->cells.volume, nc, 1, h5.t_double);
    put(cells, "center", mesh->cells.center, nc, 3, h5.t_double);
    h5.H5Gclose(cells);
    teal_free(type);
    teal_free(ghost);
    teal_free(entity);
    teal_free(rank);
    h5.H5Gclose(vtkhdf);
    h5.H5Fclose(file);
}

// ------------------------------------------------------------------ solution files
// one dataset per variable of a set, its columns cut out of the rows (stride doubles each)
static void put_variables(hid_t loc, const EquationsVariables *var, int rows)
{
    double *buf = teal_calloc((size_t)rows * (var->stride + 1), sizeof(double));
    const double *data = var->data;
    for (int i = 0, col = 0; i < var->num; col += var->dimension[i], i++) {
        const int dim = var->dimension[i];
        for (int r = 0; r < rows; r++) {
            memcpy(&buf[(size_t)r * dim], &data[((size_t)r * var->stride) + col], sizeof(double) * dim);
        }
        put(loc, var->name[i], buf, rows, dim, h5.t_double);
    }
    teal_free(buf);
}

void equations_write(const Equations *eqns, const char *name, double time, int index)
{
    assert(eqns && name);
    char fname[512];
    snprintf(fname, sizeof(fname), "%s_%05d.hdf", name, index);
    hid_t vtkhdf;
    hid_t file = new_file(fname, &vtkhdf);

    // the mesh datasets live in the mesh file, linked by its name relative to this file
    char mesh_file[512];
    const char *slash = strrchr(name, '/');
    snprintf(mesh_file, sizeof(mesh_file), "%s_mesh.hdf", slash ? slash + 1 : name);
    const char *mesh_sets[] = {"NumberOfPoints", "NumberOfCells", "NumberOfConnectivityIds",
                               "Points", "Offsets", "Connectivity", "Types"};
    char target[128];
    for (size_t i = 0; i < sizeof(mesh_sets) / sizeof(mesh_sets[0]); i++) {
        snprintf(target, sizeof(target), "/VTKHDF/%s", mesh_sets[i]);
        ok(h5.H5Lcreate_external(mesh_file, target, vtkhdf, mesh_sets[i], H5P_DEFAULT, H5P_DEFAULT),
           "H5Lcreate_external", mesh_sets[i]);
    }

    hid_t fields = ok(h5.H5Gcreate2(vtkhdf, "FieldData", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                      "H5Gcreate", "FieldData");
    put(fields, "time", &time, 1, 1, h5.t_double);
    for (int i = 0; i < eqns->properties.num; i++) {
        put(fields, eqns->properties.name[i], &eqns->properties.data[i], 1, 1, h5.t_double);
    }
    h5.H5Gclose(fields);

    hid_t cells = ok(h5.H5Gcreate2(vtkhdf, "CellData", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                     "H5Gcreate", "CellData");
    const char *mesh_cells[] = {"vtkGhostType", "entity", "rank", "volume", "center"};
    for (size_t i = 0; i < sizeof(mesh_cells) / sizeof(mesh_cells[0]); i++) {
        snprintf(target, sizeof(target), "/VTKHDF/CellData/%s", mesh_cells[i]);
        ok(h5.H5Lcreate_external(mesh_file, target, cells, mesh_cells[i], H5P_DEFAULT, H5P_DEFAULT),
           "H5Lcreate_external", mesh_cells[i]);
    }
    // the ghost states belong to the written state (src/equations/write.c:109-113): periodic
    // ghosts copy their source cell (the exchange), boundary ghosts come from the BCs
    const int stride = eqns->primitive.stride;
    double *prim_rows = eqns->primitive.data;
    int *src = teal_calloc(eqns->mesh->cells.num + 1, sizeof(int));
    for (int c = 0; c < eqns->mesh->cells.num; c++) src[c] = -1;
    teal_mesh_periodic_sources(eqns->mesh, src);
    for (int c = 0; c < eqns->mesh->cells.num; c++) {
        if (src[c] >= 0) {
            memcpy(&prim_rows[(size_t)c * stride], &prim_rows[(size_t)src[c] * stride],
                   sizeof(double) * stride);
        }
    }
    teal_free(src);
    equations_boundary(eqns, eqns->primitive.data, time);
    const int rows = eqns->mesh->cells.off_periodic;
    put_variables(cells, &eqns->primitive, rows);
    Compute *exact = eqns->reference.func.compute;
    if (exact) {
        const double *prim = eqns->primitive.data;
        double *ref = eqns->reference.data;
        for (int c = 0; c < rows; c++) {
            exact(&ref[(size_t)c * stride], eqns->properties.data, eqns->mesh->cells.center[c],
                  time, &prim[(size_t)c * stride]);
        }
        put_variables(cells, &eqns->reference, rows);
    }
    h5.H5Gclose(cells);
    h5.H5Gclose(vtkhdf);
    h5.H5Fclose(file);
}

// ------------------------------------------------------------------ restart
// Nearest file cell of each inner cell: the file's centres sorted into a uniform grid of
// buckets; a query searches shells of buckets outward until the best distance is inside the
// searched region.  Exact ties keep the lowest file index.
typedef struct {
    double lo[3], h[3];
    int n[3];
    int *start, *item;  // bucket b holds item[start[b] .. start[b + 1])
    const Vector *point;
} Buckets;

static int bucket_coord(const Buckets *b, double x, int d)
{
    int i = (int)floor((x - b->lo[d]) / b->h[d]);
    return (i < 0) ? 0 : (i >= b->n[d] ? b->n[d] - 1 : i);
}

static void buckets_build(Buckets *b, const Vector *point, int num)
{
    double hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int d = 0; d < 3; d++) b->lo[d] = DBL_MAX;
    for (int i = 0; i < num; i++) {
        const double x[3] = {point[i].x, point[i].y, point[i].z};
        for (int d = 0; d < 3; d++) {
            b->lo[d] = fmin(b->lo[d], x[d]);
            hi[d] = fmax(hi[d], x[d]);
        }
    }
    const int per_axis = (int)ceil(cbrt((double)num)) + 1;
    size_t total = 1;
    for (int d = 0; d < 3; d++) {
        const double ext = hi[d] - b->lo[d];
        b->n[d] = (ext > 0) ? per_axis : 1;
        b->h[d] = (ext > 0) ? ext / b->n[d] * (1 + 1e-12) : 1;
        total *= (size_t)b->n[d];
    }
    b->start = teal_calloc(total + 1, sizeof(int));
    b->item = teal_calloc(num + 1, sizeof(int));
    b->point = point;
    int *slot = teal_calloc(num + 1, sizeof(int));
    for (int i = 0; i < num; i++) {
        const int bx = bucket_coord(b, point[i].x, 0), by = bucket_coord(b, point[i].y, 1),
                  bz = bucket_coord(b, point[i].z, 2);
        slot[i] = bx + (b->n[0] * (by + (b->n[1] * bz)));
        b->start[slot[i] + 1] += 1;
    }
    for (size_t s = 0; s < total; s++) b->start[s + 1] += b->start[s];
    int *fill = teal_calloc(total + 1, sizeof(int));
    for (int i = 0; i < num; i++) {  // ascending index within each bucket
        b->item[b->start[slot[i]] + fill[slot[i]]++] = i;
    }
    teal_free(fill);
    teal_free(slot);
}

static int buckets_nearest(const Buckets *b, Vector q)
{
    const double x[3] = {q.x, q.y, q.z};
    int c[3];
    for (int d = 0; d < 3; d++) c[d] = bucket_coord(b, x[d], d);
    double best = DBL_MAX;
    int arg = -1;
    const int reach = (b->n[0] > b->n[1] ? (b->n[0] > b->n[2] ? b->n[0] : b->n[2])
                                         : (b->n[1] > b->n[2] ? b->n[1] : b->n[2]));
    for (int r = 0; r <= reach; r++) {
        for (int k = c[2] - r; k <= c[2] + r; k++) {
            for (int j = c[1] - r; j <= c[1] + r; j++) {
                for (int i = c[0] - r; i <= c[0] + r; i++) {
                    if (i < 0 || j < 0 || k < 0 || i >= b->n[0] || j >= b->n[1] || k >= b->n[2]) {
                        continue;
                    }
                    const int on_shell = abs(i - c[0]) == r || abs(j - c[1]) == r || abs(k - c[2]) == r;
                    if (!on_shell) {
                        continue;
                    }
                    const int s = i + (b->n[0] * (j + (b->n[1] * k)));
                    for (int e = b->start[s]; e < b->start[s + 1]; e++) {
                        const Vector p = b->point[b->item[e]];
                        const double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
                        const double d2 = ((dx * dx) + (dy * dy)) + (dz * dz);
                        if (d2 < best || (d2 == best && b->item[e] < arg)) {
                            best = d2;
                            arg = b->item[e];
                        }
                    }
                }
            }
        }
        // every point outside the searched shells is at least r bucket widths away
        double clear = DBL_MAX;
        for (int d = 0; d < 3; d++) {
            const double below = x[d] - (b->lo[d] + ((c[d] - r) * b->h[d]));
            const double above = (b->lo[d] + ((c[d] + r + 1) * b->h[d])) - x[d];
            clear = fmin(clear, fmin(below, above));
        }
        if (arg >= 0 && clear > 0 && best <= clear * clear) {
            break;
        }
    }
    return arg;
}

void equations_read(const Equations *eqns, const char *fname, double *time, int *index)
{
    assert(eqns && fname && time && index);
    const char *underscore = strrchr(fname, '_');  // "<prefix>_NNNNN.hdf": continue at NNNNN + 1
    *index = underscore ? ((int)strtol(underscore + 1, 0, 10) + 1) : 0;
    h5_load();
    hid_t file = h5.H5Fopen(fname, FILE_READ_ONLY, H5P_DEFAULT);
    if (file < 0) {
        teal_error("could not open file (%s)", fname);
    }
    hid_t vtkhdf = ok(h5.H5Gopen2(file, "VTKHDF", H5P_DEFAULT), "H5Gopen", "VTKHDF");
    hid_t fields = ok(h5.H5Gopen2(vtkhdf, "FieldData", H5P_DEFAULT), "H5Gopen", "FieldData");
    get(fields, "time", time, h5.t_double);
    for (int i = 0; i < eqns->properties.num; i++) {
        get(fields, eqns->properties.name[i], &eqns->properties.data[i], h5.t_double);
    }
    h5.H5Gclose(fields);

    long num = 0;
    get(vtkhdf, "NumberOfCells", &num, h5.t_long);
    if (num <= 0 || num > INT_MAX) {
        teal_error("invalid number of cells in %s (%ld)", fname, num);
    }
    hid_t cells = ok(h5.H5Gopen2(vtkhdf, "CellData", H5P_DEFAULT), "H5Gopen", "CellData");
    Vector *center = teal_calloc(num, sizeof(Vector));
    get(cells, "center", center, h5.t_double);
    const EquationsVariables *prim = &eqns->primitive;
    double *state = teal_calloc((size_t)num * prim->stride, sizeof(double));
    double *buf = teal_calloc((size_t)num * prim->stride, sizeof(double));
    for (int i = 0, col = 0; i < prim->num; col += prim->dimension[i], i++) {
        hsize_t cols = 1;
        if (extent(cells, prim->name[i], &cols) != (hsize_t)num || cols != (hsize_t)prim->dimension[i]) {
            teal_error("dataset dims mismatch (%s)", prim->name[i]);
        }
        get(cells, prim->name[i], buf, h5.t_double);
        for (long r = 0; r < num; r++) {
            memcpy(&state[((size_t)r * prim->stride) + col], &buf[(size_t)r * prim->dimension[i]],
                   sizeof(double) * prim->dimension[i]);
        }
    }
    h5.H5Gclose(cells);
    h5.H5Gclose(vtkhdf);
    h5.H5Fclose(file);

    Buckets b;
    buckets_build(&b, center, (int)num);
    double *dst = prim->data;
    for (int c = 0; c < eqns->mesh->cells.num_inner; c++) {
        const int src = buckets_nearest(&b, eqns->mesh->cells.center[c]);
        memcpy(&dst[(size_t)c * prim->stride], &state[(size_t)src * prim->stride],
               sizeof(double) * prim->stride);
    }
    teal_free(b.start);
    teal_free(b.item);
    teal_free(buf);
    teal_free(state);
    teal_free(center);
}

// ------------------------------------------------------------------ polar file
void teal_io_write_polar(const char *fname, double alpha_deg, const Vector *center,
                         const double *pressure_c, int num, double lift, double drag)
{
    h5_load();
    hid_t file = h5.H5Fcreate(fname, FILE_TRUNCATE, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) {
        teal_error("could not create file (%s)", fname);
    }
    put(file, "alpha", &alpha_deg, 1, 1, h5.t_double);
    put(file, "center", center, num, 3, h5.t_double);
    put(file, "pressure_c", pressure_c, num, 1, h5.t_double);
    put(file, "lift_c", &lift, 1, 1, h5.t_double);
    put(file, "drag_c", &drag, 1, 1, h5.t_double);
    h5.H5Fclose(file);
}
