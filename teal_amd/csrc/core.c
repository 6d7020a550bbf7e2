// core.c -- runtime, logging, allocation, vector/matrix math and sorting helpers.
// Implements include/teal.h, utils.h, vector.h, matrix.h with the semantics of
// src/core/teal.c, src/core/utils.c, src/core/vector.c and src/core/matrix.c for one process
// (the single rank is rank 0; there is no MPI).
#define _POSIX_C_SOURCE 200809L
#include <assert.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "matrix.h"
#include "teal.h"
#include "utils.h"
#include "vector.h"

struct teal teal = {0};

enum { ALIGN = 64 };

static void print_help(const char *prog)
{
    teal_print("usage: %s [options] ...\n\n"
               "options:\n"
               "  %-20s show this help message and exit\n"
               "  %-20s disable normal and verbose printing (errors are still shown)\n"
               "  %-20s enable verbose printing\n"
               "  %-20s enable partitioned output files\n"
               "  %-20s restart simulation from output file\n",
               prog, "-h", "-q", "-v", "-p", "-r <file>");
}

// option parsing as src/core/teal.c:37-60 (-h -q -v -p -r <file>), positional args kept
static void parse_options(int *argc, char ***argv)
{
    opterr = 0;
    optind = 1;
    int opt;
    while ((opt = getopt(*argc, *argv, "hqvpr:")) != -1) {
        switch (opt) {
            case 'h': print_help((*argv)[0]); teal_exit(EXIT_SUCCESS);
            case 'q': teal.quiet = 1; break;
            case 'v': teal.verbose = 1; break;
            case 'p': teal.partitioned = 1; break;
            case 'r': teal.restart = optarg; break;
            default: print_help((*argv)[0]); teal_exit(EXIT_FAILURE);
        }
    }
    int num = 1;
    for (int i = optind; i < *argc; i++) {
        (*argv)[num++] = (*argv)[i];
    }
    *argc = num;
    (*argv)[num] = 0;
}

static void now_string(char *buf, size_t len)
{
    time_t t = time(0);
    struct tm tm;
    localtime_r(&t, &tm);
    strftime(buf, len, "%a %b %e %T %Y", &tm);
}

void teal_init(int *argc, char ***argv)
{
    assert(argc && argv);
    parse_options(argc, argv);
    String now;
    now_string(now, sizeof(now));
    teal_print("Hello, World! This is teal!");
    teal_print("\t start time      : %s", now);
    teal_print("\t number of ranks : %d", 1);
}

void teal_deinit(void)
{
    String now;
    now_string(now, sizeof(now));
    teal_print("Goodbye, World!");
    teal_print("\t stop time : %s", now);
}

void teal_exit(int status) { exit(status); }

static void println(FILE *stream, const char *prefix, const char *fmt, va_list args)
{
    if (prefix) {
        fprintf(stream, "[0] %s: ", prefix);
    }
    vfprintf(stream, fmt, args);
    fputc('\n', stream);
    fflush(stream);
}

void teal_print(const char *fmt, ...)
{
    if (!teal.quiet) {
        va_list args;
        va_start(args, fmt);
        println(stdout, 0, fmt, args);
        va_end(args);
    }
}

void teal_verbose(const char *fmt, ...)
{
    if (!teal.quiet && teal.verbose) {
        va_list args;
        va_start(args, fmt);
        println(stdout, "VERBOSE", fmt, args);
        va_end(args);
    }
}

void teal_error(const char *fmt, ...)
{
    va_list args;
    va_start(args, fmt);
    println(stderr, "ERROR", fmt, args);
    va_end(args);
    teal_exit(EXIT_FAILURE);
}

// aligned allocation with a header that remembers the base pointer and size
typedef struct {
    void *base;
    size_t bytes;
} Alloc;

static void *teal_malloc(int num, size_t size)
{
    size_t extra = sizeof(Alloc) + (ALIGN - 1);
    if ((size_t)num > (SIZE_MAX - extra) / size) {
        teal_error("overflow (%d, %zu)", num, size);
    }
    size_t bytes = (size_t)num * size;
    char *base = malloc(bytes + extra);
    if (!base) {
        teal_error("malloc failure (%zu)", bytes + extra);
    }
    char *beg = base + sizeof(Alloc);
    char *ptr = beg + (-(uintptr_t)beg & (ALIGN - 1));
    ((Alloc *)ptr)[-1].base = base;
    ((Alloc *)ptr)[-1].bytes = bytes;
    return ptr;
}

void *teal_calloc(int num, size_t size)
{
    assert(num >= 0 && size > 0);
    if (num == 0) {
        return 0;
    }
    void *ptr = teal_malloc(num, size);
    return memset(ptr, 0, (size_t)num * size);
}

void *teal_realloc(void *ptr, int num, size_t size)
{
    assert(num >= 0 && size > 0);
    if (num == 0) {
        teal_free(ptr);
        return 0;
    }
    if (!ptr) {
        return teal_malloc(num, size);
    }
    void *fresh = teal_malloc(num, size);
    size_t bytes = (size_t)num * size;
    if (((Alloc *)ptr)[-1].bytes < bytes) {
        bytes = ((Alloc *)ptr)[-1].bytes;
    }
    memcpy(fresh, ptr, bytes);
    free(((Alloc *)ptr)[-1].base);
    return fresh;
}

void teal_free(void *ptr)
{
    if (ptr) {
        free(((Alloc *)ptr)[-1].base);
    }
}

// ---- utils (src/core/utils.c) ----
int compare_int(const void *lhs_, const void *rhs_)
{
    const int *lhs = lhs_;
    const int *rhs = rhs_;
    return (*lhs > *rhs) - (*lhs < *rhs);
}

int compare_long(const void *lhs_, const void *rhs_)
{
    const long *lhs = lhs_;
    const long *rhs = rhs_;
    return (*lhs > *rhs) - (*lhs < *rhs);
}

int compare_double(const void *lhs_, const void *rhs_)
{
    const double *lhs = lhs_;
    const double *rhs = rhs_;
    return isgreater(*lhs, *rhs) - isless(*lhs, *rhs);
}

int is_close(double lhs, double rhs)
{
    return islessequal(fabs(lhs - rhs), fmax(1e-8, 1e-5 * fmax(fabs(lhs), fabs(rhs))));
}

int is_less(double lhs, double rhs) { return !is_close(lhs, rhs) && lhs < rhs; }
int is_greater(double lhs, double rhs) { return !is_close(lhs, rhs) && lhs > rhs; }
int is_less_equal(double lhs, double rhs) { return is_close(lhs, rhs) || lhs < rhs; }
int is_greater_equal(double lhs, double rhs) { return is_close(lhs, rhs) || lhs > rhs; }

void *copy(void *dst, const void *src, int num, size_t size)
{
    if (num == 0) {
        return 0;
    }
    return memcpy(dst, src, (size_t)num * size);
}

void *move(void *dst, const void *src, int num, size_t size)
{
    if (num == 0) {
        return 0;
    }
    return memmove(dst, src, (size_t)num * size);
}

// stable merge sort (the reference relies on glibc's qsort, whose merge-sort path is stable)
static void merge_sort(char *base, char *tmp, int num, size_t size, Compare *compare)
{
    if (num < 2) {
        return;
    }
    int half = num / 2;
    merge_sort(base, tmp, half, size, compare);
    merge_sort(base + ((size_t)half * size), tmp, num - half, size, compare);
    int i = 0;
    int j = half;
    int k = 0;
    while (i < half && j < num) {
        if (compare(base + ((size_t)j * size), base + ((size_t)i * size)) < 0) {
            memcpy(tmp + ((size_t)k++ * size), base + ((size_t)j++ * size), size);
        }
        else {
            memcpy(tmp + ((size_t)k++ * size), base + ((size_t)i++ * size), size);
        }
    }
    while (i < half) {
        memcpy(tmp + ((size_t)k++ * size), base + ((size_t)i++ * size), size);
    }
    while (j < num) {
        memcpy(tmp + ((size_t)k++ * size), base + ((size_t)j++ * size), size);
    }
    memcpy(base, tmp, (size_t)num * size);
}

void sort(void *base, int num, size_t size, Compare *compare)
{
    if (num <= 1) {
        return;
    }
    char *tmp = malloc((size_t)num * size);
    if (!tmp) {
        teal_error("malloc failure (%zu)", (size_t)num * size);
    }
    merge_sort(base, tmp, num, size, compare);
    free(tmp);
}

int unique(void *base_, int num_, size_t size, Compare *compare)
{
    if (num_ <= 1) {
        return num_;
    }
    sort(base_, num_, size, compare);
    char *base = base_;
    int num = 1;
    for (int i = 1; i < num_; i++) {
        if (compare(base + ((size_t)i * size), base + ((size_t)(num - 1) * size)) != 0) {
            if (num != i) {
                memcpy(base + ((size_t)num * size), base + ((size_t)i * size), size);
            }
            num += 1;
        }
    }
    return num;
}

void *search(const void *key, const void *base, int num, size_t size, Compare *compare)
{
    if (num == 0) {
        return 0;
    }
    return bsearch(key, base, num, size, compare);
}

int digitize(const void *key, const void *base_, int num, size_t size, Compare *compare)
{
    if (num == 0) {
        return -2;
    }
    const char *base = base_;
    if (compare(key, base) < 0) {
        return -1;
    }
    if (compare(key, base + size) < 0) {
        return 0;
    }
    if (compare(key, base + ((size_t)num * size)) >= 0) {
        return num;
    }
    int left = 1;
    int right = num;
    while (left < right) {
        int middle = left + ((right - left) / 2);
        if (compare(base + ((size_t)middle * size), key) <= 0) {
            left = middle + 1;
        }
        else {
            right = middle;
        }
    }
    return left - 1;
}

// ---- vector (src/core/vector.c:10-110) ----
Vector vector_add(Vector a, Vector b) { return (Vector){a.x + b.x, a.y + b.y, a.z + b.z}; }
Vector vector_sub(Vector a, Vector b) { return (Vector){a.x - b.x, a.y - b.y, a.z - b.z}; }
Vector vector_mul(double s, Vector a) { return (Vector){s * a.x, s * a.y, s * a.z}; }
Vector vector_div(Vector a, double s) { return (Vector){a.x / s, a.y / s, a.z / s}; }

void vector_iadd(Vector *a, Vector b)
{
    a->x += b.x;
    a->y += b.y;
    a->z += b.z;
}

void vector_isub(Vector *a, Vector b)
{
    a->x -= b.x;
    a->y -= b.y;
    a->z -= b.z;
}

void vector_imul(Vector *a, double s)
{
    a->x *= s;
    a->y *= s;
    a->z *= s;
}

void vector_idiv(Vector *a, double s)
{
    a->x /= s;
    a->y /= s;
    a->z /= s;
}

Vector vector_abs(Vector a) { return (Vector){fabs(a.x), fabs(a.y), fabs(a.z)}; }
double vector_dot(Vector a, Vector b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
double vector_norm(Vector a) { return sqrt(vector_dot(a, a)); }
double vector_norm2(Vector a) { return vector_dot(a, a); }

Vector vector_cross(Vector a, Vector b)
{
    return (Vector){(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
}

// ---- matrix (src/core/matrix.c:7-54) ----
Vector matrix_vector(Matrix m, Vector v)
{
    return (Vector){vector_dot(m.x, v), vector_dot(m.y, v), vector_dot(m.z, v)};
}

Matrix matrix_transpose(Matrix m)
{
    return (Matrix){{m.x.x, m.y.x, m.z.x}, {m.x.y, m.y.y, m.z.y}, {m.x.z, m.y.z, m.z.z}};
}

Matrix matrix_matrix(Matrix a, Matrix b)
{
    Matrix bt = matrix_transpose(b);
    return (Matrix){matrix_vector(bt, a.x), matrix_vector(bt, a.y), matrix_vector(bt, a.z)};
}

double matrix_determinant(Matrix m) { return vector_dot(m.x, vector_cross(m.y, m.z)); }

Matrix matrix_inverse(Matrix m)
{
    Matrix adj = {vector_cross(m.y, m.z), vector_cross(m.z, m.x), vector_cross(m.x, m.y)};
    double det = vector_dot(m.x, adj.x);
    return (Matrix){vector_div(adj.x, det), vector_div(adj.y, det), vector_div(adj.z, det)};
}
