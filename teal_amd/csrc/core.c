// core.c -- the host runtime of the drop-in library: start-up and shut-down, command-line
// options, messages, aligned allocation, and the helpers declared in include/teal.h,
// utils.h, vector.h and matrix.h.
//
// The interface is the reference's (src/teal.h, src/utils.h, src/vector.h, src/matrix.h):
// names, signatures, option letters, message formats and the floating-point association
// of the small math helpers.  The library runs one process on one device, which is rank 0.
#define _POSIX_C_SOURCE 200809L
#include <assert.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "matrix.h"
#include "teal.h"
#include "utils.h"
#include "vector.h"

struct teal teal = {0};

// ------------------------------------------------------------------ messages
// Every line goes out whole and is flushed at once; tagged lines carry "[rank] TAG: " as the
// reference's do (src/core/teal.c:96-113 format).
static void put_line(FILE *out, const char *tag, const char *fmt, va_list args)
{
    char text[4096];
    int len = tag ? snprintf(text, sizeof(text), "[0] %s: ", tag) : 0;
    vsnprintf(text + len, sizeof(text) - (size_t)len, fmt, args);
    fputs(text, out);
    fputc('\n', out);
    fflush(out);
}

void teal_print(const char *fmt, ...)
{
    assert(fmt);
    if (teal.quiet) {
        return;
    }
    va_list args;
    va_start(args, fmt);
    put_line(stdout, NULL, fmt, args);
    va_end(args);
}

void teal_verbose(const char *fmt, ...)
{
    assert(fmt);
    if (teal.quiet || !teal.verbose) {
        return;
    }
    va_list args;
    va_start(args, fmt);
    put_line(stdout, "VERBOSE", fmt, args);
    va_end(args);
}

void teal_error(const char *fmt, ...)
{
    assert(fmt);
    va_list args;
    va_start(args, fmt);
    put_line(stderr, "ERROR", fmt, args);
    va_end(args);
    teal_exit(EXIT_FAILURE);
}

void teal_exit(int status) { exit(status); }

// ------------------------------------------------------------------ start-up
static void usage(const char *program)
{
    static const char *const option[][2] = {
        {"-h", "show this help message and exit"},
        {"-q", "disable normal and verbose printing (errors are still shown)"},
        {"-v", "enable verbose printing"},
        {"-p", "enable partitioned output files"},
        {"-r <file>", "restart simulation from output file"},
    };
    char text[1024];
    int len = snprintf(text, sizeof(text), "usage: %s [options] ...\n\noptions:\n", program);
    for (size_t i = 0; i < sizeof(option) / sizeof(option[0]); i++) {
        len += snprintf(text + len, sizeof(text) - (size_t)len, "  %-20s %s\n", option[i][0],
                        option[i][1]);
    }
    teal_print("%s", text);  // the blank last line is the reference's output too
}

// Command line: -h, -q, -v, -p and -r <file>.  Flags may be grouped (-qv), the restart file
// may follow -r directly (-rfile), and "--" ends the options.  The options are taken out of
// argv; the other arguments keep their order after argv[0] and argv stays null-terminated.
// An unknown option, or -r without a file, prints the usage and fails.
static void take_options(int *argc, char ***argv)
{
    char **arg = *argv;
    int kept = 1;
    int options_done = 0;
    for (int i = 1; i < *argc; i++) {
        const char *a = arg[i];
        if (options_done || a[0] != '-' || a[1] == 0) {
            arg[kept++] = arg[i];
            continue;
        }
        if (strcmp(a, "--") == 0) {
            options_done = 1;
            continue;
        }
        for (const char *flag = a + 1; *flag; flag++) {
            if (*flag == 'q') {
                teal.quiet = 1;
            }
            else if (*flag == 'v') {
                teal.verbose = 1;
            }
            else if (*flag == 'p') {
                teal.partitioned = 1;
            }
            else if (*flag == 'r' && (flag[1] || i + 1 < *argc)) {
                teal.restart = flag[1] ? flag + 1 : arg[++i];
                break;
            }
            else {
                usage(arg[0]);
                teal_exit((*flag == 'h') ? EXIT_SUCCESS : EXIT_FAILURE);
            }
        }
    }
    *argc = kept;
    arg[kept] = NULL;
}

static void clock_text(String out)
{
    time_t now = time(NULL);
    struct tm local;
    localtime_r(&now, &local);
    strftime(out, sizeof(String), "%a %b %e %T %Y", &local);
}

void teal_init(int *argc, char ***argv)
{
    assert(argc && argv);
    take_options(argc, argv);
    String now;
    clock_text(now);
    teal_print("Hello, World! This is teal!");
    teal_print("\t start time      : %s", now);
    teal_print("\t number of ranks : %d", 1);
}

void teal_deinit(void)
{
    String now;
    clock_text(now);
    teal_print("Goodbye, World!");
    teal_print("\t stop time : %s", now);
}

// ------------------------------------------------------------------ allocation
// Every block starts on a 64-byte boundary (whole cache lines for the SoA copies made from
// it) and is preceded by one 64-byte header holding the payload size.  The header is the
// start of the aligned_alloc block, so freeing needs only the payload pointer.
enum { LINE = 64 };

typedef union {
    size_t bytes;
    unsigned char line[LINE];
} Header;

static Header *header_of(void *payload) { return (Header *)payload - 1; }

static void *block_alloc(int num, size_t size)
{
    const size_t room = SIZE_MAX - sizeof(Header) - LINE;
    if ((size_t)num > room / size) {
        teal_error("overflow (%d, %zu)", num, size);
    }
    const size_t bytes = (size_t)num * size;
    const size_t total = ((sizeof(Header) + bytes + LINE - 1) / LINE) * LINE;
    Header *head = aligned_alloc(LINE, total);
    if (!head) {
        teal_error("malloc failure (%zu)", total);
    }
    head->bytes = bytes;
    return head + 1;
}

void *teal_calloc(int num, size_t size)
{
    assert(num >= 0 && size > 0);
    if (num == 0) {
        return NULL;
    }
    void *payload = block_alloc(num, size);
    memset(payload, 0, (size_t)num * size);
    return payload;
}

void *teal_realloc(void *ptr, int num, size_t size)
{
    assert(num >= 0 && size > 0);
    if (num == 0) {
        teal_free(ptr);
        return NULL;
    }
    void *payload = block_alloc(num, size);
    if (ptr) {
        size_t keep = header_of(ptr)->bytes;
        size_t want = (size_t)num * size;
        memcpy(payload, ptr, (keep < want) ? keep : want);
        teal_free(ptr);
    }
    return payload;
}

void teal_free(void *ptr)
{
    if (ptr) {
        free(header_of(ptr));
    }
}

// ------------------------------------------------------------------ comparisons
// three-way comparisons for sort/unique/search/digitize (NaN compares equal to everything,
// as isgreater/isless make it in the reference's compare_double)
int compare_int(const void *lhs, const void *rhs)
{
    const int a = *(const int *)lhs, b = *(const int *)rhs;
    return (a < b) ? -1 : (a > b);
}

int compare_long(const void *lhs, const void *rhs)
{
    const long a = *(const long *)lhs, b = *(const long *)rhs;
    return (a < b) ? -1 : (a > b);
}

int compare_double(const void *lhs, const void *rhs)
{
    const double a = *(const double *)lhs, b = *(const double *)rhs;
    return isless(a, b) ? -1 : (isgreater(a, b) ? 1 : 0);
}

// |lhs - rhs| <= max(1e-8, 1e-5 max(|lhs|, |rhs|)), false when either is NaN
// (src/core/utils.c:29-34); the limiter and the LSQ weights branch on it, so the bounds and
// the operations are the reference's
int is_close(double lhs, double rhs)
{
    const double bound = fmax(1e-8, 1e-5 * fmax(fabs(lhs), fabs(rhs)));
    return islessequal(fabs(lhs - rhs), bound);
}

int is_less(double lhs, double rhs) { return (lhs < rhs) && !is_close(lhs, rhs); }
int is_greater(double lhs, double rhs) { return (lhs > rhs) && !is_close(lhs, rhs); }
int is_less_equal(double lhs, double rhs) { return (lhs < rhs) || is_close(lhs, rhs); }
int is_greater_equal(double lhs, double rhs) { return (lhs > rhs) || is_close(lhs, rhs); }

// ------------------------------------------------------------------ arrays
void *copy(void *dst, const void *src, int num, size_t size)
{
    assert(num >= 0 && size > 0);
    return (num > 0) ? memcpy(dst, src, (size_t)num * size) : NULL;
}

void *move(void *dst, const void *src, int num, size_t size)
{
    assert(num >= 0 && size > 0);
    return (num > 0) ? memmove(dst, src, (size_t)num * size) : NULL;
}

// Bottom-up merge sort between the array and one scratch copy: stable, so records that
// compare equal keep their order (the mesh numbering depends on that; the reference's glibc
// qsort is a merge sort in practice).
void sort(void *base, int num, size_t size, Compare *compare)
{
    assert((base || num == 0) && num >= 0 && size > 0 && compare);
    if (num < 2) {
        return;
    }
    char *from = base;
    char *to = malloc((size_t)num * size);
    if (!to) {
        teal_error("malloc failure (%zu)", (size_t)num * size);
    }
    char *const scratch = to;
    for (int width = 1; width < num; width *= 2) {
        for (int lo = 0; lo < num; lo += 2 * width) {
            int mid = (lo + width < num) ? lo + width : num;
            int hi = (mid + width < num) ? mid + width : num;
            int a = lo, b = mid, out = lo;
            while (a < mid || b < hi) {
                int pick_b = (a == mid) ||
                             (b < hi && compare(from + (size_t)b * size, from + (size_t)a * size) < 0);
                int src = pick_b ? b++ : a++;
                memcpy(to + (size_t)out++ * size, from + (size_t)src * size, size);
            }
        }
        char *swap = from;
        from = to;
        to = swap;
    }
    if (from != (char *)base) {
        memcpy(base, from, (size_t)num * size);
    }
    free(scratch);
}

int unique(void *base, int num, size_t size, Compare *compare)
{
    assert((base || num == 0) && num >= 0 && size > 0 && compare);
    if (num < 2) {
        return num;
    }
    sort(base, num, size, compare);
    char *rec = base;
    int last = 0;  // the newest distinct record
    for (int i = 1; i < num; i++) {
        char *cur = rec + (size_t)i * size;
        if (compare(cur, rec + (size_t)last * size) == 0) {
            continue;
        }
        last += 1;
        if (last != i) {
            memcpy(rec + (size_t)last * size, cur, size);
        }
    }
    return last + 1;
}

// a record of the sorted array that compares equal to key, or NULL
void *search(const void *key, const void *base, int num, size_t size, Compare *compare)
{
    assert(key && (base || num == 0) && num >= 0 && size > 0 && compare);
    const char *rec = base;
    int lo = 0, hi = num;  // candidates [lo, hi)
    while (lo < hi) {
        int mid = lo + ((hi - lo) / 2);
        int order = compare(key, rec + (size_t)mid * size);
        if (order == 0) {
            return (void *)(rec + (size_t)mid * size);
        }
        if (order < 0) {
            hi = mid;
        }
        else {
            lo = mid + 1;
        }
    }
    return NULL;
}

// Bin of key among the num + 1 sorted edges base[0..num]: -2 with no bins, -1 below
// base[0], num at or above base[num], else the i with base[i] <= key < base[i + 1]
// (src/utils.h:56-59).  Found as the first edge above key, minus one.
int digitize(const void *key, const void *base, int num, size_t size, Compare *compare)
{
    assert(key && (base || num == 0) && num >= 0 && size > 0 && compare);
    if (num == 0) {
        return -2;
    }
    const char *edge = base;
    int lo = 0, hi = num + 1;  // the first edge above key lies in [lo, hi]
    while (lo < hi) {
        int mid = lo + ((hi - lo) / 2);
        if (compare(edge + (size_t)mid * size, key) <= 0) {
            lo = mid + 1;
        }
        else {
            hi = mid;
        }
    }
    return lo - 1;
}

// ------------------------------------------------------------------ vectors and matrices
// Component-wise arithmetic; dot products associate as (xx + yy) + zz
// (src/core/vector.c:88-91), which every geometric quantity of the mesh inherits.
Vector vector_add(Vector a, Vector b) { return (Vector){a.x + b.x, a.y + b.y, a.z + b.z}; }
Vector vector_sub(Vector a, Vector b) { return (Vector){a.x - b.x, a.y - b.y, a.z - b.z}; }
Vector vector_mul(double s, Vector a) { return (Vector){s * a.x, s * a.y, s * a.z}; }
Vector vector_div(Vector a, double s) { return (Vector){a.x / s, a.y / s, a.z / s}; }
void vector_iadd(Vector *a, Vector b) { *a = vector_add(*a, b); }
void vector_isub(Vector *a, Vector b) { *a = vector_sub(*a, b); }
void vector_imul(Vector *a, double s) { *a = (Vector){a->x * s, a->y * s, a->z * s}; }
void vector_idiv(Vector *a, double s) { *a = vector_div(*a, s); }
Vector vector_abs(Vector a) { return (Vector){fabs(a.x), fabs(a.y), fabs(a.z)}; }
double vector_dot(Vector a, Vector b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
double vector_norm2(Vector a) { return vector_dot(a, a); }
double vector_norm(Vector a) { return sqrt(vector_norm2(a)); }

Vector vector_cross(Vector a, Vector b)
{
    Vector c;
    c.x = (a.y * b.z) - (a.z * b.y);
    c.y = (a.z * b.x) - (a.x * b.z);
    c.z = (a.x * b.y) - (a.y * b.x);
    return c;
}

// rows x, y, z (src/matrix.h:5-7)
Vector matrix_vector(Matrix m, Vector v)
{
    return (Vector){vector_dot(m.x, v), vector_dot(m.y, v), vector_dot(m.z, v)};
}

Matrix matrix_transpose(Matrix m)
{
    Matrix t;
    t.x = (Vector){m.x.x, m.y.x, m.z.x};
    t.y = (Vector){m.x.y, m.y.y, m.z.y};
    t.z = (Vector){m.x.z, m.y.z, m.z.z};
    return t;
}

Matrix matrix_matrix(Matrix a, Matrix b)
{
    const Matrix cols = matrix_transpose(b);
    return (Matrix){matrix_vector(cols, a.x), matrix_vector(cols, a.y), matrix_vector(cols, a.z)};
}

double matrix_determinant(Matrix m) { return vector_dot(m.x, vector_cross(m.y, m.z)); }

// adjugate rows divided by the determinant
Matrix matrix_inverse(Matrix m)
{
    const Vector r0 = vector_cross(m.y, m.z);
    const Vector r1 = vector_cross(m.z, m.x);
    const Vector r2 = vector_cross(m.x, m.y);
    const double det = vector_dot(m.x, r0);
    return (Matrix){vector_div(r0, det), vector_div(r1, det), vector_div(r2, det)};
}
