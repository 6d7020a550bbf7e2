/* CPU ORACLE (test infrastructure only) -- vector/matrix helpers.
 * Restates src/core/vector.c:10-110, src/core/matrix.c:7-33, src/core/utils.c:29-34
 * with the same association (e.g. dot = (xx + yy) + zz, vector.c:88-91). */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>

#include "oracle.h"

OVec ov_add(OVec a, OVec b) { return (OVec){a.x + b.x, a.y + b.y, a.z + b.z}; }
OVec ov_sub(OVec a, OVec b) { return (OVec){a.x - b.x, a.y - b.y, a.z - b.z}; }
OVec ov_mul(double s, OVec a) { return (OVec){s * a.x, s * a.y, s * a.z}; }
OVec ov_div(OVec a, double s) { return (OVec){a.x / s, a.y / s, a.z / s}; }
double ov_dot(OVec a, OVec b) { return (a.x * b.x) + (a.y * b.y) + (a.z * b.z); }
double ov_norm(OVec a) { return sqrt(ov_dot(a, a)); }
double ov_norm2(OVec a) { return ov_dot(a, a); }
OVec ov_abs(OVec a) { return (OVec){fabs(a.x), fabs(a.y), fabs(a.z)}; }

OVec ov_cross(OVec a, OVec b)
{
    return (OVec){(a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x)};
}

/* src/core/matrix.c:7-14 */
OVec om_vec(OMat m, OVec v) { return (OVec){ov_dot(m.x, v), ov_dot(m.y, v), ov_dot(m.z, v)}; }

/* src/core/matrix.c:26-33 */
OMat om_transpose(OMat m)
{
    return (OMat){{m.x.x, m.y.x, m.z.x}, {m.x.y, m.y.y, m.z.y}, {m.x.z, m.y.z, m.z.z}};
}

/* src/core/matrix.c:35-38 */
double om_det(OMat m) { return ov_dot(m.x, ov_cross(m.y, m.z)); }

/* src/core/utils.c:29-34: |a-b| <= max(1e-8, 1e-5 * max(|a|,|b|)) */
int o_is_close(double a, double b)
{
    return islessequal(fabs(a - b), fmax(1e-8, 1e-5 * fmax(fabs(a), fabs(b))));
}

void *o_calloc(size_t num, size_t size)
{
    if (num == 0) {
        return 0;
    }
    void *ptr = calloc(num, size);
    if (!ptr) {
        o_fatal("oracle: calloc failure (%zu x %zu)", num, size);
    }
    return ptr;
}

void o_fatal(const char *fmt, ...)
{
    va_list args;
    va_start(args, fmt);
    fputs("[0] ERROR: ", stderr);
    vfprintf(stderr, fmt, args);
    fputc('\n', stderr);
    va_end(args);
    exit(EXIT_FAILURE);
}
