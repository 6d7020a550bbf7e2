// sync.c -- the single-process group of include/sync.h (kept apart from core.c, whose
// <unistd.h> declares the libc function named sync).
#include "sync.h"

#include <math.h>

struct sync sync = {0, 0, 1};

double sync_dot(const double *lhs, const double *rhs, int num)  // src/core/sync.c:66-74
{
    double dot = 0;
    for (int i = 0; i < num; i++) {
        dot += lhs[i] * rhs[i];
    }
    return dot;
}

double sync_norm(const double *buf, int num)  // src/core/sync.c:76-79
{
    return sqrt(sync_dot(buf, buf, num));
}
