// sync.h -- the process-group globals of the drop-in library.
// Replaces the reference interface src/sync.h:5-9 for drivers that read `sync.rank` /
// `sync.size` (e.g. output prefixes): the library drives one GPU from one process, so the
// group is rank 0 of 1.  The MPI-typed collectives of the reference header are not provided by
// this library; the double-precision dot and norm are (src/core/sync.c:66-79, serial sums).
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

extern struct sync {
    int comm;  // no communicator: always 0
    int rank;
    int size;
} sync;

// Dot product of two double arrays, summed in index order (src/sync.h:29).
double sync_dot(const double *lhs, const double *rhs, int num);

// L2 norm of a double array (src/sync.h:32).
double sync_norm(const double *buf, int num);

#ifdef __cplusplus
}
#endif
