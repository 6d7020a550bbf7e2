// teal.h -- runtime entry points of the drop-in library.
// Replaces the reference interface src/teal.h:1-39 (same names, signatures, semantics).
// One process drives one GPU; there are no MPI ranks, so "rank 0" is the only rank.
#pragma once

#include <stddef.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846 /* pi */
#endif

#ifdef __cplusplus
extern "C" {
#endif

// Global flags set by teal_init (src/teal.h:7-12).
extern struct teal {
    int quiet;
    int verbose;
    int partitioned;
    const char *restart;
} teal;

// Initialize teal; call once at program begin (src/teal.h:15). Parses -h -q -v -p -r <file>.
void teal_init(int *argc, char ***argv);

// Finalize teal; call once at program end (src/teal.h:18).
void teal_deinit(void);

// Early exit with a status code (src/teal.h:21).
void teal_exit(int status) __attribute__((noreturn));

// Print from the (single) rank to stdout unless quiet (src/teal.h:24).
void teal_print(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

// Verbose print, "[0] VERBOSE: ..." (src/teal.h:27).
void teal_verbose(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

// Print "[0] ERROR: ..." to stderr and exit with EXIT_FAILURE (src/teal.h:30).
void teal_error(const char *fmt, ...) __attribute__((format(printf, 1, 2), noreturn));

// 64-byte aligned zeroed allocation; aborts through teal_error on failure (src/teal.h:33).
void *teal_calloc(int num, size_t size) __attribute__((malloc));

// Reallocate preserving contents (src/teal.h:36).
void *teal_realloc(void *ptr, int num, size_t size);

// Free memory from teal_calloc / teal_realloc (src/teal.h:39).
void teal_free(void *ptr);

#ifdef __cplusplus
}
#endif
