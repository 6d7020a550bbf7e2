// utils.h -- small helpers of the drop-in library.
// Replaces the reference interface src/utils.h:1-59 (same names, signatures, semantics).
#pragma once

#include <stddef.h>

#define sq(val) ((val) * (val))
#define cb(val) ((val) * (val) * (val))

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int left, right;
} Pair;

typedef struct {
    int x, y, z;
} Triple;

typedef struct {
    int *off, *idx;  // compressed sparse row
} Graph;

typedef char String[128];

typedef int Compare(const void *lhs, const void *rhs);

// src/utils.h:24
Compare compare_int, compare_long, compare_double;

// |lhs - rhs| <= max(1e-8, 1e-5 max(|lhs|, |rhs|)) (src/utils.h:27, src/core/utils.c:29-34).
int is_close(double lhs, double rhs);
int is_less(double lhs, double rhs);          // src/utils.h:30
int is_greater(double lhs, double rhs);       // src/utils.h:33
int is_less_equal(double lhs, double rhs);    // src/utils.h:36
int is_greater_equal(double lhs, double rhs); // src/utils.h:39

void *copy(void *dst, const void *src, int num, size_t size);  // src/utils.h:42
void *move(void *dst, const void *src, int num, size_t size);  // src/utils.h:45
void sort(void *base, int num, size_t size, Compare *compare); // src/utils.h:48
int unique(void *base, int num, size_t size, Compare *compare); // src/utils.h:51
void *search(const void *key, const void *base, int num, size_t size, Compare *compare); // :54
int digitize(const void *key, const void *base, int num, size_t size, Compare *compare); // :59

#ifdef __cplusplus
}
#endif
