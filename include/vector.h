// vector.h -- 3-vectors. Replaces the reference interface src/vector.h:1-53.
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double x, y, z;
} Vector;

typedef struct {
    Vector left, right;
} VectorPair;

typedef struct {
    Vector unit;
    double norm;
} VectorNorm;

Vector vector_add(Vector lhs, Vector rhs);  // src/vector.h:17
Vector vector_sub(Vector lhs, Vector rhs);  // src/vector.h:20
Vector vector_mul(double lhs, Vector rhs);  // src/vector.h:23
Vector vector_div(Vector lhs, double rhs);  // src/vector.h:26
void vector_iadd(Vector *lhs, Vector rhs);  // src/vector.h:29
void vector_isub(Vector *lhs, Vector rhs);  // src/vector.h:32
void vector_imul(Vector *lhs, double rhs);  // src/vector.h:35
void vector_idiv(Vector *lhs, double rhs);  // src/vector.h:38
Vector vector_abs(Vector vec);              // src/vector.h:41
double vector_dot(Vector lhs, Vector rhs);  // (xx + yy) + zz, src/vector.h:44
double vector_norm(Vector vec);             // src/vector.h:47
double vector_norm2(Vector vec);            // src/vector.h:50
Vector vector_cross(Vector lhs, Vector rhs); // src/vector.h:53

#ifdef __cplusplus
}
#endif
