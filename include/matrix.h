// matrix.h -- row-major 3x3 matrices. Replaces the reference interface src/matrix.h:1-22.
#pragma once

#include "vector.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    Vector x, y, z;  // row-major order
} Matrix;

Vector matrix_vector(Matrix mat, Vector vec);    // src/matrix.h:10
Matrix matrix_matrix(Matrix lhs, Matrix rhs);    // src/matrix.h:13
Matrix matrix_transpose(Matrix mat);             // src/matrix.h:16
double matrix_determinant(Matrix mat);           // src/matrix.h:19
Matrix matrix_inverse(Matrix mat);               // src/matrix.h:22

#ifdef __cplusplus
}
#endif
