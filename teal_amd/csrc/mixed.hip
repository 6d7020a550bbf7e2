// mixed.hip -- FP32 instances of the residual kernels (the mixed-precision path).
// Compiled with FP contraction enabled: this path makes no bit-compatibility claim; its
// error against the FP64 path is measured and reported instead.
#include "launch.h"

template struct Kern<float>;
