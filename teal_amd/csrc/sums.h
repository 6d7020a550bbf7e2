// sums.h -- deterministic compensated reductions shared by the Krylov solver and the
// diagnostics: (sum, error) pairs accumulated with TwoSum / TwoProduct and combined in a fixed
// order, so every run gives the same bits and the result is within an ulp or so of the exact
// sum.  Workgroups of SUM_THREADS threads.  Compiled with -ffp-contract=off (the fma in
// add_product is the explicit TwoProduct error term).
#pragma once

#include <hip/hip_runtime.h>

constexpr int SUM_THREADS = 256;

__device__ inline double2 two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return make_double2(s, e);
}

__device__ inline double2 combine(double2 x, double2 y)
{
    double2 t = two_sum(x.x, y.x);
    return make_double2(t.x, (x.y + y.y) + t.y);
}

__device__ inline void add_value(double2 &acc, double x)
{
    double2 t = two_sum(acc.x, x);
    acc = make_double2(t.x, acc.y + t.y);
}

__device__ inline void add_product(double2 &acc, double a, double b)
{
    double h = a * b;
    double r = fma(a, b, -h);
    double2 t = two_sum(acc.x, h);
    acc = make_double2(t.x, acc.y + (t.y + r));
}

// block-wide reduction of (sum, error) pairs in a fixed order, the result in every thread: a
// butterfly across the 64 lanes of each wave (combine is commutative bit for bit: TwoSum's
// error term is exact, so both partners of an exchange hold the same pair), then the waves'
// pairs in wave order.  One barrier pair instead of one per tree level.
__device__ inline double2 block_reduce(double2 v)
{
    constexpr int WAVES = SUM_THREADS / 64;
    __shared__ double2 red[WAVES];
    for (int m = 32; m > 0; m >>= 1) {
        double2 o;
        o.x = __shfl_xor(v.x, m, 64);
        o.y = __shfl_xor(v.y, m, 64);
        v = combine(v, o);
    }
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = v;
    }
    __syncthreads();
    double2 out = red[0];
    for (int w = 1; w < WAVES; w++) {
        out = combine(out, red[w]);
    }
    __syncthreads();  // red is reused by the next call
    return out;
}

// total of `nb` partial pairs at stride `step`, the same in every block (fixed order)
__device__ inline double reduce_parts(const double2 *__restrict__ part, int nb, int step = 1)
{
    double2 acc = make_double2(0.0, 0.0);
    for (int b = threadIdx.x; b < nb; b += SUM_THREADS) {
        acc = combine(acc, part[(long)b * step]);
    }
    double2 tot = block_reduce(acc);
    return tot.x + tot.y;
}
