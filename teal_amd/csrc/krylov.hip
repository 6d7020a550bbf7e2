// krylov.hip -- implicit Euler with Jacobian-free Newton-GMRES on the device
// (src/simulation/advance.c:119-195, src/simulation/gmres.c:12-149).
//
// Vectors (U, U0, derivative, Newton residual, increment, the Krylov basis) are FP64
// [5][num_inner] arrays in device cell order on both precisions; the matvec's residual
// evaluations run on the configured path (FP64 or the mixed FP32 flux path).  The Newton loop
// runs on the host; the GMRES iterations, including the Hessenberg matrix, the Givens
// rotations, the convergence test and the back substitution, run on the device in the
// reference's operation order, queued by the host ahead of their results.  Every vector
// operation has the reference's per-element expression.  The reference's dot products are
// serial sums.  By default they are deterministic two-level reductions with compensated
// (TwoSum / TwoProduct) accumulation, which round differently from the reference by at most
// its own summation error; the finite-difference matvec amplifies such last-bit differences
// (a perturbed state that rounds differently changes R(P1) - R by an ulp, divided by eps).
// TEAL_REDUCE_SERIAL sums in the reference's order on one workgroup instead: bit-for-bit with
// the reference, for verification on small meshes.  Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "coherent.h"
#include "device.h"
#include "sums.h"

using namespace fv;

#define HIPCALL(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            teal_gpu_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                               __FILE__, __LINE__);                                       \
            return -1;                                                                    \
        }                                                                                 \
    } while (0)

// Finite-difference scale of the matvec when the residual is evaluated in FP32.  The
// reference scales the perturbation by the global norm |U|, which the energy dominates, so the
// momentum components see a perturbation (|E|/|rho u|) times larger than this factor.
#ifndef TEAL_MIXED_FD_RELATIVE
#define TEAL_MIXED_FD_RELATIVE 3.4526698300124393e-4
#endif

namespace {

constexpr int RT = 256;        // threads of the vector kernels
static_assert(RT == SUM_THREADS, "the reductions of sums.h assume RT threads");
constexpr int MAX_PARTS = 1024;  // blocks of a reduction (partial sums read by every block)

// ---------------------------------------------------------------- kernels
// part[block] = partial dot(a, b)
__global__ __launch_bounds__(RT) void k_dot(const double *__restrict__ a, const double *__restrict__ b,
                                            long n, double2 *__restrict__ part, const int *skip)
{
    if (skip && *skip) {
        return;
    }
    double2 acc = make_double2(0.0, 0.0);
    for (long e = (blockIdx.x * (long)RT) + threadIdx.x; e < n; e += (long)gridDim.x * RT) {
        add_product(acc, a[e], b[e]);
    }
    double2 tot = block_reduce(acc);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = tot;
    }
}

// out[idx] = total (or its square root)
__global__ __launch_bounds__(RT) void k_reduce(const double2 *__restrict__ part, int nb, double *out,
                                               int idx, int take_sqrt)
{
    double t = reduce_parts(part, nb);
    if (threadIdx.x == 0) {
        out[idx] = take_sqrt ? sqrt(t) : t;
    }
}

// out[idx] = sqrt(total), the last entry of a Hessenberg column (k_reduce with take_sqrt)
__global__ __launch_bounds__(RT) void k_reduce_column(const double2 *__restrict__ part, int nb,
                                                      double *out, int idx, const int *skip)
{
    if (*skip) {
        return;
    }
    double t = reduce_parts(part, nb);
    if (threadIdx.x == 0) {
        out[idx] = sqrt(t);
    }
}

// One modified Gram-Schmidt step (gmres.c:92-99): h = the dot whose partials are in part_in
// (hess[i-1][iter]); w -= h * v_prev; partials of dot(v_next, w) (v_next == w: the norm).
__global__ __launch_bounds__(RT) void k_mgs(double *__restrict__ w, const double *__restrict__ v_prev,
                                            const double *v_next, long n,
                                            const double2 *__restrict__ part_in, int nb,
                                            double2 *__restrict__ part_out, double *h, int hidx,
                                            const int *skip)
{
    if (*skip) {
        return;
    }
    // the operands of a thread's first PF elements are loaded before the reduction of
    // part_in, so that their latency and the reduction's overlap (same order of operations)
    constexpr int PF = 8;
    const long stride = (long)gridDim.x * RT;
    const long e0 = (blockIdx.x * (long)RT) + threadIdx.x;
    double wp[PF], vp[PF], np[PF];
#pragma unroll
    for (int k = 0; k < PF; k++) {
        const long e = e0 + (k * stride);
        if (e < n) {
            wp[k] = w[e];
            vp[k] = v_prev[e];
            np[k] = v_next ? v_next[e] : 0.0;
        }
    }
    const double hv = reduce_parts(part_in, nb);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        h[hidx] = hv;
    }
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < PF; k++) {
        const long e = e0 + (k * stride);
        if (e < n) {
            double x = wp[k] - (hv * vp[k]);
            w[e] = x;
            add_product(acc, v_next ? np[k] : x, x);
        }
    }
    for (long e = e0 + (PF * stride); e < n; e += stride) {
        double x = w[e] - (hv * v_prev[e]);
        w[e] = x;
        add_product(acc, v_next ? v_next[e] : x, x);
    }
    double2 tot = block_reduce(acc);
    if (threadIdx.x == 0) {
        part_out[blockIdx.x] = tot;
    }
}

// Serial dot product in the reference's order (src/core/sync.c:66-74: cell by cell in
// reference numbering, the 5 variables of a cell in turn, one running sum): one workgroup;
// the threads stage the products of 256 cells in LDS, thread 0 adds them in order.
// out[idx] = sum, its square root (mode 1) or fd_scale / sqrt(sum) (mode 2).
__global__ __launch_bounds__(RT) void k_serial_dot(const double *__restrict__ a,
                                                   const double *__restrict__ b,
                                                   const int *__restrict__ dev_of_ref, int ni,
                                                   double *out, int idx, int mode, double fd_scale,
                                                   const int *skip)
{
    if (skip && *skip) {
        return;
    }
    __shared__ double prod[RT * 5];
    double sum = 0.0;
    for (int base = 0; base < ni; base += RT) {
        int i = base + threadIdx.x;
        if (i < ni) {
            long d = dev_of_ref[i];
            for (int k = 0; k < 5; k++) {
                prod[(threadIdx.x * 5) + k] = a[((long)k * ni) + d] * b[((long)k * ni) + d];
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int cnt = (ni - base < RT) ? ni - base : RT;
            for (int e = 0; e < cnt * 5; e++) {
                sum = sum + prod[e];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[idx] = (mode == 0) ? sum : (mode == 1) ? sqrt(sum) : fd_scale / sqrt(sum);
    }
}

// residual norm in the reference's order (src/equations/residual.c:22-32): one running sum
// per variable over the inner cells in reference numbering; out[v] = sqrt(sum / sum_volume)
__global__ __launch_bounds__(RT) void k_serial_residual(const double *__restrict__ R,
                                                        const double *__restrict__ vol,
                                                        const int *__restrict__ dev_of_ref, int ni,
                                                        double sum_volume, double *out)
{
    __shared__ double prod[RT * 5];
    double sum = 0.0;  // thread v < 5 holds variable v
    for (int base = 0; base < ni; base += RT) {
        int i = base + threadIdx.x;
        if (i < ni) {
            long d = dev_of_ref[i];
            for (int k = 0; k < 5; k++) {
                double r = R[((long)k * ni) + d];
                prod[(k * RT) + threadIdx.x] = vol[d] * (r * r);
            }
        }
        __syncthreads();
        if (threadIdx.x < 5) {
            int cnt = (ni - base < RT) ? ni - base : RT;
            for (int e = 0; e < cnt; e++) {
                sum = sum + prod[(threadIdx.x * RT) + e];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) {
        out[threadIdx.x] = sqrt(sum / sum_volume);
    }
}

// w -= h[hidx] * v (gmres.c:96-98) with h from device memory
__global__ __launch_bounds__(RT) void k_axpy_h(double *__restrict__ w, const double *__restrict__ v,
                                               const double *h, int hidx, long n, const int *skip)
{
    if (*skip) {
        return;
    }
    const double hv = h[hidx];
    for (long e = (blockIdx.x * (long)RT) + threadIdx.x; e < n; e += (long)gridDim.x * RT) {
        w[e] = w[e] - (hv * v[e]);
    }
}

// matvec, first half (gmres.c:26-36): eps = fd_scale / |v| (from the partials of dot(v, v),
// or, without partials, already in *eps_io); P = cons2prim(U + eps v) for the inner cells
template <typename S>
__global__ __launch_bounds__(RT) void k_perturb(const double *__restrict__ U, const double *__restrict__ v,
                                                int ni, const double2 *__restrict__ part, int nb,
                                                double fd_scale, double gamma, double *eps_io,
                                                S *__restrict__ P, const int *skip)
{
    if (*skip) {
        return;
    }
    // the first cell's U and v are loaded before the reduction of the partials (overlap)
    const int c0 = (blockIdx.x * RT) + threadIdx.x;
    double u0[5], v0[5];
    if (c0 < ni) {
        for (int k = 0; k < 5; k++) {
            u0[k] = U[((long)k * ni) + c0];
            v0[k] = v[((long)k * ni) + c0];
        }
    }
    double eps;
    if (part) {
        eps = fd_scale / sqrt(reduce_parts(part, nb));
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            *eps_io = eps;
        }
    }
    else {
        eps = *eps_io;
    }
    if (c0 < ni) {
        double u[5], p[5];
        for (int k = 0; k < 5; k++) u[k] = u0[k] + (eps * v0[k]);
        cons2prim<double>(p, u, gamma);
        for (int k = 0; k < 5; k++) P[((long)c0 * 5) + k] = (S)p[k];
    }
    for (int c = c0 + (gridDim.x * RT); c < ni; c += gridDim.x * RT) {
        double u[5], p[5];
        for (int k = 0; k < 5; k++) u[k] = U[((long)k * ni) + c] + (eps * v[((long)k * ni) + c]);
        cons2prim<double>(p, u, gamma);
        for (int k = 0; k < 5; k++) P[((long)c * 5) + k] = (S)p[k];
    }
}

// matvec, second half (gmres.c:41-46): w = v - step (R(P1) - R) / eps
__global__ __launch_bounds__(RT) void k_matvec_finish(const double *__restrict__ v,
                                                      const double *__restrict__ D1,
                                                      const double *__restrict__ D, double step,
                                                      const double *eps_ptr, double *__restrict__ w,
                                                      long n, const int *skip)
{
    if (*skip) {
        return;
    }
    const double eps = *eps_ptr;
    for (long e = (blockIdx.x * (long)RT) + threadIdx.x; e < n; e += (long)gridDim.x * RT) {
        w[e] = v[e] - ((step * (D1[e] - D[e])) / eps);
    }
}

// k_matvec_finish, then part[block] = partial dot(v0, w) as k_dot(v0, w) forms it: the matvec's
// second half and the first Gram-Schmidt dot in one pass (the same values and partials)
__global__ __launch_bounds__(RT) void k_matvec_dot(const double *__restrict__ v,
                                                   const double *__restrict__ D1,
                                                   const double *__restrict__ D, double step,
                                                   const double *eps_ptr, double *__restrict__ w,
                                                   const double *__restrict__ v0, long n,
                                                   double2 *__restrict__ part, const int *skip)
{
    if (*skip) {
        return;
    }
    const double eps = *eps_ptr;
    double2 acc = make_double2(0.0, 0.0);
    for (long e = (blockIdx.x * (long)RT) + threadIdx.x; e < n; e += (long)gridDim.x * RT) {
        double x = v[e] - ((step * (D1[e] - D[e])) / eps);
        w[e] = x;
        add_product(acc, v0[e], x);
    }
    double2 tot = block_reduce(acc);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = tot;
    }
}

// basis[0] = -r / norm (gmres.c:73-77), basis[iter+1] = w / h (gmres.c:121-125):
// out = (sign * x) / d
__global__ __launch_bounds__(RT) void k_scale(double *__restrict__ out, const double *__restrict__ x,
                                              double sign, double d, long n)
{
    for (long e = (blockIdx.x * (long)RT) + threadIdx.x; e < n; e += (long)gridDim.x * RT) {
        out[e] = (sign * x[e]) / d;
    }
}

// basis[iter+1] = w / h with h read on the device (the same expression as k_scale's with sign
// 1: the next basis vector is formed before the host has read h), and part[block] = partial
// dot(basis[iter+1], basis[iter+1]) as k_dot forms it (the matvec's eps)
__global__ __launch_bounds__(RT) void k_scale_dot(double *__restrict__ out, const double *__restrict__ x,
                                                  const double *__restrict__ d, long n,
                                                  double2 *__restrict__ part, const int *skip)
{
    if (*skip) {
        return;
    }
    const double h = *d;
    double2 acc = make_double2(0.0, 0.0);
    for (long e = (blockIdx.x * (long)RT) + threadIdx.x; e < n; e += (long)gridDim.x * RT) {
        double y = x[e] / h;
        out[e] = y;
        add_product(acc, y, y);
    }
    if (part) {  // uniform across the block
        double2 tot = block_reduce(acc);
        if (threadIdx.x == 0) {
            part[blockIdx.x] = tot;
        }
    }
}

// increment = sum over i = iter-1 .. 0 of coef[i] basis[i], in that order (gmres.c:128-140),
// with iter the solve's iteration count (device memory)
__global__ __launch_bounds__(RT) void k_increment(double *__restrict__ inc, const double *__restrict__ V,
                                                  const double *__restrict__ coef, const int *iters,
                                                  long n)
{
    const int iter = *iters;
    for (long e = (blockIdx.x * (long)RT) + threadIdx.x; e < n; e += (long)gridDim.x * RT) {
        double x = 0.0;
        for (int i = iter - 1; i >= 0; i--) {
            x = x + (coef[i] * V[((long)i * n) + e]);
        }
        inc[e] = x;
    }
}

// GMRES bookkeeping on the device, so that the host need not read each Hessenberg column
// before it queues the next iteration.  gv holds cosines [dim], sines [dim], the rotated
// right-hand side [dim + 1] and the rotated Hessenberg matrix [(dim + 1) x dim], row-major;
// state[0] is 0 while the solve runs and its iteration count once it has stopped; every
// kernel of a later iteration reads it and does nothing.
__global__ void k_gmres_init(double *gv, int dim, double norm, int *state)
{
    for (int i = threadIdx.x; i < (3 * dim) + 1; i += blockDim.x) {
        gv[i] = (i == 2 * dim) ? norm : 0.0;  // rhs[0] = |r| (gmres.c:67)
    }
    if (threadIdx.x == 0) {
        state[0] = 0;
    }
}

// Column j of the Hessenberg matrix, h[i * hs] for i = 0 .. j+1, through the previous
// rotations (cs, sn), the new rotation and the update of the right-hand side's entries j and
// j+1, r[0] and r[1] (gmres.c:102-117), with the reference's operation order.  True when the
// solve stops after this column: the rotated residual is below tol_norm (gmres.c:118-119) or
// the basis is full.  One thread.
__device__ inline bool givens_column(double *h, long hs, double *cs, double *sn, double *r, int j,
                                     int dim, double tol_norm)
{
    for (int i = 0; i < j; i++) {
        double tmp = (cs[i] * h[i * hs]) + (sn[i] * h[(i + 1) * hs]);
        h[(i + 1) * hs] = (-sn[i] * h[i * hs]) + (cs[i] * h[(i + 1) * hs]);
        h[i * hs] = tmp;
    }
    double rot = sqrt((h[j * hs] * h[j * hs]) + (h[(j + 1) * hs] * h[(j + 1) * hs]));
    cs[j] = h[j * hs] / rot;
    sn[j] = h[(j + 1) * hs] / rot;
    h[j * hs] = rot;
    r[1] = -sn[j] * r[0];
    r[0] = cs[j] * r[0];
    return fabs(r[1]) < tol_norm || j + 1 == dim;
}

// the solve stopped after `iters` iterations: state for the device, status for the host
__device__ inline void gmres_stop(int *state, int *status_host, int iters)
{
    state[0] = iters;
    __atomic_store_n(status_host, iters, __ATOMIC_RELAXED);  // host-coherent memory
}

// Column j of the Hessenberg matrix (scal[0 .. j+1], from the Gram-Schmidt kernels) rotated
// into gv and the convergence test.  One thread.
__global__ void k_givens(const double *__restrict__ scal, double *gv, int j, int dim,
                         double tol_norm, int *state, int *status_host)
{
    if (state[0]) {
        return;
    }
    double *H = gv + (3 * dim) + 1;
    for (int i = 0; i <= j + 1; i++) {
        H[((long)i * dim) + j] = scal[i];
    }
    if (givens_column(H + j, dim, gv, gv + dim, gv + (2 * dim) + j, j, dim, tol_norm)) {
        gmres_stop(state, status_host, j + 1);
    }
}

// ---------------------------------------------------------------- fused Arnoldi column
// The partial sums and barrier words that cross workgroups inside one launch go through
// coherent.h's device-scope atomics only.
__device__ inline void store_part(double2 *p, double2 v)
{
    st_shared_wg(&p->x, v.x);
    st_shared_wg(&p->y, v.y);
}

__device__ inline double2 load_part(const double2 *p)
{
    return make_double2(ld_shared_wg(&p->x), ld_shared_wg(&p->y));
}

__device__ inline unsigned sync_load(const unsigned *p)
{
    return ld_shared_wg(p);
}

// Grid-wide barrier of k_arnoldi, whose workgroups are all resident; lane 0 of each group of RT
// threads has stored its partial with store_part before the call.  Every wave first waits for
// its own stores (s_waitcnt covers only the wave that issues it), then thread 0 arrives for the
// workgroup.  `sync` holds, 256 bytes apart: eight arrival counters (workgroup b counts on
// counter b % 8), the counter of completed groups, the generation and a failure flag.  The
// counters only grow (modulo 2^32): after g barriers since the words were cleared, counter c
// holds g times its members and the group counter g times the number of groups, so the last
// arrival of a counter (the one that brings it to (g + 1) times its members) counts its group,
// and the last group advances the generation, which the waiting workgroups poll; no reset
// sits on that path.  Every launch that uses the words has the same grid.  (1, 2 and 4
// counters were slower than 8.)  A wait gives up after spin_limit polls (ARNOLDI_SPIN_LIMIT,
// about a second, unless a test lowers it) and raises the failure flag, which ends every other
// wait at once and makes later launches return at their start, and err_host (host-coherent
// memory), so that the kernel ends whatever happens; the host then reports the failure and
// clears the words.
constexpr int SYNC_STRIDE = 64;  // unsigned words: 256 bytes
constexpr int SYNC_GEN = 9 * SYNC_STRIDE;
constexpr int SYNC_FAILED = 10 * SYNC_STRIDE;
constexpr int SYNC_WORDS = 11 * SYNC_STRIDE;
constexpr unsigned ARNOLDI_SPIN_LIMIT = 1u << 20;

__device__ inline void grid_barrier(unsigned *sync, unsigned &gen, int *err_host, unsigned spin_limit)
{
    wait_memory();  // this wave's partial stored before the workgroup arrives
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned nb = gridDim.x;
        const unsigned c = blockIdx.x & 7;
        const unsigned members = (nb - c + 7) / 8;
        const unsigned groups = nb < 8 ? nb : 8;
        const unsigned next = gen + 1u;
        if (atomicAdd(&sync[c * SYNC_STRIDE], 1u) + 1u == next * members &&
            atomicAdd(&sync[8 * SYNC_STRIDE], 1u) + 1u == next * groups) {
            atomicAdd(&sync[SYNC_GEN], 1u);
        }
        else {
            unsigned spins = 0;
            while (sync_load(&sync[SYNC_GEN]) == gen && sync_load(&sync[SYNC_FAILED]) == 0) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins >= spin_limit) {
                    atomicExch(&sync[SYNC_FAILED], 1u);
                    __atomic_store_n(err_host, 1, __ATOMIC_RELAXED);
                    break;
                }
            }
        }
        gen += 1;
        wait_memory();
    }
    __syncthreads();
}

constexpr int ARNOLDI_MAX_DIM = 128;
// k_arnoldi's workgroups hold ARNOLDI_GROUPS groups of RT threads; group q of workgroup b
// does the work of block b * ARNOLDI_GROUPS + q of the RT-thread kernels (its elements, its
// partial sums), so that fewer workgroups meet at each grid barrier
constexpr int ARNOLDI_GROUPS = 4;
constexpr int ARNOLDI_THREADS = ARNOLDI_GROUPS * RT;

// block_reduce of each group of RT threads separately: the result of group q in its threads
__device__ inline double2 group_reduce(double2 v)
{
    constexpr int WAVES = RT / 64;
    __shared__ double2 red[ARNOLDI_GROUPS * WAVES];
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
    const int w0 = (threadIdx.x / RT) * WAVES;
    double2 out = red[w0];
    for (int w = 1; w < WAVES; w++) {
        out = combine(out, red[w0 + w]);
    }
    __syncthreads();
    return out;
}

// reduce_parts over the partials of every workgroup of the launch (stored before the last
// grid barrier), in group 0 (the same loads and combination order as reduce_parts in an
// RT-thread block), the total in every thread of the workgroup
__device__ inline double total_of_parts(const double2 *part, int nb)
{
    __shared__ double total;
    constexpr int PER = MAX_PARTS / RT;  // partials per thread: all loads in flight at once
    double2 acc = make_double2(0.0, 0.0);
    if (threadIdx.x < RT) {
        double2 p[PER];
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int b = threadIdx.x + (u * RT);
            if (b < nb) {
                p[u] = load_part(part + b);
            }
        }
#pragma unroll
        for (int u = 0; u < PER; u++) {
            if (threadIdx.x + (u * RT) < nb) {
                acc = combine(acc, p[u]);
            }
        }
    }
    double2 tot = group_reduce(acc);
    if (threadIdx.x == 0) {
        total = tot.x + tot.y;
    }
    __syncthreads();
    const double t = total;
    __syncthreads();  // total is rewritten by the next call
    return t;
}

struct ArnoldiArgs {
    const double *V;      // basis, (dim + 1) vectors of n
    const double *R;      // residual of the perturbed state
    const double *D;      // residual at the base point
    double step;
    const double *eps;    // the matvec's eps
    long n;
    int nb;               // blocks of the RT-thread kernels: parts_of(n)
    int j, dim;
    double2 *part0, *part1;
    double *gv;
    double tol_norm;
    int *state;
    int *status_host;  // [0]: iteration count of a stopped solve, [1]: a wait timed out
    unsigned *sync;    // SYNC_WORDS: grid barrier
    unsigned spin_limit;  // polls before a barrier wait gives up
};

// One GMRES iteration's column on one launch of ceil(nb / ARNOLDI_GROUPS) workgroups, all
// resident, with w in registers (EPT elements per thread, the elements a grid-stride loop
// of the RT-thread kernels gives the thread): the matvec's second half
// w = v_j - step (R - D) / eps (gmres.c:39-49), modified Gram-Schmidt (gmres.c:92-100),
// hess[j+1][j] = |w|, the rotations and the convergence test (gmres.c:102-119), and unless the
// solve stops, basis[j+1] = w / hess[j+1][j] (gmres.c:121-125) with the partials of its norm in
// part[j % 2] for the next matvec's eps.  Every dot product is formed and reduced exactly as
// k_matvec_dot, k_mgs, k_reduce_column and k_scale_dot form and reduce it (the same partials,
// element order and combination order), so the column has the same bits as that kernel
// sequence.  Every workgroup reduces every dot and rotates the column itself (same bits
// everywhere), so that it knows whether to write the next basis vector without another
// barrier; workgroup 0 writes the rotated column.
template <int EPT>
__global__ __launch_bounds__(ARNOLDI_THREADS) void k_arnoldi(ArnoldiArgs a)
{
    // uniform: no workgroup of this launch changes either before every workgroup has read it
    if (a.state[0] || sync_load(a.sync + SYNC_FAILED)) {
        return;
    }
    __shared__ double hcol[ARNOLDI_MAX_DIM + 1], cs[ARNOLDI_MAX_DIM], sn[ARNOLDI_MAX_DIM];
    __shared__ double rhs[2];  // entries j and j+1 of the right-hand side
    __shared__ int stop;
    unsigned gen = 0;
    int *err = a.status_host + 1;
    const int j = a.j, dim = a.dim;
    const long n = a.n;
    const int lane = threadIdx.x % RT;
    const int blk = (blockIdx.x * ARNOLDI_GROUPS) + (threadIdx.x / RT);  // the RT-thread block
    const bool writes_part = lane == 0 && blk < a.nb;
    const long stride = (long)a.nb * RT;
    const long e0 = (blk * (long)RT) + lane;
    for (int i = threadIdx.x; i < j; i += ARNOLDI_THREADS) {
        cs[i] = a.gv[i];
        sn[i] = a.gv[dim + i];
    }
    if (threadIdx.x == 0) {
        rhs[0] = a.gv[(2 * dim) + j];
        gen = sync_load(&a.sync[SYNC_GEN]);
    }
    double w[EPT], vc[EPT], vn[EPT];
    const double *vj = a.V + ((long)j * n);
    const double eps = *a.eps;
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < EPT; k++) {
        const long e = e0 + (k * stride);
        if (e < n) {
            vc[k] = a.V[e];
            double x = vj[e] - ((a.step * (a.R[e] - a.D[e])) / eps);
            w[k] = x;
            add_product(acc, vc[k], x);
        }
    }
    double2 tot = group_reduce(acc);
    if (writes_part) {
        store_part(a.part0 + blk, tot);
    }
    grid_barrier(a.sync, gen, err, a.spin_limit);
    for (int i = 0; i <= j; i++) {
        const double2 *pin = (i & 1) ? a.part1 : a.part0;
        double2 *pout = (i & 1) ? a.part0 : a.part1;
        const double *vnext = a.V + ((long)(i + 1) * n);
        if (i < j) {  // the next basis vector's loads overlap the reduction
#pragma unroll
            for (int k = 0; k < EPT; k++) {
                const long e = e0 + (k * stride);
                if (e < n) {
                    vn[k] = vnext[e];
                }
            }
        }
        const double hv = total_of_parts(pin, a.nb);
        if (threadIdx.x == 0) {
            hcol[i] = hv;
        }
        acc = make_double2(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < EPT; k++) {
            const long e = e0 + (k * stride);
            if (e < n) {
                double x = w[k] - (hv * vc[k]);
                w[k] = x;
                add_product(acc, (i < j) ? vn[k] : x, x);
                if (i < j) {
                    vc[k] = vn[k];
                }
            }
        }
        tot = group_reduce(acc);
        if (writes_part) {
            store_part(pout + blk, tot);
        }
        grid_barrier(a.sync, gen, err, a.spin_limit);
    }
    const double hn = sqrt(total_of_parts(((j + 1) & 1) ? a.part1 : a.part0, a.nb));
    if (threadIdx.x == 0) {
        hcol[j + 1] = hn;
        rhs[1] = 0.0;
        stop = givens_column(hcol, 1, cs, sn, rhs, j, dim, a.tol_norm);
        if (blockIdx.x == 0) {
            double *H = a.gv + (3 * dim) + 1;
            for (int i = 0; i <= j + 1; i++) {
                H[((long)i * dim) + j] = hcol[i];
            }
            a.gv[j] = cs[j];
            a.gv[dim + j] = sn[j];
            a.gv[(2 * dim) + j] = rhs[0];
            a.gv[(2 * dim) + j + 1] = rhs[1];
            if (stop) {
                gmres_stop(a.state, a.status_host, j + 1);
            }
        }
    }
    __syncthreads();
    if (stop) {
        return;
    }
    double *vout = const_cast<double *>(a.V) + ((long)(j + 1) * n);
    acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < EPT; k++) {
        const long e = e0 + (k * stride);
        if (e < n) {
            double y = w[k] / hn;
            vout[e] = y;
            add_product(acc, y, y);
        }
    }
    tot = group_reduce(acc);
    if (writes_part) {
        ((j & 1) ? a.part1 : a.part0)[blk] = tot;
    }
}

// back substitution (gmres.c:127-140) into coef[0 .. iter).  For iter <= BACK_MAX the
// triangle is staged in shared memory by the whole workgroup, lane k of wave 0 holds coef[k]
// and forms the products H[i][k] * coef[k] of row i at once, and the subtractions run in the
// reference's order k = i+1 .. iter-1 (with contraction off, c - H * coef rounds the product
// first, so the precomputed products give the same bits); larger solves take one thread.
constexpr int BACK_MAX = 64;

__global__ __launch_bounds__(RT) void k_back_subst(const double *__restrict__ gv, int dim,
                                                   const int *state, double *__restrict__ coef)
{
    const int iter = state[0];
    const double *rhs = gv + (2 * dim), *H = gv + (3 * dim) + 1;
    if (iter > BACK_MAX) {
        if (threadIdx.x == 0) {
            for (int i = iter - 1; i >= 0; i--) {
                double c = rhs[i];
                for (int k = i + 1; k < iter; k++) {
                    c -= H[((long)i * dim) + k] * coef[k];
                }
                coef[i] = c / H[((long)i * dim) + i];
            }
        }
        return;
    }
    __shared__ double tri[BACK_MAX * BACK_MAX];
    for (int e = threadIdx.x; e < iter * iter; e += RT) {
        const int i = e / iter, k = e % iter;
        tri[(i * BACK_MAX) + k] = (k >= i) ? H[((long)i * dim) + k] : 0.0;
    }
    __syncthreads();
    if (threadIdx.x >= 64) {
        return;
    }
    const int lane = threadIdx.x;
    double mine = 0.0;  // coef[lane] once known
    for (int i = iter - 1; i >= 0; i--) {
        const double p = (lane > i && lane < iter) ? tri[(i * BACK_MAX) + lane] * mine : 0.0;
        double c = rhs[i];
        for (int k = i + 1; k < iter; k++) {
            c -= __shfl(p, k, 64);
        }
        c /= tri[(i * BACK_MAX) + i];
        if (lane == i) {
            mine = c;
        }
    }
    if (lane < iter) {
        coef[lane] = mine;
    }
}

// Newton update (advance.c:167-172): U += increment (when given), P = cons2prim(U)
template <typename S>
__global__ __launch_bounds__(RT) void k_newton_update(double *__restrict__ U, const double *__restrict__ inc,
                                                      int ni, double gamma, S *__restrict__ P)
{
    for (int c = (blockIdx.x * RT) + threadIdx.x; c < ni; c += gridDim.x * RT) {
        double u[5], p[5];
        for (int k = 0; k < 5; k++) {
            long e = ((long)k * ni) + c;
            u[k] = inc ? U[e] + inc[e] : U[e];
            if (inc) {
                U[e] = u[k];
            }
        }
        cons2prim<double>(p, u, gamma);
        for (int k = 0; k < 5; k++) P[((long)c * 5) + k] = (S)p[k];
    }
}

// Newton residual: first r = -D step (advance.c:148-153), then r = U - U0 - D step (:176-181)
__global__ __launch_bounds__(RT) void k_newton_residual(double *__restrict__ r, const double *U,
                                                        const double *U0, const double *__restrict__ D,
                                                        double step, long n)
{
    for (long e = (blockIdx.x * (long)RT) + threadIdx.x; e < n; e += (long)gridDim.x * RT) {
        r[e] = U ? ((U[e] - U0[e]) - (D[e] * step)) : (-D[e] * step);
    }
}

inline int parts_of(long n)
{
    long b = (n + RT - 1) / RT;
    return (int)(b < MAX_PARTS ? b : MAX_PARTS);
}

inline int cells_grid(int ni)
{
    long b = ((long)ni + RT - 1) / RT;
    return (int)(b < 4096 ? b : 4096);
}

}  // namespace

// device and pinned host buffers of the Newton-GMRES solver
struct Krylov {
    int dim = 0;
    long n = 0;               // 5 * num_inner
    double *V = nullptr;      // (dim + 1) basis vectors
    double *w = nullptr;      // matvec result / Gram-Schmidt vector
    double *U = nullptr;      // conserved state
    double *D = nullptr;      // derivative at U (the nonlinear residual, FP64)
    double *Dm = nullptr;     // mixed: FP32-path derivative at U, the matvec's base point
    double *Rn = nullptr;     // Newton residual
    double *inc = nullptr;    // Newton increment
    double *coef = nullptr;   // back-substitution coefficients (dim)
    double2 *part[2] = {nullptr, nullptr};
    double *scal = nullptr;   // [0..dim]: Hessenberg column, [dim+1]: eps, [dim+2..]: norms
    double *h_scal = nullptr; // pinned host copy of scal (norms read by the Newton loop)
    double *gv = nullptr;     // GMRES rotations, right-hand side, rotated Hessenberg (k_givens)
    int *state = nullptr;     // [0]: 0 while a solve runs, its iteration count once stopped
    int *h_status = nullptr;  // pinned host-coherent: [0] the count of the last solve that
                              // stopped, [1] a fused column's grid barrier timed out
    int *h_status_dev = nullptr;  // h_status's device address
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};  // one per iteration in flight
    int *dev_of_ref = nullptr; // serial reductions: device row of each inner cell (reference order)
    unsigned *sync = nullptr;  // grid barrier of the fused column kernel (k_arnoldi)
    const void *arnoldi = nullptr;  // the k_arnoldi instance for n, or null: separate kernels
    int arnoldi_ept = 0;            // its elements per thread
    long newton = 0, krylov = 0;  // iterations of the last implicit step (Newton, GMRES)
};


void teal_krylov_release(TealGpu *g)
{
    if (!g || !g->krylov) {
        return;
    }
    Krylov *k = g->krylov;
    void *ptrs[] = {k->V,    k->w,       k->U,          k->D,  k->Dm,    k->Rn,
                    k->inc,  k->coef,    k->part[0],    k->part[1], k->scal, k->dev_of_ref,
                    k->gv,   k->state,   k->sync};
    for (void *p : ptrs) {
        if (p) {
            (void)hipFree(p);
        }
    }
    for (void *p : {(void *)k->h_scal, (void *)k->h_status}) {
        if (p) {
            (void)hipHostFree(p);
        }
    }
    for (hipEvent_t e : k->ev) {
        if (e) {
            (void)hipEventDestroy(e);
        }
    }
    delete k;
    g->krylov = nullptr;
}

inline int arnoldi_grid(int nb)
{
    return (nb + ARNOLDI_GROUPS - 1) / ARNOLDI_GROUPS;
}

// The fused column kernel for k->n, when one fits: the elements a thread owns fit its
// registers (at most 8: up to 5 * 2^18 cells), the dimension fits the kernel's shared arrays,
// and every workgroup can be resident at once (its grid barrier needs all of them running).
// It is launched as an ordinary kernel: on its stream the kernels before it have finished when
// it starts, so only other processes' work could hold compute units back, and then its
// bounded waits end it with an error.  Otherwise (or with TEAL_GPU_ARNOLDI=0 in the
// environment) each column takes separate kernels.
static int arnoldi_pick(Krylov *k)
{
    k->arnoldi = nullptr;
    k->arnoldi_ept = 0;
    const char *env = getenv("TEAL_GPU_ARNOLDI");
    if ((env && atoi(env) == 0) || k->dim > ARNOLDI_MAX_DIM) {
        return 0;
    }
    const int nb = parts_of(k->n);
    const long per_thread = (k->n + ((long)nb * RT) - 1) / ((long)nb * RT);
    const void *fns[] = {(const void *)k_arnoldi<1>, (const void *)k_arnoldi<2>,
                         (const void *)k_arnoldi<4>, (const void *)k_arnoldi<8>};
    constexpr int NFN = sizeof(fns) / sizeof(fns[0]);
    int which = 0;
    while (which < NFN && (1L << which) < per_thread) which++;
    if (which == NFN) {
        return 0;
    }
    const void *fn = fns[which];
    int dev = 0, cus = 0, per_cu = 0;
    HIPCALL(hipGetDevice(&dev));
    HIPCALL(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCALL(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, ARNOLDI_THREADS, 0));
    if ((long)per_cu * cus >= arnoldi_grid(nb)) {
        k->arnoldi = fn;
        k->arnoldi_ept = 1 << which;
    }
    return 0;
}

int teal_krylov_column_ept(const TealGpu *g)
{
    const Krylov *k = g->krylov;
    return (k && k->arnoldi && !g->serial_reductions) ? k->arnoldi_ept : 0;
}

static int krylov_reserve(TealGpu *g, int dim)
{
    long n = 5L * g->ni;
    if (g->krylov && g->krylov->dim >= dim && g->krylov->n == n) {
        return 0;
    }
    teal_krylov_release(g);
    Krylov *k = new Krylov();
    g->krylov = k;
    k->dim = dim;
    k->n = n;
    size_t vec = sizeof(double) * (size_t)n;
    HIPCALL(hipMalloc((void **)&k->V, vec * (size_t)(dim + 1)));
    HIPCALL(hipMalloc((void **)&k->w, vec));
    HIPCALL(hipMalloc((void **)&k->U, vec));
    HIPCALL(hipMalloc((void **)&k->D, vec));
    HIPCALL(hipMalloc((void **)&k->Dm, vec));
    HIPCALL(hipMalloc((void **)&k->Rn, vec));
    HIPCALL(hipMalloc((void **)&k->inc, vec));
    HIPCALL(hipMalloc((void **)&k->coef, sizeof(double) * (size_t)dim));
    HIPCALL(hipMalloc((void **)&k->part[0], sizeof(double2) * MAX_PARTS));
    HIPCALL(hipMalloc((void **)&k->part[1], sizeof(double2) * MAX_PARTS));
    HIPCALL(hipMalloc((void **)&k->scal, sizeof(double) * (size_t)(dim + 8)));
    HIPCALL(hipHostMalloc((void **)&k->h_scal, sizeof(double) * (size_t)(dim + 8), hipHostMallocCoherent));
    HIPCALL(hipMalloc((void **)&k->gv, sizeof(double) * (size_t)((3 * dim) + 1 + ((dim + 1) * dim))));
    HIPCALL(hipMalloc((void **)&k->state, sizeof(int) * 4));
    HIPCALL(hipMemset(k->state, 0, sizeof(int) * 4));
    HIPCALL(hipHostMalloc((void **)&k->h_status, sizeof(int) * 4, hipHostMallocCoherent));
    HIPCALL(hipHostGetDevicePointer((void **)&k->h_status_dev, k->h_status, 0));
    for (hipEvent_t &e : k->ev) {
        HIPCALL(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    HIPCALL(hipMalloc((void **)&k->dev_of_ref, sizeof(int) * (size_t)g->ni));
    HIPCALL(hipMemcpy(k->dev_of_ref, g->cell_inv.data(), sizeof(int) * (size_t)g->ni,
                      hipMemcpyHostToDevice));
    HIPCALL(hipMalloc((void **)&k->sync, sizeof(unsigned) * SYNC_WORDS));
    HIPCALL(hipMemset(k->sync, 0, sizeof(unsigned) * SYNC_WORDS));
    return arnoldi_pick(k);
}

namespace {

// norms of two vectors into h_scal[slot], h_scal[slot + 1] (one host synchronization)
// norm of a into scal[idx] (either reduction)
void norm_into(TealGpu *g, Krylov *k, const double *a, int pbuf, int idx)
{
    if (g->serial_reductions) {
        hipLaunchKernelGGL(k_serial_dot, dim3(1), dim3(RT), 0, g->stream, a, a,
                           (const int *)k->dev_of_ref, g->ni, k->scal, idx, 1, 0.0,
                           (const int *)nullptr);
        return;
    }
    const int nb = parts_of(k->n);
    hipLaunchKernelGGL(k_dot, dim3(nb), dim3(RT), 0, g->stream, a, a, k->n, k->part[pbuf],
                       (const int *)nullptr);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(RT), 0, g->stream, (const double2 *)k->part[pbuf], nb,
                       k->scal, idx, 1);
}

int two_norms(TealGpu *g, Krylov *k, const double *a, const double *b, int slot)
{
    norm_into(g, k, a, 0, slot);
    norm_into(g, k, b, 1, slot + 1);
    HIPCALL(hipGetLastError());
    HIPCALL(hipMemcpyAsync(k->h_scal + slot, k->scal + slot, 2 * sizeof(double), hipMemcpyDeviceToHost,
                           g->stream));
    if (teal_stream_wait(g)) return -1;
    return 0;
}

// derivative of the state in d.d_P into k->D (and the boundary ghosts of d.d_P)
template <typename S>
int derivative_into(TealGpu *g, DevData<S> &d, Krylov *k, double *out)
{
    if (teal_residual_pass(g, d, 0, 0.0, 0)) return -1;
    HIPCALL(hipMemcpyAsync(out, g->d_R, sizeof(double) * (size_t)k->n, hipMemcpyDeviceToDevice,
                           g->stream));
    return 0;
}

// gmres (src/simulation/gmres.c:52-149); the increment lands in k->inc.
//
// The host queues iteration after iteration without reading them: each iteration's
// Hessenberg column is rotated and tested on the device, and once a solve has stopped every
// Krylov kernel of a later iteration does nothing.  Before it queues the next basis vector's
// residual evaluation, the host waits for the column `ahead` iterations back and stops when
// that solve has stopped.  An evaluation queued past the stop writes nothing the caller reads
// (the perturbed state and R, which the Newton update and the next residual overwrite).  The
// exact Riemann solver queues nothing ahead: its vacuum flag is sticky, and an evaluation the
// reference does not make must not raise it.
//
// A column is one launch of k_arnoldi when k->arnoldi is set (tree reductions),
// else k_matvec_dot, one k_mgs per basis vector, k_reduce_column and k_givens (or the serial
// reductions' kernels): the same bits either way.
template <typename S>
int gmres(TealGpu *g, DevData<S> &d, Krylov *k, const double *Dbase, double step, double norm,
          double fd_scale, double tol, int dim)
{
    const long n = k->n;
    const int nb = parts_of(n);
    const int cg = cells_grid(g->ni);
    const double gamma = g->cfg.gamma;
    const int *state = k->state;
    const bool fused = k->arnoldi && !g->serial_reductions;
    // TEAL_GPU_ARNOLDI_SPIN lowers the barrier's poll limit, so that a test can drive the
    // time-out path (every wait still ends)
    const char *spin_env = getenv("TEAL_GPU_ARNOLDI_SPIN");
    const unsigned spin_limit = spin_env ? (unsigned)std::max(1, atoi(spin_env)) : ARNOLDI_SPIN_LIMIT;
    *(volatile int *)k->h_status = 0;
    if (norm <= 0) {
        HIPCALL(hipMemsetAsync(k->inc, 0, sizeof(double) * (size_t)n, g->stream));
        return 0;
    }
    hipLaunchKernelGGL(k_gmres_init, dim3(1), dim3(RT), 0, g->stream, k->gv, dim, norm, k->state);
    hipLaunchKernelGGL(k_scale, dim3(nb), dim3(RT), 0, g->stream, k->V, (const double *)k->Rn, -1.0,
                       norm, n);
    const double tol_norm = tol * norm;
    double *eps_dev = k->scal + dim + 1;  // eps of the current matvec
    // first half of the matvec (gmres.c:12-50) for basis vector j: eps = fd_scale / |v| into
    // scal[dim + 1], the primitive state of U + eps v into d.d_P and its residual into d_R
    // (tree reductions: the partials of |v|^2 are in `parts` already unless it is null)
    auto perturbed_residual = [&](int j, const double2 *parts) -> int {
        const double *v = k->V + ((size_t)j * n);
        if (g->serial_reductions) {
            hipLaunchKernelGGL(k_serial_dot, dim3(1), dim3(RT), 0, g->stream, v, v,
                               (const int *)k->dev_of_ref, g->ni, k->scal, dim + 1, 2, fd_scale, state);
            parts = nullptr;
        }
        else if (!parts) {
            hipLaunchKernelGGL(k_dot, dim3(nb), dim3(RT), 0, g->stream, v, v, n, k->part[0], state);
            parts = k->part[0];
        }
        hipLaunchKernelGGL((k_perturb<S>), dim3(cg), dim3(RT), 0, g->stream, (const double *)k->U,
                           v, g->ni, parts, nb, fd_scale, gamma, eps_dev, d.d_P, state);
        HIPCALL(hipGetLastError());
        return teal_residual_pass(g, d, 0, 0.0, 0);
    };
    // second half of the matvec, w = v - step (R(cons2prim(U + eps v)) - R(U)) / eps, then
    // modified Gram-Schmidt (gmres.c:92-100): hess[i][j] lands in scal[i]; then the column's
    // rotations and the convergence test
    auto column = [&](int j) -> int {
        const double *vj = k->V + ((size_t)j * n);
        if (fused) {
            ArnoldiArgs a{k->V,       g->d_R,     Dbase, step,     eps_dev,  n,
                          nb,         j,          dim,   k->part[0], k->part[1], k->gv,
                          tol_norm,   k->state,   k->h_status_dev, k->sync, spin_limit};
            void *args[] = {&a};
            HIPCALL(hipLaunchKernel(k->arnoldi, dim3(arnoldi_grid(nb)), dim3(ARNOLDI_THREADS), args,
                                    0, g->stream));
            return 0;
        }
        if (g->serial_reductions) {
            hipLaunchKernelGGL(k_matvec_finish, dim3(nb), dim3(RT), 0, g->stream, vj,
                               (const double *)g->d_R, Dbase, step, (const double *)eps_dev, k->w, n,
                               state);
            for (int i = 0; i <= j; i++) {
                const double *vi = k->V + ((size_t)i * n);
                hipLaunchKernelGGL(k_serial_dot, dim3(1), dim3(RT), 0, g->stream, vi,
                                   (const double *)k->w, (const int *)k->dev_of_ref, g->ni, k->scal, i,
                                   0, 0.0, state);
                hipLaunchKernelGGL(k_axpy_h, dim3(nb), dim3(RT), 0, g->stream, k->w, vi,
                                   (const double *)k->scal, i, n, state);
            }
            hipLaunchKernelGGL(k_serial_dot, dim3(1), dim3(RT), 0, g->stream, (const double *)k->w,
                               (const double *)k->w, (const int *)k->dev_of_ref, g->ni, k->scal,
                               j + 1, 1, 0.0, state);
        }
        else {
            hipLaunchKernelGGL(k_matvec_dot, dim3(nb), dim3(RT), 0, g->stream, vj,
                               (const double *)g->d_R, Dbase, step, (const double *)eps_dev, k->w,
                               (const double *)k->V, n, k->part[0], state);
            int cur = 0;
            for (int i = 0; i <= j; i++) {
                const double *next = (i < j) ? k->V + ((size_t)(i + 1) * n) : nullptr;
                hipLaunchKernelGGL(k_mgs, dim3(nb), dim3(RT), 0, g->stream, k->w,
                                   (const double *)(k->V + ((size_t)i * n)), next, n,
                                   (const double2 *)k->part[cur], nb, k->part[1 - cur], k->scal, i,
                                   state);
                cur = 1 - cur;
            }
            hipLaunchKernelGGL(k_reduce_column, dim3(1), dim3(RT), 0, g->stream,
                               (const double2 *)k->part[cur], nb, k->scal, j + 1, state);
        }
        hipLaunchKernelGGL(k_givens, dim3(1), dim3(1), 0, g->stream, (const double *)k->scal, k->gv,
                           j, dim, tol_norm, k->state, k->h_status_dev);
        HIPCALL(hipGetLastError());
        return 0;
    };
    // basis vector j + 1 = w / hess[j + 1][j] (the fused column has formed it, with the
    // partials of its norm in part[j % 2]) and its perturbed residual
    auto advance = [&](int j) -> int {
        if (fused) {
            return perturbed_residual(j + 1, k->part[j % 2]);
        }
        double2 *parts = g->serial_reductions ? nullptr : k->part[0];
        hipLaunchKernelGGL(k_scale_dot, dim3(nb), dim3(RT), 0, g->stream,
                           k->V + ((size_t)(j + 1) * n), (const double *)k->w,
                           (const double *)(k->scal + j + 1), n, parts, state);
        return perturbed_residual(j + 1, parts);
    };
    const int ahead = (g->cfg.flux == TEAL_FLUX_GODUNOV) ? 0 : 1;
    constexpr int NEV = sizeof(k->ev) / sizeof(k->ev[0]);
    static_assert(NEV > 1, "events of the iterations in flight");
    auto stopped = [&]() { return *(volatile int *)k->h_status != 0; };  // no wait
    if (perturbed_residual(0, nullptr)) return -1;
    for (int j = 0; j < dim; j++) {
        if (column(j)) return -1;
        HIPCALL(hipEventRecord(k->ev[j % NEV], g->stream));
        if (j >= ahead && teal_event_wait(k->ev[(j - ahead) % NEV])) return -1;
        if (((volatile int *)k->h_status)[1]) {
            break;  // a fused column's wait timed out: reported after the next synchronization
        }
        if (stopped()) {
            break;
        }
        if (j + 1 < dim && advance(j)) return -1;
    }
    // back substitution (gmres.c:127-140) and the increment, with the solve's iteration count
    hipLaunchKernelGGL(k_back_subst, dim3(1), dim3(RT), 0, g->stream, (const double *)k->gv, dim,
                       state, k->coef);
    hipLaunchKernelGGL(k_increment, dim3(nb), dim3(RT), 0, g->stream, k->inc, (const double *)k->V,
                       (const double *)k->coef, state, n);
    HIPCALL(hipGetLastError());
    return 0;
}

// implicit_euler (src/simulation/advance.c:119-195) on the resident state.  dm is the device
// data of the configured path; the matvec's residuals run on it.  The nonlinear (Newton)
// residual always runs in FP64: on the mixed path the FP32 residual's rounding noise would
// set a floor under the Newton residual, so a mixed TealGpu keeps an FP64 mirror for it and
// the matvec differences FP32 residuals about their own FP32 base point.
template <typename S>
int implicit_step(TealGpu *g, DevData<S> &dm, double *time, double *residual, double max_step,
                  double courant, double newton_tol, double krylov_tol, int dim, double *step0_out)
{
    const bool mixed = g->precision == TEAL_MIXED;
    if (krylov_reserve(g, dim) || teal_fp64_mirror(g)) return -1;
    DevData<double> &dn = g->d64;
    Krylov *k = g->krylov;
    const long n = k->n;
    const int nb = parts_of(n);
    const int cg = cells_grid(g->ni);
    const double gamma = g->cfg.gamma;
    double dt = 0;
    // dt minimum, and U0 = prim2cons(P) unless U0 already holds the state (mixed: its FP64 copy)
    if (teal_timestep_pass(g, dm, !g->u0_valid, &dt)) return -1;
    const double step0 = courant * dt;
    const double step = fmin(step0, max_step);
    HIPCALL(hipMemcpyAsync(k->U, g->d_U0, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice,
                           g->stream));
    // residuals of the state U: FP64 (Newton) and, on the mixed path, FP32 (matvec base)
    auto residuals = [&]() -> int {
        if (mixed) {
            hipLaunchKernelGGL((k_newton_update<double>), dim3(cg), dim3(RT), 0, g->stream, k->U,
                               (const double *)nullptr, g->ni, gamma, dn.d_P);
            HIPCALL(hipGetLastError());
            if (derivative_into(g, dm, k, k->Dm)) return -1;
        }
        return derivative_into(g, dn, k, k->D);
    };
    const double *Dbase = mixed ? k->Dm : k->D;
    *time += step;
    if (residuals()) return -1;
    hipLaunchKernelGGL(k_newton_residual, dim3(nb), dim3(RT), 0, g->stream, k->Rn, (const double *)nullptr,
                       (const double *)nullptr, (const double *)k->D, step, n);
    const int slot = dim + 2;
    // a vacuum in the exact Riemann solver ends the step at the next synchronization
    if (two_norms(g, k, k->Rn, k->U, slot) || teal_riemann_check(g)) return -1;
    double norm = k->h_scal[slot];
    double cons_norm = k->h_scal[slot + 1];
    // the finite-difference step: the reference's 1e-6 on the FP64 path; on the FP32 path
    // TEAL_MIXED_FD_RELATIVE (see its definition)
    const double fd_relative = mixed ? TEAL_MIXED_FD_RELATIVE : 1e-6;
    double fd_scale = fd_relative * fmax(1, cons_norm) / fmax(1, norm);
    const double tol_norm = newton_tol * norm;
    const long max_iter = 128L * (g->ni + 1);  // advance.c:162
    k->newton = 0;
    k->krylov = 0;
    for (long iter = 0; iter < max_iter && norm > tol_norm; iter++) {
        k->newton += 1;
        if (gmres(g, dm, k, Dbase, step, norm, fd_scale, krylov_tol, dim)) return -1;
        hipLaunchKernelGGL((k_newton_update<S>), dim3(cg), dim3(RT), 0, g->stream, k->U,
                           (const double *)k->inc, g->ni, gamma, dm.d_P);
        HIPCALL(hipGetLastError());
        if (residuals()) return -1;
        hipLaunchKernelGGL(k_newton_residual, dim3(nb), dim3(RT), 0, g->stream, k->Rn,
                           (const double *)k->U, (const double *)g->d_U0, (const double *)k->D, step, n);
        if (two_norms(g, k, k->Rn, k->U, slot) || teal_riemann_check(g)) return -1;
        k->krylov += *(volatile int *)k->h_status;  // the solve's iteration count
        if (((volatile int *)k->h_status)[1]) {
            ((volatile int *)k->h_status)[1] = 0;
            HIPCALL(hipMemsetAsync(k->sync, 0, sizeof(unsigned) * SYNC_WORDS, g->stream));
            teal_gpu_set_error("implicit_euler: a GMRES column's grid barrier timed out");
            return -1;
        }
        norm = k->h_scal[slot];
        cons_norm = k->h_scal[slot + 1];
        fd_scale = fd_relative * fmax(1, cons_norm) / fmax(1, norm);
    }
    if (g->serial_reductions) {
        const double *vol64 = g->d64.d_vol64;
        hipLaunchKernelGGL(k_serial_residual, dim3(1), dim3(RT), 0, g->stream, (const double *)k->D,
                           vol64, (const int *)k->dev_of_ref, g->ni, g->sum_volume, k->scal);
        HIPCALL(hipGetLastError());
        HIPCALL(hipMemcpyAsync(k->h_scal, k->scal, 5 * sizeof(double), hipMemcpyDeviceToHost, g->stream));
        if (teal_stream_wait(g)) return -1;
        memcpy(residual, k->h_scal, 5 * sizeof(double));
    }
    else if (teal_residual_norm(g, k->D, residual)) {
        return -1;
    }
    if (mixed) {  // carry the FP64 solution into the next step
        HIPCALL(hipMemcpyAsync(g->d_U0, k->U, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice,
                               g->stream));
        g->u0_valid = 1;
    }
    else {  // the reference re-derives U from P at each step start
        g->u0_valid = 0;
    }
    g->dt_valid = 0;
    *step0_out = step0;
    return 0;
}

}  // namespace

extern "C" int teal_gpu_krylov_stats(TealGpu *g, long *stats)
{
    if (!g || !stats) {
        teal_gpu_set_error("krylov_stats: null argument");
        return -1;
    }
    stats[0] = g->krylov ? g->krylov->newton : 0;
    stats[1] = g->krylov ? g->krylov->krylov : 0;
    return 0;
}

extern "C" int teal_gpu_set_reductions(TealGpu *g, int mode)
{
    if (!g || (mode != TEAL_REDUCE_TREE && mode != TEAL_REDUCE_SERIAL)) {
        teal_gpu_set_error("set_reductions: invalid handle or mode (%d)", mode);
        return -1;
    }
    g->serial_reductions = (mode == TEAL_REDUCE_SERIAL);
    return 0;
}

extern "C" double teal_gpu_implicit_euler(TealGpu *g, double *time, double *residual,
                                          double max_step, double courant, double newton_tol,
                                          double krylov_tol, int krylov_dim)
{
    if (!g || !g->configured) {
        teal_gpu_set_error("implicit euler: TealGpu is not configured");
        return NAN;
    }
    if (g->ninst > 1) {
        teal_gpu_set_error("implicit euler: not available on a batch handle");
        return NAN;
    }
    if (!(newton_tol > 0) || !(krylov_tol > 0) || krylov_dim <= 0) {  // src/simulation/set.c:117-125
        teal_gpu_set_error("implicit euler: invalid newton tolerance (%g), krylov tolerance (%g) or "
                           "krylov dimension (%d)", newton_tol, krylov_tol, krylov_dim);
        return NAN;
    }
    double step0 = NAN;
    if (teal_riemann_clear(g)) return NAN;
    int rc = (g->precision == TEAL_FP64)
                 ? implicit_step(g, g->d64, time, residual, max_step, courant, newton_tol, krylov_tol,
                                 krylov_dim, &step0)
                 : implicit_step(g, g->d32, time, residual, max_step, courant, newton_tol, krylov_tol,
                                 krylov_dim, &step0);
    return (rc || teal_riemann_check(g)) ? NAN : step0;
}
