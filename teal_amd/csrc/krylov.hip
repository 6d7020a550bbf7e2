// krylov.hip -- implicit Euler with Jacobian-free Newton-GMRES on the device
// (src/simulation/advance.c:119-195, src/simulation/gmres.c:12-149).
//
// Vectors (U, U0, derivative, Newton residual, increment, the Krylov basis) are FP64
// [5][num_inner] arrays in device cell order on both precisions; the matvec's residual
// evaluations run on the configured path (FP64 or the mixed FP32 flux path).  The Newton and
// GMRES control flow, the Hessenberg matrix, the Givens rotations and the back substitution
// run on the host in the reference's operation order; every vector operation runs on the
// device with the reference's per-element expression.  The reference's dot products are
// serial sums.  By default they are deterministic two-level reductions with compensated
// (TwoSum / TwoProduct) accumulation, which round differently from the reference by at most
// its own summation error; the finite-difference matvec amplifies such last-bit differences
// (a perturbed state that rounds differently changes R(P1) - R by an ulp, divided by eps).
// TEAL_REDUCE_SERIAL sums in the reference's order on one workgroup instead: bit-for-bit with
// the reference, for verification on small meshes.  Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

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
                                            long n, double2 *__restrict__ part)
{
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

// out[idx] = sqrt(total), the last entry of a Hessenberg column, and host[0..idx] = out[0..idx]:
// the column written to host-visible memory (k_reduce with take_sqrt and the column's copy)
__global__ __launch_bounds__(RT) void k_reduce_column(const double2 *__restrict__ part, int nb,
                                                      double *out, int idx, double *host)
{
    double t = reduce_parts(part, nb);
    if (threadIdx.x == 0) {
        double r = sqrt(t);
        out[idx] = r;
        host[idx] = r;
    }
    for (int i = threadIdx.x; i < idx; i += RT) {
        host[i] = out[i];
    }
}

// One modified Gram-Schmidt step (gmres.c:92-99): h = the dot whose partials are in part_in
// (hess[i-1][iter]); w -= h * v_prev; partials of dot(v_next, w) (v_next == w: the norm).
__global__ __launch_bounds__(RT) void k_mgs(double *__restrict__ w, const double *__restrict__ v_prev,
                                            const double *v_next, long n,
                                            const double2 *__restrict__ part_in, int nb,
                                            double2 *__restrict__ part_out, double *h, int hidx)
{
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
                                                   double *out, int idx, int mode, double fd_scale)
{
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
                                               const double *h, int hidx, long n)
{
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
                                                S *__restrict__ P)
{
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
                                                      long n)
{
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
                                                   double2 *__restrict__ part)
{
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
                                                  double2 *__restrict__ part)
{
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

// increment = sum over i = iter-1 .. 0 of coef[i] basis[i], in that order (gmres.c:128-140)
__global__ __launch_bounds__(RT) void k_increment(double *__restrict__ inc, const double *__restrict__ V,
                                                  const double *__restrict__ coef, int iter, long n)
{
    for (long e = (blockIdx.x * (long)RT) + threadIdx.x; e < n; e += (long)gridDim.x * RT) {
        double x = 0.0;
        for (int i = iter - 1; i >= 0; i--) {
            x = x + (coef[i] * V[((long)i * n) + e]);
        }
        inc[e] = x;
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
    double *h_scal = nullptr; // pinned copy of scal (host-coherent: kernels write columns into it)
    double *h_scal_dev = nullptr;  // h_scal's device address
    int *dev_of_ref = nullptr; // serial reductions: device row of each inner cell (reference order)
    long newton = 0, krylov = 0;  // iterations of the last implicit step (Newton, GMRES)
};


void teal_krylov_release(TealGpu *g)
{
    if (!g || !g->krylov) {
        return;
    }
    Krylov *k = g->krylov;
    void *ptrs[] = {k->V, k->w, k->U, k->D, k->Dm, k->Rn, k->inc, k->coef, k->part[0], k->part[1],
                    k->scal, k->dev_of_ref};
    for (void *p : ptrs) {
        if (p) {
            (void)hipFree(p);
        }
    }
    if (k->h_scal) {
        (void)hipHostFree(k->h_scal);
    }
    delete k;
    g->krylov = nullptr;
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
    HIPCALL(hipHostGetDevicePointer((void **)&k->h_scal_dev, k->h_scal, 0));
    HIPCALL(hipMalloc((void **)&k->dev_of_ref, sizeof(int) * (size_t)g->ni));
    HIPCALL(hipMemcpy(k->dev_of_ref, g->cell_inv.data(), sizeof(int) * (size_t)g->ni,
                      hipMemcpyHostToDevice));
    return 0;
}

namespace {

// norms of two vectors into h_scal[slot], h_scal[slot + 1] (one host synchronization)
// norm of a into scal[idx] (either reduction)
void norm_into(TealGpu *g, Krylov *k, const double *a, int pbuf, int idx)
{
    if (g->serial_reductions) {
        hipLaunchKernelGGL(k_serial_dot, dim3(1), dim3(RT), 0, g->stream, a, a,
                           (const int *)k->dev_of_ref, g->ni, k->scal, idx, 1, 0.0);
        return;
    }
    const int nb = parts_of(k->n);
    hipLaunchKernelGGL(k_dot, dim3(nb), dim3(RT), 0, g->stream, a, a, k->n, k->part[pbuf]);
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

// gmres (src/simulation/gmres.c:52-149); the increment lands in k->inc
template <typename S>
int gmres(TealGpu *g, DevData<S> &d, Krylov *k, const double *Dbase, double step, double norm,
          double fd_scale, double tol, int dim)
{
    const long n = k->n;
    const int nb = parts_of(n);
    const int cg = cells_grid(g->ni);
    const double gamma = g->cfg.gamma;
    if (norm <= 0) {
        HIPCALL(hipMemsetAsync(k->inc, 0, sizeof(double) * (size_t)n, g->stream));
        return 0;
    }
    std::vector<double> rhs(dim + 1, 0.0), cosine(dim, 0.0), sine(dim, 0.0), coef(dim, 0.0);
    std::vector<double> hess((size_t)(dim + 1) * dim, 0.0);
    auto H = [&](int i, int j) -> double & { return hess[((size_t)i * dim) + j]; };
    rhs[0] = norm;
    double *eps_dev = k->scal + dim + 1;  // eps of the current matvec
    hipLaunchKernelGGL(k_scale, dim3(nb), dim3(RT), 0, g->stream, k->V, (const double *)k->Rn, -1.0,
                       norm, n);
    const double tol_norm = tol * norm;
    // first half of the matvec (gmres.c:12-50) for basis vector j: eps = fd_scale / |v| into
    // scal[dim + 1], the primitive state of U + eps v into d.d_P and its residual into d_R
    // (tree reductions: the partials of |v|^2 are in part[0] already when `have_parts`)
    auto perturbed_residual = [&](int j, bool have_parts) -> int {
        const double *v = k->V + ((size_t)j * n);
        const double2 *eps_parts = nullptr;
        if (g->serial_reductions) {
            hipLaunchKernelGGL(k_serial_dot, dim3(1), dim3(RT), 0, g->stream, v, v,
                               (const int *)k->dev_of_ref, g->ni, k->scal, dim + 1, 2, fd_scale);
        }
        else {
            if (!have_parts) {
                hipLaunchKernelGGL(k_dot, dim3(nb), dim3(RT), 0, g->stream, v, v, n, k->part[0]);
            }
            eps_parts = k->part[0];
        }
        hipLaunchKernelGGL((k_perturb<S>), dim3(cg), dim3(RT), 0, g->stream, (const double *)k->U,
                           v, g->ni, eps_parts, nb, fd_scale, gamma, eps_dev, d.d_P);
        HIPCALL(hipGetLastError());
        return teal_residual_pass(g, d, 0, 0.0, 0);
    };
    // second half of the matvec, w = v - step (R(cons2prim(U + eps v)) - R(U)) / eps, then
    // modified Gram-Schmidt (gmres.c:92-100): hess[i][j] lands in scal[i], and h_scal
    // receives the column
    auto column = [&](int j) -> int {
        const double *vj = k->V + ((size_t)j * n);
        if (g->serial_reductions) {
            hipLaunchKernelGGL(k_matvec_finish, dim3(nb), dim3(RT), 0, g->stream, vj,
                               (const double *)g->d_R, Dbase, step, (const double *)eps_dev, k->w, n);
            for (int i = 0; i <= j; i++) {
                const double *vi = k->V + ((size_t)i * n);
                hipLaunchKernelGGL(k_serial_dot, dim3(1), dim3(RT), 0, g->stream, vi,
                                   (const double *)k->w, (const int *)k->dev_of_ref, g->ni, k->scal, i,
                                   0, 0.0);
                hipLaunchKernelGGL(k_axpy_h, dim3(nb), dim3(RT), 0, g->stream, k->w, vi,
                                   (const double *)k->scal, i, n);
            }
            hipLaunchKernelGGL(k_serial_dot, dim3(1), dim3(RT), 0, g->stream, (const double *)k->w,
                               (const double *)k->w, (const int *)k->dev_of_ref, g->ni, k->scal,
                               j + 1, 1, 0.0);
            HIPCALL(hipGetLastError());
            HIPCALL(hipMemcpyAsync(k->h_scal, k->scal, sizeof(double) * (size_t)(j + 2),
                                   hipMemcpyDeviceToHost, g->stream));
        }
        else {
            hipLaunchKernelGGL(k_matvec_dot, dim3(nb), dim3(RT), 0, g->stream, vj,
                               (const double *)g->d_R, Dbase, step, (const double *)eps_dev, k->w,
                               (const double *)k->V, n, k->part[0]);
            int cur = 0;
            for (int i = 0; i <= j; i++) {
                const double *next = (i < j) ? k->V + ((size_t)(i + 1) * n) : nullptr;
                hipLaunchKernelGGL(k_mgs, dim3(nb), dim3(RT), 0, g->stream, k->w,
                                   (const double *)(k->V + ((size_t)i * n)), next, n,
                                   (const double2 *)k->part[cur], nb, k->part[1 - cur], k->scal, i);
                cur = 1 - cur;
            }
            hipLaunchKernelGGL(k_reduce_column, dim3(1), dim3(RT), 0, g->stream,
                               (const double2 *)k->part[cur], nb, k->scal, j + 1, k->h_scal_dev);
            HIPCALL(hipGetLastError());
        }
        return 0;
    };
    // basis vector j + 1 = w / hess[j + 1][j], read on the device, and its perturbed residual
    auto advance = [&](int j) -> int {
        double2 *parts = g->serial_reductions ? nullptr : k->part[0];
        hipLaunchKernelGGL(k_scale_dot, dim3(nb), dim3(RT), 0, g->stream,
                           k->V + ((size_t)(j + 1) * n), (const double *)k->w,
                           (const double *)(k->scal + j + 1), n, parts);
        return perturbed_residual(j + 1, parts != nullptr);
    };
    // The next basis vector and its perturbed residual are queued before the host waits for
    // the Hessenberg column, so the device works while the host reads it and applies the
    // rotations.  When the iteration converges that work is discarded: it writes only the next
    // basis vector, eps, the perturbed state and d_R, which the caller overwrites (the Newton
    // update rewrites d.d_P, the next residual d_R).  The exact Riemann solver is left out: its
    // vacuum flag is sticky, and an evaluation the reference does not make must not raise it.
    const bool ahead = g->cfg.flux != TEAL_FLUX_GODUNOV;
    if (perturbed_residual(0, false)) return -1;
    int iter = 0;
    for (; iter < dim; iter++) {
        const bool next = iter + 1 < dim;
        if (column(iter)) return -1;
        HIPCALL(hipEventRecord(g->ev_wait, g->stream));
        if (next && ahead && advance(iter)) return -1;
        if (teal_event_wait(g->ev_wait)) return -1;
        for (int i = 0; i <= iter + 1; i++) {
            H(i, iter) = k->h_scal[i];
        }
        // Givens rotations (gmres.c:102-119)
        for (int i = 0; i < iter; i++) {
            double tmp = (cosine[i] * H(i, iter)) + (sine[i] * H(i + 1, iter));
            H(i + 1, iter) = (-sine[i] * H(i, iter)) + (cosine[i] * H(i + 1, iter));
            H(i, iter) = tmp;
        }
        double rot = sqrt((H(iter, iter) * H(iter, iter)) + (H(iter + 1, iter) * H(iter + 1, iter)));
        cosine[iter] = H(iter, iter) / rot;
        sine[iter] = H(iter + 1, iter) / rot;
        H(iter, iter) = rot;
        rhs[iter + 1] = -sine[iter] * rhs[iter];
        rhs[iter] = cosine[iter] * rhs[iter];
        k->krylov += 1;
        if (fabs(rhs[iter + 1]) < tol_norm) {
            iter += 1;
            break;
        }
        if (next && !ahead) {
            hipLaunchKernelGGL(k_scale, dim3(nb), dim3(RT), 0, g->stream,
                               k->V + ((size_t)(iter + 1) * n), (const double *)k->w, 1.0,
                               H(iter + 1, iter), n);
            if (perturbed_residual(iter + 1, false)) return -1;
        }
    }
    // back substitution (gmres.c:127-140)
    for (int i = iter - 1; i >= 0; i--) {
        coef[i] = rhs[i];
        for (int j = i + 1; j < iter; j++) {
            coef[i] -= H(i, j) * coef[j];
        }
        coef[i] /= H(i, i);
    }
    HIPCALL(hipMemcpyAsync(k->coef, coef.data(), sizeof(double) * (size_t)dim, hipMemcpyHostToDevice,
                           g->stream));
    hipLaunchKernelGGL(k_increment, dim3(nb), dim3(RT), 0, g->stream, k->inc, (const double *)k->V,
                       (const double *)k->coef, iter, n);
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
