// diagnostics.hip -- device reductions over cells and boundary faces: the entity average
// (src/equations/average.c:8-41), which is also the Prepare step of the bulk-velocity forcing
// source (run/periodic_hill/run.c:8-21), the error norm against a reference solution
// (src/equations/norm.c:8-38) and the surface force coefficients of euler_polar
// (src/euler/polar.c:22-57).
//
// Cell sums run over an entity's cells in reference order (an inner entity is a contiguous
// range of reference rows), reading the resident device state through the device row of each
// reference row, or a host array uploaded in the reference layout.  Two modes, as for the
// Krylov dot products: deterministic compensated trees (default), or one running sum per
// variable in the reference's order (TEAL_REDUCE_SERIAL), bit for bit with the reference.
// The face sums of the polar always run serially in face order (a boundary entity's faces
// are few).  Compiled with -ffp-contract=off.
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

namespace {

constexpr int RT = SUM_THREADS;

enum { SUM_AVERAGE = 0, SUM_NORM = 1 };

// the value cell (reference row j, device row d) adds to variable k: volume * P[k]
// (average.c:26-31) or volume * (ref[k] - P[k])^2 (norm.c:26-31)
template <typename S, int KIND>
__device__ inline void cell_terms(const S *__restrict__ P, int host_rows,
                                  const double *__restrict__ ref, const int *__restrict__ cell_inv,
                                  const double *__restrict__ vol, int j, double *v, double *x)
{
    const long d = cell_inv[j];
    const long row = host_rows ? j : d;
    *v = vol[d];
    for (int k = 0; k < 5; k++) {
        const double p = (double)P[(row * 5) + k];
        if (KIND == SUM_AVERAGE) {
            x[k] = p;
        }
        else {
            const double e = ref[((long)j * 5) + k] - p;
            x[k] = e * e;
        }
    }
}

// part[k * gridDim.x + block] = this block's compensated partial sum of variable k over the
// reference rows [lo, lo + n)
template <typename S, int KIND>
__global__ __launch_bounds__(RT) void k_cell_sums(const S *__restrict__ P, int host_rows,
                                                  const double *__restrict__ ref,
                                                  const int *__restrict__ cell_inv,
                                                  const double *__restrict__ vol, int lo, int n,
                                                  double2 *__restrict__ part)
{
    double2 acc[5];
    for (int k = 0; k < 5; k++) acc[k] = make_double2(0.0, 0.0);
    for (int i = (blockIdx.x * RT) + threadIdx.x; i < n; i += gridDim.x * RT) {
        double v, x[5];
        cell_terms<S, KIND>(P, host_rows, ref, cell_inv, vol, lo + i, &v, x);
        for (int k = 0; k < 5; k++) add_product(acc[k], v, x[k]);
    }
    for (int k = 0; k < 5; k++) {
        const double2 t = block_reduce(acc[k]);
        if (threadIdx.x == 0) {
            part[((long)k * gridDim.x) + blockIdx.x] = t;
        }
    }
}

// out[k] = total / sum_volume (average) or sqrt(total / sum_volume) (norm); one workgroup
template <int KIND>
__global__ __launch_bounds__(RT) void k_cell_final(const double2 *__restrict__ part, int nb,
                                                   double sum_volume, double *out)
{
    for (int k = 0; k < 5; k++) {
        const double t = reduce_parts(part + ((long)k * nb), nb);
        if (threadIdx.x == 0) {
            out[k] = (KIND == SUM_AVERAGE) ? t / sum_volume : sqrt(t / sum_volume);
        }
    }
}

// The same sums in the reference's order: one workgroup; the threads stage 256 cells' terms
// in LDS, thread k < 5 adds variable k's in cell order (sum = sum + V x, as the reference).
template <typename S, int KIND>
__global__ __launch_bounds__(RT) void k_cell_serial(const S *__restrict__ P, int host_rows,
                                                    const double *__restrict__ ref,
                                                    const int *__restrict__ cell_inv,
                                                    const double *__restrict__ vol, int lo, int n,
                                                    double sum_volume, double *out)
{
    __shared__ double term[5 * RT];
    double sum = 0.0;
    for (int base = 0; base < n; base += RT) {
        const int i = base + threadIdx.x;
        if (i < n) {
            double v, x[5];
            cell_terms<S, KIND>(P, host_rows, ref, cell_inv, vol, lo + i, &v, x);
            for (int k = 0; k < 5; k++) term[(k * RT) + threadIdx.x] = v * x[k];
        }
        __syncthreads();
        if (threadIdx.x < 5) {
            const int cnt = (n - base < RT) ? n - base : RT;
            for (int e = 0; e < cnt; e++) sum = sum + term[(threadIdx.x * RT) + e];
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) {
        out[threadIdx.x] = (KIND == SUM_AVERAGE) ? sum / sum_volume : sqrt(sum / sum_volume);
    }
}

// euler_polar's face loop (polar.c:40-53) over `num` faces of one boundary entity: face j has
// device rows lr[j] (left inner cell, right ghost), area and normal (x, y) in geo[3 j ..];
// pc[j] = (p - p_ref) / q and coef = the force sums in face order.  One workgroup.
template <typename S>
__global__ __launch_bounds__(RT) void k_polar(const S *__restrict__ P, const int2 *__restrict__ lr,
                                              const double *__restrict__ geo, int num,
                                              double p_ref, double dyn, double length,
                                              double *__restrict__ pc, double *__restrict__ coef)
{
    __shared__ double term[2 * RT];
    double cx = 0.0, cy = 0.0;
    for (int base = 0; base < num; base += RT) {
        const int j = base + threadIdx.x;
        if (j < num) {
            const double pressure = ((double)P[((long)lr[j].x * 5) + 4] +
                                     (double)P[((long)lr[j].y * 5) + 4]) / 2;
            const double area = geo[3 * j], nx = geo[(3 * j) + 1], ny = geo[(3 * j) + 2];
            pc[j] = (pressure - p_ref) / dyn;
            term[threadIdx.x] = (pressure * area * nx) / (dyn * length);
            term[RT + threadIdx.x] = (pressure * area * ny) / (dyn * length);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int cnt = (num - base < RT) ? num - base : RT;
            for (int e = 0; e < cnt; e++) {
                cx = cx + term[e];
                cy = cy + term[RT + e];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        coef[0] = cx;
        coef[1] = cy;
    }
}

constexpr int CELL_PARTS = 1024;  // most blocks of a tree reduction

int ensure_buffers(TealGpu *g)
{
    if (!g->d_cell_inv) {
        HIPCALL(hipMalloc((void **)&g->d_cell_inv, sizeof(int) * (size_t)(g->mesh_ni + 1)));
        HIPCALL(hipMemcpy(g->d_cell_inv, g->cell_inv.data(), sizeof(int) * (size_t)g->mesh_ni,
                          hipMemcpyHostToDevice));
    }
    if (!g->d_cellpart) {
        HIPCALL(hipMalloc((void **)&g->d_cellpart, sizeof(double2) * 5 * CELL_PARTS));
    }
    if (!g->d_diag) {
        HIPCALL(hipMalloc((void **)&g->d_diag, sizeof(double) * 8));
        HIPCALL(hipMemset(g->d_diag, 0, sizeof(double) * 8));
    }
    return 0;
}

const double *vol64(const TealGpu *g)
{
    return (g->precision == TEAL_FP64) ? g->d64.d_vol64 : g->d32.d_vol64;
}

// launch the cell sums of reference rows [lo, hi) into out[0..4] (device memory)
template <typename S, int KIND>
int cell_sums(TealGpu *g, const S *P, int host_rows, const double *ref, int lo, int hi, double *out)
{
    const int n = hi - lo;
    if (g->serial_reductions) {
        hipLaunchKernelGGL((k_cell_serial<S, KIND>), dim3(1), dim3(RT), 0, g->stream, P, host_rows,
                           ref, (const int *)g->d_cell_inv, vol64(g), lo, n, g->sum_volume, out);
    }
    else {
        int nb = (n + RT - 1) / RT;
        nb = (nb < 1) ? 1 : (nb > CELL_PARTS ? CELL_PARTS : nb);
        hipLaunchKernelGGL((k_cell_sums<S, KIND>), dim3(nb), dim3(RT), 0, g->stream, P, host_rows,
                           ref, (const int *)g->d_cell_inv, vol64(g), lo, n, g->d_cellpart);
        hipLaunchKernelGGL((k_cell_final<KIND>), dim3(1), dim3(RT), 0, g->stream,
                           (const double2 *)g->d_cellpart, nb, g->sum_volume, out);
    }
    HIPCALL(hipGetLastError());
    return 0;
}

int check_single(TealGpu *g, const char *who)
{
    if (!g || !g->configured) {
        teal_gpu_set_error("%s: TealGpu is not configured", who);
        return -1;
    }
    if (g->ninst > 1) {
        teal_gpu_set_error("%s: not available on a batch handle", who);
        return -1;
    }
    return 0;
}

int inner_entity(TealGpu *g, int entity, const char *who)
{
    if (entity < 0 || entity >= g->mesh->entities.num_inner) {
        teal_gpu_set_error("%s: invalid inner entity (%d)", who, entity);
        return -1;
    }
    return 0;
}

template <typename S>
int average_impl(TealGpu *g, DevData<S> &d, int entity, const double *primitive, double *avg)
{
    const int lo = g->mesh->entities.cell_off[entity];
    const int hi = g->mesh->entities.cell_off[entity + 1];
    if (primitive) {  // a host array in the reference layout (inner rows are all it needs)
        HIPCALL(hipMemcpyAsync(g->d_stage, primitive, sizeof(double) * 5 * (size_t)g->mesh_ni,
                               hipMemcpyHostToDevice, g->stream));
        if (cell_sums<double, SUM_AVERAGE>(g, g->d_stage, 1, nullptr, lo, hi, g->d_diag)) return -1;
    }
    else if (cell_sums<S, SUM_AVERAGE>(g, d.d_P, 0, nullptr, lo, hi, g->d_diag)) {
        return -1;
    }
    HIPCALL(hipMemcpyAsync(avg, g->d_diag, sizeof(double) * 5, hipMemcpyDeviceToHost, g->stream));
    return teal_stream_wait(g);
}

template <typename S>
int polar_impl(TealGpu *g, DevData<S> &d, int entity, double *primitive, const double *reference,
               double length, double *coef, double *pc)
{
    const Mesh *m = g->mesh;
    const int f0 = m->entities.face_off[entity];
    const int num = m->entities.face_off[entity + 1] - f0;
    if (primitive && teal_upload_state(g, primitive)) return -1;
    const int nb = g->nfb - g->nfi;
    if (nb > 0) {  // equations_boundary first (polar.c:36)
        Kern<S>::boundary(g->stream, nb, g->nf, d, g->cfg.gamma);
        HIPCALL(hipGetLastError());
    }
    if (primitive && teal_download_ghosts(g, primitive)) return -1;
    // per face: device rows of its cells, area and the normal's x and y (polar.c:43-46)
    std::vector<int2> lr(num > 0 ? num : 1);
    std::vector<double> geo(3 * (size_t)(num > 0 ? num : 1));
    for (int j = 0; j < num; j++) {
        const int f = f0 + j;
        int l = m->faces.cell_idx[f].left, r = m->faces.cell_idx[f].right;
        if (r >= m->cells.off_boundary && g->periodic_src[r] >= 0) {
            r = g->periodic_src[r];  // a periodic ghost holds its source cell's state
        }
        lr[j] = make_int2(g->cell_inv[l], g->cell_inv[r]);
        geo[3 * j] = m->faces.area[f];
        geo[(3 * j) + 1] = m->faces.basis[f].x.x;
        geo[(3 * j) + 2] = m->faces.basis[f].x.y;
    }
    const double dyn = reference[0] *
                       (((reference[1] * reference[1]) + (reference[2] * reference[2])) +
                        (reference[3] * reference[3])) / 2;
    // staging: [lr (num int2)] [geo (3 num)] [pc (num)] [coef (2)]
    char *stage = (char *)g->d_stage;
    int2 *d_lr = (int2 *)stage;
    double *d_geo = (double *)(stage + sizeof(int2) * lr.size());
    double *d_pc = d_geo + geo.size();
    double *d_coef = d_pc + lr.size();
    HIPCALL(hipMemcpyAsync(d_lr, lr.data(), sizeof(int2) * lr.size(), hipMemcpyHostToDevice,
                           g->stream));
    HIPCALL(hipMemcpyAsync(d_geo, geo.data(), sizeof(double) * geo.size(), hipMemcpyHostToDevice,
                           g->stream));
    hipLaunchKernelGGL((k_polar<S>), dim3(1), dim3(RT), 0, g->stream, (const S *)d.d_P,
                       (const int2 *)d_lr, (const double *)d_geo, num, reference[4], dyn, length,
                       d_pc, d_coef);
    HIPCALL(hipGetLastError());
    HIPCALL(hipMemcpyAsync(coef, d_coef, sizeof(double) * 2, hipMemcpyDeviceToHost, g->stream));
    if (pc && num > 0) {
        HIPCALL(hipMemcpyAsync(pc, d_pc, sizeof(double) * num, hipMemcpyDeviceToHost, g->stream));
    }
    return teal_stream_wait(g);
}

}  // namespace

// ------------------------------------------------------------------ forcing source
template <typename S>
int teal_prepare_source(TealGpu *g, const S *P)
{
    if (g->source == TEAL_SOURCE_NONE) {
        return 0;
    }
    const int e = g->source_entity;
    return cell_sums<S, SUM_AVERAGE>(g, P, 0, nullptr, g->mesh->entities.cell_off[e],
                                     g->mesh->entities.cell_off[e + 1], g->d_forcing);
}

template int teal_prepare_source<double>(TealGpu *, const double *);
template int teal_prepare_source<float>(TealGpu *, const float *);

extern "C" int teal_gpu_set_source(TealGpu *g, int kind, int entity, const double *param)
{
    if (check_single(g, "set_source")) return -1;
    if (kind == TEAL_SOURCE_NONE) {
        g->source = TEAL_SOURCE_NONE;
        g->graph_key.stages = -1;  // a captured step graph no longer matches
        return 0;
    }
    if (kind != TEAL_SOURCE_FORCING || !param || !(param[1] != 0)) {
        teal_gpu_set_error("set_source: invalid source (%d) or parameters", kind);
        return -1;
    }
    if (inner_entity(g, entity, "set_source") || ensure_buffers(g)) return -1;
    if (!g->d_forcing) {
        HIPCALL(hipMalloc((void **)&g->d_forcing, sizeof(double) * 8));
    }
    const double host[8] = {0, 0, 0, 0, 0, param[0], param[1], 0};
    HIPCALL(hipMemcpy(g->d_forcing, host, sizeof(host), hipMemcpyHostToDevice));
    g->source = kind;
    g->source_entity = entity;
    g->graph_key.stages = -1;
    return 0;
}

// ------------------------------------------------------------------ diagnostics
extern "C" int teal_gpu_average(TealGpu *g, int entity, const double *primitive, double *average)
{
    if (check_single(g, "average") || inner_entity(g, entity, "average") || ensure_buffers(g)) {
        return -1;
    }
    if (!average) {
        teal_gpu_set_error("average: null output");
        return -1;
    }
    return (g->precision == TEAL_FP64) ? average_impl(g, g->d64, entity, primitive, average)
                                       : average_impl(g, g->d32, entity, primitive, average);
}

extern "C" int teal_gpu_norm(TealGpu *g, const double *primitive, const double *reference,
                             double *norm)
{
    if (check_single(g, "norm") || ensure_buffers(g)) return -1;
    if (!primitive || !reference || !norm) {
        teal_gpu_set_error("norm: null argument");
        return -1;
    }
    const size_t n = 5 * (size_t)g->mesh_ni;
    double *d_p = g->d_stage;
    double *d_ref = g->d_stage + n;  // the staging area holds 15 values per mesh row
    HIPCALL(hipMemcpyAsync(d_p, primitive, sizeof(double) * n, hipMemcpyHostToDevice, g->stream));
    HIPCALL(hipMemcpyAsync(d_ref, reference, sizeof(double) * n, hipMemcpyHostToDevice, g->stream));
    if (cell_sums<double, SUM_NORM>(g, d_p, 1, d_ref, 0, g->mesh_ni, g->d_diag)) return -1;
    HIPCALL(hipMemcpyAsync(norm, g->d_diag, sizeof(double) * 5, hipMemcpyDeviceToHost, g->stream));
    return teal_stream_wait(g);
}

extern "C" int teal_gpu_polar(TealGpu *g, int entity, double *primitive, const double *reference,
                              double length, double *coef, double *pressure_c)
{
    if (check_single(g, "polar")) return -1;
    const Mesh *m = g->mesh;
    if (entity < 0 || entity >= m->entities.off_boundary) {  // inner entities have no faces
        teal_gpu_set_error("polar: invalid entity (%d)", entity);
        return -1;
    }
    if (!reference || !coef || !(length > 0)) {
        teal_gpu_set_error("polar: invalid reference, output or length");
        return -1;
    }
    const size_t need = sizeof(int2) + 4 * sizeof(double);  // per face, see polar_impl
    const size_t num = m->entities.face_off[entity + 1] - m->entities.face_off[entity];
    if ((num + 1) * need + 2 * sizeof(double) > sizeof(double) * 15 * (size_t)m->cells.num) {
        teal_gpu_set_error("polar: entity too large for the staging area");
        return -1;
    }
    return (g->precision == TEAL_FP64)
               ? polar_impl(g, g->d64, entity, primitive, reference, length, coef, pressure_c)
               : polar_impl(g, g->d32, entity, primitive, reference, length, coef, pressure_c);
}

void teal_diagnostics_release(TealGpu *g)
{
    void *ptrs[] = {g->d_forcing, g->d_cellpart, g->d_cell_inv, g->d_diag};
    for (void *p : ptrs) {
        if (p) {
            (void)hipFree(p);
        }
    }
    g->d_forcing = nullptr;
    g->d_cellpart = nullptr;
    g->d_cell_inv = nullptr;
    g->d_diag = nullptr;
}
