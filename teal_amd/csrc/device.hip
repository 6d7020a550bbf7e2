// device.hip -- GPU state, mesh upload and the C-ABI of the residual path (teal_gpu.h).
// Compiled with -ffp-contract=off: the FP64 kernels instantiated here are bit-compatible
// with the reference.  The FP32 (mixed) kernels live in mixed.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include <immintrin.h>
#include <sched.h>

#include "coherent.h"
#include "launch.h"

using namespace fv;

extern template struct Kern<float>;  // instantiated in mixed.hip

// ------------------------------------------------------------------ errors
static char g_error[512] = "";

void teal_gpu_set_error(const char *fmt, ...)
{
    va_list args;
    va_start(args, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, args);
    va_end(args);
}

extern "C" const char *teal_gpu_last_error(void) { return g_error; }

#define HIPCALL(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            teal_gpu_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                               __FILE__, __LINE__);                                       \
            return -1;                                                                    \
        }                                                                                 \
    } while (0)

// ------------------------------------------------------------------ small kernels
// one workgroup per instance; times[i] (or `time` for all, if times is null)
__global__ void k_init_scalars(Scalars *scalars, const double *times, double time)
{
    Scalars *sc = &scalars[blockIdx.x];
    if (times) {
        time = times[blockIdx.x];
    }
    sc->dtmin_bits = 0x7ff0000000000000ULL;
    sc->step0 = 0;
    sc->step = 0;
    sc->time = time;
    for (int v = 0; v < 5; v++) sc->residual[v] = 0;
    sc->error = 0;
}

// The per-step reductions read one partial per group of 64 rows (a 128^3 mesh has 32,768).
// A single workgroup would spend tens of microseconds on the load latency, so each instance's
// groups are cut into fixed chunks, one workgroup per chunk (grid: chunks x instances).  Each
// chunk result goes to `chunk`, the workgroup that finishes an instance's chunks last (a
// ticket per instance) combines the chunk results in chunk order and writes the instance's
// scalars, then resets the ticket for the next launch.  The order of every sum is fixed by the
// chunk and group numbering alone, never by which workgroup finishes first.  Chunk results
// pass between workgroups through coherent.h's device-scope atomics.
//
// Instance i with groups [g0, g1) keeps its chunk results in rows g0 / chunk_size + i onwards
// of `chunk`: those rows never overlap the next instance's, so `chunk` needs
// ngroups / chunk_size + ninst rows.
__device__ inline long chunk_row0(int g0, int inst, int chunk_size)
{
    return (long)(g0 / chunk_size) + inst;
}

// workgroups of a chunked reduction over `n` groups
static inline unsigned chunks_of(int n, int chunk_size)
{
    return (unsigned)std::max(1, (n + chunk_size - 1) / chunk_size);
}

// true in every thread of the workgroup that arrives last of `n` at `ticket` (the ticket then
// reset to 0).  The chunk result of this workgroup was stored by wave 0 (st_shared_wg); the
// last workgroup reads the others' with ld_shared_wg after this returns.
__device__ inline bool last_workgroup(unsigned *ticket, unsigned n)
{
    __shared__ int last;
    __syncthreads();
    if (threadIdx.x == 0) {
        wait_memory();  // wave 0's chunk-result stores complete before the ticket
        last = (atomicAdd(ticket, 1u) == n - 1);
        if (last) {
            atomicExch(ticket, 0u);
        }
    }
    __syncthreads();
    return last;
}

// dt minimum over the minima of the groups of 64 rows (src/equations/timestep.c:22-29: a min,
// so the order does not matter); begin != 0 also sets step0 = courant * dt and step =
// min(step0, max_step) for the step about to run (src/simulation/advance.c:92-93).  Instance
// blockIdx.y owns the groups [blk_off[y], blk_off[y + 1]); its result goes to scalars[y].
constexpr int DT_THREADS = 1024;
constexpr int DT_CHUNK = 4 * DT_THREADS;  // groups per workgroup: four loads per thread

__device__ inline double block_min(double dt)
{
    __shared__ double red[DT_THREADS / 64];
    for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off));
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = dt;
    }
    __syncthreads();
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) dt = fmin(dt, red[w]);
    return dt;
}

__global__ __launch_bounds__(DT_THREADS) void k_dt_min(Scalars *scalars, const double *dtpart,
                                                      const int *blk_off, double *chunk,
                                                      unsigned *tickets, int begin, double courant,
                                                      double max_step)
{
    const int inst = blockIdx.y;
    const int g0 = blk_off[inst];
    const int g1 = blk_off[inst + 1];
    const int nch = std::max(1, (g1 - g0 + DT_CHUNK - 1) / DT_CHUNK);
    if ((int)blockIdx.x >= nch) return;
    const double inf = __longlong_as_double(0x7ff0000000000000LL);
    double d[4] = {inf, inf, inf, inf};  // four loads in flight per thread
    const int c0 = g0 + (blockIdx.x * DT_CHUNK);
    const int c1 = std::min(c0 + DT_CHUNK, g1);
    for (int u = 0; u < 4; u++) {
        int b = c0 + (u * DT_THREADS) + threadIdx.x;
        if (b < c1) {
            d[u] = dtpart[b];
        }
    }
    double dt = block_min(fmin(fmin(d[0], d[1]), fmin(d[2], d[3])));
    if (nch > 1) {
        double *ch = chunk + chunk_row0(g0, inst, DT_CHUNK);
        if (threadIdx.x == 0) {
            st_shared_wg(&ch[blockIdx.x], dt);
        }
        if (!last_workgroup(&tickets[inst], nch)) return;
        dt = inf;
        for (int x = threadIdx.x; x < nch; x += DT_THREADS) dt = fmin(dt, ld_shared_wg(&ch[x]));
        __syncthreads();  // block_min's shared array is reused
        dt = block_min(dt);
    }
    if (threadIdx.x == 0) {
        Scalars *sc = &scalars[inst];
        sc->dtmin_bits = (unsigned long long)__double_as_longlong(dt);
        if (begin) {
            sc->step0 = courant * dt;
            sc->step = fmin(sc->step0, max_step);
        }
    }
}

// residual = sqrt(sum V R^2 / sum V) from the partials of the groups of 64 rows [ngroups][5] in
// a fixed order (src/equations/residual.c:8-33), then time += step (advance.c:111).
// Within a chunk of STEP_END_CHUNK groups, thread t of the STEP_END_THREADS = 5 x 128 sums
// variable t % 5 of groups t / 5, t / 5 + 128, ... (coalesced reads) into STEP_END_SUMS running
// sums taken round-robin (that many loads in flight), adds those pairwise, then a fixed tree
// over the 128 partial sums.  The chunk sums are then summed the same way, in chunk order.
constexpr int STEP_END_THREADS = 640;
constexpr int STEP_END_SUMS = 16;
constexpr int STEP_END_CHUNK = STEP_END_SUMS * (STEP_END_THREADS / 5);  // one load round
static_assert(STEP_END_THREADS % 5 == 0 && (STEP_END_SUMS & (STEP_END_SUMS - 1)) == 0,
              "k_step_end layout");

// red[v] (v < 5, block shared memory of STEP_END_THREADS doubles) = the sum of variable v over
// partials[e0 .. e1) (rows of 5) in the fixed order above; SHARED: partials written by other
// workgroups of the same launch (read with ld_shared_wg)
template <bool SHARED = false>
__device__ inline void sum_partials(const double *partials, long e0, long e1, double *red)
{
    auto ld = [&](long e) { return SHARED ? ld_shared_wg(&partials[e]) : partials[e]; };
    const int t = threadIdx.x;
    const int lanes = STEP_END_THREADS / 5;
    double acc[STEP_END_SUMS] = {};
    long e = e0 + t;
    for (; e + ((STEP_END_SUMS - 1) * STEP_END_THREADS) < e1; e += STEP_END_SUMS * STEP_END_THREADS) {
        for (int u = 0; u < STEP_END_SUMS; u++) acc[u] += ld(e + (u * STEP_END_THREADS));
    }
    for (int u = 0; e < e1; e += STEP_END_THREADS, u++) {
        acc[u] += ld(e);
    }
    for (int h = STEP_END_SUMS / 2; h > 0; h >>= 1) {
        for (int u = 0; u < h; u++) acc[u] += acc[u + h];
    }
    __syncthreads();  // red may still be read by a previous call's threads
    red[t] = acc[0];
    __syncthreads();
    for (int h = lanes / 2; h > 0; h >>= 1) {
        if (t < h * 5) {
            red[t] += red[t + (h * 5)];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(STEP_END_THREADS) void k_step_end(Scalars *scalars, const double *partials,
                                                               const int *blk_off,
                                                               const double *volume, double *chunk,
                                                               unsigned *tickets, int advance_time)
{
    __shared__ double red[STEP_END_THREADS];
    const int inst = blockIdx.y;
    const int g0 = blk_off[inst];
    const int g1 = blk_off[inst + 1];
    const int nch = std::max(1, (g1 - g0 + STEP_END_CHUNK - 1) / STEP_END_CHUNK);
    if ((int)blockIdx.x >= nch) return;
    const int c0 = g0 + (blockIdx.x * STEP_END_CHUNK);
    const int c1 = std::min(c0 + STEP_END_CHUNK, g1);
    sum_partials(partials, (long)c0 * 5, (long)c1 * 5, red);
    if (nch > 1) {
        // a single chunk's sums equal their sum over that one row: adding the tree's zeros to
        // sums of squares is exact
        double *ch = chunk + (chunk_row0(g0, inst, STEP_END_CHUNK) * 5);
        if (threadIdx.x < 5) {
            st_shared_wg(&ch[(blockIdx.x * 5) + threadIdx.x], red[threadIdx.x]);
        }
        if (!last_workgroup(&tickets[inst], nch)) return;
        sum_partials<true>(ch, 0, (long)nch * 5, red);
    }
    if (threadIdx.x == 0) {
        Scalars *sc = &scalars[inst];
        const double sum_volume = volume[inst];
        for (int v = 0; v < 5; v++) sc->residual[v] = sqrt(red[v] / sum_volume);
        if (advance_time) {
            sc->time += sc->step;
        }
    }
}

// out[v] = sqrt(sum V R_v^2 / sum V) from the partials of `ng` groups of 64 rows, in exactly
// k_step_end's order for one instance, by one workgroup that keeps the chunk sums in shared
// memory (at most STEP_END_THREADS / 5 chunks) and writes a caller's array.
constexpr int RESIDUAL_OUT_MAX_GROUPS = STEP_END_CHUNK * (STEP_END_THREADS / 5);

__global__ __launch_bounds__(STEP_END_THREADS) void k_residual_out(const double *partials, int ng,
                                                                   const double *volume, double *out)
{
    __shared__ double red[STEP_END_THREADS];
    __shared__ double sums[STEP_END_THREADS];
    const int nch = std::max(1, (ng + STEP_END_CHUNK - 1) / STEP_END_CHUNK);
    for (int x = 0; x < nch; x++) {
        const long c0 = (long)x * STEP_END_CHUNK;
        sum_partials(partials, c0 * 5, std::min<long>(c0 + STEP_END_CHUNK, ng) * 5, red);
        if (threadIdx.x < 5) {
            sums[(x * 5) + threadIdx.x] = red[threadIdx.x];
        }
    }
    __syncthreads();
    if (nch > 1) {
        sum_partials(sums, 0, (long)nch * 5, red);
    }
    if (threadIdx.x < 5) {
        out[threadIdx.x] = sqrt(red[threadIdx.x] / volume[0]);
    }
}

// residual partials of a stored derivative, one per group of 64 rows: device rows [5][ni]
// (perm null), or host rows in the reference layout [rows][5], device row c reading host row
// perm[c]
__global__ void k_residual_partials(const double *__restrict__ R, const int *__restrict__ perm,
                                    const double *__restrict__ vol64, int ni, double *partials)
{
    int c = (blockIdx.x * BLOCK) + threadIdx.x;
    double s[5] = {0, 0, 0, 0, 0};
    if (c < ni) {
        const long h = perm ? perm[c] : -1;
        for (int v = 0; v < 5; v++) {
            double r = perm ? ((h >= 0) ? R[(h * 5) + v] : 0.0) : R[((long)v * ni) + c];
            s[v] = vol64[c] * (r * r);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int v = 0; v < 5; v++) s[v] += __shfl_xor(s[v], off);
    }
    if ((threadIdx.x & 63) == 0 && c < ni) {
        for (int v = 0; v < 5; v++) partials[((long)(c >> 6) * 5) + v] = s[v];
    }
}

// U0 = prim2cons(P) in FP64 from host rows (reference numbering) for the inner cells in
// device order: the step's U0 comes from the caller's FP64 state, also on the mixed path.
__global__ void k_u0_from_host(const double *__restrict__ host, const int *__restrict__ perm,
                               int ni, double gamma, double *__restrict__ U0)
{
    int d = (blockIdx.x * BLOCK) + threadIdx.x;
    if (d >= ni) return;
    const double rest[5] = {1, 0, 0, 0, 1};  // padding rows of a batch
    const double *p = (perm[d] < 0) ? rest : &host[(long)perm[d] * 5];
    double u[5];
    prim2cons<double>(u, p, gamma);
    for (int v = 0; v < 5; v++) U0[((long)v * ni) + d] = u[v];
}

// host rows of the inner cells = cons2prim(U) in FP64 (the mixed path's FP64 solution)
__global__ void k_prim_from_u(const double *__restrict__ U, const int *__restrict__ perm, int ni,
                              double gamma, double *__restrict__ host)
{
    int d = (blockIdx.x * BLOCK) + threadIdx.x;
    if (d >= ni || perm[d] < 0) return;
    double u[5], p[5];
    for (int v = 0; v < 5; v++) u[v] = U[((long)v * ni) + d];
    cons2prim<double>(p, u, gamma);
    for (int v = 0; v < 5; v++) host[((long)perm[d] * 5) + v] = p[v];
}

// ------------------------------------------------------------------ mesh upload
template <typename T>
static int upload(T **dst, const std::vector<T> &src)
{
    size_t bytes = src.size() * sizeof(T);
    if (bytes == 0) {
        *dst = nullptr;
        return 0;
    }
    HIPCALL(hipMalloc((void **)dst, bytes));
    HIPCALL(hipMemcpy(*dst, src.data(), bytes, hipMemcpyHostToDevice));
    return 0;
}

template <typename S>
static int upload_conv(S **dst, const std::vector<double> &src)
{
    std::vector<S> tmp(src.begin(), src.end());
    return upload(dst, tmp);
}

// source cell of each periodic ghost (-1 for other cells), src/mesh/create.c:409-419
static std::vector<int> periodic_sources(const Mesh *m)
{
    std::vector<int> src(m->cells.num, -1);
    teal_mesh_periodic_sources(m, src.data());
    return src;
}

// Device numbering (the integer renumbering into the device layout):
//  - inner cells in Morton (Z-curve) order of their centres, so a wavefront's 64 cells and
//    their face neighbours occupy a compact region that stays in the XCD's L2;
//  - ghost cells keep their reference index (they follow the inner cells);
//  - faces grouped by the device index of their left cell.
// Each cell's incidence list stays sorted by REFERENCE face index, which is the order the
// reference accumulates face contributions in, so results do not depend on this numbering.
static uint64_t spread_bits(uint64_t v)  // 21 bits -> every third bit of 63
{
    v &= 0x1fffff;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

void teal_device_numbering(TealGpu *g, const Mesh *m)
{
    const int ni = m->cells.num_inner;
    const int nc = m->cells.num;
    const int nf = m->faces.num;
    // device inner rows: each instance's cells in Morton order of their centres (within the
    // instance's own bounding box); a batch pads every instance to whole blocks
    g->inst_blk_off.assign(g->ninst + 1, 0);
    std::vector<std::pair<uint64_t, int>> key;
    std::vector<int> rows;  // device row of each mesh inner cell
    rows.assign(ni, -1);
    int base = 0;
    for (int i = 0; i < g->ninst; i++) {
        int c0 = g->inst_cell_off[i];
        int c1 = g->inst_cell_off[i + 1];
        double lo[3] = {INFINITY, INFINITY, INFINITY};
        double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int c = c0; c < c1; c++) {
            const double *x = &m->cells.center[c].x;
            for (int d = 0; d < 3; d++) {
                lo[d] = fmin(lo[d], x[d]);
                hi[d] = fmax(hi[d], x[d]);
            }
        }
        key.assign(c1 - c0, {0, 0});
        for (int c = c0; c < c1; c++) {
            const double *x = &m->cells.center[c].x;
            uint64_t k = 0;
            for (int d = 0; d < 3; d++) {
                double span = hi[d] - lo[d];
                double t = (span > 0) ? (x[d] - lo[d]) / span : 0.0;
                uint64_t q = (uint64_t)llround(t * 2097151.0);
                k |= spread_bits(q) << d;
            }
            key[c - c0] = {k, c};
        }
        std::sort(key.begin(), key.end());
        for (int k = 0; k < c1 - c0; k++) {
            rows[key[k].second] = base + k;
        }
        int padded = (g->ninst == 1) ? (c1 - c0) : ((c1 - c0 + BLOCK - 1) / BLOCK) * BLOCK;
        base += padded;
        g->inst_blk_off[i + 1] = (base + BLOCK - 1) / BLOCK;
    }
    const int dni = base;  // device inner rows
    g->mesh_ni = ni;
    g->ni = dni;
    g->nc = dni + (nc - ni);
    g->cell_perm.assign(g->nc, -1);
    g->cell_inv.assign(nc, 0);
    for (int c = 0; c < ni; c++) {
        g->cell_perm[rows[c]] = c;
        g->cell_inv[c] = rows[c];
    }
    for (int c = ni; c < nc; c++) {  // ghosts keep their order after the inner rows
        g->cell_perm[dni + (c - ni)] = c;
        g->cell_inv[c] = dni + (c - ni);
    }
    std::vector<std::pair<long, int>> fkey(nf);
    for (int f = 0; f < nf; f++) {
        fkey[f] = {((long)g->cell_inv[m->faces.cell_idx[f].left] << 32) | f, f};
    }
    std::sort(fkey.begin(), fkey.end());
    g->face_perm.assign(nf, 0);
    g->face_inv.assign(nf, 0);
    for (int d = 0; d < nf; d++) {
        g->face_perm[d] = fkey[d].second;
        g->face_inv[fkey[d].second] = d;
    }
}

template <typename S>
int DevData<S>::build(const Mesh *m, TealGpu *gpu)
{
    const int ni = gpu->ni;  // device rows (padding included)
    const int nc = gpu->nc;
    const int nf = m->faces.num;
    const int nfi = m->faces.num_inner;
    const int nfb = m->faces.off_boundary;
    const std::vector<int> &cinv = gpu->cell_inv;
    const std::vector<int> &cperm = gpu->cell_perm;
    const std::vector<int> &finv = gpu->face_inv;
    const std::vector<int> &fperm = gpu->face_perm;
    std::vector<int> src = periodic_sources(m);
    auto resolve = [&](int cell) {  // periodic ghost -> its source cell, then device index
        if (cell >= m->cells.off_boundary && src[cell] >= 0) {
            cell = src[cell];
        }
        return cinv[cell];
    };

    // neighbour entries pack row << 1 | side and incidences face << 2 | boundary << 1 | side
    // into an int; every array index the kernels form from these counts is 64-bit
    if (nc >= (1 << 30) || nf >= (1 << 29)) {
        teal_gpu_set_error("mesh too large for one device handle (%d rows, %d faces; the "
                           "limits are 2^30 rows and 2^29 faces)", nc, nf);
        return -1;
    }
    // incidences: walking faces in reference order keeps each list sorted by reference index
    std::vector<int> cnt(ni, 0);
    std::vector<int> inc((size_t)W * ni, -1), nbr((size_t)W * ni, -1);
    std::vector<double> w((size_t)6 * W * ni, 0.0);  // LSQ weights and cell offsets, paired
    for (int f = 0; f < nf; f++) {
        int l = m->faces.cell_idx[f].left;
        int r = m->faces.cell_idx[f].right;
        int bnd = (f >= nfi && f < nfb);
        for (int side = 0; side < 2; side++) {
            if (side == 1 && f >= nfi) {
                break;
            }
            int c = cinv[side ? r : l];
            int k = cnt[c]++;
            if (k >= W) {
                teal_gpu_set_error("cell %d has more than %d faces", side ? r : l, W);
                return -1;
            }
            size_t idx = ((size_t)k * ni) + c;
            inc[idx] = (finv[f] << 2) | (bnd << 1) | side;
            nbr[idx] = ((side ? cinv[l] : resolve(r)) << 1) | side;
            Vector wv = side ? m->faces.weight[f].right : m->faces.weight[f].left;
            Vector ov = side ? m->faces.offset[f].right : m->faces.offset[f].left;
            // pairs (wx, wy), (wz, ox), (oy, oz), each [W][ni][2]: one 2-wide load per pair
            const double six[6] = {wv.x, wv.y, wv.z, ov.x, ov.y, ov.z};
            for (int q = 0; q < 3; q++) {
                size_t at = ((((size_t)q * W) + k) * ni + c) * 2;
                w[at] = six[2 * q];
                w[at + 1] = six[(2 * q) + 1];
            }
        }
    }
    max_inc = *std::max_element(cnt.begin(), cnt.end());
    std::vector<double> vol(ni, 1.0), proj((size_t)3 * ni, 0.0);  // padding rows: 1 and 0
    for (int d = 0; d < ni; d++) {
        int c = cperm[d];
        if (c < 0) {
            continue;
        }
        vol[d] = m->cells.volume[c];
        proj[d] = m->cells.projection[c].x;
        proj[ni + d] = m->cells.projection[c].y;
        proj[(2 * ni) + d] = m->cells.projection[c].z;
    }
    std::vector<double> area(nf), basis((size_t)9 * nf), offl((size_t)3 * nf), offr((size_t)3 * nf),
        corr((size_t)4 * nf);
    std::vector<int2> flr(nf);
    std::vector<int> crow(nf);
    std::vector<double> fcen((size_t)3 * nf), ccen((size_t)3 * nc, 0.0);
    for (int d = 0; d < nc; d++) {
        if (cperm[d] < 0) {
            continue;
        }
        const double *x = &m->cells.center[cperm[d]].x;
        for (int k = 0; k < 3; k++) ccen[((size_t)d * 3) + k] = x[k];
    }
    for (int d = 0; d < nf; d++) {
        int f = fperm[d];
        int r = m->faces.cell_idx[f].right;
        flr[d].x = cinv[m->faces.cell_idx[f].left];
        flr[d].y = (f >= nfi && f < nfb) ? ~cinv[r] : resolve(r);
        crow[d] = cinv[r];  // ghosts keep their row; a periodic face's own ghost, not its source
        const double *fc = &m->faces.center[f].x;
        for (int k = 0; k < 3; k++) fcen[((size_t)k * nf) + d] = fc[k];
        area[d] = m->faces.area[f];
        const double *b = &m->faces.basis[f].x.x;
        for (int k = 0; k < 9; k++) basis[((size_t)k * nf) + d] = b[k];
        const double *ol = &m->faces.offset[f].left.x;
        const double *orr = &m->faces.offset[f].right.x;
        for (int k = 0; k < 3; k++) {
            offl[((size_t)k * nf) + d] = ol[k];
            offr[((size_t)k * nf) + d] = orr[k];
        }
        corr[d] = m->faces.correction[f].unit.x;
        corr[nf + d] = m->faces.correction[f].unit.y;
        corr[(2 * (size_t)nf) + d] = m->faces.correction[f].unit.z;
        corr[(3 * (size_t)nf) + d] = m->faces.correction[f].norm;
    }
    if (upload(&d_inc, inc) || upload(&d_nbr, nbr) || upload(&d_flr, flr) || upload_conv(&d_w, w) ||
        upload_conv(&d_vol, vol) || upload_conv(&d_proj, proj) ||
        upload_conv(&d_area, area) || upload_conv(&d_basis, basis) ||
        upload_conv(&d_offl, offl) || upload_conv(&d_offr, offr) || upload_conv(&d_corr, corr) ||
        upload(&d_vol64, vol)) {
        return -1;
    }
    // boundary faces: left cell in device numbering, ghost slot at its reference index
    int nb = nfb - nfi;
    std::vector<int> bl(nb), br(nb), bf(nb);
    for (int i = 0; i < nb; i++) {
        bf[i] = finv[nfi + i];
        bl[i] = cinv[m->faces.cell_idx[nfi + i].left];
        br[i] = cinv[m->faces.cell_idx[nfi + i].right];
    }
    if (upload(&d_bleft, bl) || upload(&d_bright, br) || upload(&d_bface, bf)) {
        return -1;
    }
    HIPCALL(hipMalloc((void **)&d_bkind, sizeof(int) * (nb + 1)));
    HIPCALL(hipMalloc((void **)&d_bctx, sizeof(S) * 5 * (nb + 1)));
    HIPCALL(hipMemset(d_bkind, 0, sizeof(int) * (nb + 1)));
    HIPCALL(hipMemset(d_bctx, 0, sizeof(S) * 5 * (nb + 1)));
    // state records
    HIPCALL(hipMalloc((void **)&d_P, sizeof(S) * 5 * (size_t)nc));
    HIPCALL(hipMemset(d_P, 0, sizeof(S) * 5 * (size_t)nc));
    HIPCALL(hipMalloc((void **)&d_P2, sizeof(S) * 5 * (size_t)nc));
    HIPCALL(hipMemset(d_P2, 0, sizeof(S) * 5 * (size_t)nc));
    HIPCALL(hipMalloc((void **)&d_GP, sizeof(S) * 20 * (size_t)ni));
    HIPCALL(hipMemset(d_GP, 0, sizeof(S) * 20 * (size_t)ni));
    HIPCALL(hipMalloc((void **)&d_F, sizeof(S) * 10 * (size_t)(nf + 1)));
    HIPCALL(hipMemset(d_F, 0, sizeof(S) * 10 * (size_t)(nf + 1)));
    HIPCALL(hipMalloc((void **)&d_eps2, sizeof(S) * (size_t)ni));
    HIPCALL(hipMemset(d_eps2, 0, sizeof(S) * (size_t)ni));
    // both block layouts; the residual passes start on the larger one
    cur = 0;
    return build_blocks(lay[0], BLOCK, inc, ni, flr, area, basis, offl, offr, corr, fcen, crow,
                        ccen) ||
           build_blocks(lay[1], SMALL_BLOCK, inc, ni, flr, area, basis, offl, offr, corr, fcen,
                        crow, ccen);
}

// true when the face offsets and corrections equal, bit for bit, their recomputation from
// the face and cell centres with the expressions of src/mesh/generate.c:754-777
static bool centres_reproduce(const std::vector<int2> &flr, const std::vector<int> &crow,
                              const std::vector<double> &fcen, const std::vector<double> &ccen,
                              const std::vector<double> &offl, const std::vector<double> &offr,
                              const std::vector<double> &corr)
{
    const size_t nf = flr.size();
    for (size_t f = 0; f < nf; f++) {
        const double *cl = &ccen[(size_t)flr[f].x * 3];
        const double *cr = &ccen[(size_t)crow[f] * 3];
        double d[3], fc[3];
        for (int k = 0; k < 3; k++) {
            fc[k] = fcen[(k * nf) + f];
            d[k] = cr[k] - cl[k];
            if (fc[k] - cl[k] != offl[(k * nf) + f] || fc[k] - cr[k] != offr[(k * nf) + f]) {
                return false;
            }
        }
        double norm = sqrt((d[0] * d[0]) + (d[1] * d[1]) + (d[2] * d[2]));
        if (norm != corr[(3 * nf) + f]) {
            return false;
        }
        for (int k = 0; k < 3; k++) {
            if (d[k] / norm != corr[(k * nf) + f]) {
                return false;
            }
        }
    }
    return true;
}

// q of face_tangents_q for every face (basis [9][nf]) when its expressions reproduce each face's
// stored tangents, or both negated; empty otherwise.  q is read from the stored basis (nqz is
// the last component of the second tangent, nqy the magnitude of the first tangent's y) and
// the reconstruction is checked with the device's expressions.
static std::vector<double> tangent_q(const std::vector<double> &basis, size_t nf)
{
    std::vector<double> q(nf);
    auto b = [&](int k, size_t f) { return basis[(k * nf) + f]; };
    for (size_t f = 0; f < nf; f++) {
        const double nx = b(0, f), ny = b(1, f), nz = b(2, f);
        const double nqz = b(8, f);
        if (nqz > 0 && -ny / nqz == b(3, f) && nx / nqz == b(4, f) && 0 == b(5, f) &&
            -nx * nz / nqz == b(6, f) && -ny * nz / nqz == b(7, f)) {
            q[f] = nqz;
            continue;
        }
        const double nqy = fabs(b(4, f));
        if (!(nqy > 0)) {
            return {};
        }
        const double t[6] = {-nx * ny / nqy, nqy, -ny * nz / nqy, -nz / nqy, 0, nx / nqy};
        bool same = true, negated = true;
        for (int k = 0; k < 6; k++) {
            same = same && t[k] == b(3 + k, f);
            negated = negated && t[k] == -b(3 + k, f);
        }
        if (!same && !negated) {
            return {};
        }
        q[f] = -nqy;
    }
    return q;
}

// Face lists of the fused stage kernel in the layout of `rows` rows per block: block b owns
// device cells [b*rows, (b+1)*rows); its list holds every face with an incidence in the block,
// once, in device face order, and slot[k][c] = (list position << 1) | side for incidence k
// of cell c (-1: none).  The face geometry is copied into list order (faces between blocks
// twice).
template <typename S>
int DevData<S>::build_blocks(BlockSet<S> &bl, int rows, const std::vector<int> &inc, int ni,
                             const std::vector<int2> &flr, const std::vector<double> &area,
                             const std::vector<double> &basis, const std::vector<double> &offl,
                             const std::vector<double> &offr, const std::vector<double> &corr,
                             const std::vector<double> &fcen, const std::vector<int> &crow,
                             const std::vector<double> &ccen)
{
    const int nf = (int)area.size();
    const int nblk = (ni + rows - 1) / rows;
    std::vector<int> off(nblk + 1, 0), list, slot((size_t)W * ni, -1);
    std::vector<int> faces;
    bl.rows = rows;
    int &max_block_faces = bl.max_block_faces;
    max_block_faces = 0;
    for (int b = 0; b < nblk; b++) {
        int c0 = b * rows;
        int c1 = std::min(ni, c0 + rows);
        faces.clear();
        for (int c = c0; c < c1; c++) {
            for (int k = 0; k < W; k++) {
                int e = inc[((size_t)k * ni) + c];
                if (e >= 0) {
                    faces.push_back(e >> 2);
                }
            }
        }
        std::sort(faces.begin(), faces.end());
        faces.erase(std::unique(faces.begin(), faces.end()), faces.end());
        for (int c = c0; c < c1; c++) {
            for (int k = 0; k < W; k++) {
                size_t idx = ((size_t)k * ni) + c;
                int e = inc[idx];
                if (e >= 0) {
                    int pos = (int)(std::lower_bound(faces.begin(), faces.end(), e >> 2) - faces.begin());
                    slot[idx] = (pos << 1) | (e & 1);
                }
            }
        }
        list.insert(list.end(), faces.begin(), faces.end());
        off[b + 1] = (int)list.size();
        max_block_faces = std::max(max_block_faces, (int)faces.size());
    }
    const int nj = (int)list.size();
    const size_t cap = (size_t)std::max(nj, 1);
    bl.num_block_faces = nj;
    // FP64: the centre form (48 fewer bytes per face) where it reproduces the stored geometry.
    // FP32 keeps the offsets rounded from FP64 (a difference of FP32 centres loses digits) but
    // derives the correction from them and the tangents from the normal (40 fewer bytes).
    // (TEAL_GPU_STORED_OFFSETS=1 at creation keeps the stored offsets and corrections, the form
    // of meshes whose centres do not reproduce them)
    bl.centre_form = sizeof(S) == 8 && !getenv("TEAL_GPU_STORED_OFFSETS") &&
                     centres_reproduce(flr, crow, fcen, ccen, offl, offr, corr);
    const bool centre_form = bl.centre_form;
    const bool fp64 = sizeof(S) == 8;
    const int brows = fp64 ? 9 : 3;
    std::vector<int2> lflr(cap);
    std::vector<int> lcrow(cap);
    std::vector<double> larea(cap), lbasis(brows * cap), loffl, loffr, lcorr, lfcen;
    if (centre_form) {
        lfcen.assign(3 * cap, 0.0);
    }
    else {
        loffl.assign(3 * cap, 0.0);
        loffr.assign(3 * cap, 0.0);
        if (fp64) {
            lcorr.assign(4 * cap, 0.0);
        }
    }
    for (int j = 0; j < nj; j++) {
        int f = list[j];
        lflr[j] = flr[f];
        lcrow[j] = crow[f];
        larea[j] = area[f];
        for (int k = 0; k < brows; k++) lbasis[((size_t)k * nj) + j] = basis[((size_t)k * nf) + f];
        if (centre_form) {
            for (int k = 0; k < 3; k++) lfcen[((size_t)k * nj) + j] = fcen[((size_t)k * nf) + f];
            continue;
        }
        for (int k = 0; k < 3; k++) {
            loffl[((size_t)k * nj) + j] = offl[((size_t)k * nf) + f];
            loffr[((size_t)k * nj) + j] = offr[((size_t)k * nf) + f];
        }
        if (fp64) {
            for (int k = 0; k < 4; k++) lcorr[((size_t)k * nj) + j] = corr[((size_t)k * nf) + f];
        }
    }
    if (upload(&bl.d_boff, off) || upload(&bl.d_bslot, slot) || upload(&bl.d_bflr, lflr)) {
        return -1;
    }
    if (!fp64) {  // (area, nx), (ny, nz), (olx, oly), (olz, orx), (ory, orz) per face
        std::vector<double> pairs(10 * cap, 0.0);
        for (int j = 0; j < nj; j++) {
            const double ten[10] = {larea[j],
                                    lbasis[j],
                                    lbasis[(size_t)nj + j],
                                    lbasis[(2 * (size_t)nj) + j],
                                    loffl[j],
                                    loffl[(size_t)nj + j],
                                    loffl[(2 * (size_t)nj) + j],
                                    loffr[j],
                                    loffr[(size_t)nj + j],
                                    loffr[(2 * (size_t)nj) + j]};
            for (int q = 0; q < 5; q++) {
                pairs[((((size_t)q * cap) + j) * 2)] = ten[2 * q];
                pairs[((((size_t)q * cap) + j) * 2) + 1] = ten[(2 * q) + 1];
            }
        }
        return upload_conv(&bl.d_bpairs, pairs);
    }
    if (centre_form) {
        // (area, n0), (n1, n2), (fcx, fcy), (fcz, q) per face where the tangents can be rebuilt,
        // else (area, b0), (b1, b2), ..., (b7, b8), (fcx, fcy), (fcz, 0)
        const std::vector<double> tq = getenv("TEAL_GPU_TANGENTS_STORED")
                                           ? std::vector<double>()
                                           : tangent_q(lbasis, (size_t)nj);
        bl.tangents = nj > 0 && !tq.empty();
        const int np = bl.tangents ? 4 : 7;
        std::vector<double> pairs(2 * np * cap, 0.0);
        for (int j = 0; j < nj; j++) {
            double rec[14] = {larea[j]};
            if (bl.tangents) {
                for (int k = 0; k < 3; k++) rec[1 + k] = lbasis[((size_t)k * nj) + j];
                for (int k = 0; k < 3; k++) rec[4 + k] = lfcen[((size_t)k * nj) + j];
                rec[7] = tq[j];
            }
            else {
                for (int k = 0; k < 9; k++) rec[1 + k] = lbasis[((size_t)k * nj) + j];
                for (int k = 0; k < 3; k++) rec[10 + k] = lfcen[((size_t)k * nj) + j];
            }
            for (int q = 0; q < np; q++) {
                pairs[((((size_t)q * cap) + j) * 2)] = rec[2 * q];
                pairs[((((size_t)q * cap) + j) * 2) + 1] = rec[(2 * q) + 1];
            }
        }
        if (upload(&bl.d_bcrow, lcrow) || upload_conv(&bl.d_bpairs, pairs) ||
            (!d_ccen && upload_conv(&d_ccen, ccen))) {
            return -1;
        }
        return 0;
    }
    if (upload_conv(&bl.d_barea, larea) || upload_conv(&bl.d_bbasis, lbasis) ||
        upload_conv(&bl.d_boffl, loffl) || upload_conv(&bl.d_boffr, loffr) ||
        upload_conv(&bl.d_bcorr, lcorr)) {
        return -1;
    }
    return 0;
}

template <typename S>
BlockTopo DevData<S>::block_topo() const
{
    const BlockSet<S> &b = lay[cur];
    return BlockTopo{b.d_boff, b.d_bslot, b.d_bflr, b.num_block_faces, b.max_block_faces};
}

template <typename S>
Geom<S> DevData<S>::block_geom() const
{
    const BlockSet<S> &b = lay[cur];
    return Geom<S>{d_vol,     d_proj,     d_eps2,    d_w,     b.d_barea,  b.d_bbasis, b.d_boffl,
                   b.d_boffr, b.d_bcorr,  b.d_bfcen, d_ccen,  b.d_bcrow,  b.d_bpairs, b.tangents};
}

template <typename S>
void DevData<S>::release()
{
    void *ptrs[] = {d_inc, d_nbr, d_flr, d_w, d_vol, d_proj, d_area, d_basis, d_offl,
                    d_offr, d_corr, d_vol64, d_bleft, d_bright, d_bface, d_bkind, d_bctx,
                    d_P, d_P2, d_GP, d_F, d_eps2, d_ccen};
    for (void *p : ptrs) {
        if (p) {
            (void)hipFree(p);
        }
    }
    for (BlockSet<S> &b : lay) {
        void *bp[] = {b.d_boff,  b.d_bslot, b.d_bflr,  b.d_barea, b.d_bbasis, b.d_boffl,
                      b.d_boffr, b.d_bcorr, b.d_bfcen, b.d_bcrow, b.d_bpairs};
        for (void *p : bp) {
            if (p) {
                (void)hipFree(p);
            }
        }
        b = BlockSet<S>();
    }
}

template <typename S>
Geom<S> DevData<S>::geom() const
{
    return Geom<S>{d_vol,  d_proj, d_eps2, d_w,     d_area,  d_basis,
                   d_offl, d_offr, d_corr, nullptr, nullptr, nullptr};
}

template struct DevData<double>;
template struct DevData<float>;


// ------------------------------------------------------------------ TealGpu
// a handle over `mesh` whose inner cells split into `ninst` independent instances at
// inst_cell_off (one instance: the whole mesh)
static TealGpu *create_handle(const Mesh *mesh, int precision, int ninst, const int *inst_cell_off,
                              const double *inst_volume, Mesh *owned)
{
    TealGpu *g = new TealGpu();
    g->mesh = mesh;
    g->owned = owned;
    g->precision = precision;
    g->ninst = ninst;
    g->inst_cell_off.assign(inst_cell_off, inst_cell_off + ninst + 1);
    g->inst_volume.assign(inst_volume, inst_volume + ninst);
    g->nf = mesh->faces.num;
    g->nfi = mesh->faces.num_inner;
    g->nfb = mesh->faces.off_boundary;
    g->sum_volume = mesh->cells.sum_volume;
    g->periodic_src = periodic_sources(mesh);
    teal_device_numbering(g, mesh);  // sets ni, nc (device rows), mesh_ni, instance blocks
    int rc = (precision == TEAL_FP64) ? g->d64.build(mesh, g) : g->d32.build(mesh, g);
    if (rc == 0) {  // the fused stage kernel needs a block's face list in LDS
        int dev = 0, lds_max = 0;
        const char *split = getenv("TEAL_GPU_SPLIT");
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
            lds_max = 0;
        }
        int want = !(split && atoi(split) != 0);
        g->lds_max = want ? lds_max : 0;
        g->d64.fused = g->d64.block_lds() > 0 && g->d64.block_lds() <= (size_t)g->lds_max;
        g->d32.fused = g->d32.block_lds() > 0 && g->d32.block_lds() <= (size_t)g->lds_max;
    }
    // the gradient kernel stages neighbour states in LDS on meshes of tetrahedra, whose
    // gathers miss the caches; the autotuner may choose otherwise
    g->grad_stage = (rc == 0 && ((precision == TEAL_FP64) ? g->d64.max_inc : g->d32.max_inc) <= 4);
    g->nblocks = grid_of(g->ni);
    // partial sums per group of 64 rows; a batch pads every instance to whole blocks, so an
    // instance's groups are its blocks' groups
    constexpr int GPB = BLOCK / 64;  // groups per block
    g->ngroups = (g->ni + 63) / 64;
    std::vector<int> grp_inst(std::max(g->ngroups, 1), 0), inst_grp_off(ninst + 1, 0);
    for (int i = 0; i < ninst; i++) {
        inst_grp_off[i + 1] = std::min(g->ngroups, g->inst_blk_off[i + 1] * GPB);
        for (int q = inst_grp_off[i]; q < inst_grp_off[i + 1]; q++) grp_inst[q] = i;
    }
    std::vector<int> all_grp = {0, g->ngroups};
    if (rc == 0 && (upload(&g->d_cell_perm, g->cell_perm) || upload(&g->d_grp_inst, grp_inst) ||
                    upload(&g->d_inst_grp_off, inst_grp_off) || upload(&g->d_all_grp, all_grp) ||
                    upload(&g->d_inst_volume, g->inst_volume))) {
        rc = -1;
    }
    if (rc == 0 && hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
        teal_gpu_set_error("teal_gpu_create: hipStreamCreate failed");
        rc = -1;
    }
    if (rc == 0 && hipEventCreateWithFlags(&g->ev_wait, hipEventDisableTiming) != hipSuccess) {
        teal_gpu_set_error("teal_gpu_create: hipEventCreate failed");
        rc = -1;
    }
    // chunk results of the per-step reductions (see chunk_row0), and their tickets
    g->max_inst_groups = 0;
    for (int i = 0; i < ninst; i++) {
        g->max_inst_groups = std::max(g->max_inst_groups, inst_grp_off[i + 1] - inst_grp_off[i]);
    }
    const size_t chunk_rows = (size_t)g->ngroups / STEP_END_CHUNK + ninst + 1;
    // host-layout staging: up to 15 values per mesh row, or 5 per device row
    size_t stage = std::max((size_t)15 * mesh->cells.num, (size_t)5 * g->nc);
    if (rc == 0 &&
        (hipMalloc((void **)&g->d_stage, sizeof(double) * stage) != hipSuccess ||
         hipMalloc((void **)&g->d_U0, sizeof(double) * 5 * (size_t)g->ni) != hipSuccess ||
         hipMalloc((void **)&g->d_R, sizeof(double) * 5 * (size_t)g->ni) != hipSuccess ||
         hipMalloc((void **)&g->d_partials, sizeof(double) * 5 * (size_t)std::max(g->ngroups, 1)) !=
             hipSuccess ||
         hipMalloc((void **)&g->d_dtpart, sizeof(double) * (size_t)std::max(g->ngroups, 1)) != hipSuccess ||
         hipMalloc((void **)&g->d_chunk, sizeof(double) * 5 * chunk_rows) != hipSuccess ||
         hipMalloc((void **)&g->d_tickets, sizeof(unsigned) * (size_t)ninst) != hipSuccess ||
         hipMalloc((void **)&g->d_scalars, sizeof(Scalars) * (size_t)ninst) != hipSuccess ||
         hipHostMalloc((void **)&g->h_scalars, sizeof(Scalars) * (size_t)ninst, hipHostMallocDefault) !=
             hipSuccess)) {
        teal_gpu_set_error("teal_gpu_create: device allocation failed");
        rc = -1;
    }
    if (rc == 0) {
        hipLaunchKernelGGL(k_init_scalars, dim3(ninst), dim3(1), 0, g->stream, g->d_scalars,
                           (const double *)nullptr, 0.0);
        if (hipMemsetAsync(g->d_tickets, 0, sizeof(unsigned) * (size_t)ninst, g->stream) != hipSuccess) {
            rc = -1;
        }
        if (hipStreamSynchronize(g->stream) != hipSuccess) {
            teal_gpu_set_error("teal_gpu_create: initialization kernel failed");
            rc = -1;
        }
    }
    if (rc != 0) {
        teal_gpu_destroy(g);
        return nullptr;
    }
    return g;
}

static int device_ready(const char *who)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        teal_gpu_set_error("%s: no HIP device is available", who);
        return -1;
    }
    return 0;
}

extern "C" TealGpu *teal_gpu_create(const Mesh *mesh, int precision)
{
    if (!mesh || !mesh->faces.area) {
        teal_gpu_set_error("teal_gpu_create: mesh is not generated");
        return nullptr;
    }
    if (precision != TEAL_FP64 && precision != TEAL_MIXED) {
        teal_gpu_set_error("teal_gpu_create: invalid precision (%d)", precision);
        return nullptr;
    }
    if (device_ready("teal_gpu_create")) {
        return nullptr;
    }
    int off[2] = {0, mesh->cells.num_inner};
    double vol = mesh->cells.sum_volume;
    return create_handle(mesh, precision, 1, off, &vol, nullptr);
}

extern "C" TealGpu *teal_gpu_create_batch(const Mesh *const *meshes, int num, int precision)
{
    if (!meshes || num < 1) {
        teal_gpu_set_error("teal_gpu_create_batch: no meshes (%d)", num);
        return nullptr;
    }
    if (precision != TEAL_FP64 && precision != TEAL_MIXED) {
        teal_gpu_set_error("teal_gpu_create_batch: invalid precision (%d)", precision);
        return nullptr;
    }
    for (int i = 0; i < num; i++) {
        if (!meshes[i] || !meshes[i]->faces.area) {
            teal_gpu_set_error("teal_gpu_create_batch: mesh %d is not generated", i);
            return nullptr;
        }
    }
    if (device_ready("teal_gpu_create_batch")) {
        return nullptr;
    }
    std::vector<int> off(num + 1, 0);
    std::vector<double> vol(num);
    for (int i = 0; i < num; i++) {
        off[i + 1] = off[i] + meshes[i]->cells.num_inner;
        vol[i] = meshes[i]->cells.sum_volume;
    }
    Mesh *u = teal_mesh_union(meshes, num);
    TealGpu *g = create_handle(u, precision, num, off.data(), vol.data(), u);
    if (!g) {
        return nullptr;
    }
    // host rows of each instance in the union (the group order of teal_mesh_union)
    g->inst_rows.assign((size_t)6 * num, 0);
    int inner = 0, boundary = u->cells.num_inner, periodic = u->cells.off_boundary;
    for (int i = 0; i < num; i++) {
        const Mesh *m = meshes[i];
        int *r = &g->inst_rows[(size_t)6 * i];
        r[0] = inner;
        r[1] = m->cells.num_inner;
        r[2] = boundary;
        r[3] = m->cells.off_boundary - m->cells.num_inner;
        r[4] = periodic;
        r[5] = m->cells.num - m->cells.off_boundary;
        inner += r[1];
        boundary += r[3];
        periodic += r[5];
    }
    return g;
}

extern "C" void teal_gpu_destroy(TealGpu *g)
{
    if (!g) {
        return;
    }
    if (g->stream) {
        (void)hipStreamSynchronize(g->stream);
    }
    g->d64.release();
    g->d32.release();
    void *ptrs[] = {g->d_stage,     g->d_U0,        g->d_R,          g->d_partials,
                    g->d_dtpart,    g->d_chunk,     g->d_tickets,    g->d_scalars,
                    g->d_cell_perm, g->d_grp_inst,  g->d_inst_grp_off, g->d_all_grp,
                    g->d_inst_volume};
    for (void *p : ptrs) {
        if (p) {
            (void)hipFree(p);
        }
    }
    if (g->h_scalars) {
        (void)hipHostFree(g->h_scalars);
    }
    g->ev_flux.release();
    g->ev_grad.release();
    g->ev_face.release();
    teal_krylov_release(g);
    teal_diagnostics_release(g);
    if (g->ev_wait) {
        (void)hipEventDestroy(g->ev_wait);
    }
    if (g->stream) {
        (void)hipStreamDestroy(g->stream);
    }
    if (g->graph) {
        (void)hipGraphExecDestroy(g->graph);
    }
    if (g->owned) {
        mesh_destroy(g->owned);
    }
    delete g;
}

template <typename S>
static int configure_impl(TealGpu *g, DevData<S> &d, const TealGpuConfig *cfg)
{
    int nb = g->nfb - g->nfi;
    const Mesh *m = g->mesh;
    int nent = m->entities.off_boundary - m->entities.num_inner;
    if (cfg->num_boundary != nent) {
        teal_gpu_set_error("configure: %d boundary conditions given, mesh has %d", cfg->num_boundary,
                           nent);
        return -1;
    }
    std::vector<int> kind(nb + 1, 0);
    std::vector<S> ctx((size_t)5 * (nb + 1), S(0));
    for (int e = 0; e < nent; e++) {
        int beg = m->entities.face_off[m->entities.num_inner + e];
        int end = m->entities.face_off[m->entities.num_inner + e + 1];
        for (int f = beg; f < end; f++) {
            int i = f - g->nfi;
            kind[i] = cfg->bc ? cfg->bc[e] : 0;
            if (kind[i] <= TEAL_BC_NONE || kind[i] > TEAL_BC_ADIABATIC_WALL) {
                teal_gpu_set_error("configure: boundary entity %s has no device condition",
                                   m->entities.name[m->entities.num_inner + e]);
                return -1;
            }
            for (int v = 0; v < 5; v++) {
                ctx[((size_t)i * 5) + v] = cfg->bc_context ? (S)cfg->bc_context[(e * 5) + v] : S(0);
            }
        }
    }
    HIPCALL(hipMemcpy(d.d_bkind, kind.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice));
    HIPCALL(hipMemcpy(d.d_bctx, ctx.data(), sizeof(S) * 5 * (nb + 1), hipMemcpyHostToDevice));
    if (cfg->limiter == VENKATAKRISHNAN_ID) {
        std::vector<S> e2(g->ni);
        for (int d = 0; d < g->ni; d++) {  // src/equations/limiter.c:8-21
            int c = g->cell_perm[d];
            if (c < 0) {
                e2[d] = S(0);
                continue;
            }
            double v = cfg->coefficient * cbrt(m->cells.volume[c]);
            e2[d] = cfg->eps2 ? (S)cfg->eps2[c] : (S)(v * v * v);
        }
        HIPCALL(hipMemcpy(d.d_eps2, e2.data(), sizeof(S) * g->ni, hipMemcpyHostToDevice));
    }
    return 0;
}

extern "C" int teal_gpu_configure(TealGpu *g, const TealGpuConfig *cfg)
{
    if (!g || !cfg) {
        teal_gpu_set_error("configure: null argument");
        return -1;
    }
    if (cfg->space_order != 1 && cfg->space_order != 2) {
        teal_gpu_set_error("configure: invalid space order (%d)", cfg->space_order);
        return -1;
    }
    if (cfg->flux < TEAL_FLUX_NONE || cfg->flux > TEAL_FLUX_CENTRAL) {
        teal_gpu_set_error("configure: invalid convective flux (%d)", cfg->flux);
        return -1;
    }
    if (cfg->limiter < 0 || cfg->limiter > 2) {
        teal_gpu_set_error("configure: invalid limiter (%d)", cfg->limiter);
        return -1;
    }
    if (cfg->viscous && cfg->space_order != 2) {
        teal_gpu_set_error("configure: the viscous flux needs space order 2");
        return -1;
    }
    g->cfg = *cfg;
    g->cfg.bc = nullptr;
    g->cfg.bc_context = nullptr;
    g->cfg.eps2 = nullptr;
    int nent = cfg->num_boundary > 0 ? cfg->num_boundary : 0;
    g->cfg_bc.assign(cfg->bc ? cfg->bc : nullptr, cfg->bc ? cfg->bc + nent : nullptr);
    g->cfg_ctx.assign(cfg->bc_context ? cfg->bc_context : nullptr,
                      cfg->bc_context ? cfg->bc_context + (5 * nent) : nullptr);
    g->cfg_eps2.assign(cfg->eps2 ? cfg->eps2 : nullptr, cfg->eps2 ? cfg->eps2 + g->mesh_ni : nullptr);
    g->configured = 1;
    g->u0_valid = 0;
    if (g->graph) {  // captured launches carry the old configuration
        (void)hipGraphExecDestroy(g->graph);
        g->graph = nullptr;
    }

    if (g->precision == TEAL_FP64) {
        return configure_impl(g, g->d64, cfg);
    }
    if (g->d64.d_P && configure_impl(g, g->d64, cfg)) {  // the FP64 mirror, when built
        return -1;
    }
    return configure_impl(g, g->d32, cfg);
}

int teal_fp64_mirror(TealGpu *g)
{
    if (g->precision == TEAL_FP64 || g->d64.d_P) {
        return 0;
    }
    if (g->d64.build(g->mesh, g)) {
        g->d64.release();
        g->d64 = DevData<double>();
        return -1;
    }
    g->d64.fused = g->d64.block_lds() > 0 && g->d64.block_lds() <= (size_t)g->lds_max;
    TealGpuConfig c = g->cfg;
    c.bc = g->cfg_bc.empty() ? nullptr : g->cfg_bc.data();
    c.bc_context = g->cfg_ctx.empty() ? nullptr : g->cfg_ctx.data();
    c.eps2 = g->cfg_eps2.empty() ? nullptr : g->cfg_eps2.data();
    return configure_impl(g, g->d64, &c);
}

// ------------------------------------------------------------------ kernel launches
Topo TealGpu::topo() const
{
    if (precision == TEAL_FP64) {
        return Topo{ni, nc, nf, d64.d_inc, d64.d_nbr, d64.d_flr, d64.max_inc};
    }
    return Topo{ni, nc, nf, d32.d_inc, d32.d_nbr, d32.d_flr, d32.max_inc};
}

Phys TealGpu::phys() const { return Phys{cfg.gamma, cfg.mu, cfg.prandtl, cfg.navierstokes}; }

// one residual evaluation of the state in P: ghosts, gradient + limiter, face fluxes,
// accumulation.  mode 0 writes R; mode 1 is an RK stage updating P (last: + U0, dt, residual)
template <typename S>
int teal_residual_pass(TealGpu *g, DevData<S> &d, int mode, double alpha, int last)
{
    const int nb = g->nfb - g->nfi;
    if (nb > 0) {
        Kern<S>::boundary(g->stream, nb, g->nf, d, g->cfg.gamma);
    }
    if (g->cfg.space_order == 2) {
        bool timed = g->timing && g->ev_grad.begin(g->stream);
        // inviscid passes keep only the limited gradient (15 values, in rows of 16)
        Kern<S>::grad_limit(g->stream, g->cfg.limiter, g->cfg.viscous ? 0 : 2, g->grad_block,
                            g->grad_stage, g->topo(), d.geom(), d.d_P, d.d_GP);
        if (timed) {
            g->ev_grad.end(g->stream);
        }
    }
    if (teal_prepare_source(g, (const S *)d.d_P)) return -1;
    const int conv = g->cfg.flux != TEAL_FLUX_NONE;
    StageArgs a;
    a.forcing = g->source ? g->d_forcing : nullptr;
    a.R = g->d_R;
    a.U0 = g->d_U0;
    a.U0_next = g->d_U0;
    a.alpha = alpha;
    a.scalars = g->d_scalars;
    a.grp_inst = (g->ninst > 1) ? g->d_grp_inst : nullptr;
    a.partials = g->d_partials;
    a.dtpart = g->d_dtpart;
    a.keep_u = (g->precision == TEAL_MIXED);
    int kmode = (mode == 0) ? 0 : (last ? 2 : 1);
    if (d.fused) {  // one kernel: face fluxes in LDS, accumulation, stage update into d_P2
        bool timed = g->timing && g->ev_flux.begin(g->stream);
        int rc = Kern<S>::flux_block(g->stream, g->cfg.flux, g->cfg.viscous, g->cfg.space_order, kmode,
                                     d.block_rows(), d.block_lds(g->cfg.viscous, conv), g->topo(),
                                     d.block_topo(),
                                     d.block_geom(),
                                     g->phys(),
                                     d.d_P, d.d_GP, d.d_P2, a);
        if (timed) {
            g->ev_flux.end(g->stream);
        }
        if (rc) {
            return rc;
        }
        HIPCALL(hipGetLastError());
        if (kmode != 0) {  // neighbouring blocks read the old state: the stage writes the other buffer
            std::swap(d.d_P, d.d_P2);
        }
        return 0;
    }
    if (conv || g->cfg.viscous) {
        bool timed = g->timing && g->ev_face.begin(g->stream);
        int rc = Kern<S>::face(g->stream, g->cfg.flux, g->cfg.viscous, g->cfg.space_order, g->topo(),
                               d.geom(), g->phys(), d.d_P, d.d_GP, d.d_F, &g->d_scalars->error);
        if (timed) {
            g->ev_face.end(g->stream);
        }
        if (rc) {
            return rc;
        }
    }
    bool timed = g->timing && g->ev_flux.begin(g->stream);
    Kern<S>::accum(g->stream, g->cfg.viscous, conv, kmode, g->topo(), d.geom(), g->phys(), d.d_F, d.d_P,
                   a);
    if (timed) {
        g->ev_flux.end(g->stream);
    }
    HIPCALL(hipGetLastError());
    return 0;
}

template int teal_residual_pass<double>(TealGpu *, DevData<double> &, int, double, int);
template int teal_residual_pass<float>(TealGpu *, DevData<float> &, int, double, int);

// host primitive rows (reference numbering) -> device P records, and U0 = prim2cons(P) in FP64
// (valid for the next explicit step; the time step still has to be computed)
// host rows of a device-row count: ni or nc device rows hold mesh_ni or the mesh's number of
// cells (padding rows of a batch have no host row)
static inline size_t host_rows(const TealGpu *g, int device_rows)
{
    return (size_t)device_rows - (size_t)(g->ni - g->mesh_ni);
}

template <typename S>
static int upload_primitive(TealGpu *g, DevData<S> &d, const double *primitive)
{
    size_t bytes = sizeof(double) * 5 * host_rows(g, g->nc);
    HIPCALL(hipMemcpyAsync(g->d_stage, primitive, bytes, hipMemcpyHostToDevice, g->stream));
    Kern<S>::to_device(g->stream, g->d_stage, d.d_P, g->nc, 5, 5, 5, g->d_cell_perm, 0);
    if (g->configured) {
        hipLaunchKernelGGL(k_u0_from_host, dim3(g->nblocks), dim3(BLOCK), 0, g->stream,
                           (const double *)g->d_stage, (const int *)g->d_cell_perm, g->ni,
                           g->cfg.gamma, g->d_U0);
        g->u0_valid = 1;
        g->dt_valid = 0;
    }
    HIPCALL(hipGetLastError());
    return 0;
}

// device rows (records or [width][num] arrays) -> host rows of `host_stride` doubles
template <typename S>
static int download_rows(TealGpu *g, const S *src, int num, int src_stride, int width, int soa,
                         int host_stride, double *host)
{
    Kern<S>::to_host(g->stream, src, g->d_stage, num, src_stride, width, host_stride, g->d_cell_perm,
                     soa);
    HIPCALL(hipGetLastError());
    HIPCALL(hipMemcpyAsync(host, g->d_stage, sizeof(double) * host_stride * host_rows(g, num),
                           hipMemcpyDeviceToHost, g->stream));
    HIPCALL(hipStreamSynchronize(g->stream));
    return 0;
}

static int check_ready(TealGpu *g)
{
    if (!g) {
        teal_gpu_set_error("null TealGpu");
        return -1;
    }
    if (!g->configured) {
        teal_gpu_set_error("TealGpu is not configured");
        return -1;
    }
    return 0;
}

// ghost entries the reference writes into the caller's primitive array: boundary ghosts by
// the BC pass, periodic ghosts by the exchange (src/equations/exchange.c:8-51)
static void fill_periodic_ghosts(TealGpu *g, double *buf, int len)
{
    for (int c = g->mesh->cells.off_boundary; c < g->mesh->cells.num; c++) {
        int s = g->periodic_src[c];
        if (s >= 0) {
            memcpy(&buf[(size_t)c * len], &buf[(size_t)s * len], sizeof(double) * len);
        }
    }
}

// boundary-ghost rows of the device state -> the caller's host array (the rows the reference's
// boundary pass writes, src/equations/boundary.c:5-43)
template <typename S>
static int download_ghosts(TealGpu *g, DevData<S> &d, double *primitive)
{
    if (g->nfb - g->nfi <= 0) {
        return 0;
    }
    std::vector<double> tmp((size_t)5 * g->mesh->cells.num);
    if (download_rows(g, d.d_P, g->nc, 5, 5, 0, 5, tmp.data())) return -1;
    for (int c = g->mesh_ni; c < g->mesh->cells.off_boundary; c++) {
        memcpy(&primitive[(size_t)c * 5], &tmp[(size_t)c * 5], sizeof(double) * 5);
    }
    return 0;
}

int teal_upload_state(TealGpu *g, const double *primitive)
{
    return (g->precision == TEAL_FP64) ? upload_primitive(g, g->d64, primitive)
                                       : upload_primitive(g, g->d32, primitive);
}

int teal_download_ghosts(TealGpu *g, double *primitive)
{
    return (g->precision == TEAL_FP64) ? download_ghosts(g, g->d64, primitive)
                                       : download_ghosts(g, g->d32, primitive);
}

template <typename S>
static int derivative_impl(TealGpu *g, DevData<S> &d, double *primitive, double *derivative)
{
    if (upload_primitive(g, d, primitive) || teal_riemann_clear(g)) return -1;
    if (teal_residual_pass(g, d, 0, 0.0, 0)) return -1;
    // ghost states back to the caller (boundary ghosts computed on the device)
    if (download_ghosts(g, d, primitive)) return -1;
    fill_periodic_ghosts(g, primitive, 5);
    if (download_rows(g, (const double *)g->d_R, g->ni, 5, 5, 1, 5, derivative)) return -1;
    g->u0_valid = 0;
    return teal_riemann_check(g);
}

extern "C" int teal_gpu_derivative(TealGpu *g, double *primitive, double *derivative)
{
    if (check_ready(g)) return -1;
    return (g->precision == TEAL_FP64) ? derivative_impl(g, g->d64, primitive, derivative)
                                       : derivative_impl(g, g->d32, primitive, derivative);
}

template <typename S>
static int boundary_impl(TealGpu *g, DevData<S> &d, double *primitive)
{
    if (upload_primitive(g, d, primitive)) return -1;
    int nb = g->nfb - g->nfi;
    if (nb > 0) {
        Kern<S>::boundary(g->stream, nb, g->nf, d, g->cfg.gamma);
        HIPCALL(hipGetLastError());
        if (download_ghosts(g, d, primitive)) return -1;
    }
    g->u0_valid = 0;
    return 0;
}

extern "C" int teal_gpu_boundary(TealGpu *g, double *primitive)
{
    if (check_ready(g)) return -1;
    return (g->precision == TEAL_FP64) ? boundary_impl(g, g->d64, primitive)
                                       : boundary_impl(g, g->d32, primitive);
}

template <typename S>
static int gradient_impl(TealGpu *g, DevData<S> &d, double *primitive, double *gradient, int mode)
{
    if (upload_primitive(g, d, primitive)) return -1;
    int lim = (mode == 0) ? 0 : g->cfg.limiter;
    if (mode == 1) {  // limiter on the caller's gradient
        if (lim == 0) {
            teal_gpu_set_error("invalid limiter kind (0)");
            return -1;
        }
        size_t bytes = sizeof(double) * 15 * (size_t)g->mesh->cells.num;
        HIPCALL(hipMemcpyAsync(g->d_stage, gradient, bytes, hipMemcpyHostToDevice, g->stream));
        Kern<S>::to_device(g->stream, g->d_stage, d.d_GP, g->ni, 15, 15, 20, g->d_cell_perm, 0);
    }
    Kern<S>::grad_limit(g->stream, lim, mode, g->grad_block, g->grad_stage, g->topo(), d.geom(),
                        d.d_P, d.d_GP);
    HIPCALL(hipGetLastError());
    // rows of 20: gradient (15) then psi (5); the limited gradient is G * psi (limiter.c:61-63)
    std::vector<double> GP((size_t)20 * g->mesh_ni);
    if (download_rows(g, d.d_GP, g->ni, 20, 20, 0, 20, GP.data())) return -1;
    for (int c = 0; c < g->mesh_ni; c++) {
        const double *row = &GP[(size_t)c * 20];
        for (int i = 0; i < 15; i++) {
            gradient[((size_t)c * 15) + i] = (mode == 1) ? row[i] * row[15 + (i / 3)] : row[i];
        }
    }
    if (mode == 0) {
        for (int c = g->mesh_ni; c < g->mesh->cells.num; c++) {
            memset(&gradient[(size_t)c * 15], 0, sizeof(double) * 15);
        }
        fill_periodic_ghosts(g, gradient, 15);
        fill_periodic_ghosts(g, primitive, 5);
    }
    g->u0_valid = 0;
    return 0;
}

extern "C" int teal_gpu_gradient(TealGpu *g, double *primitive, double *gradient)
{
    if (check_ready(g)) return -1;
    return (g->precision == TEAL_FP64) ? gradient_impl(g, g->d64, primitive, gradient, 0)
                                       : gradient_impl(g, g->d32, primitive, gradient, 0);
}

extern "C" int teal_gpu_limiter(TealGpu *g, const double *primitive, double *gradient)
{
    if (check_ready(g)) return -1;
    std::vector<double> p(primitive, primitive + (size_t)5 * g->mesh->cells.num);
    return (g->precision == TEAL_FP64) ? gradient_impl(g, g->d64, p.data(), gradient, 1)
                                       : gradient_impl(g, g->d32, p.data(), gradient, 1);
}

template <typename S>
int teal_timestep_pass(TealGpu *g, DevData<S> &d, int write_u0, double *out)
{
    Kern<S>::timestep_u0(g->stream, g->topo(), d.geom(), g->phys(), d.d_P, g->d_U0, g->d_dtpart,
                         write_u0);
    HIPCALL(hipGetLastError());
    if (out) {
        // one minimum over all rows (a batch handle answers for the union of its meshes)
        hipLaunchKernelGGL(k_dt_min, dim3(chunks_of(g->ngroups, DT_CHUNK), 1), dim3(DT_THREADS), 0,
                           g->stream, g->d_scalars, (const double *)g->d_dtpart,
                           (const int *)g->d_all_grp, g->d_chunk, g->d_tickets, 0, 0.0, 0.0);
        HIPCALL(hipMemcpyAsync(g->h_scalars, g->d_scalars, sizeof(Scalars), hipMemcpyDeviceToHost,
                               g->stream));
        HIPCALL(hipStreamSynchronize(g->stream));
        *out = __builtin_bit_cast(double, g->h_scalars->dtmin_bits);
    }
    return 0;
}

extern "C" double teal_gpu_timestep(TealGpu *g, const double *primitive)
{
    if (check_ready(g)) return NAN;
    std::vector<double> p(primitive, primitive + (size_t)5 * g->mesh->cells.num);
    double dt = NAN;
    int rc = (g->precision == TEAL_FP64)
                 ? (upload_primitive(g, g->d64, p.data()) || teal_timestep_pass(g, g->d64, 0, &dt))
                 : (upload_primitive(g, g->d32, p.data()) || teal_timestep_pass(g, g->d32, 0, &dt));
    g->u0_valid = 0;
    return rc ? NAN : dt;
}

// Waits of the Krylov iterations and step read-backs last tens of microseconds, where a
// blocking synchronization's wake-up would dominate: poll, with a pause between queries; after
// 200 us give the core away between queries, after 2 ms sleep 50 us between them.
int teal_event_wait(hipEvent_t ev)
{
    hipError_t e;
    const auto t0 = std::chrono::steady_clock::now();
    while ((e = hipEventQuery(ev)) == hipErrorNotReady) {
        const auto us = std::chrono::duration_cast<std::chrono::microseconds>(
                            std::chrono::steady_clock::now() - t0).count();
        if (us < 200) {
            _mm_pause();
        }
        else if (us < 2000) {
            sched_yield();
        }
        else {
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
    }
    HIPCALL(e);
    return 0;
}

int teal_stream_wait(TealGpu *g)
{
    HIPCALL(hipEventRecord(g->ev_wait, g->stream));
    return teal_event_wait(g->ev_wait);
}

// per-instance residual norms of the partials in d_partials: residual[5 * i + v]
static int read_residuals(TealGpu *g, double *residual)
{
    hipLaunchKernelGGL(k_step_end, dim3(chunks_of(g->max_inst_groups, STEP_END_CHUNK), g->ninst),
                       dim3(STEP_END_THREADS), 0, g->stream, g->d_scalars,
                       (const double *)g->d_partials, (const int *)g->d_inst_grp_off,
                       (const double *)g->d_inst_volume, g->d_chunk, g->d_tickets, 0);
    HIPCALL(hipGetLastError());
    HIPCALL(hipMemcpyAsync(g->h_scalars, g->d_scalars, sizeof(Scalars) * (size_t)g->ninst,
                           hipMemcpyDeviceToHost, g->stream));
    if (teal_stream_wait(g)) return -1;
    for (int i = 0; i < g->ninst; i++) {
        memcpy(&residual[(size_t)i * 5], g->h_scalars[i].residual, sizeof(double) * 5);
    }
    return 0;
}

int teal_residual_norm(TealGpu *g, const double *d_derivative, double *residual)
{
    const double *vol64 = (g->precision == TEAL_FP64) ? g->d64.d_vol64 : g->d32.d_vol64;
    hipLaunchKernelGGL(k_residual_partials, dim3(g->nblocks), dim3(BLOCK), 0, g->stream, d_derivative,
                       (const int *)nullptr, vol64, g->ni, g->d_partials);
    return read_residuals(g, residual);
}

extern "C" int teal_gpu_residual(TealGpu *g, const double *derivative, double *residual)
{
    if (!g) {
        teal_gpu_set_error("null TealGpu");
        return -1;
    }
    HIPCALL(hipMemcpyAsync(g->d_stage, derivative, sizeof(double) * 5 * (size_t)g->mesh_ni,
                           hipMemcpyHostToDevice, g->stream));
    Kern<double>::to_device(g->stream, g->d_stage, g->d_R, g->ni, 5, 5, 5, g->d_cell_perm, 1);
    const double *vol64 = (g->precision == TEAL_FP64) ? g->d64.d_vol64 : g->d32.d_vol64;
    hipLaunchKernelGGL(k_residual_partials, dim3(g->nblocks), dim3(BLOCK), 0, g->stream, g->d_R,
                       (const int *)nullptr, vol64, g->ni, g->d_partials);
    return read_residuals(g, residual);
}

// ------------------------------------------------------------------ residual, no allocation
// teal_gpu_residual_noalloc keeps its partial sums (one per 64 rows) in this fixed array rather
// than in memory it allocates.  In host code the name `scratch` is not a device address: the address
// of the current device's copy comes from hipGetSymbolAddress, once per device, and only
// that address is passed to kernels.  Every call on a device shares its copy; each call
// overwrites the partials it reads (nothing has to start at zero), and calls that overlap on
// one device race on it.
constexpr int NOALLOC_MAX_GROUPS = 262144;  // inner cells <= 262144 * 64 (16,777,216)
static __device__ __align__(256) double scratch[(size_t)NOALLOC_MAX_GROUPS * 5];  // 10 MiB
static_assert(NOALLOC_MAX_GROUPS <= RESIDUAL_OUT_MAX_GROUPS, "k_residual_out keeps its chunk sums in LDS");
constexpr int NOALLOC_MAX_DEVICES = 64;
static double *scratch_dev[NOALLOC_MAX_DEVICES];

static int scratch_of_current_device(double **out)
{
    int dev = -1;
    HIPCALL(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        teal_gpu_set_error("residual_noalloc: device %d beyond the %d supported", dev,
                           NOALLOC_MAX_DEVICES);
        return -1;
    }
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        HIPCALL(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        scratch_dev[dev] = (double *)p;
    }
    *out = scratch_dev[dev];
    return 0;
}

extern "C" int teal_gpu_residual_noalloc(TealGpu *g, const double *derivative, double *residual,
                                         void *stream)
{
    if (check_ready(g)) return -1;
    if (g->ninst > 1) {
        teal_gpu_set_error("residual_noalloc: not available on a batch handle");
        return -1;
    }
    if (!derivative || !residual) {
        teal_gpu_set_error("residual_noalloc: null device array");
        return -1;
    }
    if (g->ngroups > NOALLOC_MAX_GROUPS) {
        teal_gpu_set_error("residual_noalloc: %d inner cells exceed the limit of %ld", g->ni,
                           (long)NOALLOC_MAX_GROUPS * 64);
        return -1;
    }
    double *part = nullptr;
    if (scratch_of_current_device(&part)) return -1;
    hipStream_t s = stream ? (hipStream_t)stream : g->stream;
    const double *vol64 = (g->precision == TEAL_FP64) ? g->d64.d_vol64 : g->d32.d_vol64;
    hipLaunchKernelGGL(k_residual_partials, dim3(g->nblocks), dim3(BLOCK), 0, s, derivative,
                       (const int *)g->d_cell_perm, vol64, g->ni, part);
    hipLaunchKernelGGL(k_residual_out, dim3(1), dim3(STEP_END_THREADS), 0, s, (const double *)part,
                       g->ngroups, (const double *)g->d_inst_volume, residual);
    HIPCALL(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------ resident state
extern "C" int teal_gpu_upload(TealGpu *g, const double *primitive)
{
    if (check_ready(g)) return -1;
    int rc = (g->precision == TEAL_FP64) ? upload_primitive(g, g->d64, primitive)
                                         : upload_primitive(g, g->d32, primitive);
    if (rc) return rc;
    HIPCALL(hipStreamSynchronize(g->stream));
    return 0;
}

extern "C" int teal_gpu_download(TealGpu *g, double *primitive)
{
    if (check_ready(g)) return -1;
    int rc = (g->precision == TEAL_FP64)
                 ? download_rows(g, g->d64.d_P, g->nc, 5, 5, 0, 5, primitive)
                 : download_rows(g, g->d32.d_P, g->nc, 5, 5, 0, 5, primitive);
    if (rc) return rc;
    if (g->precision == TEAL_MIXED && g->u0_valid) {
        // inner states from the FP64 solution U instead of its FP32 copy
        hipLaunchKernelGGL(k_prim_from_u, dim3(g->nblocks), dim3(BLOCK), 0, g->stream,
                           (const double *)g->d_U0, (const int *)g->d_cell_perm, g->ni,
                           g->cfg.gamma, g->d_stage);
        HIPCALL(hipGetLastError());
        std::vector<double> tmp((size_t)5 * g->mesh_ni);
        HIPCALL(hipMemcpyAsync(tmp.data(), g->d_stage, sizeof(double) * 5 * (size_t)g->mesh_ni,
                               hipMemcpyDeviceToHost, g->stream));
        HIPCALL(hipStreamSynchronize(g->stream));
        for (int d = 0; d < g->ni; d++) {
            int c = g->cell_perm[d];
            if (c < 0) continue;  // padding row of a batch
            memcpy(&primitive[(size_t)c * 5], &tmp[(size_t)c * 5], sizeof(double) * 5);
        }
    }
    fill_periodic_ghosts(g, primitive, 5);
    return 0;
}

// lserk coefficients, src/simulation/advance.c:50-75 (4-digit literals, kept as written)
static const double lserk_alpha[3][6][7] = {
    {
        {0.0000, 1.0000, 0.0000, 0.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.3333, 1.0000, 0.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.1481, 0.4000, 1.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.0833, 0.2069, 0.4265, 1.0000, 0.0000, 0.0000},
        {0.0000, 0.0533, 0.1263, 0.2375, 0.4414, 1.0000, 0.0000},
        {0.0000, 0.0370, 0.0851, 0.1521, 0.2562, 0.4512, 1.0000},
    },
    {
        {0.0000, 1.0000, 0.0000, 0.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.4242, 1.0000, 0.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.1918, 0.4929, 1.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.1084, 0.2602, 0.5052, 1.0000, 0.0000, 0.0000},
        {0.0000, 0.0695, 0.1602, 0.2898, 0.5060, 1.0000, 0.0000},
        {0.0000, 0.0482, 0.1085, 0.1885, 0.3050, 0.5063, 1.0000},
    },
    {
        {0.0000, 1.0000, 0.0000, 0.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.6612, 1.0000, 0.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.2884, 0.5010, 1.0000, 0.0000, 0.0000, 0.0000},
        {0.0000, 0.1666, 0.3027, 0.5275, 1.0000, 0.0000, 0.0000},
        {0.0000, 0.1067, 0.1979, 0.3232, 0.5201, 1.0000, 0.0000},
        {0.0000, 0.0742, 0.1393, 0.2198, 0.3302, 0.5181, 1.0000},
    },
};

// the launches of one explicit step once U0 and the per-block dt minima are current: per
// instance dt, the stages, per-instance residual and time (what a graph of a step holds)
template <typename S>
static int explicit_stages(TealGpu *g, DevData<S> &d, double max_step, double courant,
                           const double *alpha, int num_stages)
{
    hipLaunchKernelGGL(k_dt_min, dim3(chunks_of(g->max_inst_groups, DT_CHUNK), g->ninst),
                       dim3(DT_THREADS), 0, g->stream, g->d_scalars, (const double *)g->d_dtpart,
                       (const int *)g->d_inst_grp_off, g->d_chunk, g->d_tickets, 1, courant,
                       max_step);
    for (int s = 0; s < num_stages; s++) {
        if (teal_residual_pass(g, d, 1, alpha[s + 1], s == num_stages - 1)) return -1;
    }
    hipLaunchKernelGGL(k_step_end, dim3(chunks_of(g->max_inst_groups, STEP_END_CHUNK), g->ninst),
                       dim3(STEP_END_THREADS), 0, g->stream, g->d_scalars,
                       (const double *)g->d_partials, (const int *)g->d_inst_grp_off,
                       (const double *)g->d_inst_volume, g->d_chunk, g->d_tickets, 1);
    HIPCALL(hipGetLastError());
    return 0;
}

// one explicit step on the device; nothing is read back
template <typename S>
static int explicit_step(TealGpu *g, DevData<S> &d, double max_step, double courant,
                         const double *alpha, int num_stages)
{
    if (!g->u0_valid || !g->dt_valid) {  // U0 = prim2cons(P) unless the upload provided it
        if (teal_timestep_pass(g, d, !g->u0_valid, nullptr)) return -1;
        g->u0_valid = 1;
        g->dt_valid = 1;
    }
    return explicit_stages(g, d, max_step, courant, alpha, num_stages);
}

// The exact Riemann solver's flag (Scalars::error) covers one entry-point call: cleared when
// the call starts, checked before it returns (the reference stops the program at the first
// vacuum or non-converged state, src/euler/riemann.c:80,179).  Only Godunov sets it.
__global__ void k_clear_error(Scalars *scalars, int num)
{
    int i = (blockIdx.x * BLOCK) + threadIdx.x;
    if (i < num) {
        scalars[i].error = 0;
    }
}

int teal_riemann_clear(TealGpu *g)
{
    if (g->cfg.flux != TEAL_FLUX_GODUNOV) {
        return 0;
    }
    hipLaunchKernelGGL(k_clear_error, dim3(grid_of(g->ninst)), dim3(BLOCK), 0, g->stream,
                       g->d_scalars, g->ninst);
    HIPCALL(hipGetLastError());
    return 0;
}

static int scalars_error(TealGpu *g)
{
    for (int i = 0; i < g->ninst; i++) {
        if (g->h_scalars[i].error == 1) {
            teal_gpu_set_error("vacuum state detected in riemann solver");
            return -1;
        }
        if (g->h_scalars[i].error == 2) {
            teal_gpu_set_error("newton solver did not converge (20)");
            return -1;
        }
    }
    return 0;
}

int teal_riemann_check(TealGpu *g)
{
    if (g->cfg.flux != TEAL_FLUX_GODUNOV) {
        return 0;
    }
    HIPCALL(hipMemcpyAsync(g->h_scalars, g->d_scalars, sizeof(Scalars) * (size_t)g->ninst,
                           hipMemcpyDeviceToHost, g->stream));
    if (teal_stream_wait(g)) return -1;
    return scalars_error(g);
}

static int read_scalars(TealGpu *g)
{
    HIPCALL(hipMemcpyAsync(g->h_scalars, g->d_scalars, sizeof(Scalars) * (size_t)g->ninst,
                           hipMemcpyDeviceToHost, g->stream));
    if (teal_stream_wait(g)) return -1;
    return scalars_error(g);
}

static int check_single(TealGpu *g, const char *who)
{
    if (check_ready(g)) return -1;
    if (g->ninst > 1) {
        teal_gpu_set_error("%s: a batch handle advances with teal_gpu_run_lserk_batch", who);
        return -1;
    }
    return 0;
}

extern "C" double teal_gpu_lserk(TealGpu *g, double *time, double *residual, double max_step,
                                 double courant, int time_order, int num_stages)
{
    if (check_single(g, "lserk")) return NAN;
    if (time_order < 1 || time_order > 3 || num_stages < 1 || num_stages > 6) {
        teal_gpu_set_error("lserk: invalid time order (%d) / stages (%d)", time_order, num_stages);
        return NAN;
    }
    const double *alpha = lserk_alpha[time_order - 1][num_stages - 1];
    if (teal_riemann_clear(g)) return NAN;
    int rc = (g->precision == TEAL_FP64)
                 ? explicit_step(g, g->d64, max_step, courant, alpha, num_stages)
                 : explicit_step(g, g->d32, max_step, courant, alpha, num_stages);
    if (rc || read_scalars(g)) return NAN;
    *time += g->h_scalars->step;
    if (residual) memcpy(residual, g->h_scalars->residual, sizeof(double) * 5);
    return g->h_scalars->step0;
}

// forward Euler (src/simulation/advance.c:11-45) is the one-stage scheme U = U0 + R*step,
// i.e. lserk with alpha = {0, 1} and U0 = prim2cons(P)
extern "C" double teal_gpu_euler(TealGpu *g, double *time, double *residual, double max_step,
                                 double courant)
{
    if (check_single(g, "euler")) return NAN;
    static const double alpha[2] = {0.0, 1.0};
    if (teal_riemann_clear(g)) return NAN;
    int rc = (g->precision == TEAL_FP64) ? explicit_step(g, g->d64, max_step, courant, alpha, 1)
                                         : explicit_step(g, g->d32, max_step, courant, alpha, 1);
    if (rc || read_scalars(g)) return NAN;
    *time += g->h_scalars->step;
    if (residual) memcpy(residual, g->h_scalars->residual, sizeof(double) * 5);
    return g->h_scalars->step0;
}

static void drop_graph(TealGpu *g)
{
    if (g->graph) {
        (void)hipGraphExecDestroy(g->graph);
        g->graph = nullptr;
    }
}

// the graph of two explicit steps with these coefficients, captured on first use (capture
// records the launches without running them; the stage buffers end where they started)
template <typename S>
static int step_graph(TealGpu *g, DevData<S> &d, double courant, const double *alpha,
                      int num_stages)
{
    if (g->graph && g->graph_key.courant == courant && g->graph_key.alpha == alpha &&
        g->graph_key.stages == num_stages && g->graph_key.P == (const void *)d.d_P) {
        return 0;
    }
    drop_graph(g);
    const void *P = d.d_P;
    HIPCALL(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int k = 0; k < 2 && rc == 0; k++) {
        rc = explicit_stages(g, d, INFINITY, courant, alpha, num_stages);
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(g->stream, &graph);
    if (rc == 0 && e == hipSuccess && (const void *)d.d_P != P) {
        teal_gpu_set_error("step graph: stage buffers did not return to their start");
        rc = -1;
    }
    if (rc == 0 && e == hipSuccess) {
        e = hipGraphInstantiate(&g->graph, graph, nullptr, nullptr, 0);
    }
    if (graph) {
        (void)hipGraphDestroy(graph);
    }
    if (rc || e != hipSuccess) {
        if (rc == 0) {
            teal_gpu_set_error("step graph capture failed: %s", hipGetErrorString(e));
        }
        g->graph = nullptr;
        return -1;
    }
    g->graph_key.courant = courant;
    g->graph_key.alpha = alpha;
    g->graph_key.stages = num_stages;
    g->graph_key.P = P;
    return 0;
}

// `steps` explicit steps without read-back: the first eagerly when U0 or the dt minima have
// to be computed, then graph replays of step pairs (when enabled), the rest eagerly
template <typename S>
static int run_steps(TealGpu *g, DevData<S> &d, int steps, double courant, const double *alpha,
                     int num_stages)
{
    int done = 0;
    if (steps > 0 && (!g->u0_valid || !g->dt_valid)) {
        if (explicit_step(g, d, INFINITY, courant, alpha, num_stages)) return -1;
        done = 1;
    }
    int pairs = (g->graphs && !g->timing) ? (steps - done) / 2 : 0;
    if (pairs > 0) {
        if (step_graph(g, d, courant, alpha, num_stages)) return -1;
        for (int k = 0; k < pairs; k++) {
            HIPCALL(hipGraphLaunch(g->graph, g->stream));
        }
        done += 2 * pairs;
    }
    for (; done < steps; done++) {
        if (explicit_step(g, d, INFINITY, courant, alpha, num_stages)) return -1;
    }
    return 0;
}

static const double *lserk_coefficients(int time_order, int num_stages)
{
    if (time_order < 1 || time_order > 3 || num_stages < 1 || num_stages > 6) {
        teal_gpu_set_error("lserk: invalid time order (%d) / stages (%d)", time_order, num_stages);
        return nullptr;
    }
    return lserk_alpha[time_order - 1][num_stages - 1];
}

extern "C" int teal_gpu_run_lserk(TealGpu *g, int steps, double *time, double courant,
                                  int time_order, int num_stages)
{
    if (check_single(g, "run_lserk")) return -1;
    const double *alpha = lserk_coefficients(time_order, num_stages);
    if (!alpha) return -1;
    hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(1), 0, g->stream, g->d_scalars,
                       (const double *)nullptr, *time);
    g->dt_valid = 0;
    int rc = (g->precision == TEAL_FP64) ? run_steps(g, g->d64, steps, courant, alpha, num_stages)
                                         : run_steps(g, g->d32, steps, courant, alpha, num_stages);
    if (rc || read_scalars(g)) return -1;
    *time = g->h_scalars->time;
    return 0;
}

// ------------------------------------------------------------------ batches
extern "C" int teal_gpu_run_lserk_batch(TealGpu *g, int steps, double *time, double *residual,
                                        double courant, int time_order, int num_stages)
{
    if (check_ready(g)) return -1;
    if (!time) {
        teal_gpu_set_error("run_lserk_batch: null time array");
        return -1;
    }
    const double *alpha = lserk_coefficients(time_order, num_stages);
    if (!alpha) return -1;
    HIPCALL(hipMemcpyAsync(g->d_stage, time, sizeof(double) * (size_t)g->ninst,
                           hipMemcpyHostToDevice, g->stream));
    hipLaunchKernelGGL(k_init_scalars, dim3(g->ninst), dim3(1), 0, g->stream, g->d_scalars,
                       (const double *)g->d_stage, 0.0);
    HIPCALL(hipGetLastError());
    g->dt_valid = 0;
    int rc = (g->precision == TEAL_FP64) ? run_steps(g, g->d64, steps, courant, alpha, num_stages)
                                         : run_steps(g, g->d32, steps, courant, alpha, num_stages);
    if (rc || read_scalars(g)) return -1;
    for (int i = 0; i < g->ninst; i++) {
        time[i] = g->h_scalars[i].time;
        if (residual) memcpy(&residual[(size_t)i * 5], g->h_scalars[i].residual, sizeof(double) * 5);
    }
    return 0;
}

// instance i's host rows <-> its rows of the union's host layout
static void batch_rows(const TealGpu *g, int i, double *uni, double *inst, int to_union)
{
    const int *r = &g->inst_rows[(size_t)6 * i];
    int local = 0;
    for (int grp = 0; grp < 3; grp++) {
        double *u = &uni[(size_t)r[2 * grp] * 5];
        double *m = &inst[(size_t)local * 5];
        size_t bytes = sizeof(double) * 5 * (size_t)r[(2 * grp) + 1];
        if (to_union) {
            memcpy(u, m, bytes);
        }
        else {
            memcpy(m, u, bytes);
        }
        local += r[(2 * grp) + 1];
    }
}

static int check_batch(TealGpu *g, const void *const *arrays, const char *who)
{
    if (check_ready(g)) return -1;
    if (g->inst_rows.empty() || !arrays) {
        teal_gpu_set_error("%s: not a batch handle, or null arrays", who);
        return -1;
    }
    for (int i = 0; i < g->ninst; i++) {
        if (!arrays[i]) {
            teal_gpu_set_error("%s: null array for instance %d", who, i);
            return -1;
        }
    }
    return 0;
}

extern "C" int teal_gpu_upload_batch(TealGpu *g, const double *const *primitive)
{
    if (check_batch(g, (const void *const *)primitive, "upload_batch")) return -1;
    std::vector<double> uni((size_t)5 * g->mesh->cells.num, 0.0);
    for (int i = 0; i < g->ninst; i++) {
        batch_rows(g, i, uni.data(), const_cast<double *>(primitive[i]), 1);
    }
    return teal_gpu_upload(g, uni.data());
}

extern "C" int teal_gpu_download_batch(TealGpu *g, double *const *primitive)
{
    if (check_batch(g, (const void *const *)primitive, "download_batch")) return -1;
    std::vector<double> uni((size_t)5 * g->mesh->cells.num, 0.0);
    if (teal_gpu_download(g, uni.data())) return -1;
    for (int i = 0; i < g->ninst; i++) {
        batch_rows(g, i, uni.data(), primitive[i], 0);
    }
    return 0;
}

extern "C" const Mesh *teal_gpu_mesh(TealGpu *g) { return g ? g->mesh : nullptr; }

extern "C" int teal_gpu_batch_size(TealGpu *g)
{
    if (!g) {
        teal_gpu_set_error("null TealGpu");
        return -1;
    }
    return g->ninst;
}

extern "C" int teal_gpu_set_graphs(TealGpu *g, int enable)
{
    if (!g) {
        teal_gpu_set_error("null TealGpu");
        return -1;
    }
    g->graphs = enable != 0;
    if (!g->graphs) {
        drop_graph(g);
    }
    return 0;
}

// ------------------------------------------------------------------ autotuning
// the state an explicit run changes, saved around trial runs
template <typename S>
struct Snapshot {
    S *P = nullptr;
    double *U0 = nullptr;
    Scalars *sc = nullptr;
    int u0_valid = 0, dt_valid = 0;

    int save(TealGpu *g, DevData<S> &d)
    {
        HIPCALL(hipMalloc((void **)&P, sizeof(S) * 5 * (size_t)g->nc));
        HIPCALL(hipMalloc((void **)&U0, sizeof(double) * 5 * (size_t)g->ni));
        HIPCALL(hipMalloc((void **)&sc, sizeof(Scalars) * (size_t)g->ninst));
        HIPCALL(hipMemcpyAsync(P, d.d_P, sizeof(S) * 5 * (size_t)g->nc, hipMemcpyDeviceToDevice,
                               g->stream));
        HIPCALL(hipMemcpyAsync(U0, g->d_U0, sizeof(double) * 5 * (size_t)g->ni,
                               hipMemcpyDeviceToDevice, g->stream));
        HIPCALL(hipMemcpyAsync(sc, g->d_scalars, sizeof(Scalars) * (size_t)g->ninst,
                               hipMemcpyDeviceToDevice, g->stream));
        u0_valid = g->u0_valid;
        dt_valid = g->dt_valid;
        return 0;
    }
    int restore(TealGpu *g, DevData<S> &d)
    {
        HIPCALL(hipMemcpyAsync(d.d_P, P, sizeof(S) * 5 * (size_t)g->nc, hipMemcpyDeviceToDevice,
                               g->stream));
        HIPCALL(hipMemcpyAsync(g->d_U0, U0, sizeof(double) * 5 * (size_t)g->ni,
                               hipMemcpyDeviceToDevice, g->stream));
        HIPCALL(hipMemcpyAsync(g->d_scalars, sc, sizeof(Scalars) * (size_t)g->ninst,
                               hipMemcpyDeviceToDevice, g->stream));
        HIPCALL(hipStreamSynchronize(g->stream));
        g->u0_valid = u0_valid;
        g->dt_valid = dt_valid;
        return 0;
    }
    ~Snapshot()
    {
        (void)hipFree(P);
        (void)hipFree(U0);
        (void)hipFree(sc);
    }
};

// times `steps` explicit steps (after one untimed step) in each launch shape the handle
// supports -- fused stage kernel or face/accumulate pair, eager or graph replay -- keeps the
// fastest, and restores the state; ms[4] receives the per-step times (NaN: not available)
template <typename S>
static int autotune(TealGpu *g, DevData<S> &d, int steps, double courant, const double *alpha,
                    int num_stages, double *ms)
{
    auto fits = [&](int layout) {
        const int keep = d.cur;
        d.cur = layout;
        const size_t lds = d.block_lds();
        d.cur = keep;
        return d.lay[layout].rows > 0 && lds > 0 && lds <= (size_t)g->lds_max;
    };
    Snapshot<S> snap;
    if (snap.save(g, d)) return -1;
    const int graphs0 = g->graphs, cur0 = d.cur, gb0 = g->grad_block, st0 = g->grad_stage;
    hipEvent_t t0, t1;
    HIPCALL(hipEventCreate(&t0));
    HIPCALL(hipEventCreate(&t1));
    // ms per step of the current configuration, from the saved state
    auto time_steps = [&](double *out) -> int {
        drop_graph(g);
        if (snap.restore(g, d) || run_steps(g, d, 1, courant, alpha, num_stages) ||
            (g->graphs && step_graph(g, d, courant, alpha, num_stages))) {
            return -1;
        }
        (void)hipEventRecord(t0, g->stream);
        int rc = run_steps(g, d, steps, courant, alpha, num_stages);
        (void)hipEventRecord(t1, g->stream);
        float t = 0;
        if (rc || hipEventSynchronize(t1) != hipSuccess || hipEventElapsedTime(&t, t0, t1) != hipSuccess) {
            return -1;
        }
        *out = t / steps;
        return 0;
    };
    double best = INFINITY;
    int best_shape = -1;
    int rc = 0;
    // stage shapes: fused with BLOCK-row blocks, fused with SMALL_BLOCK-row blocks, split;
    // each with eager launches, then with graph replay
    for (int shape = 0; shape < TEAL_AUTOTUNE_STAGE_SHAPES && rc == 0; shape++) {
        const bool fused = shape < 4;
        const int layout = (shape < 2) ? 0 : 1;
        ms[shape] = NAN;
        if (fused && !fits(layout)) continue;
        d.fused = fused;
        d.cur = fused ? layout : cur0;
        g->graphs = shape & 1;
        rc = time_steps(&ms[shape]);
        if (rc == 0 && ms[shape] < best) {
            best = ms[shape];
            best_shape = shape;
        }
    }
    if (rc == 0 && best_shape >= 0) {
        d.fused = best_shape < 4;
        d.cur = (best_shape < 2) ? 0 : (best_shape < 4) ? 1 : cur0;
        g->graphs = best_shape & 1;
        // gradient kernel workgroup sizes, direct or LDS-staged gathers, on the chosen stage
        // shape
        static const int sizes[3] = {BLOCK, 128, 64};
        constexpr int NGRAD = TEAL_AUTOTUNE_SHAPES - TEAL_AUTOTUNE_STAGE_SHAPES;
        static_assert(NGRAD == 6, "gradient shapes: 3 workgroup sizes x direct / staged");
        double best_gb = INFINITY;
        int pick = gb0, pick_st = st0;
        for (int k = 0; k < NGRAD && rc == 0; k++) {
            double *out = &ms[TEAL_AUTOTUNE_STAGE_SHAPES + k];
            *out = NAN;
            g->grad_block = sizes[k % 3];
            g->grad_stage = k / 3;
            rc = time_steps(out);
            if (rc == 0 && *out < best_gb) {
                best_gb = *out;
                pick = sizes[k % 3];
                pick_st = k / 3;
            }
        }
        g->grad_stage = (rc == 0) ? pick_st : st0;
        g->grad_block = (rc == 0) ? pick : gb0;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    drop_graph(g);
    if (rc != 0 || best_shape < 0) {
        d.cur = cur0;
        d.fused = fits(cur0);
        g->graphs = graphs0;
        g->grad_block = gb0;
        g->grad_stage = st0;
    }
    if (snap.restore(g, d)) return -1;
    return rc;
}

extern "C" int teal_gpu_autotune(TealGpu *g, int steps, double courant, int time_order,
                                 int num_stages, double *ms)
{
    if (check_ready(g)) return -1;
    const double *alpha = lserk_coefficients(time_order, num_stages);
    if (!alpha) return -1;
    if (steps < 2) {
        teal_gpu_set_error("autotune: at least 2 timed steps (%d)", steps);
        return -1;
    }
    double tmp[TEAL_AUTOTUNE_SHAPES];
    double *out = ms ? ms : tmp;
    return (g->precision == TEAL_FP64) ? autotune(g, g->d64, steps, courant, alpha, num_stages, out)
                                       : autotune(g, g->d32, steps, courant, alpha, num_stages, out);
}

// one of the autotuner's stage shapes (0 .. TEAL_AUTOTUNE_STAGE_SHAPES - 1) and a gradient
// workgroup size, set directly
template <typename S>
static int set_shape(TealGpu *g, DevData<S> &d, int shape, int grad_block, int grad_stage)
{
    const bool fused = shape < 4;
    const int layout = (shape < 2) ? 0 : 1;
    if (fused) {
        const int keep = d.cur;
        d.cur = layout;
        const size_t lds = d.block_lds();
        if (d.lay[layout].rows == 0 || lds == 0 || lds > (size_t)g->lds_max) {
            d.cur = keep;
            teal_gpu_set_error("set_shape: the %d-row block layout does not fit (%zu bytes of LDS)",
                               d.lay[layout].rows, lds);
            return -1;
        }
    }
    d.fused = fused;
    g->graphs = shape & 1;
    g->grad_block = grad_block;
    g->grad_stage = grad_stage;
    drop_graph(g);
    return 0;
}

extern "C" int teal_gpu_set_shape(TealGpu *g, int shape, int grad_block, int grad_stage)
{
    if (check_ready(g)) return -1;
    if (shape < 0 || shape >= TEAL_AUTOTUNE_STAGE_SHAPES ||
        (grad_block != 64 && grad_block != 128 && grad_block != BLOCK) ||
        (grad_stage != 0 && grad_stage != 1)) {
        teal_gpu_set_error("set_shape: invalid shape (%d), gradient workgroup (%d) or staging (%d)",
                           shape, grad_block, grad_stage);
        return -1;
    }
    return (g->precision == TEAL_FP64) ? set_shape(g, g->d64, shape, grad_block, grad_stage)
                                       : set_shape(g, g->d32, shape, grad_block, grad_stage);
}

// ------------------------------------------------------------------ timing
int EventPool::reserve(int n)
{
    while ((int)start.size() < n) {
        hipEvent_t a, b;
        HIPCALL(hipEventCreate(&a));
        HIPCALL(hipEventCreate(&b));
        start.push_back(a);
        stop.push_back(b);
    }
    used = 0;
    return 0;
}

void EventPool::release()
{
    for (size_t i = 0; i < start.size(); i++) {
        (void)hipEventDestroy(start[i]);
        (void)hipEventDestroy(stop[i]);
    }
    start.clear();
    stop.clear();
    used = 0;
}

bool EventPool::begin(hipStream_t s)
{
    if (used >= (int)start.size()) {
        return false;
    }
    return hipEventRecord(start[used], s) == hipSuccess;
}

void EventPool::end(hipStream_t s)
{
    (void)hipEventRecord(stop[used], s);
    used += 1;
}

double EventPool::total_ms()
{
    double total = 0;
    for (int i = 0; i < used; i++) {
        float ms = 0;
        if (hipEventSynchronize(stop[i]) == hipSuccess &&
            hipEventElapsedTime(&ms, start[i], stop[i]) == hipSuccess) {
            total += ms;
        }
    }
    return total;
}

// capacity > 0: time up to `capacity` launches of each timed kernel from now on (events are
// preallocated here, outside any timed region); capacity == 0: stop timing
extern "C" int teal_gpu_timing(TealGpu *g, int capacity)
{
    if (!g) return -1;
    g->timing = capacity > 0;
    if (capacity > 0) {
        if (g->ev_grad.reserve(capacity) || g->ev_face.reserve(capacity) ||
            g->ev_flux.reserve(capacity)) {
            return -1;
        }
    }
    return 0;
}

// mean duration (ms) of the timed launches of kernel 0 (gradient + limiter), 1 (face fluxes)
// or 2 (accumulation + RK update)
extern "C" double teal_gpu_timing_ms(TealGpu *g, int kernel, int *launches)
{
    EventPool *pool = (kernel == 0) ? &g->ev_grad : (kernel == 1) ? &g->ev_face : &g->ev_flux;
    if (launches) *launches = pool->used;
    return pool->used ? pool->total_ms() / pool->used : 0.0;
}

extern "C" int teal_gpu_info(TealGpu *g, long *counts)
{
    if (!g) return -1;
    counts[0] = g->ni;
    counts[1] = g->nc;
    counts[2] = g->nf;
    counts[3] = g->nfi;
    counts[4] = g->nfb;
    counts[5] = W;
    const bool fused = (g->precision == TEAL_FP64) ? g->d64.fused : g->d32.fused;
    const bool centre = (g->precision == TEAL_FP64) ? g->d64.lay[g->d64.cur].centre_form
                                                    : g->d32.lay[g->d32.cur].centre_form;
    const int rows = (g->precision == TEAL_FP64) ? g->d64.block_rows() : g->d32.block_rows();
    counts[6] = fused;
    counts[7] = fused && centre;
    counts[8] = fused ? rows : 0;
    counts[9] = g->grad_block;
    counts[11] = g->grad_stage;
    counts[10] = teal_krylov_column_ept(g);
    counts[12] = g->precision == TEAL_FP64 && fused && g->d64.lay[g->d64.cur].tangents;
    return 0;
}

// ------------------------------------------------------------------ FP64 kernel instances
template struct Kern<double>;
template int teal_timestep_pass<double>(TealGpu *, DevData<double> &, int, double *);
template int teal_timestep_pass<float>(TealGpu *, DevData<float> &, int, double *);
