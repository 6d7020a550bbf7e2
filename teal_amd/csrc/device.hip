// device.hip -- GPU state, mesh upload and the C-ABI of the residual path (teal_gpu.h).
// Compiled with -ffp-contract=off: the FP64 kernels instantiated here are bit-compatible
// with the reference.  The FP32 (mixed) kernels live in mixed.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

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
__global__ void k_init_scalars(Scalars *sc, double time)
{
    sc->dtmin_bits = 0x7ff0000000000000ULL;
    sc->step0 = 0;
    sc->step = 0;
    sc->time = time;
    for (int v = 0; v < 5; v++) sc->residual[v] = 0;
    sc->error = 0;
}

__global__ void k_reset_dt(Scalars *sc) { sc->dtmin_bits = 0x7ff0000000000000ULL; }

// step0 = courant * min dt (advance.c:92-93); the accumulator is reset for the next step
__global__ void k_step_begin(Scalars *sc, double courant, double max_step)
{
    double dtmin = __longlong_as_double((long long)sc->dtmin_bits);
    sc->step0 = courant * dtmin;
    sc->step = fmin(sc->step0, max_step);
    sc->dtmin_bits = 0x7ff0000000000000ULL;
}

// residual = sqrt(sum V R^2 / sum V) from per-block partials in a fixed order
// (src/equations/residual.c:8-33), then time += step (advance.c:111)
__global__ void k_step_end(Scalars *sc, const double *partials, int nblocks, double sum_volume,
                           int advance_time)
{
    __shared__ double red[BLOCK][5];
    double s[5] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += BLOCK) {
        for (int v = 0; v < 5; v++) s[v] += partials[(b * 5) + v];
    }
    for (int v = 0; v < 5; v++) red[threadIdx.x][v] = s[v];
    __syncthreads();
    for (int h = BLOCK / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            for (int v = 0; v < 5; v++) red[threadIdx.x][v] += red[threadIdx.x + h][v];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int v = 0; v < 5; v++) sc->residual[v] = sqrt(red[0][v] / sum_volume);
        if (advance_time) {
            sc->time += sc->step;
        }
    }
}

// residual partials of a stored derivative [5][ni]
__global__ void k_residual_partials(const double *__restrict__ R, const double *__restrict__ vol64,
                                    int ni, double *partials)
{
    __shared__ double red[BLOCK / 64][5];
    int c = (blockIdx.x * BLOCK) + threadIdx.x;
    double s[5] = {0, 0, 0, 0, 0};
    if (c < ni) {
        for (int v = 0; v < 5; v++) {
            double r = R[(v * ni) + c];
            s[v] = vol64[c] * (r * r);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int v = 0; v < 5; v++) s[v] += __shfl_xor(s[v], off);
    }
    if ((threadIdx.x & 63) == 0) {
        for (int v = 0; v < 5; v++) red[threadIdx.x >> 6][v] = s[v];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 0; v < 5; v++) {
            double b = red[0][v];
            for (int w = 1; w < BLOCK / 64; w++) b += red[w][v];
            partials[(blockIdx.x * 5) + v] = b;
        }
    }
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

// the ghost block of neighbour entry e receives the send list of the entry whose send tag is
// e's receive tag (src/equations/exchange.c:19-33, src/mesh/create.c:409-419)
static std::vector<int> periodic_sources(const Mesh *m)
{
    std::vector<int> src(m->cells.num, -1);
    const MeshNeighbors *nb = &m->neighbors;
    for (int e = 0; e < nb->num; e++) {
        int s = -1;
        for (int k = 0; k < nb->num; k++) {
            if (nb->tag[k][1] == nb->tag[e][0]) {
                s = k;
            }
        }
        if (s < 0) {
            continue;
        }
        for (int j = nb->recv_off[e]; j < nb->recv_off[e + 1]; j++) {
            src[j] = nb->send.idx[nb->send.off[s] + (j - nb->recv_off[e])];
        }
    }
    return src;
}

template <typename S>
int DevData<S>::build(const Mesh *m, TealGpu *gpu)
{
    const int ni = m->cells.num_inner;
    const int nc = m->cells.num;
    const int nf = m->faces.num;
    const int nfi = m->faces.num_inner;
    const int nfb = m->faces.off_boundary;
    std::vector<int> src = periodic_sources(m);
    auto resolve = [&](int cell) {
        if (cell >= m->cells.off_boundary && src[cell] >= 0) {
            return src[cell];
        }
        return cell;
    };

    // incidences sorted by face index (face order == accumulation order)
    std::vector<int> cnt(ni, 0);
    std::vector<int> inc((size_t)W * ni, -1), nbr((size_t)W * ni, 0);
    std::vector<double> w((size_t)3 * W * ni, 0.0), o((size_t)3 * W * ni, 0.0);
    for (int f = 0; f < nf; f++) {
        int l = m->faces.cell_idx[f].left;
        int r = m->faces.cell_idx[f].right;
        for (int side = 0; side < 2; side++) {
            int c = side ? r : l;
            if (side == 1 && f >= nfi) {
                break;
            }
            int k = cnt[c]++;
            if (k >= W) {
                teal_gpu_set_error("cell %d has more than %d faces", c, W);
                return -1;
            }
            size_t idx = ((size_t)k * ni) + c;
            inc[idx] = (f << 1) | side;
            nbr[idx] = side ? l : resolve(r);
            Vector wv = side ? m->faces.weight[f].right : m->faces.weight[f].left;
            Vector ov = side ? m->faces.offset[f].right : m->faces.offset[f].left;
            double wa[3] = {wv.x, wv.y, wv.z};
            double oa[3] = {ov.x, ov.y, ov.z};
            for (int d = 0; d < 3; d++) {
                w[((((size_t)d * W) + k) * ni) + c] = wa[d];
                o[((((size_t)d * W) + k) * ni) + c] = oa[d];
            }
        }
    }
    std::vector<double> vol(ni), proj((size_t)3 * ni);
    for (int c = 0; c < ni; c++) {
        vol[c] = m->cells.volume[c];
        proj[c] = m->cells.projection[c].x;
        proj[ni + c] = m->cells.projection[c].y;
        proj[(2 * ni) + c] = m->cells.projection[c].z;
    }
    std::vector<double> area(nf), basis((size_t)9 * nf), offl((size_t)3 * nf), offr((size_t)3 * nf),
        corr((size_t)4 * nf);
    for (int f = 0; f < nf; f++) {
        area[f] = m->faces.area[f];
        const double *b = &m->faces.basis[f].x.x;
        for (int k = 0; k < 9; k++) basis[((size_t)k * nf) + f] = b[k];
        const double *ol = &m->faces.offset[f].left.x;
        const double *orr = &m->faces.offset[f].right.x;
        for (int d = 0; d < 3; d++) {
            offl[((size_t)d * nf) + f] = ol[d];
            offr[((size_t)d * nf) + f] = orr[d];
        }
        corr[f] = m->faces.correction[f].unit.x;
        corr[nf + f] = m->faces.correction[f].unit.y;
        corr[(2 * (size_t)nf) + f] = m->faces.correction[f].unit.z;
        corr[(3 * (size_t)nf) + f] = m->faces.correction[f].norm;
    }
    if (upload(&d_inc, inc) || upload(&d_nbr, nbr) || upload_conv(&d_w, w) ||
        upload_conv(&d_o, o) || upload_conv(&d_vol, vol) || upload_conv(&d_proj, proj) ||
        upload_conv(&d_area, area) || upload_conv(&d_basis, basis) ||
        upload_conv(&d_offl, offl) || upload_conv(&d_offr, offr) || upload_conv(&d_corr, corr) ||
        upload(&d_vol64, vol)) {
        return -1;
    }
    // boundary faces
    int nb = nfb - nfi;
    std::vector<int> bl(nb), br(nb), bf(nb);
    for (int i = 0; i < nb; i++) {
        bf[i] = nfi + i;
        bl[i] = m->faces.cell_idx[nfi + i].left;
        br[i] = m->faces.cell_idx[nfi + i].right;
    }
    if (upload(&d_bleft, bl) || upload(&d_bright, br) || upload(&d_bface, bf)) {
        return -1;
    }
    HIPCALL(hipMalloc((void **)&d_bkind, sizeof(int) * (nb + 1)));
    HIPCALL(hipMalloc((void **)&d_bctx, sizeof(S) * 5 * (nb + 1)));
    HIPCALL(hipMemset(d_bkind, 0, sizeof(int) * (nb + 1)));
    HIPCALL(hipMemset(d_bctx, 0, sizeof(S) * 5 * (nb + 1)));
    // state
    for (int i = 0; i < 2; i++) {
        HIPCALL(hipMalloc((void **)&d_P[i], sizeof(S) * 5 * (size_t)nc));
        HIPCALL(hipMemset(d_P[i], 0, sizeof(S) * 5 * (size_t)nc));
    }
    HIPCALL(hipMalloc((void **)&d_G, sizeof(S) * 15 * (size_t)ni));
    HIPCALL(hipMalloc((void **)&d_psi, sizeof(S) * 5 * (size_t)ni));
    HIPCALL(hipMemset(d_G, 0, sizeof(S) * 15 * (size_t)ni));
    HIPCALL(hipMemset(d_psi, 0, sizeof(S) * 5 * (size_t)ni));
    HIPCALL(hipMalloc((void **)&d_eps2, sizeof(S) * (size_t)ni));
    HIPCALL(hipMemset(d_eps2, 0, sizeof(S) * (size_t)ni));
    (void)gpu;
    return 0;
}

template <typename S>
void DevData<S>::release()
{
    void *ptrs[] = {d_inc, d_nbr, d_w, d_o, d_vol, d_proj, d_area, d_basis, d_offl, d_offr,
                    d_corr, d_vol64, d_bleft, d_bright, d_bface, d_bkind, d_bctx, d_P[0],
                    d_P[1], d_G, d_psi, d_eps2};
    for (void *p : ptrs) {
        if (p) {
            (void)hipFree(p);
        }
    }
}

template <typename S>
Geom<S> DevData<S>::geom() const
{
    return Geom<S>{d_vol, d_proj, d_eps2, d_w, d_o, d_area, d_basis, d_offl, d_offr, d_corr};
}

template struct DevData<double>;
template struct DevData<float>;


// ------------------------------------------------------------------ TealGpu
static int grid_of(int n) { return (n + BLOCK - 1) / BLOCK; }

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
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        teal_gpu_set_error("teal_gpu_create: no HIP device is available");
        return nullptr;
    }
    TealGpu *g = new TealGpu();
    g->mesh = mesh;
    g->precision = precision;
    g->ni = mesh->cells.num_inner;
    g->nc = mesh->cells.num;
    g->nf = mesh->faces.num;
    g->nfi = mesh->faces.num_inner;
    g->nfb = mesh->faces.off_boundary;
    g->sum_volume = mesh->cells.sum_volume;
    g->periodic_src = periodic_sources(mesh);
    int rc = (precision == TEAL_FP64) ? g->d64.build(mesh, g) : g->d32.build(mesh, g);
    if (rc == 0 && hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
        teal_gpu_set_error("teal_gpu_create: hipStreamCreate failed");
        rc = -1;
    }
    size_t stage = (size_t)15 * (g->nc > g->ni ? g->nc : g->ni);
    g->nblocks = grid_of(g->ni);
    if (rc == 0 &&
        (hipMalloc((void **)&g->d_stage, sizeof(double) * stage) != hipSuccess ||
         hipMalloc((void **)&g->d_U0, sizeof(double) * 5 * (size_t)g->ni) != hipSuccess ||
         hipMalloc((void **)&g->d_R, sizeof(double) * 5 * (size_t)g->ni) != hipSuccess ||
         hipMalloc((void **)&g->d_partials, sizeof(double) * 5 * (size_t)g->nblocks) != hipSuccess ||
         hipMalloc((void **)&g->d_scalars, sizeof(Scalars)) != hipSuccess ||
         hipHostMalloc((void **)&g->h_scalars, sizeof(Scalars), hipHostMallocDefault) != hipSuccess)) {
        teal_gpu_set_error("teal_gpu_create: device allocation failed");
        rc = -1;
    }
    if (rc == 0) {
        hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(1), 0, g->stream, g->d_scalars, 0.0);
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
    void *ptrs[] = {g->d_stage, g->d_U0, g->d_R, g->d_partials, g->d_scalars};
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
    teal_krylov_release(g);
    if (g->stream) {
        (void)hipStreamDestroy(g->stream);
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
        for (int c = 0; c < g->ni; c++) {  // src/equations/limiter.c:8-21
            double v = cfg->coefficient * cbrt(m->cells.volume[c]);
            e2[c] = cfg->eps2 ? (S)cfg->eps2[c] : (S)(v * v * v);
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
    g->configured = 1;
    g->u0_valid = 0;
    return (g->precision == TEAL_FP64) ? configure_impl(g, g->d64, cfg) : configure_impl(g, g->d32, cfg);
}

// ------------------------------------------------------------------ kernel launches
Topo TealGpu::topo() const
{
    const int *inc = (precision == TEAL_FP64) ? d64.d_inc : d32.d_inc;
    const int *nbr = (precision == TEAL_FP64) ? d64.d_nbr : d32.d_nbr;
    return Topo{ni, nc, nf, nfi, nfb, inc, nbr};
}

Phys TealGpu::phys() const { return Phys{cfg.gamma, cfg.mu, cfg.prandtl, cfg.navierstokes}; }

// one residual evaluation on state P[cur]: ghosts, gradient + limiter, fluxes.
// mode 0 writes R; mode 1 is an RK stage writing P[1-cur] (last: + U0, dt, residual)
template <typename S>
int teal_residual_pass(TealGpu *g, DevData<S> &d, int cur, int mode, double alpha, int last)
{
    const int nb = g->nfb - g->nfi;
    if (nb > 0) {
        Kern<S>::boundary(g->stream, nb, g->nc, g->nf, d, d.d_P[cur], g->cfg.gamma);
    }
    if (g->cfg.space_order == 2) {
        bool timed = g->timing && g->ev_grad.begin(g->stream);
        Kern<S>::grad_limit(g->stream, g->cfg.limiter, 0, g->topo(), d.geom(), d.d_P[cur], d.d_G,
                            d.d_psi);
        if (timed) {
            g->ev_grad.end(g->stream);
        }
    }
    StageArgs a;
    a.R = g->d_R;
    a.U0 = g->d_U0;
    a.U0_next = g->d_U0;
    a.alpha = alpha;
    a.scalars = g->d_scalars;
    a.partials = g->d_partials;
    a.keep_u = (g->precision == TEAL_MIXED);
    int kmode = (mode == 0) ? 0 : (last ? 2 : 1);
    bool timed = g->timing && g->ev_flux.begin(g->stream);
    int rc = Kern<S>::flux(g->stream, g->cfg.flux, g->cfg.viscous, g->cfg.space_order, kmode,
                           g->topo(), d.geom(), g->phys(), d.d_P[cur], d.d_G, d.d_psi,
                           d.d_P[1 - cur], a);
    if (timed) {
        g->ev_flux.end(g->stream);
    }
    if (rc) {
        return rc;
    }
    HIPCALL(hipGetLastError());
    return 0;
}

template int teal_residual_pass<double>(TealGpu *, DevData<double> &, int, int, double, int);
template int teal_residual_pass<float>(TealGpu *, DevData<float> &, int, int, double, int);

template <typename S>
static int upload_primitive(TealGpu *g, DevData<S> &d, const double *primitive)
{
    size_t bytes = sizeof(double) * 5 * (size_t)g->nc;
    HIPCALL(hipMemcpyAsync(g->d_stage, primitive, bytes, hipMemcpyHostToDevice, g->stream));
    Kern<S>::aos_to_soa(g->stream, g->d_stage, d.d_P[g->cur], g->nc, 5, 5, g->nc);
    HIPCALL(hipGetLastError());
    return 0;
}

template <typename S>
static int download_soa(TealGpu *g, const S *src, int num, int stride, int width, int src_stride,
                        double *host)
{
    Kern<S>::soa_to_aos(g->stream, src, g->d_stage, num, stride, width, src_stride);
    HIPCALL(hipGetLastError());
    HIPCALL(hipMemcpyAsync(host, g->d_stage, sizeof(double) * stride * (size_t)num,
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
    for (int c = g->mesh->cells.off_boundary; c < g->nc; c++) {
        int s = g->periodic_src[c];
        if (s >= 0) {
            memcpy(&buf[(size_t)c * len], &buf[(size_t)s * len], sizeof(double) * len);
        }
    }
}

template <typename S>
static int derivative_impl(TealGpu *g, DevData<S> &d, double *primitive, double *derivative)
{
    if (upload_primitive(g, d, primitive)) return -1;
    if (teal_residual_pass(g, d, g->cur, 0, 0.0, 0)) return -1;
    // ghost states back to the caller (boundary ghosts computed on the device)
    int nb = g->nfb - g->nfi;
    if (nb > 0) {
        std::vector<double> tmp((size_t)5 * g->nc);
        if (download_soa(g, d.d_P[g->cur], g->nc, 5, 5, g->nc, tmp.data())) return -1;
        for (int c = g->ni; c < g->mesh->cells.off_boundary; c++) {
            memcpy(&primitive[(size_t)c * 5], &tmp[(size_t)c * 5], sizeof(double) * 5);
        }
    }
    fill_periodic_ghosts(g, primitive, 5);
    Kern<double>::soa_to_aos(g->stream, g->d_R, g->d_stage, g->ni, 5, 5, g->ni);
    HIPCALL(hipGetLastError());
    HIPCALL(hipMemcpyAsync(derivative, g->d_stage, sizeof(double) * 5 * (size_t)g->ni,
                           hipMemcpyDeviceToHost, g->stream));
    HIPCALL(hipStreamSynchronize(g->stream));
    g->u0_valid = 0;
    return 0;
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
        Kern<S>::boundary(g->stream, nb, g->nc, g->nf, d, d.d_P[g->cur], g->cfg.gamma);
        HIPCALL(hipGetLastError());
        std::vector<double> tmp((size_t)5 * g->nc);
        if (download_soa(g, d.d_P[g->cur], g->nc, 5, 5, g->nc, tmp.data())) return -1;
        for (int c = g->ni; c < g->mesh->cells.off_boundary; c++) {
            memcpy(&primitive[(size_t)c * 5], &tmp[(size_t)c * 5], sizeof(double) * 5);
        }
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
        size_t bytes = sizeof(double) * 15 * (size_t)g->nc;
        HIPCALL(hipMemcpyAsync(g->d_stage, gradient, bytes, hipMemcpyHostToDevice, g->stream));
        Kern<S>::aos_to_soa(g->stream, g->d_stage, d.d_G, g->ni, 15, 15, g->ni);
        if (lim == 0) {
            teal_gpu_set_error("invalid limiter kind (0)");
            return -1;
        }
    }
    Kern<S>::grad_limit(g->stream, lim, mode, g->topo(), d.geom(), d.d_P[g->cur], d.d_G, d.d_psi);
    HIPCALL(hipGetLastError());
    // limited gradient = G * psi (src/equations/limiter.c:61-63)
    std::vector<double> G((size_t)15 * g->ni), psi((size_t)5 * g->ni);
    if (download_soa(g, d.d_G, g->ni, 15, 15, g->ni, G.data())) return -1;
    if (mode == 1 && download_soa(g, d.d_psi, g->ni, 5, 5, g->ni, psi.data())) return -1;
    for (int c = 0; c < g->ni; c++) {
        for (int i = 0; i < 15; i++) {
            double v = G[((size_t)c * 15) + i];
            gradient[((size_t)c * 15) + i] = (mode == 1) ? v * psi[((size_t)c * 5) + (i / 3)] : v;
        }
    }
    if (mode == 0) {
        for (int c = g->ni; c < g->nc; c++) {
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
    std::vector<double> p(primitive, primitive + (size_t)5 * g->nc);
    return (g->precision == TEAL_FP64) ? gradient_impl(g, g->d64, p.data(), gradient, 1)
                                       : gradient_impl(g, g->d32, p.data(), gradient, 1);
}

template <typename S>
int teal_timestep_pass(TealGpu *g, DevData<S> &d, int write_u0, double *out)
{
    hipLaunchKernelGGL(k_reset_dt, dim3(1), dim3(1), 0, g->stream, g->d_scalars);
    Kern<S>::timestep_u0(g->stream, g->topo(), d.geom(), g->phys(), d.d_P[g->cur], g->d_U0,
                         g->d_scalars, write_u0);
    HIPCALL(hipGetLastError());
    if (out) {
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
    std::vector<double> p(primitive, primitive + (size_t)5 * g->nc);
    double dt = NAN;
    int rc = (g->precision == TEAL_FP64)
                 ? (upload_primitive(g, g->d64, p.data()) || teal_timestep_pass(g, g->d64, 0, &dt))
                 : (upload_primitive(g, g->d32, p.data()) || teal_timestep_pass(g, g->d32, 0, &dt));
    g->u0_valid = 0;
    return rc ? NAN : dt;
}

extern "C" int teal_gpu_residual(TealGpu *g, const double *derivative, double *residual)
{
    if (!g) {
        teal_gpu_set_error("null TealGpu");
        return -1;
    }
    HIPCALL(hipMemcpyAsync(g->d_stage, derivative, sizeof(double) * 5 * (size_t)g->ni,
                           hipMemcpyHostToDevice, g->stream));
    Kern<double>::aos_to_soa(g->stream, g->d_stage, g->d_R, g->ni, 5, 5, g->ni);
    const double *vol64 = (g->precision == TEAL_FP64) ? g->d64.d_vol64 : g->d32.d_vol64;
    hipLaunchKernelGGL(k_residual_partials, dim3(g->nblocks), dim3(BLOCK), 0, g->stream, g->d_R,
                       vol64, g->ni, g->d_partials);
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(BLOCK), 0, g->stream, g->d_scalars,
                       (const double *)g->d_partials, g->nblocks, g->sum_volume, 0);
    HIPCALL(hipGetLastError());
    HIPCALL(hipMemcpyAsync(g->h_scalars, g->d_scalars, sizeof(Scalars), hipMemcpyDeviceToHost,
                           g->stream));
    HIPCALL(hipStreamSynchronize(g->stream));
    memcpy(residual, g->h_scalars->residual, sizeof(double) * 5);
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
    g->u0_valid = 0;
    return 0;
}

extern "C" int teal_gpu_download(TealGpu *g, double *primitive)
{
    if (check_ready(g)) return -1;
    int rc = (g->precision == TEAL_FP64)
                 ? download_soa(g, g->d64.d_P[g->cur], g->nc, 5, 5, g->nc, primitive)
                 : download_soa(g, g->d32.d_P[g->cur], g->nc, 5, 5, g->nc, primitive);
    if (rc) return rc;
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

// one explicit step on the device; nothing is read back
template <typename S>
static int explicit_step(TealGpu *g, DevData<S> &d, double max_step, double courant,
                         const double *alpha, int num_stages)
{
    if (!g->u0_valid) {
        if (teal_timestep_pass(g, d, 1, nullptr)) return -1;
        g->u0_valid = 1;
    }
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, g->stream, g->d_scalars, courant,
                       max_step);
    for (int s = 0; s < num_stages; s++) {
        if (teal_residual_pass(g, d, g->cur, 1, alpha[s + 1], s == num_stages - 1)) return -1;
        g->cur = 1 - g->cur;
    }
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(BLOCK), 0, g->stream, g->d_scalars,
                       (const double *)g->d_partials, g->nblocks, g->sum_volume, 1);
    HIPCALL(hipGetLastError());
    return 0;
}

static int read_scalars(TealGpu *g)
{
    HIPCALL(hipMemcpyAsync(g->h_scalars, g->d_scalars, sizeof(Scalars), hipMemcpyDeviceToHost,
                           g->stream));
    HIPCALL(hipStreamSynchronize(g->stream));
    if (g->h_scalars->error == 1) {
        teal_gpu_set_error("vacuum state detected in riemann solver");
        return -1;
    }
    if (g->h_scalars->error == 2) {
        teal_gpu_set_error("newton solver did not converge (20)");
        return -1;
    }
    return 0;
}

extern "C" double teal_gpu_lserk(TealGpu *g, double *time, double *residual, double max_step,
                                 double courant, int time_order, int num_stages)
{
    if (check_ready(g)) return NAN;
    if (time_order < 1 || time_order > 3 || num_stages < 1 || num_stages > 6) {
        teal_gpu_set_error("lserk: invalid time order (%d) / stages (%d)", time_order, num_stages);
        return NAN;
    }
    const double *alpha = lserk_alpha[time_order - 1][num_stages - 1];
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
    if (check_ready(g)) return NAN;
    static const double alpha[2] = {0.0, 1.0};
    int rc = (g->precision == TEAL_FP64) ? explicit_step(g, g->d64, max_step, courant, alpha, 1)
                                         : explicit_step(g, g->d32, max_step, courant, alpha, 1);
    if (rc || read_scalars(g)) return NAN;
    *time += g->h_scalars->step;
    if (residual) memcpy(residual, g->h_scalars->residual, sizeof(double) * 5);
    return g->h_scalars->step0;
}

extern "C" int teal_gpu_run_lserk(TealGpu *g, int steps, double *time, double courant,
                                  int time_order, int num_stages)
{
    if (check_ready(g)) return -1;
    if (time_order < 1 || time_order > 3 || num_stages < 1 || num_stages > 6) {
        teal_gpu_set_error("lserk: invalid time order (%d) / stages (%d)", time_order, num_stages);
        return -1;
    }
    const double *alpha = lserk_alpha[time_order - 1][num_stages - 1];
    hipLaunchKernelGGL(k_init_scalars, dim3(1), dim3(1), 0, g->stream, g->d_scalars, *time);
    g->u0_valid = 0;
    for (int i = 0; i < steps; i++) {
        int rc = (g->precision == TEAL_FP64)
                     ? explicit_step(g, g->d64, INFINITY, courant, alpha, num_stages)
                     : explicit_step(g, g->d32, INFINITY, courant, alpha, num_stages);
        if (rc) return -1;
    }
    if (read_scalars(g)) return -1;
    *time = g->h_scalars->time;
    return 0;
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
        if (g->ev_flux.reserve(capacity) || g->ev_grad.reserve(capacity)) return -1;
    }
    return 0;
}

extern "C" double teal_gpu_timing_flux_ms(TealGpu *g, int *launches)
{
    if (launches) *launches = g->ev_flux.used;
    return g->ev_flux.used ? g->ev_flux.total_ms() / g->ev_flux.used : 0.0;
}

extern "C" double teal_gpu_timing_grad_ms(TealGpu *g, int *launches)
{
    if (launches) *launches = g->ev_grad.used;
    return g->ev_grad.used ? g->ev_grad.total_ms() / g->ev_grad.used : 0.0;
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
    return 0;
}

// ------------------------------------------------------------------ FP64 kernel instances
template struct Kern<double>;
template int teal_timestep_pass<double>(TealGpu *, DevData<double> &, int, double *);
template int teal_timestep_pass<float>(TealGpu *, DevData<float> &, int, double *);
