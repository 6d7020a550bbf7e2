// kernels.h -- HIP kernels of the residual path, templated on the storage/arithmetic type S
// of states, gradients and geometry (double: FP64 path; float: mixed path).
//
// Device layout (all structure-of-arrays, one array per component):
//   P   [5][num_cells]   primitive {rho,u,v,w,p}; boundary-ghost slots hold BC states,
//                        periodic ghosts are never stored: they resolve to their source cell
//   G   [15][num_inner]  LSQ gradient, component (var*3 + xyz)
//   psi [5][num_inner]   limiter factor (1 when no limiter)
//   U0  [5][num_inner]   FP64 conserved state at the start of the RK step
//   incidences, slot-major [W][num_inner]: inc = face<<1 | side (side 1: cell is the
//   face's right cell), sorted by face index so the per-cell accumulation order equals
//   the reference's face-order scatter (src/equations/derivative.c:116-224).
//   per-incidence geometry [3][W][num_inner]: LSQ weight and face offset of that side
//   face geometry [..][num_faces]: area, basis (9), offset left/right (3+3), correction (4)
//
// Kernels (one thread per inner cell, owner-gather, no atomics on state):
//   k_boundary      ghost states of boundary faces (src/equations/boundary.c:5-43)
//   k_grad_limit    LSQ gradient (gradient.c:7-58) fused with the limiter (limiter.c:23-123)
//   k_flux          viscous then convective face fluxes in face order, finalize
//                   (derivative.c:91-267) and, in RK mode, the stage update + cons2prim
//                   (src/simulation/advance.c:101-109); the last stage also produces the
//                   next step's U0, the dt minimum and residual partial sums
#pragma once

#include <hip/hip_runtime.h>

#include "physics.h"

namespace fv {

constexpr int W = 6;        // MAX_CELL_FACES (src/mesh.h:8-12)
constexpr int BLOCK = 256;  // 4 wavefronts of 64
constexpr int NUM_XCD = 8;

struct Topo {
    int ni, nc, nf, nfi, nfb;
    const int *__restrict__ inc;  // [W][ni]
    const int *__restrict__ nbr;  // [W][ni] resolved neighbour cell
};

template <typename S>
struct Geom {
    const S *__restrict__ vol;    // [ni]
    const S *__restrict__ proj;   // [3][ni]
    const S *__restrict__ eps2;   // [ni] (VK) or null
    const S *__restrict__ w;      // [3][W][ni]
    const S *__restrict__ o;      // [3][W][ni]
    const S *__restrict__ area;   // [nf]
    const S *__restrict__ basis;  // [9][nf]
    const S *__restrict__ offl;   // [3][nf]
    const S *__restrict__ offr;   // [3][nf]
    const S *__restrict__ corr;   // [4][nf] unit xyz, norm
};

struct Phys {
    double gamma, mu, prandtl;
    int navierstokes;
};

// Device scalars of the resident time loop.
struct Scalars {
    unsigned long long dtmin_bits;  // min over cells of the local time step (bits of a double)
    double step0;                   // courant * dtmin
    double step;                    // min(step0, max_step)
    double time;
    double residual[5];
    int error;                      // 1: vacuum in the exact Riemann solver, 2: no convergence
};

// XCD-aware block order: the hardware deals workgroups round-robin to the 8 XCDs; this
// bijection gives every XCD one contiguous range of cells so face neighbours (close in
// the BFS numbering) share that XCD's L2.
__device__ inline int xcd_block()
{
    int nb = gridDim.x;
    int b = blockIdx.x;
    int q = nb / NUM_XCD;
    int r = nb % NUM_XCD;
    int xcd = b % NUM_XCD;
    int idx = b / NUM_XCD;
    return (xcd < r) ? (xcd * (q + 1)) + idx : (r * (q + 1)) + ((xcd - r) * q) + idx;
}

// Face-contribution accumulator.  FP64: plain sequential sums (the reference's order);
// FP32: compensated (TwoSum) accumulation, read out in FP64.
template <typename S>
struct Acc;

template <>
struct Acc<double> {
    double v[5];
    __device__ void zero()
    {
        for (int i = 0; i < 5; i++) v[i] = 0;
    }
    __device__ void add(int i, double x) { v[i] = v[i] + x; }
    __device__ void sub(int i, double x) { v[i] = v[i] - x; }
    __device__ double get(int i) const { return v[i]; }
};

template <>
struct Acc<float> {
    float hi[5], lo[5];
    __device__ void zero()
    {
        for (int i = 0; i < 5; i++) hi[i] = lo[i] = 0;
    }
    __device__ void add(int i, float x)
    {
        float s = hi[i] + x;
        float bp = s - hi[i];
        float err = (hi[i] - (s - bp)) + (x - bp);
        hi[i] = s;
        lo[i] += err;
    }
    __device__ void sub(int i, float x) { add(i, -x); }
    __device__ double get(int i) const { return (double)hi[i] + (double)lo[i]; }
};

// ---------------------------------------------------------------- boundary ghosts
template <typename S>
__global__ __launch_bounds__(BLOCK) void k_boundary(int num, int nc, int nf,
                                                    const int *__restrict__ bleft,
                                                    const int *__restrict__ bright,
                                                    const int *__restrict__ bface,
                                                    const int *__restrict__ bkind,
                                                    const S *__restrict__ bctx,
                                                    const S *__restrict__ basis, S *P,
                                                    double gamma)
{
    int i = (blockIdx.x * BLOCK) + threadIdx.x;
    if (i >= num) {
        return;
    }
    int l = bleft[i];
    int r = bright[i];
    int f = bface[i];
    S b9[9];
    for (int k = 0; k < 9; k++) b9[k] = basis[(k * nf) + f];
    S in[5], out[5], ctx[5];
    for (int v = 0; v < 5; v++) {
        in[v] = P[(v * nc) + l];
        ctx[v] = bctx[(i * 5) + v];
    }
    boundary<S>(bkind[i], out, in, (S)gamma, m3(b9), ctx);
    for (int v = 0; v < 5; v++) P[(v * nc) + r] = out[v];
}

// ---------------------------------------------------------------- gradient + limiter
// MODE 0: compute the LSQ gradient then limit; MODE 1: limiter only on a given gradient.
template <typename S, int LIM, int MODE>
__global__ __launch_bounds__(BLOCK) void k_grad_limit(Topo t, Geom<S> g,
                                                      const S *__restrict__ P,
                                                      S *__restrict__ G, S *__restrict__ psi)
{
    int c = (xcd_block() * BLOCK) + threadIdx.x;
    if (c >= t.ni) {
        return;
    }
    const int ni = t.ni;
    const int nc = t.nc;
    S pc[5], mn[5], mx[5];
    for (int v = 0; v < 5; v++) {
        pc[v] = P[(v * nc) + c];
        mn[v] = mx[v] = pc[v];
    }
    S gr[15];
    if (MODE == 0) {
        for (int i = 0; i < 15; i++) gr[i] = 0;
    }
    else {
        for (int i = 0; i < 15; i++) gr[i] = G[(i * ni) + c];
    }
#pragma unroll
    for (int k = 0; k < W; k++) {
        int e = t.inc[(k * ni) + c];
        if (e < 0) {
            continue;
        }
        int n = t.nbr[(k * ni) + c];
        int side = e & 1;
        S wx = g.w[(((0 * W) + k) * ni) + c];
        S wy = g.w[(((1 * W) + k) * ni) + c];
        S wz = g.w[(((2 * W) + k) * ni) + c];
        for (int v = 0; v < 5; v++) {
            S pn = P[(v * nc) + n];
            mn[v] = tmin(mn[v], pn);
            mx[v] = tmax(mx[v], pn);
            if (MODE == 0) {
                S diff = side ? (pc[v] - pn) : (pn - pc[v]);  // P[right] - P[left]
                S s = side ? -diff : diff;
                gr[(v * 3) + 0] += s * wx;
                gr[(v * 3) + 1] += s * wy;
                gr[(v * 3) + 2] += s * wz;
            }
        }
    }
    S ps[5];
    for (int v = 0; v < 5; v++) ps[v] = 1;
    if (LIM != 0) {
        S eps2 = (LIM == 2) ? g.eps2[c] : S(0);
#pragma unroll
        for (int k = 0; k < W; k++) {
            int e = t.inc[(k * ni) + c];
            if (e < 0) {
                continue;
            }
            V3<S> o = v3(g.o[(((0 * W) + k) * ni) + c], g.o[(((1 * W) + k) * ni) + c],
                         g.o[(((2 * W) + k) * ni) + c]);
            for (int v = 0; v < 5; v++) {
                S d2 = dot(v3(gr[(v * 3) + 0], gr[(v * 3) + 1], gr[(v * 3) + 2]), o);
                if (tabs(d2) <= S(1e-8)) {  // is_close(d2, 0), src/core/utils.c:29-34
                    continue;
                }
                S d1 = ((d2 > 0) ? mx[v] : mn[v]) - pc[v];
                if (LIM == 1) {
                    ps[v] = tmin(ps[v], tmin(S(1), d1 / d2));
                }
                else {
                    S top = (((d1 * d1) + eps2) * d2) + (2 * (d2 * d2) * d1);
                    S bot = (d1 * d1) + (2 * (d2 * d2)) + (d1 * d2) + eps2;
                    ps[v] = tmin(ps[v], top / bot / d2);
                }
            }
        }
    }
    if (MODE == 0) {
        for (int i = 0; i < 15; i++) G[(i * ni) + c] = gr[i];
    }
    for (int v = 0; v < 5; v++) psi[(v * ni) + c] = ps[v];
}

// ---------------------------------------------------------------- face fluxes + update
struct StageArgs {
    double *R;            // MODE 0: derivative out [5][ni]
    const double *U0;     // MODE 1/2
    double *U0_next;      // MODE 2: next step's U0 (may alias U0)
    double alpha;         // stage coefficient alpha[i+1]
    Scalars *scalars;     // step, dt reduction
    double *partials;     // MODE 2: [num_blocks][5] residual partial sums
    int keep_u;           // MODE 2: next U0 = U (mixed) instead of prim2cons(P_new) (FP64)
};

template <typename S>
__device__ inline void load5(S *dst, const S *__restrict__ src, int stride, int idx)
{
    for (int v = 0; v < 5; v++) dst[v] = src[(v * stride) + idx];
}

// MODE 0: write R = dU/dt; MODE 1: RK stage, write P_new; MODE 2: last RK stage (+ U0,
// dt min, residual partials).
template <typename S, int FLUX, int VISC, int ORDER, int MODE>
__global__ __launch_bounds__(BLOCK) void k_flux(Topo t, Geom<S> g, Phys ph,
                                                const S *__restrict__ P,
                                                const S *__restrict__ G,
                                                const S *__restrict__ psi,
                                                S *__restrict__ Pout, StageArgs a)
{
    const int blk = xcd_block();
    const int c = (blk * BLOCK) + threadIdx.x;
    const bool active = c < t.ni;
    const int ni = t.ni;
    const int nc = t.nc;
    const int nf = t.nf;
    const S gamma = (S)ph.gamma;
    double r[5] = {0, 0, 0, 0, 0};
    double pnew[5] = {0, 0, 0, 0, 0};
    double unew[5] = {0, 0, 0, 0, 0};
    if (active) {
        S pc[5];
        load5(pc, P, nc, c);
        Acc<S> acc;
        acc.zero();
        if (VISC) {
            const S mu = (S)ph.mu;
            const S pr = (S)ph.prandtl;
            S gc[15];
            for (int i = 0; i < 15; i++) gc[i] = G[(i * ni) + c];
#pragma unroll
            for (int k = 0; k < W; k++) {
                int e = t.inc[(k * ni) + c];
                if (e < 0) {
                    continue;
                }
                int f = e >> 1;
                int side = e & 1;
                int n = t.nbr[(k * ni) + c];
                bool bnd = (f >= t.nfi) && (f < t.nfb);
                S pn[5];
                load5(pn, P, nc, n);
                S pm[5];
                for (int v = 0; v < 5; v++) pm[v] = (pc[v] + pn[v]) / 2;
                V3<S> gm[5];
                if (bnd) {  // G[left] only (derivative.c:129-139); c is the left cell
                    for (int v = 0; v < 5; v++) {
                        gm[v] = v3(gc[(v * 3) + 0], gc[(v * 3) + 1], gc[(v * 3) + 2]);
                    }
                }
                else {
                    for (int v = 0; v < 5; v++) {
                        gm[v].x = (gc[(v * 3) + 0] + G[(((v * 3) + 0) * ni) + n]) / 2;
                        gm[v].y = (gc[(v * 3) + 1] + G[(((v * 3) + 1) * ni) + n]) / 2;
                        gm[v].z = (gc[(v * 3) + 2] + G[(((v * 3) + 2) * ni) + n]) / 2;
                    }
                }
                V3<S> cu = v3(g.corr[(0 * nf) + f], g.corr[(1 * nf) + f], g.corr[(2 * nf) + f]);
                S cn = g.corr[(3 * nf) + f];
                for (int v = 0; v < 5; v++) {  // correct, derivative.c:81-89
                    S dlr = side ? (pc[v] - pn[v]) : (pn[v] - pc[v]);
                    S proj = dot(gm[v], cu);
                    S adjust = proj - (dlr / cn);
                    V3<S> a3 = mul(adjust, cu);
                    gm[v].x -= a3.x;
                    gm[v].y -= a3.y;
                    gm[v].z -= a3.z;
                }
                V3<S> nrm = v3(g.basis[(0 * nf) + f], g.basis[(1 * nf) + f], g.basis[(2 * nf) + f]);
                S fl[5];
                viscous_central<S>(fl, pm, gm, gamma, mu, pr, nrm);
                S area = g.area[f];
                for (int v = 0; v < 5; v++) {
                    S x = fl[v] * area;
                    if (side) {
                        acc.add(v, x);
                    }
                    else {
                        acc.sub(v, x);
                    }
                }
            }
        }
        if (FLUX != FLUX_NONE) {
            S gc[15];
            if (ORDER == 2) {
                S psc[5];
                load5(psc, psi, ni, c);
                for (int v = 0; v < 5; v++) {
                    for (int d = 0; d < 3; d++) gc[(v * 3) + d] = G[(((v * 3) + d) * ni) + c] * psc[v];
                }
            }
            int err = 0;
#pragma unroll
            for (int k = 0; k < W; k++) {
                int e = t.inc[(k * ni) + c];
                if (e < 0) {
                    continue;
                }
                int f = e >> 1;
                int side = e & 1;
                int n = t.nbr[(k * ni) + c];
                bool bnd = (f >= t.nfi) && (f < t.nfb);
                S pn[5];
                load5(pn, P, nc, n);
                S qc[5], qn[5];
                if (ORDER == 2) {
                    const S *offc = side ? g.offr : g.offl;
                    const S *offn = side ? g.offl : g.offr;
                    V3<S> oc = v3(offc[(0 * nf) + f], offc[(1 * nf) + f], offc[(2 * nf) + f]);
                    for (int v = 0; v < 5; v++) {
                        qc[v] = pc[v] + dot(v3(gc[(v * 3) + 0], gc[(v * 3) + 1], gc[(v * 3) + 2]), oc);
                    }
                    if (bnd) {  // raw ghost state (derivative.c:203-211)
                        for (int v = 0; v < 5; v++) qn[v] = pn[v];
                    }
                    else {
                        V3<S> on = v3(offn[(0 * nf) + f], offn[(1 * nf) + f], offn[(2 * nf) + f]);
                        for (int v = 0; v < 5; v++) {
                            S ps = psi[(v * ni) + n];
                            V3<S> gn = v3(G[(((v * 3) + 0) * ni) + n] * ps,
                                          G[(((v * 3) + 1) * ni) + n] * ps,
                                          G[(((v * 3) + 2) * ni) + n] * ps);
                            qn[v] = pn[v] + dot(gn, on);
                        }
                    }
                }
                else {
                    for (int v = 0; v < 5; v++) {
                        qc[v] = pc[v];
                        qn[v] = pn[v];
                    }
                }
                S b9[9];
                for (int i = 0; i < 9; i++) b9[i] = g.basis[(i * nf) + f];
                S fl[5];
                if (side) {
                    convective<FLUX, S>(fl, qn, qc, gamma, m3(b9), &err);
                }
                else {
                    convective<FLUX, S>(fl, qc, qn, gamma, m3(b9), &err);
                }
                S area = g.area[f];
                for (int v = 0; v < 5; v++) {
                    S x = fl[v] * area;
                    if (side) {
                        acc.sub(v, x);
                    }
                    else {
                        acc.add(v, x);
                    }
                }
            }
            if (FLUX == FLUX_GODUNOV && err) {
                atomicMax(&a.scalars->error, err);
            }
        }
        double vol = (double)g.vol[c];
        for (int v = 0; v < 5; v++) r[v] = -acc.get(v) / vol;  // derivative.c:253-258
        if (MODE == 0) {
            for (int v = 0; v < 5; v++) a.R[(v * ni) + c] = r[v];
        }
        else {
            double step = a.scalars->step;
            for (int v = 0; v < 5; v++) unew[v] = a.U0[(v * ni) + c] + ((r[v] * a.alpha) * step);
            cons2prim<double>(pnew, unew, ph.gamma);
            for (int v = 0; v < 5; v++) Pout[(v * nc) + c] = (S)pnew[v];
        }
    }
    if (MODE == 2) {
        __shared__ double red[BLOCK / 64][6];
        double dt = __longlong_as_double(0x7ff0000000000000LL);  // +inf
        double sums[5] = {0, 0, 0, 0, 0};
        if (active) {
            double pst[5];  // the stored state, as the next step's timestep() reads it
            for (int v = 0; v < 5; v++) pst[v] = (double)(S)pnew[v];
            double vol = (double)g.vol[c];
            V3<double> pj = v3((double)g.proj[c], (double)g.proj[ni + c], (double)g.proj[(2 * ni) + c]);
            dt = ph.navierstokes ? ns_dt<double>(pst, ph.gamma, ph.mu, ph.prandtl, vol, pj)
                                 : euler_dt<double>(pst, ph.gamma, vol, pj);
            double u0n[5];
            if (a.keep_u) {
                for (int v = 0; v < 5; v++) u0n[v] = unew[v];
            }
            else {
                prim2cons<double>(u0n, pst, ph.gamma);
            }
            for (int v = 0; v < 5; v++) {
                a.U0_next[(v * ni) + c] = u0n[v];
                sums[v] = vol * (r[v] * r[v]);
            }
        }
        // deterministic block reduction: wavefront butterfly, then LDS
        for (int off = 32; off > 0; off >>= 1) {
            dt = fmin(dt, __shfl_xor(dt, off));
            for (int v = 0; v < 5; v++) sums[v] += __shfl_xor(sums[v], off);
        }
        int lane = threadIdx.x & 63;
        int wave = threadIdx.x >> 6;
        if (lane == 0) {
            red[wave][0] = dt;
            for (int v = 0; v < 5; v++) red[wave][v + 1] = sums[v];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double bdt = red[0][0];
            double bs[5];
            for (int v = 0; v < 5; v++) bs[v] = red[0][v + 1];
            for (int w = 1; w < BLOCK / 64; w++) {
                bdt = fmin(bdt, red[w][0]);
                for (int v = 0; v < 5; v++) bs[v] += red[w][v + 1];
            }
            atomicMin(&a.scalars->dtmin_bits, (unsigned long long)__double_as_longlong(bdt));
            for (int v = 0; v < 5; v++) a.partials[(blk * 5) + v] = bs[v];
        }
    }
}

// ---------------------------------------------------------------- time step + U0
template <typename S>
__global__ __launch_bounds__(BLOCK) void k_timestep_u0(Topo t, Geom<S> g, Phys ph,
                                                       const S *__restrict__ P, double *U0,
                                                       Scalars *sc, int write_u0)
{
    int c = (blockIdx.x * BLOCK) + threadIdx.x;
    double dt = __longlong_as_double(0x7ff0000000000000LL);
    if (c < t.ni) {
        double p[5];
        for (int v = 0; v < 5; v++) p[v] = (double)P[(v * t.nc) + c];
        double vol = (double)g.vol[c];
        V3<double> pj = v3((double)g.proj[c], (double)g.proj[t.ni + c], (double)g.proj[(2 * t.ni) + c]);
        dt = ph.navierstokes ? ns_dt<double>(p, ph.gamma, ph.mu, ph.prandtl, vol, pj)
                             : euler_dt<double>(p, ph.gamma, vol, pj);
        if (write_u0) {
            double u[5];
            prim2cons<double>(u, p, ph.gamma);
            for (int v = 0; v < 5; v++) U0[(v * t.ni) + c] = u[v];
        }
    }
    for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off));
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&sc->dtmin_bits, (unsigned long long)__double_as_longlong(dt));
    }
}

// ---------------------------------------------------------------- layout conversion
template <typename S>
__global__ void k_aos_to_soa(const double *__restrict__ in, S *__restrict__ out, int num,
                             int stride, int width, int out_stride)
{
    int i = (blockIdx.x * BLOCK) + threadIdx.x;
    if (i >= num) return;
    for (int v = 0; v < width; v++) out[(v * out_stride) + i] = (S)in[((long)i * stride) + v];
}

template <typename S>
__global__ void k_soa_to_aos(const S *__restrict__ in, double *__restrict__ out, int num,
                             int stride, int width, int in_stride)
{
    int i = (blockIdx.x * BLOCK) + threadIdx.x;
    if (i >= num) return;
    for (int v = 0; v < width; v++) out[((long)i * stride) + v] = (double)in[(v * in_stride) + i];
}

}  // namespace fv
