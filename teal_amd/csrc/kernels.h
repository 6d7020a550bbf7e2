// kernels.h -- HIP kernels of the residual path, templated on the storage/arithmetic type S
// of states, gradients, geometry and face fluxes (double: FP64 path; float: mixed path).
//
// Device layout.  Cells and faces are stored in the device numbering (inner cells in Morton
// order, faces grouped by their left cell; see teal_device_numbering):
//   P    2x [num_cells][5] primitive {rho,u,v,w,p} records (read / written alternately by
//                         the RK stages); boundary-ghost rows hold the BC
//                         states, periodic ghosts are never stored (they resolve to their
//                         source cell)
//   GP   [num_inner][20]  LSQ gradient (var*3 + xyz, 15 values) then limiter factor psi (5)
//   F    [num_faces][NF]  face contributions F*A: FP64 viscous (5, when present) then
//                         convective (5); FP32 one merged record (Fc - Fv)*A (FaceRec)
//   U0   [5][num_inner]   FP64 conserved state at the start of the RK step
//   incidences, slot-major [W][num_inner]: inc = face<<2 | boundary<<1 | side (side 1: the cell
//   is the face's right cell), slots sorted by the REFERENCE face index, so each cell sums
//   its face contributions in the order the reference scatters them
//   (src/equations/derivative.c:116-224); nbr = neighbour row << 1 | side for the gradient
//   kernel; per-incidence LSQ weight and face offset [3][W][ni]
//   face geometry, structure of arrays [..][num_faces]: area, basis (9), offset l/r (3+3),
//   correction (4); face cells [num_faces][2] (right < 0: ~ghost of a boundary face).  The
//   fused kernel reads a copy in block-list order: FP64 in centre form (face and cell centres
//   instead of offsets and corrections) where that reproduces the stored values; FP32 with
//   the normal alone (tangents rebuilt) and the correction derived from the offsets
// Records (P, GP, F) are arrays of structures because they are gathered through cell and
// face indices: one record is one or two cache lines instead of one line per component.
//
// Kernels per residual evaluation:
//   k_boundary   ghost states of boundary faces (src/equations/boundary.c:5-43)
//   k_grad_limit LSQ gradient (gradient.c:7-58) fused with the limiter (limiter.c:23-123);
//                one thread per inner cell, workgroups of 64, 128 or 256 (autotuned)
//   k_flux_block one workgroup per block of 256 or 128 device cells (autotuned block layout):
//                the viscous and convective flux
//                of every face touching the block (derivative.c:91-227) into LDS, then per
//                cell the viscous contributions in face order, the convective ones (FP32:
//                the merged records, summed in FP64), R = -R/V (derivative.c:229-267) and, in
//                RK mode, the stage update + cons2prim (src/simulation/advance.c:101-109);
//                the last stage also produces the next step's U0, the dt minima and the
//                residual partial sums, one per 64 device rows (so the sums do not depend on
//                the block layout)
//   k_face + k_accum  the same work as two kernels through a face flux array in HBM, used
//                when a block's face list does not fit in LDS
#pragma once

#include <hip/hip_runtime.h>

#include "physics.h"

namespace fv {

constexpr int W = 6;        // MAX_CELL_FACES (src/mesh.h:8-12)
// cells per workgroup (and threads of the per-cell kernels)
#ifndef TEAL_BLOCK
#define TEAL_BLOCK 256
#endif
constexpr int BLOCK = TEAL_BLOCK;
// Device rows are padded (batches) and numbered in blocks of BLOCK; the fused stage kernel runs
// one workgroup per block of BS = BLOCK or BLOCK / 2 rows (two block layouts, one thread per
// row: the face phase spreads the block's faces over the same threads).
constexpr int SMALL_BLOCK = BLOCK / 2;
constexpr int NUM_XCD = 8;

struct Topo {
    int ni, nc, nf;
    const int *__restrict__ inc;   // [W][ni]
    const int *__restrict__ nbr;   // [W][ni] neighbour row << 1 | side of the cell (-1: none);
                                   // the gradient kernel reads it instead of inc
    const int2 *__restrict__ flr;  // [nf] left, right (~ghost for boundary faces)
    int wmax = W;                  // most incidences of any inner cell (4: tetrahedra)
};

template <typename S>
struct Geom {
    const S *__restrict__ vol;    // [ni]
    const S *__restrict__ proj;   // [3][ni]
    const S *__restrict__ eps2;   // [ni] (VK) or null
    const S *__restrict__ w;      // [3][W][ni][2]: per incidence the LSQ weight and the
                                  // cell offset as pairs (wx, wy), (wz, ox), (oy, oz)
    const S *__restrict__ area;   // [nf]
    const S *__restrict__ basis;  // [9][nf]
    const S *__restrict__ offl;   // [3][nf]
    const S *__restrict__ offr;   // [3][nf]
    const S *__restrict__ corr;   // [4][nf] unit xyz, norm; null with offl set: derived from
                                  // the offsets (cr - cl = offl - offr)
    // centre form (offl, offr, corr null): the face offsets and corrections are recomputed
    // from the face centre and the cell centres with the expressions that produced them
    // (src/mesh/generate.c:754-777), so they are bit-identical
    const S *__restrict__ fcen;   // [3][nf] face centres
    const S *__restrict__ ccen;   // [num_cells][3] cell centres (device rows)
    const int *__restrict__ crow; // [nf] centre row of the right cell (the ghost row of a
                                  // periodic face, whose flr.y is its source cell)
    // face records as 2-wide pairs; when set, area, basis, offl, offr and fcen are not read.
    // FP32 [5][nf][2]: (area, nx), (ny, nz), (olx, oly), (olz, orx), (ory, orz);
    // FP64 centre form [7][nf][2]: (area, b0), (b1, b2), ..., (b7, b8), (fcx, fcy), (fcz, 0),
    // or with `tangents` set [4][nf][2]: (area, n0), (n1, n2), (fcx, fcy), (fcz, q), the
    // tangents rebuilt from the normal and q (face_tangents_q)
    const S *__restrict__ pairs = nullptr;
    int tangents = 0;
};

// FP64 face tangents from the stored normal n and q, with the expressions of
// src/mesh/generate.c:420-435.  q > 0: the first branch, q = nqz; q < 0: the second, -q =
// nqy.  The reference builds the tangents before it flips x and y of the basis (437-451), so
// from the stored, possibly flipped, normal these expressions give the stored tangents, or in
// the second branch both negated.  Every convective flux is odd in the tangential velocity
// components (they enter as products with the tangents or squared), and IEEE negation is
// exact, so both give the same flux bits.  The device layout is built with this form only
// where a host check found that one of the two holds for every face.
__device__ inline void face_tangents_q(double *b9, double q)
{
    const double nx = b9[0], ny = b9[1], nz = b9[2];
    if (q > 0) {
        b9[3] = -ny / q;
        b9[4] = nx / q;
        b9[5] = 0;
        b9[6] = -nx * nz / q;
        b9[7] = -ny * nz / q;
        b9[8] = q;
    }
    else {
        const double nqy = -q;
        b9[3] = -nx * ny / nqy;
        b9[4] = nqy;
        b9[5] = -ny * nz / nqy;
        b9[6] = -nz / nqy;
        b9[7] = 0;
        b9[8] = nx / nqy;
    }
}

// Face tangents from the unit normal n with the construction of src/mesh/generate.c:420-435
// (FP32 face records keep only n).  The stored basis may hold both tangents negated (the
// reference flips x and y of faces whose normal pointed into the left cell, 437-451); the
// convective fluxes are invariant under that rotation of the tangent plane.
template <typename S>
__device__ inline void face_tangents(S *b9)
{
    const S nx = b9[0], ny = b9[1], nz = b9[2];
    const S nqz = sqrt((nx * nx) + (ny * ny));
    const S nqy = sqrt((nx * nx) + (nz * nz));
    if (nqz > nqy) {
        b9[3] = -ny / nqz;
        b9[4] = nx / nqz;
        b9[5] = 0;
        b9[6] = (-nx * nz) / nqz;
        b9[7] = (-ny * nz) / nqz;
        b9[8] = nqz;
    }
    else {
        b9[3] = (-nx * ny) / nqy;
        b9[4] = nqy;
        b9[5] = (-ny * nz) / nqy;
        b9[6] = -nz / nqy;
        b9[7] = 0;
        b9[8] = nx / nqy;
    }
}

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
// bijection gives every XCD one contiguous range of rows so neighbouring rows (close in the
// Morton numbering) share that XCD's L2.
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
// FP32: the FP32 face values summed in FP64 (a conversion and one FP64 addition per value,
// the same registers as an FP32 TwoSum pair and fewer instructions; the sum of a cell's few
// FP32 values is exact unless their magnitudes span more than 29 binary orders).
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
    double v[5];
    __device__ void zero()
    {
        for (int i = 0; i < 5; i++) v[i] = 0;
    }
    __device__ void add(int i, float x) { v[i] = v[i] + (double)x; }
    __device__ void sub(int i, float x) { v[i] = v[i] - (double)x; }
    __device__ double get(int i) const { return v[i]; }
};

// Values kept per face between the face phase and the accumulation.  FP64 keeps the viscous
// and convective contributions apart, for the reference's order (every viscous face of a
// cell, then every convective one: derivative.c:116-153, 191-224).  FP32 merges them into one
// record (Fc - Fv) A per face, summed into the cell with one FP64 addition.
template <typename S, int VISC, int CONV>
struct FaceRec {
    static constexpr bool MERGE = sizeof(S) == 4 && VISC && CONV;
    static constexpr int NV = (VISC && CONV && !MERGE) ? 10 : 5;  // values per face
    static constexpr int CONV_OFF = (VISC && !MERGE) ? 5 : 0;      // first convective value
    static constexpr int AVISC = MERGE ? 0 : VISC;                  // accumulation passes
};

// Arrays streamed once per launch (face records, per-incidence weights, offsets and
// indices, cell volumes and projections, U0): loaded non-temporally, so they do not displace
// the gathered states in L2 and MALL.
template <typename T>
__device__ inline T ld_stream(const T *p)
{
    return __builtin_nontemporal_load(p);
}

template <typename S>
__device__ inline void load_rec(S *dst, const S *__restrict__ src, long row, int width)
{
    for (int v = 0; v < width; v++) dst[v] = src[(row * width) + v];
}

// ---------------------------------------------------------------- boundary ghosts
template <typename S>
__global__ __launch_bounds__(BLOCK) void k_boundary(int num, int nf, const int *__restrict__ bleft,
                                                    const int *__restrict__ bright,
                                                    const int *__restrict__ bface,
                                                    const int *__restrict__ bkind,
                                                    const S *__restrict__ bctx,
                                                    const S *__restrict__ basis, S *P,
                                                    S *P2, double gamma)
{
    int i = (blockIdx.x * BLOCK) + threadIdx.x;
    if (i >= num) {
        return;
    }
    int f = bface[i];
    S b9[9];
    for (int k = 0; k < 9; k++) b9[k] = basis[((long)k * nf) + f];
    S in[5], out[5], ctx[5];
    load_rec(in, P, bleft[i], 5);
    for (int v = 0; v < 5; v++) ctx[v] = bctx[((long)i * 5) + v];
    boundary<S>(bkind[i], out, in, (S)gamma, m3(b9), ctx);
    long r = bright[i];
    for (int v = 0; v < 5; v++) P[(r * 5) + v] = out[v];
    if (P2) {  // the other state buffer: the stage writes its inner rows only
        for (int v = 0; v < 5; v++) P2[(r * 5) + v] = out[v];
    }
}

// ---------------------------------------------------------------- gradient + limiter
// MODE 0: compute the LSQ gradient then limit, GP rows [ni][20] = gradient, psi;
// MODE 1: limiter only on the gradient in GP; MODE 2 (inviscid stages, which use the limited
// gradient alone): compute and limit, GP rows [ni][16] = gradient * psi (limiter.c:61-63)
// and one unused value, so rows stay 16-byte aligned.
// WK: incidence rows visited (W, or 4 on meshes of tetrahedra, whose rows 4 and 5 are empty).
// STAGE: the workgroup's own GB consecutive rows (a Morton-ordered sub-cube) are staged in LDS,
// and a neighbour inside that range is read from there, the others from global memory: the
// same values, fewer gathers.  It pays where the gathers miss (tetrahedra) and costs a barrier
// where they hit (structured hexahedra); the autotuner picks.
template <typename S, int LIM, int MODE, int WK = W, int GB = BLOCK, int STAGE = 0>
__global__ __launch_bounds__(GB) void k_grad_limit(Topo t, Geom<S> g, const S *__restrict__ P,
                                                   S *__restrict__ GP)
{
    __shared__ S sp[STAGE ? 5 : 1][STAGE ? GB : 1];
    const int c0 = xcd_block() * GB;
    const bool active = c0 + (int)threadIdx.x < t.ni;
    if (!STAGE && !active) {
        return;
    }
    // staged: the workgroup's threads past the last inner row work on row c0 (c0 < ni) and
    // store nothing, so that every thread reaches the barrier
    const int c = active ? c0 + (int)threadIdx.x : c0;
    const int ni = t.ni;
    S pc[5], mn[5], mx[5];
    load_rec(pc, P, c, 5);
    if constexpr (STAGE) {
        for (int v = 0; v < 5; v++) sp[v][threadIdx.x] = pc[v];
    }
    for (int v = 0; v < 5; v++) mn[v] = mx[v] = pc[v];
    S gr[15];
    if (MODE != 1) {
        for (int i = 0; i < 15; i++) gr[i] = 0;
    }
    else {
        for (int i = 0; i < 15; i++) gr[i] = GP[((long)c * 20) + i];
    }
    // every incidence's loads are issued before any is consumed (absent incidences load the
    // cell's own row and are masked), so a thread waits on memory twice, not twice per face
    int nb[WK];
#pragma unroll
    for (int k = 0; k < WK; k++) nb[k] = ld_stream(&t.nbr[((long)k * ni) + c]);
    typedef S S2 __attribute__((ext_vector_type(2)));
    const S2 *wo = reinterpret_cast<const S2 *>(g.w);
    S pn[WK][5], wk[WK][3], ok[WK][3];
    const int last = min(c0 + GB, ni);  // staged rows [c0, last): inner rows of this workgroup
    if constexpr (STAGE) {
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < WK; k++) {
        const int n = (nb[k] < 0) ? c : (nb[k] >> 1);
        if (STAGE && n >= c0 && n < last) {
            for (int v = 0; v < 5; v++) pn[k][v] = sp[v][n - c0];
        }
        else {
            load_rec(pn[k], P, n, 5);
        }
        const S2 a = ld_stream(&wo[((long)((0 * W) + k) * ni) + c]);
        const S2 b = ld_stream(&wo[((long)((1 * W) + k) * ni) + c]);
        wk[k][0] = a.x;
        wk[k][1] = a.y;
        wk[k][2] = b.x;
        ok[k][0] = b.y;
        if (LIM != 0) {
            const S2 e = ld_stream(&wo[((long)((2 * W) + k) * ni) + c]);
            ok[k][1] = e.x;
            ok[k][2] = e.y;
        }
    }
#pragma unroll
    for (int k = 0; k < WK; k++) {
        if (nb[k] < 0) {
            continue;
        }
        const int side = nb[k] & 1;
        for (int v = 0; v < 5; v++) {
            mn[v] = tmin(mn[v], pn[k][v]);
            mx[v] = tmax(mx[v], pn[k][v]);
            if (MODE != 1) {
                S diff = side ? (pc[v] - pn[k][v]) : (pn[k][v] - pc[v]);  // P[right] - P[left]
                S s = side ? -diff : diff;
                gr[(v * 3) + 0] += s * wk[k][0];
                gr[(v * 3) + 1] += s * wk[k][1];
                gr[(v * 3) + 2] += s * wk[k][2];
            }
        }
    }
    S ps[5];
    for (int v = 0; v < 5; v++) ps[v] = 1;
    if (LIM != 0) {
        S eps2 = (LIM == 2) ? g.eps2[c] : S(0);
#pragma unroll
        for (int k = 0; k < WK; k++) {
            if (nb[k] < 0) {
                continue;
            }
            V3<S> o = v3(ok[k][0], ok[k][1], ok[k][2]);
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
    if (!active) {
        return;
    }
    if (MODE == 2) {
        S *out = &GP[(long)c * 16];
        for (int i = 0; i < 15; i++) out[i] = gr[i] * ps[i / 3];
        out[15] = 0;
        return;
    }
    S *out = &GP[(long)c * 20];
    if (MODE == 0) {
        for (int i = 0; i < 15; i++) out[i] = gr[i];
    }
    for (int v = 0; v < 5; v++) out[15 + v] = ps[v];
}

// ---------------------------------------------------------------- face fluxes
// Viscous (fv) and convective (fc) contribution F*A of face f, whose cells are lr (left,
// right or ~ghost) and, in the centre form, whose right cell's centre is row `crow` (g.crow[f],
// loaded by the caller).  GPRE: FP64 records with rebuilt tangents (g.tangents), their 4 pairs
// loaded by the caller into `pre` (one face ahead).
template <typename S, int FLUX, int VISC, int ORDER, bool GPRE = false>
__device__ inline void face_flux(const Topo &t, const Geom<S> &g, const Phys &ph,
                                 const S *__restrict__ P, const S *__restrict__ GP, int f,
                                 int2 lr, int crow, S *fv, S *fc, int *err,
                                 const S *pre = nullptr)
{
    const int nf = t.nf;
    const S gamma = (S)ph.gamma;
    const int l = lr.x;
    const bool bnd = lr.y < 0;
    const int r = bnd ? ~lr.y : lr.y;
    // face -> cell offsets and the centre correction (generate.c:754-777)
    V3<S> ol, orr, cu, nrm;
    S cn, area;
    bool paired = false;
    if constexpr (sizeof(S) == 4) {
        if (g.pairs) {
            typedef float f2 __attribute__((ext_vector_type(2)));
            const f2 *q = reinterpret_cast<const f2 *>(g.pairs);
            const f2 a = ld_stream(&q[f]), b = ld_stream(&q[(long)nf + f]), c = ld_stream(&q[(2 * (long)nf) + f]),
                     d = ld_stream(&q[(3 * (long)nf) + f]), e = ld_stream(&q[(4 * (long)nf) + f]);
            area = a.x;
            nrm = v3(a.y, b.x, b.y);
            ol = v3(c.x, c.y, d.x);
            orr = v3(d.y, e.x, e.y);
            V3<S> dl = v3(ol.x - orr.x, ol.y - orr.y, ol.z - orr.z);
            cn = sqrt(dot(dl, dl));
            cu = v3(dl.x / cn, dl.y / cn, dl.z / cn);
            paired = true;
        }
    }
    S bp[9];  // FP64 paired records: the full basis
    if constexpr (sizeof(S) == 8) {
        if (g.pairs) {
            typedef double d2 __attribute__((ext_vector_type(2)));
            const d2 *q = reinterpret_cast<const d2 *>(g.pairs);
            V3<S> fc;
            if (GPRE || g.tangents) {  // (area, n0), (n1, n2), (fcx, fcy), (fcz, q)
                d2 x[4];
                if constexpr (GPRE) {
                    for (int k = 0; k < 4; k++) x[k] = d2{pre[2 * k], pre[(2 * k) + 1]};
                }
                else {
                    for (int k = 0; k < 4; k++) x[k] = ld_stream(&q[((long)k * nf) + f]);
                }
                area = x[0].x;
                bp[0] = x[0].y;
                bp[1] = x[1].x;
                bp[2] = x[1].y;
                face_tangents_q(bp, x[3].y);
                fc = v3(x[2].x, x[2].y, x[3].x);
            }
            else {
                d2 x[7];
                for (int k = 0; k < 7; k++) x[k] = ld_stream(&q[((long)k * nf) + f]);
                area = x[0].x;
                bp[0] = x[0].y;
                for (int k = 1; k < 5; k++) {
                    bp[(2 * k) - 1] = x[k].x;
                    bp[2 * k] = x[k].y;
                }
                fc = v3(x[5].x, x[5].y, x[6].x);
            }
            nrm = v3(bp[0], bp[1], bp[2]);
            const long rr = crow;
            V3<S> cl = v3(g.ccen[((long)l * 3) + 0], g.ccen[((long)l * 3) + 1], g.ccen[((long)l * 3) + 2]);
            V3<S> cr = v3(g.ccen[(rr * 3) + 0], g.ccen[(rr * 3) + 1], g.ccen[(rr * 3) + 2]);
            ol = v3(fc.x - cl.x, fc.y - cl.y, fc.z - cl.z);
            orr = v3(fc.x - cr.x, fc.y - cr.y, fc.z - cr.z);
            V3<S> dl = v3(cr.x - cl.x, cr.y - cl.y, cr.z - cl.z);
            cn = sqrt(dot(dl, dl));
            cu = v3(dl.x / cn, dl.y / cn, dl.z / cn);
            paired = true;
        }
    }
    if (paired) {
    }
    else if (g.offl) {
        ol = v3(ld_stream(&g.offl[(0 * (long)nf) + f]), ld_stream(&g.offl[(1 * (long)nf) + f]), ld_stream(&g.offl[(2 * (long)nf) + f]));
        orr = v3(ld_stream(&g.offr[(0 * (long)nf) + f]), ld_stream(&g.offr[(1 * (long)nf) + f]), ld_stream(&g.offr[(2 * (long)nf) + f]));
        if (g.corr) {
            cu = v3(ld_stream(&g.corr[(0 * (long)nf) + f]), ld_stream(&g.corr[(1 * (long)nf) + f]), ld_stream(&g.corr[(2 * (long)nf) + f]));
            cn = ld_stream(&g.corr[(3 * (long)nf) + f]);
        }
        else {
            V3<S> dl = v3(ol.x - orr.x, ol.y - orr.y, ol.z - orr.z);
            cn = sqrt(dot(dl, dl));
            cu = v3(dl.x / cn, dl.y / cn, dl.z / cn);
        }
    }
    else {
        const long rr = crow;
        V3<S> fc = v3(ld_stream(&g.fcen[(0 * (long)nf) + f]), ld_stream(&g.fcen[(1 * (long)nf) + f]), ld_stream(&g.fcen[(2 * (long)nf) + f]));
        V3<S> cl = v3(g.ccen[((long)l * 3) + 0], g.ccen[((long)l * 3) + 1], g.ccen[((long)l * 3) + 2]);
        V3<S> cr = v3(g.ccen[(rr * 3) + 0], g.ccen[(rr * 3) + 1], g.ccen[(rr * 3) + 2]);
        ol = v3(fc.x - cl.x, fc.y - cl.y, fc.z - cl.z);
        orr = v3(fc.x - cr.x, fc.y - cr.y, fc.z - cr.z);
        V3<S> dl = v3(cr.x - cl.x, cr.y - cl.y, cr.z - cl.z);
        cn = sqrt(dot(dl, dl));
        cu = v3(dl.x / cn, dl.y / cn, dl.z / cn);
    }
    S pl[5], pr[5];
    load_rec(pl, P, l, 5);
    load_rec(pr, P, r, 5);
    // gradient records: viscous passes read [G (15), psi (5)] rows; inviscid passes read the
    // limited gradient G * psi alone, from rows of 16 (k_grad_limit MODE 2)
    S gl[20], gr[20];
    if (ORDER == 2) {
        if (VISC) {
            load_rec(gl, GP, l, 20);
            if (!bnd) {
                load_rec(gr, GP, r, 20);
            }
        }
        else {
            for (int i = 0; i < 15; i++) gl[i] = GP[((long)l * 16) + i];
            if (!bnd) {
                for (int i = 0; i < 15; i++) gr[i] = GP[((long)r * 16) + i];
            }
        }
    }
    if (!paired) {
        area = ld_stream(&g.area[f]);
        nrm = v3(ld_stream(&g.basis[(0 * (long)nf) + f]), ld_stream(&g.basis[(1 * (long)nf) + f]),
                 ld_stream(&g.basis[(2 * (long)nf) + f]));
    }
    if (VISC) {
        const S mu = (S)ph.mu;
        const S prl = (S)ph.prandtl;
        S pm[5];
        for (int v = 0; v < 5; v++) pm[v] = (pl[v] + pr[v]) / 2;
        V3<S> gm[5];
        for (int v = 0; v < 5; v++) {
            if (bnd) {  // G[left] only (derivative.c:129-139)
                gm[v] = v3(gl[(v * 3) + 0], gl[(v * 3) + 1], gl[(v * 3) + 2]);
            }
            else {
                gm[v] = v3((gl[(v * 3) + 0] + gr[(v * 3) + 0]) / 2, (gl[(v * 3) + 1] + gr[(v * 3) + 1]) / 2,
                           (gl[(v * 3) + 2] + gr[(v * 3) + 2]) / 2);
            }
        }
        for (int v = 0; v < 5; v++) {  // correct, derivative.c:81-89
            S proj = dot(gm[v], cu);
            S adjust = proj - ((pr[v] - pl[v]) / cn);
            V3<S> a3 = mul(adjust, cu);
            gm[v].x -= a3.x;
            gm[v].y -= a3.y;
            gm[v].z -= a3.z;
        }
        S fl[5];
        if constexpr (sizeof(S) == 4) {
            viscous_central_normal<S>(fl, pm, gm, gamma, mu, prl, nrm);
        }
        else {
            viscous_central<S>(fl, pm, gm, gamma, mu, prl, nrm);
        }
        for (int v = 0; v < 5; v++) fv[v] = fl[v] * area;
    }
    if (FLUX != FLUX_NONE) {
        S ql[5], qr[5];
        if (ORDER == 2) {  // reconstruct with the limited gradient G*psi (derivative.c:158-164)
            auto limited = [&](const S *gx, int v) {
                if (VISC) {
                    S ps = gx[15 + v];
                    return v3(gx[(v * 3) + 0] * ps, gx[(v * 3) + 1] * ps, gx[(v * 3) + 2] * ps);
                }
                return v3(gx[(v * 3) + 0], gx[(v * 3) + 1], gx[(v * 3) + 2]);
            };
            // FP32: (G psi).o = psi (G.o), the offset product first
            auto delta = [&](const S *gx, int v, V3<S> o) {
                if (sizeof(S) == 4 && VISC) {
                    return gx[15 + v] * dot(v3(gx[(v * 3) + 0], gx[(v * 3) + 1], gx[(v * 3) + 2]), o);
                }
                return dot(limited(gx, v), o);
            };
            for (int v = 0; v < 5; v++) ql[v] = pl[v] + delta(gl, v, ol);
            if (bnd) {  // raw ghost state (derivative.c:203-211)
                for (int v = 0; v < 5; v++) qr[v] = pr[v];
            }
            else {
                for (int v = 0; v < 5; v++) qr[v] = pr[v] + delta(gr, v, orr);
            }
        }
        else {
            for (int v = 0; v < 5; v++) {
                ql[v] = pl[v];
                qr[v] = pr[v];
            }
        }
        S flx[5];
        if constexpr (sizeof(S) == 4 && FLUX == FLUX_HLLC) {  // FP32: the flux from the normal alone
            flux_hllc_normal(flx, ql, qr, gamma, nrm);
        }
        else {
            S b9[9];
            if (sizeof(S) == 8 && paired) {
                for (int i = 0; i < 9; i++) b9[i] = bp[i];
            }
            else if (paired) {  // FP32 records keep the normal alone
                b9[0] = nrm.x;
                b9[1] = nrm.y;
                b9[2] = nrm.z;
                face_tangents(b9);
            }
            else {
                for (int i = 0; i < 9; i++) b9[i] = ld_stream(&g.basis[((long)i * nf) + f]);
            }
            convective<FLUX, S>(flx, ql, qr, gamma, m3(b9), err);
        }
        for (int v = 0; v < 5; v++) fc[v] = flx[v] * area;
    }
}

// One thread per face: F[f] = {Fv*A (5, viscous), Fc*A (5, convective)}.
template <typename S, int FLUX, int VISC, int ORDER>
__global__ __launch_bounds__(BLOCK) void k_face(Topo t, Geom<S> g, Phys ph, const S *__restrict__ P,
                                                const S *__restrict__ GP, S *__restrict__ F,
                                                int *error)
{
    using R = FaceRec<S, VISC, FLUX != FLUX_NONE>;
    constexpr int NF = R::NV;
    const int f = (xcd_block() * BLOCK) + threadIdx.x;
    if (f >= t.nf) {
        return;
    }
    S fv[5], fc[5];
    int err = 0;
    face_flux<S, FLUX, VISC, ORDER>(t, g, ph, P, GP, f, t.flr[f], g.crow ? g.crow[f] : 0, fv, fc,
                                    &err);
    S *out = &F[(long)f * NF];
    if (R::MERGE) {
        for (int v = 0; v < 5; v++) out[v] = fc[v] - fv[v];
    }
    else {
        if (VISC) {
            for (int v = 0; v < 5; v++) out[v] = fv[v];
        }
        if (FLUX != FLUX_NONE) {
            for (int v = 0; v < 5; v++) out[R::CONV_OFF + v] = fc[v];
        }
    }
    if (FLUX == FLUX_GODUNOV && err) {
        atomicMax(error, err);
    }
}

// ---------------------------------------------------------------- accumulate + update
struct StageArgs {
    double *R;            // MODE 0: derivative out [5][ni]
    const double *U0;     // MODE 1/2
    double *U0_next;      // MODE 2: next step's U0 (may alias U0)
    double alpha;         // stage coefficient alpha[i+1]
    Scalars *scalars;     // step, dt reduction: one record per instance
    const int *grp_inst;  // instance of each group of 64 rows (null: one instance)
    double *partials;     // MODE 2: [num_groups][5] residual partial sums, one per 64 rows
    double *dtpart;       // MODE 2: [num_groups] minima of the local time step
    int keep_u;           // MODE 2: next U0 = U (mixed) instead of prim2cons(P_new) (FP64)
    // bulk-velocity forcing source, or null: {entity average of the primitive state (5),
    // target velocity U, relaxation time tau}; the cell's source is s1 = rho (U - <u>) / tau
    // in the x momentum and u s1 in the energy (the source and Prepare hook of
    // run/periodic_hill/run.c:8-21, evaluated on the state the residual is taken of)
    const double *forcing = nullptr;
};

// Per-cell inputs of stage_epilogue: the volume, U0 (RK stages) and the projections (last
// stage).  The FP64 stage kernel loads them before its face loop, where their latency hides
// behind the face fluxes (its LDS allows 2 waves per SIMD, so the registers are free).
struct CellIn {
    double vol = 0, u0[5] = {0, 0, 0, 0, 0}, pj[3] = {0, 0, 0};
};

template <typename S, int MODE>
__device__ inline CellIn cell_loads(const Topo &t, const Geom<S> &g, const StageArgs &a, int c,
                                    bool active)
{
    CellIn in;
    if (active) {
        const int ni = t.ni;
        in.vol = (double)ld_stream(&g.vol[c]);
        if (MODE != 0) {
            for (int v = 0; v < 5; v++) in.u0[v] = ld_stream(&a.U0[((long)v * ni) + c]);
        }
        if (MODE == 2) {
            for (int k = 0; k < 3; k++) in.pj[k] = (double)ld_stream(&g.proj[((long)k * ni) + c]);
        }
    }
    return in;
}

// Stage epilogue of inner cell c (active) given its accumulated face sums: finalize
// R = -R/V (derivative.c:253-258); MODE 0 writes R; MODE 1 writes P_new = cons2prim(U0 +
// (R alpha) step) (advance.c:103-108); MODE 2 also produces the next step's U0, and the dt
// minimum and residual partial sums of the 64 rows of this wavefront (c >> 6: every lane of
// the wave must call it, with consecutive c).
// Pin is the state the residual is taken of (the source term's context); it may equal Pout.
template <typename S, int MODE>
__device__ inline void stage_epilogue(const Topo &t, const Geom<S> &g, const Phys &ph, const Acc<S> &acc,
                                      int c, bool active, const S *Pin, S *Pout, const StageArgs &a,
                                      const CellIn &in)
{
    const int ni = t.ni;
    double r[5] = {0, 0, 0, 0, 0};
    double pnew[5] = {0, 0, 0, 0, 0};
    double unew[5] = {0, 0, 0, 0, 0};
    if (active) {
        const double vol = in.vol;
        if constexpr (sizeof(S) == 8) {  // the reference's divisions (derivative.c:253-258)
            for (int v = 0; v < 5; v++) r[v] = -acc.get(v) / vol;
        }
        else {  // FP32 path: one reciprocal, FP64 products
            const double iv = 1.0 / vol;
            for (int v = 0; v < 5; v++) r[v] = -acc.get(v) * iv;
        }
        if (a.forcing) {  // R = (-R/V) + source, every component (derivative.c:260-265)
            const double rho = (double)Pin[(long)c * 5];
            const double u = (double)Pin[((long)c * 5) + 1];
            const double s1 = (rho * (a.forcing[5] - a.forcing[1])) / a.forcing[6];
            const double src[5] = {0.0, s1, 0.0, 0.0, u * s1};
            for (int v = 0; v < 5; v++) r[v] = r[v] + src[v];
        }
        if (MODE == 0) {
            for (int v = 0; v < 5; v++) a.R[((long)v * ni) + c] = r[v];
        }
        else {
            double step = a.scalars[a.grp_inst ? a.grp_inst[c >> 6] : 0].step;
            for (int v = 0; v < 5; v++) unew[v] = in.u0[v] + ((r[v] * a.alpha) * step);
            if constexpr (sizeof(S) == 8) {
                cons2prim<double>(pnew, unew, ph.gamma);
            }
            else {
                cons2prim_recip<double>(pnew, unew, ph.gamma);
            }
            for (int v = 0; v < 5; v++) Pout[((long)c * 5) + v] = (S)pnew[v];
        }
    }
    if (MODE == 2) {
        double dt = __longlong_as_double(0x7ff0000000000000LL);  // +inf
        double sums[5] = {0, 0, 0, 0, 0};
        if (active) {
            double pst[5];  // the stored state, as the next step's timestep() reads it
            for (int v = 0; v < 5; v++) pst[v] = (double)(S)pnew[v];
            const double vol = in.vol;
            V3<double> pj = v3(in.pj[0], in.pj[1], in.pj[2]);
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
                __builtin_nontemporal_store(u0n[v], &a.U0_next[((long)v * ni) + c]);  // read next step
                sums[v] = vol * (r[v] * r[v]);
            }
        }
        // deterministic wavefront butterfly; lane 0 (row c = the group's first) writes
        for (int off = 32; off > 0; off >>= 1) {
            dt = fmin(dt, __shfl_xor(dt, off));
            for (int v = 0; v < 5; v++) sums[v] += __shfl_xor(sums[v], off);
        }
        if ((threadIdx.x & 63) == 0 && c < ni) {
            const long grp = c >> 6;
            a.dtpart[grp] = dt;
            for (int v = 0; v < 5; v++) a.partials[(grp * 5) + v] = sums[v];
        }
    }
}

// Sum of the face contributions of inner cell c in reference face order: viscous
// (R[left] -= Fv*A, R[right] += Fv*A, derivative.c:123-126) then convective (R[left] += Fc*A,
// R[right] -= Fc*A, derivative.c:197-200).  `slot(k)` returns incidence k of the cell (< 0:
// none), `fetch(e, part)` the 5 values of the face of incidence e (part 0: viscous, 1:
// convective).
template <typename S, int VISC, int CONV, typename Slot, typename Fetch>
__device__ inline void accumulate(Acc<S> &acc, Slot slot, Fetch fetch)
{
    acc.zero();
    if (VISC) {
#pragma unroll
        for (int k = 0; k < W; k++) {
            int e = slot(k);
            if (e < 0) {
                continue;
            }
            S x[5];
            fetch(e, 0, x);
            for (int v = 0; v < 5; v++) {
                if (e & 1) {
                    acc.add(v, x[v]);
                }
                else {
                    acc.sub(v, x[v]);
                }
            }
        }
    }
    if (CONV) {
#pragma unroll
        for (int k = 0; k < W; k++) {
            int e = slot(k);
            if (e < 0) {
                continue;
            }
            S x[5];
            fetch(e, 1, x);
            for (int v = 0; v < 5; v++) {
                if (e & 1) {
                    acc.sub(v, x[v]);
                }
                else {
                    acc.add(v, x[v]);
                }
            }
        }
    }
}

// Accumulate from the face flux array F written by k_face.
// Pout is also the state the residual is taken of (the face fluxes are already in F): the
// source term reads a cell's row before its update overwrites it, so Pout is not restrict.
template <typename S, int VISC, int CONV, int MODE>
__global__ __launch_bounds__(BLOCK) void k_accum(Topo t, Geom<S> g, Phys ph, const S *__restrict__ F,
                                                 S *Pout, StageArgs a)
{
    using R = FaceRec<S, VISC, CONV>;
    const int blk = xcd_block();
    const int c = (blk * BLOCK) + threadIdx.x;
    const bool active = c < t.ni;
    Acc<S> acc;
    acc.zero();
    if (active) {
        accumulate<S, R::AVISC, CONV>(
            acc, [&](int k) { return ld_stream(&t.inc[((long)k * t.ni) + c]); },
            [&](int e, int part, S *x) {
                const S *rec = &F[((long)(e >> 2) * R::NV) + (part ? R::CONV_OFF : 0)];
                for (int v = 0; v < 5; v++) x[v] = rec[v];
            });
    }
    stage_epilogue<S, MODE>(t, g, ph, acc, c, active, Pout, Pout, a,
                            cell_loads<S, MODE>(t, g, a, c, active));
}

// Per-block face lists of the fused kernel: the workgroup of device cells
// [b*BS, (b+1)*BS) computes the fluxes of every face touching them into LDS, then each
// cell sums its faces from there.  The faces of block b are list entries [off[b], off[b+1]);
// a face between two blocks appears in both lists, and the face geometry (Geom face arrays,
// stride nj) is stored in list order, so each block streams its faces contiguously.
// slot = position in the block's list << 1 | side.
struct BlockTopo {
    const int *__restrict__ off;    // [num_blocks + 1]
    const int *__restrict__ slot;   // [W][ni]
    const int2 *__restrict__ flr;   // [nj] left, right (~ghost row) of each list entry
    int nj;                         // list entries
    int max_faces;                  // longest list (LDS rows)
};

// GEO (FP64 with viscous and convective records and rebuilt tangents, g.tangents): the next
// face's geometry records are loaded while this face is computed.
template <typename S, int FLUX, int VISC, int ORDER, int MODE, int BS, int GEO = 0>
__global__ __launch_bounds__(BS) void k_flux_block(Topo t, BlockTopo bt, Geom<S> g, Phys ph,
                                                   const S *__restrict__ P,
                                                   const S *__restrict__ GP,
                                                   S *__restrict__ Pout, StageArgs a)
{
    constexpr int CONV = FLUX != FLUX_NONE;
    using R = FaceRec<S, VISC, CONV>;
    extern __shared__ unsigned char lds_raw[];
    S *lds = (S *)lds_raw;  // [R::NV][max_faces]
    const int mf = bt.max_faces;
    const int blk = xcd_block();
    const int beg = bt.off[blk];
    const int num = bt.off[blk + 1] - beg;
    int err = 0;
    Topo tf = t;  // face arrays in list order
    tf.nf = bt.nj;
    tf.flr = bt.flr;
    const int c = (blk * BS) + threadIdx.x;
    const bool active = c < t.ni;
    // FP64 with viscous and convective records: the cell's incidence slots and epilogue inputs
    // are loaded before the face loop (registers to spare at the 2 waves per SIMD its LDS
    // allows); the other kernels load them after it
    constexpr bool EARLY = sizeof(S) == 8 && R::NV == 10;
    CellIn in;
    int sl[W];
    if constexpr (EARLY) {
        in = cell_loads<S, MODE>(t, g, a, c, active);
        for (int k = 0; k < W; k++) sl[k] = active ? ld_stream(&bt.slot[((long)k * t.ni) + c]) : -1;
    }
    // the next face's cell pair (and, in the centre form, the row of its right cell's centre)
    // is loaded while this face is computed (its gathers need it); with GEO, so are its
    // geometry records (streamed from HBM, the longest wait)
    int2 lr_next = (threadIdx.x < num) ? bt.flr[beg + threadIdx.x] : int2{0, 0};
    int cr_next = (threadIdx.x < num && g.crow) ? g.crow[beg + threadIdx.x] : 0;
    constexpr bool gpre = EARLY && GEO;
    typedef S S2 __attribute__((ext_vector_type(2)));
    const S2 *q2 = reinterpret_cast<const S2 *>(g.pairs);
    S gx_next[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (gpre && threadIdx.x < num) {
        for (int k = 0; k < 4; k++) {
            const S2 x = ld_stream(&q2[((long)k * bt.nj) + beg + threadIdx.x]);
            gx_next[2 * k] = x.x;
            gx_next[(2 * k) + 1] = x.y;
        }
    }
    for (int j = threadIdx.x; j < num; j += BS) {
        const int2 lr = lr_next;
        const int cr = cr_next;
        S gx[8];
        for (int i = 0; i < 8; i++) gx[i] = gx_next[i];
        if (j + BS < num) {
            lr_next = bt.flr[beg + j + BS];
            if (g.crow) {
                cr_next = g.crow[beg + j + BS];
            }
            if (gpre) {
                for (int k = 0; k < 4; k++) {
                    const S2 x = ld_stream(&q2[((long)k * bt.nj) + beg + j + BS]);
                    gx_next[2 * k] = x.x;
                    gx_next[(2 * k) + 1] = x.y;
                }
            }
        }
        S fv[5], fc[5];
        face_flux<S, FLUX, VISC, ORDER, gpre>(tf, g, ph, P, GP, beg + j, lr, cr, fv, fc, &err, gx);
        if (R::MERGE) {
            for (int v = 0; v < 5; v++) lds[(v * mf) + j] = fc[v] - fv[v];
        }
        else {
            if (VISC) {
                for (int v = 0; v < 5; v++) lds[(v * mf) + j] = fv[v];
            }
            if (CONV) {
                for (int v = 0; v < 5; v++) lds[((R::CONV_OFF + v) * mf) + j] = fc[v];
            }
        }
    }
    if (FLUX == FLUX_GODUNOV && err) {
        atomicMax(&a.scalars->error, err);
    }
    __syncthreads();
    if constexpr (!EARLY) {
        in = cell_loads<S, MODE>(t, g, a, c, active);
    }
    Acc<S> acc;
    acc.zero();
    if (active) {
        accumulate<S, R::AVISC, CONV>(
            acc,
            [&](int k) { return EARLY ? sl[k] : ld_stream(&bt.slot[((long)k * t.ni) + c]); },
            [&](int e, int part, S *x) {
                const int j = e >> 1;
                for (int v = 0; v < 5; v++) x[v] = lds[(((part ? R::CONV_OFF : 0) + v) * mf) + j];
            });
    }
    stage_epilogue<S, MODE>(t, g, ph, acc, c, active, P, Pout, a, in);
}

// ---------------------------------------------------------------- time step + U0
// local time step (src/equations/timestep.c:7-30) and, if asked, U0 = prim2cons(P)
// (advance.c:95-98); dtpart[group] = the minimum of each group of 64 rows
template <typename S>
__global__ __launch_bounds__(BLOCK) void k_timestep_u0(Topo t, Geom<S> g, Phys ph,
                                                       const S *__restrict__ P, double *U0,
                                                       double *dtpart, int write_u0)
{
    int c = (blockIdx.x * BLOCK) + threadIdx.x;
    double dt = __longlong_as_double(0x7ff0000000000000LL);
    if (c < t.ni) {
        double p[5];
        for (int v = 0; v < 5; v++) p[v] = (double)P[((long)c * 5) + v];
        double vol = (double)ld_stream(&g.vol[c]);
        V3<double> pj = v3((double)ld_stream(&g.proj[c]), (double)ld_stream(&g.proj[(long)t.ni + c]),
                            (double)ld_stream(&g.proj[(2 * (long)t.ni) + c]));
        dt = ph.navierstokes ? ns_dt<double>(p, ph.gamma, ph.mu, ph.prandtl, vol, pj)
                             : euler_dt<double>(p, ph.gamma, vol, pj);
        if (write_u0) {
            double u[5];
            prim2cons<double>(u, p, ph.gamma);
            for (int v = 0; v < 5; v++) U0[((long)v * t.ni) + c] = u[v];
        }
    }
    for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off));
    if ((threadIdx.x & 63) == 0 && c < t.ni) {
        dtpart[c >> 6] = dt;
    }
}

// ---------------------------------------------------------------- layout conversion
// host rows (reference numbering, out/in_stride doubles each) <-> device rows (device
// numbering): device row i is host row perm[i] (identity when perm is null).  SOA selects a
// [width][num] device array instead of [num] records of `stride` values.
template <typename S, int SOA>
__global__ void k_to_device(const double *__restrict__ in, S *__restrict__ out, int num,
                            int in_stride, int width, int out_stride, const int *__restrict__ perm)
{
    int i = (blockIdx.x * BLOCK) + threadIdx.x;
    if (i >= num) return;
    long src = perm ? perm[i] : i;
    for (int v = 0; v < width; v++) {
        // a padding row of a batch (no host row) holds a resting state, rho = p = 1, in the
        // primitive records (row-major, 5 wide), and zeros in every other array
        S x = (src >= 0) ? (S)in[(src * in_stride) + v]
                         : (S)((!SOA && width == 5 && (v == 0 || v == 4)) ? 1 : 0);
        if (SOA) {
            out[((long)v * num) + i] = x;
        }
        else {
            out[((long)i * out_stride) + v] = x;
        }
    }
}

template <typename S, int SOA>
__global__ void k_to_host(const S *__restrict__ in, double *__restrict__ out, int num,
                          int in_stride, int width, int out_stride, const int *__restrict__ perm)
{
    int i = (blockIdx.x * BLOCK) + threadIdx.x;
    if (i >= num) return;
    long dst = perm ? perm[i] : i;
    if (dst < 0) return;  // padding row of a batch
    for (int v = 0; v < width; v++) {
        S x = SOA ? in[((long)v * num) + i] : in[((long)i * in_stride) + v];
        out[(dst * out_stride) + v] = (double)x;
    }
}

}  // namespace fv
