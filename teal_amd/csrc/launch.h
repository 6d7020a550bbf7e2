// launch.h -- kernel launchers and the compile-time flux dispatch table.
// Included by device.hip (instantiates Kern<double>) and mixed.hip (Kern<float>).
#pragma once

#include "device.h"

namespace fv {

inline int grid_of(long n) { return (int)((n + BLOCK - 1) / BLOCK); }

template <typename S>
using FaceLaunch = void (*)(hipStream_t, Topo, Geom<S>, Phys, const S *, const S *, S *, int *);

template <typename S, int FLUX, int VISC, int ORDER>
static void launch_face(hipStream_t s, Topo t, Geom<S> g, Phys ph, const S *P, const S *GP, S *F,
                        int *error)
{
    hipLaunchKernelGGL((k_face<S, FLUX, VISC, ORDER>), dim3(grid_of(t.nf)), dim3(BLOCK), 0, s, t, g,
                       ph, P, GP, F, error);
}

// one face kernel per (flux, viscous, order): the flux is a compile-time choice, as the
// reference's name table (src/euler/convective.c:593-623) is a link-time one
template <typename S, int FLUX>
static FaceLaunch<S> face_row(int visc, int order)
{
    if (order == 1) {
        return launch_face<S, FLUX, 0, 1>;
    }
    return visc ? launch_face<S, FLUX, 1, 2> : launch_face<S, FLUX, 0, 2>;
}

template <typename S>
static FaceLaunch<S> face_table(int flux, int visc, int order)
{
    switch (flux) {
        case FLUX_NONE: return face_row<S, FLUX_NONE>(visc, order);
        case FLUX_GODUNOV: return face_row<S, FLUX_GODUNOV>(visc, order);
        case FLUX_ROE: return face_row<S, FLUX_ROE>(visc, order);
        case FLUX_HLL: return face_row<S, FLUX_HLL>(visc, order);
        case FLUX_HLLC: return face_row<S, FLUX_HLLC>(visc, order);
        case FLUX_HLLE: return face_row<S, FLUX_HLLE>(visc, order);
        case FLUX_AUSMD: return face_row<S, FLUX_AUSMD>(visc, order);
        case FLUX_AUSMDV: return face_row<S, FLUX_AUSMDV>(visc, order);
        case FLUX_LXF: return face_row<S, FLUX_LXF>(visc, order);
        case FLUX_CENTRAL: return face_row<S, FLUX_CENTRAL>(visc, order);
        default: return nullptr;
    }
}

template <typename S, int VISC, int CONV, int MODE>
static void launch_accum(hipStream_t s, Topo t, Geom<S> g, Phys ph, const S *F, S *P, StageArgs a)
{
    hipLaunchKernelGGL((k_accum<S, VISC, CONV, MODE>), dim3(grid_of(t.ni)), dim3(BLOCK), 0, s, t, g,
                       ph, F, P, a);
}

template <typename S, int VISC, int CONV>
static void accum_row(int mode, hipStream_t s, Topo t, Geom<S> g, Phys ph, const S *F, S *P,
                      StageArgs a)
{
    switch (mode) {
        case 0: launch_accum<S, VISC, CONV, 0>(s, t, g, ph, F, P, a); return;
        case 1: launch_accum<S, VISC, CONV, 1>(s, t, g, ph, F, P, a); return;
        default: launch_accum<S, VISC, CONV, 2>(s, t, g, ph, F, P, a); return;
    }
}

template <typename S>
using BlockLaunch = void (*)(hipStream_t, size_t, Topo, BlockTopo, Geom<S>, Phys, const S *,
                             const S *, S *, StageArgs);

template <typename S, int FLUX, int VISC, int ORDER, int MODE, int BS, int GEO>
static void launch_block_geo(hipStream_t s, size_t lds, Topo t, BlockTopo bt, Geom<S> g, Phys ph,
                             const S *P, const S *GP, S *Pout, StageArgs a)
{
    static size_t granted = 64 * 1024;  // dynamic LDS above 64 KiB has to be requested
    if (lds > granted &&
        hipFuncSetAttribute((const void *)k_flux_block<S, FLUX, VISC, ORDER, MODE, BS, GEO>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess) {
        granted = lds;
    }
    hipLaunchKernelGGL((k_flux_block<S, FLUX, VISC, ORDER, MODE, BS, GEO>),
                       dim3((unsigned)((t.ni + BS - 1) / BS)), dim3(BS), lds, s, t, bt, g, ph, P,
                       GP, Pout, a);
}

// BS rows (and threads) per workgroup: the block layout the face lists were built for.  FP64
// with viscous and convective records and rebuilt tangents: the instance that loads the next
// face's geometry records ahead.
template <typename S, int FLUX, int VISC, int ORDER, int MODE, int BS>
static void launch_block(hipStream_t s, size_t lds, Topo t, BlockTopo bt, Geom<S> g, Phys ph,
                         const S *P, const S *GP, S *Pout, StageArgs a)
{
    if constexpr (sizeof(S) == 8 && VISC && FLUX != FLUX_NONE) {
        if (g.pairs && g.tangents) {
            launch_block_geo<S, FLUX, VISC, ORDER, MODE, BS, 1>(s, lds, t, bt, g, ph, P, GP, Pout, a);
            return;
        }
    }
    launch_block_geo<S, FLUX, VISC, ORDER, MODE, BS, 0>(s, lds, t, bt, g, ph, P, GP, Pout, a);
}

template <typename S, int FLUX, int VISC, int ORDER, int BS>
static BlockLaunch<S> block_mode(int mode)
{
    switch (mode) {
        case 0: return launch_block<S, FLUX, VISC, ORDER, 0, BS>;
        case 1: return launch_block<S, FLUX, VISC, ORDER, 1, BS>;
        default: return launch_block<S, FLUX, VISC, ORDER, 2, BS>;
    }
}

template <typename S, int FLUX, int BS>
static BlockLaunch<S> block_row(int visc, int order, int mode)
{
    if (order == 1) {
        return block_mode<S, FLUX, 0, 1, BS>(mode);
    }
    return visc ? block_mode<S, FLUX, 1, 2, BS>(mode) : block_mode<S, FLUX, 0, 2, BS>(mode);
}

template <typename S, int BS>
static BlockLaunch<S> block_table_bs(int flux, int visc, int order, int mode)
{
    switch (flux) {
        case FLUX_NONE: return block_row<S, FLUX_NONE, BS>(visc, order, mode);
        case FLUX_GODUNOV: return block_row<S, FLUX_GODUNOV, BS>(visc, order, mode);
        case FLUX_ROE: return block_row<S, FLUX_ROE, BS>(visc, order, mode);
        case FLUX_HLL: return block_row<S, FLUX_HLL, BS>(visc, order, mode);
        case FLUX_HLLC: return block_row<S, FLUX_HLLC, BS>(visc, order, mode);
        case FLUX_HLLE: return block_row<S, FLUX_HLLE, BS>(visc, order, mode);
        case FLUX_AUSMD: return block_row<S, FLUX_AUSMD, BS>(visc, order, mode);
        case FLUX_AUSMDV: return block_row<S, FLUX_AUSMDV, BS>(visc, order, mode);
        case FLUX_LXF: return block_row<S, FLUX_LXF, BS>(visc, order, mode);
        case FLUX_CENTRAL: return block_row<S, FLUX_CENTRAL, BS>(visc, order, mode);
        default: return nullptr;
    }
}

template <typename S>
static BlockLaunch<S> block_table(int flux, int visc, int order, int mode, int bs)
{
    if (bs == BLOCK) {
        return block_table_bs<S, BLOCK>(flux, visc, order, mode);
    }
    if (bs == SMALL_BLOCK) {
        return block_table_bs<S, SMALL_BLOCK>(flux, visc, order, mode);
    }
    return nullptr;
}

}  // namespace fv

template <typename S>
void Kern<S>::boundary(hipStream_t s, int num, int nf, const DevData<S> &d, double gamma)
{
    hipLaunchKernelGGL(fv::k_boundary<S>, dim3(fv::grid_of(num)), dim3(fv::BLOCK), 0, s, num, nf,
                       (const int *)d.d_bleft, (const int *)d.d_bright, (const int *)d.d_bface,
                       (const int *)d.d_bkind, (const S *)d.d_bctx, (const S *)d.d_basis, d.d_P,
                       d.d_P2, gamma);
}

template <typename S>
int Kern<S>::flux_block(hipStream_t s, int flux, int visc, int order, int mode, int bs, size_t lds,
                        fv::Topo t, fv::BlockTopo bt, fv::Geom<S> g, fv::Phys ph, const S *P,
                        const S *GP, S *Pout, fv::StageArgs a)
{
    fv::BlockLaunch<S> fn = fv::block_table<S>(flux, visc, order, mode, bs);
    if (!fn) {
        teal_gpu_set_error("no device kernel for flux %d (viscous %d, order %d, %d-row blocks)", flux,
                           visc, order, bs);
        return -1;
    }
    fn(s, lds, t, bt, g, ph, P, GP, Pout, a);
    return 0;
}

template <typename S, int WK, int GB, int ST>
static void grad_limit_wk(hipStream_t s, int lim, int mode, fv::Topo t, fv::Geom<S> g, const S *P,
                          S *GP)
{
    dim3 grid((unsigned)((t.ni + GB - 1) / GB)), block(GB);
    if (mode == 0) {
        switch (lim) {
            case 0: hipLaunchKernelGGL((fv::k_grad_limit<S, 0, 0, WK, GB, ST>), grid, block, 0, s, t, g, P, GP); return;
            case 1: hipLaunchKernelGGL((fv::k_grad_limit<S, 1, 0, WK, GB, ST>), grid, block, 0, s, t, g, P, GP); return;
            default: hipLaunchKernelGGL((fv::k_grad_limit<S, 2, 0, WK, GB, ST>), grid, block, 0, s, t, g, P, GP); return;
        }
    }
    if (mode == 2) {
        switch (lim) {
            case 0: hipLaunchKernelGGL((fv::k_grad_limit<S, 0, 2, WK, GB, ST>), grid, block, 0, s, t, g, P, GP); return;
            case 1: hipLaunchKernelGGL((fv::k_grad_limit<S, 1, 2, WK, GB, ST>), grid, block, 0, s, t, g, P, GP); return;
            default: hipLaunchKernelGGL((fv::k_grad_limit<S, 2, 2, WK, GB, ST>), grid, block, 0, s, t, g, P, GP); return;
        }
    }
    switch (lim) {
        case 1: hipLaunchKernelGGL((fv::k_grad_limit<S, 1, 1, WK, GB, ST>), grid, block, 0, s, t, g, P, GP); return;
        default: hipLaunchKernelGGL((fv::k_grad_limit<S, 2, 1, WK, GB, ST>), grid, block, 0, s, t, g, P, GP); return;
    }
}

template <typename S, int WK, int ST>
static void grad_limit_gb(hipStream_t s, int gb, int lim, int mode, fv::Topo t, fv::Geom<S> g,
                          const S *P, S *GP)
{
    switch (gb) {
        case 64: grad_limit_wk<S, WK, 64, ST>(s, lim, mode, t, g, P, GP); return;
        case 128: grad_limit_wk<S, WK, 128, ST>(s, lim, mode, t, g, P, GP); return;
        default: grad_limit_wk<S, WK, fv::BLOCK, ST>(s, lim, mode, t, g, P, GP); return;
    }
}

// gb: threads per workgroup (64, 128 or BLOCK); stage: neighbour states staged in LDS (both
// autotuned)
template <typename S>
void Kern<S>::grad_limit(hipStream_t s, int lim, int mode, int gb, int stage, fv::Topo t,
                         fv::Geom<S> g, const S *P, S *GP)
{
    if (t.wmax <= 4) {
        if (stage) {
            grad_limit_gb<S, 4, 1>(s, gb, lim, mode, t, g, P, GP);
        }
        else {
            grad_limit_gb<S, 4, 0>(s, gb, lim, mode, t, g, P, GP);
        }
    }
    else if (stage) {
        grad_limit_gb<S, fv::W, 1>(s, gb, lim, mode, t, g, P, GP);
    }
    else {
        grad_limit_gb<S, fv::W, 0>(s, gb, lim, mode, t, g, P, GP);
    }
}

template <typename S>
int Kern<S>::face(hipStream_t s, int flux, int visc, int order, fv::Topo t, fv::Geom<S> g,
                  fv::Phys ph, const S *P, const S *GP, S *F, int *error)
{
    fv::FaceLaunch<S> fn = fv::face_table<S>(flux, visc, order);
    if (!fn) {
        teal_gpu_set_error("no device kernel for flux %d (viscous %d, order %d)", flux, visc, order);
        return -1;
    }
    fn(s, t, g, ph, P, GP, F, error);
    return 0;
}

template <typename S>
void Kern<S>::accum(hipStream_t s, int visc, int conv, int mode, fv::Topo t, fv::Geom<S> g,
                    fv::Phys ph, const S *F, S *P, fv::StageArgs a)
{
    if (visc && conv) {
        fv::accum_row<S, 1, 1>(mode, s, t, g, ph, F, P, a);
    }
    else if (visc) {
        fv::accum_row<S, 1, 0>(mode, s, t, g, ph, F, P, a);
    }
    else if (conv) {
        fv::accum_row<S, 0, 1>(mode, s, t, g, ph, F, P, a);
    }
    else {
        fv::accum_row<S, 0, 0>(mode, s, t, g, ph, F, P, a);
    }
}

template <typename S>
void Kern<S>::timestep_u0(hipStream_t s, fv::Topo t, fv::Geom<S> g, fv::Phys ph, const S *P,
                          double *U0, double *dtpart, int write_u0)
{
    hipLaunchKernelGGL(fv::k_timestep_u0<S>, dim3(fv::grid_of(t.ni)), dim3(fv::BLOCK), 0, s, t, g,
                       ph, P, U0, dtpart, write_u0);
}

template <typename S>
void Kern<S>::to_device(hipStream_t s, const double *in, S *out, int num, int in_stride, int width,
                        int out_stride, const int *perm, int soa)
{
    if (soa) {
        hipLaunchKernelGGL((fv::k_to_device<S, 1>), dim3(fv::grid_of(num)), dim3(fv::BLOCK), 0, s, in,
                           out, num, in_stride, width, out_stride, perm);
    }
    else {
        hipLaunchKernelGGL((fv::k_to_device<S, 0>), dim3(fv::grid_of(num)), dim3(fv::BLOCK), 0, s, in,
                           out, num, in_stride, width, out_stride, perm);
    }
}

template <typename S>
void Kern<S>::to_host(hipStream_t s, const S *in, double *out, int num, int in_stride, int width,
                      int out_stride, const int *perm, int soa)
{
    if (soa) {
        hipLaunchKernelGGL((fv::k_to_host<S, 1>), dim3(fv::grid_of(num)), dim3(fv::BLOCK), 0, s, in,
                           out, num, in_stride, width, out_stride, perm);
    }
    else {
        hipLaunchKernelGGL((fv::k_to_host<S, 0>), dim3(fv::grid_of(num)), dim3(fv::BLOCK), 0, s, in,
                           out, num, in_stride, width, out_stride, perm);
    }
}
