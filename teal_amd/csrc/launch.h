// launch.h -- kernel launchers and the compile-time flux dispatch table.
// Included by device.hip (instantiates Kern<double>) and mixed.hip (Kern<float>).
#pragma once

#include "device.h"

namespace fv {

template <typename S, int FLUX, int VISC, int ORDER, int MODE>
static void launch_flux(hipStream_t s, int grid, Topo t, Geom<S> g, Phys ph, const S *P,
                        const S *G, const S *psi, S *Pout, StageArgs a)
{
    hipLaunchKernelGGL((k_flux<S, FLUX, VISC, ORDER, MODE>), dim3(grid), dim3(BLOCK), 0, s, t, g,
                       ph, P, G, psi, Pout, a);
}

template <typename S>
using FluxLaunch = void (*)(hipStream_t, int, Topo, Geom<S>, Phys, const S *, const S *,
                            const S *, S *, StageArgs);

// one kernel per (flux, viscous, order, mode): the flux is a compile-time choice, as the
// reference's name table (src/euler/convective.c:593-623) is a link-time one
template <typename S, int FLUX>
struct FluxRow {
    static FluxLaunch<S> get(int visc, int order, int mode)
    {
        if (order == 1) {
            switch (mode) {
                case 0: return launch_flux<S, FLUX, 0, 1, 0>;
                case 1: return launch_flux<S, FLUX, 0, 1, 1>;
                default: return launch_flux<S, FLUX, 0, 1, 2>;
            }
        }
        if (visc) {
            switch (mode) {
                case 0: return launch_flux<S, FLUX, 1, 2, 0>;
                case 1: return launch_flux<S, FLUX, 1, 2, 1>;
                default: return launch_flux<S, FLUX, 1, 2, 2>;
            }
        }
        switch (mode) {
            case 0: return launch_flux<S, FLUX, 0, 2, 0>;
            case 1: return launch_flux<S, FLUX, 0, 2, 1>;
            default: return launch_flux<S, FLUX, 0, 2, 2>;
        }
    }
};

template <typename S>
static FluxLaunch<S> flux_table(int flux, int visc, int order, int mode)
{
    switch (flux) {
        case FLUX_NONE: return FluxRow<S, FLUX_NONE>::get(visc, order, mode);
        case FLUX_GODUNOV: return FluxRow<S, FLUX_GODUNOV>::get(visc, order, mode);
        case FLUX_ROE: return FluxRow<S, FLUX_ROE>::get(visc, order, mode);
        case FLUX_HLL: return FluxRow<S, FLUX_HLL>::get(visc, order, mode);
        case FLUX_HLLC: return FluxRow<S, FLUX_HLLC>::get(visc, order, mode);
        case FLUX_HLLE: return FluxRow<S, FLUX_HLLE>::get(visc, order, mode);
        case FLUX_AUSMD: return FluxRow<S, FLUX_AUSMD>::get(visc, order, mode);
        case FLUX_AUSMDV: return FluxRow<S, FLUX_AUSMDV>::get(visc, order, mode);
        case FLUX_LXF: return FluxRow<S, FLUX_LXF>::get(visc, order, mode);
        case FLUX_CENTRAL: return FluxRow<S, FLUX_CENTRAL>::get(visc, order, mode);
        default: return nullptr;
    }
}

}  // namespace fv

template <typename S>
void Kern<S>::boundary(hipStream_t s, int num, int nc, int nf, const DevData<S> &d, S *P,
                       double gamma)
{
    hipLaunchKernelGGL(fv::k_boundary<S>, dim3((num + fv::BLOCK - 1) / fv::BLOCK),
                       dim3(fv::BLOCK), 0, s, num, nc, nf, (const int *)d.d_bleft,
                       (const int *)d.d_bright, (const int *)d.d_bface, (const int *)d.d_bkind,
                       (const S *)d.d_bctx, (const S *)d.d_basis, P, gamma);
}

template <typename S>
void Kern<S>::grad_limit(hipStream_t s, int lim, int mode, fv::Topo t, fv::Geom<S> g,
                         const S *P, S *G, S *psi)
{
    int grid = (t.ni + fv::BLOCK - 1) / fv::BLOCK;
    if (mode == 0) {
        switch (lim) {
            case 0: hipLaunchKernelGGL((fv::k_grad_limit<S, 0, 0>), dim3(grid), dim3(fv::BLOCK), 0, s, t, g, P, G, psi); return;
            case 1: hipLaunchKernelGGL((fv::k_grad_limit<S, 1, 0>), dim3(grid), dim3(fv::BLOCK), 0, s, t, g, P, G, psi); return;
            default: hipLaunchKernelGGL((fv::k_grad_limit<S, 2, 0>), dim3(grid), dim3(fv::BLOCK), 0, s, t, g, P, G, psi); return;
        }
    }
    switch (lim) {
        case 1: hipLaunchKernelGGL((fv::k_grad_limit<S, 1, 1>), dim3(grid), dim3(fv::BLOCK), 0, s, t, g, P, G, psi); return;
        default: hipLaunchKernelGGL((fv::k_grad_limit<S, 2, 1>), dim3(grid), dim3(fv::BLOCK), 0, s, t, g, P, G, psi); return;
    }
}

template <typename S>
int Kern<S>::flux(hipStream_t s, int flux, int visc, int order, int mode, fv::Topo t,
                  fv::Geom<S> g, fv::Phys ph, const S *P, const S *G, const S *psi, S *Pout,
                  fv::StageArgs a)
{
    fv::FluxLaunch<S> fn = fv::flux_table<S>(flux, visc, order, mode);
    if (!fn) {
        teal_gpu_set_error("no device kernel for flux %d (viscous %d, order %d)", flux, visc, order);
        return -1;
    }
    fn(s, (t.ni + fv::BLOCK - 1) / fv::BLOCK, t, g, ph, P, G, psi, Pout, a);
    return 0;
}

template <typename S>
void Kern<S>::timestep_u0(hipStream_t s, fv::Topo t, fv::Geom<S> g, fv::Phys ph, const S *P,
                          double *U0, fv::Scalars *sc, int write_u0)
{
    hipLaunchKernelGGL(fv::k_timestep_u0<S>, dim3((t.ni + fv::BLOCK - 1) / fv::BLOCK),
                       dim3(fv::BLOCK), 0, s, t, g, ph, P, U0, sc, write_u0);
}

template <typename S>
void Kern<S>::aos_to_soa(hipStream_t s, const double *in, S *out, int num, int stride, int width,
                         int out_stride)
{
    hipLaunchKernelGGL(fv::k_aos_to_soa<S>, dim3((num + fv::BLOCK - 1) / fv::BLOCK),
                       dim3(fv::BLOCK), 0, s, in, out, num, stride, width, out_stride);
}

template <typename S>
void Kern<S>::soa_to_aos(hipStream_t s, const S *in, double *out, int num, int stride, int width,
                         int in_stride)
{
    hipLaunchKernelGGL(fv::k_soa_to_aos<S>, dim3((num + fv::BLOCK - 1) / fv::BLOCK),
                       dim3(fv::BLOCK), 0, s, in, out, num, stride, width, in_stride);
}
