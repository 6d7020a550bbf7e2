// device.h -- internal declarations of the GPU module (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "kernels.h"
#include "teal_gpu.h"

enum { VENKATAKRISHNAN_ID = 2 };

void teal_gpu_set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

struct TealGpu;

// device arrays of one precision (see kernels.h for the layout)
template <typename S>
struct DevData {
    int *d_inc = nullptr, *d_nbr = nullptr;
    S *d_w = nullptr, *d_o = nullptr, *d_vol = nullptr, *d_proj = nullptr;
    S *d_area = nullptr, *d_basis = nullptr, *d_offl = nullptr, *d_offr = nullptr,
      *d_corr = nullptr;
    double *d_vol64 = nullptr;
    int *d_bleft = nullptr, *d_bright = nullptr, *d_bface = nullptr, *d_bkind = nullptr;
    S *d_bctx = nullptr;
    S *d_P[2] = {nullptr, nullptr};
    S *d_G = nullptr, *d_psi = nullptr, *d_eps2 = nullptr;

    int build(const Mesh *mesh, TealGpu *gpu);
    void release();
    fv::Geom<S> geom() const;
};

// kernel launchers of one precision; instantiated in device.hip (double) and mixed.hip (float)
template <typename S>
struct Kern {
    static void boundary(hipStream_t s, int num, int nc, int nf, const DevData<S> &d, S *P,
                         double gamma);
    static void grad_limit(hipStream_t s, int lim, int mode, fv::Topo t, fv::Geom<S> g,
                           const S *P, S *G, S *psi);
    static int flux(hipStream_t s, int flux, int visc, int order, int mode, fv::Topo t,
                    fv::Geom<S> g, fv::Phys ph, const S *P, const S *G, const S *psi, S *Pout,
                    fv::StageArgs a);
    static void timestep_u0(hipStream_t s, fv::Topo t, fv::Geom<S> g, fv::Phys ph,
                            const S *P, double *U0, fv::Scalars *sc, int write_u0);
    static void aos_to_soa(hipStream_t s, const double *in, S *out, int num, int stride, int width,
                           int out_stride);
    static void soa_to_aos(hipStream_t s, const S *in, double *out, int num, int stride, int width,
                           int in_stride);
};

// a preallocated pool of event pairs, recorded around launches of one kernel
struct EventPool {
    std::vector<hipEvent_t> start, stop;
    int used = 0;
    int reserve(int n);
    void release();
    bool begin(hipStream_t s);
    void end(hipStream_t s);
    double total_ms();
};

struct Krylov;

struct TealGpu {
    const Mesh *mesh = nullptr;
    int precision = 0;
    int ni = 0, nc = 0, nf = 0, nfi = 0, nfb = 0;
    double sum_volume = 0;
    std::vector<int> periodic_src;
    DevData<double> d64;
    DevData<float> d32;
    TealGpuConfig cfg{};
    int configured = 0;
    int u0_valid = 0;
    int cur = 0;
    int nblocks = 0;
    hipStream_t stream = nullptr;
    double *d_stage = nullptr;
    double *d_U0 = nullptr;
    double *d_R = nullptr;
    double *d_partials = nullptr;
    fv::Scalars *d_scalars = nullptr;
    fv::Scalars *h_scalars = nullptr;
    int timing = 0;
    EventPool ev_flux, ev_grad;
    Krylov *krylov = nullptr;

    fv::Topo topo() const;
    fv::Phys phys() const;
};

// residual pass helpers shared with krylov.hip
template <typename S>
int teal_residual_pass(TealGpu *g, DevData<S> &d, int cur, int mode, double alpha, int last);
template <typename S>
int teal_timestep_pass(TealGpu *g, DevData<S> &d, int write_u0, double *out);
void teal_krylov_release(TealGpu *g);
