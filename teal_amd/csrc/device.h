// device.h -- internal declarations of the GPU module (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "kernels.h"
#include "teal_gpu.h"

enum { VENKATAKRISHNAN_ID = 2 };

void teal_gpu_set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

struct TealGpu;

// face lists of the fused stage kernel for one block layout (fv::BlockTopo): blocks of `rows`
// device rows, each with the faces touching it, their geometry in list order
template <typename S>
struct BlockSet {
    int rows = 0;  // device rows per block (0: not built)
    int *d_boff = nullptr, *d_bslot = nullptr;
    int2 *d_bflr = nullptr;
    S *d_bpairs = nullptr;  // face records in block-list order as pairs (Geom::pairs)
    S *d_barea = nullptr, *d_bbasis = nullptr, *d_boffl = nullptr, *d_boffr = nullptr,
      *d_bcorr = nullptr;  // face geometry in block-list order (stride num_block_faces)
    S *d_bfcen = nullptr;  // centre form: face centres in block-list order
    int *d_bcrow = nullptr;  // centre form: right-cell centre row per list entry
    bool centre_form = false;
    bool tangents = false;  // FP64 centre form: tangents rebuilt from the normal (Geom::tangents;
                            // TEAL_GPU_TANGENTS_STORED=1 at creation keeps the stored basis)
    int num_block_faces = 0;
    int max_block_faces = 0;
};

// device arrays of one precision (see kernels.h for the layout)
template <typename S>
struct DevData {
    int *d_inc = nullptr, *d_nbr = nullptr;
    int2 *d_flr = nullptr;
    S *d_w = nullptr, *d_vol = nullptr, *d_proj = nullptr;
    S *d_area = nullptr, *d_basis = nullptr, *d_offl = nullptr, *d_offr = nullptr,
      *d_corr = nullptr;
    double *d_vol64 = nullptr;
    int *d_bleft = nullptr, *d_bright = nullptr, *d_bface = nullptr, *d_bkind = nullptr;
    S *d_bctx = nullptr;
    S *d_P = nullptr;    // [num_cells][5] current state
    S *d_P2 = nullptr;   // [num_cells][5] the other state buffer of the fused stage kernel
    S *d_GP = nullptr;   // [num_inner][20]
    S *d_F = nullptr;    // [num_faces][10] (split path only)
    S *d_eps2 = nullptr;
    // the two block layouts of the fused kernel: lay[0] blocks of fv::BLOCK rows, lay[1] of
    // fv::SMALL_BLOCK; `cur` is the one the residual passes use (autotuned)
    BlockSet<S> lay[2];
    int cur = 0;
    S *d_ccen = nullptr;   // centre form: cell centres [num_cells][3], device rows
    int max_inc = fv::W;  // most incidences of any inner cell
    int fused = 0;  // 1: residual passes use k_flux_block (its face lists fit in LDS and the
                    // environment does not set TEAL_GPU_SPLIT=1)

    int build(const Mesh *mesh, TealGpu *gpu);
    int build_blocks(BlockSet<S> &b, int rows, const std::vector<int> &inc, int ni,
                     const std::vector<int2> &flr, const std::vector<double> &area,
                     const std::vector<double> &basis, const std::vector<double> &offl,
                     const std::vector<double> &offr, const std::vector<double> &corr,
                     const std::vector<double> &fcen, const std::vector<int> &crow,
                     const std::vector<double> &ccen);
    void release();
    fv::Geom<S> geom() const;
    fv::BlockTopo block_topo() const;
    fv::Geom<S> block_geom() const;  // cell geometry + face geometry in block-list order
    int block_rows() const { return lay[cur].rows; }
    // dynamic LDS of the fused stage kernel: the face records of the largest block list
    // (values per face as FaceRec: 10 for FP64 with viscous and convective fluxes, else 5)
    size_t block_lds(int visc = 1, int conv = 1) const
    {
        const int nv = (sizeof(S) == 8 && visc && conv) ? 10 : 5;
        return sizeof(S) * nv * (size_t)lay[cur].max_block_faces;
    }
};

// kernel launchers of one precision; instantiated in device.hip (double) and mixed.hip (float)
template <typename S>
struct Kern {
    static void boundary(hipStream_t s, int num, int nf, const DevData<S> &d, double gamma);
    static void grad_limit(hipStream_t s, int lim, int mode, int gb, int stage, fv::Topo t,
                           fv::Geom<S> g, const S *P, S *GP);
    static int face(hipStream_t s, int flux, int visc, int order, fv::Topo t, fv::Geom<S> g,
                    fv::Phys ph, const S *P, const S *GP, S *F, int *error);
    static void accum(hipStream_t s, int visc, int conv, int mode, fv::Topo t, fv::Geom<S> g,
                      fv::Phys ph, const S *F, S *P, fv::StageArgs a);
    static int flux_block(hipStream_t s, int flux, int visc, int order, int mode, int bs,
                          size_t lds, fv::Topo t, fv::BlockTopo bt, fv::Geom<S> g, fv::Phys ph,
                          const S *P, const S *GP, S *Pout, fv::StageArgs a);
    static void timestep_u0(hipStream_t s, fv::Topo t, fv::Geom<S> g, fv::Phys ph,
                            const S *P, double *U0, double *dtpart, int write_u0);
    // host layout (reference numbering) <-> device layout (device numbering);
    // perm[d] = reference index of device row d (null: identity)
    static void to_device(hipStream_t s, const double *in, S *out, int num, int in_stride,
                          int width, int out_stride, const int *perm, int soa);
    static void to_host(hipStream_t s, const S *in, double *out, int num, int in_stride,
                        int width, int out_stride, const int *perm, int soa);
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
    const Mesh *mesh = nullptr;  // the mesh, or the union of a batch's meshes
    int precision = 0;
    int ni = 0, nc = 0, nf = 0, nfi = 0, nfb = 0;  // device rows (inner rows include padding)
    int mesh_ni = 0;                               // inner cells of the mesh
    double sum_volume = 0;
    // independent instances (a batch): instance i owns mesh inner cells
    // [inst_cell_off[i], inst_cell_off[i+1]) and device blocks [inst_blk_off[i], ...[i+1]);
    // its device rows are padded to whole blocks (padding rows have no faces, volume 1 and a
    // resting state, so they add nothing to the dt minimum or the residual)
    int ninst = 1;
    Mesh *owned = nullptr;  // the union mesh of a batch, destroyed with the handle
    std::vector<int> inst_cell_off, inst_blk_off;
    std::vector<double> inst_volume;
    // host rows of instance i in the union: inner, boundary-ghost and periodic-ghost bases and
    // counts, inst_rows[6 * i + 0..5]
    std::vector<int> inst_rows;
    // partial sums (dt minima, residual sums) are kept per group of 64 device rows, whatever
    // the block layout: instance i owns groups [inst_grp_off[i], inst_grp_off[i+1])
    int *d_grp_inst = nullptr;     // instance of each group
    int *d_inst_grp_off = nullptr;
    int *d_all_grp = nullptr;      // {0, ngroups}: every group as one range
    double *d_inst_volume = nullptr;
    std::vector<int> periodic_src;
    std::vector<int> cell_perm, cell_inv, face_perm, face_inv;  // device <-> reference numbering
    int *d_cell_perm = nullptr;
    DevData<double> d64;
    DevData<float> d32;
    TealGpuConfig cfg{};
    std::vector<int> cfg_bc;        // the configuration's per-entity arrays, kept for mirrors
    std::vector<double> cfg_ctx;    // built later (the FP64 mirror of a mixed TealGpu)
    std::vector<double> cfg_eps2;
    int lds_max = 0;                // dynamic LDS a workgroup may use
    int configured = 0;
    int u0_valid = 0;  // d_U0 holds the conserved state of the current step start
    int dt_valid = 0;  // the scalars hold the dt minimum of that state
    int nblocks = 0;   // blocks of fv::BLOCK rows
    int ngroups = 0;   // groups of 64 rows
    int grad_block = fv::BLOCK;  // threads per workgroup of the gradient kernel (autotuned)
    int grad_stage = 0;          // gradient kernel stages neighbour states in LDS (autotuned;
                                 // by default on meshes of tetrahedra)
    hipStream_t stream = nullptr;
    hipEvent_t ev_wait = nullptr;  // host waits on the stream by polling this event
    double *d_stage = nullptr;
    double *d_U0 = nullptr;
    double *d_R = nullptr;
    double *d_partials = nullptr;  // [ngroups][5] residual partial sums
    double *d_dtpart = nullptr;    // [ngroups] time-step minima
    double *d_chunk = nullptr;     // chunk results of the per-step reductions
    unsigned *d_tickets = nullptr; // [ninst] arrivals of a reduction's workgroups, 0 between launches
    int max_inst_groups = 0;       // groups of the largest instance
    fv::Scalars *d_scalars = nullptr;
    fv::Scalars *h_scalars = nullptr;
    int timing = 0;
    EventPool ev_flux, ev_grad, ev_face;
    // graph replay of explicit steps: a captured pair of steps (an even number of stage buffer
    // swaps, so every replay starts from the buffers it was captured with), valid for the key
    int graphs = 0;
    hipGraphExec_t graph = nullptr;
    struct {
        double courant;
        const double *alpha;
        int stages;
        const void *P;
    } graph_key{};
    Krylov *krylov = nullptr;
    int serial_reductions = 0;  // TEAL_REDUCE_SERIAL: GMRES/Newton dot products, entity averages
                                // and norms in reference order
    // source term (teal_gpu_set_source): its kind, the inner entity its average runs over,
    // and {average (5), velocity, relaxation} on the device for the stage epilogue
    int source = 0;
    int source_entity = 0;
    double *d_forcing = nullptr;
    // diagnostics (diagnostics.hip): partial sums, the device row of each mesh inner cell
    // (reference order) and the last result
    double2 *d_cellpart = nullptr;
    int *d_cell_inv = nullptr;
    double *d_diag = nullptr;

    fv::Topo topo() const;
    fv::Phys phys() const;
};

// residual pass helpers shared with krylov.hip
template <typename S>
int teal_residual_pass(TealGpu *g, DevData<S> &d, int mode, double alpha, int last);
template <typename S>
int teal_timestep_pass(TealGpu *g, DevData<S> &d, int write_u0, double *out);
void teal_krylov_release(TealGpu *g);
// elements per thread of the fused GMRES column kernel in use, 0: separate kernels (or no
// implicit step yet)
int teal_krylov_column_ept(const TealGpu *g);
void teal_diagnostics_release(TealGpu *g);
// average of the forcing source's entity over the state P into d_forcing (no-op without a
// source): the Prepare step of every residual evaluation
template <typename S>
int teal_prepare_source(TealGpu *g, const S *P);
// host primitive (reference layout, num_cells x 5) -> the resident state; boundary-ghost rows
// of the resident state -> the host array
int teal_upload_state(TealGpu *g, const double *primitive);
int teal_download_ghosts(TealGpu *g, double *primitive);
// wait for the handle's stream by polling an event recorded on it (a blocking
// synchronization costs tens of microseconds of wake-up per call; the Krylov solver and the
// single-step entry points synchronize every iteration)
int teal_stream_wait(TealGpu *g);
// the exact Riemann solver's error flag of one entry-point call: cleared at its start,
// read back and turned into an error before it returns (no-ops unless the flux is Godunov)
int teal_riemann_clear(TealGpu *g);
int teal_riemann_check(TealGpu *g);
// poll an event until the work before it has completed
int teal_event_wait(hipEvent_t ev);
// residual = sqrt(sum V R^2 / sum V) of a device derivative [5][ni] (src/equations/residual.c)
int teal_residual_norm(TealGpu *g, const double *d_derivative, double *residual);
void teal_device_numbering(TealGpu *g, const Mesh *m);
// FP64 device arrays and configuration on a mixed TealGpu (its implicit solver evaluates the
// nonlinear residual in FP64); no-op on the FP64 path or when already built
int teal_fp64_mirror(TealGpu *g);
