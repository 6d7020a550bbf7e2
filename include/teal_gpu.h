// teal_gpu.h -- C-ABI of the MI355X residual path (plain pointers and sizes, no torch types).
//
// These entry points sit behind the reference's residual interface: each one replaces the
// CPU routine cited next to it and takes the same arrays in the reference's host layouts
// (primitive num_cells x 5 AoS, derivative num_inner x 5, gradient num_cells x 5 x 3).
// The teal API in equations.h / simulation.h is implemented on top of them; tests and
// bench.py call them directly through ctypes.
//
// Errors: every entry point returns 0 (or a finite value) on success; on failure it returns
// nonzero (or NaN) and teal_gpu_last_error() describes the failure.  Calls on one TealGpu
// are serialized on its own HIP stream and are not thread-safe (like the reference).
#pragma once

#include "mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct TealGpu TealGpu;

// Disjoint union of generated meshes (no counterpart in the reference, which advances one mesh
// per program): cells [inner | boundary ghosts | periodic ghosts] and faces [inner | boundary |
// periodic], each group in instance order; entities [inner | boundary | periodic] likewise;
// periodic neighbour tags made distinct per instance.  Geometry is copied, indices renumbered,
// so each instance's part is bit-identical to the instance.  Free with mesh_destroy.
Mesh *teal_mesh_union(const Mesh *const *meshes, int num);

enum TealPrecision {
    TEAL_FP64 = 0,  // FP64 storage and arithmetic: bit-compatible with the reference
    TEAL_MIXED = 1, // FP32 states/gradients/geometry/flux math, FP64 face accumulation,
                    // FP64 solution update, dt and residual norm
};

// flux ids follow the reference's name table (src/euler/convective.c:593-623)
enum TealFlux {
    TEAL_FLUX_NONE = 0,
    TEAL_FLUX_GODUNOV,
    TEAL_FLUX_ROE,
    TEAL_FLUX_HLL,
    TEAL_FLUX_HLLC,
    TEAL_FLUX_HLLE,
    TEAL_FLUX_AUSMD,
    TEAL_FLUX_AUSMDV,
    TEAL_FLUX_LXF,
    TEAL_FLUX_CENTRAL,
};

// boundary ids follow src/euler/boundary.c:273-297 and src/navierstokes/boundary.c:20-26
enum TealBoundary {
    TEAL_BC_NONE = 0,
    TEAL_BC_SYMMETRY,
    TEAL_BC_SUPERSONIC_INFLOW,
    TEAL_BC_SUPERSONIC_OUTFLOW,
    TEAL_BC_SUBSONIC_INFLOW,
    TEAL_BC_SUBSONIC_OUTFLOW,
    TEAL_BC_PRESSURE_OUTFLOW,
    TEAL_BC_FARFIELD,
    TEAL_BC_ADIABATIC_WALL,
};

typedef struct {
    int space_order;           // 1 or 2 (src/equations.h:134)
    int flux;                  // TealFlux (src/equations.h:137)
    int viscous;               // 1: central viscous flux (src/navierstokes/viscous.c:11-39)
    int limiter;               // 0 none, 1 Barth-Jespersen, 2 Venkatakrishnan (limiter.h:5-9)
    int navierstokes;          // 1: navierstokes_timestep, 0: euler_timestep
    double gamma, mu, prandtl; // properties (src/euler.h:19-21, src/navierstokes.h:19-23)
    int num_boundary;          // entities.off_boundary - entities.num_inner
    const int *bc;             // [num_boundary] TealBoundary per boundary entity
    const double *bc_context;  // [num_boundary][5] primitive context state per entity
    const double *eps2;        // [num_inner] Venkatakrishnan parameter (limiter.c:8-21), or
                               // NULL: computed here as (coefficient cbrt(V))^3
    double coefficient;        // Venkatakrishnan K (used when eps2 is NULL)
} TealGpuConfig;

// Upload a generated mesh (src/mesh.h:13-72 layout) into the device layout.
// Replaces nothing in the reference (it has no device); called once per Equations.
TealGpu *teal_gpu_create(const Mesh *mesh, int precision);
void teal_gpu_destroy(TealGpu *gpu);
int teal_gpu_configure(TealGpu *gpu, const TealGpuConfig *config);
const char *teal_gpu_last_error(void);

// Host-pointer residual path.  Replaces, one for one:
int teal_gpu_boundary(TealGpu *gpu, double *primitive);          // src/equations/boundary.c:5
int teal_gpu_gradient(TealGpu *gpu, double *primitive, double *gradient); // gradient.c:7
int teal_gpu_limiter(TealGpu *gpu, const double *primitive, double *gradient); // limiter.c:115
int teal_gpu_derivative(TealGpu *gpu, double *primitive, double *derivative); // derivative.c:269
double teal_gpu_timestep(TealGpu *gpu, const double *primitive); // src/equations/timestep.c:7
int teal_gpu_residual(TealGpu *gpu, const double *derivative, double *residual); // residual.c:8
// No counterpart in the reference.  teal_gpu_residual for callers that cannot allocate (code
// under HIP graph capture, serving loops): it allocates nothing and does not synchronize.
// `derivative` is a DEVICE array in the reference layout (inner cells x 5 doubles), `residual`
// a DEVICE array of 5 doubles, written on `stream` (a hipStream_t; null: the handle's stream)
// with the same values as teal_gpu_residual.  Partial sums live in a fixed 10 MiB device
// array: meshes of more than 16,777,216 inner cells and batch handles are refused.
// Calls that overlap on one GPU race on that array.
int teal_gpu_residual_noalloc(TealGpu *gpu, const double *derivative, double *residual,
                              void *stream);

// Resident state: the device copy of the primitive variables is authoritative between
// teal_gpu_upload and teal_gpu_download.
int teal_gpu_upload(TealGpu *gpu, const double *primitive);
int teal_gpu_download(TealGpu *gpu, double *primitive);

// One step of the reference's Advance methods on the resident state; returns step0
// (src/simulation/advance.c:11 euler, :47 lserk, :119 implicit_euler).
double teal_gpu_euler(TealGpu *gpu, double *time, double *residual, double max_step,
                      double courant);
double teal_gpu_lserk(TealGpu *gpu, double *time, double *residual, double max_step,
                      double courant, int time_order, int num_stages);
double teal_gpu_implicit_euler(TealGpu *gpu, double *time, double *residual, double max_step,
                               double courant, double newton_tolerance,
                               double krylov_tolerance, int krylov_dimension);

// Dot products of implicit_euler / GMRES (src/core/sync.c:66-79 sums serially in cell
// order): TEAL_REDUCE_TREE (default) is a deterministic compensated tree reduction across the
// GPU; TEAL_REDUCE_SERIAL adds in the reference's order on one workgroup, reproducing the
// reference bit for bit (slow: for verification on small meshes).  No counterpart in the
// reference.
enum TealReduction { TEAL_REDUCE_TREE = 0, TEAL_REDUCE_SERIAL = 1 };
int teal_gpu_set_reductions(TealGpu *gpu, int mode);

// Source term on the device (no counterpart in the reference's library, whose sources are user
// callbacks, src/equations.h:27-29, 41): TEAL_SOURCE_FORCING is the bulk-velocity forcing of
// run/periodic_hill/run.c:8-21.  Before every residual evaluation its Prepare step averages the
// state being evaluated over inner entity `entity` (src/equations/average.c:8-41); the
// finalize step then adds s1 = rho (U - <u>) / tau to the x momentum and u s1 to the energy of
// each cell (src/equations/derivative.c:260-265), param = {U, tau}.  TEAL_SOURCE_NONE removes
// it.  Single-mesh handles only.
enum TealSource { TEAL_SOURCE_NONE = 0, TEAL_SOURCE_FORCING = 1 };
int teal_gpu_set_source(TealGpu *gpu, int kind, int entity, const double *param);

// Diagnostics as device reductions (single-mesh handles):
// teal_gpu_average -- src/equations/average.c:8-41: volume-weighted sum over inner entity
//   `entity` divided by the mesh volume; primitive = host array in the reference layout, or
//   null for the resident state.
// teal_gpu_norm -- src/equations/norm.c:8-38: sqrt(sum V (reference - primitive)^2 / sum V)
//   over the inner cells, both host arrays [num_inner][5] in the reference layout.
// teal_gpu_polar -- the face loop of src/euler/polar.c:36-53 on boundary entity `entity`:
//   the boundary pass on primitive (host array, its ghost rows updated as the reference's
//   boundary pass does; null: the resident state), then pressure_c[j] per face (may be null)
//   and the force coefficients coef[0..1] summed in face order; reference = the free-stream
//   primitive state, length = the reference length.
// The cell sums use the reduction mode of teal_gpu_set_reductions (serial: bit for bit).
int teal_gpu_average(TealGpu *gpu, int entity, const double *primitive, double *average);
int teal_gpu_norm(TealGpu *gpu, const double *primitive, const double *reference, double *norm);
int teal_gpu_polar(TealGpu *gpu, int entity, double *primitive, const double *reference,
                   double length, double *coef, double *pressure_c);

// Iteration counts of the last teal_gpu_implicit_euler call: stats[0] Newton iterations,
// stats[1] GMRES iterations (one matvec, i.e. one residual evaluation, each).  No counterpart
// in the reference, which does not report them.
int teal_gpu_krylov_stats(TealGpu *gpu, long *stats);

// `steps` lserk steps with no host synchronization between them (the inner loop of
// src/simulation/run.c:68-102 without output); the final time is returned in *time.
int teal_gpu_run_lserk(TealGpu *gpu, int steps, double *time, double courant, int time_order,
                       int num_stages);

// Batches (no counterpart in the reference, which advances one mesh per program): many
// independent generated meshes of any sizes advanced together, every kernel launch covering
// all of them, each instance with its own time step, time and residual.  The handle works
// on teal_mesh_union(meshes, num): teal_gpu_configure takes the union's boundary entities
// (each mesh's in turn), teal_gpu_upload/download and the per-call routines the union's
// layout.  teal_gpu_upload_batch / teal_gpu_download_batch take one primitive array per mesh
// in its own layout.  teal_gpu_run_lserk_batch runs `steps` lserk steps on every instance
// (time[i] in/out, residual[5 * i + v] of the last step when not null); the single-mesh
// Advance methods and teal_gpu_implicit_euler refuse a batch handle.  teal_gpu_residual on a
// batch writes 5 values per instance; teal_gpu_timestep returns the minimum over the union.
TealGpu *teal_gpu_create_batch(const Mesh *const *meshes, int num, int precision);
int teal_gpu_batch_size(TealGpu *gpu);
const Mesh *teal_gpu_mesh(TealGpu *gpu);  // the mesh the handle works on (a batch: the union)
int teal_gpu_upload_batch(TealGpu *gpu, const double *const *primitives);
int teal_gpu_download_batch(TealGpu *gpu, double *const *primitives);
int teal_gpu_run_lserk_batch(TealGpu *gpu, int steps, double *time, double *residual,
                             double courant, int time_order, int num_stages);

// Graph replay (no counterpart in the reference): with graphs enabled, teal_gpu_run_lserk
// and teal_gpu_run_lserk_batch capture two explicit steps into a HIP graph on first use and
// replay it, one launch per pair of steps; results are identical to eager launches.  The graph
// is recaptured when the configuration or the coefficients change.
int teal_gpu_set_graphs(TealGpu *gpu, int enable);

// Launch-shape autotuner (no counterpart in the reference): times `steps` lserk steps of the
// current state in each shape the handle supports and keeps the fastest, then restores the
// state, time and residual it started from.  ms[] receives per-step milliseconds (NaN where a
// shape is unavailable), TEAL_AUTOTUNE_SHAPES values:
//   ms[0], ms[1]  fused stage kernel on blocks of 256 device rows; eager, graph replay
//   ms[2], ms[3]  fused stage kernel on blocks of 128 rows (its own face lists); eager, graph
//   ms[4], ms[5]  face-flux + accumulation kernels; eager, graph replay
//   ms[6..8]      the chosen stage shape with gradient-kernel workgroups of 256, 128, 64
//   ms[9..11]     the same with the workgroup's own cell states staged in LDS for the
//                 neighbour gathers
// Every shape gives the same results, bit for bit in FP64.
enum { TEAL_AUTOTUNE_STAGE_SHAPES = 6, TEAL_AUTOTUNE_SHAPES = 12 };
// Sets a stage shape (the index of ms[0..5] above), a gradient workgroup size (64, 128 or 256)
// and whether the gradient kernel stages states in LDS (0 or 1; by default 1 on meshes of
// tetrahedra) directly; an error if the shape does not fit the device.
int teal_gpu_set_shape(TealGpu *gpu, int shape, int grad_block, int grad_stage);
int teal_gpu_autotune(TealGpu *gpu, int steps, double courant, int time_order, int num_stages,
                      double *ms);

// Per-kernel timing with HIP events recorded on the library's stream around each launch:
// teal_gpu_timing(gpu, capacity) preallocates events for `capacity` launches per kernel
// (0 stops timing); teal_gpu_timing_ms returns the mean duration in ms of kernel 0 (gradient
// + limiter), 1 (face fluxes; split path only) or 2 (the stage kernel: face fluxes in LDS +
// accumulation + RK update, or the accumulation alone on the split path) and the launch count.
// While timing is on, steps run eagerly (no graph replay).
int teal_gpu_timing(TealGpu *gpu, int capacity);
double teal_gpu_timing_ms(TealGpu *gpu, int kernel, int *launches);

// Device-layout facts for measurement: inner cells (device rows: a batch counts the padding
// that makes every instance whole 256-row blocks), cells, faces, inner faces, offset of the
// periodic faces, maximum faces per cell, whether residual passes run the fused stage kernel,
// and whether it reads the centre form of the face geometry (offsets and corrections
// recomputed from face and cell centres), and how the implicit solver forms a GMRES column.
// counts[13]: ni, nc, nf, nfi, nfb, W, fused, centre form, rows per fused block (0: split),
// gradient workgroup size, vector elements per thread of the one-launch GMRES column (0:
// separate kernels per Gram-Schmidt step, or no implicit step yet), gradient states staged
// in LDS, FP64 face tangents rebuilt from the normal (4 instead of 7 face-record pairs)
int teal_gpu_info(TealGpu *gpu, long *counts);

// Precision of the device mirror behind an Equations (no counterpart in the reference,
// whose path is FP64 only): TEAL_FP64 (the default) or TEAL_MIXED.  Takes effect at the
// next device call; the mirror is rebuilt when the precision changes.
struct Equations;
void teal_equations_set_precision(struct Equations *eqns, int precision);

// The bulk-velocity forcing source (run/periodic_hill/run.c:8-21) for the teal API: sets the
// equations' source to the pair teal_forcing_source / teal_forcing_prepare, which the device
// path runs as TEAL_SOURCE_FORCING with the average over inner entity `entity`, target
// velocity `velocity` and relaxation time `relaxation`.  The pair are ordinary callbacks with
// the reference signatures (src/equations.h:27-29, 41) for host callers; other source
// callbacks have no device counterpart and are refused (teal_error) when the device is used.
void teal_equations_set_forcing(struct Equations *eqns, const char *entity, double velocity,
                                double relaxation);
void teal_forcing_source(void *source, const double *property, Vector center, double time,
                         const void *context);
void teal_forcing_prepare(const struct Equations *eqns, const void *primitive);

// The configured device mirror of an Equations (created on first use; teal_error when one of
// its plugins has no device counterpart).  For callers that drive the C-ABI above directly.
TealGpu *teal_equations_device(const struct Equations *eqns);

#ifdef __cplusplus
}
#endif
