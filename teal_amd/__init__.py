"""teal_amd -- MI355X residual path of the teal finite-volume solver.

The product is the C library `libteal_amd.so` (teal API of include/*.h + the device C-ABI
of include/teal_gpu.h, HIP kernels for gfx950).  This module only loads it through ctypes
and declares the signatures; it raises if the library is missing or has no GPU to run on.
"""
import ctypes as C
import os

import numpy as np

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libteal_amd.so")

D = C.c_double
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)
LP = C.POINTER(C.c_long)
VP = C.c_void_p

FP64, MIXED = 0, 1
SOURCE_NONE, SOURCE_FORCING = 0, 1
REDUCE_TREE, REDUCE_SERIAL = 0, 1
# the values teal_gpu_autotune reports (include/teal_gpu.h): six stage shapes, then the
# chosen one with three gradient-kernel workgroup sizes
AUTOTUNE_STAGE_SHAPES = ("fused256", "fused256+graph", "fused128", "fused128+graph", "split",
                         "split+graph")
AUTOTUNE_SHAPES = AUTOTUNE_STAGE_SHAPES + ("grad256", "grad128", "grad64", "grad256lds",
                                           "grad128lds", "grad64lds")
FLUX = {"none": 0, "godunov": 1, "roe": 2, "hll": 3, "hllc": 4, "hlle": 5, "ausmd": 6,
        "ausmdv": 7, "lxf": 8, "central": 9}
BC = {"symmetry": 1, "slipwall": 1, "supersonic inflow": 2, "supersonic outflow": 3,
      "subsonic inflow": 4, "subsonic outflow": 5, "pressure outflow": 6, "farfield": 7,
      "adiabatic wall": 8, "wall": 8}
LIMITER = {None: 0, "minmod": 1, "barth jespersen": 1, "venkatakrishnan": 2}


class Vector(C.Structure):
    _fields_ = [("x", D), ("y", D), ("z", D)]


class Triple(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


class GpuConfig(C.Structure):
    _fields_ = [("space_order", C.c_int), ("flux", C.c_int), ("viscous", C.c_int),
                ("limiter", C.c_int), ("navierstokes", C.c_int), ("gamma", D), ("mu", D),
                ("prandtl", D), ("num_boundary", C.c_int), ("bc", IP), ("bc_context", DP),
                ("eps2", DP), ("coefficient", D)]


def _sig(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


_lib = None


def lib():
    """The loaded product library (raises when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        l = C.CDLL(LIB_PATH)
        _sig(l, "mesh_create", VP, Vector, Vector, Triple, Triple)
        _sig(l, "mesh_create_tets", VP, C.c_int, D, C.c_ulonglong)
        _sig(l, "mesh_generate", None, VP)
        _sig(l, "mesh_read", VP, C.c_char_p)
        _sig(l, "mesh_validate", None, VP)
        _sig(l, "mesh_destroy", None, VP)
        _sig(l, "teal_gpu_create", VP, VP, C.c_int)
        _sig(l, "teal_gpu_destroy", None, VP)
        _sig(l, "teal_gpu_configure", C.c_int, VP, C.POINTER(GpuConfig))
        _sig(l, "teal_gpu_last_error", C.c_char_p)
        _sig(l, "teal_gpu_boundary", C.c_int, VP, DP)
        _sig(l, "teal_gpu_gradient", C.c_int, VP, DP, DP)
        _sig(l, "teal_gpu_limiter", C.c_int, VP, DP, DP)
        _sig(l, "teal_gpu_derivative", C.c_int, VP, DP, DP)
        _sig(l, "teal_gpu_timestep", D, VP, DP)
        _sig(l, "teal_gpu_residual", C.c_int, VP, DP, DP)
        _sig(l, "teal_gpu_residual_noalloc", C.c_int, VP, VP, VP, VP)
        _sig(l, "teal_gpu_upload", C.c_int, VP, DP)
        _sig(l, "teal_gpu_download", C.c_int, VP, DP)
        _sig(l, "teal_gpu_euler", D, VP, DP, DP, D, D)
        _sig(l, "teal_gpu_lserk", D, VP, DP, DP, D, D, C.c_int, C.c_int)
        _sig(l, "teal_gpu_implicit_euler", D, VP, DP, DP, D, D, D, D, C.c_int)
        _sig(l, "teal_gpu_run_lserk", C.c_int, VP, C.c_int, DP, D, C.c_int, C.c_int)
        _sig(l, "teal_gpu_set_reductions", C.c_int, VP, C.c_int)
        _sig(l, "teal_gpu_krylov_stats", C.c_int, VP, LP)
        _sig(l, "teal_gpu_timing", C.c_int, VP, C.c_int)
        _sig(l, "teal_gpu_timing_ms", D, VP, C.c_int, IP)
        _sig(l, "teal_gpu_info", C.c_int, VP, LP)
        _sig(l, "teal_mesh_union", VP, C.POINTER(VP), C.c_int)
        _sig(l, "teal_gpu_create_batch", VP, C.POINTER(VP), C.c_int, C.c_int)
        _sig(l, "teal_gpu_batch_size", C.c_int, VP)
        _sig(l, "teal_gpu_mesh", VP, VP)
        _sig(l, "teal_gpu_upload_batch", C.c_int, VP, C.POINTER(DP))
        _sig(l, "teal_gpu_download_batch", C.c_int, VP, C.POINTER(DP))
        _sig(l, "teal_gpu_run_lserk_batch", C.c_int, VP, C.c_int, DP, DP, D, C.c_int, C.c_int)
        _sig(l, "teal_gpu_set_graphs", C.c_int, VP, C.c_int)
        _sig(l, "teal_gpu_autotune", C.c_int, VP, C.c_int, D, C.c_int, C.c_int, DP)
        _sig(l, "teal_gpu_set_shape", C.c_int, VP, C.c_int, C.c_int, C.c_int)
        _sig(l, "teal_gpu_set_source", C.c_int, VP, C.c_int, C.c_int, DP)
        _sig(l, "teal_gpu_average", C.c_int, VP, C.c_int, DP, DP)
        _sig(l, "teal_gpu_norm", C.c_int, VP, DP, DP, DP)
        _sig(l, "teal_gpu_polar", C.c_int, VP, C.c_int, DP, DP, D, DP, DP)
        _lib = l
    return _lib


def last_error():
    return lib().teal_gpu_last_error().decode()


def check(rc, what="call"):
    if rc:
        raise RuntimeError(f"{what} failed: {last_error()}")


def dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(DP)


def mesh_box(n, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), periodic=(1, 1, 1)):
    """mesh_create + mesh_generate of an n[0] x n[1] x n[2] box (library mesh code)."""
    l = lib()
    m = l.mesh_create(Vector(*lo), Vector(*hi), Triple(*n), Triple(*periodic))
    l.mesh_generate(m)
    return m


class Gpu:
    """Thin owner of a TealGpu handle (device mirror of one mesh + configuration)."""

    def __init__(self, mesh_ptr, precision=FP64):
        self.lib = lib()
        self.mesh = mesh_ptr
        self.arrays = _abi.mesh_arrays(mesh_ptr)
        self.h = self.lib.teal_gpu_create(mesh_ptr, precision)
        if not self.h:
            raise RuntimeError(f"teal_gpu_create failed: {last_error()}")
        self.precision = precision

    def configure(self, space_order=2, flux="hllc", viscous=0, limiter="minmod", coefficient=0.0,
                  navierstokes=None, gamma=1.4, mu=0.0, prandtl=0.71, bcs=None):
        """bcs: {entity name: (condition name, context 5-tuple or None)}"""
        a = self.arrays
        names = a["entity_names"]
        first, off = a["entities_num_inner"], a["entities_off_boundary"]
        nb = off - first
        kind = (C.c_int * max(nb, 1))()
        ctx = (C.c_double * (5 * max(nb, 1)))()
        bcs = bcs or {}
        for i in range(nb):
            name, c = bcs[names[first + i]]
            kind[i] = BC[name]
            for v in range(5):
                ctx[5 * i + v] = 0.0 if c is None else c[v]
        lim = LIMITER[limiter]
        cfg = GpuConfig(space_order, FLUX[flux] if flux else 0, viscous, lim,
                        viscous if navierstokes is None else navierstokes, gamma, mu, prandtl, nb,
                        kind, ctx, None, coefficient)
        check(self.lib.teal_gpu_configure(self.h, C.byref(cfg)), "configure")
        return self

    def derivative(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64).copy()
        R = np.zeros((self.arrays["num_inner"], 5))
        check(self.lib.teal_gpu_derivative(self.h, dp(P), dp(R)), "derivative")
        return R, P

    def residual(self, R):
        """sqrt(sum V R^2 / sum V) per variable of a derivative in host layout: 5 values, or
        one row of 5 per instance on a batch handle."""
        R = np.ascontiguousarray(R, dtype=np.float64)
        n = self.lib.teal_gpu_batch_size(self.h)
        out = np.zeros((n, 5))
        check(self.lib.teal_gpu_residual(self.h, dp(R), dp(out)), "residual")
        return out[0] if n == 1 else out

    def upload(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        check(self.lib.teal_gpu_upload(self.h, dp(P)), "upload")

    def download(self):
        P = np.zeros((self.arrays["num_cells"], 5))
        check(self.lib.teal_gpu_download(self.h, dp(P)), "download")
        return P

    def lserk(self, time, courant, max_step=np.inf, time_order=2, num_stages=3):
        t = C.c_double(time)
        res = (C.c_double * 5)()
        step0 = self.lib.teal_gpu_lserk(self.h, C.byref(t), res, max_step, courant, time_order,
                                        num_stages)
        if np.isnan(step0):
            raise RuntimeError(f"lserk failed: {last_error()}")
        return step0, t.value, np.array(res[:])

    def implicit_euler(self, time, courant, newton_tol=0.1, krylov_tol=0.3, krylov_dim=32,
                       max_step=np.inf):
        t = C.c_double(time)
        res = (C.c_double * 5)()
        step0 = self.lib.teal_gpu_implicit_euler(self.h, C.byref(t), res, max_step, courant,
                                                 newton_tol, krylov_tol, krylov_dim)
        if np.isnan(step0):
            raise RuntimeError(f"implicit_euler failed: {last_error()}")
        return step0, t.value, np.array(res[:])

    def run_lserk(self, steps, time, courant, time_order=2, num_stages=3):
        t = C.c_double(time)
        check(self.lib.teal_gpu_run_lserk(self.h, steps, C.byref(t), courant, time_order,
                                          num_stages), "run_lserk")
        return t.value

    def set_graphs(self, enable):
        check(self.lib.teal_gpu_set_graphs(self.h, int(enable)), "set_graphs")
        return self

    def autotune(self, steps, courant, time_order=2, num_stages=3):
        """Per-step ms of the launch shapes named in AUTOTUNE_SHAPES (NaN: unavailable); the
        handle keeps the fastest stage shape, then the fastest gradient workgroup size and
        gather (direct or LDS-staged)."""
        ms = np.zeros(len(AUTOTUNE_SHAPES))
        check(self.lib.teal_gpu_autotune(self.h, steps, courant, time_order, num_stages, dp(ms)),
              "autotune")
        return ms

    def set_shape(self, shape, grad_block=256, grad_stage=None):
        """Stage shape (an index into AUTOTUNE_STAGE_SHAPES or its name), gradient workgroup
        size and LDS staging of the gradient's gathers (None: keep the handle's), set
        directly."""
        if isinstance(shape, str):
            shape = AUTOTUNE_STAGE_SHAPES.index(shape)
        if grad_stage is None:
            grad_stage = self.info()["grad_stage"]
        check(self.lib.teal_gpu_set_shape(self.h, shape, grad_block, int(grad_stage)),
              "set_shape")
        return self

    def set_source(self, kind, entity=0, params=(0.0, 1.0)):
        """Bulk-velocity forcing (kind SOURCE_FORCING): params = (U, tau), averaged over inner
        entity `entity`; SOURCE_NONE removes it."""
        p = np.ascontiguousarray(params, dtype=np.float64)
        check(self.lib.teal_gpu_set_source(self.h, kind, entity, dp(p)), "set_source")
        return self

    def average(self, entity=0, P=None):
        out = np.zeros(5)
        src = None if P is None else dp(np.ascontiguousarray(P, dtype=np.float64))
        check(self.lib.teal_gpu_average(self.h, entity, src, dp(out)), "average")
        return out

    def norm(self, P, ref):
        out = np.zeros(5)
        P = np.ascontiguousarray(P, dtype=np.float64)
        ref = np.ascontiguousarray(ref, dtype=np.float64)
        check(self.lib.teal_gpu_norm(self.h, dp(P), dp(ref), dp(out)), "norm")
        return out

    def polar(self, entity, P, reference, length):
        """(coef[2], pressure coefficient per face, P with its ghost rows updated)"""
        P = np.ascontiguousarray(P, dtype=np.float64).copy()
        ref = np.ascontiguousarray(reference, dtype=np.float64)
        off = self.arrays["entity_face_off"]
        pc = np.zeros(max(int(off[entity + 1] - off[entity]), 1))
        coef = np.zeros(2)
        check(self.lib.teal_gpu_polar(self.h, entity, dp(P), dp(ref), length, dp(coef), dp(pc)),
              "polar")
        return coef, pc[:int(off[entity + 1] - off[entity])], P

    def krylov_stats(self):
        st = (C.c_long * 2)()
        check(self.lib.teal_gpu_krylov_stats(self.h, st), "krylov_stats")
        return {"newton": int(st[0]), "krylov": int(st[1])}

    def set_reductions(self, mode):
        check(self.lib.teal_gpu_set_reductions(self.h, mode), "set_reductions")
        return self

    def info(self):
        c = (C.c_long * 13)()
        check(self.lib.teal_gpu_info(self.h, c), "info")
        keys = ("num_inner", "num_cells", "num_faces", "num_inner_faces", "off_periodic_faces",
                "max_cell_faces", "fused", "centre_form", "block_rows", "grad_block",
                "column_ept", "grad_stage", "tangents")
        return dict(zip(keys, [int(x) for x in c]))

    def close(self):
        if self.h:
            self.lib.teal_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch(Gpu):
    """A TealGpu over many independent meshes (teal_gpu_create_batch); configure() sees the
    union's boundary entities, so a {name: condition} map applies to every instance."""

    def __init__(self, mesh_ptrs, precision=FP64):
        self.lib = lib()
        self.meshes = list(mesh_ptrs)
        arr = (VP * len(self.meshes))(*self.meshes)
        self.h = self.lib.teal_gpu_create_batch(arr, len(self.meshes), precision)
        if not self.h:
            raise RuntimeError(f"teal_gpu_create_batch failed: {last_error()}")
        self.mesh = self.lib.teal_gpu_mesh(self.h)
        self.arrays = _abi.mesh_arrays(self.mesh)
        self.inst_arrays = [_abi.mesh_arrays(m) for m in self.meshes]
        self.precision = precision

    def upload_batch(self, Ps):
        Ps = [np.ascontiguousarray(P, dtype=np.float64) for P in Ps]
        arr = (DP * len(Ps))(*[dp(P) for P in Ps])
        check(self.lib.teal_gpu_upload_batch(self.h, arr), "upload_batch")

    def download_batch(self):
        Ps = [np.zeros((a["num_cells"], 5)) for a in self.inst_arrays]
        arr = (DP * len(Ps))(*[dp(P) for P in Ps])
        check(self.lib.teal_gpu_download_batch(self.h, arr), "download_batch")
        return Ps

    def run_lserk_batch(self, steps, times, courant, time_order=2, num_stages=3):
        t = np.ascontiguousarray(times, dtype=np.float64).copy()
        res = np.zeros((len(self.meshes), 5))
        check(self.lib.teal_gpu_run_lserk_batch(self.h, steps, dp(t), dp(res), courant,
                                                time_order, num_stages), "run_lserk_batch")
        return t, res
