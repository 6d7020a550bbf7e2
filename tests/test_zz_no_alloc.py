"""teal_gpu_residual_noalloc, the entry point that allocates nothing (no counterpart in the
reference; for callers that cannot allocate), against teal_gpu_residual
(src/equations/residual.c:8-33) through the C-ABI: the same partial sums in the same order, so
the same bits, on two calls in a row, the second on a smaller mesh (the fixed scratch array
still holds the first call's partials).  The caller's device arrays come from the HIP runtime
(the test is the caller and may allocate; the entry point does not)."""
import ctypes as C
import math

import numpy as np
import pytest

import teal_amd

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


def _device_copy(hip, host):
    ptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(ptr), host.nbytes) == 0
    assert hip.hipMemcpy(ptr, host.ctypes.data, host.nbytes, H2D) == 0
    return ptr


@pytest.mark.parametrize("precision", [teal_amd.FP64, teal_amd.MIXED], ids=["fp64", "mixed"])
def test_residual_noalloc_matches_residual(precision):
    hip = _hip()
    lib = teal_amd.lib()
    # the first mesh has 2,893 groups of 64 rows (two chunks of the residual's chunked sum),
    # the second call is on a smaller mesh
    for n, seed in ((56, 3), (9, 4)):
        print(f"residual_noalloc {n}x{n + 1}x{n + 2} precision={precision}", flush=True)
        mesh = teal_amd.mesh_box((n, n + 1, n + 2), periodic=(1, 0, 1))
        gpu = teal_amd.Gpu(mesh, precision).configure(
            flux="hllc", bcs={"bottom": ("symmetry", None), "top": ("symmetry", None)})
        ni = gpu.arrays["num_inner"]
        R = np.ascontiguousarray(np.random.default_rng(seed).standard_normal((ni, 5)) *
                                 np.array([1.0, 10.0, 10.0, 10.0, 100.0]))
        want = gpu.residual(R)
        d_R = _device_copy(hip, R)
        out = np.full(5, np.nan)
        d_out = _device_copy(hip, out)
        rc = lib.teal_gpu_residual_noalloc(gpu.h, d_R, d_out, None)
        assert rc == 0, teal_amd.last_error()
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(out.ctypes.data, d_out, out.nbytes, D2H) == 0
        hip.hipFree(d_R)
        hip.hipFree(d_out)
        np.testing.assert_array_equal(out, want)
        # the norm itself, from its definition (volume-weighted RMS per variable, divided by
        # the mesh's sum_volume as src/equations/residual.c:14,31 does; that sum is sequential,
        # generate.c:499-514, and over 185k equal volumes differs from the exact one by ~1e-12)
        vol = gpu.arrays["volume"][:ni]
        ref = np.array([math.sqrt(math.fsum(vol * R[:, v] * R[:, v]) / gpu.arrays["sum_volume"])
                        for v in range(5)])
        np.testing.assert_allclose(out, ref, rtol=1e-12)
        gpu.close()
        lib.mesh_destroy(mesh)


def test_residual_noalloc_refuses_bad_arguments():
    lib = teal_amd.lib()
    mesh = teal_amd.mesh_box((4, 4, 4))
    gpu = teal_amd.Gpu(mesh, teal_amd.FP64).configure(flux="hllc")
    assert lib.teal_gpu_residual_noalloc(gpu.h, None, None, None) != 0
    assert "null device array" in teal_amd.last_error()
    gpu.close()
    lib.mesh_destroy(mesh)
