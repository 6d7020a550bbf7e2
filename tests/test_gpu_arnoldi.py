"""The one-launch GMRES column (k_arnoldi: matvec, every modified Gram-Schmidt step, the
norm, the Givens rotation and the next basis vector on one launch whose workgroups meet at
grid barriers, with the vector in registers) against the kernel-per-step sequence it
replaces, bit for bit.

Both form each dot product with the same workgroup count, element order and combination
order, so an implicit step takes the same Newton and GMRES iterations and ends in the same
state, to the last bit.  TEAL_GPU_ARNOLDI=0 in the environment when a handle's first implicit
step runs selects the kernel-per-step sequence.  The sizes cover every register tile the
channel meshes reach (1, 2, 4 and 8 vector elements per thread; 64^3 is the benchmark's)."""
import os

import numpy as np
import pytest

import teal_amd
from teal_amd import _abi

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def channel(n):
    """Channel [0,2] x [-1,1] x [0,1], periodic in x and z, y levels stretched by 1.05 toward
    the walls."""
    lib = teal_amd.lib()
    mesh = lib.mesh_create(teal_amd.Vector(0, -1, 0), teal_amd.Vector(2, 1, 1), teal_amd.Triple(*n),
                           teal_amd.Triple(1, 0, 1))
    half = n[1] // 2
    widths = 1.05 ** np.arange(half)
    edges = np.concatenate([[0.0], np.cumsum(widths)]) / widths.sum()
    levels = np.concatenate([-1 + edges, (1 - edges[::-1])[1:]])
    coord = _abi.mesh_arrays(mesh)["coord"]
    coord[:, 1] = levels[np.rint((coord[:, 1] + 1) / (2.0 / n[1])).astype(int)]
    lib.mesh_generate(mesh)
    return mesh


def run(mesh, P0, precision, fused, steps, courant):
    saved = os.environ.pop("TEAL_GPU_ARNOLDI", None)
    if not fused:
        os.environ["TEAL_GPU_ARNOLDI"] = "0"
    try:
        g = teal_amd.Gpu(mesh, precision)
        g.configure(space_order=2, flux="hllc", viscous=1, limiter=None, navierstokes=1,
                    gamma=GAMMA, mu=0.01, prandtl=0.72,
                    bcs={"bottom": ("wall", None), "top": ("wall", None)})
        g.upload(P0)
        t, out = 0.0, []
        for _ in range(steps):
            s, t, res = g.implicit_euler(t, courant, 0.1, 0.3, 30)
            out.append((s, t, tuple(res), tuple(g.krylov_stats().values())))
        ept = g.info()["column_ept"]
        P = g.download()
    finally:
        os.environ.pop("TEAL_GPU_ARNOLDI", None)
        if saved is not None:
            os.environ["TEAL_GPU_ARNOLDI"] = saved
    g.close()
    return out, P, ept


@pytest.mark.parametrize("n,precision,ept", [
    ((16, 16, 8), teal_amd.FP64, 1),
    ((64, 32, 32), teal_amd.FP64, 2),
    ((64, 64, 32), teal_amd.FP64, 4),
    ((64, 64, 64), teal_amd.FP64, 8),
    ((64, 64, 64), teal_amd.MIXED, 8),
])
def test_fused_column_matches_kernel_sequence(n, precision, ept):
    name = "fp64" if precision == teal_amd.FP64 else "mixed"
    print(f"arnoldi column {n} {name}", flush=True)
    mesh = channel(n)
    a = _abi.mesh_arrays(mesh)
    nc, ni = a["num_cells"], a["num_inner"]
    y = a["center"][:nc, 1]
    P0 = np.zeros((nc, 5))
    P0[:, 0] = 1.0
    P0[:, 1] = (1 - y ** 2) * (1 + 0.01 * np.random.default_rng(5).standard_normal(nc))
    P0[:, 4] = 1.0 / (GAMMA * 0.2 ** 2)
    P0[ni:] = 0
    fused, P_f, ept_f = run(mesh, P0, precision, True, 3, 200.0)
    plain, P_p, ept_p = run(mesh, P0, precision, False, 3, 200.0)
    print(f"  fused: {fused[-1]}", flush=True)
    assert ept_f == ept and ept_p == 0
    assert fused == plain
    assert sum(s[3][1] for s in fused) > 3 * 3  # several GMRES iterations per solve
    np.testing.assert_array_equal(P_f[:ni], P_p[:ni])


def test_barrier_time_out_is_reported_and_leaves_no_trace():
    """A lowered poll limit (TEAL_GPU_ARNOLDI_SPIN=1) makes the column kernel's barrier waits
    give up: every wait still ends, the implicit step reports the time-out, and after the
    state is uploaded again the same handle steps exactly as a fresh one."""
    print("arnoldi barrier time-out", flush=True)
    mesh = channel((64, 32, 32))
    a = _abi.mesh_arrays(mesh)
    nc, ni = a["num_cells"], a["num_inner"]
    y = a["center"][:nc, 1]
    P0 = np.zeros((nc, 5))
    P0[:, 0] = 1.0
    P0[:, 1] = (1 - y ** 2) * (1 + 0.01 * np.random.default_rng(5).standard_normal(nc))
    P0[:, 4] = 1.0 / (GAMMA * 0.2 ** 2)
    P0[ni:] = 0
    clean, P_clean, _ = run(mesh, P0, teal_amd.FP64, True, 2, 200.0)
    g = teal_amd.Gpu(mesh, teal_amd.FP64)
    g.configure(space_order=2, flux="hllc", viscous=1, limiter=None, navierstokes=1,
                gamma=GAMMA, mu=0.01, prandtl=0.72,
                bcs={"bottom": ("wall", None), "top": ("wall", None)})
    g.upload(P0)
    os.environ["TEAL_GPU_ARNOLDI_SPIN"] = "1"
    try:
        with pytest.raises(RuntimeError, match="grid barrier timed out"):
            g.implicit_euler(0.0, 200.0, 0.1, 0.3, 30)
    finally:
        os.environ.pop("TEAL_GPU_ARNOLDI_SPIN", None)
    assert g.info()["column_ept"] == 2
    g.upload(P0)
    t, out = 0.0, []
    for _ in range(2):
        s, t, res = g.implicit_euler(t, 200.0, 0.1, 0.3, 30)
        out.append((s, t, tuple(res), tuple(g.krylov_stats().values())))
    np.testing.assert_array_equal(g.download()[:ni], P_clean[:ni])
    assert out == clean
    g.close()
