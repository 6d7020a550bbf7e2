"""Properties at the benchmark's full size, where the oracle would take minutes per step:
the Navier-Stokes Taylor-Green vortex on the periodic 128^3 mesh (2,097,152 cells).

On a periodic mesh every face flux leaves one cell and enters the other, so the volume
integrals of the conserved variables do not change from step to step.  The FP64 path sums
each cell's face contributions in the reference's order and updates in FP64: the integrals
move only by the rounding of the per-cell updates.  The mixed path adds each face's FP32
record to both cells in FP64 and updates U in FP64: the same holds, to the rounding of
those sums and updates.  Both are checked over 10 steps, the integrals summed
exactly (a naive sum of 2 M terms alone drifts by about 4e-11); the bounds are relative to
the integral of |U| (for momentum, of its magnitude: those integrals start at zero).
Measured: FP64 0 for mass and below 1e-18 otherwise; mixed 6.5e-12 for mass, 1.9e-12 and
below for momentum and energy."""
import math

import numpy as np
import pytest

import teal_amd
from teal_amd import _abi

pytestmark = pytest.mark.gpu

GAMMA, MACH, RE, PR = 1.4, 0.1, 1600.0, 0.71
N = 128
STEPS = 10
BOUND = {"fp64": 1e-13, "mixed": 1e-11}


def tgv(center):
    x, y, z = center[:, 0], center[:, 1], center[:, 2]
    P = np.empty((len(center), 5))
    P[:, 0] = 1.0
    P[:, 1] = np.sin(x) * np.cos(y) * np.cos(z)
    P[:, 2] = -np.cos(x) * np.sin(y) * np.cos(z)
    P[:, 3] = 0.0
    P[:, 4] = 1 / (GAMMA * MACH ** 2) + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2) / 16
    return P


def integrals(P, volume):
    """Exactly rounded sums (math.fsum) of V U per conserved variable, and the scale: the
    integral of |U|, of the momentum magnitude for the three momentum components."""
    U = np.empty_like(P)
    U[:, 0] = P[:, 0]
    U[:, 1:4] = P[:, :1] * P[:, 1:4]
    U[:, 4] = P[:, 4] / (GAMMA - 1) + 0.5 * P[:, 0] * (P[:, 1:4] ** 2).sum(axis=1)
    VU = volume[:, None] * U
    total = np.array([math.fsum(VU[:, v]) for v in range(5)])
    scale = np.abs(VU).sum(axis=0)
    scale[1:4] = np.sqrt((VU[:, 1:4] ** 2).sum(axis=1)).sum()
    return total, scale


@pytest.fixture(scope="module")
def mesh():
    print(f"full-size TGV {N}^3 mesh", flush=True)
    m = teal_amd.mesh_box((N, N, N), (0.0, 0.0, 0.0), (2 * np.pi,) * 3, (1, 1, 1))
    return m, _abi.mesh_arrays(m)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tgv_128_conserves_integrals(mesh, precision):
    m, a = mesh
    print(f"full-size TGV {N}^3 conservation, {precision}, {STEPS} steps", flush=True)
    nc, ni = a["num_cells"], a["num_inner"]
    P0 = tgv(a["center"][:nc])
    P0[ni:] = 0
    vol = a["volume"][:ni].copy()
    g = teal_amd.Gpu(m, teal_amd.FP64 if precision == "fp64" else teal_amd.MIXED)
    g.configure(space_order=2, flux="hllc", viscous=1, limiter="minmod", navierstokes=1,
                gamma=GAMMA, mu=1 / RE, prandtl=PR)
    g.upload(P0)
    before, scale = integrals(g.download()[:ni], vol)
    t = g.run_lserk(STEPS, 0.0, 0.5)
    after, _ = integrals(g.download()[:ni], vol)
    drift = np.abs(after - before) / scale
    print(f"  time {t:.6g}; integral drift / integral |U|: {drift}", flush=True)
    assert t > 0
    assert (drift <= BOUND[precision]).all(), drift
    g.close()


def conserved(P):
    U = np.empty_like(P)
    U[:, 0] = P[:, 0]
    U[:, 1:4] = P[:, :1] * P[:, 1:4]
    U[:, 4] = P[:, 4] / (GAMMA - 1) + 0.5 * P[:, 0] * (P[:, 1:4] ** 2).sum(axis=1)
    return U


def test_tgv_128_mixed_error_after_1000_steps(mesh):
    """The mixed path's stated error at the benchmark configuration: after 1000 LSERK steps at
    CFL 0.5 from the same state, with the FP64 path's step sizes (the mixed path capped at the
    FP64 step, its own CFL limit 5% looser), the volume-weighted L2 norm of mixed - FP64
    (src/equations/norm.c:8-38's formula) per conserved variable, relative to the FP64 field's
    volume-weighted RMS (the momentum components to the RMS momentum magnitude, since rho w
    starts at zero), is at most 1e-6.  Measured: 8.0e-7 (rho), 8.5e-7 (rho u, rho v), 7.9e-7
    (rho w), 5.6e-8 (E)."""
    m, a = mesh
    print(f"full-size TGV {N}^3 mixed vs FP64 error, 1000 steps", flush=True)
    nc, ni = a["num_cells"], a["num_inner"]
    P0 = tgv(a["center"][:nc])
    P0[ni:] = 0
    vol = a["volume"][:ni].copy()
    gs = {}
    for name, prec in (("fp64", teal_amd.FP64), ("mixed", teal_amd.MIXED)):
        g = teal_amd.Gpu(m, prec)
        g.configure(space_order=2, flux="hllc", viscous=1, limiter="minmod", navierstokes=1,
                    gamma=GAMMA, mu=1 / RE, prandtl=PR)
        g.upload(P0)
        gs[name] = g
    t64 = tmx = 0.0
    for _ in range(1000):
        _, t_new, _ = gs["fp64"].lserk(t64, 0.5)
        dt = t_new - t64
        t64 = t_new
        _, tmx, _ = gs["mixed"].lserk(tmx, 0.5 * 1.05, max_step=dt)
    assert tmx == pytest.approx(t64, rel=1e-12)
    U64 = conserved(gs["fp64"].download()[:ni])
    Umx = conserved(gs["mixed"].download()[:ni])
    for g in gs.values():
        g.close()
    w = vol / vol.sum()
    d = Umx - U64
    l2 = np.sqrt((w[:, None] * d * d).sum(axis=0))
    scale = np.sqrt((w[:, None] * U64 * U64).sum(axis=0))
    scale[1:4] = np.sqrt((w * (U64[:, 1:4] ** 2).sum(axis=1)).sum())
    rel = l2 / scale
    print(f"  time {t64:.6g}; relative L2 error per conserved variable: {rel}", flush=True)
    assert (rel <= 1e-6).all(), rel
