"""The exact Riemann solver's failure flag (vacuum or no convergence) through every entry point
that evaluates the Godunov flux: the reference stops at the first such state
(src/euler/riemann.c:80, 179); the device reports it as the call's error, and the flag covers
one call only, so a later call on a clean state succeeds."""
import numpy as np
import pytest

import teal_amd


def _states(arrays):
    nc = arrays["num_cells"]
    x = arrays["center"][:nc, 0]
    # two halves moving apart at Mach ~40: 2/(gamma-1) (a_l + a_r) < u_r - u_l opens a vacuum
    # at x = 0.5 (Toro, Riemann Solvers, sec. 4.3.3)
    bad = np.zeros((nc, 5))
    bad[:, 0] = 1.0
    bad[:, 1] = np.where(x < 0.5, -5.0, 5.0)
    bad[:, 4] = 1e-2
    # a smooth subsonic wave; it must move and carry a pressure variation, or its residual is
    # exactly zero (a stationary contact with uniform pressure is a steady state)
    clean = np.zeros((nc, 5))
    clean[:, 0] = 1.0 + 0.1 * np.sin(2 * np.pi * x)
    clean[:, 1] = 0.1
    clean[:, 4] = 1.0 + 0.05 * np.cos(2 * np.pi * x)
    return bad, clean


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [teal_amd.FP64, teal_amd.MIXED])
def test_vacuum_is_the_calls_error(precision):
    mesh = teal_amd.mesh_box((16, 2, 2))
    gpu = teal_amd.Gpu(mesh, precision).configure(space_order=1, flux="godunov", limiter=None)
    bad, clean = _states(gpu.arrays)
    with pytest.raises(RuntimeError, match="vacuum"):
        gpu.derivative(bad)
    R, _ = gpu.derivative(clean)
    assert np.isfinite(R).all() and np.abs(R).max() > 0

    gpu.upload(bad)
    with pytest.raises(RuntimeError, match="vacuum"):
        gpu.lserk(0.0, 0.5)
    gpu.upload(clean)
    step0, _, res = gpu.lserk(0.0, 0.5)
    assert np.isfinite(step0) and np.isfinite(res).all()

    gpu.upload(bad)
    with pytest.raises(RuntimeError, match="vacuum"):
        gpu.implicit_euler(0.0, 10.0)
    gpu.upload(clean)
    step0, _, res = gpu.implicit_euler(0.0, 10.0)
    assert np.isfinite(step0) and np.isfinite(res).all()

    gpu.upload(bad)
    with pytest.raises(RuntimeError, match="run_lserk failed"):
        gpu.run_lserk(3, 0.0, 0.5)
    gpu.upload(clean)
    assert np.isfinite(gpu.run_lserk(3, 0.0, 0.5))
    gpu.close()
    teal_amd.lib().mesh_destroy(mesh)
