"""Mixed path (FP32 flux / gradient / geometry, FP64 accumulation and update)
against the FP64 oracle, on both the fused stage kernel and the face + accumulation kernels.  Its error is stated, not zero: the bounds below are the ones this
path is held to (relative to the largest magnitude of each quantity)."""
import ctypes as C

import numpy as np
import pytest

import teal_amd
from cases import LARGE_CASES, RESIDUAL_CASES, initial_state
from oracle_lib import dp, oracle
from test_gpu_residual import path, setup_pair  # noqa: F401 (path: fused / split fixture)

pytestmark = pytest.mark.gpu

# derivative: R is a small difference of face fluxes of size p A / V (the TGV carries a
# background pressure 1/(gamma M^2) = 71.4), so FP32 rounding of the fluxes shows up as
# about 1.5e-4 of max |R| on the 24^3 TGV; the bound keeps a margin over that
DERIVATIVE_RTOL = 1e-3
# state after 4 LSERK steps (conserved variables, FP64 update)
STATE_RTOL = 2e-6

CASES = RESIDUAL_CASES[:3] + RESIDUAL_CASES[3:4] + RESIDUAL_CASES[12:15] + LARGE_CASES


@pytest.mark.parametrize("name,case,kind", CASES, ids=[c[0] for c in CASES])
def test_mixed_derivative_error(name, case, kind, path):
    print(f"mixed {name} {path}", flush=True)
    om, oe, gpu, arrays, P = setup_pair(case, kind, precision=teal_amd.MIXED)
    Po = P.copy()
    Ro = np.zeros((arrays["num_inner"], 5))
    oracle().oracle_derivative(oe, dp(Po), dp(Ro))
    Rg, _ = gpu.derivative(P)
    scale = np.abs(Ro).max(axis=0)
    scale[1:4] = np.abs(Ro[:, 1:4]).max()
    err = np.abs(Rg - Ro).max(axis=0) / scale
    print(f"  derivative max error / max |R| per variable: {err}", flush=True)
    assert (err <= DERIVATIVE_RTOL).all(), err


@pytest.mark.parametrize("name,case,kind", CASES, ids=[c[0] for c in CASES])
def test_mixed_lserk_error(name, case, kind, path):
    print(f"mixed lserk {name} {path}", flush=True)
    om, oe, gpu, arrays, P = setup_pair(case, kind, precision=teal_amd.MIXED)
    ni, nc = arrays["num_inner"], arrays["num_cells"]
    po = np.ctypeslib.as_array(oracle().oracle_eqns_primitive(oe), shape=(nc * 5,))
    po[:] = P.ravel()
    gpu.upload(P)
    t_o = C.c_double(0.0)
    res_o = (C.c_double * 5)()
    t_g = 0.0
    for _ in range(4):
        oracle().oracle_lserk(oe, C.byref(t_o), res_o, 1e300, 0.5, 2, 3)
        dt = t_o.value - t_g
        # same step sizes as the FP64 run (the mixed dt differs in the last bits)
        _, t_g, _ = gpu.lserk(t_g, 0.5 * 1.05, max_step=dt)
    prop = np.array([case.gamma, case.mu, case.prandtl, 0, 0, 0, 0, 0])

    def cons(Pv):
        U = np.zeros((ni, 5))
        for i in range(ni):
            oracle().oracle_prim2cons(dp(U[i]), dp(np.ascontiguousarray(Pv[i])), dp(prop))
        return U

    Ug = cons(gpu.download()[:ni])
    Uo = cons(po.reshape(-1, 5)[:ni])
    scale = np.abs(Uo).max(axis=0)
    scale[1:4] = np.abs(Uo[:, 1:4]).max()  # momentum components against the momentum scale
    err = np.abs(Ug - Uo).max(axis=0) / scale
    print(f"  state max error / max |U| per variable: {err}", flush=True)
    assert (err <= STATE_RTOL).all(), err
