"""Device residual path vs the CPU oracle on identical meshes and states (FP64: bit-exact).

Every case runs on both device paths: the fused stage kernel (face fluxes of a 256-cell block
in LDS) and the split face-kernel + accumulation-kernel path (TEAL_GPU_SPLIT=1)."""
import ctypes as C

import numpy as np
import pytest

import teal_amd
from teal_amd import _abi
from cases import LARGE_CASES, RESIDUAL_CASES, gpu_bcs, initial_state
from oracle_lib import dp, oracle, oracle_eqns, oracle_mesh

pytestmark = pytest.mark.gpu


PATHS = ["fused", "split"]


@pytest.fixture(params=PATHS)
def path(request, monkeypatch):
    if request.param == "split":
        monkeypatch.setenv("TEAL_GPU_SPLIT", "1")
    else:
        monkeypatch.delenv("TEAL_GPU_SPLIT", raising=False)
    return request.param


def setup_pair(case, kind, precision=teal_amd.FP64):
    om = oracle_mesh(case)
    oe = oracle_eqns(case, om)
    pm = teal_amd.mesh_box(case.n, case.lo, case.hi, case.periodic)
    gpu = teal_amd.Gpu(pm, precision).configure(
        space_order=case.space_order, flux=case.flux, viscous=case.viscous, limiter=case.limiter,
        coefficient=case.coefficient, navierstokes=1 if case.kind == "navierstokes" else 0,
        gamma=case.gamma, mu=case.mu, prandtl=case.prandtl, bcs=gpu_bcs(case))
    arrays = _abi.mesh_arrays(om)
    P = initial_state(case, arrays, kind)
    return om, oe, gpu, arrays, P


@pytest.mark.parametrize("name,case,kind", RESIDUAL_CASES, ids=[c[0] for c in RESIDUAL_CASES])
def test_derivative_bitwise(name, case, kind, path):
    print(f"case {name} ({path})", flush=True)
    om, oe, gpu, arrays, P = setup_pair(case, kind)
    Po = P.copy()
    Ro = np.zeros((arrays["num_inner"], 5))
    oracle().oracle_derivative(oe, dp(Po), dp(Ro))
    Rg, Pg = gpu.derivative(P)
    assert np.abs(Ro).max() > 0
    if case.flux == "godunov":
        # the exact Riemann solver calls pow(): the device pow and glibc's may differ in the
        # last ulp, so this flux is held to the FP64 target of 1e-12 relative instead of bits
        np.testing.assert_allclose(Rg, Ro, rtol=1e-12, atol=1e-12 * np.abs(Ro).max())
    else:
        # bit-for-bit equal residual (and ghost states written back like the reference)
        np.testing.assert_array_equal(Rg, Ro)
    np.testing.assert_array_equal(Pg, Po)


@pytest.mark.parametrize("name,case,kind", RESIDUAL_CASES, ids=[c[0] for c in RESIDUAL_CASES])
def test_rebuilt_tangents_match_stored_basis(name, case, kind, monkeypatch):
    """FP64 face records that keep the normal alone, the tangents rebuilt on the device with
    the reference's expressions (possibly both negated), give the residual of the stored basis
    bit for bit, the exact Riemann solver included.  TEAL_GPU_TANGENTS_STORED=1 when a handle
    is created keeps the stored basis."""
    print(f"case {name} (rebuilt tangents)", flush=True)
    monkeypatch.delenv("TEAL_GPU_SPLIT", raising=False)
    monkeypatch.delenv("TEAL_GPU_TANGENTS_STORED", raising=False)
    _, _, rebuilt, _, P = setup_pair(case, kind)
    monkeypatch.setenv("TEAL_GPU_TANGENTS_STORED", "1")
    _, _, stored, _, _ = setup_pair(case, kind)
    info = rebuilt.info()
    assert info["fused"] and info["centre_form"] and info["tangents"] == 1
    assert stored.info()["tangents"] == 0
    R1, P1 = rebuilt.derivative(P.copy())
    R2, P2 = stored.derivative(P.copy())
    assert np.abs(R1).max() > 0
    np.testing.assert_array_equal(R1, R2)
    np.testing.assert_array_equal(P1, P2)
    # RK stages (the stage kernels' update and last-stage modes) on both forms
    rebuilt.upload(P)
    stored.upload(P)
    t1 = t2 = 0.0
    for _ in range(2):
        _, t1, res1 = rebuilt.lserk(t1, 0.5)
        _, t2, res2 = stored.lserk(t2, 0.5)
        assert t1 == t2 and tuple(res1) == tuple(res2)
    np.testing.assert_array_equal(rebuilt.download(), stored.download())


@pytest.mark.parametrize("name,case,kind", RESIDUAL_CASES, ids=[c[0] for c in RESIDUAL_CASES])
def test_stored_offsets_form_matches_oracle(name, case, kind, monkeypatch):
    """FP64 face records with the stored offsets and corrections (the form of meshes whose cell
    and face centres do not reproduce them; TEAL_GPU_STORED_OFFSETS=1 forces it here): the
    residual equals the oracle's, and two RK steps equal those of the centre form, bit for bit."""
    print(f"case {name} (stored offsets)", flush=True)
    monkeypatch.delenv("TEAL_GPU_SPLIT", raising=False)
    monkeypatch.delenv("TEAL_GPU_STORED_OFFSETS", raising=False)
    _, _, centre, _, _ = setup_pair(case, kind)
    monkeypatch.setenv("TEAL_GPU_STORED_OFFSETS", "1")
    om, oe, stored, arrays, P = setup_pair(case, kind)
    assert centre.info()["centre_form"] == 1
    assert stored.info()["centre_form"] == 0 and stored.info()["tangents"] == 0
    Ro = np.zeros((arrays["num_inner"], 5))
    oracle().oracle_derivative(oe, dp(P.copy()), dp(Ro))
    Rg, _ = stored.derivative(P.copy())
    if case.flux == "godunov":  # device pow (see test_derivative_bitwise)
        np.testing.assert_allclose(Rg, Ro, rtol=1e-12, atol=1e-12 * np.abs(Ro).max())
    else:
        np.testing.assert_array_equal(Rg, Ro)
    centre.upload(P)
    stored.upload(P)
    t1 = t2 = 0.0
    for _ in range(2):
        _, t1, res1 = centre.lserk(t1, 0.5)
        _, t2, res2 = stored.lserk(t2, 0.5)
        assert t1 == t2 and tuple(res1) == tuple(res2)
    np.testing.assert_array_equal(centre.download(), stored.download())


LSERK_CASES = RESIDUAL_CASES[:3] + RESIDUAL_CASES[12:15] + LARGE_CASES


@pytest.mark.parametrize("name,case,kind", LSERK_CASES, ids=[c[0] for c in LSERK_CASES])
def test_lserk_steps_bitwise(name, case, kind, path):
    print(f"case {name} ({path})", flush=True)
    om, oe, gpu, arrays, P = setup_pair(case, kind)
    po = np.ctypeslib.as_array(oracle().oracle_eqns_primitive(oe), shape=(arrays["num_cells"] * 5,))
    po[:] = P.ravel()
    gpu.upload(P)
    t_o = C.c_double(0.0)
    res_o = (C.c_double * 5)()
    t_g = 0.0
    for step in range(4):
        s_o = oracle().oracle_lserk(oe, C.byref(t_o), res_o, 1e300, 0.5, 2, 3)
        s_g, t_g, res_g = gpu.lserk(t_g, 0.5)
        assert s_g == s_o
        assert t_g == t_o.value
        np.testing.assert_allclose(res_g, np.array(res_o[:]), rtol=1e-12, atol=0)
    Pg = gpu.download()
    # inner states and the boundary ghosts of the last stage (periodic ghosts are refreshed
    # by the next exchange, so they are not part of the state)
    nb = arrays["off_boundary"]
    np.testing.assert_array_equal(Pg[:nb], po.reshape(-1, 5)[:nb])


def test_tgv_layout_uses_fused_centre_form():
    """The TGV mesh's stored face offsets and corrections are reproduced bit for bit from the
    centres, so the FP64 stage kernel reads the centre form; the mixed path keeps the arrays."""
    pm = teal_amd.mesh_box((16, 16, 16), (0, 0, 0), (2 * np.pi,) * 3, (1, 1, 1))
    g64 = teal_amd.Gpu(pm, teal_amd.FP64)
    g32 = teal_amd.Gpu(pm, teal_amd.MIXED)
    assert g64.info()["fused"] == 1 and g64.info()["centre_form"] == 1
    assert g32.info()["fused"] == 1 and g32.info()["centre_form"] == 0
    g64.close()
    g32.close()
