"""Source terms and the diagnostics reductions on the device against the oracle:

- the bulk-velocity forcing of run/periodic_hill/run.c:8-21 (Prepare: the entity average of
  the state being evaluated; source: rho (U - <u>) / tau in the x momentum, u times that in the
  energy), through equations_derivative's finalize step (src/equations/derivative.c:242-265),
  LSERK steps and implicit Euler;
- equations_average (src/equations/average.c:8-41), equations_norm (norm.c:8-38) and the face
  sums of euler_polar (src/euler/polar.c:36-53).

With serial reductions (the reference's summation order) the FP64 results are bit for bit;
with the default compensated trees the averages differ from the reference's serial sums by
at most that sum's own rounding error, and the results are compared at relative 1e-12."""
import ctypes as C

import numpy as np
import pytest

import teal_amd
from cases import gpu_bcs
from oracle_lib import Case, dp, oracle, oracle_eqns, oracle_mesh, random_state
from teal_amd import _abi

pytestmark = pytest.mark.gpu

U_BULK, TAU = 1.0, 0.05

CHANNEL = Case((8, 6, 4), lo=(0, -1, 0), hi=(4, 1, 2), periodic=(1, 0, 1), kind="navierstokes",
               mu=0.02, bcs=[("bottom", "wall", None), ("top", "wall", None)])
HILL_2D = Case((10, 6, 1), lo=(0, 0, 0), hi=(9, 3, 1), periodic=(1, 0, 0), kind="navierstokes",
               mu=0.01, limiter="venkatakrishnan", coefficient=1.0,
               bcs=[("bottom", "wall", None), ("top", "wall", None)])


def channel_state(a, seed=5):
    nc, ni = a["num_cells"], a["num_inner"]
    P = random_state(nc, seed, vel=0.2)
    y = a["center"][:nc, 1]
    P[:, 1] += 0.8 * (1 - y * y)  # parabolic streamwise profile plus noise
    P[:, 4] = 1 / (1.4 * 0.2 ** 2) * (1 + 0.01 * P[:, 4])
    P[ni:] = 0.0
    return np.ascontiguousarray(P)


def setup(case, precision=teal_amd.FP64, serial=True):
    om = oracle_mesh(case)
    oe = oracle_eqns(case, om)
    oracle().oracle_eqns_set_forcing(oe, U_BULK, TAU)
    a = _abi.mesh_arrays(om)
    pm = teal_amd.mesh_box(case.n, case.lo, case.hi, case.periodic)
    gpu = teal_amd.Gpu(pm, precision).configure(
        space_order=case.space_order, flux=case.flux, viscous=case.viscous, limiter=case.limiter,
        coefficient=case.coefficient, navierstokes=1, gamma=case.gamma, mu=case.mu,
        prandtl=case.prandtl, bcs=gpu_bcs(case))
    gpu.set_source(teal_amd.SOURCE_FORCING, 0, (U_BULK, TAU))
    if serial:
        gpu.set_reductions(teal_amd.REDUCE_SERIAL)
    return om, oe, a, gpu


@pytest.mark.parametrize("case", [CHANNEL, HILL_2D], ids=["channel", "hill2d"])
@pytest.mark.parametrize("serial", [True, False], ids=["serial", "tree"])
def test_forcing_derivative(case, serial):
    print(f"forcing derivative {case.n} serial={serial}", flush=True)
    om, oe, a, gpu = setup(case, serial=serial)
    ni = a["num_inner"]
    P = channel_state(a)
    Ro = np.zeros((ni, 5))
    oracle().oracle_derivative(oe, dp(P.copy()), dp(Ro))
    Rg, _ = gpu.derivative(P)
    if serial:
        np.testing.assert_array_equal(Rg, Ro)
    else:
        np.testing.assert_allclose(Rg, Ro, rtol=0, atol=1e-12 * np.abs(Ro).max())
    # the source is really there: without it the x momentum differs by about rho (U - <u>)/tau
    gpu.set_source(teal_amd.SOURCE_NONE)
    R0, _ = gpu.derivative(P)
    assert np.abs(R0[:, 1] - Ro[:, 1]).max() > 1e-3
    gpu.close()


def test_forcing_lserk_steps():
    print("forcing lserk", flush=True)
    om, oe, a, gpu = setup(CHANNEL)
    nc, ni = a["num_cells"], a["num_inner"]
    P = channel_state(a)
    po = np.ctypeslib.as_array(oracle().oracle_eqns_primitive(oe), shape=(nc * 5,))
    po[:] = P.ravel()
    gpu.upload(P)
    t_o, res_o, t_g = C.c_double(0.0), (C.c_double * 5)(), 0.0
    for _ in range(5):
        s_o = oracle().oracle_lserk(oe, C.byref(t_o), res_o, 1e300, 0.5, 2, 3)
        s_g, t_g, res_g = gpu.lserk(t_g, 0.5)
        assert s_g == s_o and t_g == t_o.value
        np.testing.assert_allclose(res_g, np.array(res_o[:]), rtol=1e-12)
    np.testing.assert_array_equal(gpu.download()[:ni], po.reshape(-1, 5)[:ni])
    gpu.close()


# FP64: the default solver at CFL 20, bit for bit with serial reductions.  Mixed (FP32 matvec
# residuals, FP64 Newton residual): both paths stop Newton at the same relative tolerance, so
# they differ by about that tolerance times the step's change (0.1 x ~5e-2 at CFL 20 with the
# default solver); the mixed case runs the tight solver at CFL 5 and is held to 1e-3 of the
# largest |U| per conserved variable, the bound of the mixed implicit tests.
IMPLICIT = {teal_amd.FP64: (20.0, (0.1, 0.3, 30)), teal_amd.MIXED: (5.0, (1e-3, 1e-2, 10))}


@pytest.mark.parametrize("precision", [teal_amd.FP64, teal_amd.MIXED], ids=["fp64", "mixed"])
def test_forcing_implicit(precision):
    """Implicit Euler with the forcing source (the matvec's residuals carry the source too)."""
    from test_gpu_implicit import MIXED_RTOL, conserved, relative_error

    print(f"forcing implicit precision={precision}", flush=True)
    courant, solver = IMPLICIT[precision]
    om, oe, a, gpu = setup(CHANNEL, precision, serial=True)
    nc, ni = a["num_cells"], a["num_inner"]
    P = channel_state(a)
    po = np.ctypeslib.as_array(oracle().oracle_eqns_primitive(oe), shape=(nc * 5,))
    po[:] = P.ravel()
    gpu.upload(P)
    t_o, res_o, t_g = C.c_double(0.0), (C.c_double * 5)(), 0.0
    for _ in range(3):
        s_o = oracle().oracle_implicit_euler(oe, C.byref(t_o), res_o, 1e300, courant, *solver)
        s_g, t_g, res_g = gpu.implicit_euler(t_g, courant, *solver)
        if precision == teal_amd.FP64:
            assert s_g == s_o
            np.testing.assert_allclose(res_g, np.array(res_o[:]), rtol=1e-12)
    Po, Pg = po.reshape(-1, 5)[:ni], gpu.download()[:ni]
    if precision == teal_amd.FP64:
        np.testing.assert_array_equal(Pg, Po)
    else:
        err = relative_error(conserved(Pg, CHANNEL.gamma), conserved(Po, CHANNEL.gamma))
        print(f"  state error / max |U| per variable: {err}", flush=True)
        assert (err <= MIXED_RTOL).all(), err
    gpu.close()


@pytest.mark.parametrize("serial", [True, False], ids=["serial", "tree"])
def test_average_and_norm(serial):
    case = CHANNEL
    om, oe, a, gpu = setup(case, serial=serial)
    gpu.set_source(teal_amd.SOURCE_NONE)
    ni = a["num_inner"]
    P = channel_state(a, seed=11)
    want = np.zeros(5)
    oracle().oracle_average(oe, b"domain", dp(P), dp(want))
    host = gpu.average(0, P)
    gpu.upload(P)
    resident = gpu.average(0)
    ref = np.ascontiguousarray(P[:ni] + np.random.default_rng(1).normal(0, 1e-3, (ni, 5)))
    nw = np.zeros(5)
    oracle().oracle_norm(oe, dp(ref), dp(P), dp(nw))
    ng = gpu.norm(P[:ni].copy(), ref)
    if serial:
        np.testing.assert_array_equal(host, want)
        np.testing.assert_array_equal(resident, want)
        np.testing.assert_array_equal(ng, nw)
    else:
        np.testing.assert_allclose(host, want, rtol=1e-14)
        np.testing.assert_allclose(resident, want, rtol=1e-14)
        np.testing.assert_allclose(ng, nw, rtol=1e-14)
    gpu.close()


def test_polar():
    case = Case((6, 5, 3), lo=(0, 0, 0), hi=(3.0, 2.0, 1.0), periodic=(0, 0, 1), flux="roe",
                bcs=[("left", "supersonic inflow", (1.0, 2.0, 0.3, 0.0, 0.7)),
                     ("right", "supersonic outflow", None), ("bottom", "symmetry", None),
                     ("top", "farfield", (1.0, 2.0, 0.3, 0.0, 0.7))])
    om = oracle_mesh(case)
    oe = oracle_eqns(case, om)
    a = _abi.mesh_arrays(om)
    gpu = teal_amd.Gpu(teal_amd.mesh_box(case.n, case.lo, case.hi, case.periodic)).configure(
        flux="roe", bcs=gpu_bcs(case))
    P = random_state(a["num_cells"], 4)
    P[a["num_inner"]:] = 0.0
    ref = np.array([1.0, 2.0, 0.3, 0.0, 0.7])
    for name in ("bottom", "top"):
        entity = a["entity_names"].index(name)
        num = int(a["entity_face_off"][entity + 1] - a["entity_face_off"][entity])
        Po = P.copy()
        coef, pc = np.zeros(2), np.zeros(num)
        oracle().oracle_polar(oe, entity, dp(Po), dp(ref), 1.5, dp(coef), dp(pc))
        cg, pcg, Pg = gpu.polar(entity, P, ref, 1.5)
        np.testing.assert_array_equal(cg, coef)
        np.testing.assert_array_equal(pcg, pc)
        np.testing.assert_array_equal(Pg, Po)  # ghost rows as the reference's boundary pass
    gpu.close()
