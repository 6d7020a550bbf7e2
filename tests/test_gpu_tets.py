"""Device path on jittered tetrahedral meshes (6 tets per hex, boundary triangles as ghost
cells) against the oracle on the same mesh: FP64 bit-for-bit, both device paths."""
import ctypes as C

import numpy as np
import pytest

import teal_amd
from cases import gpu_bcs
from oracle_lib import Case, dp, oracle, oracle_eqns, random_state
from teal_amd import _abi

pytestmark = pytest.mark.gpu

SIDES = ("left", "right", "bottom", "top", "back", "front")
TET_CASES = [
    ("tets_euler_roe", Case((5, 5, 5), periodic=(0, 0, 0), flux="roe", bcs=[
        ("left", "supersonic inflow", (1.0, 2.0, 0.0, 0.0, 1 / 1.4))]
        + [(s, "supersonic outflow", None) for s in SIDES[1:]])),
    ("tets_euler_ausmd", Case((5, 5, 5), periodic=(0, 0, 0), flux="ausmd", bcs=[
        ("left", "supersonic inflow", (1.0, 2.0, 0.0, 0.0, 1 / 1.4))]
        + [(s, "supersonic outflow", None) for s in SIDES[1:]])),
    ("tets_euler_ausmdv_walls", Case((5, 5, 5), periodic=(0, 0, 0), flux="ausmdv", bcs=[
        ("left", "subsonic inflow", (1.1, 0.1, 0.0, 0.2, 1.0)),
        ("right", "subsonic outflow", (1.0, 0.0, 0.0, 0.0, 0.9)),
        ("bottom", "symmetry", None), ("top", "farfield", (1.0, 0.2, 0.0, 0.1, 1.0)),
        ("back", "pressure outflow", (1, 0, 0, 0, 0.95)), ("front", "slipwall", None)])),
    ("tets_ns_hllc_vk", Case((5, 5, 5), periodic=(0, 0, 0), kind="navierstokes", mu=0.02,
                             limiter="venkatakrishnan", coefficient=2.0,
                             bcs=[(s, "wall", None) for s in SIDES])),
    ("tets_euler_lxf_o1", Case((5, 5, 5), periodic=(0, 0, 0), flux="lxf", space_order=1,
                               bcs=[(s, "supersonic outflow", None) for s in SIDES])),
]


@pytest.fixture(params=["fused", "split"])
def path(request, monkeypatch):
    if request.param == "split":
        monkeypatch.setenv("TEAL_GPU_SPLIT", "1")
    else:
        monkeypatch.delenv("TEAL_GPU_SPLIT", raising=False)
    return request.param


def setup(case, n=5, jitter=0.2, seed=77):
    lib = teal_amd.lib()
    o = oracle()
    o.oracle_mesh_copy_created.restype = C.c_void_p
    o.oracle_mesh_copy_created.argtypes = [C.c_void_p]
    pm = lib.mesh_create_tets(n, jitter, seed)
    om = o.oracle_mesh_copy_created(pm)
    o.oracle_mesh_generate(om)
    lib.mesh_generate(pm)
    oe = oracle_eqns(case, om)
    gpu = teal_amd.Gpu(pm).configure(
        space_order=case.space_order, flux=case.flux, viscous=case.viscous, limiter=case.limiter,
        coefficient=case.coefficient, navierstokes=1 if case.kind == "navierstokes" else 0,
        gamma=case.gamma, mu=case.mu, prandtl=case.prandtl, bcs=gpu_bcs(case))
    a = _abi.mesh_arrays(om)
    P = random_state(a["num_cells"], 21)
    P[a["num_inner"]:] = 0
    return oe, gpu, a, P


@pytest.mark.parametrize("name,case", TET_CASES, ids=[c[0] for c in TET_CASES])
def test_tets_derivative_and_lserk(name, case, path):
    print(f"tets {name} ({path})", flush=True)
    oe, gpu, a, P = setup(case)
    nc, ni = a["num_cells"], a["num_inner"]
    Po = P.copy()
    Ro = np.zeros((ni, 5))
    oracle().oracle_derivative(oe, dp(Po), dp(Ro))
    Rg, Pg = gpu.derivative(P)
    np.testing.assert_array_equal(Rg, Ro)
    np.testing.assert_array_equal(Pg, Po)
    po = np.ctypeslib.as_array(oracle().oracle_eqns_primitive(oe), shape=(nc * 5,))
    po[:] = P.ravel()
    gpu.upload(P)
    t_o = C.c_double(0.0)
    res_o = (C.c_double * 5)()
    t_g = 0.0
    for _ in range(3):
        s_o = oracle().oracle_lserk(oe, C.byref(t_o), res_o, 1e300, 0.5, 2, 3)
        s_g, t_g, res_g = gpu.lserk(t_g, 0.5)
        assert s_g == s_o and t_g == t_o.value
        np.testing.assert_allclose(res_g, np.array(res_o[:]), rtol=1e-12, atol=0)
    nb = a["off_boundary"]
    np.testing.assert_array_equal(gpu.download()[:nb], po.reshape(-1, 5)[:nb])
    gpu.close()
