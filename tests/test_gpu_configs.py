"""The benchmark workloads at parity-test sizes, device vs oracle:

- Euler Taylor-Green vortex, periodic 32^3 on [0, 2 pi]^3, HLLC, LSQ + Barth-Jespersen,
  LSERK order 2 / 3 stages, CFL 0.5 (the reference CPU workload), bit-for-bit over 20 steps
  (the full 100 steps take the oracle ~10 s);
- Euler on a jittered tetrahedral mesh (6 tets per hex, interior nodes moved by +-0.2 h),
  uniform Mach 2 flow with a Gaussian density bump, supersonic inflow at x = 0 and outflow
  elsewhere, Roe and AUSMD, at 10^3 hexes, bit-for-bit in FP64 and within a stated bound in
  the mixed path;
- Navier-Stokes implicit Euler (Newton-GMRES, Krylov dimension 30) in a channel with
  stretched no-slip walls, at 8 x 8 x 4, bit-for-bit with serial reductions."""
import ctypes as C

import numpy as np
import pytest

import teal_amd
from cases import gpu_bcs
from oracle_lib import Case, dp, oracle, oracle_eqns, oracle_mesh
from teal_amd import _abi

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def configure(gpu, case):
    return gpu.configure(space_order=case.space_order, flux=case.flux, viscous=case.viscous,
                         limiter=case.limiter, coefficient=case.coefficient,
                         navierstokes=1 if case.kind == "navierstokes" else 0, gamma=case.gamma,
                         mu=case.mu, prandtl=case.prandtl, bcs=gpu_bcs(case))


def step_both(oe, gpu, P, nc, steps, courant=0.5):
    po = np.ctypeslib.as_array(oracle().oracle_eqns_primitive(oe), shape=(nc * 5,))
    po[:] = P.ravel()
    gpu.upload(P)
    t_o = C.c_double(0.0)
    res_o = (C.c_double * 5)()
    t_g = 0.0
    for _ in range(steps):
        s_o = oracle().oracle_lserk(oe, C.byref(t_o), res_o, 1e300, courant, 2, 3)
        s_g, t_g, res_g = gpu.lserk(t_g, courant)
        yield s_o, s_g, t_o.value, t_g, np.array(res_o[:]), res_g
    yield po.reshape(-1, 5), gpu.download()


def test_euler_tgv_32():
    print("euler tgv 32^3", flush=True)
    case = Case((32, 32, 32), lo=(0, 0, 0), hi=(2 * np.pi,) * 3)
    om = oracle_mesh(case)
    oe = oracle_eqns(case, om)
    a = _abi.mesh_arrays(om)
    nc, ni = a["num_cells"], a["num_inner"]
    from cases import initial_state
    P = initial_state(case, a, "tgv")
    gpu = configure(teal_amd.Gpu(teal_amd.mesh_box(case.n, case.lo, case.hi, case.periodic)), case)
    *steps, (Po, Pg) = step_both(oe, gpu, P, nc, 20)
    for s_o, s_g, t_o, t_g, r_o, r_g in steps:
        assert s_g == s_o and t_g == t_o
        np.testing.assert_allclose(r_g, r_o, rtol=1e-12)
    np.testing.assert_array_equal(Pg[:ni], Po[:ni])
    gpu.close()


def tet_setup(flux, precision, n=10, seed=2024):
    lib = teal_amd.lib()
    o = oracle()
    o.oracle_mesh_copy_created.restype = C.c_void_p
    o.oracle_mesh_copy_created.argtypes = [C.c_void_p]
    pm = lib.mesh_create_tets(n, 0.2, seed)
    om = o.oracle_mesh_copy_created(pm)
    o.oracle_mesh_generate(om)
    lib.mesh_generate(pm)
    free = (1.0, 2.0, 0.0, 0.0, 1 / GAMMA)  # Mach 2: c = 1
    case = Case((n, n, n), periodic=(0, 0, 0), flux=flux, bcs=[("left", "supersonic inflow", free)]
                + [(s, "supersonic outflow", None) for s in ("right", "bottom", "top", "back", "front")])
    oe = oracle_eqns(case, om)
    gpu = configure(teal_amd.Gpu(pm, precision), case)
    a = _abi.mesh_arrays(om)
    nc, ni = a["num_cells"], a["num_inner"]
    x = a["center"][:nc]
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(0.3, 0.7, 3)
    sigma = 0.1
    P = np.tile(np.array(free), (nc, 1))
    P[:, 0] += 0.1 * np.exp(-((x - x0) ** 2).sum(axis=1) / (2 * sigma ** 2))
    P[ni:] = 0
    return oe, gpu, P, nc, ni


@pytest.mark.parametrize("flux", ["roe", "ausmd"])
def test_tet_bump_fp64(flux):
    print(f"tet bump {flux} fp64", flush=True)
    oe, gpu, P, nc, ni = tet_setup(flux, teal_amd.FP64)
    *steps, (Po, Pg) = step_both(oe, gpu, P, nc, 10)
    for s_o, s_g, t_o, t_g, r_o, r_g in steps:
        assert s_g == s_o and t_g == t_o
    np.testing.assert_array_equal(Pg[:ni], Po[:ni])
    gpu.close()


@pytest.mark.parametrize("flux", ["roe", "ausmd"])
def test_tet_bump_mixed(flux):
    """Mixed path with the FP64 step sizes: state error after 10 steps within 1e-6 of max |P|."""
    print(f"tet bump {flux} mixed", flush=True)
    oe, gpu, P, nc, ni = tet_setup(flux, teal_amd.MIXED)
    po = np.ctypeslib.as_array(oracle().oracle_eqns_primitive(oe), shape=(nc * 5,))
    po[:] = P.ravel()
    gpu.upload(P)
    t_o = C.c_double(0.0)
    res_o = (C.c_double * 5)()
    t_g = 0.0
    for _ in range(10):
        oracle().oracle_lserk(oe, C.byref(t_o), res_o, 1e300, 0.5, 2, 3)
        _, t_g, _ = gpu.lserk(t_g, 0.5 * 1.05, max_step=t_o.value - t_g)
    Po = po.reshape(-1, 5)[:ni]
    err = np.abs(gpu.download()[:ni] - Po).max(axis=0) / np.abs(Po).max()
    print(f"  mixed tet state error / max |P|: {err}", flush=True)
    assert (err <= 1e-6).all(), err
    gpu.close()


def stretched_channel(n, ratio=1.05):
    """Created (not generated) channel mesh on [0,2] x [-1,1] x [0,1], periodic in x and z,
    its y node levels spaced geometrically (ratio per cell) away from both walls."""
    lib = teal_amd.lib()
    pm = lib.mesh_create(teal_amd.Vector(0, -1, 0), teal_amd.Vector(2, 1, 1), teal_amd.Triple(*n),
                         teal_amd.Triple(1, 0, 1))
    half = n[1] // 2
    widths = ratio ** np.arange(half)
    edges = np.concatenate([[0.0], np.cumsum(widths)]) / widths.sum()  # wall (0) .. centre (1)
    levels = np.concatenate([-1 + edges, (1 - edges[::-1])[1:]])
    coord = _abi.mesh_arrays(pm)["coord"]
    j = np.rint((coord[:, 1] + 1) / (2.0 / n[1])).astype(int)
    coord[:, 1] = levels[j]
    return pm


def test_channel_implicit_serial():
    """No-slip walls in y (nodes stretched geometrically by 1.05 toward the walls), periodic x
    and z, parabolic profile + 1% seeded noise, Re = 100 on the half height, M = 0.2, implicit
    Euler NewtonKrylov {0.1, 0.3, 30}."""
    print("channel implicit", flush=True)
    n = (8, 8, 4)
    case = Case(n, lo=(0, -1, 0), hi=(2, 1, 1), periodic=(1, 0, 1), kind="navierstokes",
                mu=0.01, bcs=[("bottom", "wall", None), ("top", "wall", None)])
    o = oracle()
    o.oracle_mesh_copy_created.restype = C.c_void_p
    o.oracle_mesh_copy_created.argtypes = [C.c_void_p]
    pm = stretched_channel(n)
    om = o.oracle_mesh_copy_created(pm)
    o.oracle_mesh_generate(om)
    teal_amd.lib().mesh_generate(pm)
    oe = oracle_eqns(case, om)
    a = _abi.mesh_arrays(om)
    nc, ni = a["num_cells"], a["num_inner"]
    y = a["center"][:nc, 1]
    rng = np.random.default_rng(5)
    p0 = 1 / (GAMMA * 0.2 ** 2)
    P = np.zeros((nc, 5))
    P[:, 0] = 1.0
    P[:, 1] = (1 - y ** 2) * (1 + 0.01 * rng.standard_normal(nc))
    P[:, 4] = p0
    P[ni:] = 0
    gpu = configure(teal_amd.Gpu(pm), case)
    gpu.set_reductions(teal_amd.REDUCE_SERIAL)
    po = np.ctypeslib.as_array(oracle().oracle_eqns_primitive(oe), shape=(nc * 5,))
    po[:] = P.ravel()
    gpu.upload(P)
    t_o = C.c_double(0.0)
    res_o = (C.c_double * 5)()
    t_g = 0.0
    for _ in range(3):
        s_o = oracle().oracle_implicit_euler(oe, C.byref(t_o), res_o, 1e300, 20.0, 0.1, 0.3, 30)
        s_g, t_g, r_g = gpu.implicit_euler(t_g, 20.0, 0.1, 0.3, 30)
        assert s_g == s_o and t_g == t_o.value
        np.testing.assert_array_equal(r_g, np.array(res_o[:]))
    nb = a["off_boundary"]
    np.testing.assert_array_equal(gpu.download()[:nb], po.reshape(-1, 5)[:nb])
    gpu.close()
