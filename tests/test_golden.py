"""The committed golden fixtures (tests/golden/*.npz, written from the reference's own sources
by tests/golden/make_golden.py) against the CPU oracle, the library's host mesh code and,
on a GPU, the device path through the C-ABI."""
import ctypes as C
import os

import numpy as np
import pytest

import teal_amd
from cases import LARGE_CASES, RESIDUAL_CASES, gpu_bcs
from oracle_lib import arr3, dp, oracle, oracle_eqns, oracle_mesh
from teal_amd import _abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {name: (case, kind) for name, case, kind in RESIDUAL_CASES + LARGE_CASES}
FLUXES = ["godunov", "roe", "hll", "hllc", "hlle", "ausmd", "ausmdv", "lxf", "central"]
BCS = ["symmetry", "supersonic inflow", "supersonic outflow", "subsonic inflow",
       "subsonic outflow", "pressure outflow", "farfield", "adiabatic wall"]


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


RESIDUAL = load("residual.npz")
STEPPING = load("stepping.npz")
POINTWISE = load("pointwise.npz")
MESH = load("mesh.npz")
RESIDUAL_NAMES = sorted({k.split("/")[0] for k in RESIDUAL})
STEPPING_KEYS = sorted({k.rsplit("/", 1)[0] for k in STEPPING})
MESH_KEYS = sorted({k.split("/")[0] for k in MESH})


def assert_residual(name, got, want):
    if CASES[name][0].flux == "godunov":  # device pow() may differ from glibc's in the last ulp
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    else:
        np.testing.assert_array_equal(got, want)


# ---------------------------------------------------------------- oracle (CPU)
@pytest.mark.parametrize("name", RESIDUAL_NAMES)
def test_oracle_derivative_golden(name):
    case, _ = CASES[name]
    m = oracle_mesh(case)
    e = oracle_eqns(case, m)
    P = RESIDUAL[f"{name}/state"].copy()
    R = np.zeros_like(RESIDUAL[f"{name}/derivative"])
    oracle().oracle_derivative(e, dp(P), dp(R))
    np.testing.assert_array_equal(R, RESIDUAL[f"{name}/derivative"])
    np.testing.assert_array_equal(P, RESIDUAL[f"{name}/state_out"])


def run_scheme(lib, pre, e, scheme, args, t, res):
    if scheme == "lserk":
        return getattr(lib, pre + "lserk")(e, C.byref(t), res, 1e300, 0.5, int(args[0]),
                                           int(args[1]))
    if scheme == "euler":
        return getattr(lib, pre + "euler")(e, C.byref(t), res, 1e300, 0.5)
    return getattr(lib, pre + "implicit_euler")(e, C.byref(t), res, 1e300, 5.0, args[0], args[1],
                                                int(args[2]))


def parse_key(key):
    name, rest = key.split("/")
    parts = rest.split("_")
    if parts[0] == "implicit":
        return name, "implicit_euler", [float(x) for x in parts[2:]]
    return name, parts[0], [float(x) for x in parts[1:]]


@pytest.mark.parametrize("key", STEPPING_KEYS)
def test_oracle_stepping_golden(key):
    name, scheme, args = parse_key(key)
    case, _ = CASES[name]
    m = oracle_mesh(case)
    e = oracle_eqns(case, m)
    nc = _abi.mesh_arrays(m)["num_cells"]
    prim = np.ctypeslib.as_array(oracle().oracle_eqns_primitive(e), shape=(nc * 5,))
    prim[:] = STEPPING[key + "/state"].ravel()
    t = C.c_double(0.0)
    res = (C.c_double * 5)()
    for i in range(len(STEPPING[key + "/steps"])):
        s = run_scheme(oracle(), "oracle_", e, scheme, args, t, res)
        assert s == STEPPING[key + "/steps"][i]
        assert t.value == STEPPING[key + "/times"][i]
        np.testing.assert_array_equal(np.array(res[:]), STEPPING[key + "/residuals"][i])
    np.testing.assert_array_equal(prim.reshape(nc, 5), STEPPING[key + "/state_out"])


def test_oracle_pointwise_golden():
    g = POINTWISE
    o = oracle()
    prop = g["property"]
    n = len(g["left"])
    for f in FLUXES:
        out = np.zeros((n, 5))
        for i in range(n):
            o.oracle_flux(f.encode(), dp(out[i]), dp(g["left"][i]), dp(g["right"][i]), dp(prop),
                          dp(g["basis"][i]))
        np.testing.assert_array_equal(out, g[f"flux/{f}"], err_msg=f)
    out = np.zeros((n, 5))
    for i in range(n):
        o.oracle_viscous(dp(out[i]), dp(g["left"][i]), dp(g["gradient"][i]), dp(prop),
                         dp(g["basis"][i]))
    np.testing.assert_array_equal(out, g["viscous"])
    for b in BCS:
        out = np.zeros((n, 5))
        for i in range(n):
            o.oracle_bc(b.encode(), dp(out[i]), dp(g["left"][i]), dp(prop), dp(g["basis"][i]),
                        dp(g["right"][i]))
        np.testing.assert_array_equal(out, g[f"bc/{b}"], err_msg=b)
    out = np.zeros((n, 5))
    for i in range(n):
        o.oracle_riemann(dp(g["left"][i]), dp(g["right"][i]), 1.4, g["riemann_loc"][i],
                         dp(out[i]))
    np.testing.assert_array_equal(out, g["riemann"])
    cons, back = np.zeros((n, 5)), np.zeros((n, 5))
    for i in range(n):
        o.oracle_prim2cons(dp(cons[i]), dp(g["left"][i]), dp(prop))
        o.oracle_cons2prim(dp(back[i]), dp(cons[i]), dp(prop))
    np.testing.assert_array_equal(cons, g["prim2cons"])
    np.testing.assert_array_equal(back, g["cons2prim"])


def mesh_args(key):
    dims, per = key.split("_p")
    return tuple(int(x) for x in dims.split("x")), tuple(int(x) for x in per)


def check_created_mesh(m, key):
    a = _abi.mesh_arrays(m)
    counts = [a["num_nodes"], a["num_cells"], a["num_inner"], a["off_boundary"], a["off_periodic"],
              a["entities_num_inner"], a["entities_off_boundary"], a["neighbors_num"]]
    np.testing.assert_array_equal(counts, MESH[f"{key}/counts"])
    for k in ("coord", "cell_node_off", "cell_node_idx", "entity_cell_off", "rotation",
              "translation", "neighbors_tag", "neighbors_recv_off", "neighbors_send_off",
              "neighbors_send_idx"):
        np.testing.assert_array_equal(a[k], MESH[f"{key}/{k}"], err_msg=k)
    assert a["entity_names"] == list(MESH[f"{key}/entity_names"])


@pytest.mark.parametrize("key", MESH_KEYS)
def test_oracle_mesh_create_golden(key):
    n, per = mesh_args(key)
    m = oracle().oracle_mesh_create(arr3((0.0, -1.0, 0.5)), arr3((2.0, 1.0, 1.5)),
                                    arr3(n, C.c_int), arr3(per, C.c_int))
    check_created_mesh(m, key)
    oracle().oracle_mesh_destroy(m)


# ---------------------------------------------------------------- library host code (CPU)
@pytest.mark.parametrize("key", MESH_KEYS)
def test_library_mesh_create_golden(key):
    n, per = mesh_args(key)
    lib = teal_amd.lib()
    m = lib.mesh_create(teal_amd.Vector(0.0, -1.0, 0.5), teal_amd.Vector(2.0, 1.0, 1.5),
                        teal_amd.Triple(*n), teal_amd.Triple(*per))
    check_created_mesh(m, key)
    lib.mesh_destroy(m)


# ---------------------------------------------------------------- device path (GPU)
def device(case, precision=teal_amd.FP64):
    pm = teal_amd.mesh_box(case.n, case.lo, case.hi, case.periodic)
    return teal_amd.Gpu(pm, precision).configure(
        space_order=case.space_order, flux=case.flux, viscous=case.viscous, limiter=case.limiter,
        coefficient=case.coefficient, navierstokes=1 if case.kind == "navierstokes" else 0,
        gamma=case.gamma, mu=case.mu, prandtl=case.prandtl, bcs=gpu_bcs(case))


@pytest.mark.gpu
@pytest.mark.parametrize("name", RESIDUAL_NAMES)
def test_device_derivative_golden(name):
    print(f"golden derivative {name}", flush=True)
    case, _ = CASES[name]
    gpu = device(case)
    R, P = gpu.derivative(RESIDUAL[f"{name}/state"])
    assert_residual(name, R, RESIDUAL[f"{name}/derivative"])
    np.testing.assert_array_equal(P, RESIDUAL[f"{name}/state_out"])
    gpu.close()


@pytest.mark.gpu
@pytest.mark.parametrize("key", [k for k in STEPPING_KEYS if "implicit" not in k])
def test_device_explicit_stepping_golden(key):
    print(f"golden stepping {key}", flush=True)
    name, scheme, args = parse_key(key)
    case, _ = CASES[name]
    gpu = device(case)
    gpu.upload(STEPPING[key + "/state"])
    t = 0.0
    lib = gpu.lib
    for i in range(len(STEPPING[key + "/steps"])):
        tc = C.c_double(t)
        res = (C.c_double * 5)()
        if scheme == "lserk":
            s = lib.teal_gpu_lserk(gpu.h, C.byref(tc), res, 1e300, 0.5, int(args[0]), int(args[1]))
        else:
            s = lib.teal_gpu_euler(gpu.h, C.byref(tc), res, 1e300, 0.5)
        t = tc.value
        assert s == STEPPING[key + "/steps"][i]
        assert t == STEPPING[key + "/times"][i]
        np.testing.assert_allclose(np.array(res[:]), STEPPING[key + "/residuals"][i], rtol=1e-12)
    nb = _abi.mesh_arrays(gpu.mesh)["off_boundary"]
    np.testing.assert_array_equal(gpu.download()[:nb], STEPPING[key + "/state_out"][:nb])
    gpu.close()


@pytest.mark.gpu
@pytest.mark.parametrize("key", [k for k in STEPPING_KEYS if "implicit" in k])
def test_device_implicit_golden(key):
    """Serial (reference-order) reductions: bit-for-bit."""
    print(f"golden implicit {key}", flush=True)
    name, scheme, args = parse_key(key)
    case, _ = CASES[name]
    gpu = device(case).set_reductions(teal_amd.REDUCE_SERIAL)
    gpu.upload(STEPPING[key + "/state"])
    t = C.c_double(0.0)
    res = (C.c_double * 5)()
    for i in range(len(STEPPING[key + "/steps"])):
        s = gpu.lib.teal_gpu_implicit_euler(gpu.h, C.byref(t), res, 1e300, 5.0, args[0], args[1],
                                            int(args[2]))
        assert s == STEPPING[key + "/steps"][i], teal_amd.last_error()
        assert t.value == STEPPING[key + "/times"][i]
        np.testing.assert_array_equal(np.array(res[:]), STEPPING[key + "/residuals"][i])
    nb = _abi.mesh_arrays(gpu.mesh)["off_boundary"]
    np.testing.assert_array_equal(gpu.download()[:nb], STEPPING[key + "/state_out"][:nb])
    gpu.close()
