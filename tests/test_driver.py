"""Drop-in check at the level of a teal program: a driver written against the reference API
(tests/drivers/tgv.c, built by tests/drivers/build.sh) links libteal_amd.so, runs
simulation_run on the device and is compared with the oracle stepping the same state; the
reference's own Taylor-Green driver compiles and links unchanged against include/."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle_lib import Case, oracle, oracle_eqns, oracle_mesh
from teal_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVERS = os.path.join(ROOT, "tests", "drivers")
TGV = os.path.join(DRIVERS, "bin", "tgv")
REF_TGV = "/root/reference/run/taylor_green_vortex/run.c"


def driver():
    if not os.path.exists(TGV):
        subprocess.run(["bash", os.path.join(DRIVERS, "build.sh")], check=True)
    return TGV


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device error path")
def test_driver_fails_loudly_without_a_device(tmp_path):
    r = subprocess.run([driver(), "6", "2", str(tmp_path / "out")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "no HIP device" in r.stderr + r.stdout


@pytest.mark.skipif(not os.path.exists(REF_TGV), reason="reference tree absent")
def test_reference_driver_links_unchanged(tmp_path):
    out = tmp_path / "ref_tgv"
    subprocess.run(["gcc", "-std=c99", "-O1", "-w", "-I", os.path.join(ROOT, "include"), REF_TGV,
                    "-o", str(out), "-L", os.path.join(ROOT, "teal_amd"), "-lteal_amd", "-lm"],
                   check=True, timeout=120)
    assert out.exists()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_driver_simulation_run(tmp_path, precision):
    n, iters = 12, 6
    print(f"driver tgv {n}^3 {iters} iterations {precision}", flush=True)
    prefix = str(tmp_path / "tgv")
    r = subprocess.run([driver(), str(n), str(iters), prefix, precision], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    case = Case((n, n, n), lo=(0, 0, 0), hi=(2 * np.pi,) * 3, kind="navierstokes", mu=1 / 1600)
    om = oracle_mesh(case)
    oe = oracle_eqns(case, om)
    a = _abi.mesh_arrays(om)
    nc, ni = a["num_cells"], a["num_inner"]
    P0 = np.fromfile(prefix + "_initial.bin").reshape(nc, 5)
    P1 = np.fromfile(prefix + "_final.bin").reshape(nc, 5)
    po = np.ctypeslib.as_array(oracle().oracle_eqns_primitive(oe), shape=(nc * 5,))
    po[:] = P0.ravel()
    t = C.c_double(0.0)
    res = (C.c_double * 5)()
    for _ in range(iters):  # simulation_run: max_step = DBL_MAX - time (run.c:69)
        oracle().oracle_lserk(oe, C.byref(t), res, np.finfo(float).max - t.value, 0.5, 2, 3)
    want = po.reshape(nc, 5)[:ni]
    assert f"final time {t.value:.17g}" in r.stdout or precision == "mixed"
    if precision == "fp64":
        np.testing.assert_array_equal(P1[:ni], want)
    else:
        err = np.abs(P1[:ni] - want).max(axis=0) / np.abs(want).max(axis=0).clip(1e-300)
        print(f"  mixed driver state error / max |P|: {err}", flush=True)
        assert (err[[0, 4]] <= 1e-6).all(), err


REF_TGV_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_tgv")


def _h5_cells(path, name, width):
    from test_io import dump, numbers

    return numbers(dump(path, f"/VTKHDF/CellData/{name}")).reshape(-1, width)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(REF_TGV_BIN),
                    reason="the reference's driver is built only where the reference tree is")
@pytest.mark.skipif(not os.path.exists("/opt/conda/bin/h5dump"), reason="h5dump absent")
def test_reference_tgv_driver_on_device(tmp_path):
    """The reference's own run/taylor_green_vortex/run.c, unchanged and linked against
    libteal_amd.so (oracle/Makefile), runs simulation_run on the device: NS on [-pi, pi]^3,
    Venkatakrishnan K = 1, mu = 1/1600, the default LSERK order 2 / 3 stages at CFL 0.99
    (src/simulation/create.c:23).  Its last table row (iter, time, step, residual; %g) and its
    final VTKHDF state must equal the oracle's lserk from the driver's own initial state (read
    back from its first output file, so the host libm of its initial-condition callback does
    not enter the comparison), bit for bit."""
    n, iters = 8, 5
    print(f"reference tgv driver {n}^3 {iters} iterations", flush=True)
    exe = tmp_path / "ref_tgv"
    exe.symlink_to(REF_TGV_BIN)
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(ROOT, "teal_amd"))
    r = subprocess.run([str(exe), str(n), str(iters)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=300, env=env)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    prefix = str(exe) + f"_{n:04d}_0001"  # run.c: "%s_%04d_%04d", argv[0], num, sync.size
    first, last = prefix + "_00000.hdf", prefix + "_00001.hdf"
    assert os.path.exists(first) and os.path.exists(last)

    case = Case((n, n, n), lo=(-np.pi,) * 3, hi=(np.pi,) * 3, kind="navierstokes", mu=1 / 1600,
                limiter="venkatakrishnan", coefficient=1.0)
    om = oracle_mesh(case)
    oe = oracle_eqns(case, om)
    a = _abi.mesh_arrays(om)
    nc, ni = a["num_cells"], a["num_inner"]

    def state(path):
        return np.hstack([_h5_cells(path, "density", 1), _h5_cells(path, "velocity", 3),
                          _h5_cells(path, "pressure", 1)])[:ni]

    P0 = np.zeros((nc, 5))
    P0[:ni] = state(first)
    po = np.ctypeslib.as_array(oracle().oracle_eqns_primitive(oe), shape=(nc * 5,))
    po[:] = P0.ravel()
    t = C.c_double(0.0)
    res = (C.c_double * 5)()
    for _ in range(iters):
        step0 = oracle().oracle_lserk(oe, C.byref(t), res, np.finfo(float).max - t.value, 0.99, 2, 3)
    row = "%13d %13g %13g %13g" % (iters, t.value, step0, max(res[:]))
    assert any(line.strip().startswith(row.strip()) for line in r.stdout.splitlines()), row
    np.testing.assert_array_equal(state(last), po.reshape(nc, 5)[:ni])
