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
