"""VTKHDF output and restart (src/mesh/write.c:185-211, src/equations/write.c:123-144,
src/equations/read.c:180-197) through the teal API, with the C driver tests/drivers/io.c on a
fully periodic box (no boundary ghosts: the whole path runs on the host).  The reference's
writers need parallel HDF5 and cannot be built here, so the files are checked against the
layout its writers define: every dataset and attribute they write, with the values taken from
the mesh and the state, read back with the image's h5dump; and the restart round trip."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVERS = os.path.join(ROOT, "tests", "drivers")
IO = os.path.join(DRIVERS, "bin", "io")
H5DUMP = "/opt/conda/bin/h5dump"

pytestmark = pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump absent")


def driver():
    if not os.path.exists(IO):
        subprocess.run(["bash", os.path.join(DRIVERS, "build.sh")], check=True)
    return IO


def run(*args):
    r = subprocess.run([driver(), *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def dump(path, obj):
    """Values of one dataset or attribute as a flat float array (h5dump -y text)."""
    flag = "-a" if obj.startswith("attr:") else "-d"
    name = obj[5:] if obj.startswith("attr:") else obj
    out = subprocess.run([H5DUMP, "-y", "-w", "0", "-m", "%.17g", flag, name, path], capture_output=True,
                         text=True, timeout=60, check=True).stdout
    body = out[out.index("DATA {") + 6:out.rindex("}")]
    return body


def numbers(text):
    return np.array([float(x) for x in re.findall(r"-?[0-9.]+(?:e[-+]?[0-9]+)?", text)])


def mesh_of(n):
    import teal_amd
    from teal_amd import _abi

    m = teal_amd.mesh_box((n, n, n), (0, 0, 0), (1, 2, 0.5), (1, 1, 1))
    return m, _abi.mesh_arrays(m)


def test_mesh_file_layout(tmp_path):
    prefix = str(tmp_path / "run")
    run("write", prefix, 4)
    path = prefix + "_mesh.hdf"
    m, a = mesh_of(4)
    nc = a["off_periodic"]
    assert '"UnstructuredGrid"' in dump(path, "attr:/VTKHDF/Type")
    np.testing.assert_array_equal(numbers(dump(path, "attr:/VTKHDF/Version")), [1, 0])
    assert numbers(dump(path, "/VTKHDF/NumberOfCells"))[0] == nc
    assert numbers(dump(path, "/VTKHDF/NumberOfPoints"))[0] == a["num_nodes"]
    off = a["cell_node_off"][:nc + 1]
    np.testing.assert_array_equal(numbers(dump(path, "/VTKHDF/Offsets")), off)
    np.testing.assert_array_equal(numbers(dump(path, "/VTKHDF/Connectivity")),
                                  a["cell_node_idx"][:off[-1]])
    types = numbers(dump(path, "/VTKHDF/Types"))
    ni = a["num_inner"]
    assert (types[:ni] == 12).all() and (types[ni:] == 9).all()  # hexahedra, quad ghosts
    np.testing.assert_array_equal(numbers(dump(path, "/VTKHDF/Points")).reshape(-1, 3), a["coord"])
    ghost = numbers(dump(path, "/VTKHDF/CellData/vtkGhostType"))
    assert (ghost[:ni] == 0).all() and (ghost[ni:] == 48).all()
    np.testing.assert_array_equal(numbers(dump(path, "/VTKHDF/CellData/volume")), a["volume"][:nc])
    entity = numbers(dump(path, "/VTKHDF/CellData/entity"))
    eoff = a["entity_cell_off"]
    for e in range(len(a["entity_names"])):
        assert (entity[eoff[e]:eoff[e + 1]] == e).all()


def test_solution_file_and_restart(tmp_path):
    prefix = str(tmp_path / "run")
    run("write", prefix, 4)
    path = prefix + "_00003.hdf"
    m, a = mesh_of(4)
    nc, ni = a["off_periodic"], a["num_inner"]
    P = np.fromfile(prefix + "_state.bin").reshape(-1, 5)
    # periodic ghosts carry their source cell's state (the exchange before the write)
    assert numbers(dump(path, "/VTKHDF/FieldData/time"))[0] == 1.25
    assert numbers(dump(path, "/VTKHDF/FieldData/dynamic viscosity"))[0] == 0.0125
    rho = numbers(dump(path, "/VTKHDF/CellData/density"))
    vel = numbers(dump(path, "/VTKHDF/CellData/velocity")).reshape(-1, 3)
    np.testing.assert_array_equal(rho[:ni], P[:ni, 0])
    np.testing.assert_array_equal(vel[:ni], P[:ni, 1:4])
    from teal_amd import lib

    src = np.full(a["num_cells"], -1, dtype=np.int32)
    lib().teal_mesh_periodic_sources(C.c_void_p(m), src.ctypes.data_as(C.POINTER(C.c_int)))
    g = np.arange(ni, nc)
    assert (src[g] >= 0).all()
    np.testing.assert_array_equal(rho[g], P[src[g], 0])
    # restart on the same mesh: the state comes back bit for bit, the output index continues
    out = run("read", path, 4, str(tmp_path / "back"))
    assert "time 1.25 index 4 mu 0.0125" in out
    back = np.fromfile(str(tmp_path / "back") + ".bin").reshape(-1, 5)
    np.testing.assert_array_equal(back[:ni], P[:ni])


def test_restart_onto_another_mesh(tmp_path):
    """Each inner cell of a finer mesh takes the state of the nearest written centre (the
    reference's kd-tree search, read.c:114-151), checked by brute force."""
    prefix = str(tmp_path / "run")
    run("write", prefix, 4)
    run("read", prefix + "_00003.hdf", 7, str(tmp_path / "fine"))
    _, coarse = mesh_of(4)
    _, fine = mesh_of(7)
    nc = coarse["off_periodic"]
    P = np.fromfile(prefix + "_state.bin").reshape(-1, 5)
    written = P.copy()
    cen_w = coarse["center"][:nc]
    back = np.fromfile(str(tmp_path / "fine") + ".bin").reshape(-1, 5)
    for c in range(fine["num_inner"]):
        d2 = ((cen_w - fine["center"][c]) ** 2).sum(axis=1)
        k = int(np.argmin(d2))
        if k < coarse["num_inner"]:
            np.testing.assert_array_equal(back[c], written[k])

