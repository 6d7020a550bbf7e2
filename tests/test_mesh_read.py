"""mesh_read (src/mesh/read.c:901-918, src/mesh/gmsh.c:135-904) on the CPU.

The reference ships no .msh file, so the files are written by tests/gmsh_file.py from meshes
whose reading is known: the library's own Cartesian box (mesh_create, pinned against the
reference's build) and jittered tet mesh, written node for node and cell for cell, with
physical groups in the mesh's entity order.  mesh_read followed by mesh_generate must then
give the generated mesh exactly: every connectivity and geometry array, bit for bit, and the
same periodic exchange sources (the exchange tags are entity indices in mesh_read, side
numbers in mesh_create: src/mesh/read.c:310-372 against src/mesh/create.c:388-440).  ASCII,
binary and byte-swapped binary files read alike; malformed files end the program through
teal_error with the reference's messages."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

import teal_amd
from gmsh_file import Msh, from_mesh
from teal_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def box(n, periodic, lo=(0.0, 0.0, 0.0), hi=(1.0, 2.0, 0.5)):
    lib = teal_amd.lib()
    return lib.mesh_create(teal_amd.Vector(*lo), teal_amd.Vector(*hi), teal_amd.Triple(*n),
                           teal_amd.Triple(*periodic))


def sources(mesh):
    a = _abi.mesh_arrays(mesh)
    src = np.full(a["num_cells"], -1, dtype=np.int32)
    teal_amd.lib().teal_mesh_periodic_sources(C.c_void_p(mesh), src.ctypes.data_as(C.c_void_p))
    return src


def assert_same_mesh(got_ptr, want_ptr, skip=()):
    got, want = _abi.mesh_arrays(got_ptr), _abi.mesh_arrays(want_ptr)
    assert set(got) == set(want)
    for key, w in want.items():
        if key in skip:
            continue
        g = got[key]
        if isinstance(w, np.ndarray):
            np.testing.assert_array_equal(g, w, err_msg=key)
        else:
            assert g == w, key
    np.testing.assert_array_equal(sources(got_ptr), sources(want_ptr))


def read(path):
    mesh = teal_amd.lib().mesh_read(str(path).encode())
    assert mesh
    return mesh


def box_file(tmp_path, n, periodic, name, binary=False, swap=False):
    lib = teal_amd.lib()
    raw = box(n, periodic)
    span = (1.0, 2.0, 0.5)
    links = []
    for axis, (lo_side, hi_side) in enumerate((("left", "right"), ("bottom", "top"), ("back", "front"))):
        if periodic[axis]:
            t = [0.0, 0.0, 0.0]
            t[axis] = span[axis]
            links.append((hi_side, lo_side, t))
    path = tmp_path / name
    from_mesh(_abi.mesh_arrays(raw), periodic_links=links).write(path, binary=binary, swap=swap)
    lib.mesh_generate(raw)
    return path, raw


@pytest.mark.parametrize("periodic", [(0, 0, 0), (1, 0, 1), (1, 1, 1)], ids=["walls", "xz", "xyz"])
@pytest.mark.parametrize("encoding", ["ascii", "binary", "swapped"])
def test_box_reads_as_mesh_create(tmp_path, periodic, encoding):
    lib = teal_amd.lib()
    path, want = box_file(tmp_path, (4, 3, 2), periodic, "box.msh", binary=encoding != "ascii",
                          swap=encoding == "swapped")
    got = read(path)
    lib.mesh_generate(got)
    assert_same_mesh(got, want, skip=("neighbors_tag",))
    a = _abi.mesh_arrays(got)
    # exchange entries: tag = (entity, partner entity), read.c:310-372
    ne, nb = len(a["entity_names"]), a["entities_off_boundary"]
    assert a["neighbors_num"] == ne - nb
    for m, (e, p) in enumerate(a["neighbors_tag"]):
        assert e == nb + m and nb <= p < ne and a["neighbors_tag"][p - nb][1] == e
    lib.mesh_destroy(got)
    lib.mesh_destroy(want)


def test_tets_read_as_generated(tmp_path):
    lib = teal_amd.lib()
    raw = lib.mesh_create_tets(5, 0.2, 11)
    path = tmp_path / "tets.msh"
    from_mesh(_abi.mesh_arrays(raw)).write(path, binary=True)
    lib.mesh_generate(raw)
    got = read(path)
    lib.mesh_generate(got)
    assert_same_mesh(got, raw)
    lib.mesh_destroy(got)
    lib.mesh_destroy(raw)


def test_node_order_and_unused_nodes(tmp_path):
    """Nodes are those the cells use, in file order: an extra unused node block and a
    shuffled tag numbering change nothing."""
    lib = teal_amd.lib()
    raw = box((3, 2, 2), (0, 0, 0))
    a = _abi.mesh_arrays(raw)
    tags = np.random.default_rng(3).permutation(np.arange(1000, 1000 + a["num_nodes"]))
    m = from_mesh(a, node_tags=tags)
    m.node_blocks.insert(0, (0, 7, np.array([5, 6]), np.array([[9.0, 9, 9], [8, 8, 8]])))
    path = tmp_path / "shuffled.msh"
    m.write(path)
    lib.mesh_generate(raw)
    got = read(path)
    lib.mesh_generate(got)
    assert_same_mesh(got, raw)
    lib.mesh_destroy(got)
    lib.mesh_destroy(raw)


def _fails_with(path, message):
    code = ("import sys; sys.path.insert(0, %r); import teal_amd; "
            "teal_amd.lib().mesh_read(%r.encode())" % (ROOT, str(path)))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert message in r.stderr + r.stdout, r.stderr + r.stdout


def _small(tmp_path):
    raw = box((2, 2, 1), (0, 0, 0))
    m = from_mesh(_abi.mesh_arrays(raw))
    teal_amd.lib().mesh_destroy(raw)
    return m


def test_errors(tmp_path):
    m = _small(tmp_path)
    m.write(tmp_path / "ok.msh")
    text = (tmp_path / "ok.msh").read_text()
    bad = tmp_path / "version.msh"
    bad.write_text(text.replace("4.1 0 8", "4 0 8", 1))
    _fails_with(bad, "unsupported version (4)")
    bad = tmp_path / "size.msh"
    bad.write_text(text.replace("4.1 0 8", "4.1 0 4", 1))
    _fails_with(bad, "unsupported data-size (4)")
    bad = tmp_path / "section.msh"
    bad.write_text(text + "$Bogus\n$EndBogus\n")
    _fails_with(bad, "unsupported mesh section ($Bogus)")
    _fails_with(tmp_path / "missing.msh", "could not open file")
    _fails_with(tmp_path / "mesh.vtk", "unsupported file extension (.vtk)")

    m = _small(tmp_path)
    m.surfaces[0] = (m.surfaces[0][0], [2, 3])
    m.write(tmp_path / "tags.msh")
    _fails_with(tmp_path / "tags.msh", "unsupported number of physical tags (2)")

    m = _small(tmp_path)
    m.physicals = m.physicals[:-1]
    m.write(tmp_path / "name.msh")
    _fails_with(tmp_path / "name.msh", "could not find physical name")

    m = _small(tmp_path)
    dim, tag, _, tags, nodes = m.blocks[1]
    m.blocks[1] = (dim, tag, 1, tags, nodes[:, :2])  # 2-node lines
    m.write(tmp_path / "type.msh")
    _fails_with(tmp_path / "type.msh", "unsupported element type (1)")

    m = _small(tmp_path)
    m.blocks = m.blocks[1:]  # faces only
    m.write(tmp_path / "novol.msh")
    _fails_with(tmp_path / "novol.msh", "mesh does not contain any volumes")


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [teal_amd.FP64], ids=["fp64"])
def test_read_mesh_runs_on_device(tmp_path, precision):
    """A periodic-x Navier-Stokes channel box written to a binary .msh and read back feeds the
    device layout: its residual and three LSERK steps on the device equal the oracle's on the
    box it was written from, bit for bit."""
    from cases import gpu_bcs
    from oracle_lib import Case, dp, oracle, oracle_eqns, oracle_mesh, random_state

    lib = teal_amd.lib()
    case = Case((6, 5, 4), lo=(0.0, 0.0, 0.0), hi=(1.0, 2.0, 0.5), periodic=(1, 0, 1),
                kind="navierstokes", mu=0.02, bcs=[("bottom", "wall", None), ("top", "wall", None)])
    print("read mesh on device", flush=True)
    path, want = box_file(tmp_path, case.n, case.periodic, "dev.msh", binary=True)
    lib.mesh_destroy(want)
    mesh = read(path)
    lib.mesh_generate(mesh)
    om = oracle_mesh(case)
    oe = oracle_eqns(case, om)
    a = _abi.mesh_arrays(om)
    nc, ni = a["num_cells"], a["num_inner"]
    gpu = teal_amd.Gpu(mesh, precision).configure(
        space_order=2, flux="hllc", viscous=1, limiter="minmod", navierstokes=1, gamma=1.4,
        mu=case.mu, prandtl=case.prandtl, bcs=gpu_bcs(case))
    P = random_state(nc, 21, vel=0.3)
    P[ni:] = 0.0
    Ro = np.zeros((ni, 5))
    oracle().oracle_derivative(oe, dp(P.copy()), dp(Ro))
    Rg, _ = gpu.derivative(P.copy())
    np.testing.assert_array_equal(Rg, Ro)
    po = np.ctypeslib.as_array(oracle().oracle_eqns_primitive(oe), shape=(nc * 5,))
    po[:] = P.ravel()
    gpu.upload(P)
    t_o, res_o, t_g = C.c_double(0.0), (C.c_double * 5)(), 0.0
    for _ in range(3):
        s_o = oracle().oracle_lserk(oe, C.byref(t_o), res_o, 1e300, 0.5, 2, 3)
        s_g, t_g, _ = gpu.lserk(t_g, 0.5)
        assert s_g == s_o
    np.testing.assert_array_equal(gpu.download()[:ni], po.reshape(nc, 5)[:ni])
    gpu.close()
    lib.mesh_destroy(mesh)


def test_mixed_hex_prism_mesh_against_oracle(tmp_path):
    """A box with columns split into prisms (two element types in the volume, quads and
    triangles on the z sides) written with one element block per type: mesh_read keeps the
    blocks' order, and after mesh_generate equals the oracle's mesh_generate on the same cells
    in that order, bit for bit."""
    from mixed_mesh import mixed_box, oracle_mixed_mesh
    from oracle_lib import oracle
    from gmsh_file import file_order, from_cells

    lib = teal_amd.lib()
    coord, cells, ni, names, off = mixed_box((4, 3, 2), (0.5, 0.4, 0.3), {(1, 0), (2, 1), (0, 2)})
    path = tmp_path / "mixed.msh"
    from_cells(coord, cells, 1, names, off).write(path)
    order = file_order(cells, 1, off)
    om = oracle_mixed_mesh(coord, [cells[c] for c in order], ni, names, off)
    oracle().oracle_mesh_generate(om)
    got = read(path)
    lib.mesh_generate(got)
    want = _abi.mesh_arrays(om)
    a = _abi.mesh_arrays(got)
    assert a["num_inner"] == ni and len(set(len(c) for c in cells[:ni])) == 2
    for key, w in want.items():
        if isinstance(w, np.ndarray):
            np.testing.assert_array_equal(a[key], w, err_msg=key)
        else:
            assert a[key] == w, key
    lib.mesh_destroy(got)
