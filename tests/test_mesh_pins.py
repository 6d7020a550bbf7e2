"""Mesh generation (src/mesh/generate.c, which the reference build cannot compile without METIS)
pinned by closed forms and invariants: uniform and stretched hex geometry, closed cells and
volume sums on jittered tets, the BFS order of the cell renumbering (visit distance never
decreases), the face order and orientation.  The library's host mesh code is checked
bit-for-bit against the oracle on the same inputs."""
import ctypes as C

import numpy as np
import pytest

import teal_amd
from oracle_lib import Case, arr3, oracle, oracle_mesh
from teal_amd import _abi


def hex_mesh(n, lo, hi, periodic):
    return _abi.mesh_arrays(oracle_mesh(Case(n, lo=lo, hi=hi, periodic=periodic)))


def tet_arrays(n=4, jitter=0.25, seed=3):
    lib = teal_amd.lib()
    o = oracle()
    o.oracle_mesh_copy_created.restype = C.c_void_p
    o.oracle_mesh_copy_created.argtypes = [C.c_void_p]
    pm = lib.mesh_create_tets(n, jitter, seed)
    om = o.oracle_mesh_copy_created(pm)
    lib.mesh_destroy(pm)
    o.oracle_mesh_generate(om)
    return _abi.mesh_arrays(om)


BOXES = [((5, 4, 3), (0.0, -1.0, 2.0), (1.0, 1.0, 2.6), (0, 0, 0)),
         ((4, 4, 4), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (1, 1, 1)),
         ((6, 3, 2), (0.0, 0.0, 0.0), (3.0, 1.5, 0.5), (1, 0, 0))]


@pytest.mark.parametrize("n,lo,hi,periodic", BOXES)
def test_hex_geometry_closed_form(n, lo, hi, periodic):
    a = hex_mesh(n, lo, hi, periodic)
    h = (np.array(hi) - np.array(lo)) / np.array(n)
    ni, nfi = a["num_inner"], a["num_inner_faces"]
    vol = a["volume"][:ni]
    np.testing.assert_allclose(vol, np.prod(h), rtol=1e-14)
    assert abs(a["sum_volume"] - np.prod(np.array(hi) - np.array(lo))) <= 1e-13 * a["sum_volume"]
    x = a["center"][:ni]
    idx = np.rint((x - np.array(lo)) / h - 0.5)
    np.testing.assert_allclose(x, np.array(lo) + (idx + 0.5) * h, rtol=0, atol=1e-14)
    # projection = 1/2 sum_f A |n| = (hy hz, hx hz, hx hy)
    np.testing.assert_allclose(a["projection"][:ni],
                               np.broadcast_to([h[1] * h[2], h[0] * h[2], h[0] * h[1]], (ni, 3)),
                               rtol=1e-14)
    basis = a["basis"].reshape(-1, 3, 3)
    normal = basis[:, 0]
    lr = a["cell_idx"]
    # unit axis-aligned normals pointing from left to right, area = product of the other sides
    axis = np.argmax(np.abs(normal), axis=1)
    np.testing.assert_allclose(np.abs(normal)[np.arange(len(axis)), axis], 1.0, rtol=1e-15)
    areas = np.prod(h) / h[axis]
    np.testing.assert_allclose(a["area"], areas, rtol=1e-14)
    d = a["center"][lr[:, 1]] - a["center"][lr[:, 0]]
    assert ((d * normal).sum(axis=1) > 0).all()
    # orthonormal right-handed bases
    np.testing.assert_allclose(np.einsum("fij,fkj->fik", basis, basis),
                               np.broadcast_to(np.eye(3), basis.shape), atol=1e-15)
    np.testing.assert_allclose(np.linalg.det(basis), 1.0, rtol=1e-14)
    # corrections: unit vector along the centre difference, norm = the cell spacing
    corr = a["correction"]
    np.testing.assert_allclose(corr[:, 3], h[axis], rtol=1e-14)
    np.testing.assert_allclose(corr[:, :3], normal, atol=1e-15)
    # face offsets: +-h/2 along the normal
    off = a["face_offset"].reshape(-1, 2, 3)
    np.testing.assert_allclose(off[:, 0], 0.5 * h[axis][:, None] * normal, atol=1e-14)
    np.testing.assert_allclose(off[:nfi, 1], -0.5 * h[axis][:nfi, None] * normal[:nfi], atol=1e-14)


def test_tet_geometry_invariants():
    a = tet_arrays()
    ni, nf = a["num_inner"], a["num_faces"]
    assert abs(a["sum_volume"] - 1.0) <= 1e-14
    assert (a["volume"][:ni] > 0).all()
    lr = a["cell_idx"]
    n = a["basis"].reshape(-1, 3, 3)[:, 0]
    area = a["area"]
    # closed cells: sum of outward A n over each inner cell's faces vanishes
    closed = np.zeros((ni, 3))
    np.add.at(closed, lr[:, 0], area[:, None] * n)
    inner = lr[:, 1] < ni
    np.add.at(closed, lr[inner, 1], -area[inner, None] * n[inner])
    assert np.abs(closed).max() <= 1e-15
    # projection = 1/2 sum_f A |n|
    proj = np.zeros((ni, 3))
    np.add.at(proj, lr[:, 0], 0.5 * area[:, None] * np.abs(n))
    np.add.at(proj, lr[inner, 1], 0.5 * area[inner, None] * np.abs(n[inner]))
    np.testing.assert_allclose(a["projection"][:ni], proj, rtol=1e-13)
    # triangle faces: area and centroid from the face nodes
    coord = a["coord"]
    fo, fi = a["face_node_off"], a["face_node_idx"]
    for f in range(0, nf, 7):
        p = coord[fi[fo[f]:fo[f + 1]]]
        assert len(p) == 3
        cr = np.cross(p[1] - p[0], p[2] - p[0])
        np.testing.assert_allclose(area[f], 0.5 * np.linalg.norm(cr), rtol=1e-13)
        np.testing.assert_allclose(abs(cr @ n[f]) / np.linalg.norm(cr), 1.0, rtol=1e-13)
        np.testing.assert_allclose(a["face_center"][f], p.mean(axis=0), atol=1e-15)
    # normals point away from the left cell
    assert (((a["face_center"] - a["center"][lr[:, 0]]) * n).sum(axis=1) > 0).all()


@pytest.mark.parametrize("mesh", ["hex", "tets"])
def test_cell_numbering_is_a_bfs_order(mesh):
    """src/mesh/generate.c:68-133: breadth-first from cell 0, so the graph distance from cell 0
    never decreases along the numbering."""
    a = hex_mesh((6, 5, 4), (0, 0, 0), (1, 1, 1), (1, 0, 1)) if mesh == "hex" else tet_arrays()
    ni = a["num_inner"]
    off, idx = a["cell_cell_off"], a["cell_cell_idx"]
    dist = np.full(ni, -1)
    dist[0] = 0
    queue = [0]
    for c in queue:
        for k in idx[off[c]:off[c + 1]]:
            if k < ni and dist[k] < 0:
                dist[k] = dist[c] + 1
                queue.append(k)
    assert (dist >= 0).all()
    assert (np.diff(dist) >= 0).all()
    # within a BFS the first-numbered neighbour of each cell > 0 is its parent, numbered earlier
    for c in range(1, ni):
        nb = idx[off[c]:off[c + 1]]
        assert nb[nb < ni].min() < c


@pytest.mark.parametrize("mesh", ["hex", "tets"])
def test_face_order(mesh):
    """src/mesh/generate.c:214-225, 294-315: inner faces first, sorted by left cell; boundary and
    periodic faces sorted by right (ghost) cell; left < right."""
    a = hex_mesh((6, 5, 4), (0, 0, 0), (1, 1, 1), (1, 0, 1)) if mesh == "hex" else tet_arrays()
    lr = a["cell_idx"]
    ni, nfi, nfb = a["num_inner"], a["num_inner_faces"], a["off_boundary_faces"]
    assert (lr[:, 0] < lr[:, 1]).all()
    assert (lr[:nfi, 1] < ni).all() and (lr[nfi:, 1] >= ni).all()
    assert (np.diff(lr[:nfi, 0]) >= 0).all()
    assert (np.diff(lr[nfi:nfb, 1]) > 0).all() and (np.diff(lr[nfb:, 1]) > 0).all()
    # every inner cell's face list matches its neighbour list
    off, idx = a["cell_cell_off"], a["cell_cell_idx"]
    for c in range(ni):
        faces = np.concatenate([lr[lr[:, 0] == c, 1], lr[lr[:, 1] == c, 0]])
        assert sorted(faces) == sorted(idx[off[c]:off[c + 1]])


@pytest.mark.parametrize("kind", ["hex", "tets"])
def test_library_mesh_generate_matches_oracle(kind):
    """The library's mesh_generate (teal_amd/csrc/mesh_generate.c) vs the oracle's, bitwise."""
    lib = teal_amd.lib()
    o = oracle()
    o.oracle_mesh_copy_created.restype = C.c_void_p
    o.oracle_mesh_copy_created.argtypes = [C.c_void_p]
    if kind == "hex":
        pm = lib.mesh_create(teal_amd.Vector(0, 0, 0), teal_amd.Vector(1, 2, 3),
                             teal_amd.Triple(7, 5, 4), teal_amd.Triple(1, 0, 1))
    else:
        pm = lib.mesh_create_tets(5, 0.2, 99)
    om = o.oracle_mesh_copy_created(pm)
    o.oracle_mesh_generate(om)
    lib.mesh_generate(pm)
    lib.mesh_validate(pm)
    a, b = _abi.mesh_arrays(pm), _abi.mesh_arrays(om)
    for k in a:
        if isinstance(a[k], np.ndarray):
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)
        else:
            assert a[k] == b[k], k
    lib.mesh_destroy(pm)
