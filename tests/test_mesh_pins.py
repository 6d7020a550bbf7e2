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


# ---------------------------------------------------------------------------------------------
# The integer renumbering of mesh_generate, restated step by step from the reference and
# compared array for array with the oracle.  The dual graph comes from the published METIS rule
# (brute_dual, itself checked against the oracle's dual); everything after it is the
# reference's own logic, so a mistake in the oracle's visit order within a BFS level, in the
# degree sort, in the node renumbering or in the order of faces that share a left cell fails
# here.

def restate_connectivity(coord, cells, ni):
    """src/mesh/generate.c:15-66 (connect_cells), 68-133 (reorder_cells), 171-204
    (reorder_nodes), 206-321 (create_faces) and src/mesh/reorder.c:34-135 on one rank."""
    from test_oracle_closed_form import brute_dual

    nc = len(cells)
    nn = len(coord)
    # connect_cells: one dummy node per ghost enforces ncommon = 3 (generate.c:29-32)
    eptr, eind, dummy = [0], [], nn
    for i, c in enumerate(cells):
        eind += list(c)
        if i >= ni:
            eind.append(dummy)
            dummy += 1
        eptr.append(len(eind))
    dual = brute_dual(eptr, eind, 3)
    nbr = [[j for j in dual[i] if i < ni or j < ni] for i in range(nc)]  # generate.c:51
    # reorder_cells: breadth-first from cell 0; each cell's newly found neighbours in list
    # order, then stably sorted by descending degree (the insertion sort of generate.c:106-115)
    degree = [len(nbr[i]) for i in range(ni)]
    new = [-1] * ni
    num, seed = 0, 0
    moved = 0  # lists whose order the degree sort changed
    while num < ni:
        while new[seed] != -1:
            seed = (seed + 1) % ni
        new[seed] = num
        num += 1
        queue = [seed]
        for cur in queue:
            found = [j for j in nbr[cur] if j < ni and new[j] == -1]
            for j in found:
                new[j] = -2
            before = list(found)
            found.sort(key=lambda j: -degree[j])
            moved += found != before
            for j in found:
                new[j] = num
                num += 1
                queue.append(j)
    # mesh_reorder_cells(map, 0, ni): rows move, each list keeps its order, indices remapped
    remap = lambda j: new[j] if j < ni else j  # noqa: E731
    order = sorted(range(ni), key=lambda i: new[i]) + list(range(ni, nc))
    cell_nodes = [list(cells[i]) for i in order]
    cell_nbr = [[remap(j) for j in nbr[i]] for i in order]
    # reorder_nodes: numbered by first appearance over the cells' node lists
    nmap = {}
    for c in cell_nodes:
        for v in c:
            if v not in nmap:
                nmap[v] = len(nmap)
    assert len(nmap) == nn
    cell_nodes = [[nmap[v] for v in c] for c in cell_nodes]
    new_coord = np.zeros_like(coord)
    for v, w in nmap.items():
        new_coord[w] = coord[v]
    # create_faces: left < right over the inner cells' lists, nodes in the left cell's order
    faces = []
    for i in range(ni):
        for j in cell_nbr[i]:
            if i < j:
                common = [v for v in cell_nodes[i] for w in cell_nodes[j] if v == w]
                if len(common) == 4:  # correct_face_node_order, generate.c:239-265
                    for t in range(3):
                        x = new_coord[common]
                        a2b, a2c, a2d = x[1] - x[0], x[2] - x[0], x[3] - x[0]
                        if np.dot(np.cross(a2b, a2c), np.cross(a2c, a2d)) > 0:
                            break
                        common[t + 1], common[t + 2] = common[t + 2], common[t + 1]
                    else:
                        raise AssertionError("quad vertex order could not be made valid")
                faces.append((j < ni, i, j, common))
    # inner faces first by left, the others by right; ties keep creation order
    faces.sort(key=lambda f: (0, f[1]) if f[0] else (1, f[2]))
    return {
        "cell_node_off": np.cumsum([0] + [len(c) for c in cell_nodes]),
        "cell_node_idx": np.array([v for c in cell_nodes for v in c]),
        "cell_cell_off": np.cumsum([0] + [len(c) for c in cell_nbr]),
        "cell_cell_idx": np.array([v for c in cell_nbr for v in c]),
        "coord": new_coord,
        "cell_idx": np.array([[f[1], f[2]] for f in faces]),
        "face_node_off": np.cumsum([0] + [len(f[3]) for f in faces]),
        "face_node_idx": np.array([v for f in faces for v in f[3]]),
        "num_inner_faces": sum(1 for f in faces if f[0]),
        "degrees": sorted(set(degree)),
        "moved": moved,
    }


def _created_hex(n, lo, hi):
    """mesh_create's cells of a non-periodic box (oracle), as node lists."""
    m = oracle().oracle_mesh_create(arr3(lo), arr3(hi), arr3(n, C.c_int), arr3((0, 0, 0), C.c_int))
    a = _abi.mesh_arrays(m)
    off, idx = a["cell_node_off"], a["cell_node_idx"]
    cells = [list(idx[off[i]:off[i + 1]]) for i in range(a["num_cells"])]
    return m, a["coord"].copy(), cells, a["num_inner"]


MIXED = [((4, 3, 2), (1.0, 0.5, 0.25), {(1, 1), (2, 0)}),
         ((3, 3, 1), (0.5, 0.5, 1.0), {(0, 0), (1, 1), (2, 2)}),
         ((2, 4, 3), (1.0, 1.0, 0.5), {(1, 2)})]


@pytest.mark.parametrize("n,h,split", MIXED)
def test_renumbering_restated_on_mixed_elements(n, h, split):
    """Cells of degree 5 (prisms) and 6 (hexahedra) in one mesh: the degree sort reorders each
    BFS level, and faces that share a left cell come out in that cell's neighbour order."""
    from mixed_mesh import mixed_box, oracle_mixed_mesh

    coord, cells, ni, names, off = mixed_box(n, h, split)
    want = restate_connectivity(coord, cells, ni)
    assert want["degrees"] == [5, 6] and want["moved"] > 0
    m = oracle_mixed_mesh(coord, cells, ni, names, off)
    oracle().oracle_mesh_generate(m)
    a = _abi.mesh_arrays(m)
    for k in ("cell_node_off", "cell_node_idx", "cell_cell_off", "cell_cell_idx", "coord",
              "cell_idx", "face_node_off", "face_node_idx", "num_inner_faces"):
        np.testing.assert_array_equal(a[k], want[k], err_msg=k)
    oracle().oracle_mesh_destroy(m)


@pytest.mark.parametrize("n", [(3, 2, 2), (4, 1, 3), (2, 3, 1)])
def test_renumbering_restated_on_boxes(n):
    """The same restatement on mesh_create's boxes (every cell of degree 6, so the BFS order
    is the neighbour-list order alone)."""
    m, coord, cells, ni = _created_hex(n, (0, 0, 0), (1, 1, 1))
    want = restate_connectivity(coord, cells, ni)
    oracle().oracle_mesh_generate(m)
    a = _abi.mesh_arrays(m)
    for k in ("cell_node_off", "cell_node_idx", "cell_cell_off", "cell_cell_idx", "coord",
              "cell_idx", "face_node_off", "face_node_idx", "num_inner_faces"):
        np.testing.assert_array_equal(a[k], want[k], err_msg=k)
    oracle().oracle_mesh_destroy(m)


@pytest.mark.parametrize("h,periodic", [((3.0, 2.0, 1.0), (0, 0, 0)), ((1.0, 3.0, 0.25), (1, 0, 1)),
                                        ((6.0, 0.5, 2.0), (0, 1, 0))])
def test_hex_geometry_bit_exact(h, periodic):
    """Spacings whose product is a multiple of 3: every tetrahedron of the reference's volume
    and centre decomposition (src/mesh/generate.c:468-550) then has a dyadic volume, and each
    geometric quantity the reference computes is exactly representable, so the oracle must
    produce it bit for bit: volumes, cell centres (ghosts mirrored or translated), face areas,
    centres and normals, cell->face offsets, corrections and projections."""
    n = (3, 2, 2)
    hh = np.array(h)
    hi = tuple(hh * np.array(n))
    a = hex_mesh(n, (0.0, 0.0, 0.0), hi, periodic)
    ni, nc, nf = a["num_inner"], a["num_cells"], a["num_faces"]
    assert (a["volume"][:ni] == np.prod(hh)).all()
    assert a["sum_volume"] == np.prod(hi)
    x = a["center"]
    idx = np.floor(x[:ni] / hh)
    np.testing.assert_array_equal(x[:ni], (idx + 0.5) * hh)
    # ghosts (boundary: mirrored across the face; periodic: translated by the domain length)
    # sit one spacing beyond the face, on the same lattice
    gidx = np.floor(x[ni:] / hh)
    np.testing.assert_array_equal(x[ni:], (gidx + 0.5) * hh)
    np.testing.assert_array_equal(a["projection"][:ni],
                                  np.broadcast_to([h[1] * h[2], h[0] * h[2], h[0] * h[1]], (ni, 3)))
    normal = a["basis"].reshape(-1, 3, 3)[:, 0]
    axis = np.argmax(np.abs(normal), axis=1)
    unit = np.eye(3)[axis]
    np.testing.assert_array_equal(np.abs(normal), unit)
    np.testing.assert_array_equal(a["area"], np.prod(hh) / hh[axis])
    lr = a["cell_idx"]
    fc = a["face_center"]
    # the face centre is the midpoint of its two cells' centres, exactly
    np.testing.assert_array_equal(fc, (x[lr[:, 0]] + x[lr[:, 1]]) / 2)
    off = a["face_offset"].reshape(-1, 2, 3)
    np.testing.assert_array_equal(off[:, 0], fc - x[lr[:, 0]])
    np.testing.assert_array_equal(off[:, 0], 0.5 * hh[axis][:, None] * normal)
    np.testing.assert_array_equal(off[:, 1], fc - x[lr[:, 1]])
    corr = a["correction"]
    np.testing.assert_array_equal(corr[:, 3], hh[axis])
    np.testing.assert_array_equal(corr[:, :3], normal)
    # cell -> face offsets per neighbour entry (generate.c:629-667): half the way to the
    # neighbour's centre
    coff, cidx = a["cell_cell_off"], a["cell_cell_idx"]
    cof = a["cell_offset"].reshape(-1, 3)
    for c in range(ni):
        for e in range(coff[c], coff[c + 1]):
            np.testing.assert_array_equal(cof[e], (x[cidx[e]] - x[c]) / 2)


def _tangents_from_normal(n):
    """The reference's tangent construction (src/mesh/generate.c:420-435) from normals n [k][3],
    and which branch each face takes."""
    nx, ny, nz = n[:, 0], n[:, 1], n[:, 2]
    nqz, nqy = np.hypot(nx, ny), np.hypot(nx, nz)
    first = nqz > nqy
    with np.errstate(divide="ignore", invalid="ignore"):
        t1 = np.where(first[:, None], np.stack([-ny / nqz, nx / nqz, 0 * nx], 1),
                      np.stack([-nx * ny / nqy, nqy, -ny * nz / nqy], 1))
        t2 = np.where(first[:, None], np.stack([-nx * nz / nqz, -ny * nz / nqz, nqz], 1),
                      np.stack([-nz / nqy, 0 * nx, nx / nqy], 1))
    return t1, t2, first


@pytest.mark.parametrize("kind", ["box", "periodic", "tets", "mixed"])
def test_face_tangents_follow_the_construction(kind):
    """The stored face bases follow generate.c:420-451: tangents built from the unit normal, then
    the normal and the first tangent negated where the normal pointed into the left cell.  So
    the construction applied to the stored normal gives the stored tangents, or (second branch
    of a flipped face) both negated.  The FP64 device records rebuild the tangents this way."""
    if kind == "tets":
        a = tet_arrays()
    elif kind == "mixed":
        from mixed_mesh import mixed_box, oracle_mixed_mesh

        coord, cells, ni, names, off = mixed_box((4, 3, 2), (1.0, 0.5, 0.25), {(1, 1), (2, 0)})
        m = oracle_mixed_mesh(coord, cells, ni, names, off)
        oracle().oracle_mesh_generate(m)
        a = _abi.mesh_arrays(m)
    else:
        a = hex_mesh(*BOXES[0 if kind == "box" else 1])
    b = np.asarray(a["basis"]).reshape(-1, 9)
    t1, t2, first = _tangents_from_normal(b[:, 0:3])
    same = np.all(t1 == b[:, 3:6], 1) & np.all(t2 == b[:, 6:9], 1)
    negated = np.all(t1 == -b[:, 3:6], 1) & np.all(t2 == -b[:, 6:9], 1)
    assert np.all(same | (negated & ~first))
    assert np.any(negated & ~same)  # some faces are flipped second-branch faces
    if kind == "mixed":
        oracle().oracle_mesh_destroy(m)
