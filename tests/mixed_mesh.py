"""Test meshes of mixed elements that mesh_create cannot make: a box of hexahedra in which
chosen columns are split into two prisms each (the hexahedron's own prism split,
src/mesh/generate.c:492-495), with one ghost cell per boundary face.  Hexahedra have six
face neighbours and prisms five, so the breadth-first renumbering's sort by degree
(src/mesh/generate.c:106-115) has work to do."""
import ctypes as C

import numpy as np

from oracle_lib import oracle

SIDES = ("left", "right", "bottom", "top", "back", "front")

HEX_FACES = ((0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7))
PRISM_FACES = ((0, 2, 1), (3, 4, 5), (0, 1, 4, 3), (1, 2, 5, 4), (2, 0, 3, 5))


def mixed_box(n, h, split, lo=(0.0, 0.0, 0.0)):
    """n cells per axis of spacing h; `split` = set of (i, j) columns cut into prisms.
    Returns (coord [nn][3], cells: list of node lists, num_inner, entity names, cell_off)."""
    nx, ny, nz = n
    node = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)  # noqa: E731
    coord = np.array([[lo[0] + i * h[0], lo[1] + j * h[1], lo[2] + k * h[2]]
                      for k in range(nz + 1) for j in range(ny + 1) for i in range(nx + 1)])
    cells = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                p = [node(i, j, k), node(i + 1, j, k), node(i + 1, j + 1, k), node(i, j + 1, k),
                     node(i, j, k + 1), node(i + 1, j, k + 1), node(i + 1, j + 1, k + 1),
                     node(i, j + 1, k + 1)]
                if (i, j) in split:
                    cells.append([p[0], p[1], p[2], p[4], p[5], p[6]])
                    cells.append([p[0], p[2], p[3], p[4], p[6], p[7]])
                else:
                    cells.append(p)
    ni = len(cells)
    count = {}
    faces = []
    for c in cells:
        for f in (HEX_FACES if len(c) == 8 else PRISM_FACES):
            nodes = [c[v] for v in f]
            key = frozenset(nodes)
            count[key] = count.get(key, 0) + 1
            faces.append(nodes)
    hi = [lo[d] + n[d] * h[d] for d in range(3)]
    ghosts = {s: [] for s in SIDES}
    for nodes in faces:
        if count[frozenset(nodes)] != 1:
            continue
        x = coord[nodes]
        for d in range(3):
            if (x[:, d] == lo[d]).all():
                ghosts[SIDES[2 * d]].append(nodes)
            elif (x[:, d] == hi[d]).all():
                ghosts[SIDES[2 * d + 1]].append(nodes)
    names = ["domain"]
    off = [0, ni]
    for s in SIDES:
        if ghosts[s]:
            cells.extend(ghosts[s])
            names.append(s)
            off.append(off[-1] + len(ghosts[s]))
    return coord, cells, ni, names, off


def oracle_mixed_mesh(coord, cells, ni, names, off):
    """The created mesh as an oracle OMesh (not generated)."""
    node_off = np.cumsum([0] + [len(c) for c in cells]).astype(np.int32)
    node_idx = np.array([v for c in cells for v in c], dtype=np.int32)
    buf = b"".join(nm.encode().ljust(128, b"\0") for nm in names)
    xyz = np.ascontiguousarray(coord, dtype=np.float64)
    eoff = np.array(off, dtype=np.int32)
    IP = C.POINTER(C.c_int)
    return oracle().oracle_mesh_from_cells(
        len(coord), xyz.ctypes.data_as(C.POINTER(C.c_double)), len(cells), ni,
        node_off.ctypes.data_as(IP), node_idx.ctypes.data_as(IP), len(names), buf,
        eoff.ctypes.data_as(IP))
