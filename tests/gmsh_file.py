"""Writer of Gmsh 4.1 mesh files (the format src/mesh/gmsh.c reads), ASCII or binary, for the
mesh_read tests: the reference ships no .msh file (test/1_mesh/mesh.geo needs gmsh, which the
image lacks), so the tests write meshes whose expected reading is known -- the library's own
generated box and tet meshes, node for node and cell for cell.

Layout written (Gmsh 4.1 reference manual, section "MSH file format"):
  $MeshFormat  4.1 <0|1> 8 [int 1 in binary]
  $PhysicalNames (always text)
  $Entities    points, curves (none), surfaces and volumes with their physical tags
  $Nodes       one block per given node block
  $Elements    one block per given element block
  $Periodic    links: dim, tag, master tag, 16 affine values, node pairs (tag, master tag)
"""
import struct

import numpy as np

ELEMENT_NODES = {2: 3, 3: 4, 4: 4, 5: 8, 6: 6, 7: 5}


class Msh:
    def __init__(self):
        self.physicals = []   # (dim, tag, name)
        self.surfaces = []    # (tag, [physical tags])
        self.volumes = []     # (tag, [physical tags])
        self.node_blocks = [] # (dim, tag, tags[n], coords[n, 3])
        self.blocks = []      # (dim, entity tag, element type, element tags[m], node tags[m, len])
        self.links = []       # (dim, tag, master tag, affine[16], pairs[k, 2])

    def write(self, path, binary=False, swap=False):
        end = ">" if swap else "<"
        out = bytearray()

        def text(s):
            out.extend(s.encode())

        def pack(fmt, *vals):
            out.extend(struct.pack(end + fmt, *vals))

        def arr(dtype, values):
            a = np.ascontiguousarray(values, dtype=np.dtype(dtype).newbyteorder(end))
            out.extend(a.tobytes())

        text(f"$MeshFormat\n4.1 {1 if binary else 0} 8\n")
        if binary:
            pack("i", 1)
            text("\n")
        text("$EndMeshFormat\n$PhysicalNames\n")
        text(f"{len(self.physicals)}\n")
        for dim, tag, name in self.physicals:
            text(f'{dim} {tag} "{name}"\n')
        text("$EndPhysicalNames\n$Entities\n")
        lo, hi = np.zeros(3), np.ones(3)
        if binary:
            arr("u8", [0, 0, len(self.surfaces), len(self.volumes)])
            for tag, phys in self.surfaces + self.volumes:
                pack("i", tag)
                arr("f8", np.concatenate([lo, hi]))
                pack("Q", len(phys))
                arr("i4", phys)
                pack("Q", 0)  # bounding entities
        else:
            text(f"0 0 {len(self.surfaces)} {len(self.volumes)}\n")
            for tag, phys in self.surfaces + self.volumes:
                text(f"{tag} 0 0 0 1 1 1 {len(phys)} {' '.join(map(str, phys))} 0\n")
        text("\n$EndEntities\n$Nodes\n" if binary else "$EndEntities\n$Nodes\n")
        num_nodes = sum(len(b[2]) for b in self.node_blocks)
        all_tags = np.concatenate([np.asarray(b[2]) for b in self.node_blocks])
        head = [len(self.node_blocks), num_nodes, int(all_tags.min()), int(all_tags.max())]
        if binary:
            arr("u8", head)
        else:
            text(" ".join(map(str, head)) + "\n")
        for dim, tag, tags, coords in self.node_blocks:
            if binary:
                pack("iiiQ", dim, tag, 0, len(tags))
                arr("u8", tags)
                arr("f8", np.asarray(coords).reshape(-1))
            else:
                text(f"{dim} {tag} 0 {len(tags)}\n")
                text("".join(f"{t}\n" for t in tags))
                text("".join(f"{x!r} {y!r} {z!r}\n" for x, y, z in np.asarray(coords).tolist()))
        text("\n$EndNodes\n$Elements\n" if binary else "$EndNodes\n$Elements\n")
        num_el = sum(len(b[3]) for b in self.blocks)
        el_tags = np.concatenate([np.asarray(b[3]) for b in self.blocks])
        head = [len(self.blocks), num_el, int(el_tags.min()), int(el_tags.max())]
        if binary:
            arr("u8", head)
        else:
            text(" ".join(map(str, head)) + "\n")
        for dim, etag, etype, tags, nodes in self.blocks:
            rec = np.hstack([np.asarray(tags).reshape(-1, 1), np.asarray(nodes)])
            if binary:
                pack("iiiQ", dim, etag, etype, len(tags))
                arr("u8", rec.reshape(-1))
            else:
                text(f"{dim} {etag} {etype} {len(tags)}\n")
                text("".join(" ".join(map(str, r)) + "\n" for r in rec.tolist()))
        text("\n$EndElements\n" if binary else "$EndElements\n")
        if self.links:
            text("$Periodic\n")
            if binary:
                pack("Q", len(self.links))
            else:
                text(f"{len(self.links)}\n")
            for dim, tag, master, affine, pairs in self.links:
                pairs = np.asarray(pairs).reshape(-1, 2)
                if binary:
                    pack("iii", dim, tag, master)
                    pack("Q", len(affine))
                    arr("f8", affine)
                    pack("Q", len(pairs))
                    arr("u8", pairs.reshape(-1))
                else:
                    text(f"{dim} {tag} {master}\n")
                    text(f"{len(affine)} " + " ".join(repr(float(a)) for a in affine) + "\n")
                    text(f"{len(pairs)}\n")
                    text("".join(f"{a} {b}\n" for a, b in pairs.tolist()))
            text("\n$EndPeriodic\n" if binary else "$EndPeriodic\n")
        with open(path, "wb") as f:
            f.write(bytes(out))


ELEMENT_TYPE = {(3, 4): 4, (3, 8): 5, (3, 6): 6, (3, 5): 7, (2, 3): 2, (2, 4): 3}


def from_cells(coord, cells, num_inner_entities, names, off, node_tags=None, periodic_links=()):
    """A Msh of cells given as node lists (indices into coord), entity e holding cells
    off[e] .. off[e+1] (the first `num_inner_entities` entities volumes, the rest faces): one
    node block, and per entity one element block per element type, in the order the types
    first appear (the reader keeps that block order within an entity).  Physical tags follow
    the entity order.  `periodic_links` are (slave entity name, master entity name,
    translation) triples."""
    m = Msh()
    coord = np.asarray(coord, dtype=float)
    nn = len(coord)
    tags = np.arange(1, nn + 1) * 3 + 7 if node_tags is None else np.asarray(node_tags)
    m.node_blocks.append((3, 1, tags, coord))
    el = 1
    for e, name in enumerate(names):
        dim = 3 if e < num_inner_entities else 2
        m.physicals.append((dim, e + 1, name))
        ent_tag = 100 + e
        (m.volumes if dim == 3 else m.surfaces).append((ent_tag, [e + 1]))
        by_type = {}
        for c in range(int(off[e]), int(off[e + 1])):
            nodes = list(cells[c])
            by_type.setdefault(ELEMENT_TYPE[(dim, len(nodes))], []).append(tags[nodes])
        for etype, rows in by_type.items():
            m.blocks.append((dim, ent_tag, etype, np.arange(el, el + len(rows)), np.array(rows)))
            el += len(rows)
    for slave, master, t in periodic_links:
        s, mm = names.index(slave), names.index(master)
        snodes = np.unique([v for c in range(off[s], off[s + 1]) for v in cells[c]])
        mnodes = np.unique([v for c in range(off[mm], off[mm + 1]) for v in cells[c]])
        pairs = []
        for n in snodes:  # master node at slave node - t
            d = np.abs(coord[mnodes] - (coord[n] - np.asarray(t))).max(axis=1)
            j = int(np.argmin(d))
            assert d[j] < 1e-9
            pairs.append((tags[n], tags[mnodes[j]]))
        affine = np.eye(4)
        affine[:3, 3] = t
        m.links.append((2, 100 + s, 100 + mm, affine.reshape(-1), np.array(pairs)))
    return m


def file_order(cells, num_inner_entities, off):
    """The cell order mesh_read gives a file written by from_cells: per entity, grouped by
    element type in first-appearance order."""
    order = []
    for e in range(len(off) - 1):
        dim = 3 if e < num_inner_entities else 2
        groups = {}
        for c in range(int(off[e]), int(off[e + 1])):
            groups.setdefault(ELEMENT_TYPE[(dim, len(cells[c]))], []).append(c)
        for g in groups.values():
            order.extend(g)
    return order


def from_mesh(a, node_tags=None, periodic_links=()):
    """from_cells of an ungenerated library mesh (the arrays of _abi.mesh_arrays before
    mesh_generate): the same node, cell and entity order when each entity has one element
    type."""
    noff, nidx = a["cell_node_off"], a["cell_node_idx"]
    cells = [nidx[noff[c]:noff[c + 1]] for c in range(a["num_cells"])]
    return from_cells(a["coord"], cells, a["entities_num_inner"], a["entity_names"],
                      a["entity_cell_off"], node_tags, periodic_links)
