"""ctypes mirror of the public C structs in include/mesh.h (same layout as the reference's
src/mesh.h:13-72) plus numpy views over their arrays.

Used by the Python host helpers, the tests and bench.py to read meshes built by the C
library without copying."""
import ctypes as C

import numpy as np

c_int_p = C.POINTER(C.c_int)
c_long_p = C.POINTER(C.c_long)
c_double_p = C.POINTER(C.c_double)


class Graph(C.Structure):
    _fields_ = [("off", c_int_p), ("idx", c_int_p)]


class MeshNodes(C.Structure):
    _fields_ = [("num", C.c_int), ("num_inner", C.c_int), ("global_", c_long_p),
                ("coord", c_double_p)]


class MeshCells(C.Structure):
    _fields_ = [("num", C.c_int), ("num_inner", C.c_int), ("off_boundary", C.c_int),
                ("off_periodic", C.c_int), ("node", Graph), ("cell", Graph),
                ("volume", c_double_p), ("sum_volume", C.c_double), ("center", c_double_p),
                ("projection", c_double_p), ("offset", c_double_p)]


class MeshFaces(C.Structure):
    _fields_ = [("num", C.c_int), ("num_inner", C.c_int), ("off_boundary", C.c_int),
                ("node", Graph), ("cell_idx", c_int_p), ("area", c_double_p),
                ("center", c_double_p), ("basis", c_double_p), ("weight", c_double_p),
                ("offset", c_double_p), ("correction", c_double_p)]


class MeshEntities(C.Structure):
    _fields_ = [("num", C.c_int), ("num_inner", C.c_int), ("off_boundary", C.c_int),
                ("name", C.POINTER(C.c_char * 128)), ("cell_off", c_int_p),
                ("face_off", c_int_p), ("rotation", c_double_p), ("translation", c_double_p)]


class MeshNeighbors(C.Structure):
    _fields_ = [("num", C.c_int), ("tag", c_int_p), ("rank", c_int_p), ("recv_off", c_int_p),
                ("send", Graph)]


class Mesh(C.Structure):
    _fields_ = [("nodes", MeshNodes), ("cells", MeshCells), ("faces", MeshFaces),
                ("entities", MeshEntities), ("neighbors", MeshNeighbors)]


def view(ptr, *shape):
    """numpy view of `prod(shape)` elements behind a ctypes pointer (no copy)."""
    n = int(np.prod(shape))
    if n == 0 or not ptr:
        base = {c_int_p: np.int32, c_long_p: np.int64}.get(type(ptr), np.float64)
        return np.zeros(shape, dtype=base)
    return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(shape)


def mesh_arrays(mesh_ptr):
    """Dictionary of numpy views (no copies) of every array of a generated mesh."""
    m = C.cast(mesh_ptr, C.POINTER(Mesh)).contents
    nc, ni = m.cells.num, m.cells.num_inner
    nf = m.faces.num
    ne = m.entities.num
    nn = m.nodes.num
    cell_off = view(m.cells.cell.off, nc + 1) if m.cells.cell.off else None
    out = {
        "num_nodes": nn, "num_cells": nc, "num_inner": ni,
        "off_boundary": m.cells.off_boundary, "off_periodic": m.cells.off_periodic,
        "num_faces": nf, "num_inner_faces": m.faces.num_inner,
        "off_boundary_faces": m.faces.off_boundary,
        "coord": view(m.nodes.coord, nn, 3),
        "cell_node_off": view(m.cells.node.off, nc + 1),
        "sum_volume": m.cells.sum_volume,
        "entity_names": [m.entities.name[i].value.decode() for i in range(ne)],
        "entity_cell_off": view(m.entities.cell_off, ne + 1),
        "entities_num_inner": m.entities.num_inner,
        "entities_off_boundary": m.entities.off_boundary,
        "rotation": view(m.entities.rotation, ne, 9),
        "translation": view(m.entities.translation, ne, 3),
        "neighbors_num": m.neighbors.num,
    }
    out["cell_node_idx"] = view(m.cells.node.idx, int(out["cell_node_off"][-1]))
    nb = m.neighbors.num
    out["neighbors_tag"] = view(m.neighbors.tag, nb, 2)
    out["neighbors_recv_off"] = view(m.neighbors.recv_off, nb + 1)
    out["neighbors_send_off"] = view(m.neighbors.send.off, nb + 1)
    out["neighbors_send_idx"] = view(m.neighbors.send.idx, int(out["neighbors_send_off"][-1]))
    if cell_off is not None:
        out["cell_cell_off"] = cell_off
        out["cell_cell_idx"] = view(m.cells.cell.idx, int(cell_off[-1]))
    if nf > 0 and m.faces.area:
        out.update({
            "face_node_off": view(m.faces.node.off, nf + 1),
            "cell_idx": view(m.faces.cell_idx, nf, 2),
            "area": view(m.faces.area, nf),
            "face_center": view(m.faces.center, nf, 3),
            "basis": view(m.faces.basis, nf, 9),
            "weight": view(m.faces.weight, nf, 6),
            "face_offset": view(m.faces.offset, nf, 6),
            "correction": view(m.faces.correction, nf, 4),
            "volume": view(m.cells.volume, nc),
            "center": view(m.cells.center, nc, 3),
            "projection": view(m.cells.projection, nc, 3),
            "cell_offset": view(m.cells.offset, int(cell_off[ni])),
            "entity_face_off": view(m.entities.face_off, ne + 1),
        })
        out["face_node_idx"] = view(m.faces.node.idx, int(out["face_node_off"][-1]))
        out["cell_offset"] = view(m.cells.offset, int(cell_off[ni]), 3)
    return out
