"""teal_mesh_union (the batch's one mesh): group layout, index renumbering and copied geometry,
checked on the CPU against the instances it is built from -- a periodic box, a box with walls
and a jittered tetrahedral mesh.  Also the argument checks of teal_gpu_create_batch, which
fail before any device call."""
import ctypes as C

import numpy as np

import teal_amd
from teal_amd import _abi


def instances():
    lib = teal_amd.lib()
    tets = lib.mesh_create_tets(3, 0.2, 7)
    lib.mesh_generate(tets)
    return [teal_amd.mesh_box((4, 3, 5), periodic=(1, 1, 1)),
            teal_amd.mesh_box((3, 4, 2), periodic=(0, 1, 0)),
            tets]


def group_map(n_inner, off_boundary, n, base):
    """Union index of each instance index for one group layout [inner | boundary | rest]."""
    idx = np.arange(n)
    out = np.empty(n, dtype=np.int64)
    out[:n_inner] = base[0] + idx[:n_inner]
    out[n_inner:off_boundary] = base[1] + idx[n_inner:off_boundary] - n_inner
    out[off_boundary:] = base[2] + idx[off_boundary:] - off_boundary
    return out


def test_union_layout_and_geometry():
    lib = teal_amd.lib()
    ms = instances()
    arr = (C.c_void_p * len(ms))(*ms)
    u = lib.teal_mesh_union(arr, len(ms))
    ua = _abi.mesh_arrays(u)
    ia = [_abi.mesh_arrays(m) for m in ms]
    ni = [a["num_inner"] for a in ia]
    nb = [a["off_boundary"] - a["num_inner"] for a in ia]
    nc = [a["num_cells"] for a in ia]
    fi = [a["num_inner_faces"] for a in ia]
    fb = [a["off_boundary_faces"] - a["num_inner_faces"] for a in ia]
    nf = [a["num_faces"] for a in ia]
    assert ua["num_inner"] == sum(ni) and ua["num_cells"] == sum(nc)
    assert ua["off_boundary"] == sum(ni) + sum(nb)
    assert ua["num_faces"] == sum(nf) and ua["num_inner_faces"] == sum(fi)
    assert ua["off_boundary_faces"] == sum(fi) + sum(fb)
    cbase = [sum(ni), ua["off_boundary"]]
    fbase = [sum(fi), ua["off_boundary_faces"]]
    ci = [0, cbase[0], cbase[1]]
    fo = [0, fbase[0], fbase[1]]
    for a in ia:
        cmap = group_map(a["num_inner"], a["off_boundary"], a["num_cells"], ci)
        fmap = group_map(a["num_inner_faces"], a["off_boundary_faces"], a["num_faces"], fo)
        # geometry copied bit for bit
        np.testing.assert_array_equal(ua["volume"][cmap], a["volume"][:a["num_cells"]])
        np.testing.assert_array_equal(ua["center"][cmap], a["center"][:a["num_cells"]])
        for key in ("area", "face_center", "basis", "weight", "face_offset", "correction"):
            np.testing.assert_array_equal(ua[key][fmap], a[key])
        # face cells renumbered through the cell map
        np.testing.assert_array_equal(ua["cell_idx"][fmap], cmap[a["cell_idx"]])
        ci = [ci[0] + a["num_inner"], ci[1] + a["off_boundary"] - a["num_inner"],
              ci[2] + a["num_cells"] - a["off_boundary"]]
        fo = [fo[0] + a["num_inner_faces"], fo[1] + a["off_boundary_faces"] - a["num_inner_faces"],
              fo[2] + a["num_faces"] - a["off_boundary_faces"]]
    # every instance's entities, in the union's [inner | boundary | periodic] order
    names = sorted(n for a in ia for n in a["entity_names"])
    assert sorted(ua["entity_names"]) == names
    # periodic neighbour entries: each instance's, in order, tags offset per instance so that
    # no two instances pair up; send lists renumbered through the cell map
    want = np.concatenate([a["neighbors_tag"] + (i << 16) for i, a in enumerate(ia)])
    np.testing.assert_array_equal(ua["neighbors_tag"], want)
    ci = [0, cbase[0], cbase[1]]
    sends = []
    for a in ia:
        cmap = group_map(a["num_inner"], a["off_boundary"], a["num_cells"], ci)
        sends.append(cmap[a["neighbors_send_idx"]])
        ci = [ci[0] + a["num_inner"], ci[1] + a["off_boundary"] - a["num_inner"],
              ci[2] + a["num_cells"] - a["off_boundary"]]
    np.testing.assert_array_equal(ua["neighbors_send_idx"], np.concatenate(sends))
    np.testing.assert_allclose(ua["sum_volume"], sum(a["sum_volume"] for a in ia), rtol=1e-15)
    lib.mesh_destroy(u)


def test_create_batch_argument_errors():
    lib = teal_amd.lib()
    assert not lib.teal_gpu_create_batch(None, 0, teal_amd.FP64)
    assert "no meshes" in teal_amd.last_error()
    ms = instances()
    arr = (C.c_void_p * len(ms))(*ms)
    assert not lib.teal_gpu_create_batch(arr, len(ms), 7)
    assert "precision" in teal_amd.last_error()
