"""The drop-in boundary on the CPU: the library loads, exports every function that
include/*.h declares, and fails loudly (an error, no fallback) where no HIP device exists."""
import ctypes as C
import glob
import os
import re

import pytest

import teal_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _clean(path):
    s = open(path).read()
    s = re.sub(r"//[^\n]*", "", s)
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    s = re.sub(r"#[^\n]*", "", s)
    s = re.sub(r'extern\s+"C"\s*\{', "", s)
    s = re.sub(r"__attribute__\s*\(\(.*?\)\)", "", s)
    while True:  # drop struct / enum bodies
        t = re.sub(r"\{[^{}]*\}", "", s)
        if t == s:
            return s
        s = t


def declared_functions():
    """(header, name) of every function prototype in include/*.h."""
    paths = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))
    text = {p: _clean(p) for p in paths}
    # function types (typedef double Advance(...)) declare functions as `Advance a, b;`
    ftypes = set()
    for s in text.values():
        ftypes |= set(re.findall(r"typedef\s+[\w\s\*]*?\b(\w+)\s*\(", s))
    out = []
    for path, s in text.items():
        for stmt in s.split(";"):
            st = " ".join(stmt.replace("}", " ").split())
            head = st.split(" ", 1)
            if len(head) == 2 and head[0] in ftypes:
                out += [(os.path.basename(path), n.strip()) for n in head[1].split(",")]
                continue
            if not st or st.startswith("typedef") or "(" not in st:
                continue
            m = re.match(r"^(?:const\s+)?[A-Za-z_]\w*(?:\s+|\s*\*+\s*)(?:\*\s*)*([A-Za-z_]\w*)\s*\(", st)
            if m:
                out.append((os.path.basename(path), m.group(1)))
    return out


DECLARED = declared_functions()


def test_headers_declare_the_interface():
    names = {n for _, n in DECLARED}
    headers = {h for h, _ in DECLARED}
    assert len(DECLARED) > 80
    assert {"teal.h", "mesh.h", "equations.h", "euler.h", "navierstokes.h", "simulation.h",
            "teal_gpu.h", "vector.h", "matrix.h", "utils.h"} <= headers
    for key in ("mesh_create", "mesh_generate", "equations_derivative", "lserk", "implicit_euler",
                "simulation_run", "teal_gpu_create", "teal_gpu_lserk", "euler_convective"):
        assert key in names, key


@pytest.mark.parametrize("header,name", DECLARED, ids=[f"{h}:{n}" for h, n in DECLARED])
def test_library_exports(header, name):
    lib = C.CDLL(teal_amd.LIB_PATH)
    assert hasattr(lib, name), f"{name} ({header}) is not exported by libteal_amd.so"


def gpu_present():
    return os.path.exists("/dev/kfd")


@pytest.mark.skipif(gpu_present(), reason="checks the no-device error path")
def test_no_device_is_an_error():
    lib = teal_amd.lib()
    m = teal_amd.mesh_box((4, 4, 4))
    assert not lib.teal_gpu_create(m, teal_amd.FP64)
    assert "no HIP device" in teal_amd.last_error()
    with pytest.raises(RuntimeError, match="teal_gpu_create failed"):
        teal_amd.Gpu(m)
    lib.mesh_destroy(m)


def test_missing_library_is_an_error(monkeypatch, tmp_path):
    monkeypatch.setattr(teal_amd, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(teal_amd, "_lib", None)
    with pytest.raises(RuntimeError, match="is not built"):
        teal_amd.lib()


def test_invalid_precision_is_an_error():
    lib = teal_amd.lib()
    m = teal_amd.mesh_box((3, 3, 3))
    assert not lib.teal_gpu_create(m, 7)
    assert "precision" in teal_amd.last_error()
    lib.mesh_destroy(m)
