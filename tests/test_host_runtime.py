"""The host runtime of the library (teal_amd/csrc/core.c) on the CPU: allocation, the array
helpers of utils.h, the closeness tests and the command-line options of teal_init.

Where the reference build exists (oracle/_ref, compiled from the reference's own
src/core/utils.c), the helpers are compared with it on the same random inputs; otherwise with
numpy's stable sort and searchsorted, which state the same contracts."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

import teal_amd
from oracle_lib import REF_SO, have_ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CMP = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)


def lib():
    l = teal_amd.lib()
    l.teal_calloc.restype = C.c_void_p
    l.teal_calloc.argtypes = [C.c_int, C.c_size_t]
    l.teal_realloc.restype = C.c_void_p
    l.teal_realloc.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    l.teal_free.argtypes = [C.c_void_p]
    return l


def helpers(l):
    for name in ("sort",):
        getattr(l, name).argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
    l.unique.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
    l.unique.restype = C.c_int
    l.search.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
    l.search.restype = C.c_void_p
    l.digitize.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
    l.digitize.restype = C.c_int
    for name in ("is_close", "is_less", "is_greater", "is_less_equal", "is_greater_equal"):
        getattr(l, name).argtypes = [C.c_double, C.c_double]
        getattr(l, name).restype = C.c_int
    return l


def cmp_ptr(l, name):
    return C.cast(getattr(l, name), C.c_void_p)


LIBS = [("library", lambda: helpers(lib()))]
if have_ref():
    LIBS.append(("reference", lambda: helpers(C.CDLL(REF_SO, mode=C.RTLD_GLOBAL))))


def test_allocation_is_aligned_zeroed_and_preserved_by_realloc():
    l = lib()
    assert not l.teal_calloc(0, 8)
    p = l.teal_calloc(1000, 8)
    assert p % 64 == 0
    a = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(1000,))
    assert not a.any()
    a[:] = np.arange(1000)
    q = l.teal_realloc(p, 3000, 8)  # grow: the old values are kept
    assert q % 64 == 0
    b = np.ctypeslib.as_array(C.cast(q, C.POINTER(C.c_double)), shape=(3000,))
    np.testing.assert_array_equal(b[:1000], np.arange(1000))
    r = l.teal_realloc(q, 10, 8)  # shrink: the first values are kept
    c = np.ctypeslib.as_array(C.cast(r, C.POINTER(C.c_double)), shape=(10,))
    np.testing.assert_array_equal(c, np.arange(10))
    assert not l.teal_realloc(r, 0, 8)  # frees
    s = l.teal_realloc(None, 5, 24)  # from nothing
    assert s and s % 64 == 0
    l.teal_free(s)
    l.teal_free(None)


@pytest.mark.parametrize("which,make", LIBS, ids=[n for n, _ in LIBS])
def test_sort_is_stable(which, make):
    l = make()
    rng = np.random.default_rng(5)
    for n in (0, 1, 2, 3, 17, 1000, 4099):
        rec = np.zeros((n, 2), dtype=np.int32)
        rec[:, 0] = rng.integers(0, 9, n)  # many equal keys
        rec[:, 1] = np.arange(n)
        want = rec[np.argsort(rec[:, 0], kind="stable")]
        l.sort(rec.ctypes.data, n, 8, cmp_ptr(l, "compare_int"))  # compares the key only
        np.testing.assert_array_equal(rec, want)


@pytest.mark.parametrize("which,make", LIBS, ids=[n for n, _ in LIBS])
def test_unique_search_digitize(which, make):
    l = make()
    rng = np.random.default_rng(6)
    for n in (1, 2, 50, 777):
        x = rng.integers(-40, 40, n).astype(np.int64)
        k = l.unique(x.ctypes.data, n, 8, cmp_ptr(l, "compare_long"))
        np.testing.assert_array_equal(x[:k], np.unique(x))
        u = x[:k].copy()
        for key in range(-45, 45):
            kk = C.c_long(key)
            hit = l.search(C.byref(kk), u.ctypes.data, k, 8, cmp_ptr(l, "compare_long"))
            if key in u:
                assert hit and C.cast(hit, C.POINTER(C.c_long))[0] == key
            else:
                assert not hit
        edges = np.sort(rng.uniform(-1, 1, n + 1))
        for key in np.concatenate([rng.uniform(-1.2, 1.2, 200), edges]):
            kd = C.c_double(key)
            got = l.digitize(C.byref(kd), edges.ctypes.data, n, 8, cmp_ptr(l, "compare_double"))
            want = int(np.searchsorted(edges, key, side="right")) - 1
            assert got == min(want, n), (key, got, want)
    assert l.digitize(C.byref(C.c_double(0)), None, 0, 8, cmp_ptr(l, "compare_double")) == -2


def test_closeness_matches_the_reference():
    mine = helpers(lib())
    others = [helpers(C.CDLL(REF_SO, mode=C.RTLD_GLOBAL))] if have_ref() else []
    rng = np.random.default_rng(7)
    base = rng.standard_normal(400) * 10.0 ** rng.integers(-10, 4, 400)
    pairs = [(a, a * (1 + d)) for a, d in zip(base, rng.choice([0, 1e-9, 1e-6, 1e-5, 2e-5, 1e-3], 400))]
    pairs += [(0.0, 1e-8), (0.0, 1.1e-8), (1e-9, -1e-9), (np.nan, 1.0), (np.inf, np.inf),
              (1.0, np.nan), (-0.0, 0.0)]
    for a, b in pairs:
        want = bool(abs(a - b) <= max(1e-8, 1e-5 * max(abs(a), abs(b))))
        assert bool(mine.is_close(a, b)) == want, (a, b)
        for name in ("is_close", "is_less", "is_greater", "is_less_equal", "is_greater_equal"):
            for o in others:
                assert getattr(mine, name)(a, b) == getattr(o, name)(a, b), (name, a, b)


OPTIONS_SCRIPT = r"""
import ctypes as C, sys
l = C.CDLL(sys.argv[1])
args = sys.argv[2:]
argv = (C.c_char_p * (len(args) + 1))(*[a.encode() for a in args], None)
argc = C.c_int(len(args))
pargv = C.pointer(argv)
l.teal_init(C.byref(argc), C.byref(pargv))
class Teal(C.Structure):
    _fields_ = [("quiet", C.c_int), ("verbose", C.c_int), ("partitioned", C.c_int),
                ("restart", C.c_char_p)]
t = Teal.in_dll(l, "teal")
rest = [pargv[0][i].decode() for i in range(argc.value)]
assert pargv[0][argc.value] is None
print("RESULT", t.quiet, t.verbose, t.partitioned, t.restart, rest, flush=True)
"""


def run_options(*args):
    r = subprocess.run([sys.executable, "-c", OPTIONS_SCRIPT, teal_amd.LIB_PATH, *args],
                       capture_output=True, text=True, timeout=60)
    return r


@pytest.mark.parametrize("args,want", [
    (["prog", "a", "-q", "b", "-r", "state.h5", "c", "-v"], "1 1 0 b'state.h5' ['prog', 'a', 'b', 'c']"),
    (["prog", "-qvp", "-rout.h5", "x"], "1 1 1 b'out.h5' ['prog', 'x']"),
    (["prog", "-q", "--", "-v", "y"], "1 0 0 None ['prog', '-v', 'y']"),
    (["prog", "-", "-q"], "1 0 0 None ['prog', '-']"),
])
def test_teal_init_options(args, want):
    r = run_options(*args)
    assert r.returncode == 0, r.stderr
    assert f"RESULT {want}" in r.stdout
    if "-q" not in args and "-qvp" not in args:
        assert "Hello, World! This is teal!" in r.stdout


def test_teal_init_help_and_unknown_option():
    r = run_options("prog", "-h")
    assert r.returncode == 0 and "usage: prog [options] ..." in r.stdout
    assert "-r <file>" in r.stdout and "RESULT" not in r.stdout
    r = run_options("prog", "-x")
    assert r.returncode != 0 and "usage: prog" in r.stdout
    r = run_options("prog", "-r")  # no file
    assert r.returncode != 0 and "usage: prog" in r.stdout
