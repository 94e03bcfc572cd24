"""The C-ABI libraries load and export every symbol their headers declare
(CPU only: no compute call needs a GPU here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from fastsolver_amd import _lib, gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hodlr_[a-z0-9_]+)\s*\(", text)))


def test_device_library_exports_header():
    names = _declared("fastsolver_amd.h")
    assert len(names) >= 30
    lib = ctypes.CDLL(_lib.DEVICE_LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding table covers the header exactly
    assert sorted(_lib.DEVICE_SYMBOLS) == names


def test_gen_library_exports_header():
    names = _declared("fastsolver_amd_gen.h")
    lib = ctypes.CDLL(_lib.GEN_LIB)
    assert all(hasattr(lib, n) for n in names)
    assert sorted(_lib.GEN_SYMBOLS) == names


def test_error_strings_without_gpu():
    lib = _lib.device_lib()
    assert lib.hodlr_error_string(0) == b"ok"
    assert b"zero pivot" in lib.hodlr_error_string(-6)
    assert lib.hodlr_version().startswith(b"fastsolver_amd")
    assert lib.hodlr_profile_category(5) == b"leaf_trsm"


def test_plan_create_rejects_bad_arguments_before_touching_the_gpu():
    lib = _lib.device_lib()
    h = ctypes.c_void_p()
    assert lib.hodlr_plan_create(ctypes.byref(h), None, 0, 0, None) == -1
    ranks = np.array([30], np.int32)
    s = _lib.System(100, 10, 1, 1, ranks.ctypes.data, 0, 100, 0, 100, 0, 0, 0, 0, 0, 100)
    assert lib.hodlr_plan_create(ctypes.byref(h), ctypes.byref(s), 1, 0, None) == -1   # null arrays
    assert lib.hodlr_plan_create(ctypes.byref(h), ctypes.byref(s), 1, 7, None) == -1   # bad precision


@pytest.mark.parametrize("cfg,leaves,depth,xcols", [(1, 64, 6, 1 + 6 * 30), (2, 256, 8, 4 + 8 * 64),
                                                    (3, 256, 8, 8 + 8 * 100), (5, 1024, 10, 32 + 10 * 128)])
def test_workload_trees(cfg, leaves, depth, xcols):
    spec = gen.config(cfg)[0]
    nl, md, xc, vc, kp = gen.tree_info(spec)
    assert (nl, md, xc) == (leaves, depth, xcols)
    assert vc == xc - spec.nrhs
    assert kp == spec.n * (spec.n // leaves)


def test_uneven_batch_specs():
    specs = gen.config(4)
    assert len(specs) == 512
    for s in specs[:64]:
        assert s.n % 256 == 0 and 4096 <= s.n <= 65536
        assert 1 <= s.nrhs <= 16 and all(8 <= r <= 64 for r in s.ranks)
        nl, md, *_ = gen.tree_info(s)
        assert md == len(s.ranks) and s.n // nl <= 256


def test_tree_errors():
    with pytest.raises(ValueError):
        gen.tree_info(gen.SystemSpec(200, 50, 1, [60]))   # leaf rows (50) <= rank


def test_cxx_facade_compiles_and_example_links():
    """The C++ facade (the reference's HodlrMatrix / FastSolver driver API)
    compiles against both headers, and the in-tree example resolves both
    libraries without LD_LIBRARY_PATH."""
    import subprocess
    src = "#include \"fastsolver_amd.hpp\"\nint main(){fastsolver_amd::HodlrMatrix m(2,480,3,3,30,60,1,\"g\");" \
          "return m.get_num_rhs()==2 ? 0 : 1;}\n"
    p = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        "-x", "c++", "-"], input=src, text=True, capture_output=True)
    assert p.returncode == 0, p.stderr
    exe = os.path.join(ROOT, "fastsolver_amd", "examples", "single_launch")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    ldd = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libfastsolver_amd.so" in ldd and "libfastsolver_amd_gen.so" in ldd
    assert "not found" not in ldd, ldd


def _sys(n, threshold, nrhs, ranks):
    r = np.array(ranks, np.int32)
    # dummy non-null device pointers: validation fails before any is touched
    return _lib.System(n, threshold, nrhs, len(r), r.ctypes.data, 8, n, 8, n, 8, 0, 0, 0, 8, n), r


def test_plan_create_rejects_unsupported_trees_before_touching_the_gpu():
    lib = _lib.device_lib()
    h = ctypes.c_void_p()
    s, keep = _sys(4096, 2048, 1, [30])          # leaves of 2048 rows: above the 1024 the kernels hold
    assert lib.hodlr_plan_create(ctypes.byref(h), ctypes.byref(s), 1, 0, None) == -3
    s, keep = _sys(256, 64, 1, [64])             # leaf rows <= rank (hodlr_matrix.cc:167 asserts N > rank)
    assert lib.hodlr_plan_create(ctypes.byref(h), ctypes.byref(s), 1, 0, None) == -2
    s, keep = _sys(1 << 12, 256, 2, [20, 20])    # per-depth ranks must cover every level
    assert lib.hodlr_plan_create(ctypes.byref(h), ctypes.byref(s), 1, 0, None) == -2


def test_plan_create_rejects_empty_and_degenerate_systems():
    lib = _lib.device_lib()
    h = ctypes.c_void_p()
    for (n, thr, c) in [(0, 64, 1), (256, 0, 1), (256, 64, 0)]:
        s, keep = _sys(n, thr, c, [8])
        assert lib.hodlr_plan_create(ctypes.byref(h), ctypes.byref(s), 1, 0, None) == -1, (n, thr, c)
