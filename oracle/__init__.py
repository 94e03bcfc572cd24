"""HODLR CPU oracle -- test infrastructure only.

ctypes bindings for ``oracle/build/liboracle.so`` (built from
``oracle/hodlr_oracle.cc`` by ``oracle/Makefile``).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` load
it, and only as the checker / the timed CPU reference.  See the header of
``hodlr_oracle.cc`` for the reference file:line each function restates.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_ptr = ctypes.c_void_p


def build():
    """Compile the oracle library (g++, no GPU needed)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_lapack_name.argtypes = [ctypes.c_char_p, _int]
        L.oracle_lapack_version.argtypes = [ctypes.c_char_p, _int]
        L.oracle_set_blas_threads.argtypes = [_int]
        L.oracle_set_blas_threads.restype = None
        L.oracle_hodlr_solve.argtypes = [_i64, _i64, _int, _int, _ptr, _ptr, _i64, _ptr, _i64, _ptr, _i64, _int,
                                         _int, _int, _int]
        L.oracle_hodlr_matvec.argtypes = [_i64, _i64, _int, _int, _ptr, _ptr, _i64, _ptr, _i64, _ptr, _i64, _int,
                                          _int, _ptr, _i64, _ptr, _i64]
        L.oracle_hodlr_dense.argtypes = [_i64, _i64, _int, _int, _ptr, _ptr, _i64, _ptr, _i64, _ptr, _i64, _int,
                                         _ptr]
        L.oracle_dense_dgesv.argtypes = [_int, _int, _ptr, _int, _ptr, _int]
        L.oracle_dense_solve_ld.argtypes = [_int, _int, _ptr, _int, _ptr, _int, _ptr, _int]
        L.oracle_woodbury_circulant.argtypes = [_i64, _int, _dbl, _int, _ptr, _i64, _ptr, _i64]
        L.oracle_drand48_rows.argtypes = [_i64, _i64, _int, _ptr, _i64]
        L.oracle_drand48_rows.restype = None
        L.oracle_glibc_drand48_rows.argtypes = [_i64, _i64, _int, _ptr, _i64]
        L.oracle_glibc_drand48_rows.restype = None
        L.oracle_rel_l2.argtypes = [_i64, _int, _ptr, _i64, _ptr, _i64]
        L.oracle_rel_l2.restype = _dbl
        _lib = L
    return _lib


def _f64(a):
    assert a.dtype == np.float64 and a.flags.f_contiguous, "column-major float64 arrays expected"
    return a.ctypes.data


def _ranks(ranks):
    r = np.ascontiguousarray(np.atleast_1d(ranks), dtype=np.int32)
    return r, len(r)


def lapack_name():
    buf = ctypes.create_string_buffer(512)
    if lib().oracle_lapack_name(buf, 512) != 0:
        raise RuntimeError("no LAPACK found for the oracle")
    return buf.value.decode()


def lapack_version():
    buf = ctypes.create_string_buffer(512)
    if lib().oracle_lapack_version(buf, 512) != 0:
        raise RuntimeError("no LAPACK found for the oracle")
    return buf.value.decode().strip()


def set_blas_threads(n):
    lib().oracle_set_blas_threads(int(n))


def hodlr_solve(sysd, X, V, K, order=0, node_form=0, nthreads=0):
    """Reference fused factor+solve, in place on X (and K, like the reference).

    ``sysd`` carries n, threshold, nrhs, ranks, klayout, ldk.  Returns the
    LAPACK-style status (0 = ok)."""
    r, nr = _ranks(sysd["ranks"])
    return lib().oracle_hodlr_solve(sysd["n"], sysd["threshold"], sysd["nrhs"], nr, r.ctypes.data,
                                    _f64(X), X.shape[0], _f64(V), V.shape[0], _f64(K),
                                    sysd.get("ldk", K.shape[0]), sysd.get("klayout", 0), order, node_form, nthreads)


def hodlr_matvec(sysd, X, V, K, x):
    r, nr = _ranks(sysd["ranks"])
    x = np.asfortranarray(x, dtype=np.float64)
    y = np.zeros_like(x, order="F")
    rc = lib().oracle_hodlr_matvec(sysd["n"], sysd["threshold"], sysd["nrhs"], nr, r.ctypes.data,
                                   _f64(X), X.shape[0], _f64(V), V.shape[0], _f64(K),
                                   sysd.get("ldk", K.shape[0]), sysd.get("klayout", 0),
                                   x.shape[1], _f64(x), x.shape[0], _f64(y), y.shape[0])
    if rc:
        raise RuntimeError(f"oracle_hodlr_matvec failed: {rc}")
    return y


def hodlr_dense(sysd, X, V, K):
    n = sysd["n"]
    r, nr = _ranks(sysd["ranks"])
    A = np.zeros((n, n), order="F")
    rc = lib().oracle_hodlr_dense(n, sysd["threshold"], sysd["nrhs"], nr, r.ctypes.data, _f64(X), X.shape[0],
                                  _f64(V), V.shape[0], _f64(K), sysd.get("ldk", K.shape[0]),
                                  sysd.get("klayout", 0), _f64(A))
    if rc:
        raise RuntimeError(f"oracle_hodlr_dense failed: {rc}")
    return A


def dense_dgesv(A, B):
    A = np.array(A, order="F", dtype=np.float64)
    B = np.array(B, order="F", dtype=np.float64)
    if B.ndim == 1:
        B = B.reshape(-1, 1, order="F")
    info = lib().oracle_dense_dgesv(A.shape[0], B.shape[1], _f64(A), A.shape[0], _f64(B), B.shape[0])
    return info, B


def dense_solve_ld(A, B):
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    X = np.zeros_like(B, order="F")
    info = lib().oracle_dense_solve_ld(A.shape[0], B.shape[1], _f64(A), A.shape[0], _f64(B), B.shape[0],
                                       _f64(X), X.shape[0])
    return info, X


def woodbury_circulant(n, r, diag, B):
    B = np.asfortranarray(B, dtype=np.float64)
    X = np.zeros_like(B, order="F")
    lib().oracle_woodbury_circulant(n, r, diag, B.shape[1], _f64(B), B.shape[0], _f64(X), X.shape[0])
    return X


def drand48_rows(seed, nrow, ncol, glibc=False):
    out = np.zeros((nrow, ncol), order="F")
    f = lib().oracle_glibc_drand48_rows if glibc else lib().oracle_drand48_rows
    f(seed, nrow, ncol, _f64(out), nrow)
    return out


def rel_l2(x, xref):
    x = np.asfortranarray(x, dtype=np.float64)
    xref = np.asfortranarray(xref, dtype=np.float64)
    if x.ndim == 1:
        x = x.reshape(-1, 1, order="F")
        xref = xref.reshape(-1, 1, order="F")
    return lib().oracle_rel_l2(x.shape[0], x.shape[1], _f64(x), x.shape[0], _f64(xref), xref.shape[0])
