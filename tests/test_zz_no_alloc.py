"""hodlr_dgesv_batched_noalloc: the batched dgesv without device allocations
(its workspace is the library's fixed scratch array) gives exactly the
results of hodlr_dgesv_batched, on two calls in a row, the second smaller;
inputs beyond the scratch size are refused before any GPU work."""
import ctypes

import numpy as np
import pytest

from fastsolver_amd import _lib, hodlr

pytestmark = pytest.mark.gpu

HODLR_ERR_SCRATCH = -8


def _arr(ctype, vals):
    return (ctype * len(vals))(*vals)


def _call(fn, As, Bs):
    n = len(As)
    info = np.full(n, -99, np.int32)
    rc = fn(n, _arr(ctypes.c_int32, [a.rows for a in As]), _arr(ctypes.c_int32, [b.cols for b in Bs]),
            _arr(ctypes.c_void_p, [a.ptr for a in As]), _arr(ctypes.c_int64, [a.ld for a in As]),
            _arr(ctypes.c_void_p, [b.ptr for b in Bs]), _arr(ctypes.c_int64, [b.ld for b in Bs]),
            info.ctypes.data, None)
    return rc, info


def _problems(rng, sizes):
    out = []
    for n, k in sizes:
        A = rng.standard_normal((n, n)) / np.sqrt(n) + 4.0 * np.eye(n)
        if n > 8:
            A[[1, n - 2]] = A[[n - 2, 1]]      # forces a row interchange
        B = rng.standard_normal((n, k))
        out.append((np.asfortranarray(A), np.asfortranarray(B)))
    return out


@pytest.mark.parametrize("sizes_first,sizes_second", [
    ([(1024, 96), (300, 5), (64, 1), (777, 40)], [(130, 3), (1000, 17)]),
    ([(512, 64)] * 6, [(33, 2)]),
])
def test_noalloc_matches_allocating_entry_point(sizes_first, sizes_second):
    lib = _lib.device_lib()
    rng = np.random.default_rng(11)
    for sizes in (sizes_first, sizes_second):          # two calls in a row, the second smaller
        print(f"noalloc dgesv batch {sizes}", flush=True)
        probs = _problems(rng, sizes)
        A1 = [hodlr.DeviceArray.from_host(a) for a, _ in probs]
        B1 = [hodlr.DeviceArray.from_host(b) for _, b in probs]
        A2 = [hodlr.DeviceArray.from_host(a) for a, _ in probs]
        B2 = [hodlr.DeviceArray.from_host(b) for _, b in probs]
        rc1, info1 = _call(lib.hodlr_dgesv_batched, A1, B1)
        rc2, info2 = _call(lib.hodlr_dgesv_batched_noalloc, A2, B2)
        assert rc1 == 0 and rc2 == 0, (rc1, rc2)
        assert np.array_equal(info1, info2) and np.all(info2 == 0)
        for (a, b), a1, b1, a2, b2 in zip(probs, A1, B1, A2, B2):
            assert np.array_equal(a1.download(), a2.download())    # LU factors
            x = b2.download()
            assert np.array_equal(b1.download(), x)                # solutions
            assert np.linalg.norm(a @ x - b) / np.linalg.norm(b) <= 1e-12


def test_noalloc_rejects_inputs_beyond_the_scratch():
    lib = _lib.device_lib()
    # 64 systems of 1024 x 1024 with 128 right-hand sides need 128 MiB of workspace
    n = 64
    ns = _arr(ctypes.c_int32, [1024] * n)
    ks = _arr(ctypes.c_int32, [128] * n)
    ptrs = _arr(ctypes.c_void_p, [8] * n)          # never dereferenced: refused on the host
    lds = _arr(ctypes.c_int64, [1024] * n)
    info = np.zeros(n, np.int32)
    rc = lib.hodlr_dgesv_batched_noalloc(n, ns, ks, ptrs, lds, ptrs, lds, info.ctypes.data, None)
    assert rc == HODLR_ERR_SCRATCH
    assert b"scratch" in lib.hodlr_error_string(rc)
