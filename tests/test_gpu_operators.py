"""Standalone batched operators (the task bodies' dgemm_/dgesv_) against LAPACK
through the oracle's library and numpy, on the GPU."""
import numpy as np
import pytest

import oracle
from fastsolver_amd import hodlr

pytestmark = pytest.mark.gpu


def test_dgemm_batched_nn_and_tn():
    rng = np.random.default_rng(11)
    ops, ref = [], []
    for (tr, m, n, k) in [("n", 70, 33, 129), ("t", 30, 62, 60), ("n", 1, 5, 3), ("t", 128, 160, 4099)]:
        A = np.asfortranarray(rng.standard_normal((m, k) if tr == "n" else (k, m)))
        B = np.asfortranarray(rng.standard_normal((k, n)))
        C = np.asfortranarray(rng.standard_normal((m, n)))
        alpha, beta = -1.0, (1.0 if tr == "n" else 0.0)
        ops.append((tr, alpha, hodlr.DeviceArray.from_host(A), hodlr.DeviceArray.from_host(B), beta,
                    hodlr.DeviceArray.from_host(C)))
        ref.append(alpha * ((A if tr == "n" else A.T) @ B) + beta * C)
    hodlr.gemm_batched(ops)
    for o, r in zip(ops, ref):
        got = o[5].download()
        assert np.abs(got - r).max() <= 1e-12 * max(1.0, np.abs(r).max()), o[0]


def test_sgemm_batched():
    rng = np.random.default_rng(12)
    A = np.asfortranarray(rng.standard_normal((300, 200)).astype(np.float32))
    B = np.asfortranarray(rng.standard_normal((200, 100)).astype(np.float32))
    C = np.zeros((300, 100), np.float32, order="F")
    dC = hodlr.DeviceArray.from_host(C)
    hodlr.gemm_batched([("n", 1.0, hodlr.DeviceArray.from_host(A), hodlr.DeviceArray.from_host(B), 0.0, dC)],
                       dtype=np.float32)
    r = A.astype(np.float64) @ B.astype(np.float64)
    assert np.abs(dC.download() - r).max() <= 1e-4 * np.abs(r).max()


@pytest.mark.parametrize("n,nrhs", [(60, 92), (256, 181), (1000, 7), (1024, 300), (33, 1)])
def test_dgesv_batched_vs_lapack(n, nrhs):
    rng = np.random.default_rng(n)
    A = np.asfortranarray(rng.standard_normal((n, n)))      # general matrix: pivoting happens
    B = np.asfortranarray(rng.standard_normal((n, nrhs)))
    info_ref, X_ref = oracle.dense_dgesv(A, B)
    dA, dB = hodlr.DeviceArray.from_host(A), hodlr.DeviceArray.from_host(B)
    info = hodlr.dgesv_batched([dA], [dB])
    assert info[0] == info_ref == 0
    X = dB.download()
    # same LAPACK semantics: residual at rounding level and the same solution
    assert np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X)) < 1e-14
    assert np.linalg.norm(X - X_ref) / np.linalg.norm(X_ref) < 1e-9 * np.linalg.cond(A) / 1e3


def test_dgesv_batched_singular_info():
    A = np.asfortranarray(np.ones((8, 8)))
    dA, dB = hodlr.DeviceArray.from_host(A), hodlr.DeviceArray.from_host(np.ones((8, 1)))
    info = hodlr.dgesv_batched([dA], [dB])
    assert info[0] == 2   # LAPACK: U(2,2) is exactly zero


def test_integer_known_answers_fp64_and_fp32():
    """Raw circulant factors (integer entries) through the reduce (t,n) and
    broadcast (n,n; alpha -1, beta 1) products: every partial sum is an
    integer below 2^24, so FP64 and FP32 results must equal the fixture bitwise."""
    import os
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "circulant_int_gemm.npz"))
    A, B, W, eta, C, D = (np.asfortranarray(d[k].astype(np.float64)) for k in ("A", "B", "W", "eta", "C", "D"))
    dW = hodlr.DeviceArray(W.shape[0], W.shape[1])
    dC = hodlr.DeviceArray.from_host(C)
    hodlr.gemm_batched([("t", 1.0, hodlr.DeviceArray.from_host(A), hodlr.DeviceArray.from_host(B), 0.0, dW),
                        ("n", -1.0, hodlr.DeviceArray.from_host(A), hodlr.DeviceArray.from_host(eta), 1.0, dC)])
    assert np.array_equal(dW.download(), W)
    assert np.array_equal(dC.download(), D)
    f = lambda x: np.asfortranarray(x.astype(np.float32))
    sC = hodlr.DeviceArray.from_host(f(C))
    hodlr.gemm_batched([("n", -1.0, hodlr.DeviceArray.from_host(f(A)), hodlr.DeviceArray.from_host(f(eta)), 1.0, sC)],
                       dtype=np.float32)
    assert np.array_equal(sC.download().astype(np.float64), D)


def _round_bf16(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(np.shape(x))


def test_bf16_gemm_integer_known_answer_and_rounding():
    """BF16 MFMA products: exact on the integer circulant fixture (entries < 256
    are exact in BF16), and on random inputs equal to the product of the
    BF16-rounded operands up to FP32 accumulation."""
    import os
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "circulant_int_gemm.npz"))
    f = lambda x: np.asfortranarray(np.asarray(x, dtype=np.float32))
    A, eta, C, D = (f(d[k]) for k in ("A", "eta", "C", "D"))
    dC = hodlr.DeviceArray.from_host(C)
    hodlr.gemm_batched([("n", -1.0, hodlr.DeviceArray.from_host(A), hodlr.DeviceArray.from_host(eta), 1.0, dC)],
                       dtype=np.float32, bf16=True)
    assert np.array_equal(dC.download(), D)
    rng = np.random.default_rng(5)
    for (m, n, k) in [(300, 100, 200), (64, 32, 64), (1000, 77, 129)]:
        A = f(rng.standard_normal((m, k)))
        B = f(rng.standard_normal((k, n)))
        dC = hodlr.DeviceArray(m, n, np.float32)
        hodlr.gemm_batched([("n", 1.0, hodlr.DeviceArray.from_host(A), hodlr.DeviceArray.from_host(B), 0.0, dC)],
                           dtype=np.float32, bf16=True)
        ref = _round_bf16(A).astype(np.float64) @ _round_bf16(B).astype(np.float64)
        scale = np.abs(_round_bf16(A)).astype(np.float64) @ np.abs(_round_bf16(B)).astype(np.float64)
        assert np.all(np.abs(dC.download() - ref) <= 1e-6 * scale + 1e-30), (m, n, k)
        # and it is really BF16: the unrounded product differs by far more
        assert np.abs(dC.download() - A.astype(np.float64) @ B.astype(np.float64)).max() > 1e-4


@pytest.mark.parametrize("n", [5, 17, 64, 100, 128, 130, 200, 513, 777, 1024])
def test_dgesv_batched_lu_factors_and_pivots_match_lapack(n):
    """The in-place LU factors equal LAPACK getrf's: the pivot search picks the
    same rows (first largest |a|, dgetf2's idamax), so L and U agree to rounding
    (a different pivot anywhere would change them completely)."""
    import scipy.linalg
    rng = np.random.default_rng(100 + n)
    A = np.asfortranarray(rng.standard_normal((n, n)))
    lu, piv = scipy.linalg.lu_factor(A)
    dA, dB = hodlr.DeviceArray.from_host(A), hodlr.DeviceArray.from_host(np.ones((n, 1)))
    info = hodlr.dgesv_batched([dA], [dB])
    assert info[0] == 0
    got = dA.download()
    assert np.abs(got - lu).max() <= 1e-10 * np.abs(lu).max()


@pytest.mark.parametrize("n,zero_col", [(100, 70), (128, 0), (72, 71), (64, 10), (40, 39), (9, 0)])
def test_dgesv_batched_zero_pivot_matches_lapack_getrf(n, zero_col):
    """Blocks of at most 128 rows take the register-resident LU (one wave per
    block up to 64 rows, 256 threads above): a zero column gives an exactly
    zero pivot there; info and the factors (the column left unscaled, the
    elimination continued) equal LAPACK getrf's."""
    import scipy.linalg.lapack as lapack
    rng = np.random.default_rng(300 + n)
    A = np.asfortranarray(rng.standard_normal((n, n)))
    A[:, zero_col] = 0.0
    lu, piv, info_ref = lapack.dgetrf(A)
    assert info_ref == zero_col + 1
    dA, dB = hodlr.DeviceArray.from_host(A), hodlr.DeviceArray.from_host(np.ones((n, 1)))
    info = hodlr.dgesv_batched([dA], [dB])
    assert info[0] == info_ref
    assert np.abs(dA.download() - lu).max() <= 1e-10 * np.abs(lu).max()


def test_dgesv_batched_uneven_sizes_in_one_call():
    """One batched call over blocks of different orders (the recursive LU runs
    each step over the largest block; smaller ones drop out of the late steps):
    every block's factors and solution match LAPACK's."""
    import scipy.linalg
    sizes = [33, 130, 513, 777, 1024, 64]
    rng = np.random.default_rng(7)
    As = [np.asfortranarray(rng.standard_normal((n, n))) for n in sizes]
    Bs = [np.asfortranarray(rng.standard_normal((n, 3))) for n in sizes]
    dA = [hodlr.DeviceArray.from_host(A) for A in As]
    dB = [hodlr.DeviceArray.from_host(B) for B in Bs]
    info = hodlr.dgesv_batched(dA, dB)
    assert list(info) == [0] * len(sizes)
    for n, A, B, a, b in zip(sizes, As, Bs, dA, dB):
        lu, _ = scipy.linalg.lu_factor(A)
        assert np.abs(a.download() - lu).max() <= 1e-10 * np.abs(lu).max(), n
        X = b.download()
        assert np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X)) < 1e-14, n


def test_dgesv_batched_many_small_blocks():
    """A batch of 4096 blocks of 16..256 rows (the shape of many-system plans
    with small leaves) takes the one-wave-per-block panel path: every block
    solves to rounding level and sampled factors match LAPACK's."""
    import scipy.linalg
    rng = np.random.default_rng(11)
    sizes = [int(x) for x in rng.integers(16, 257, 4096)]
    As = [np.asfortranarray(rng.standard_normal((n, n))) for n in sizes]
    Bs = [np.asfortranarray(rng.standard_normal((n, 1))) for n in sizes]
    dA = [hodlr.DeviceArray.from_host(A) for A in As]
    dB = [hodlr.DeviceArray.from_host(B) for B in Bs]
    info = hodlr.dgesv_batched(dA, dB)
    assert not np.any(info)
    worst = 0.0
    for A, B, b in zip(As, Bs, dB):
        X = b.download()
        worst = max(worst, np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X)))
    assert worst < 1e-14
    for i in range(0, 4096, 409):
        lu, _ = scipy.linalg.lu_factor(As[i])
        assert np.abs(dA[i].download() - lu).max() <= 1e-10 * np.abs(lu).max(), sizes[i]
