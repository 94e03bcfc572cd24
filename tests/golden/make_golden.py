"""Writes the committed golden fixtures of tests/golden/ (run from the repo root:
``python tests/golden/make_golden.py``).

The reference holds no data files or asserted values, so every fixture here
comes from a method independent of the HODLR algorithm under test:

* ``single_launch_circulant.npz``, ``accuracy_circulant.npz`` -- the
  reference's own test matrices (test/single_launch.cc:21-38,
  test/main.cc:174-185): A = diag*I + U U^T, U(i,j) = (i+j) mod r, with the
  drand48 right-hand side.  x_ref from the Woodbury closed form in long double.
* ``cfg1_n2048.npz``, ``cfg1_n16384.npz`` -- the random leaf family of the
  N=16384 / rank 30 workload (and its N=2048 cut): x_ref from LAPACK dgesv on
  the assembled dense matrix, the check direct_solve.cc:188-225 performs.
* ``circulant_int_gemm.npz`` -- exact-integer known answers for the reduce
  (dgemm_('t','n'), src/solver/gemm.cc:376) and broadcast (dgemm_('n','n'),
  gemm.cc:473) products on the raw, pre-solve circulant factors of the
  single_launch case (U(i,j) = (i+j) mod 30, leaf of 60 rows, 2 + 2*30
  columns).  Every product and partial sum is an integer below 2^24, so the
  answer is exact in FP64 and FP32 arithmetic alike (int64 reference).
* every fixture stores exact checksums of the generated inputs, pinning the
  generator (Philox4x32-10, Box-Muller, drand48) across machines.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from fastsolver_amd import gen  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def checksums(X, V, K):
    return np.array([X.sum(), np.abs(X).sum(), V.sum(), np.abs(V).sum(), K.sum(), np.abs(K).sum()])


def spec_array(spec):
    return np.array([spec.n, spec.threshold, spec.nrhs] + list(spec.ranks), dtype=np.int64)


def circulant(name, spec):
    X, V, K = gen.generate(spec)
    b = X[:, :spec.nrhs].copy(order="F")
    x = oracle.woodbury_circulant(spec.n, spec.ranks[0], spec.diag, b)
    np.savez(os.path.join(OUT, name), x_ref=x, b=b, checksums=checksums(X, V, K), spec=spec_array(spec),
             diag=np.array([spec.diag]), seed=np.array([spec.seed]))


def dense(name, spec):
    X, V, K = gen.generate(spec)
    b = X[:, :spec.nrhs].copy(order="F")
    A = oracle.hodlr_dense(spec.as_dict(), X, V, K)
    info, x = oracle.dense_dgesv(A, b)
    assert info == 0
    np.savez(os.path.join(OUT, name), x_ref=x, checksums=checksums(X, V, K), spec=spec_array(spec),
             seed=np.array([spec.seed]))


def integer_gemm(name):
    r, m, ncol = 30, 60, 2 + 2 * 30
    rows = np.arange(180, 180 + m)                     # leaf 3 of the N = 480 tree
    A = ((rows[:, None] + np.arange(r)[None, :]) % r).astype(np.int64)        # V_leaf (= U_leaf), 60 x 30
    B = ((rows[:, None] + np.arange(ncol)[None, :]) % r).astype(np.int64)     # X_leaf columns, 60 x 62
    W = A.T @ B                                                                 # reduce: 30 x 62
    eta = ((np.arange(r)[:, None] * np.arange(ncol)[None, :]) % 7).astype(np.int64)   # 30 x 62
    C = ((3 * rows[:, None] + np.arange(ncol)[None, :]) % 11).astype(np.int64)         # 60 x 62
    D = C - A @ eta                                                             # broadcast: d - u * eta
    assert np.abs(W).max() < 2 ** 24 and np.abs(D).max() < 2 ** 24
    np.savez_compressed(os.path.join(OUT, name), **{k: v.astype(np.int32) for k, v in
                                                    dict(A=A, B=B, W=W, eta=eta, C=C, D=D).items()})


if __name__ == "__main__":
    integer_gemm("circulant_int_gemm.npz")
    circulant("single_launch_circulant.npz", gen.single_launch_case())
    circulant("accuracy_circulant.npz", gen.accuracy_case())
    dense("cfg1_n2048.npz", gen.config(1, scale=8)[0])
    dense("cfg1_n16384.npz", gen.config(1)[0])
    print("fixtures written to", OUT)
