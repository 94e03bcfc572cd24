"""Helpers shared by the GPU parity tests: build inputs on the host, run the
oracle on the same arrays, upload to the device."""
import numpy as np

import oracle
from fastsolver_amd import gen


def host_system(spec):
    X, V, K = gen.generate(spec)
    return X, V, K


def oracle_solve(spec, X, V, K, **kw):
    """Reference fused factor+solve on copies; returns the full X after the solve."""
    Xo = X.copy(order="F")
    Ko = K.copy()
    rc = oracle.hodlr_solve(spec.as_dict(), Xo, V, Ko, **kw)
    assert rc == 0, rc
    return Xo


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def colblock_rel(a, b, tree_cols):
    """max over column ranges of ||a - b|| / ||b||"""
    worst = 0.0
    for lo, hi in tree_cols:
        nb = np.linalg.norm(b[:, lo:hi])
        if nb > 0:
            worst = max(worst, float(np.linalg.norm(a[:, lo:hi] - b[:, lo:hi]) / nb))
    return worst
