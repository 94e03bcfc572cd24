"""Host-side input generation (ctypes over libfastsolver_amd_gen.so).

Produces the arrays the solver consumes, in column-major (Fortran) order:
``X`` = [b | U(1) | ... | U(L)] (N x (c+R)), ``V`` (N x R) and the leaf blocks
``K`` (packed, leaf after leaf).  Presets ``config(i)`` are the five workload
families of the benchmark (i = 1..5) plus the reference's own circulant tests.
"""
import ctypes
import dataclasses
import math
import os
from typing import List, Optional

import numpy as np

from ._lib import gen_lib

K_PACKED = 0
K_STACKED = 1
UV_UNIFORM = 0
UV_DECAY = 1
SEED = 1245667  # test/single_launch.cc:25


class Family(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("system_id", ctypes.c_uint32), ("d_scale", ctypes.c_double),
                ("d_shift", ctypes.c_double), ("uv_mode", ctypes.c_int), ("decay", ctypes.c_double)]


@dataclasses.dataclass
class SystemSpec:
    n: int
    threshold: int
    nrhs: int
    ranks: List[int]
    kind: str = "random"          # "random" or "circulant"
    seed: int = SEED
    system_id: int = 0
    d_scale: float = 1.0 / 16     # <= 0 means 1/sqrt(leaf rows)
    d_shift: float = 4.0
    uv_mode: int = UV_UNIFORM
    decay: float = 25.0
    diag: float = 1.0e4           # circulant family only
    rhs_region_rows: int = 0      # circulant family only (0 = the leaves)

    def ranks_array(self):
        return np.ascontiguousarray(self.ranks, dtype=np.int32)

    def as_dict(self):
        return {"n": self.n, "threshold": self.threshold, "nrhs": self.nrhs, "ranks": list(self.ranks),
                "klayout": K_PACKED}


def tree_info(spec: SystemSpec):
    """(nleaves, maxdepth, xcols, vcols, kpacked) of the spec's tree."""
    L = gen_lib()
    r = spec.ranks_array()
    nl, md = ctypes.c_int(), ctypes.c_int()
    xc, vc, kp = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    rc = L.hodlr_gen_tree(spec.n, spec.threshold, spec.nrhs, len(r), r.ctypes.data, ctypes.byref(nl),
                          ctypes.byref(md), ctypes.byref(xc), ctypes.byref(vc), ctypes.byref(kp))
    if rc:
        raise ValueError(f"invalid HODLR tree for {spec}: code {rc}")
    return nl.value, md.value, xc.value, vc.value, kp.value


def leaves(spec: SystemSpec):
    nl = tree_info(spec)[0]
    r = spec.ranks_array()
    rb = np.zeros(nl, np.int64)
    nr = np.zeros(nl, np.int64)
    dp = np.zeros(nl, np.int32)
    rc = gen_lib().hodlr_gen_leaves(spec.n, spec.threshold, spec.nrhs, len(r), r.ctypes.data, rb.ctypes.data,
                                    nr.ctypes.data, dp.ctypes.data)
    if rc:
        raise ValueError(rc)
    return rb, nr, dp


def generate(spec: SystemSpec, X=None, V=None, K=None):
    """Fill (or allocate) X, V, K for one system; returns (X, V, K)."""
    nl, md, xc, vc, kp = tree_info(spec)
    n = spec.n
    if X is None:
        X = np.empty((n, xc), order="F")
    if V is None:
        V = np.empty((n, vc), order="F")
    if K is None:
        K = np.empty(kp)
    for a in (X, V, K):
        assert a.dtype == np.float64 and a.flags.f_contiguous
    L = gen_lib()
    if spec.kind == "circulant":
        assert len(spec.ranks) == 1
        rc = L.hodlr_gen_circulant(n, spec.threshold, spec.nrhs, spec.ranks[0], spec.diag, spec.seed,
                                   spec.rhs_region_rows, X.ctypes.data, X.shape[0], V.ctypes.data, V.shape[0],
                                   K.ctypes.data, 0, K_PACKED)
    else:
        fam = Family(spec.seed, spec.system_id, spec.d_scale, spec.d_shift, spec.uv_mode, spec.decay)
        r = spec.ranks_array()
        rc = L.hodlr_gen_random(n, spec.threshold, spec.nrhs, len(r), r.ctypes.data, ctypes.byref(fam),
                                X.ctypes.data, X.shape[0], V.ctypes.data, V.shape[0], K.ctypes.data, 0, K_PACKED)
    if rc:
        raise RuntimeError(f"generation failed for {spec}: code {rc}")
    return X, V, K


def stream_values(seed, system, kind, node, normal, first, count):
    out = np.empty(count)
    gen_lib().hodlr_gen_stream_values(seed, system, kind, node, int(normal), first, count, out.ctypes.data)
    return out


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    gen_lib().hodlr_gen_philox4x32_10(c, k, o)
    return list(o)


def lapack_name():
    buf = ctypes.create_string_buffer(512)
    if gen_lib().hodlr_gen_lapack_name(buf, 512) != 0:
        return None
    return buf.value.decode()


# ------------------------------------------------------------------ presets
def config(i: int, seed: int = SEED, scale: Optional[int] = None) -> List[SystemSpec]:
    """Workload families 1..5 (the benchmark's configurations).

    ``scale`` (a power of two) divides N while keeping the leaf size, for
    quick parity runs of the same family at a smaller depth."""
    def cut(n):
        return n if not scale else n // scale
    if i == 1:
        return [SystemSpec(cut(16384), 256, 1, [30], seed=seed, d_scale=1 / 16)]
    if i == 2:
        return [SystemSpec(cut(131072), 512, 4, [64], seed=seed, d_scale=-1.0)]
    if i == 3:
        return [SystemSpec(cut(262144), 1024, 8, [100], seed=seed, d_scale=1 / 32, uv_mode=UV_DECAY, decay=25.0)]
    if i == 4:
        count = 512 if not scale else max(1, 512 // scale)
        return batched_uneven(count, seed)
    if i == 5:
        return [SystemSpec(cut(1 << 20), 1024, 32, [128], seed=seed, d_scale=1 / 32, uv_mode=UV_DECAY, decay=32.0)]
    raise ValueError(i)


def batched_uneven(count, seed=SEED, nmin_blocks=16, nmax_blocks=256, rmin=8, rmax=64, cmax=16):
    """Independent systems of uneven N (multiples of 256), per-level ranks and nrhs."""
    specs = []
    for s in range(count):
        u = stream_values(seed, s, 5, 0, False, 0, 32)
        nb = nmin_blocks + int(u[0] * (nmax_blocks - nmin_blocks + 1))
        n = 256 * nb
        c = 1 + int(u[1] * cmax)
        depth = max(1, math.ceil(math.log2(nb)))
        ranks = [rmin + int(u[2 + d] * (rmax - rmin + 1)) for d in range(depth)]
        specs.append(SystemSpec(n, 256, c, ranks, seed=seed, system_id=s, d_scale=1 / 16))
    return specs


def single_launch_case():
    """test/single_launch.cc:21-38: 3 levels, rank 30, threshold 60, 2 rhs, diag 1e4."""
    return SystemSpec(60 * 8, 60, 2, [30], kind="circulant", seed=1245667, diag=1.0e4)


def accuracy_case():
    """test/main.cc:174-185,257-259: rank 6, N = 480, threshold 15, 2 rhs, diag 10, seed 1123."""
    return SystemSpec(15 * 32, 15, 2, [6], kind="circulant", seed=1123, diag=10.0)


def test3_1_case():
    """test/main.cc:157-171,245-281 (test 3.1, the reference's timed run): rank 300,
    N = 400 * 2^8 = 102,400, threshold 600 (leaves of 400, 8 levels), 2 rhs,
    diag 1e5, seed 1123."""
    return SystemSpec(400 * (1 << 8), 600, 2, [300], kind="circulant", seed=1123, diag=1.0e5)
