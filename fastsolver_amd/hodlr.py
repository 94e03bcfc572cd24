"""Python host layer over the C-ABI (include/fastsolver_amd.h).

``Plan`` binds device arrays of one or more HODLR systems; ``HodlrMatrix`` /
``FastSolver`` mirror the reference driver API (src/htree/hodlr_matrix.cc,
src/solver/fast_solver.cc) without the Legion context arguments.

Device memory, copies, streams and events go through the library's own HIP
runtime (``DeviceArray``); numpy holds the host side.  Column-major storage:
an N x k array has leading dimension N.
"""
import ctypes
import time
from typing import List, Sequence

import numpy as np

from . import gen as _gen
from ._lib import PlanInfo, SolveReport, System, check, device_lib, gen_lib

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16 = 2
PROFILE_NCAT = 22   # HODLR_PROFILE_NCAT
K_PACKED = 0
K_STACKED = 1


class DeviceArray:
    """A column-major device array (rows x cols) owned by this object."""

    def __init__(self, rows, cols=1, dtype=np.float64):
        self.rows, self.cols = int(rows), int(cols)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.rows * self.cols * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(device_lib().hodlr_device_malloc(ctypes.byref(p), max(self.nbytes, 1)), "hodlr_device_malloc")
        self.ptr = p.value

    @property
    def ld(self):
        return max(self.rows, 1)

    @classmethod
    def from_host(cls, a: np.ndarray):
        a = np.asarray(a)
        rows, cols = (a.shape[0], 1) if a.ndim == 1 else a.shape
        d = cls(rows, cols, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.asfortranarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(device_lib().hodlr_memcpy(self.ptr, a.ctypes.data, self.nbytes, 0), "upload")

    def download(self) -> np.ndarray:
        """Host copy (hodlr_memcpy waits for every stream of the device first)."""
        out = np.empty((self.rows, self.cols), dtype=self.dtype, order="F")
        check(device_lib().hodlr_memcpy(out.ctypes.data, self.ptr, self.nbytes, 1), "download")
        return out

    def upload_cols(self, a: np.ndarray, col0: int):
        """Write host columns a (rows x k) into columns col0 .. col0+k-1."""
        a = np.asfortranarray(a, dtype=self.dtype)
        assert a.shape[0] == self.rows and col0 + a.shape[1] <= self.cols
        off = col0 * self.rows * self.dtype.itemsize
        check(device_lib().hodlr_memcpy(self.ptr + off, a.ctypes.data, a.nbytes, 0), "upload")

    def free(self):
        if self.ptr:
            device_lib().hodlr_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(device_lib().hodlr_stream_create(ctypes.byref(s)), "hodlr_stream_create")
        self.handle = s.value

    def synchronize(self):
        check(device_lib().hodlr_stream_synchronize(self.handle), "hodlr_stream_synchronize")


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(device_lib().hodlr_event_create(ctypes.byref(e)), "hodlr_event_create")
        self.handle = e.value

    def record(self, stream):
        check(device_lib().hodlr_event_record(self.handle, stream.handle if stream else None))

    def elapsed_ms(self, end):
        ms = ctypes.c_double()
        check(device_lib().hodlr_event_elapsed_ms(self.handle, end.handle, ctypes.byref(ms)))
        return ms.value


def device_synchronize():
    check(device_lib().hodlr_device_synchronize(), "hodlr_device_synchronize")


def device_name():
    buf = ctypes.create_string_buffer(128)
    check(device_lib().hodlr_device_name(buf, 128))
    return buf.value.decode()


class DeviceSystem:
    """Device arrays of one system: X (N x (c+R)), V (N x R), K (packed), x (N x c)."""

    def __init__(self, spec: _gen.SystemSpec, X, V, K, x=None, klayout=K_PACKED, ldk=0):
        self.spec = spec
        self.X, self.V, self.K = X, V, K
        self.x = x if x is not None else DeviceArray(spec.n, spec.nrhs)
        self.klayout = klayout
        self.ldk = ldk
        self._ranks = np.ascontiguousarray(spec.ranks, dtype=np.int32)

    @classmethod
    def from_host(cls, spec, X, V, K):
        return cls(spec, DeviceArray.from_host(X), DeviceArray.from_host(V), DeviceArray.from_host(K))

    def cstruct(self):
        s = self.spec
        return System(s.n, s.threshold, s.nrhs, len(self._ranks), self._ranks.ctypes.data,
                      self.X.ptr, self.X.ld, self.V.ptr, max(self.V.ld, s.n),
                      self.K.ptr, self.ldk, self.klayout, 0, self.x.ptr, self.x.ld)

    def solution(self):
        return self.x.download()


class Plan:
    """A solver plan over a batch of systems (one launch per phase and height)."""

    def __init__(self, systems: Sequence[DeviceSystem], precision=PREC_FP64, stream=None):
        self.lib = device_lib()
        self.systems = list(systems)
        self.stream = stream if stream is not None else Stream()
        arr = (System * len(self.systems))(*[s.cstruct() for s in self.systems])
        self._arr = arr
        h = ctypes.c_void_p()
        check(self.lib.hodlr_plan_create(ctypes.byref(h), arr, len(self.systems), precision,
                                         ctypes.c_void_p(self.stream.handle)), "hodlr_plan_create")
        self.h = h
        self.precision = precision

    def close(self):
        if getattr(self, "h", None):
            self.lib.hodlr_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -------------------------------------------------------------- phases
    def info(self) -> PlanInfo:
        i = PlanInfo()
        check(self.lib.hodlr_plan_query(self.h, ctypes.byref(i)), "hodlr_plan_query")
        return i

    def factor_solve(self):
        check(self.lib.hodlr_factor_solve(self.h), "hodlr_factor_solve")

    def leaf_solve(self):
        check(self.lib.hodlr_leaf_solve(self.h), "hodlr_leaf_solve")

    def level_reduce(self, h):
        check(self.lib.hodlr_level_reduce(self.h, h), "hodlr_level_reduce")

    def level_node_solve(self, h):
        check(self.lib.hodlr_level_node_solve(self.h, h), "hodlr_level_node_solve")

    def level_broadcast(self, h):
        check(self.lib.hodlr_level_broadcast(self.h, h), "hodlr_level_broadcast")

    def extract_solution(self):
        check(self.lib.hodlr_extract_solution(self.h), "hodlr_extract_solution")

    def solve(self, tol=1e-12, maxiter=20, raise_on_fail=True) -> SolveReport:
        r = SolveReport()
        rc = self.lib.hodlr_solve(self.h, tol, maxiter, ctypes.byref(r))
        if raise_on_fail:
            check(rc, "hodlr_solve")
        return r

    def residual(self):
        cmax = max(s.spec.nrhs for s in self.systems)
        rel = np.zeros(len(self.systems) * cmax)
        mx = ctypes.c_double()
        check(self.lib.hodlr_residual(self.h, rel.ctypes.data, ctypes.byref(mx)), "hodlr_residual")
        return mx.value, rel.reshape(len(self.systems), cmax)

    def _ptrs(self, arrays):
        p = (ctypes.c_void_p * len(arrays))(*[a.ptr for a in arrays])
        ld = (ctypes.c_int64 * len(arrays))(*[a.ld for a in arrays])
        return p, ld

    def synchronize(self):
        self.stream.synchronize()

    def solve_only(self, rhs: List, out: List):
        pr, lr = self._ptrs(rhs)
        po, lo = self._ptrs(out)
        check(self.lib.hodlr_solve_only(self.h, pr, lr, po, lo), "hodlr_solve_only")

    def matvec(self, x: List, y: List):
        px, lx = self._ptrs(x)
        py, ly = self._ptrs(y)
        check(self.lib.hodlr_matvec(self.h, px, lx, py, ly), "hodlr_matvec")

    def status(self):
        info = self.info()
        li = np.zeros(info.nleaves, np.int32)
        nn = ctypes.c_int32()
        self.lib.hodlr_status(self.h, li.ctypes.data, None, ctypes.byref(nn))
        ni = np.zeros(max(nn.value, 1), np.int32)
        rc = self.lib.hodlr_status(self.h, li.ctypes.data, ni.ctypes.data, ctypes.byref(nn))
        return rc, li, ni[: nn.value]

    def workspace(self, s=0) -> np.ndarray:
        """Host copy of system s's workspace, N x (c+R): [x | A_sub^-1 U]."""
        p, ld, dt = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_int32()
        check(self.lib.hodlr_workspace(self.h, s, ctypes.byref(p), ctypes.byref(ld), ctypes.byref(dt)))
        xcols = _gen.tree_info(self.systems[s].spec)[2]
        out = np.empty((ld.value, xcols), dtype=np.float64 if dt.value == 0 else np.float32, order="F")
        self.synchronize()
        check(self.lib.hodlr_memcpy(out.ctypes.data, p, out.nbytes, 1), "workspace download")
        return out

    def autotune(self, reps=2):
        """Pick 64x64 or 128x128 tiles per GEMM launch by timing both; returns
        how many launches chose 128x128.  Factor again afterwards."""
        big = ctypes.c_int32()
        check(self.lib.hodlr_autotune(self.h, reps, ctypes.byref(big)), "hodlr_autotune")
        return big.value

    def force_tiling(self, cfg):
        """0: 64x64 GEMM tiles and 64-row TRSM steps everywhere, 1: the larger
        choices, -1: built-in defaults."""
        check(self.lib.hodlr_tune_force(self.h, cfg), "hodlr_tune_force")

    def use_graph(self, enable=True):
        check(self.lib.hodlr_use_graph(self.h, int(enable)), "hodlr_use_graph")

    def categories(self):
        return [self.lib.hodlr_profile_category(i).decode() for i in range(PROFILE_NCAT)]

    def profile(self, categories=None):
        """Time the launches of the named categories (None = all, [] = off)."""
        names = self.categories()
        if categories is None:
            mask = (1 << PROFILE_NCAT) - 1
        else:
            mask = 0
            for c in categories:
                mask |= 1 << names.index(c)
        check(self.lib.hodlr_profile_enable(self.h, mask))

    def profile_reset(self):
        check(self.lib.hodlr_profile_reset(self.h))

    def profile_read(self):
        n = PROFILE_NCAT
        t = np.zeros(n)
        c = np.zeros(n, np.int32)
        f = np.zeros(n)
        b = np.zeros(n)
        check(self.lib.hodlr_profile_read(self.h, t.ctypes.data, c.ctypes.data, f.ctypes.data, b.ctypes.data))
        names = [self.lib.hodlr_profile_category(i).decode() for i in range(n)]
        return {names[i]: {"ms": t[i], "launches": int(c[i]), "flops": f[i], "bytes": b[i]} for i in range(n)}


# ---------------------------------------------------------------------------
# Mirror of the reference driver API (test/single_launch.cc usage)
# ---------------------------------------------------------------------------
class Range:
    """Placement hint of the reference (src/core/range.h); ignored on one GPU."""

    def __init__(self, begin_or_size=0, size=None):
        self.begin, self.size = (0, begin_or_size) if size is None else (begin_or_size, size)


class HodlrMatrix:
    """HodlrMatrix(col, row, gl, sl, r, t, leaf, name) as hodlr_matrix.h:19-21."""

    def __init__(self, col, row, gl, sl, r, t, leaf=1, name="global"):
        self.rhs_cols, self.rhs_rows = col, row
        self.gloLevel, self.subLevel = gl, sl
        self.rank, self.threshold, self.leafSize = r, t, leaf
        self.file_rhs = name + "_rhs.txt"
        self.file_soln = name + "_soln.txt"
        self.spec = None
        self.dev = None
        self.seed = None
        self.timeInit = 0.0

    def create_tree(self):
        """Tree and device arrays (hodlr_matrix.cc:44-82); the extra rank*(gl-sl)
        columns of a two-stage run (hodlr_matrix.cc:48) are not used on one GPU."""
        self.spec = _gen.SystemSpec(self.rhs_rows, self.threshold, self.rhs_cols, [self.rank], kind="circulant")
        nl, md, xc, vc, kp = _gen.tree_info(self.spec)
        self.dev = DeviceSystem(self.spec, DeviceArray(self.rhs_rows, xc), DeviceArray(self.rhs_rows, max(vc, 1)),
                                DeviceArray(kp))

    def init_rhs(self, seed, procs=None):
        """drand48 right-hand side reseeded per leaf region (hodlr_matrix.cc:175-209),
        written into the first rhs_cols columns of X; either order with
        init_circulant_matrix."""
        t0 = time.perf_counter()
        self.seed = seed
        self.spec.seed = seed
        b = np.empty((self.rhs_rows, self.rhs_cols), order="F")
        rc = gen_lib().hodlr_gen_rhs_drand48(self.rhs_rows, self.threshold, self.rhs_cols, self.rank, seed, 0,
                                             b.ctypes.data, self.rhs_rows)
        if rc:
            raise RuntimeError(f"hodlr_gen_rhs_drand48 failed: {rc}")
        self.dev.X.upload_cols(b, 0)
        self.timeInit += time.perf_counter() - t0

    def init_circulant_matrix(self, diag, procs=None, skipU=False):
        """A = diag*I + U U^T, U(i,j) = (i+j) mod rank (hodlr_matrix.cc:94-109): U
        into X's columns rhs_cols.. (unless skipU), V and K; b is left alone."""
        t0 = time.perf_counter()
        self.spec.diag = diag
        X, V, K = _gen.generate(self.spec)
        if not skipU:
            self.dev.X.upload_cols(X[:, self.rhs_cols:], self.rhs_cols)
        self.dev.V.upload(V)
        self.dev.K.upload(K)
        self.timeInit += time.perf_counter() - t0

    def display_launch_time(self):
        print(f"Time cost for launching init-tasks :{self.timeInit} s")

    def get_num_leaf(self):
        return _gen.tree_info(self.spec)[0]

    def get_num_rhs(self):
        return self.rhs_cols

    def get_file_soln(self):
        return self.file_soln

    def solution(self):
        return self.dev.solution()

    def rhs(self):
        return self.dev.X.download()[:, : self.rhs_cols]

    def save_rhs(self):
        _save_rows(self.file_rhs, self.rhs(), self.spec, seed=self.seed)

    def save_solution(self):
        _save_rows(self.file_soln, self.solution(), self.spec)


def _save_rows(path, arr, spec, seed=None):
    """Text format of SaveRegionTask (save_region_task.cc:37-67): one row per line,
    tab-terminated values at 20 significant digits, one block per leaf region
    (preceded by the seed when given)."""
    rb, nr, _ = _gen.leaves(spec)
    with open(path, "w") as f:
        for b, n in zip(rb, nr):
            if seed is not None:
                f.write(f"{seed}\n")
            for i in range(b, b + n):
                f.write("".join(f"{float(v):.20g}\t" for v in arr[i]) + "\n")


class FastSolver:
    """FastSolver::bfs_solve (fast_solver.h:19-20) on one GPU."""

    def __init__(self, precision=PREC_FP64):
        self.precision = precision
        self.plan = None

    def bfs_solve(self, hmat: HodlrMatrix, procs=None, tol=1e-12, maxiter=20):
        """Factor + solve; the result is in hmat's solution array when this returns
        (FP32/BF16 plans refine in FP64 to `tol`)."""
        self.plan = Plan([hmat.dev], self.precision)
        if self.precision == PREC_FP64:
            self.plan.factor_solve()
        else:
            self.report = self.plan.solve(tol, maxiter)
        self.plan.synchronize()
        return self.plan


# ---------------------------------------------------------------------------
# Standalone batched operators (the task bodies' BLAS/LAPACK calls)
# ---------------------------------------------------------------------------
def _arr(ctype, vals):
    return (ctype * len(vals))(*vals)


def gemm_batched(ops, dtype=np.float64, stream=None, bf16=False):
    """ops: list of (transa, alpha, A, B, beta, C) with DeviceArray operands;
    C = alpha op(A) B + beta C with op(A) = A ('n') or A^T ('t').  float32 with
    bf16=True runs the products on BF16 MFMA (FP32 accumulation)."""
    lib = device_lib()
    n = len(ops)
    tr = [o[0].encode() for o in ops]
    m = [o[5].rows for o in ops]
    nn = [o[5].cols for o in ops]
    k = [(o[2].cols if o[0] in "nN" else o[2].rows) for o in ops]
    args = (_arr(ctypes.c_int32, m), _arr(ctypes.c_int32, nn), _arr(ctypes.c_int32, k),
            _arr(ctypes.c_double, [o[1] for o in ops]), _arr(ctypes.c_void_p, [o[2].ptr for o in ops]),
            _arr(ctypes.c_int64, [o[2].ld for o in ops]), _arr(ctypes.c_void_p, [o[3].ptr for o in ops]),
            _arr(ctypes.c_int64, [o[3].ld for o in ops]), _arr(ctypes.c_double, [o[4] for o in ops]),
            _arr(ctypes.c_void_p, [o[5].ptr for o in ops]), _arr(ctypes.c_int64, [o[5].ld for o in ops]),
            stream.handle if stream else None)
    if dtype == np.float64:
        check(lib.hodlr_dgemm_batched(n, ctypes.c_char_p(b"".join(tr)), *args), "hodlr_dgemm_batched")
    elif bf16:
        check(lib.hodlr_sgemm_batched_bf16(n, *args), "hodlr_sgemm_batched_bf16")
    else:
        check(lib.hodlr_sgemm_batched(n, *args), "hodlr_sgemm_batched")


def dgesv_batched(As, Bs, stream=None):
    """Solve A_i X_i = B_i in place (A_i <- LU, B_i <- X_i); returns LAPACK info per system."""
    lib = device_lib()
    n = len(As)
    info = np.zeros(n, np.int32)
    check(lib.hodlr_dgesv_batched(n, _arr(ctypes.c_int32, [a.rows for a in As]),
                                  _arr(ctypes.c_int32, [b.cols for b in Bs]), _arr(ctypes.c_void_p, [a.ptr for a in As]),
                                  _arr(ctypes.c_int64, [a.ld for a in As]), _arr(ctypes.c_void_p, [b.ptr for b in Bs]),
                                  _arr(ctypes.c_int64, [b.ld for b in Bs]), info.ctypes.data,
                                  stream.handle if stream else None), "hodlr_dgesv_batched")
    return info
