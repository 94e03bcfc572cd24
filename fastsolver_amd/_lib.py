"""ctypes loaders for the in-tree native libraries.

The device library must be the one that runs: there is no CPU fallback, and a
missing or unloadable library raises immediately.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
DEVICE_LIB = os.path.join(HERE, "libfastsolver_amd.so")
GEN_LIB = os.path.join(HERE, "libfastsolver_amd_gen.so")

_dev = None
_gen = None

_i32, _i64, _dbl, _ptr = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p


class System(ctypes.Structure):
    _fields_ = [("n", _i64), ("threshold", _i64), ("nrhs", _i32), ("nranks", _i32), ("ranks", _ptr),
                ("X", _ptr), ("ldx", _i64), ("V", _ptr), ("ldv", _i64), ("K", _ptr), ("ldk", _i64),
                ("klayout", _i32), ("pad0", _i32), ("x", _ptr), ("ldxs", _i64)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("nsys", _i32), ("precision", _i32), ("height", _i32), ("nleaves", _i32),
                ("workspace_bytes", _i64), ("flops_factor", _dbl), ("flops_leaf", _dbl), ("flops_reduce", _dbl),
                ("flops_node", _dbl), ("flops_bcast", _dbl), ("bytes_leaf", _dbl), ("bytes_reduce", _dbl),
                ("bytes_bcast", _dbl), ("flops_matvec", _dbl), ("bytes_matvec", _dbl), ("flops_solve_only", _dbl),
                ("bytes_solve_only", _dbl), ("launches_factor", _i32), ("pad0", _i32)]


class SolveReport(ctypes.Structure):
    _fields_ = [("iterations", _i32), ("converged", _i32), ("relres", _dbl), ("relres_first", _dbl)]


# every exported symbol of include/fastsolver_amd.h with its argument types
DEVICE_SYMBOLS = {
    "hodlr_plan_create": ([ctypes.POINTER(_ptr), ctypes.POINTER(System), _i32, _i32, _ptr], ctypes.c_int),
    "hodlr_plan_destroy": ([_ptr], ctypes.c_int),
    "hodlr_plan_query": ([_ptr, ctypes.POINTER(PlanInfo)], ctypes.c_int),
    "hodlr_factor_solve": ([_ptr], ctypes.c_int),
    "hodlr_leaf_solve": ([_ptr], ctypes.c_int),
    "hodlr_level_reduce": ([_ptr, _i32], ctypes.c_int),
    "hodlr_level_node_solve": ([_ptr, _i32], ctypes.c_int),
    "hodlr_level_broadcast": ([_ptr, _i32], ctypes.c_int),
    "hodlr_extract_solution": ([_ptr], ctypes.c_int),
    "hodlr_solve_only": ([_ptr, _ptr, _ptr, _ptr, _ptr], ctypes.c_int),
    "hodlr_matvec": ([_ptr, _ptr, _ptr, _ptr, _ptr], ctypes.c_int),
    "hodlr_residual": ([_ptr, _ptr, ctypes.POINTER(_dbl)], ctypes.c_int),
    "hodlr_solve": ([_ptr, _dbl, _i32, ctypes.POINTER(SolveReport)], ctypes.c_int),
    "hodlr_status": ([_ptr, _ptr, _ptr, ctypes.POINTER(_i32)], ctypes.c_int),
    "hodlr_workspace": ([_ptr, _i32, ctypes.POINTER(_ptr), ctypes.POINTER(_i64), ctypes.POINTER(_i32)], ctypes.c_int),
    "hodlr_use_graph": ([_ptr, _i32], ctypes.c_int),
    "hodlr_autotune": ([_ptr, _i32, ctypes.POINTER(_i32)], ctypes.c_int),
    "hodlr_tune_force": ([_ptr, _i32], ctypes.c_int),
    "hodlr_profile_enable": ([_ptr, _i32], ctypes.c_int),
    "hodlr_profile_reset": ([_ptr], ctypes.c_int),
    "hodlr_profile_read": ([_ptr, _ptr, _ptr, _ptr, _ptr], ctypes.c_int),
    "hodlr_profile_category": ([_i32], ctypes.c_char_p),
    "hodlr_error_string": ([_i32], ctypes.c_char_p),
    "hodlr_version": ([], ctypes.c_char_p),
    "hodlr_dgemm_batched": ([_i32, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr],
                            ctypes.c_int),
    "hodlr_sgemm_batched": ([_i32, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr],
                            ctypes.c_int),
    "hodlr_sgemm_batched_bf16": ([_i32, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr],
                            ctypes.c_int),
    "hodlr_dgesv_batched": ([_i32, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr], ctypes.c_int),
    "hodlr_dgesv_batched_noalloc": ([_i32, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr], ctypes.c_int),
    "hodlr_device_count": ([ctypes.POINTER(_i32)], ctypes.c_int),
    "hodlr_device_name": ([ctypes.c_char_p, _i32], ctypes.c_int),
    "hodlr_device_malloc": ([ctypes.POINTER(_ptr), _i64], ctypes.c_int),
    "hodlr_device_free": ([_ptr], ctypes.c_int),
    "hodlr_memcpy": ([_ptr, _ptr, _i64, _i32], ctypes.c_int),
    "hodlr_stream_create": ([ctypes.POINTER(_ptr)], ctypes.c_int),
    "hodlr_stream_destroy": ([_ptr], ctypes.c_int),
    "hodlr_stream_synchronize": ([_ptr], ctypes.c_int),
    "hodlr_device_synchronize": ([], ctypes.c_int),
    "hodlr_event_create": ([ctypes.POINTER(_ptr)], ctypes.c_int),
    "hodlr_event_destroy": ([_ptr], ctypes.c_int),
    "hodlr_event_record": ([_ptr, _ptr], ctypes.c_int),
    "hodlr_event_elapsed_ms": ([_ptr, _ptr, ctypes.POINTER(_dbl)], ctypes.c_int),
}

GEN_SYMBOLS = {
    "hodlr_gen_philox4x32_10": ([_ptr, _ptr, _ptr], None),
    "hodlr_gen_stream_values": ([ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                 _i64, _i64, _ptr], ctypes.c_int),
    "hodlr_gen_lapack_name": ([ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
    "hodlr_gen_circulant": ([_i64, _i64, ctypes.c_int, ctypes.c_int, _dbl, _i64, _i64, _ptr, _i64, _ptr, _i64, _ptr,
                             _i64, ctypes.c_int], ctypes.c_int),
    "hodlr_gen_rhs_drand48": ([_i64, _i64, ctypes.c_int, ctypes.c_int, _i64, _i64, _ptr, _i64], ctypes.c_int),
    "hodlr_gen_random": ([_i64, _i64, ctypes.c_int, ctypes.c_int, _ptr, _ptr, _ptr, _i64, _ptr, _i64, _ptr, _i64,
                          ctypes.c_int], ctypes.c_int),
    "hodlr_gen_tree": ([_i64, _i64, ctypes.c_int, ctypes.c_int, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr], ctypes.c_int),
    "hodlr_gen_leaves": ([_i64, _i64, ctypes.c_int, ctypes.c_int, _ptr, _ptr, _ptr, _ptr], ctypes.c_int),
}


def _bind(lib, table):
    for name, (args, res) in table.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


def device_lib():
    """The HIP library; raises if it was not built (no fallback path exists)."""
    global _dev
    if _dev is None:
        if not os.path.exists(DEVICE_LIB):
            raise RuntimeError(f"{DEVICE_LIB} is missing: build it with `python -m fastsolver_amd.build` "
                               "(the solver has no CPU fallback)")
        _dev = _bind(ctypes.CDLL(DEVICE_LIB), DEVICE_SYMBOLS)
    return _dev


def gen_lib():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_LIB):
            from .build import build_gen
            build_gen()
        _gen = _bind(ctypes.CDLL(GEN_LIB), GEN_SYMBOLS)
    return _gen


def check(rc, what="call"):
    if rc != 0:
        msg = device_lib().hodlr_error_string(rc).decode()
        raise HodlrError(rc, f"{what} failed: {msg} ({rc})")
    return rc


class HodlrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code
