"""ctypes binding of ``libgraphgp_amd.so`` (the C ABI of ``include/graphgp_amd.h``).

The library is built in place by ``graphgp_amd.build``.  There is no fallback: if the
library is missing or no GPU is visible, every compute call raises.
"""
from __future__ import annotations

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgraphgp_amd.so")

F32, F64 = 0, 1
PREC = {"fp64": 0, "fp32": 1, "fp32_increment": 2}

_c_size = ctypes.c_size_t
_i64 = ctypes.c_int64
_int = ctypes.c_int
_p = ctypes.c_void_p
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)

_SIGS = {
    "ggp_last_error": (ctypes.c_char_p, []),
    "ggp_abi_version": (_int, []),
    "ggp_build_tree_workspace": (_c_size, [_i64, _int, _int]),
    "ggp_build_tree": (_int, [_p, _i64, _int, _int, _p, _p, _p, _p, _c_size, _p]),
    "ggp_query_neighbors_workspace": (_c_size, [_i64, _int, _i64, _int, _int]),
    "ggp_query_preceding_neighbors": (_int, [_p, _p, _i64, _int, _int, _i64, _int, _p, _p, _c_size, _p]),
    "ggp_query_neighbors": (_int, [_p, _p, _i64, _int, _int, _p, _p, _i64, _int, _p, _p, _c_size, _p]),
    "ggp_compute_depths_workspace": (_c_size, [_i64, _int]),
    "ggp_compute_depths": (_int, [_p, _i64, _int, _i64, _p, _p, _c_size, _p]),
    "ggp_compute_depths_method": (_int, [_p, _i64, _int, _i64, _p, _int, _p, _c_size, _p]),
    "ggp_compute_depths_status": (_int, [_p, _u32p, _p]),
    "ggp_order_by_depth_workspace": (_c_size, [_i64]),
    "ggp_order_by_depth": (_int, [_p, _i64, _int, _int, _p, _p, _i64, _int, _p, ctypes.c_int32, _p, _p, _p, _p,
                                  _p, _c_size, _p]),
    "ggp_depth_offsets": (_int, [_p, _i64, _i64, _p, _p]),
    "ggp_build_graph_workspace": (_c_size, [_i64, _int, _int, _i64, _int]),
    "ggp_build_graph": (_int, [_p, _i64, _int, _int, _i64, _int, _p, _p, _p, _p, _i32p, _p, _c_size, _p]),
    "ggp_check_graph": (_int, [_p, _i64, _int, _i64, _p, _i64, _p, _p]),
    "ggp_refine_workspace": (_c_size, [_i64, _int, _i64, _int, _int, _int, _int]),
    "ggp_generate": (_int, [_p, _int, _i64, _int, _p, _i64, _int, _p, _p, _p, _int, _p, _int, _p, _int, _int, _p,
                            _c_size, _p]),
    "ggp_generate_sched": (_int, [_p, _int, _i64, _int, _p, _i64, _int, _p, _p, _p, _int, _p, _int, _p, _int, _int,
                                  _p, _p, _p, _p, _c_size, _p]),
    "ggp_graph_schedule_workspace": (_c_size, [_i64, _i64]),
    "ggp_graph_schedule": (_int, [_p, _int, _i64, _int, _p, _i64, _int, _p, _i64, _p, _p, _p, _p, _c_size, _p]),
    "ggp_generate_vjp_workspace": (_c_size, [_i64, _int, _i64, _int, _int, _int]),
    "ggp_generate_vjp": (_int, [_p, _int, _i64, _int, _p, _i64, _int, _p, _p, _p, _int, _p, _int, _p, _int, _p,
                                _int, _p, _int, _p, _int, _p, _c_size, _p]),
    "ggp_generate_inv_vjp": (_int, [_p, _int, _i64, _int, _p, _i64, _int, _p, _p, _p, _int, _p, _int, _p, _int, _p,
                                    _int, _p, _int, _p, _c_size, _p]),
    "ggp_generate_logdet_vjp": (_int, [_p, _int, _i64, _int, _p, _i64, _int, _p, _p, _int, _p, _p, _int, _p, _c_size,
                                       _p]),
    "ggp_generate_inv": (_int, [_p, _int, _i64, _int, _p, _i64, _int, _p, _p, _p, _int, _p, _int, _p, _int, _int,
                                _p, _c_size, _p]),
    "ggp_generate_logdet": (_int, [_p, _int, _i64, _int, _p, _i64, _int, _p, _p, _int, _p, _int, _p, _c_size, _p]),
    "ggp_refine": (_int, [_p, _int, _i64, _int, _p, _i64, _int, _p, _p, _int, _p, _int, _p, _int, _p, _int, _p,
                          _c_size, _p]),
    "ggp_refine_inv": (_int, [_p, _int, _i64, _int, _p, _i64, _int, _p, _p, _int, _p, _int, _p, _p, _int, _int, _p,
                              _c_size, _p]),
    "ggp_refine_logdet": (_int, [_p, _int, _i64, _int, _p, _i64, _int, _p, _p, _int, _p, _int, _p, _c_size, _p]),
    "ggp_dense_workspace": (_c_size, [_i64]),
    "ggp_generate_dense": (_int, [_p, _int, _i64, _int, _p, _p, _int, _p, _int, _p, _int, _p, _c_size, _p]),
    "ggp_generate_dense_inv": (_int, [_p, _int, _i64, _int, _p, _p, _int, _p, _int, _p, _int, _p, _c_size, _p]),
    "ggp_generate_dense_logdet": (_int, [_p, _int, _i64, _int, _p, _p, _int, _p, _p, _c_size, _p]),
    "ggp_generate_logdet_noalloc": (_int, [_p, _int, _i64, _int, _p, _i64, _int, _p, _p, _int, _p, _int, _p]),
    "ggp_cov_matrix": (_int, [_p, _i64, _p, _i64, _int, _int, _p, _p, _int, _p, _p]),
    "ggp_profile_events": (_int, [_p, _p]),
    "ggp_stage_profile": (_int, [_int]),
    "ggp_stage_times": (_int, [_int, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_int)]),
    "ggp_refine_counters": (_int, [_p, _u32p, _p]),
    "ggp_set_cu_budget": (_int, [_int]),
    "ggp_set_yield": (_int, [_p, _int]),
}

_lib = None


class HipError(RuntimeError):
    """An error reported by the HIP library."""


def lib():
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"graphgp_amd: {LIB_PATH} is missing; build it with `python -m graphgp_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().ggp_last_error().decode()
        if rc == -1:
            raise ValueError(msg)
        raise HipError(f"{what}: {msg} (status {rc})")


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("graphgp_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def to_host(t):
    """A small device tensor copied to the host on the current stream, after waiting for
    that stream.  The copy lands in pinned memory: the runtime stages copies into pageable
    memory through a buffer that all host threads share, so one could wait on another
    thread's unrelated work."""
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t, non_blocking=True)
    torch.cuda.current_stream().synchronize()
    return h


def stage_profile(enable: bool) -> None:
    """Switch the library's stage timing on (clearing its sums) or off, for this thread."""
    check(lib().ggp_stage_profile(1 if enable else 0), "stage_profile")


def stage_times() -> dict:
    """{stage name: (total device ms, occurrences)} recorded since stage_profile(True)."""
    L = lib()
    n = L.ggp_stage_times(0, None, None, None)
    check(min(n, 0), "stage_times")
    names = (ctypes.c_char_p * max(n, 1))()
    ms = (ctypes.c_double * max(n, 1))()
    cnt = (ctypes.c_int * max(n, 1))()
    n = L.ggp_stage_times(n, names, ms, cnt)
    check(min(n, 0), "stage_times")
    return {names[i].decode(): (ms[i], cnt[i]) for i in range(n)}


def workspace(nbytes: int):
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device())


def dtype_code(t) -> int:
    if t.dtype == torch.float32:
        return F32
    if t.dtype == torch.float64:
        return F64
    raise TypeError(f"unsupported dtype {t.dtype} (float32 or float64 expected)")
