"""ctypes binding of libisai_amd.so (the C ABI declared in include/*.h).

The library is the product; this module only declares argument types, turns
non-zero status codes into exceptions and moves numpy arrays to and from device
memory.  There is no Python or CPU fallback: importing fails loudly when the
shared library has not been built.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libisai_amd.so"

if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
    )

# One HIP runtime per process.  PyTorch-ROCm ships its own libamdhip64 (soname
# libamdhip64.so.7, loaded by the name libamdhip64.so from torch/lib); if this
# library were loaded first, its dependency would pull /opt/rocm's copy and torch
# would then load a second runtime that finds no device.  Loading torch first makes
# the dynamic linker resolve libamdhip64.so.7 to the copy already in the process.
try:
    import torch  # noqa: F401
except ImportError:
    pass

lib = C.CDLL(str(LIB_PATH))

P = C.c_void_p
I = C.c_int
SZ = C.c_size_t
U64 = C.c_uint64
D = C.c_double
F = C.c_float

# name -> argument types (every function returns int status unless listed in _RESTYPE)
SIGNATURES: dict[str, list] = {
    # isai_types.h
    "isai_last_error": [],
    "isai_status_string": [I],
    # isai_runtime.h
    "isai_device_count": [P],
    "isai_device_info": [P, I, P, P],
    "isai_set_device": [I],
    "isai_device_malloc": [P, SZ],
    "isai_device_free": [P],
    "isai_host_free": [P],
    "isai_memcpy_htod": [P, P, SZ],
    "isai_memcpy_dtoh": [P, P, SZ],
    "isai_memcpy_dtod_async": [P, P, SZ, P],
    "isai_fill_u32_async": [P, C.c_uint32, SZ, P],
    "isai_fill_f64_async": [P, D, SZ, P],
    "isai_copy_f64_async": [P, P, SZ, P],
    "isai_device_synchronize": [],
    "isai_stream_create": [P],
    "isai_stream_destroy": [P],
    "isai_stream_synchronize": [P],
    "isai_event_create": [P],
    "isai_event_destroy": [P],
    "isai_event_record": [P, P],
    "isai_event_synchronize": [P],
    "isai_event_elapsed_ms": [P, P, P],
    # isai_gen.h
    "isai_gen_problem_count": [I, I, P, P],
    "isai_gen_problem_fill": [I, I, U64, P, P, P],
    "isai_gen_normal": [U64, SZ, P],
    # csr2bsr.h
    "isai_csr_to_bsr_count": [I, I, P, P, P, P],
    "isai_csr_to_bsr_fill": [I, I, P, P, P, P, P, P],
    "isai_csr_to_bsr_device_count": [I, I, P, P, P, P, P],
    "isai_csr_to_bsr_device_fill": [I, I, P, P, P, P, P, P, P],
    # read_matrix.h
    "isai_read_matrix_market": [C.c_char_p, P, P, P, P, P, P],
    "isai_rel_errors_csr": [I, P, P, P, P, P, P, P, P],
    # bsr2csr.h
    "isai_bsr_to_csr": [I, I, I, P, P, P, P, P, P, P],
    # bcsc2bcsr.h
    "isai_tri_bcsr_pattern_count": [I, P, P, I, P],
    "isai_tri_bcsr_pattern": [I, P, P, I, P, P],
    "isai_map_bcsc2bcsr": [I, P, P, P, P, P],
    "isai_reorder_bcsc_to_bcsr": [I, I, P, P, P, P],
    # isai_bsr.h
    "isai_bcsc_pattern_count": [I, P, P, I, P],
    "isai_bcsc_pattern": [I, P, P, I, P, P],
    "isai_build_dinv": [I, I, P, P, P, I, P, P],
    "isai_pairs_count": [I, I, P, P, P, P, I, P, P, P],
    "isai_pairs_count_noalloc": [I, I, P, P, P, P, I, P, P, P],
    "isai_pairs_fill": [I, P, P, P, P, I, P, P, P, P],
    "isai_build_lower": [I, I, P, P, P, P, P, P, P, P, I, P],
    "isai_build_upper": [I, I, P, P, P, P, P, P, P, P, I, P],
    "isai_build_upper_lookup": [I, I, P, P, P, P, P, I, P, P, I, P],
    "isai_build_lu": [I, I, P] + [P] * 7 + [P] * 7 + [I, P],
    "isai_pair_stats": [I, P, P, P, P],
    "isai_pair_column_stats": [I, P, P, P, P, P, P],
    # bsrilu.h
    "isai_bsrilu0_inplace": [I, I, I, P, P, P, P, P],
    "isai_bsrilu0_analysis": [I, I, I, P, P, P, P, P],
    "isai_bsrilu0_factor": [P, P, P, P],
    "isai_bsrilu0_plan_info": [P, P, P],
    "isai_bsrilu0_destroy": [P],
    "isai_bsr_levels": [I, P, P, I, P, P],
    # cape_sptrsv.h
    "isai_csr_sptrsv_lu": [I, P, P, P, I, P],
    "isai_csr_sptrsv_work_size": [I, P],
    "isai_csr_sptrsv_lu_device": [I, I, P, P, P, P, P, P, I, I, P],
    "isai_bsr_sptrsv_analysis": [I, I, I, P, P, P, I, P, P],
    "isai_bsr_sptrsv_update_values": [P, P, P],
    "isai_bsr_sptrsv_solve": [P, P, P, P, P],
    "isai_bsr_sptrsv_set_spin_limit": [P, I],
    "isai_sptrsv_status": [P, P],
    "isai_bsr_sptrsv_destroy": [P],
    "isai_bsr_sptrsv_schedule": [P, P],
    # isai_solve.h
    "isai_setup": [I, I, I, P, P, P, P, P, P],
    "isai_setup_times": [P, P, P, P],
    "isai_set_preconditioner": [P, I, I, P],
    "isai_handle_info": [P, P, P, P],
    "isai_handle_factors": [P, P, P, P, P, P, P, P],
    "isai_precond_apply": [P, P, P, P],
    "isai_handle_status": [P, P],
    "isai_handle_set_spin_limit": [P, I],
    "isai_handle_build_plan": [P, P, P, P],
    "isai_solve": [P, I, I, D, I, P, P, P, P, P],
    "isai_destroy": [P],
    "isai_set_inner_solve": [P, D, I],
    "isai_last_solve_stats": [P, P, P],
    "isai_fgmres_arnoldi": [P, P, P, P, P, P, I],
    # isai_adjoint.h
    "isai_bsr_transpose": [I, I, I, P, P, P, P, P, P, P],
    "isai_adjoint_handle": [P, P, P],
    "isai_bsr_outer_grad": [I, I, I, P, P, P, P, D, P, P],
    # isai_batch.h
    "isai_batch_setup": [I, P, P, I, P, P, P, I, P, P],
    "isai_batch_vector_length": [P, P],
    "isai_batch_vector_offset": [P, I, I, P],
    "isai_batch_pairs": [P, P],
    "isai_batch_tuning": [P, P, P, P],
    "isai_batch_tuning_table": [P, I, P, I, P],
    "isai_batch_select": [P, I, I],
    "isai_batch_setup_times": [P, P, P],
    "isai_batch_spmm": [P, P, P, P],
    "isai_batch_solve": [P, D, I, P, P, I, P, P, P, P],
    "isai_batch_set_spin_limit": [P, C.c_longlong],
    "isai_set_solve_graph": [P, I],
    "isai_batch_destroy": [P],
    # bsr_tensor_SpMV.h
    "isai_bsr_spmv": [I, I, I, P, P, P, P, P, P],
    "isai_bsr_spmv_prec": [I, I, I, I, P, P, P, P, P, P],
    "isai_convert_values": [SZ, I, P, P, P],
    "isai_apply": [I, I, I, P, P, P, I, P, P, P, I, P, P, P, P],
    "isai_bsr_spmv_f32_dot": [I, I, P, P, P, P, P, P, I, P, P, P],
    "isai_handle_spmv_f32_dot": [P, P, P, P, I, P, P, P, P],
    "isai_bsr_spmv_bench": [P, P, P, I, I, I, P, P, I, I, P],
}
_RESTYPE = {"isai_last_error": C.c_char_p, "isai_status_string": C.c_char_p}

for _name, _args in SIGNATURES.items():
    _fn = getattr(lib, _name)
    _fn.argtypes = _args
    _fn.restype = _RESTYPE.get(_name, I)


class SetupOpts(C.Structure):
    """isai_setup_opts_t"""
    _fields_ = [("precond", I), ("precision", I), ("sptrsv_chunk", I), ("reserved", I * 5)]


class PairStats(C.Structure):
    """isai_pair_stats_t"""
    _fields_ = ([("nb", I), ("mnnz", I), ("total_pairs", C.c_longlong),
                 ("n_avg", D), ("q_avg", D), ("qlen_avg", D)]
                + [(f"{k}_{p}", I) for k in ("n", "q", "qlen") for p in ("max", "p50", "p90", "p95", "p99")]
                + [("bin_cols", I * 10), ("bin_max_q", I * 10),
                   ("qlen_hist", C.c_longlong * 10), ("bin_sum_q", C.c_longlong * 10)])

    def as_dict(self) -> dict:
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if isinstance(v, C.Array) else v
        return d


class IsaiError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        msg = lib.isai_last_error().decode(errors="replace")
        name = lib.isai_status_string(status).decode()
        super().__init__(f"{where}: {name} ({status}) {msg}")


def check(status: int, where: str = "isai") -> int:
    if status != 0:
        raise IsaiError(status, where)
    return status


def call(name: str, *args) -> int:
    """Call a C-ABI function and raise IsaiError on a non-zero status."""
    return check(getattr(lib, name)(*args), name)


def ptr(a) -> int | None:
    """Address of a numpy array (host) or DeviceArray (device); None for None."""
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    return int(a)


class DeviceArray:
    """A device allocation owned from Python (hipMalloc through the C ABI)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = P()
        call("isai_device_malloc", C.byref(p), max(self.nbytes, 1))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            call("isai_memcpy_htod", d.ptr, a.ctypes.data, a.nbytes)
        return d

    @classmethod
    def zeros(cls, shape, dtype) -> "DeviceArray":
        return cls.from_numpy(np.zeros(shape, dtype))

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            call("isai_memcpy_dtoh", out.ctypes.data, self.ptr, self.nbytes)
        return out

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        call("isai_memcpy_htod", self.ptr, a.ctypes.data, a.nbytes)

    def free(self) -> None:
        if getattr(self, "ptr", None):
            lib.isai_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def declared_symbols(include_dir: Path) -> list[str]:
    """Every isai_* function declared in the public headers."""
    import re

    names = set()
    for h in sorted(Path(include_dir).glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(isai_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)
