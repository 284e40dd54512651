"""Thin Python view of the isai_solve.h handle (setup + preconditioned solve)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._ffi import DeviceArray, SetupOpts, call, lib

PRECOND = {"none": 0, "isai": 1, "ilu": 2}
PRECISION = {"fp64": 0, "fp32": 1, "bf16": 2}
SOLVER = {"bicgstab": 0, "fgmres": 1}


class Matrix:
    """A BSR matrix resident in device memory."""

    def __init__(self, rp: np.ndarray, ci: np.ndarray, val: np.ndarray):
        self.nb, self.nnzb, self.bs = len(rp) - 1, len(ci), val.shape[1]
        self.n = self.nb * self.bs
        self.d_rp = DeviceArray.from_numpy(np.ascontiguousarray(rp, np.int32))
        self.d_ci = DeviceArray.from_numpy(np.ascontiguousarray(ci, np.int32))
        self.d_val = DeviceArray.from_numpy(np.ascontiguousarray(val, np.float64))

    def spmv(self, d_x: DeviceArray, d_y: DeviceArray, stream=None) -> None:
        call("isai_bsr_spmv", self.nb, self.nnzb, self.bs, self.d_rp.ptr, self.d_ci.ptr, self.d_val.ptr,
             d_x.ptr, d_y.ptr, stream)


class Handle:
    """isai_setup / isai_solve / isai_destroy."""

    def __init__(self, A: Matrix, precond: str = "isai", precision: str = "fp64", chunk: int = 0, stream=None):
        self.A = A
        self.stream = stream
        opts = SetupOpts(PRECOND[precond], PRECISION[precision], chunk)
        h = C.c_void_p()
        call("isai_setup", A.nb, A.nnzb, A.bs, A.d_rp.ptr, A.d_ci.ptr, A.d_val.ptr, C.byref(opts), C.byref(h),
             stream)
        self.h = h

    def times(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        call("isai_setup_times", self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def info(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        call("isai_handle_info", self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def factors(self):
        ptrs = [C.c_void_p() for _ in range(7)]
        call("isai_handle_factors", self.h, *[C.byref(p) for p in ptrs])
        return [p.value for p in ptrs]

    def set_preconditioner(self, precond: str, precision: str = "fp64") -> None:
        call("isai_set_preconditioner", self.h, PRECOND[precond], PRECISION[precision], self.stream)

    def apply(self, d_v: DeviceArray, d_z: DeviceArray) -> None:
        call("isai_precond_apply", self.h, d_v.ptr, d_z.ptr, self.stream)

    def solve(self, d_b: DeviceArray, d_x: DeviceArray, solver: str = "bicgstab", tol: float = 1e-10,
              maxit: int = 1000, restart: int = 30, check: bool = True):
        it, rr = C.c_int(), C.c_double()
        st = lib.isai_solve(self.h, SOLVER[solver], restart, tol, maxit, d_b.ptr, d_x.ptr, C.byref(it),
                            C.byref(rr), self.stream)
        if check and st != 0:
            from ._ffi import check as _check
            _check(st, "isai_solve")
        return st, it.value, rr.value

    def close(self) -> None:
        if getattr(self, "h", None):
            lib.isai_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
