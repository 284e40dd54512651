"""Thin Python view of the isai_solve.h handle (setup + preconditioned solve)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._ffi import DeviceArray, SetupOpts, call, lib

PRECOND = {"none": 0, "isai": 1, "ilu": 2}
PRECISION = {"fp64": 0, "fp32": 1, "bf16": 2}
SOLVER = {"bicgstab": 0, "fgmres": 1, "fgmres_mixed": 2, "fgmres_mixed_gmres": 3}


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

    def build_plan(self) -> dict:
        """The ISAI build plan chosen from the pair statistics (isai_handle_build_plan)."""
        n, ol, ou = C.c_int(), C.c_int(), C.c_int()
        call("isai_handle_build_plan", self.h, C.byref(n), C.byref(ol), C.byref(ou))
        return {"ncache": n.value, "ordered_L": ol.value, "ordered_U": ou.value}

    def set_graph(self, use_graph: bool = True) -> None:
        """BiCGSTAB iterations replayed from a captured hipGraph (isai_set_solve_graph)."""
        call("isai_set_solve_graph", self.h, int(use_graph))

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

    def set_inner(self, inner_tol: float = 3e-5, inner_maxit: int = 40) -> None:
        """Inner-solve parameters of the mixed-precision solver (isai_set_inner_solve)."""
        call("isai_set_inner_solve", self.h, inner_tol, inner_maxit)

    def last_stats(self):
        """(outer FGMRES steps, inner iterations) of the last solve."""
        o, i = C.c_int(), C.c_int()
        call("isai_last_solve_stats", self.h, C.byref(o), C.byref(i))
        return o.value, i.value

    def close(self) -> None:
        if getattr(self, "h", None):
            lib.isai_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pair_stats(d_mbrow: DeviceArray, d_pair_ptr: DeviceArray, nb: int, stream=None) -> dict:
    """isai_pair_stats: the analyze_pair.h figures of one factor's pair lists."""
    from ._ffi import PairStats
    out = PairStats()
    call("isai_pair_stats", nb, d_mbrow.ptr, d_pair_ptr.ptr, C.byref(out), stream)
    return out.as_dict()


def pair_column_stats(d_mbrow: DeviceArray, d_pair_ptr: DeviceArray, nb: int, stream=None):
    """isai_pair_column_stats: per-column N_j, Q_j and largest qlen (numpy arrays)."""
    outs = [DeviceArray((max(nb, 1),), np.int32) for _ in range(3)]
    call("isai_pair_column_stats", nb, d_mbrow.ptr, d_pair_ptr.ptr, *[o.ptr for o in outs], stream)
    call("isai_device_synchronize")
    return [o.to_numpy()[:nb] for o in outs]


class Batch:
    """isai_batch.h: many block-diagonal systems with several right-hand sides each."""

    def __init__(self, A: Matrix, sys_ptr: np.ndarray, nrhs: np.ndarray, precision: str = "fp32", stream=None):
        self.A = A
        self.stream = stream
        self.sys_ptr = np.ascontiguousarray(sys_ptr, np.int32)
        self.nrhs = np.ascontiguousarray(nrhs, np.int32)
        h = C.c_void_p()
        call("isai_batch_setup", len(self.nrhs), self.sys_ptr.ctypes.data, self.nrhs.ctypes.data, A.nnzb,
             A.d_rp.ptr, A.d_ci.ptr, A.d_val.ptr, PRECISION[precision], C.byref(h), stream)
        self.h = h
        n, p = C.c_size_t(), C.c_int()
        call("isai_batch_vector_length", self.h, C.byref(n))
        call("isai_batch_pairs", self.h, C.byref(p))
        self.length, self.npairs = n.value, p.value

    def offset(self, sys: int, rhs: int) -> int:
        o = C.c_size_t()
        call("isai_batch_vector_offset", self.h, sys, rhs, C.byref(o))
        return o.value

    def pack(self, vectors) -> np.ndarray:
        """vectors[s][j] (length 4*nb_s) -> one zero-padded multi-vector."""
        out = np.zeros(self.length)
        for s, vs in enumerate(vectors):
            for j, v in enumerate(vs):
                o = self.offset(s, j)
                out[o:o + len(v)] = v
        return out

    def unpack(self, multi: np.ndarray):
        out = []
        for s in range(len(self.nrhs)):
            n = 4 * int(self.sys_ptr[s + 1] - self.sys_ptr[s])
            out.append([multi[self.offset(s, j):self.offset(s, j) + n].copy() for j in range(self.nrhs[s])])
        return out

    def tuning(self):
        v, us, nc = C.c_int(), C.c_double(), C.c_int()
        call("isai_batch_tuning", self.h, C.byref(v), C.byref(us), C.byref(nc))
        return v.value, us.value, nc.value

    def tuning_table(self, operand: int):
        us = np.zeros(32)
        n = C.c_int()
        call("isai_batch_tuning_table", self.h, operand, us.ctypes.data, len(us), C.byref(n))
        return us[:n.value]

    def select(self, operand: int, candidate: int) -> None:
        call("isai_batch_select", self.h, operand, candidate)

    def setup_times(self):
        a, b = C.c_double(), C.c_double()
        call("isai_batch_setup_times", self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def spmm(self, d_X: DeviceArray, d_Y: DeviceArray) -> None:
        call("isai_batch_spmm", self.h, d_X.ptr, d_Y.ptr, self.stream)

    def solve(self, d_B: DeviceArray, d_X: DeviceArray, tol: float = 1e-10, maxit: int = 1000,
              use_graph: int = True, check: bool = True):
        """use_graph: 0 / False per-step launches, 1 / True their captured graph,
        2 the persistent kernel's captured graph, 3 the persistent kernel."""
        it = np.zeros(self.npairs, np.int32)
        rr = np.zeros(self.npairs)
        mx = C.c_int()
        st = lib.isai_batch_solve(self.h, tol, maxit, d_B.ptr, d_X.ptr, int(use_graph), it.ctypes.data,
                                  rr.ctypes.data, C.byref(mx), self.stream)
        if check and st != 0:
            from ._ffi import check as _check
            _check(st, "isai_batch_solve")
        return st, it, rr

    def close(self) -> None:
        if getattr(self, "h", None):
            lib.isai_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
