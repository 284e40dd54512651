"""torch.autograd binding of the differentiation form (include/isai_adjoint.h).

``solve(val, b, rowptr, colind)`` returns x = A^{-1} b for the BSR matrix A
(bs = 4, values ``val`` of shape (nnzb, 4, 4)) through the library's
ISAI-preconditioned BiCGSTAB.  Its backward pass builds the adjoint handle from
the forward factors (no refactorization), solves A^T lambda = dL/dx, and returns
dL/db = lambda and dL/dA = -lambda x^T restricted to the BSR pattern (the
library's outer-product kernel).  Torch supplies device memory and the stream;
every numerical step runs in libisai_amd.so.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._ffi import SetupOpts, call, check, lib
from .solver import PRECISION, PRECOND, SOLVER

# iteration counts and residuals of the most recent forward / adjoint solves
last_stats: dict = {}


class _Handle:
    def __init__(self, h: C.c_void_p):
        self.h = h

    def __del__(self):
        if self.h:
            lib.isai_destroy(self.h)
            self.h = None


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _solve(h, b: torch.Tensor, tol: float, maxit: int, stream: int, what: str):
    x = torch.zeros_like(b)
    it, rr = C.c_int(), C.c_double()
    st = lib.isai_solve(h, SOLVER["bicgstab"], 30, tol, maxit, b.data_ptr(), x.data_ptr(), C.byref(it), C.byref(rr),
                        stream)
    check(st, what)
    return x, it.value, rr.value


def _check_inputs(val, b, rowptr, colind):
    for name, t, dt in (("val", val, torch.float64), ("b", b, torch.float64), ("rowptr", rowptr, torch.int32),
                        ("colind", colind, torch.int32)):
        if not t.is_cuda or t.dtype != dt or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous {dt} tensor on the GPU")
    nb = rowptr.numel() - 1
    if val.shape != (colind.numel(), 4, 4) or b.numel() != 4 * nb:
        raise ValueError("shapes: val (nnzb, 4, 4), b (4 * nb,), rowptr (nb + 1,), colind (nnzb,)")


class BsrSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, val, b, rowptr, colind, tol, maxit, precond, precision):
        _check_inputs(val, b, rowptr, colind)
        nb, nnzb = rowptr.numel() - 1, colind.numel()
        stream = _stream()
        h = C.c_void_p()
        opts = SetupOpts(PRECOND[precond], PRECISION[precision], 0)
        call("isai_setup", nb, nnzb, 4, rowptr.data_ptr(), colind.data_ptr(), val.data_ptr(), C.byref(opts),
             C.byref(h), stream)
        handle = _Handle(h)
        x, it, rr = _solve(h, b, tol, maxit, stream, "isai_solve (forward)")
        last_stats["forward"] = (it, rr)
        ctx.handle = handle
        ctx.tol, ctx.maxit = tol, maxit
        ctx.save_for_backward(val, x, rowptr, colind)
        return x

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gx):
        val, x, rowptr, colind = ctx.saved_tensors
        nb, nnzb = rowptr.numel() - 1, colind.numel()
        stream = _stream()
        adj = C.c_void_p()
        call("isai_adjoint_handle", ctx.handle.h, C.byref(adj), stream)
        adj_handle = _Handle(adj)
        lam, it, rr = _solve(adj_handle.h, gx.contiguous(), ctx.tol, ctx.maxit, stream, "isai_solve (adjoint)")
        last_stats["adjoint"] = (it, rr)
        grad_val = None
        if ctx.needs_input_grad[0]:
            grad_val = torch.empty_like(val)
            call("isai_bsr_outer_grad", nb, nnzb, 4, rowptr.data_ptr(), colind.data_ptr(), lam.data_ptr(),
                 x.data_ptr(), -1.0, grad_val.data_ptr(), stream)
        del adj_handle
        grad_b = lam if ctx.needs_input_grad[1] else None
        return grad_val, grad_b, None, None, None, None, None, None


def solve(val: torch.Tensor, b: torch.Tensor, rowptr: torch.Tensor, colind: torch.Tensor, tol: float = 1e-12,
          maxit: int = 2000, precond: str = "isai", precision: str = "fp64") -> torch.Tensor:
    """x = A^{-1} b, differentiable in val and b."""
    return BsrSolve.apply(val, b, rowptr, colind, tol, maxit, precond, precision)
