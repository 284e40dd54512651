"""GPU parity of the differentiation form (isai_adjoint.h and the autograd harness).

The specified differentiation matrix (2-D 9-pt 256^2 with the 2-D 5-pt problem's
coefficients, generator problem 4) is indefinite: the oracle's ISAI-BiCGSTAB
diverges on it (relres 5.5e3 after 3000 iterations), so the forward solve that
the gradient differentiates does not exist there.  The gradient checks run on
the same 9-pt 256^2 stencil with the well-posed coefficients of generator
problem 3 (diagonal 8.5 I + U(-0.2, 0.2), off-diagonal -I U(0.5, 1.5)); one test
keeps the specified matrix and checks that its forward solve reports failure.
"""
import ctypes

import numpy as np
import pytest

import _oracle as O
from gpu_pipeline import dev, reduced_apply_bound, rel_err
from isai_kernel_amd import DeviceArray, IsaiError, call, problems
from isai_kernel_amd.solver import Handle, Matrix

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(404)


def _download(ptr, shape, dtype):
    out = np.empty(shape, dtype)
    call("isai_memcpy_dtoh", out.ctypes.data, ptr, out.nbytes)
    return out


def _adjoint(h):
    g = ctypes.c_void_p()
    call("isai_adjoint_handle", h.h, ctypes.byref(g), None)
    adj = Handle.__new__(Handle)  # wrap the adjoint handle without a new setup
    adj.h, adj.stream, adj.A = g, None, None
    return adj


def _nonsymmetric_bsr(nb, density, rng):
    mask = rng.random((nb, nb)) < density
    np.fill_diagonal(mask, True)
    rp = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int32)
    ci = np.concatenate([np.nonzero(mask[i])[0] for i in range(nb)]).astype(np.int32)
    val = rng.uniform(-1, 1, (len(ci), 4, 4))
    return rp, ci, val


@pytest.mark.parametrize("case", ["stencil", "nonsymmetric"])
def test_bsr_transpose_bitexact(gpu, case):
    print(f"bsr transpose {case}", flush=True)
    if case == "stencil":
        rp, ci, val = problems.generate(2, 5, seed=1)
    else:
        rp, ci, val = _nonsymmetric_bsr(37, 0.15, RNG)
    nb, nnzb = len(rp) - 1, len(ci)
    d_rp, d_ci, d_v = dev(rp), dev(ci), dev(val)
    t_rp, t_ci, t_v = DeviceArray((nb + 1,), np.int32), DeviceArray((nnzb,), np.int32), DeviceArray(val.shape, np.float64)
    call("isai_bsr_transpose", nb, nnzb, 4, d_rp.ptr, d_ci.ptr, d_v.ptr, t_rp.ptr, t_ci.ptr, t_v.ptr, None)
    trp, tci, tval = O.bsr_transpose(rp, ci, val)
    assert np.array_equal(t_rp.to_numpy(), trp) and np.array_equal(t_ci.to_numpy(), tci)
    assert np.array_equal(t_v.to_numpy(), tval)


def test_adjoint_factors_bitexact(gpu):
    """Derived packed ILU(0) of A^T and the transposed ISAI factors equal the
    oracle's (the forward factors are bit-exact, the transforms are single ops)."""
    print("adjoint factors problem 3 L=14", flush=True)
    rp, ci, val = problems.generate(3, 14, seed=2)
    A = Matrix(rp, ci, val)
    h = Handle(A)
    adj = _adjoint(h)
    try:
        lu, L, U = O.isai_factors(rp, ci, val)
        trp, tci, lut = O.transpose_packed_lu(rp, ci, lu)
        MUt = O.bsr_transpose(*U)
        MLt = O.bsr_transpose(*L)
        Lrow, Lcol, Lval, Urow, Ucol, Uval, LU = adj.factors()
        nnzbL, nnzbU, _ = adj.info()
        assert nnzbL == len(MUt[1]) and nnzbU == len(MLt[1])
        assert np.array_equal(_download(LU, lut.shape, np.float64), lut)
        assert np.array_equal(_download(Lrow, MUt[0].shape, np.int32), MUt[0])
        assert np.array_equal(_download(Lcol, MUt[1].shape, np.int32), MUt[1])
        assert np.array_equal(_download(Lval, MUt[2].shape, np.float64), MUt[2])
        assert np.array_equal(_download(Urow, MLt[0].shape, np.int32), MLt[0])
        assert np.array_equal(_download(Ucol, MLt[1].shape, np.int32), MLt[1])
        assert np.array_equal(_download(Uval, MLt[2].shape, np.float64), MLt[2])
    finally:
        adj.close()
        h.close()


@pytest.mark.parametrize("kind,prec", [("isai", "fp64"), ("isai", "fp32"), ("isai", "bf16"), ("ilu", "fp64")])
def test_adjoint_precond_apply_and_solve(gpu, kind, prec):
    """Transposed ISAI apply z = M_L^T (M_U^T v) and transposed SpTRSV
    z = (L U)^{-T} v against the oracle; adjoint solve to 1e-12."""
    print(f"adjoint apply/solve {kind} {prec}", flush=True)
    rp, ci, val = problems.generate(3, 24, seed=3)
    A = Matrix(rp, ci, val)
    h = Handle(A, precond=kind, precision=prec)
    adj = _adjoint(h)
    try:
        lu, L, U = O.isai_factors(rp, ci, val)
        v = RNG.standard_normal(A.n)
        d_v, d_z = DeviceArray.from_numpy(v), DeviceArray((A.n,), np.float64)
        adj.apply(d_v, d_z)
        call("isai_device_synchronize")
        z = d_z.to_numpy()
        if kind == "isai":
            MUt, MLt = O.bsr_transpose(*U), O.bsr_transpose(*L)
            P = O.make_precond(1, A.nb, 4, L=MUt, U=MLt, f32=prec == "fp32", bf16=prec == "bf16")
            ref = O.precond_apply(P, v)
            if prec == "fp64":
                assert rel_err(z, ref) <= 1e-12, rel_err(z, ref)
            else:  # the derived FP32-accumulation bound, row by row
                y_ref = O.spmv_f32acc(*MUt, v, bf16=prec == "bf16")
                bound = reduced_apply_bound(MUt, MLt, v, y_ref, prec == "bf16")
                assert np.all(np.abs(z - ref) <= bound + 1e-300), np.max(np.abs(z - ref) / (bound + 1e-300))
        else:
            trp, tci, lut = O.transpose_packed_lu(rp, ci, lu)
            assert rel_err(z, O.sptrsv_lu(trp, tci, lut, v)) <= 1e-12
            # (L U)^T z = v with the forward factor
            LUd = O.bsr_to_dense(rp, ci, lu)
            Lf, Uf = np.tril(LUd, -1) + np.eye(A.n), np.triu(LUd)
            assert rel_err((Lf @ Uf).T @ z, v) <= 1e-12
        g = RNG.standard_normal(A.n)
        d_g, d_lam = DeviceArray.from_numpy(g), DeviceArray.zeros((A.n,), np.float64)
        st, it, rr = adj.solve(d_g, d_lam, "bicgstab", tol=1e-12, maxit=1000)
        trp, tci, tval = O.bsr_transpose(rp, ci, val)
        assert st == 0 and O.rel_residual(trp, tci, tval, g, d_lam.to_numpy()) <= 1e-12
        print(f"  adjoint iterations {it}", flush=True)
    finally:
        adj.close()
        h.close()


def test_outer_grad_bitexact(gpu):
    print("outer-product gradient problem 3 L=30", flush=True)
    rp, ci, val = problems.generate(3, 30, seed=4)
    nb, nnzb = len(rp) - 1, len(ci)
    lam, x = RNG.standard_normal(4 * nb), RNG.standard_normal(4 * nb)
    d_rp, d_ci, d_l, d_x = dev(rp), dev(ci), dev(lam), dev(x)
    d_g = DeviceArray((nnzb, 4, 4), np.float64)
    call("isai_bsr_outer_grad", nb, nnzb, 4, d_rp.ptr, d_ci.ptr, d_l.ptr, d_x.ptr, -1.0, d_g.ptr, None)
    call("isai_device_synchronize")
    assert np.array_equal(d_g.to_numpy(), O.outer_grad(rp, ci, lam, x, -1.0))


def test_specified_indefinite_matrix_reports_failure(gpu):
    print("problem 4 (indefinite 9-pt 256^2) forward solve", flush=True)
    rp, ci, val = problems.generate(4, 256, seed=0)
    A = Matrix(rp, ci, val)
    h = Handle(A)
    try:
        b = RNG.standard_normal(A.n)
        d_b, d_x = DeviceArray.from_numpy(b), DeviceArray.zeros((A.n,), np.float64)
        st, it, rr = h.solve(d_b, d_x, "bicgstab", tol=1e-12, maxit=300, check=False)
        assert st == -6 and rr > 1e-6
    finally:
        h.close()


# ------------------------------------------------------------ autograd harness
torch = pytest.importorskip("torch")


def _torch_problem(L, seed):
    from isai_kernel_amd import autograd  # noqa: F401  (imports torch)
    rp, ci, val = problems.generate(3, L, seed=seed)
    n = 4 * (len(rp) - 1)
    g = np.random.default_rng(seed)
    b, t = g.standard_normal(n), g.standard_normal(n)
    cu = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cu)  # noqa: E731
    return (rp, ci, val, b, t), (T(rp), T(ci), T(val), T(b), T(t))


def _loss_and_grads(Tv, Tb, Tt, Trp, Tci, tol=1e-12):
    from isai_kernel_amd.autograd import solve
    v = Tv.clone().requires_grad_(True)
    bb = Tb.clone().requires_grad_(True)
    x = solve(v, bb, Trp, Tci, tol=tol)
    loss = ((x - Tt) ** 2).sum()
    loss.backward()
    return loss.item(), v.grad.cpu().numpy(), bb.grad.cpu().numpy(), x.detach().cpu().numpy()


def test_autograd_matches_cpu_reference_gradient(gpu):
    """9-pt 256^2: dL/db and dL/dA against the oracle's FP64 gradient (its own
    BiCGSTAB on A and on A^T with its own ISAI setup of A^T)."""
    print("autograd vs CPU reference gradient 9-pt 256^2", flush=True)
    (rp, ci, val, b, t), (Trp, Tci, Tv, Tb, Tt) = _torch_problem(256, 7)
    loss, gA, gb, x = _loss_and_grads(Tv, Tb, Tt, Trp, Tci)
    from isai_kernel_amd.autograd import last_stats
    print(f"  forward {last_stats['forward']} adjoint {last_stats['adjoint']}", flush=True)
    lu, L, U = O.isai_factors(rp, ci, val)
    P = O.make_precond(1, len(rp) - 1, 4, L=L, U=U)
    xo, st, _, _ = O.bicgstab(rp, ci, val, P, b, tol=1e-12, maxit=2000)
    assert st == 0
    trp, tci, tval = O.bsr_transpose(rp, ci, val)
    _, Lt, Ut = O.isai_factors(trp, tci, tval)
    Pt = O.make_precond(1, len(rp) - 1, 4, L=Lt, U=Ut)
    lam, st, _, _ = O.bicgstab(trp, tci, tval, Pt, 2.0 * (xo - t), tol=1e-12, maxit=2000)
    assert st == 0
    gA_ref = O.outer_grad(rp, ci, lam, xo, -1.0)
    assert rel_err(x, xo) <= 1e-10
    assert rel_err(gb, lam) <= 1e-9, rel_err(gb, lam)
    assert rel_err(gA.ravel(), gA_ref.ravel()) <= 1e-9, rel_err(gA.ravel(), gA_ref.ravel())


def test_autograd_matches_finite_differences(gpu):
    """Central differences, h = 1e-6, on 64 random entries of A and 16 of b."""
    print("autograd vs finite differences 9-pt 256^2", flush=True)
    from isai_kernel_amd.autograd import solve
    (rp, ci, val, b, t), (Trp, Tci, Tv, Tb, Tt) = _torch_problem(256, 11)
    loss, gA, gb, _ = _loss_and_grads(Tv, Tb, Tt, Trp, Tci)
    h = 1e-6
    rng = np.random.default_rng(5)

    def L(v, bb):
        with torch.no_grad():
            x = solve(v, bb, Trp, Tci, tol=1e-12)
            return float(((x - Tt) ** 2).sum())

    errs = []
    for k, r, c in zip(rng.integers(0, len(ci), 64), rng.integers(0, 4, 64), rng.integers(0, 4, 64)):
        vp, vm = Tv.clone(), Tv.clone()
        vp[k, r, c] += h
        vm[k, r, c] -= h
        fd = (L(vp, Tb) - L(vm, Tb)) / (2 * h)
        errs.append(abs(fd - gA[k, r, c]) / max(abs(gA).max(), 1e-30))
    for i in rng.integers(0, len(b), 16):
        bp, bm = Tb.clone(), Tb.clone()
        bp[i] += h
        bm[i] -= h
        fd = (L(Tv, bp) - L(Tv, bm)) / (2 * h)
        errs.append(abs(fd - gb[i]) / max(abs(gb).max(), 1e-30))
    print(f"  max FD deviation / max|grad| = {max(errs):.3e}", flush=True)
    assert max(errs) <= 1e-4
