"""GPU parity of the setup handle and the preconditioned Krylov solves."""
import ctypes

import numpy as np
import pytest

import _oracle as O
from gpu_pipeline import reduced_apply_bound, rel_err
from isai_kernel_amd import DeviceArray, call, problems
from isai_kernel_amd.solver import Handle, Matrix

pytestmark = pytest.mark.gpu


def _download(ptr, shape, dtype):
    out = np.empty(shape, dtype)
    call("isai_memcpy_dtoh", out.ctypes.data, ptr, out.nbytes)
    return out


def test_setup_factors_bitexact_and_apply(gpu):
    print("setup factors problem 0 L=20", flush=True)
    rp, ci, val = problems.generate(0, 20, seed=1)
    A = Matrix(rp, ci, val)
    h = Handle(A)
    lu, L, U = O.isai_factors(rp, ci, val)
    Lrow, Lcol, Lval, Urow, Ucol, Uval, LU = h.factors()
    nnzbL, nnzbU, nlev = h.info()
    assert nnzbL == len(L[1]) and nnzbU == len(U[1]) and nlev == O.levels(rp, ci, True)[1]
    assert np.array_equal(_download(LU, lu.shape, np.float64), lu)
    assert np.array_equal(_download(Lval, L[2].shape, np.float64), L[2])
    assert np.array_equal(_download(Uval, U[2].shape, np.float64), U[2])
    assert np.array_equal(_download(Lrow, L[0].shape, np.int32), L[0])
    assert np.array_equal(_download(Ucol, U[1].shape, np.int32), U[1])
    v = np.random.default_rng(0).standard_normal(A.n)
    d_v, d_z = DeviceArray.from_numpy(v), DeviceArray((A.n,), np.float64)
    for kind, prec in (("isai", "fp64"), ("isai", "fp32"), ("isai", "bf16"), ("ilu", "fp64")):
        h.set_preconditioner(kind, prec)
        h.apply(d_v, d_z)
        call("isai_device_synchronize")
        P = O.make_precond(1 if kind == "isai" else 2, A.nb, 4, L=L, U=U, LU=(rp, ci, lu),
                           f32=prec == "fp32", bf16=prec == "bf16")
        ref = O.precond_apply(P, v)
        z = d_z.to_numpy()
        if kind == "isai" and prec != "fp64":
            # same rounded operands as the oracle's reduced path, FP32 accumulation in
            # another order: the derived row-wise bound
            y_ref = O.spmv_f32acc(*L, v, bf16=prec == "bf16")
            bound = reduced_apply_bound(L, U, v, y_ref, prec == "bf16")
            assert np.all(np.abs(z - ref) <= bound + 1e-300), (prec, np.max(np.abs(z - ref) / (bound + 1e-300)))
            # the handle keeps M_L v in FP32 between the two products; the public
            # two-product apply keeps it in FP64: the products read it rounded to
            # FP32 either way, so the results are identical
            p = 1 if prec == "fp32" else 2
            dt = np.float32 if p == 1 else np.uint16
            lo_L, lo_U = DeviceArray((L[2].size,), dt), DeviceArray((U[2].size,), dt)
            call("isai_convert_values", L[2].size, p, Lval, lo_L.ptr, None)
            call("isai_convert_values", U[2].size, p, Uval, lo_U.ptr, None)
            d_y, d_x = DeviceArray((A.n,), np.float64), DeviceArray((A.n,), np.float64)
            call("isai_apply", A.nb, 4, p, Lrow, Lcol, lo_L.ptr, nnzbL,
                 Urow, Ucol, lo_U.ptr, nnzbU, d_v.ptr, d_y.ptr, d_x.ptr, None)
            call("isai_device_synchronize")
            assert np.array_equal(d_x.to_numpy(), z), prec
        else:
            assert rel_err(z, ref) <= 1e-12, (kind, rel_err(z, ref))
    h.close()


@pytest.mark.parametrize("precond", ["isai", "ilu", "none"])
def test_bicgstab_plain_small(gpu, precond):
    """Problem 0 at its own size (2-D 5-pt 64^2, 4096 block rows), b = A*1."""
    print(f"bicgstab problem 0 precond={precond}", flush=True)
    rp, ci, val = problems.generate(0, 0, seed=0)
    A = Matrix(rp, ci, val)
    ones = np.ones(A.n)
    b = O.spmv(rp, ci, val, ones)
    h = Handle(A, precond=precond)
    d_b, d_x = DeviceArray.from_numpy(b), DeviceArray.zeros((A.n,), np.float64)
    st, it, rr = h.solve(d_b, d_x, "bicgstab", tol=1e-10, maxit=2000)
    x = d_x.to_numpy()
    assert st == 0 and rr <= 1e-10
    assert O.rel_residual(rp, ci, val, b, x) <= 1e-10  # independent FP64 check
    assert np.max(np.abs(x - 1.0)) <= 1e-7
    # iteration count within a couple of the oracle's textbook BiCGSTAB
    lu, L, U = O.isai_factors(rp, ci, val)
    P = O.make_precond({"isai": 1, "ilu": 2, "none": 0}[precond], A.nb, 4, L=L, U=U, LU=(rp, ci, lu))
    _, ost, oit, orr = O.bicgstab(rp, ci, val, P, b, tol=1e-10, maxit=2000)
    assert ost == 0 and abs(it - oit) <= 2, (it, oit)
    h.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
def test_fgmres_mixed_precision(gpu, precision):
    """27-pt convection-diffusion (small grid): FGMRES(30) with ISAI stored in
    FP64 / FP32 / BF16 reaches 1e-10 in FP64."""
    print(f"fgmres problem 2 L=12 precision={precision}", flush=True)
    rp, ci, val = problems.generate(2, 12, seed=0)
    A = Matrix(rp, ci, val)
    b = O.spmv(rp, ci, val, np.ones(A.n))
    h = Handle(A, precond="isai", precision=precision)
    d_b, d_x = DeviceArray.from_numpy(b), DeviceArray.zeros((A.n,), np.float64)
    st, it, rr = h.solve(d_b, d_x, "fgmres", tol=1e-10, maxit=3000, restart=30)
    x = d_x.to_numpy()
    assert st == 0 and rr <= 1e-10
    assert O.rel_residual(rp, ci, val, b, x) <= 1.5e-10
    lu, L, U = O.isai_factors(rp, ci, val)
    P = O.make_precond(1, A.nb, 4, L=L, U=U, f32=precision == "fp32", bf16=precision == "bf16")
    _, ost, oit, orr = O.fgmres(rp, ci, val, P, b, m=30, tol=1e-10, maxit=3000)
    assert ost == 0 and abs(it - oit) <= max(3, oit // 10), (it, oit)
    print(f"  gpu iters {it} oracle iters {oit}", flush=True)
    h.close()


def test_bicgstab_full_size_fp64_and_mixed(gpu):
    """96^3 7-point, b = A*1: FP64 ISAI and FP32 ISAI BiCGSTAB reach 1e-10."""
    print("bicgstab full size 96^3", flush=True)
    rp, ci, val = problems.generate(1, 96, seed=0)
    A = Matrix(rp, ci, val)
    d_one = DeviceArray.from_numpy(np.ones(A.n))
    d_b = DeviceArray((A.n,), np.float64)
    A.spmv(d_one, d_b)
    h = Handle(A)
    res = {}
    for prec in ("fp64", "fp32", "bf16"):
        h.set_preconditioner("isai", prec)
        d_x = DeviceArray.zeros((A.n,), np.float64)
        st, it, rr = h.solve(d_b, d_x, "bicgstab", tol=1e-10, maxit=500)
        x = d_x.to_numpy()
        assert st == 0 and rr <= 1e-10, (prec, st, rr)
        assert np.max(np.abs(x - 1.0)) <= 1e-7
        res[prec] = it
    print(f"  iterations {res}", flush=True)
    h.close()


@pytest.mark.parametrize("solver,prec", [("fgmres", "fp64"), ("fgmres", "bf16"), ("fgmres_mixed", "bf16"),
                                         ("fgmres_mixed_gmres", "bf16")])
def test_fgmres_arnoldi_relation(gpu, solver, prec):
    """One full FGMRES cycle (m = 10, no early stop) satisfies the flexible Arnoldi
    relation A Z_m = V_{m+1} Hbar_m (Saad, Iterative Methods, 2nd ed., Prop. 9.2)
    to 1e-12, with the basis V orthonormal to 1e-12 (CGS2 with the DGKS criterion);
    Z holds the preconditioned vectors (the mixed solver: the inner solutions)."""
    print(f"arnoldi relation {solver} {prec}", flush=True)
    rp, ci, val = problems.generate(2, 8, seed=4)
    A = Matrix(rp, ci, val)
    h = Handle(A, precond="isai", precision=prec)
    b = problems.normal(3, A.n)
    d_b, d_x = DeviceArray.from_numpy(b), DeviceArray.zeros((A.n,), np.float64)
    m = 10
    if solver.startswith("fgmres_mixed"):
        # two inner iterations per step keep the outer estimate |g_m| above tol ||b||
        # through the cycle (with three it falls below 1e-30 ||b|| at the last step, the
        # cycle ends as converged and V_m is not formed)
        h.set_inner(1e-5, 2)
        maxit = 2 * m
    else:
        maxit = m
    h.solve(d_b, d_x, solver, tol=1e-30, maxit=maxit, restart=m, check=False)
    mm, steps = ctypes.c_int(), ctypes.c_int()
    pv, pz = ctypes.c_void_p(), ctypes.c_void_p()
    hbar = np.zeros((m + 1) * m)
    call("isai_fgmres_arnoldi", h.h, ctypes.byref(mm), ctypes.byref(steps), ctypes.byref(pv), ctypes.byref(pz),
         hbar.ctypes.data, hbar.size)
    assert mm.value == m and steps.value == m
    V = _download(pv.value, (m + 1, A.n), np.float64)
    Z = _download(pz.value, (m, A.n), np.float64)
    Hb = hbar.reshape(m + 1, m)
    AZ = np.stack([O.spmv(rp, ci, val, Z[j]) for j in range(m)])
    VH = Hb.T @ V  # row j = sum_i Hbar[i, j] V_i
    for j in range(m):
        assert np.linalg.norm(AZ[j] - VH[j]) <= 1e-12 * np.linalg.norm(AZ[j]), j
    assert np.max(np.abs(V @ V.T - np.eye(m + 1))) <= 1e-12
    assert np.all(np.tril(Hb, -2) == 0) and np.all(np.diag(Hb, -1) > 0)  # upper Hessenberg, h_{j+1,j} = ||w||
    h.close()


@pytest.mark.parametrize("precond,precision", [("isai", "fp64"), ("isai", "bf16"), ("ilu", "fp64")])
def test_bicgstab_graph_replay_bit_identical(gpu, precond, precision):
    """isai_set_solve_graph: the iterations captured once as a hipGraph and replayed
    give the launched iterations' x, count and residual bit for bit; the second
    graph solve reuses the capture, a preconditioner switch recaptures."""
    print(f"bicgstab graph replay precond={precond} precision={precision}", flush=True)
    rp, ci, val = problems.generate(0, 0, seed=0)
    A = Matrix(rp, ci, val)
    b = O.spmv(rp, ci, val, np.random.default_rng(3).standard_normal(A.n))
    s = ctypes.c_void_p()
    call("isai_stream_create", ctypes.byref(s))
    try:
        h = Handle(A, precond=precond, precision=precision, stream=s)
        d_b = DeviceArray.from_numpy(b)

        def run():
            d_x = DeviceArray.zeros((A.n,), np.float64)
            st, it, rr = h.solve(d_b, d_x, "bicgstab", tol=1e-10, maxit=2000)
            assert st == 0 and rr <= 1e-10
            return d_x.to_numpy(), it, rr

        x0, it0, rr0 = run()
        h.set_graph(True)
        for _ in range(2):
            x1, it1, rr1 = run()
            assert it1 == it0 and rr1 == rr0 and np.array_equal(x1, x0), (it0, it1, rr0, rr1)
        # a different preconditioner between graph solves: recaptured, still equal
        h.set_preconditioner("none")
        h.set_graph(False)
        xn, itn, rrn = run()
        h.set_graph(True)
        xg, itg, rrg = run()
        assert itg == itn and rrg == rrn and np.array_equal(xg, xn)
        # an iteration limit that is not a multiple of the replayed group stops exactly there
        h.set_preconditioner(precond, precision)
        d_x = DeviceArray.zeros((A.n,), np.float64)
        st, it, _ = h.solve(d_b, d_x, "bicgstab", tol=1e-30, maxit=7, check=False)
        h.set_graph(False)
        d_y = DeviceArray.zeros((A.n,), np.float64)
        st2, it2, _ = h.solve(d_b, d_y, "bicgstab", tol=1e-30, maxit=7, check=False)
        assert (st, it) == (st2, it2) and it == 7
        assert np.array_equal(d_x.to_numpy(), d_y.to_numpy())
        h.close()
    finally:
        call("isai_stream_synchronize", s)
        call("isai_stream_destroy", s)
