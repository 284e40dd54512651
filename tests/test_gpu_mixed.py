"""GPU parity of the mixed-precision form (ISAI_SOLVER_FGMRES_MIXED): an FP64 FGMRES
outer loop whose flexible preconditioner is an inner BiCGSTAB in FP32 storage, against
the oracle's restatement of the same inner-outer scheme (ora_fgmres_mixed) and an
independent FP64 residual."""
import ctypes

import numpy as np
import pytest

import _oracle as O
from isai_kernel_amd import DeviceArray, call, problems
from isai_kernel_amd.solver import Handle, Matrix

pytestmark = pytest.mark.gpu


def _oracle_precond(rp, ci, val, prec):
    _, L, U = O.isai_factors(rp, ci, val)
    return O.make_precond(1, len(rp) - 1, 4, L=L, U=U, f32=prec == "fp32", bf16=prec == "bf16")


@pytest.mark.parametrize("prec", ["bf16", "fp32", "fp64"])
@pytest.mark.parametrize("problem,L", [(1, 24), (2, 12), (0, 64)])
def test_mixed_solve_matches_oracle_scheme(gpu, prec, problem, L):
    print(f"mixed fgmres problem={problem} L={L} isai={prec}", flush=True)
    rp, ci, val = problems.generate(problem, L, seed=2)
    A = Matrix(rp, ci, val)
    b = problems.normal(7, A.n)
    h = Handle(A, precond="isai", precision=prec)
    h.set_inner(1e-5, 40)
    d_b, d_x = DeviceArray.from_numpy(b), DeviceArray.zeros((A.n,), np.float64)
    st, it, rr = h.solve(d_b, d_x, "fgmres_mixed", tol=1e-10, maxit=400, restart=10)
    outer, inner = h.last_stats()
    x = d_x.to_numpy()
    assert st == 0 and rr <= 1e-10 and inner == it
    rr_ind = O.rel_residual(rp, ci, val, b, x)  # independent FP64 residual
    assert rr_ind <= 1e-10, rr_ind
    # the oracle's restatement: FP64 ISAI handles run the inner apply with FP32 factors
    P = _oracle_precond(rp, ci, val, "fp32" if prec == "fp64" else prec)
    xo, sto, ito, outo, rro = O.fgmres_mixed(rp, ci, val, P, b, m=10, tol=1e-10, maxit=400, inner_tol=1e-5,
                                             inner_maxit=40)
    assert sto == 0 and rro <= 1e-10
    # the same outer steps; inner counts may differ by the FP32 rounding order of the
    # products (one iteration per inner solve)
    assert outer == outo, (outer, outo)
    assert abs(it - ito) <= outer, (it, ito)
    # both solutions are within the 1e-10 residual of the exact one
    xd = O.bicgstab(rp, ci, val, _oracle_precond(rp, ci, val, "fp64"), b, tol=1e-13, maxit=500)[0]
    scale = np.max(np.abs(xd))
    assert np.max(np.abs(x - xd)) <= 1e-8 * scale
    assert np.max(np.abs(xo - xd)) <= 1e-8 * scale
    h.close()


@pytest.mark.parametrize("prec", ["bf16", "fp32"])
@pytest.mark.parametrize("problem,L", [(1, 24), (2, 12), (0, 64)])
def test_mixed_gmres_solve_matches_oracle_scheme(gpu, prec, problem, L):
    """ISAI_SOLVER_FGMRES_MIXED_GMRES (inner FP32-storage GMRES cycles of at most 30
    steps) against the oracle's restatement (ora_fgmres_mixed_gmres): the same outer
    steps, inner steps within one per outer step (a cycle that stops on its tolerance
    may stop one step apart under a different FP32 rounding order), FP64 residual
    <= 1e-10 checked independently, and the solution of a tight FP64 solve."""
    print(f"mixed fgmres/gmres problem={problem} L={L} isai={prec}", flush=True)
    rp, ci, val = problems.generate(problem, L, seed=2)
    A = Matrix(rp, ci, val)
    b = problems.normal(7, A.n)
    h = Handle(A, precond="isai", precision=prec)
    h.set_inner(1e-5, 30)
    d_b, d_x = DeviceArray.from_numpy(b), DeviceArray.zeros((A.n,), np.float64)
    st, it, rr = h.solve(d_b, d_x, "fgmres_mixed_gmres", tol=1e-10, maxit=600, restart=10)
    outer, inner = h.last_stats()
    x = d_x.to_numpy()
    assert st == 0 and rr <= 1e-10 and inner == it
    assert O.rel_residual(rp, ci, val, b, x) <= 1e-10
    P = _oracle_precond(rp, ci, val, prec)
    xo, sto, ito, outo, rro = O.fgmres_mixed(rp, ci, val, P, b, m=10, tol=1e-10, maxit=600, inner_tol=1e-5,
                                             inner_maxit=30, inner="gmres")
    assert sto == 0 and rro <= 1e-10
    print(f"  outer {outer} / {outo}, inner steps {it} / {ito}", flush=True)
    assert outer == outo, (outer, outo)
    assert abs(it - ito) <= outer, (it, ito)
    xd = O.bicgstab(rp, ci, val, _oracle_precond(rp, ci, val, "fp64"), b, tol=1e-13, maxit=500)[0]
    scale = np.max(np.abs(xd))
    assert np.max(np.abs(x - xd)) <= 1e-8 * scale
    assert np.max(np.abs(xo - xd)) <= 1e-8 * scale
    h.close()


def test_mixed_inner_product_matches_oracle_rounding(gpu):
    """The inner solve's FP32 A product (A rounded to FP32, FP32 x, FP32 accumulation)
    against the oracle's rounded product, bounded by the FP32 accumulation error, and
    its fused FP64 dot partials against the FP64 dot of the FP32 results."""
    rp, ci, val = problems.generate(2, 10, seed=3)
    A = Matrix(rp, ci, val)
    n = A.n
    rng = np.random.default_rng(5)
    x = rng.standard_normal(n).astype(np.float32)
    a = rng.standard_normal(n).astype(np.float32)
    A32 = DeviceArray((len(ci) * 16,), np.float32)
    call("isai_convert_values", len(ci) * 16, 1, A.d_val.ptr, A32.ptr, None)
    d_x, d_a = DeviceArray.from_numpy(x), DeviceArray.from_numpy(a)
    d_y = DeviceArray((n,), np.float32)
    d_part = DeviceArray((2 * 4096,), np.float64)
    groups = ctypes.c_int()
    call("isai_bsr_spmv_f32_dot", A.nb, A.nnzb, A.d_rp.ptr, A.d_ci.ptr, A32.ptr, d_x.ptr, d_y.ptr, d_a.ptr, 2,
         d_part.ptr, ctypes.byref(groups), None)
    call("isai_device_synchronize")
    groups = groups.value
    assert groups == int(np.ceil(A.nb / 32))
    y = d_y.to_numpy().astype(np.float64)
    ref = O.spmv_f32acc(rp, ci, val, x.astype(np.float64))
    absA = np.abs(O.bsr_to_dense(rp, ci, val))
    bound = 2 * 27 * 4 * 2.0**-24 * (absA @ np.abs(x.astype(np.float64)))
    assert np.all(np.abs(y - ref) <= bound + 1e-30)
    part = d_part.to_numpy()
    assert abs(part[:groups].sum() - float(np.dot(y, a.astype(np.float64)))) <= 1e-12 * np.sum(np.abs(y * a))
    assert abs(part[groups:2 * groups].sum() - float(np.dot(y, y))) <= 1e-12 * np.dot(y, y)


@pytest.mark.parametrize("problem,L,rows", [(1, 24, 2), (2, 10, 1), (2, 9, 1), (0, 33, 2), (3, 21, 2)])
def test_handle_fp32_product_line_aligned_layouts(gpu, problem, L, rows):
    """The product the mixed solves run, on the handle's line-aligned FP32 copy of A
    (one row per lane group with every row padded to an even block count, or row
    pairs): y equal to the public product on the unpadded FP32 values, and the fused
    FP64 dots equal to the host dots of y up to FP64 rounding, for ndot = 0, 1, 2.
    Odd block-row counts (9^3, 33^2, 21^2) leave the last pair with one row."""
    print(f"handle fp32 product problem={problem} L={L}", flush=True)
    rp, ci, val = problems.generate(problem, L, seed=4)
    A = Matrix(rp, ci, val)
    n = A.n
    rng = np.random.default_rng(9)
    x = rng.standard_normal(n).astype(np.float32)
    a = rng.standard_normal(n).astype(np.float32)
    A32 = DeviceArray((len(ci) * 16,), np.float32)
    call("isai_convert_values", len(ci) * 16, 1, A.d_val.ptr, A32.ptr, None)
    d_x, d_a = DeviceArray.from_numpy(x), DeviceArray.from_numpy(a)
    d_y0 = DeviceArray((n,), np.float32)
    d_p0, d_part = DeviceArray((2 * 4096,), np.float64), DeviceArray((2 * 4096,), np.float64)
    g0, g, rpg = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    call("isai_bsr_spmv_f32_dot", A.nb, A.nnzb, A.d_rp.ptr, A.d_ci.ptr, A32.ptr, d_x.ptr, d_y0.ptr, d_a.ptr, 2,
         d_p0.ptr, ctypes.byref(g0), None)
    h = Handle(A, precond="isai", precision="bf16")
    y0 = None
    for ndot in (0, 1, 2):
        d_y = DeviceArray.from_numpy(np.full(n, np.nan, np.float32))  # no stale result can pass
        call("isai_handle_spmv_f32_dot", h.h, d_x.ptr, d_y.ptr, d_a.ptr, ndot, d_part.ptr, ctypes.byref(g),
             ctypes.byref(rpg), None)
        call("isai_device_synchronize")
        assert rpg.value == rows and g.value == g0.value
        y0 = d_y0.to_numpy() if y0 is None else y0
        y = d_y.to_numpy()
        assert np.array_equal(y, y0), ndot
        G = g.value
        yd = y.astype(np.float64)
        part = d_part.to_numpy()
        if ndot >= 1:
            ya = float(np.dot(yd, a.astype(np.float64)))
            assert abs(part[:G].sum() - ya) <= 1e-12 * np.sum(np.abs(yd * a))
        if ndot == 2:
            assert abs(part[G:2 * G].sum() - float(np.dot(yd, yd))) <= 1e-12 * np.dot(yd, yd)
    h.close()


def test_mixed_solve_full_size_96cube(gpu):
    """The north-star workload (3-D 7-pt 96^3, b = A*1) with BF16 ISAI: FP64 relative
    residual <= 1e-10 checked on the host, and the solution of all ones."""
    print("mixed fgmres 96^3 bf16", flush=True)
    rp, ci, val = problems.generate(1, 0, seed=0)
    A = Matrix(rp, ci, val)
    d_one = DeviceArray.from_numpy(np.ones(A.n))
    d_b, d_x = DeviceArray((A.n,), np.float64), DeviceArray.zeros((A.n,), np.float64)
    A.spmv(d_one, d_b)
    h = Handle(A, precond="isai", precision="bf16")
    st, it, rr = h.solve(d_b, d_x, "fgmres_mixed", tol=1e-10, maxit=400, restart=10)
    outer, inner = h.last_stats()
    print(f"  outer {outer} inner {inner} relres {rr:.3e}", flush=True)
    x = d_x.to_numpy()
    b = d_b.to_numpy()
    assert st == 0 and O.rel_residual(rp, ci, val, b, x) <= 1e-10
    assert np.max(np.abs(x - 1.0)) <= 1e-8
    h.close()


@pytest.mark.parametrize("form", [("fp32", "fgmres"), ("bf16", "fgmres"), ("bf16", "fgmres_mixed"),
                                  ("bf16", "fgmres_mixed_gmres")])
def test_precision_form_64cube_against_oracle_reference(gpu, form):
    """The precision form at its own size (3-D 27-pt convection-diffusion 64^3,
    FGMRES(30) to 1e-10, b = isai_gen_normal(1)) against the oracle's FP64
    FGMRES(30) solution (sampled entries and projections recorded by
    tests/golden/make_precision_reference.py).  Both residuals are <= 1e-10, so the
    solutions differ by at most about cond * 2e-10; the oracle's own FGMRES and
    BiCGSTAB solutions differ by 3.1e-8 relative here, and 1e-6 leaves a factor 30."""
    import json
    from pathlib import Path
    prec, solver = form
    print(f"precision form 64^3 {prec} {solver}", flush=True)
    ref = json.loads((Path(__file__).resolve().parent / "golden" / "precision_64cube_oracle.json").read_text())
    rp, ci, val = problems.generate(2, 64, seed=0)
    A = Matrix(rp, ci, val)
    assert A.n == ref["n"]
    d_b, d_x = DeviceArray.from_numpy(problems.normal(1, A.n)), DeviceArray.zeros((A.n,), np.float64)
    h = Handle(A, precond="isai", precision=prec)
    st, it, rr = h.solve(d_b, d_x, solver, tol=1e-10, maxit=3000, restart=30)
    x = d_x.to_numpy()
    assert st == 0 and rr <= 1e-10
    idx = np.asarray(ref["sample_index"])
    err = np.max(np.abs(x[idx] - np.asarray(ref["x_samples"]))) / ref["x_norm_inf"]
    rng = np.random.default_rng(ref["sample_seed"] + 1)
    proj = [float(np.dot(rng.standard_normal(A.n), x)) for _ in range(len(ref["projections"]))]
    perr = max(abs(a - b) / abs(b) for a, b in zip(proj, ref["projections"]))
    print(f"  {it} it, sample err {err:.2e}, projection err {perr:.2e}", flush=True)
    assert err <= 1e-6 and perr <= 1e-6
    h.close()
