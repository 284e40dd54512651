"""Pair-list statistics on the GPU (the analyze_pair.h figures) against the oracle,
and the ISAI build plan they drive (LDS Y-cache size, column order), which must
leave the factor values bit-identical."""
import ctypes

import numpy as np
import pytest

import _oracle as O
from isai_kernel_amd import DeviceArray, call, problems
from isai_kernel_amd.solver import Handle, Matrix, pair_column_stats, pair_stats

pytestmark = pytest.mark.gpu


def _device_pairs(rp, ci, upper):
    """BCSC pattern on the host, pair pointers by isai_pairs_count on the device."""
    mb, mc = problems.bcsc_pattern(rp, ci, not upper)
    d_rp, d_ci = DeviceArray.from_numpy(rp), DeviceArray.from_numpy(ci)
    d_mb, d_mc = DeviceArray.from_numpy(mb), DeviceArray.from_numpy(mc)
    d_pp = DeviceArray((len(mc) + 1,), np.int32)
    total = ctypes.c_int()
    call("isai_pairs_count", len(rp) - 1, len(mc), d_rp.ptr, d_ci.ptr, d_mb.ptr, d_mc.ptr, int(upper), d_pp.ptr,
         ctypes.byref(total), None)
    return mb, mc, d_mb, d_pp, total.value, (d_rp, d_ci, d_mc)


def _arrow(nb, arms, seed):
    """Block-tridiagonal matrix whose every block row also couples to the first
    `arms` block columns, strictly block-diagonally dominant.  ISAI(L) gets `arms`
    columns of length ~nb and a skewed pair-count distribution, while every block
    row stays short (the ILU(0) kernel takes at most 64 blocks per row)."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(nb):
        cols = {i, max(i - 1, 0), min(i + 1, nb - 1)} | set(range(arms))
        rows.append(sorted(cols))
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    ci = np.concatenate(rows).astype(np.int32)
    val = rng.uniform(-0.05, 0.05, (len(ci), 4, 4))
    for i in range(nb):
        k = rp[i] + rows[i].index(i)
        off = np.abs(val[rp[i]:rp[i + 1]]).sum(axis=(0, 2)) - np.abs(val[k]).sum(axis=1)
        val[k] += np.diag(off + 1.0)
    return rp, ci, val


CASES = [(0, 12, False), (1, 7, False), (1, 7, True), (2, 5, False), (2, 5, True), (4, 9, True)]


@pytest.mark.parametrize("problem,L,upper", CASES)
def test_pair_stats_match_oracle(gpu, problem, L, upper):
    print(f"pair stats problem={problem} L={L} upper={upper}", flush=True)
    rp, ci, _ = problems.generate(problem, L, seed=1)
    mb, mc, d_mb, d_pp, total, _keep = _device_pairs(rp, ci, upper)
    got = pair_stats(d_mb, d_pp, len(rp) - 1)
    pp_host = d_pp.to_numpy()
    # the device pair pointers are the oracle's (brute-force pinned) pair lists
    assert np.array_equal(pp_host, O.pairs(rp, ci, mb, mc, upper=upper)[0])
    exp, cn, cq, cm = O.pair_stats(mb, pp_host)
    assert got["total_pairs"] == total
    for k, v in exp.items():
        assert got[k] == v, (k, got[k], v)
    N, Q, MQ = pair_column_stats(d_mb, d_pp, len(rp) - 1)
    assert np.array_equal(N, cn) and np.array_equal(Q, cq) and np.array_equal(MQ, cm)


def test_pair_stats_skewed_pattern(gpu):
    """Columns of length up to nb next to columns of length 2, and a Q distribution
    whose maximum is ~120 times its mean."""
    print("pair stats arrow nb=300 arms=3", flush=True)
    rp, ci, _ = _arrow(300, 3, seed=4)
    for upper in (False, True):
        mb, mc, d_mb, d_pp, total, _keep = _device_pairs(rp, ci, upper)
        got = pair_stats(d_mb, d_pp, len(rp) - 1)
        exp, *_ = O.pair_stats(mb, d_pp.to_numpy())
        for k, v in exp.items():
            assert got[k] == v, (upper, k, got[k], v)


def test_build_plan_and_bitexact_factors(gpu):
    """Regular stencil: every column fits the LDS cache, natural order.  Arrow: the
    cache covers 99 % of the columns (the arms read Y from global memory) and ISAI(L)
    columns are scheduled by descending pair count.  Both give the oracle's factors
    bit for bit."""
    print("build plan: 7-pt 9^3 and arrow nb=400", flush=True)
    rp, ci, val = problems.generate(1, 9, seed=2)
    h = Handle(Matrix(rp, ci, val))
    assert h.build_plan() == {"ncache": 4, "ordered_L": 0, "ordered_U": 0}
    h.close()

    rp, ci, val = _arrow(400, 3, seed=5)
    lu, L, U = O.isai_factors(rp, ci, val)
    h = Handle(Matrix(rp, ci, val))
    plan = h.build_plan()
    mbL, _ = O.bcsc_pattern(rp, ci, True)
    nL = np.diff(mbL)
    assert nL.max() > 31 and 0 < plan["ncache"] <= 31 and (nL > plan["ncache"]).sum() == 3, (plan, nL.max())
    assert plan["ordered_L"] == 1 and plan["ordered_U"] == 0, plan
    Lrow, Lcol, Lval, Urow, Ucol, Uval, LU = h.factors()
    got = {}
    for name, p, ref in (("LU", LU, lu), ("L", Lval, L[2]), ("U", Uval, U[2])):
        out = np.empty(ref.shape)
        call("isai_memcpy_dtoh", out.ctypes.data, p, out.nbytes)
        got[name] = out
        assert np.array_equal(out, ref), name
    # and the preconditioned solve works on it
    A = h.A
    b = O.spmv(rp, ci, val, np.ones(A.n))
    d_b, d_x = DeviceArray.from_numpy(b), DeviceArray.zeros((A.n,), np.float64)
    st, it, rr = h.solve(d_b, d_x, "bicgstab", tol=1e-10, maxit=200)
    assert st == 0 and O.rel_residual(rp, ci, val, b, d_x.to_numpy()) <= 1e-10
    h.close()
