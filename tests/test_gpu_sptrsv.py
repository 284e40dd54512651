"""GPU parity of the sync-free triangular solves (BSR-native and CSR/cape forms)."""
import ctypes

import numpy as np
import pytest

import _oracle as O
from gpu_pipeline import dev, gpu_ilu, rel_err, split_lu
from isai_kernel_amd import DeviceArray, call, problems, ptr

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(77)
TOL = 1e-12


def _plan(rp, ci, d_rp, d_ci, d_lu, chunk):
    plan = ctypes.c_void_p()
    call("isai_bsr_sptrsv_analysis", len(rp) - 1, len(ci), 4, d_rp.ptr, d_ci.ptr, d_lu.ptr, chunk,
         ctypes.byref(plan), None)
    return plan


@pytest.mark.parametrize("problem,L", [(0, 16), (1, 8), (2, 5), (3, 12)])
@pytest.mark.parametrize("chunk", [1, 5, 64, 128])
def test_bsr_sptrsv_matches_oracle(gpu, problem, L, chunk):
    print(f"bsr sptrsv problem={problem} L={L} chunk={chunk}", flush=True)
    rp, ci, val = problems.generate(problem, L, seed=2)
    d_rp, d_ci, d_lu = gpu_ilu(rp, ci, val)
    lu = d_lu.to_numpy()
    n = (len(rp) - 1) * 4
    b = RNG.standard_normal(n)
    plan = _plan(rp, ci, d_rp, d_ci, d_lu, chunk)
    try:
        d_b = dev(b)
        d_x = DeviceArray((n,), np.float64)
        ref = O.sptrsv_lu(rp, ci, lu, b)
        for rep in range(3):  # counters reset between solves
            call("isai_bsr_sptrsv_solve", plan, d_lu.ptr, d_b.ptr, d_x.ptr, None)
            call("isai_sptrsv_status", plan, None)
            x = d_x.to_numpy()
            assert rel_err(x, ref) <= TOL, rep
        # in place: x overwrites b
        call("isai_bsr_sptrsv_solve", plan, d_lu.ptr, d_b.ptr, d_b.ptr, None)
        call("isai_sptrsv_status", plan, None)
        assert np.array_equal(d_b.to_numpy(), x)
    finally:
        call("isai_bsr_sptrsv_destroy", plan)


def test_csr_sptrsv_host_form_matches_oracle(gpu):
    print("csr sptrsv host form", flush=True)
    rp, ci, val = problems.generate(0, 12, seed=3)
    lu, _, _ = O.bsrilu0(rp, ci, val)
    crp, cci, cv = O.bsr2csr(rp, ci, lu)
    m = len(crp) - 1
    b = RNG.standard_normal(m)
    st, ref = O.csr_sptrsv_lu(crp, cci, cv, b)
    assert st == 0
    bb = b.copy()
    call("isai_csr_sptrsv_lu", m, ptr(crp), ptr(cci), ptr(cv), len(cci), ptr(bb))
    assert rel_err(bb, ref) <= TOL


@pytest.mark.parametrize("chunk", [1, 16, 1024])
def test_csr_sptrsv_device_form(gpu, chunk):
    print(f"csr sptrsv device chunk={chunk}", flush=True)
    rp, ci, val = problems.generate(1, 6, seed=4)
    lu, _, _ = O.bsrilu0(rp, ci, val)
    crp, cci, cv = O.bsr2csr(rp, ci, lu)
    m = len(crp) - 1
    b = RNG.standard_normal(m)
    _, ref = O.csr_sptrsv_lu(crp, cci, cv, b)
    ws = ctypes.c_size_t()
    call("isai_csr_sptrsv_work_size", m, ctypes.byref(ws))
    d = [dev(a) for a in (crp, cci, cv, b)]
    d_x = DeviceArray((m,), np.float64)
    d_w = DeviceArray((ws.value,), np.uint8)
    call("isai_csr_sptrsv_lu_device", m, len(cci), d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, d_x.ptr, d_w.ptr,
         chunk, 0, None)
    assert rel_err(d_x.to_numpy(), ref) <= TOL


def test_bsr_sptrsv_full_size_residual(gpu):
    """96^3 7-point: || L (U x) - b || / || b || with the GPU factor and solve."""
    print("bsr sptrsv full size 96^3", flush=True)
    rp, ci, val = problems.generate(1, 96, seed=0)
    d_rp, d_ci, d_lu = gpu_ilu(rp, ci, val)
    n = (len(rp) - 1) * 4
    b = RNG.standard_normal(n)
    plan = _plan(rp, ci, d_rp, d_ci, d_lu, 0)
    try:
        d_b, d_x = dev(b), DeviceArray((n,), np.float64)
        call("isai_bsr_sptrsv_solve", plan, d_lu.ptr, d_b.ptr, d_x.ptr, None)
        call("isai_sptrsv_status", plan, None)
        x = d_x.to_numpy()
    finally:
        call("isai_bsr_sptrsv_destroy", plan)
    Ls, Us = split_lu(rp, ci, d_lu.to_numpy())
    r = Ls @ (Us @ x) - b
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-13
