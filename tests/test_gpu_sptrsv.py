"""GPU parity of the sync-free triangular solves (BSR-native and CSR/cape forms)."""
import ctypes

import numpy as np
import pytest

import _oracle as O
from gpu_pipeline import dev, gpu_ilu, rel_err, split_lu
from isai_kernel_amd import DeviceArray, call, lib, problems, ptr

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(77)
TOL = 1e-12


def _plan(rp, ci, d_rp, d_ci, d_lu, chunk):
    plan = ctypes.c_void_p()
    call("isai_bsr_sptrsv_analysis", len(rp) - 1, len(ci), 4, d_rp.ptr, d_ci.ptr, d_lu.ptr, chunk,
         ctypes.byref(plan), None)
    return plan


@pytest.mark.parametrize("problem,L", [(0, 16), (1, 8), (2, 5), (3, 12)])
@pytest.mark.parametrize("chunk", [1, 5, 64, 128, -1, -2, -3, -4])
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


def _schedule(plan):
    s = ctypes.c_int()
    call("isai_bsr_sptrsv_schedule", plan, ctypes.byref(s))
    return s.value


@pytest.mark.parametrize("problem,L,schedule", [(1, 16, -4), (1, 24, -4), (1, 20, -4), (1, 9, -4), (0, 64, -4),
                                                (0, 45, -4), (2, 6, -1), (3, 20, -1)])
def test_bsr_sptrsv_level_units(gpu, problem, L, schedule):
    """Level-synchronous units (chunk = -4) against the oracle: 7-point grids of
    16^3 and 24^3 (8 and 27 units of 8x8x8 rows), 20^3 and 9^3 (8,000 and 729 rows:
    a partial last unit, and U's order grown by its own walk since the row count is
    no multiple of 512), 2-D 5-point grids of 64^2 and 45^2 (two dependencies per
    row); the 27-point and 9-point stencils (more than three dependencies per row)
    fall back to the unit kernel.  Three solves in a row (ticket reset) and the
    schedule the plan reports."""
    print(f"bsr sptrsv level units problem={problem} L={L}", flush=True)
    rp, ci, val = problems.generate(problem, L, seed=5)
    d_rp, d_ci, d_lu = gpu_ilu(rp, ci, val)
    n = (len(rp) - 1) * 4
    b = RNG.standard_normal(n)
    plan = _plan(rp, ci, d_rp, d_ci, d_lu, -4)
    try:
        assert _schedule(plan) == schedule
        d_b, d_x = dev(b), DeviceArray((n,), np.float64)
        ref = O.sptrsv_lu(rp, ci, d_lu.to_numpy(), b)
        for rep in range(3):
            call("isai_bsr_sptrsv_solve", plan, d_lu.ptr, d_b.ptr, d_x.ptr, None)
            call("isai_sptrsv_status", plan, None)
            assert rel_err(d_x.to_numpy(), ref) <= TOL, rep
    finally:
        call("isai_bsr_sptrsv_destroy", plan)


@pytest.mark.parametrize("problem,L,schedule", [(1, 16, -4), (0, 64, -1), (2, 16, -1)])
def test_bsr_sptrsv_automatic_schedule(gpu, problem, L, schedule):
    """chunk = 0 picks the level-synchronous units on a 3-D 7-point grid (23 rows per
    local level), the unit kernel on a 2-D 5-point grid (512-row units of about 10
    rows per level) and on the 27-point stencil (13 dependencies per row); every
    choice solves to the oracle."""
    print(f"bsr sptrsv automatic schedule problem={problem} L={L}", flush=True)
    rp, ci, val = problems.generate(problem, L, seed=8)
    d_rp, d_ci, d_lu = gpu_ilu(rp, ci, val)
    n = (len(rp) - 1) * 4
    b = RNG.standard_normal(n)
    plan = _plan(rp, ci, d_rp, d_ci, d_lu, 0)
    try:
        assert _schedule(plan) == schedule
        d_b, d_x = dev(b), DeviceArray((n,), np.float64)
        call("isai_bsr_sptrsv_solve", plan, d_lu.ptr, d_b.ptr, d_x.ptr, None)
        call("isai_sptrsv_status", plan, None)
        assert rel_err(d_x.to_numpy(), O.sptrsv_lu(rp, ci, d_lu.to_numpy(), b)) <= TOL
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


@pytest.mark.parametrize("chunk", [128, -1, -2, -3, -4])
def test_bsr_sptrsv_full_size_residual(gpu, chunk):
    """96^3 7-point: || L (U x) - b || / || b || with the GPU factor and solve, for
    the chunk schedule (128), the unit kernel (-1) and the row kernel in level (-2)
    and unit (-3) order."""
    print(f"bsr sptrsv full size 96^3 chunk={chunk}", flush=True)
    rp, ci, val = problems.generate(1, 96, seed=0)
    d_rp, d_ci, d_lu = gpu_ilu(rp, ci, val)
    n = (len(rp) - 1) * 4
    b = RNG.standard_normal(n)
    plan = _plan(rp, ci, d_rp, d_ci, d_lu, chunk)
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


@pytest.mark.parametrize("chunk", [128, -1, -2, -4])
def test_spin_limit_timeout_is_reported_and_cleared(gpu, chunk):
    """A wait budget of one poll makes dependent chunks (or rows) give up (bounded
    waits: every kernel still exits); the plan reports ISAI_ERR_TIMEOUT once, then
    solves correctly again with the default budget."""
    print(f"bsr sptrsv forced timeout (spin limit 1), problem 1 L=48 chunk={chunk}", flush=True)
    rp, ci, val = problems.generate(1, 48, seed=6)
    d_rp, d_ci, d_lu = gpu_ilu(rp, ci, val)
    n = (len(rp) - 1) * 4
    b = RNG.standard_normal(n)
    plan = _plan(rp, ci, d_rp, d_ci, d_lu, chunk)
    try:
        d_b, d_x = dev(b), DeviceArray((n,), np.float64)
        call("isai_bsr_sptrsv_set_spin_limit", plan, 1)
        statuses = []
        for _ in range(3):
            call("isai_bsr_sptrsv_solve", plan, d_lu.ptr, d_b.ptr, d_x.ptr, None)
            statuses.append(lib.isai_sptrsv_status(plan, None))
            if statuses[-1] != 0:
                break
        assert statuses[-1] == -7, statuses
        assert lib.isai_sptrsv_status(plan, None) == 0  # cleared by the read
        call("isai_bsr_sptrsv_set_spin_limit", plan, 1 << 16)
        call("isai_bsr_sptrsv_solve", plan, d_lu.ptr, d_b.ptr, d_x.ptr, None)
        call("isai_sptrsv_status", plan, None)
        assert rel_err(d_x.to_numpy(), O.sptrsv_lu(rp, ci, d_lu.to_numpy(), b)) <= TOL
    finally:
        call("isai_bsr_sptrsv_destroy", plan)


def test_solver_reports_triangular_solve_timeout(gpu):
    """isai_solve with the ILU(0) preconditioner returns ISAI_ERR_TIMEOUT when a wait
    of its triangular solves ran out (forced from the host with a one-poll budget)."""
    print("solver ILU-SpTRSV forced timeout, problem 1 L=48", flush=True)
    from isai_kernel_amd.solver import Handle, Matrix
    rp, ci, val = problems.generate(1, 48, seed=6)
    A = Matrix(rp, ci, val)
    h = Handle(A, precond="ilu")
    try:
        call("isai_handle_set_spin_limit", h.h, 1)
        d_b = dev(RNG.standard_normal(A.n))
        st = None
        for _ in range(3):
            d_x = DeviceArray.zeros((A.n,), np.float64)
            st, it, rr = h.solve(d_b, d_x, "bicgstab", tol=1e-10, maxit=5, check=False)
            if st == -7:
                break
        assert st == -7
        assert lib.isai_handle_status(h.h, None) == 0
        call("isai_handle_set_spin_limit", h.h, 1 << 16)
        d_x = DeviceArray.zeros((A.n,), np.float64)
        st, it, rr = h.solve(d_b, d_x, "bicgstab", tol=1e-10, maxit=500)
        assert st == 0 and rr <= 1e-10
    finally:
        h.close()


def _random_wide_pattern(nb, per_row, seed, symmetric=True):
    """Structurally symmetric block pattern with about `per_row` random partners per block row on each
    side of the diagonal, diagonally dominant 4x4 blocks.  Rows depend on rows all over the matrix, so
    a 64-row unit of the unit order imports more than 512 rows (up to about 1,000 here)."""
    rng = np.random.default_rng(seed)
    cols = [{i} for i in range(nb)]
    for i in range(nb):
        for j in rng.choice(nb, per_row, replace=False):
            cols[i].add(int(j))
            if symmetric:
                cols[int(j)].add(i)
    rp = np.zeros(nb + 1, np.int32)
    ci = []
    for i in range(nb):
        c = sorted(cols[i])
        ci += c
        rp[i + 1] = rp[i] + len(c)
    ci = np.array(ci, np.int32)
    val = rng.standard_normal((len(ci), 4, 4)) * 0.01
    for i in range(nb):
        val[rp[i] + np.searchsorted(ci[rp[i]:rp[i + 1]], i)] += 4.0 * np.eye(4)
    return rp, ci, val


@pytest.mark.parametrize("chunk", [0, -1, -2])
def test_bsr_sptrsv_unit_import_cap_fallback(gpu, chunk):
    """Unit order on a pattern whose units would import more than 512 rows: the plan falls back to the
    row kernel in unit order, with the same results as the level order and the oracle."""
    print(f"bsr sptrsv random wide pattern nb=2048 chunk={chunk}", flush=True)
    rp, ci, val = _random_wide_pattern(2048, 12, seed=9)
    d_rp, d_ci, d_lu = gpu_ilu(rp, ci, val)
    n = (len(rp) - 1) * 4
    b = RNG.standard_normal(n)
    plan = _plan(rp, ci, d_rp, d_ci, d_lu, chunk)
    try:
        d_b, d_x = dev(b), DeviceArray((n,), np.float64)
        for _ in range(2):
            call("isai_bsr_sptrsv_solve", plan, d_lu.ptr, d_b.ptr, d_x.ptr, None)
            call("isai_sptrsv_status", plan, None)
            assert rel_err(d_x.to_numpy(), O.sptrsv_lu(rp, ci, d_lu.to_numpy(), b)) <= TOL
    finally:
        call("isai_bsr_sptrsv_destroy", plan)


@pytest.mark.parametrize("chunk", [0, -1])
def test_bsr_sptrsv_unit_many_imports(gpu, chunk):
    """Unit schedule on a random pattern whose units import up to ~440 rows (below the 512 cap): every
    item group of the import wave is used, with the 8-dependency batch, against the oracle."""
    print(f"bsr sptrsv random pattern nb=4096 (many imports) chunk={chunk}", flush=True)
    rp, ci, val = _random_wide_pattern(4096, 4, seed=11)
    d_rp, d_ci, d_lu = gpu_ilu(rp, ci, val)
    n = (len(rp) - 1) * 4
    b = RNG.standard_normal(n)
    plan = _plan(rp, ci, d_rp, d_ci, d_lu, chunk)
    try:
        d_b, d_x = dev(b), DeviceArray((n,), np.float64)
        ref = O.sptrsv_lu(rp, ci, d_lu.to_numpy(), b)
        for _ in range(3):
            call("isai_bsr_sptrsv_solve", plan, d_lu.ptr, d_b.ptr, d_x.ptr, None)
            call("isai_sptrsv_status", plan, None)
            assert rel_err(d_x.to_numpy(), ref) <= TOL
    finally:
        call("isai_bsr_sptrsv_destroy", plan)


@pytest.mark.parametrize("chunk", [-1, -3])
def test_bsr_sptrsv_unit_nonsymmetric_pattern(gpu, chunk):
    """Unit order on a structurally non-symmetric pattern: U's order is grown by its own greedy walk
    (a symmetric pattern reuses L's order reversed), against the oracle."""
    print(f"bsr sptrsv random non-symmetric pattern nb=4096 chunk={chunk}", flush=True)
    rp, ci, val = _random_wide_pattern(4096, 3, seed=13, symmetric=False)
    d_rp, d_ci, d_lu = gpu_ilu(rp, ci, val)
    n = (len(rp) - 1) * 4
    b = RNG.standard_normal(n)
    plan = _plan(rp, ci, d_rp, d_ci, d_lu, chunk)
    try:
        d_b, d_x = dev(b), DeviceArray((n,), np.float64)
        ref = O.sptrsv_lu(rp, ci, d_lu.to_numpy(), b)
        for _ in range(2):
            call("isai_bsr_sptrsv_solve", plan, d_lu.ptr, d_b.ptr, d_x.ptr, None)
            call("isai_sptrsv_status", plan, None)
            assert rel_err(d_x.to_numpy(), ref) <= TOL
    finally:
        call("isai_bsr_sptrsv_destroy", plan)
