"""GPU parity of the batching form (isai_batch.h): block-diagonal systems with
1..16 right-hand sides each, batched setup, batched BiCGSTAB replayed as a graph."""
import ctypes

import numpy as np
import pytest

import _oracle as O
from gpu_pipeline import rel_err
from isai_kernel_amd import DeviceArray, IsaiError, call, lib, problems
from isai_kernel_amd.solver import Batch, Matrix

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stream(gpu):
    s = ctypes.c_void_p()
    call("isai_stream_create", ctypes.byref(s))
    yield s
    call("isai_stream_destroy", s)


def _system(rp, ci, val, sys_ptr, s):
    r0, r1 = int(sys_ptr[s]), int(sys_ptr[s + 1])
    k0, k1 = int(rp[r0]), int(rp[r1])
    return (rp[r0:r1 + 1] - k0).astype(np.int32), (ci[k0:k1] - r0).astype(np.int32), val[k0:k1]


def _rhs(sys_ptr, nrhs, seed):
    rng = np.random.default_rng(seed)
    return [[rng.standard_normal(4 * int(sys_ptr[s + 1] - sys_ptr[s])) for _ in range(nrhs[s])]
            for s in range(len(nrhs))]


def _make(nsys, lmin, lmax, seed, rhs_max, precision, stream):
    rp, ci, val, sys_ptr, Ls, nrhs = problems.batch(nsys, lmin, lmax, seed=seed, rhs_max=rhs_max)
    A = Matrix(rp, ci, val)
    return (rp, ci, val, sys_ptr, nrhs), Batch(A, sys_ptr, nrhs, precision=precision, stream=stream)


def test_batch_layout_and_spmm(stream):
    print("batch spmm 7 systems L 4..13", flush=True)
    (rp, ci, val, sys_ptr, nrhs), bt = _make(7, 4, 13, 5, 16, "fp64", stream)
    try:
        assert bt.npairs == int(nrhs.sum())
        # segments: 1024-aligned, in (system, rhs) order, non-overlapping
        offs = [bt.offset(s, j) for s in range(len(nrhs)) for j in range(nrhs[s])]
        assert all(o % 1024 == 0 for o in offs) and offs == sorted(offs) and bt.length % 1024 == 0
        X = _rhs(sys_ptr, nrhs, 1)
        d_X = DeviceArray.from_numpy(bt.pack(X))
        d_Y = DeviceArray.from_numpy(np.full(bt.length, 0.0))
        bt.spmm(d_X, d_Y)
        call("isai_stream_synchronize", stream)
        Y = d_Y.to_numpy()
        got = bt.unpack(Y)
        for s in range(len(nrhs)):
            srp, sci, sval = _system(rp, ci, val, sys_ptr, s)
            for j in range(nrhs[s]):
                assert rel_err(got[s][j], O.spmv(srp, sci, sval, X[s][j])) <= 1e-14, (s, j)
        # padding untouched
        mask = np.ones(bt.length, bool)
        for s in range(len(nrhs)):
            for j in range(nrhs[s]):
                o = bt.offset(s, j)
                mask[o:o + 4 * int(sys_ptr[s + 1] - sys_ptr[s])] = False
        assert not Y[mask].any()
        v, us, nc = bt.tuning()
        assert nc == 48 and 0 <= v < 24 and us > 0
        tab = bt.tuning_table(0)
        assert len(tab) == 24 and (tab > 0).all() and tab[v] == pytest.approx(us) and tab.min() == pytest.approx(us)
    finally:
        bt.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
def test_batch_solve_matches_oracle(stream, precision):
    """Every pair reaches 1e-10 true FP64 residual; per-pair iteration counts and
    solutions agree with the oracle's single-system BiCGSTAB on the same system."""
    print(f"batch solve 6 systems L 5..16 precision={precision}", flush=True)
    (rp, ci, val, sys_ptr, nrhs), bt = _make(6, 5, 16, 11, 4, precision, stream)
    try:
        B = _rhs(sys_ptr, nrhs, 2)
        d_B = DeviceArray.from_numpy(bt.pack(B))
        d_X = DeviceArray.zeros((bt.length,), np.float64)
        st, iters, rr = bt.solve(d_B, d_X, tol=1e-10, maxit=500)
        assert st == 0 and (rr <= 1e-10).all()
        X = bt.unpack(d_X.to_numpy())
        q = 0
        for s in range(len(nrhs)):
            srp, sci, sval = _system(rp, ci, val, sys_ptr, s)
            lu, L, U = O.isai_factors(srp, sci, sval)
            P = O.make_precond(1, len(srp) - 1, 4, L=L, U=U, f32=precision == "fp32", bf16=precision == "bf16")
            for j in range(nrhs[s]):
                assert O.rel_residual(srp, sci, sval, B[s][j], X[s][j]) <= 1e-10, (s, j)
                xo, sto, ito, _ = O.bicgstab(srp, sci, sval, P, B[s][j], tol=1e-10, maxit=500)
                assert sto == 0
                assert abs(int(iters[q]) - ito) <= 2, (s, j, iters[q], ito)
                assert rel_err(X[s][j], xo) <= 1e-8, (s, j)
                q += 1
    finally:
        bt.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64", "bf16"])
def test_batch_graph_replay_equals_eager(stream, precision):
    """Every schedule gives bit-identical iterates: the per-step launches eagerly and
    replayed as a captured graph (host reads and synchronisations between replays),
    and the persistent kernel launched directly and replayed as a captured graph
    (twice, so the second replay reuses the captured launch)."""
    print(f"batch schedules 9 systems L 6..24 {precision}", flush=True)
    (rp, ci, val, sys_ptr, nrhs), bt = _make(9, 6, 24, 3, 16, precision, stream)
    try:
        # the persistent kernel holds the row-pair, MFMA and whole-row lane mappings
        # or forms equal to them bit for bit (a column-split choice runs row-pair
        # there): keep the fastest of those
        for op in (0, 1):
            tab = bt.tuning_table(op)
            bt.select(op, min((c for c in range(len(tab)) if c // 6 != 1), key=lambda c: tab[c]))
        B = DeviceArray.from_numpy(bt.pack(_rhs(sys_ptr, nrhs, 9)))
        xs, its = [], []
        for mode in (1, 0, 1, 3, 2, 2):
            d_X = DeviceArray.zeros((bt.length,), np.float64)
            st, iters, rr = bt.solve(B, d_X, tol=1e-10, maxit=500, use_graph=mode)
            assert st == 0 and (rr <= 1e-10).all(), mode
            xs.append(d_X.to_numpy())
            its.append(iters)
        for k in range(1, len(xs)):
            assert np.array_equal(xs[0], xs[k]) and np.array_equal(its[0], its[k]), k
        # restarting from the solution: every pair stops before its first iteration
        d_X = DeviceArray.from_numpy(xs[0])
        st, iters, rr = bt.solve(B, d_X, tol=1e-10, maxit=500)
        assert st == 0 and not iters.any() and np.array_equal(d_X.to_numpy(), xs[0])
    finally:
        bt.close()


def test_batch_persistent_barrier_timeout_is_reported(stream):
    """A grid-barrier wait beyond the bound (1 us here, set from the host) makes every
    workgroup of the persistent kernel leave and the solve return ISAI_ERR_TIMEOUT;
    with the bound restored the same batch object solves again."""
    print("batch persistent timeout 6 systems L 20..40", flush=True)
    (rp, ci, val, sys_ptr, nrhs), bt = _make(6, 20, 40, 8, 16, "fp32", stream)
    try:
        B = DeviceArray.from_numpy(bt.pack(_rhs(sys_ptr, nrhs, 2)))
        d_X = DeviceArray.zeros((bt.length,), np.float64)
        call("isai_batch_set_spin_limit", bt.h, 1)
        st, iters, rr = bt.solve(B, d_X, tol=1e-10, maxit=500, use_graph=3, check=False)
        assert st == -7, st  # ISAI_ERR_TIMEOUT
        call("isai_batch_set_spin_limit", bt.h, 5_000_000)
        d_X = DeviceArray.zeros((bt.length,), np.float64)
        st, iters, rr = bt.solve(B, d_X, tol=1e-10, maxit=500, use_graph=2)
        assert st == 0 and (rr <= 1e-10).all()
        with pytest.raises(IsaiError):
            bt.solve(B, d_X, tol=1e-10, maxit=500, use_graph=4)
    finally:
        bt.close()


def test_batch_iteration_limit_and_errors(stream):
    print("batch limits and argument errors", flush=True)
    (rp, ci, val, sys_ptr, nrhs), bt = _make(4, 8, 14, 4, 3, "fp64", stream)
    try:
        d_B = DeviceArray.from_numpy(bt.pack(_rhs(sys_ptr, nrhs, 3)))
        d_X = DeviceArray.zeros((bt.length,), np.float64)
        st, iters, rr = bt.solve(d_B, d_X, tol=1e-10, maxit=3, check=False)
        assert st == -6 and iters.max() <= 3 and (rr > 1e-10).any()
        d_X2 = DeviceArray.zeros((bt.length,), np.float64)  # the persistent kernel stops the same way
        st2, iters2, rr2 = bt.solve(d_B, d_X2, tol=1e-10, maxit=3, use_graph=2, check=False)
        assert st2 == -6 and np.array_equal(iters2, iters) and np.array_equal(d_X2.to_numpy(), d_X.to_numpy())
        with pytest.raises(IsaiError):
            bt.offset(0, int(nrhs[0]))
    finally:
        bt.close()
    A = Matrix(rp, ci, val)
    with pytest.raises(IsaiError):  # 17 right-hand sides
        Batch(A, sys_ptr, np.full(len(nrhs), 17), stream=stream)
    # a block that couples two systems
    merged = np.array([0, sys_ptr[2], sys_ptr[4]], np.int32)
    bad = np.array([0, sys_ptr[1], sys_ptr[4]], np.int32)
    Batch(A, merged, np.array([1, 2]), stream=stream).close()  # merging systems is still block-diagonal
    call("isai_device_synchronize")
    ci2 = ci.copy()
    ci2[int(rp[sys_ptr[1]]) - 1] = sys_ptr[1]  # last block of system 0 now points into system 1
    A2 = Matrix(rp, ci2, val)
    h = ctypes.c_void_p()
    nr = np.ones(2, np.int32)
    st = lib.isai_batch_setup(2, bad.ctypes.data, nr.ctypes.data, len(ci2), A2.d_rp.ptr, A2.d_ci.ptr,
                              A2.d_val.ptr, 0, ctypes.byref(h), stream)
    assert st == -4 and not h.value


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
def test_batch_every_spmm_candidate(stream, precision):
    """Each lane mapping / workgroup / remap candidate: the A product against the
    oracle, and a solve whose ISAI products use that candidate."""
    print(f"batch spmm candidates {precision}", flush=True)
    (rp, ci, val, sys_ptr, nrhs), bt = _make(5, 3, 12, 21, 16, precision, stream)
    try:
        X = _rhs(sys_ptr, nrhs, 4)
        d_X = DeviceArray.from_numpy(bt.pack(X))
        d_Y = DeviceArray.zeros((bt.length,), np.float64)
        B = DeviceArray.from_numpy(bt.pack(_rhs(sys_ptr, nrhs, 5)))
        refs = [[O.spmv(*_system(rp, ci, val, sys_ptr, s), X[s][j]) for j in range(nrhs[s])] for s in range(len(nrhs))]
        its = []
        row_pair = None
        for c in range(24):
            bt.select(0, c)
            bt.select(1, c)
            bt.spmm(d_X, d_Y)
            call("isai_stream_synchronize", stream)
            y = d_Y.to_numpy()
            got = bt.unpack(y)
            for s in range(len(nrhs)):
                for j in range(nrhs[s]):
                    assert rel_err(got[s][j], refs[s][j]) <= 1e-14, (c, s, j)
            # the whole-row mapping (3) accumulates in the row-pair order
            if c == 0:
                row_pair = y.copy()
            elif c // 6 >= 3:
                assert np.array_equal(y, row_pair), c
            d_x = DeviceArray.zeros((bt.length,), np.float64)
            st, iters, rr = bt.solve(B, d_x, tol=1e-10, maxit=500)
            assert st == 0 and (rr <= 1e-10).all(), c
            its.append(iters)
            if c // 6 != 1:  # the persistent kernel holds the row-pair, MFMA and whole-row mappings
                d_p = DeviceArray.zeros((bt.length,), np.float64)
                st, it_p, rr = bt.solve(B, d_p, tol=1e-10, maxit=500, use_graph=3)
                assert st == 0 and np.array_equal(it_p, iters) and np.array_equal(d_p.to_numpy(), d_x.to_numpy()), c
        # FP64 factors: every candidate's products agree to rounding, so the
        # iteration counts agree within one
        if precision == "fp64":
            for it in its[1:]:
                assert np.abs(it.astype(int) - its[0].astype(int)).max() <= 1
    finally:
        bt.close()


@pytest.mark.parametrize("problem", [1, 2])
def test_batch_whole_row_product_other_stencils(stream, problem):
    """The whole-row mapping on 3-D 7-pt systems (rows of <= 7 blocks, all held at
    once) and 27-pt systems (rows of 27 blocks, longer than the 16 a lane group holds:
    the block-by-block path): equal to the row-pair product bit for bit, and to the
    oracle's per-pair SpMV."""
    print(f"batch whole-row product problem {problem}", flush=True)
    rp, ci, val, sys_ptr, Ls, nrhs = problems.batch(3, 3, 6, seed=11, rhs_max=16, problem=problem)
    bt = Batch(Matrix(rp, ci, val), sys_ptr, nrhs, precision="fp64", stream=stream)
    try:
        X = _rhs(sys_ptr, nrhs, 8)
        d_X = DeviceArray.from_numpy(bt.pack(X))
        d_Y = DeviceArray.zeros((bt.length,), np.float64)
        refs = [[O.spmv(*_system(rp, ci, val, sys_ptr, s), X[s][j]) for j in range(nrhs[s])] for s in range(len(nrhs))]
        row_pair = None
        for c in [0] + list(range(18, 24)):
            bt.select(0, c)
            bt.spmm(d_X, d_Y)
            call("isai_stream_synchronize", stream)
            y = d_Y.to_numpy()
            if c == 0:
                row_pair = y.copy()
                got = bt.unpack(y)
                for s in range(len(nrhs)):
                    for j in range(nrhs[s]):
                        assert rel_err(got[s][j], refs[s][j]) <= 1e-14, (s, j)
            else:
                assert np.array_equal(y, row_pair), c
    finally:
        bt.close()
    # FP32 factors: the whole-row MFMA factor products (candidate 18; rows longer than
    # 8 blocks fall back per wave) give the MFMA ones' iterates (candidate 12) bit for
    # bit, with the same A product (candidate 18) in both
    bt = Batch(Matrix(rp, ci, val), sys_ptr, nrhs, precision="fp32", stream=stream)
    try:
        B = DeviceArray.from_numpy(bt.pack(_rhs(sys_ptr, nrhs, 9)))
        xs = []
        for c in (12, 18):
            bt.select(0, 18)
            bt.select(1, c)
            d_x = DeviceArray.zeros((bt.length,), np.float64)
            st, iters, rr = bt.solve(B, d_x, tol=1e-10, maxit=500)
            assert st == 0 and (rr <= 1e-10).all(), c
            xs.append((iters, d_x.to_numpy()))
        assert np.array_equal(xs[0][0], xs[1][0]) and np.array_equal(xs[0][1], xs[1][1])
    finally:
        bt.close()


def test_batch_full_size_config(stream):
    """The batching form at its own size: 256 systems, 2-D 9-pt L x L with L in
    [24, 200] (4.1M block rows, 36.9M blocks), 1..16 right-hand sides N(0,1) each,
    FP32 ISAI, graph-replayed BiCGSTAB to 1e-10.  Checked with the oracle: the FP64
    residual of 16 pairs spanning L and the right-hand-side count, and the oracle's
    single-system BiCGSTAB on three systems (smallest, median and largest L)."""
    print("batch full size 256 systems L 24..200", flush=True)
    rp, ci, val, sys_ptr, Ls, nrhs = problems.batch(256, 24, 200, seed=0, rhs_max=16)
    A = Matrix(rp, ci, val)
    bt = Batch(A, sys_ptr, nrhs, precision="fp32", stream=stream)
    try:
        B = _rhs(sys_ptr, nrhs, 1)
        d_B = DeviceArray.from_numpy(bt.pack(B))
        d_X = DeviceArray.zeros((bt.length,), np.float64)
        st, iters, rr = bt.solve(d_B, d_X, tol=1e-10, maxit=1000, use_graph=True)
        assert st == 0 and (rr <= 1e-10).all() and len(rr) == int(nrhs.sum())
        X = bt.unpack(d_X.to_numpy())
        first = np.concatenate([[0], np.cumsum(nrhs)[:-1]])  # pair index of (s, 0)
        order = np.argsort(Ls, kind="stable")
        picks = [int(order[k]) for k in np.linspace(0, len(order) - 1, 8).round().astype(int)]
        checked = 0
        for s in picks:
            srp, sci, sval = _system(rp, ci, val, sys_ptr, s)
            for j in sorted({0, int(nrhs[s]) - 1}):
                assert O.rel_residual(srp, sci, sval, B[s][j], X[s][j]) <= 1e-10, (s, j)
                checked += 1
        assert checked >= 9
        # sixteen pairs at least: the first right-hand side of eight more systems
        for s in [int(order[k]) for k in np.linspace(1, len(order) - 2, 8).round().astype(int)]:
            srp, sci, sval = _system(rp, ci, val, sys_ptr, s)
            assert O.rel_residual(srp, sci, sval, B[s][0], X[s][0]) <= 1e-10, s
            checked += 1
        assert checked >= 16
        for s in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
            srp, sci, sval = _system(rp, ci, val, sys_ptr, s)
            lu, L, U = O.isai_factors(srp, sci, sval)
            P = O.make_precond(1, len(srp) - 1, 4, L=L, U=U, f32=True)
            xo, sto, ito, _ = O.bicgstab(srp, sci, sval, P, B[s][0], tol=1e-10, maxit=1000)
            print(f"  system {s} L={Ls[s]}: batch {iters[first[s]]} it, oracle {ito} it", flush=True)
            assert sto == 0 and rel_err(X[s][0], xo) <= 1e-8
            # BiCGSTAB's count moves with the rounding of the preconditioner: on these
            # three systems the oracle itself takes 20-27 iterations over FP64 / FP32 /
            # BF16 factors and three right-hand sides (e.g. 22 / 23 / 20 for the L = 24
            # system), so the batch (another FP32 summation order) is held to 20 %
            assert abs(int(iters[first[s]]) - ito) <= max(2, 0.2 * ito)
    finally:
        bt.close()
