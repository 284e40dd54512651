"""GPU parity of the BSR SpMV family against the oracle (through the C ABI)."""
import ctypes

import numpy as np
import pytest

import _oracle as O
from isai_kernel_amd import DeviceArray, call, lib, problems, ptr

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(2024)
TOL64 = 1e-12  # FP64 relative tolerance of the path


def _dev(*arrays):
    return [DeviceArray.from_numpy(a) for a in arrays]


def _spmv_gpu(rp, ci, val, x, prec=0, dval=None):
    nb, bs = len(rp) - 1, val.shape[1]
    d_rp, d_ci, d_x = _dev(rp, ci, x)
    d_v = dval if dval is not None else DeviceArray.from_numpy(val)
    d_y = DeviceArray((nb * bs,), np.float64)
    if prec == 0:
        call("isai_bsr_spmv", nb, len(ci), bs, d_rp.ptr, d_ci.ptr, d_v.ptr, d_x.ptr, d_y.ptr, None)
    else:
        call("isai_bsr_spmv_prec", nb, len(ci), bs, prec, d_rp.ptr, d_ci.ptr, d_v.ptr, d_x.ptr, d_y.ptr, None)
    call("isai_device_synchronize")
    return d_y.to_numpy()


def _scale(rp, ci, val, x):
    return O.spmv(rp, ci, np.abs(val), np.abs(x)).max()


@pytest.mark.parametrize("problem,L", [(0, 16), (1, 9), (2, 6), (3, 13), (4, 20)])
def test_spmv_fp64_matches_oracle(gpu, problem, L):
    print(f"spmv fp64 problem={problem} L={L}", flush=True)
    rp, ci, val = problems.generate(problem, L, seed=11)
    x = RNG.standard_normal((len(rp) - 1) * 4)
    y = _spmv_gpu(rp, ci, val, x)
    ref = O.spmv(rp, ci, val, x)
    assert np.max(np.abs(y - ref)) <= TOL64 * _scale(rp, ci, val, x)
    assert np.array_equal(y, _spmv_gpu(rp, ci, val, x))  # deterministic


def _ragged(nb, maxlen, seed):
    """Random block pattern with row lengths 0..maxlen (sorted unique columns)."""
    rng = np.random.default_rng(seed)
    rows = [np.sort(rng.choice(nb, size=rng.integers(0, maxlen + 1), replace=False)) for _ in range(nb)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    ci = np.concatenate(rows).astype(np.int32)
    val = rng.standard_normal((len(ci), 4, 4))
    return rp, ci, val


@pytest.mark.parametrize("maxlen", [6, 12, 60])
def test_spmv_fp64_ragged_rows(gpu, maxlen):
    """Row lengths across the product's batch widths (4, 8, 16 blocks in flight,
    picked from the average length) and several batches per row."""
    print(f"spmv fp64 ragged maxlen={maxlen}", flush=True)
    rp, ci, val = _ragged(300, maxlen, seed=maxlen)
    x = RNG.standard_normal(300 * 4)
    y = _spmv_gpu(rp, ci, val, x)
    assert np.max(np.abs(y - O.spmv(rp, ci, val, x))) <= TOL64 * _scale(rp, ci, val, x)
    assert np.all(y[np.repeat(np.diff(rp) == 0, 4)] == 0)


@pytest.mark.parametrize("bs", [1, 2, 3, 5, 7, 8])
def test_spmv_generic_block_sizes(gpu, bs):
    print(f"spmv generic bs={bs}", flush=True)
    nb = 37
    mask = RNG.random((nb, nb)) < 0.2
    np.fill_diagonal(mask, True)
    rp = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int32)
    ci = np.concatenate([np.nonzero(r)[0] for r in mask]).astype(np.int32)
    val = RNG.standard_normal((len(ci), bs, bs))
    x = RNG.standard_normal(nb * bs)
    y = _spmv_gpu(rp, ci, val, x)
    assert np.max(np.abs(y - O.spmv(rp, ci, val, x))) <= TOL64 * _scale(rp, ci, val, x)


def test_spmv_empty_rows_and_zero_rows(gpu):
    print("spmv empty rows", flush=True)
    rp = np.array([0, 0, 2, 2, 3], np.int32)
    ci = np.array([0, 3, 1], np.int32)
    val = RNG.standard_normal((3, 4, 4))
    x = RNG.standard_normal(16)
    y = _spmv_gpu(rp, ci, val, x)
    assert np.max(np.abs(y - O.spmv(rp, ci, val, x))) <= TOL64 * _scale(rp, ci, val, x)
    assert np.all(y[0:4] == 0) and np.all(y[8:12] == 0)


@pytest.mark.parametrize("case", ["27pt", "7pt", "lower7pt", "ragged"])
@pytest.mark.parametrize("prec,bf16", [(1, False), (2, True)])
def test_spmv_reduced_precision_within_fp32_accumulation_bound(gpu, prec, bf16, case):
    """Long rows (27 blocks), short rows (7 and 4 blocks: several rows per lane group)
    and ragged rows with empty ones, each against the oracle's rounded product."""
    print(f"spmv prec={prec} case={case}", flush=True)
    if case == "ragged":
        rp, ci, val = _ragged(300, 40, seed=5)
    else:
        rp, ci, val = problems.generate(2 if case == "27pt" else 1, 7, seed=3)
    if case == "lower7pt":
        keep = ci <= np.repeat(np.arange(len(rp) - 1), np.diff(rp))
        rp = np.concatenate([[0], np.cumsum(np.add.reduceat(keep, rp[:-1]))]).astype(np.int32)
        ci, val = ci[keep].copy(), val[keep].copy()
    x = RNG.standard_normal((len(rp) - 1) * 4)
    d_val64 = DeviceArray.from_numpy(val)
    d_low = DeviceArray((val.size,), np.float32 if prec == 1 else np.uint16)
    call("isai_convert_values", val.size, prec, d_val64.ptr, d_low.ptr, None)
    call("isai_device_synchronize")
    low = d_low.to_numpy()
    if prec == 1:
        assert np.array_equal(low, val.reshape(-1).astype(np.float32))
    else:
        exp = np.array([O.ora.ora_round_bf16(v) for v in val.reshape(-1)[:4096]], np.float32)
        got = (low[:4096].astype(np.uint32) << 16).view(np.float32)
        assert np.array_equal(got, exp)
    y = _spmv_gpu(rp, ci, val, x, prec=prec, dval=d_low)
    ref = O.spmv_f32acc(rp, ci, val, x, bf16=bf16)
    # same rounded operands as the oracle, FP32 accumulation in another order:
    # |y - ref| <= 2 * len * 2^-24 * (|A_rounded| |x|) row by row
    absval = np.abs(val.astype(np.float32).astype(np.float64)) if not bf16 else np.abs(
        (low.astype(np.uint32) << 16).view(np.float32).astype(np.float64).reshape(val.shape))
    bound = O.spmv(rp, ci, absval, np.abs(x)) * (2 * int(np.diff(rp).max()) * 4 * 2.0**-24)
    assert np.all(np.abs(y - ref) <= bound + 1e-300)
    assert np.array_equal(y, _spmv_gpu(rp, ci, val, x, prec=prec, dval=d_low))  # deterministic


def test_isai_apply_two_products(gpu):
    print("isai apply", flush=True)
    rp, ci, val = problems.generate(0, 12, seed=4)
    trpL, tciL = problems.tri_bcsr_pattern(rp, ci, True)
    trpU, tciU = problems.tri_bcsr_pattern(rp, ci, False)
    ML = RNG.standard_normal((len(tciL), 4, 4))
    MU = RNG.standard_normal((len(tciU), 4, 4))
    nb = len(rp) - 1
    b = RNG.standard_normal(nb * 4)
    dl = _dev(trpL, tciL, ML, trpU, tciU, MU, b)
    d_y = DeviceArray((nb * 4,), np.float64)
    d_x = DeviceArray((nb * 4,), np.float64)
    call("isai_apply", nb, 4, 0, dl[0].ptr, dl[1].ptr, dl[2].ptr, len(tciL), dl[3].ptr, dl[4].ptr,
         dl[5].ptr, len(tciU), dl[6].ptr, d_y.ptr, d_x.ptr, None)
    call("isai_device_synchronize")
    ref = O.spmv(trpU, tciU, MU, O.spmv(trpL, tciL, ML, b))
    assert np.allclose(d_x.to_numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_reorder_and_bsr2csr_bitexact(gpu):
    print("reorder + bsr2csr", flush=True)
    rp, ci, val = problems.generate(1, 6, seed=5)
    mb, mc = problems.bcsc_pattern(rp, ci, True)
    trp, tci = problems.tri_bcsr_pattern(rp, ci, True)
    mp = problems.map_bcsc2bcsr(mb, mc, trp, tci)
    src = RNG.standard_normal((len(mc), 4, 4))
    d_map, d_src = _dev(mp, src)
    d_dst = DeviceArray(src.shape, np.float64)
    call("isai_reorder_bcsc_to_bcsr", len(mc), 4, d_map.ptr, d_src.ptr, d_dst.ptr, None)
    call("isai_device_synchronize")
    assert d_dst.to_numpy().tobytes() == O.reorder(mp, src).tobytes()
    nb = len(rp) - 1
    nnz = len(ci) * 16
    d_rp, d_ci, d_v = _dev(rp, ci, val)
    o_rp = DeviceArray((nb * 4 + 1,), np.int32)
    o_ci = DeviceArray((nnz,), np.int32)
    o_v = DeviceArray((nnz,), np.float64)
    call("isai_bsr_to_csr", nb, len(ci), 4, d_rp.ptr, d_ci.ptr, d_v.ptr, o_rp.ptr, o_ci.ptr, o_v.ptr, None)
    call("isai_device_synchronize")
    crp, cci, cv = O.bsr2csr(rp, ci, val)
    assert np.array_equal(o_rp.to_numpy(), crp) and np.array_equal(o_ci.to_numpy(), cci)
    assert o_v.to_numpy().tobytes() == cv.tobytes()


def test_spmv_full_size_properties(gpu):
    """96^3 7-point (the benchmark matrix): linearity and the column-sum checksum."""
    print("spmv full-size 96^3", flush=True)
    rp, ci, val = problems.generate(1, 96, seed=0)
    nb = len(rp) - 1
    n = nb * 4
    x = RNG.standard_normal(n)
    z = RNG.standard_normal(n)
    d_rp, d_ci, d_v = _dev(rp, ci, val)
    outs = []
    for v in (x, z, x + 2.0 * z):
        d_x = DeviceArray.from_numpy(v)
        d_y = DeviceArray((n,), np.float64)
        call("isai_bsr_spmv", nb, len(ci), 4, d_rp.ptr, d_ci.ptr, d_v.ptr, d_x.ptr, d_y.ptr, None)
        call("isai_device_synchronize")
        outs.append(d_y.to_numpy())
    scale = np.abs(outs[0]).max() + 2 * np.abs(outs[1]).max()
    assert np.max(np.abs(outs[2] - (outs[0] + 2.0 * outs[1]))) <= 1e-12 * scale * 10
    # 1^T (A x) == (A^T 1)^T x, with A^T 1 summed on the host from the BSR arrays
    colsum = np.zeros(n)
    cols = (ci[:, None] * 4 + np.arange(4)[None, :])  # [nnzb, 4]
    np.add.at(colsum, cols.reshape(-1), val.sum(axis=1).reshape(-1))
    lhs, rhs = outs[0].sum(), colsum @ x
    assert abs(lhs - rhs) <= 1e-10 * (np.abs(colsum) @ np.abs(x))
    # spot rows against the oracle restricted to a slice of block rows
    sl = slice(400000, 400200)
    ref = np.array([sum(val[k] @ x[ci[k] * 4:ci[k] * 4 + 4] for k in range(rp[i], rp[i + 1])) for i in range(sl.start, sl.stop)])
    assert np.allclose(outs[0].reshape(-1, 4)[sl], ref, rtol=1e-12, atol=1e-12)


def test_spmv_bench_host_form(gpu):
    print("spmv bench host form", flush=True)
    rp, ci, val = problems.generate(0, 32, seed=1)
    nb = len(rp) - 1
    x = RNG.standard_normal(nb * 4)
    y = np.empty(nb * 4)
    ms = ctypes.c_double()
    call("isai_bsr_spmv_bench", ptr(rp), ptr(ci), ptr(val), nb, len(ci), 4, ptr(x), ptr(y), 2, 5,
         ctypes.byref(ms))
    assert ms.value > 0
    assert np.allclose(y, O.spmv(rp, ci, val, x), rtol=1e-12, atol=1e-12)
