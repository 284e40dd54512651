"""GPU parity of the device CSR -> BSR conversion (include/csr2bsr.h device form):
bit-identical to the host form and to the oracle."""
import ctypes
import json
from pathlib import Path

import numpy as np
import pytest

import _oracle as O
from gpu_pipeline import dev
from isai_kernel_amd import DeviceArray, IsaiError, call, lib, problems

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
RNG = np.random.default_rng(515)


def device_csr2bsr(n, bs, rp, ci, v):
    d_rp, d_ci, d_v = dev(np.asarray(rp, np.int32)), dev(np.asarray(ci, np.int32)), dev(np.asarray(v, np.float64))
    nb = n // bs
    d_brp = DeviceArray((nb + 1,), np.int32)
    nnzb = ctypes.c_int()
    call("isai_csr_to_bsr_device_count", n, bs, d_rp.ptr, d_ci.ptr, d_brp.ptr, ctypes.byref(nnzb), None)
    d_bci = DeviceArray((max(nnzb.value, 1),), np.int32)
    d_bv = DeviceArray((max(nnzb.value, 1), bs, bs), np.float64)
    call("isai_csr_to_bsr_device_fill", n, bs, d_rp.ptr, d_ci.ptr, d_v.ptr, d_brp.ptr, d_bci.ptr, d_bv.ptr, None)
    return d_brp.to_numpy(), d_bci.to_numpy()[:nnzb.value], d_bv.to_numpy()[:nnzb.value]


def shuffled_duplicated(crp, cci, cv, frac=0.05):
    n = len(crp) - 1
    rows = np.repeat(np.arange(n), np.diff(crp))
    dup = RNG.random(len(cci)) < frac
    rows2 = np.concatenate([rows, rows[dup]])
    cols2 = np.concatenate([cci, cci[dup]])
    vals2 = np.concatenate([cv, 0.125 * cv[dup]])
    order = np.lexsort((RNG.random(len(rows2)), rows2))
    rows2, cols2, vals2 = rows2[order], cols2[order], vals2[order]
    rp2 = np.zeros(n + 1, np.int64)
    np.add.at(rp2, rows2 + 1, 1)
    return np.cumsum(rp2).astype(np.int32), cols2.astype(np.int32), vals2


def test_device_csr2bsr_hand_fixture(gpu):
    print("device csr2bsr hand fixture", flush=True)
    d = json.loads((ROOT / "tests/golden/csr2bsr_small.json").read_text())
    v = np.array([float(x) for x in d["csr_val"]])
    exp = np.array([[[float(x) for x in r] for r in b] for b in d["bsr_val"]])
    brp, bci, bv = device_csr2bsr(d["n"], d["bs"], d["csr_rowptr"], d["csr_colind"], v)
    assert brp.tolist() == d["bsr_rowptr"] and bci.tolist() == d["bsr_colind"]
    assert bv.tobytes() == exp.tobytes()  # includes the -0.0 first-entry rule


@pytest.mark.parametrize("problem,L", [(0, 16), (1, 7), (2, 5), (3, 11)])
def test_device_csr2bsr_shuffled_duplicated(gpu, problem, L):
    print(f"device csr2bsr problem {problem} L={L}", flush=True)
    rp, ci, val = problems.generate(problem, L, seed=7)
    crp, cci, cv = O.bsr2csr(rp, ci, val)
    rp2, ci2, v2 = shuffled_duplicated(crp, cci, cv)
    n = len(rp2) - 1
    got = device_csr2bsr(n, 4, rp2, ci2, v2)
    for x, y in zip(got, O.csr2bsr(n, 4, rp2, ci2, v2)):
        assert x.tobytes() == y.tobytes()
    for x, y in zip(got, problems.csr_to_bsr(n, 4, rp2, ci2, v2)):
        assert x.tobytes() == y.tobytes()


@pytest.mark.parametrize("bs", [1, 2, 3, 5, 8])
def test_device_csr2bsr_block_sizes_and_empty_rows(gpu, bs):
    """Random scalar pattern with empty rows, explicit zeros and -0.0 values."""
    print(f"device csr2bsr random bs={bs}", flush=True)
    n = bs * 37
    rows, cols, vals = [], [], []
    for r in range(n):
        if r % 7 == 3:
            continue  # empty scalar row (and some empty block rows for bs = 1)
        k = RNG.integers(1, 12)
        for c in RNG.integers(0, n, k):
            rows.append(r)
            cols.append(c)
            vals.append(RNG.choice([0.0, -0.0, RNG.standard_normal()]))
    rows, cols, vals = np.array(rows), np.array(cols, np.int32), np.array(vals)
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    got = device_csr2bsr(n, bs, rp, cols, vals)
    for x, y in zip(got, O.csr2bsr(n, bs, rp, cols, vals)):
        assert x.tobytes() == y.tobytes()


def test_device_csr2bsr_full_size_round_trip(gpu):
    """96^3 7-pt: BSR -> CSR (device) -> BSR (device) reproduces the generator's
    arrays bit for bit."""
    print("device csr2bsr 96^3 round trip", flush=True)
    rp, ci, val = problems.generate(1, 96, seed=0)
    nb, nnzb = len(rp) - 1, len(ci)
    d_rp, d_ci, d_v = dev(rp), dev(ci), dev(val)
    n = nb * 4
    d_crp = DeviceArray((n + 1,), np.int32)
    d_cci = DeviceArray((nnzb * 16,), np.int32)
    d_cv = DeviceArray((nnzb * 16,), np.float64)
    call("isai_bsr_to_csr", nb, nnzb, 4, d_rp.ptr, d_ci.ptr, d_v.ptr, d_crp.ptr, d_cci.ptr, d_cv.ptr, None)
    d_brp = DeviceArray((nb + 1,), np.int32)
    cnt = ctypes.c_int()
    call("isai_csr_to_bsr_device_count", n, 4, d_crp.ptr, d_cci.ptr, d_brp.ptr, ctypes.byref(cnt), None)
    assert cnt.value == nnzb
    d_bci = DeviceArray((nnzb,), np.int32)
    d_bv = DeviceArray((nnzb, 4, 4), np.float64)
    call("isai_csr_to_bsr_device_fill", n, 4, d_crp.ptr, d_cci.ptr, d_cv.ptr, d_brp.ptr, d_bci.ptr, d_bv.ptr, None)
    assert np.array_equal(d_brp.to_numpy(), rp) and np.array_equal(d_bci.to_numpy(), ci)
    assert d_bv.to_numpy().tobytes() == val.tobytes()


def test_device_csr2bsr_errors(gpu):
    print("device csr2bsr errors", flush=True)
    rp = dev(np.zeros(6, np.int32))
    ci = dev(np.zeros(1, np.int32))
    brp = DeviceArray((2,), np.int32)
    nnzb = ctypes.c_int()
    assert lib.isai_csr_to_bsr_device_count(5, 4, rp.ptr, ci.ptr, brp.ptr, ctypes.byref(nnzb), None) == -1
    # a row pointer that does not come from the count pass is refused
    crp = np.array([0, 1, 2, 2, 3], np.int32)
    cci = np.array([0, 5, 2], np.int32)
    d = [dev(a) for a in (crp, cci, np.ones(3))]
    bad = dev(np.array([0, 1, 2], np.int32))
    out_c, out_v = DeviceArray((4,), np.int32), DeviceArray((4, 2, 2), np.float64)
    with pytest.raises(IsaiError):
        call("isai_csr_to_bsr_device_fill", 4, 2, d[0].ptr, d[1].ptr, d[2].ptr, bad.ptr, out_c.ptr, out_v.ptr, None)
